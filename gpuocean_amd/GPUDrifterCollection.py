"""Lagrangian drifters advected by a CDKLM16 simulator on the GPU.

Drop-in for the reference's ``drifters/GPUDrifterCollection.py`` (with the
position bookkeeping of ``drifters/BaseDrifterCollection.py``): N drifters plus
one observation drifter stored as an (N+1) x 2 float32 array in HBM, advected
by ``gpo_drift`` (bilinear cell-centre velocity, optional wind drag, forward
Euler, periodic wrap).
"""

import ctypes

import numpy as np
import torch

from . import Common
from . import _lib as L
from .forcing import TimeLevels, WindStress


class GPUDrifterCollection:
    def __init__(self, gpu_ctx, numDrifters, observation_variance=0.01, boundaryConditions=None,
                 initialization_cov_drifters=None, domain_size_x=1.0, domain_size_y=1.0, gpu_stream=None,
                 initialize=False, wind=None, wind_drift_factor=0.0, block_width=64):
        L.lib()
        self.gpu_ctx = gpu_ctx
        self.numDrifters = int(numDrifters)
        self.obs_index = self.numDrifters
        self.observation_variance = observation_variance
        self.boundaryConditions = Common.BoundaryConditions() if boundaryConditions is None else boundaryConditions
        self.domain_size_x = domain_size_x
        self.domain_size_y = domain_size_y
        self.wind = WindStress() if wind is None else wind
        self.wind_drift_factor = np.float32(wind_drift_factor)
        self.sensitivity = 1.0
        self.device = gpu_ctx.device if gpu_ctx is not None and hasattr(gpu_ctx, "device") \
            else torch.device("cuda", torch.cuda.current_device())
        self.gpu_stream = gpu_stream if gpu_stream is not None else torch.cuda.Stream(device=self.device)
        with torch.cuda.stream(self.gpu_stream):
            self.positions = torch.zeros((self.numDrifters + 1, 2), dtype=torch.float32, device=self.device)
        self._wind_levels = TimeLevels(self.wind.t)
        self._wind_bufs = {}
        if initialize:
            self.uniformly_distribute_drifters(initialization_cov_drifters=initialization_cov_drifters)

    # -------------------------------------------------------------- bookkeeping
    def setGPUStream(self, gpu_stream):
        self.gpu_stream = gpu_stream

    def getNumDrifters(self):
        return self.numDrifters

    def getObservationVariance(self):
        return self.observation_variance

    def getBoundaryConditions(self):
        return self.boundaryConditions

    def getDomainSizeX(self):
        return self.domain_size_x

    def getDomainSizeY(self):
        return self.domain_size_y

    def setDomainSize(self, size_x, size_y):
        self.domain_size_x, self.domain_size_y = size_x, size_y

    def _all(self):
        self.gpu_stream.synchronize()
        return self.positions.cpu().numpy()

    def getDrifterPositions(self):
        return self._all()[:-1, :]

    def getObservationPosition(self):
        return self._all()[self.obs_index, :]

    def setDrifterPositions(self, newDrifterPositions):
        allp = np.concatenate((np.asarray(newDrifterPositions, dtype=np.float32).reshape(-1, 2),
                               np.asarray([self.getObservationPosition()], dtype=np.float32)), axis=0)
        assert allp.shape == (self.numDrifters + 1, 2)
        with torch.cuda.stream(self.gpu_stream):
            self.positions.copy_(torch.from_numpy(np.ascontiguousarray(allp)))

    def setObservationPosition(self, newObservationPosition):
        allp = np.concatenate((self.getDrifterPositions(),
                               np.asarray([newObservationPosition], dtype=np.float32)), axis=0)
        with torch.cuda.stream(self.gpu_stream):
            self.positions.copy_(torch.from_numpy(np.ascontiguousarray(allp, dtype=np.float32)))

    def setSensitivity(self, sensitivity):
        self.sensitivity = sensitivity

    def uniformly_distribute_drifters(self, initialization_cov_drifters=None):
        """Drifters at the centres of a near-square grid of sub-domains, optionally
        perturbed (drifters/BaseDrifterCollection.py:141-173)."""
        sub_y = int(np.round(np.sqrt(self.numDrifters)))
        sub_x = int(np.ceil(1.0 * self.numDrifters / sub_y))
        pos = np.empty((self.numDrifters, 2))
        for sy in range(sub_y):
            for sx in range(sub_x):
                d = sy * sub_x + sx
                if d >= self.numDrifters:
                    break
                pos[d, 0] = (sx + 0.5) * self.domain_size_x / sub_x
                pos[d, 1] = (sy + 0.5) * self.domain_size_y / sub_y
        if initialization_cov_drifters is None:
            initialization_cov_drifters = np.zeros((2, 2))
        elif np.isscalar(initialization_cov_drifters):
            initialization_cov_drifters = np.eye(2) * initialization_cov_drifters
        for d in range(self.numDrifters):
            pos[d, :] = np.random.multivariate_normal(pos[d, :], initialization_cov_drifters)
        self.setDrifterPositions(pos)

    def copy(self):
        c = GPUDrifterCollection(self.gpu_ctx, self.numDrifters, observation_variance=self.observation_variance,
                                 boundaryConditions=self.boundaryConditions, domain_size_x=self.domain_size_x,
                                 domain_size_y=self.domain_size_y, gpu_stream=self.gpu_stream, wind=self.wind,
                                 wind_drift_factor=self.wind_drift_factor)
        with torch.cuda.stream(self.gpu_stream):
            c.positions.copy_(self.positions)
        return c

    def cleanUp(self):
        self.positions = None

    # ------------------------------------------------------------------ drift
    def _wind_args(self, t):
        if not self.wind_drift_factor:
            return None, None, None, None, 1, 1, 0.0
        t0, t1, c0, c1, w = self._wind_levels.update(t)
        with torch.cuda.stream(self.gpu_stream):
            if c0 or 'xc' not in self._wind_bufs:
                self._wind_bufs['xc'] = torch.from_numpy(np.ascontiguousarray(self.wind.wind_u[t0], np.float32)).to(self.device)
                self._wind_bufs['yc'] = torch.from_numpy(np.ascontiguousarray(self.wind.wind_v[t0], np.float32)).to(self.device)
            if c1 or 'xn' not in self._wind_bufs:
                self._wind_bufs['xn'] = torch.from_numpy(np.ascontiguousarray(self.wind.wind_u[t1], np.float32)).to(self.device)
                self._wind_bufs['yn'] = torch.from_numpy(np.ascontiguousarray(self.wind.wind_v[t1], np.float32)).to(self.device)
        ny_, nx_ = np.asarray(self.wind.wind_u[0]).shape
        b = self._wind_bufs
        return (L.ptr(b['xc']), L.ptr(b['xn']), L.ptr(b['yc']), L.ptr(b['yn']), nx_, ny_, float(np.float32(w)))

    def drift(self, eta, hu, hv, Hm, nx, ny, t, dx, dy, dt, x_zero_ref=2, y_zero_ref=2):
        """One Euler step of all drifters (GPUDrifterCollection.py:210-232);
        eta/hu/hv/Hm are DeviceArray2D buffers of the simulator."""
        wxc, wxn, wyc, wyn, wnx, wny, wt = self._wind_args(t)
        bc = self.boundaryConditions
        L.check(L.lib().gpo_drift(
            int(nx), int(ny), float(dx), float(dy), float(dt), int(x_zero_ref), int(y_zero_ref),
            ctypes.c_void_p(eta.data_ptr()), ctypes.c_void_p(hu.data_ptr()), ctypes.c_void_p(hv.data_ptr()),
            eta.pitch, ctypes.c_void_p(Hm.data_ptr()), Hm.pitch, int(bc.isPeriodicNorthSouth()),
            int(bc.isPeriodicEastWest()), self.numDrifters, ctypes.c_void_p(self.positions.data_ptr()), 8,
            float(np.float32(self.sensitivity)), wxc, wxn, wyc, wyn, int(wnx), int(wny), wt,
            float(self.wind_drift_factor), ctypes.c_void_p(self.gpu_stream.cuda_stream)), "gpo_drift")

    def driftFromSim(self, sim, dt):
        self.drift(sim.h0[0], sim.h0[1], sim.h0[2], sim.Bm, sim.nx, sim.ny, sim.t, sim.dx, sim.dy, dt, 2, 2)

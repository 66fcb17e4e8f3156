"""SOAR-correlated, geostrophically balanced model error for CDKLM16.

Drop-in for the reference's ``SWEsimulators/OceanStateNoise.py`` (the parts on
the CDKLM16 path): same constructor arguments, ``fromsim``, ``perturbSim``,
``perturbOceanState``, ``generateNormalDistribution`` and the buffer getters.

The random field comes from a counter-based Philox4x32-10 generator keyed by
the seed (``xorwow_seed``/``np_seed``), with counter (element, draw, stream), so
every draw is reproducible and any part of it can be regenerated on the fly.
The reference's XORWOW (curand) and LCG streams are not reproduced; the
statistical properties it tests are (zero mean, unit variance, seeded
reproducibility).

With ``interpolation_factor`` > 1 the random field and SOAR live on a grid
that factor times coarser, and ``gpo_bicubic_interpolation`` carries the
perturbation to the computational grid, optionally aligned with a given cell
(``align_with_cell_i/j``), with its balanced momentum
(OceanStateNoise.py:418-441, 461-475).
"""

import ctypes

import numpy as np
import torch

from . import _lib as L

SOAR_CUTOFF = 2  # utils/config.py:31


def soar_weights(dx, dy, q0, L_):
    """5x5 SOAR covariance weights q0 (1 + d/L) exp(-d/L) of the stencil offsets
    (gpu_kernels/ocean_noise.cu:126-132), row-major from offset (-2, -2)."""
    w = np.zeros((5, 5), dtype=np.float64)
    for k in range(-2, 3):
        for l_ in range(-2, 3):
            d = np.sqrt((float(dx) * l_) ** 2 + (float(dy) * k) ** 2)
            w[k + 2, l_ + 2] = float(q0) * (1.0 + d / float(L_)) * np.exp(-d / float(L_))
    return np.ascontiguousarray(w.astype(np.float32))


class OceanStateNoise:
    """Small-scale perturbations of (eta, hu, hv): SOAR-correlated d_eta plus the
    momentum that keeps it in geostrophic balance (OceanStateNoise.py:40-459)."""

    def __init__(self, gpu_ctx, gpu_stream, nx, ny, dx, dy, boundaryConditions, staggered,
                 soar_q0=None, soar_L=None, interpolation_factor=1, use_lcg=False, xorwow_seed=None,
                 np_seed=None, angle=np.array([[0]], dtype=np.float32),
                 coriolis_f=np.array([[0]], dtype=np.float32), block_width=16, block_height=16,
                 device=None, stream_id=0):
        assert interpolation_factor > 0 and interpolation_factor % 2 == 1, \
            'interpolation_factor must be a positive odd integer'
        if boundaryConditions.isPeriodicNorthSouth():
            assert ny % interpolation_factor == 0, 'ny must be divisible by the interpolation factor'
        if boundaryConditions.isPeriodicEastWest():
            assert nx % interpolation_factor == 0, 'nx must be divisible by the interpolation factor'
        if staggered:
            raise NotImplementedError("staggered grids (FBL/CTCS) are not part of gpuocean_amd")
        L.lib()
        self.gpu_ctx = gpu_ctx
        self.gpu_stream = gpu_stream
        self.device = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.nx, self.ny = np.int32(nx), np.int32(ny)
        self.dx, self.dy = np.float32(dx), np.float32(dy)
        self.use_lcg = use_lcg
        self.cutoff = np.int32(SOAR_CUTOFF)
        self.interpolation_factor = np.int32(interpolation_factor)
        self.periodicNorthSouth = np.int32(boundaryConditions.isPeriodicNorthSouth())
        self.periodicEastWest = np.int32(boundaryConditions.isPeriodicEastWest())
        self.coarse_nx = np.int32(np.ceil(nx / self.interpolation_factor))
        self.coarse_ny = np.int32(np.ceil(ny / self.interpolation_factor))
        self.coarse_dx = np.float32(dx * self.interpolation_factor)
        self.coarse_dy = np.float32(dy * self.interpolation_factor)
        self.rand_ghost_cells_x = np.int32(0 if self.periodicEastWest else 2 + self.cutoff)
        self.rand_ghost_cells_y = np.int32(0 if self.periodicNorthSouth else 2 + self.cutoff)
        self.rand_nx = np.int32(self.coarse_nx + 2 * self.rand_ghost_cells_x)
        self.rand_ny = np.int32(self.coarse_ny + 2 * self.rand_ghost_cells_y)

        seed = xorwow_seed if xorwow_seed is not None else np_seed
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 31 - 1))
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.stream_id = int(stream_id)
        self.draw = 0
        # Draw index whose normal field random_numbers holds, and the
        # (draw, q0_scale, perturbation_scale) coarse_buffer was formed from.
        # Batched ensemble steps apply the model error without filling these
        # buffers; the getters regenerate them from the counter-based stream.
        self._xi_draw = None
        self._coarse_key = None
        self._last_scales = (1.0, 1.0)

        self.soar_q0 = np.float32(self.dx / 100000) if soar_q0 is None else np.float32(soar_q0)
        self.soar_L = np.float32(0.75 * self.coarse_dx) if soar_L is None else np.float32(soar_L)

        with torch.cuda.stream(gpu_stream):
            self.random_numbers = torch.zeros((int(self.rand_ny), int(self.rand_nx)), dtype=torch.float32,
                                              device=self.device)
            self.coarse_buffer = torch.zeros((int(self.coarse_ny) + 4, int(self.coarse_nx) + 4),
                                             dtype=torch.float32, device=self.device)
        self.angle = np.asarray(angle)
        self._f_cells = None  # per-cell coriolis f (DeviceArray2D) when f is a texture; set by fromsim

    @classmethod
    def fromsim(cls, sim, soar_q0=None, soar_L=None, interpolation_factor=1, use_lcg=False, xorwow_seed=None,
                block_width=16, block_height=16, stream_id=None):
        obj = cls(sim.gpu_ctx, sim.gpu_stream, sim.nx, sim.ny, sim.dx, sim.dy, sim.boundary_conditions, False,
                  soar_q0=soar_q0, soar_L=soar_L, interpolation_factor=interpolation_factor, use_lcg=use_lcg,
                  xorwow_seed=xorwow_seed, angle=sim.angle, device=sim.device,
                  stream_id=sim.local_particle_id if stream_id is None else stream_id)
        obj._f_cells = sim.f_cells
        return obj

    def _stream(self):
        return ctypes.c_void_p(self.gpu_stream.cuda_stream)

    def getSeed(self):
        return self.seed

    def resetSeed(self):
        self.draw = 0

    def _refresh_random_numbers(self):
        """random_numbers = the normal field of the latest draw."""
        if self.draw > 0 and self._xi_draw != self.draw - 1:
            self._fill(self.draw - 1)

    def _refresh_coarse(self):
        if self.draw > 0 and self._coarse_key != (self.draw - 1,) + tuple(self._last_scales):
            self._refresh_random_numbers()
            self._soar(*self._last_scales)

    def getRandomNumbers(self):
        self._refresh_random_numbers()
        self.gpu_stream.synchronize()
        return self.random_numbers.cpu().numpy()

    def getCoarseBuffer(self):
        self._refresh_coarse()
        self.gpu_stream.synchronize()
        return self.coarse_buffer.cpu().numpy()

    def _fill(self, draw):
        rows, cols = self.random_numbers.shape
        L.check(L.lib().gpo_random_normal(ctypes.c_void_p(self.random_numbers.data_ptr()), rows, cols, 4 * cols,
                                          self.seed, self.stream_id, int(draw), self._stream()),
                "gpo_random_normal")
        self._xi_draw = int(draw)

    def _soar(self, q0_scale, perturbation_scale):
        W = self.weights(q0_scale)
        cols = int(self.rand_nx)
        L.check(L.lib().gpo_soar(int(self.coarse_nx), int(self.coarse_ny), W.ctypes.data,
                                 float(np.float32(perturbation_scale)), int(self.periodicNorthSouth),
                                 int(self.periodicEastWest), ctypes.c_void_p(self.random_numbers.data_ptr()),
                                 4 * cols, ctypes.c_void_p(self.coarse_buffer.data_ptr()),
                                 4 * (int(self.coarse_nx) + 4), 0, self._stream()), "gpo_soar")
        self._coarse_key = (self._xi_draw, q0_scale, perturbation_scale)

    def generateNormalDistribution(self):
        """Next standard normal field into random_numbers (one draw)."""
        self._fill(self.draw)
        self.draw += 1

    def weights(self, q0_scale=1.0):
        return soar_weights(self.coarse_dx, self.coarse_dy, np.float32(self.soar_q0 * q0_scale), self.soar_L)

    def perturbSim(self, sim, q0_scale=1.0, update_random_field=True, perturbation_scale=1.0,
                   perpendicular_scale=0.0, align_with_cell_i=None, align_with_cell_j=None, stream=None):
        """Add a balanced perturbation to sim's latest state (OceanStateNoise.py:333-353)."""
        self.perturbOceanState(sim.h0[0], sim.h0[1], sim.h0[2], sim.Hi, sim.f, beta=sim.coriolis_beta, g=sim.g,
                               q0_scale=q0_scale, update_random_field=update_random_field,
                               perturbation_scale=perturbation_scale, perpendicular_scale=perpendicular_scale,
                               align_with_cell_i=align_with_cell_i, align_with_cell_j=align_with_cell_j,
                               land_mask_value=sim.mask_value, north=sim.params)

    def perturbSimSimilarAs(self, simToPerturb, modelError=None, perturbation_scale=1.0):
        """Perturb simToPerturb with modelError's current random field (the
        SOAR counterpart of ModelErrorKL.perturbSimSimilarAs, ModelErrorKL.py:450)."""
        assert modelError.random_numbers.shape == self.random_numbers.shape
        modelError._refresh_random_numbers()
        modelError.gpu_stream.synchronize()
        with torch.cuda.stream(self.gpu_stream):
            self.random_numbers.copy_(modelError.random_numbers)
        self.perturbSim(simToPerturb, update_random_field=False, perturbation_scale=perturbation_scale)

    def perturbOceanState(self, eta, hu, hv, H, f, beta=0.0, g=9.81, y0_reference_cell=0, ghost_cells_x=0,
                          ghost_cells_y=0, q0_scale=1.0, update_random_field=True, perturbation_scale=1.0,
                          perpendicular_scale=0.0, align_with_cell_i=None, align_with_cell_j=None,
                          land_mask_value=np.float32(1.0e20), stream=None, north=None):
        """Random field -> SOAR -> geostrophic balance, added to eta/hu/hv
        (OceanStateNoise.py:356-459).  eta/hu/hv/H are DeviceArray2D buffers."""
        if perpendicular_scale > 0:
            raise NotImplementedError("perpendicular perturbations are not implemented")
        if update_random_field:
            self.generateNormalDistribution()
        else:
            self._refresh_random_numbers()
        self._last_scales = (q0_scale, perturbation_scale)
        self._soar(q0_scale, perturbation_scale)
        lib = L.lib()
        nxv = ny_ = None
        if north is not None:
            nxv, ny_ = north.north_x, north.north_y
        else:
            nxv, ny_ = float(np.sin(self.angle.reshape(-1)[0])), float(np.cos(self.angle.reshape(-1)[0]))
        fc = self._f_cells
        f_scalar = float(np.float32(f)) if np.asarray(f).size == 1 else 0.0
        if self.interpolation_factor > 1:
            oi, oj = self._obtain_coarse_grid_offset(align_with_cell_i, align_with_cell_j)
            L.check(lib.gpo_bicubic_interpolation(
                int(self.nx), int(self.ny), float(self.dx), float(self.dy), int(self.coarse_nx), int(self.coarse_ny),
                float(self.coarse_dx), float(self.coarse_dy), int(oi), int(oj), float(np.float32(g)), f_scalar,
                float(np.float32(beta)), float(nxv), float(ny_),
                None if fc is None else ctypes.c_void_p(fc.data_ptr()), 0 if fc is None else fc.pitch,
                ctypes.c_void_p(self.coarse_buffer.data_ptr()), 4 * (int(self.coarse_nx) + 4),
                ctypes.c_void_p(eta.data_ptr()), ctypes.c_void_p(hu.data_ptr()), ctypes.c_void_p(hv.data_ptr()),
                eta.pitch, ctypes.c_void_p(H.data_ptr()), H.pitch, float(land_mask_value), self._stream()),
                "gpo_bicubic_interpolation")
            return
        L.check(lib.gpo_geostrophic_balance(
            int(self.nx), int(self.ny), float(self.dx), float(self.dy), float(np.float32(g)), f_scalar,
            float(np.float32(beta)), float(nxv), float(ny_),
            None if fc is None else ctypes.c_void_p(fc.data_ptr()), 0 if fc is None else fc.pitch,
            ctypes.c_void_p(self.coarse_buffer.data_ptr()), 4 * (int(self.coarse_nx) + 4),
            ctypes.c_void_p(eta.data_ptr()), ctypes.c_void_p(hu.data_ptr()), ctypes.c_void_p(hv.data_ptr()),
            eta.pitch, ctypes.c_void_p(H.data_ptr()), H.pitch, float(land_mask_value), self._stream()),
            "gpo_geostrophic_balance")

    def _obtain_coarse_grid_offset(self, fine_index_i, fine_index_j):
        """Fine-cell shift of the coarse grid that puts a coarse cell centre on
        the given fine cell (OceanStateNoise.py:461-475); 0 when not given."""
        default_offset = int(self.interpolation_factor) // 2
        offset_i = offset_j = 0
        if fine_index_i is not None:
            offset_i = -int(fine_index_i % int(self.interpolation_factor) - default_offset)
        if fine_index_j is not None:
            offset_j = -int(fine_index_j % int(self.interpolation_factor) - default_offset)
        return offset_i, offset_j

    def cleanUp(self, do_gc=True):
        self.random_numbers = None
        self.coarse_buffer = None

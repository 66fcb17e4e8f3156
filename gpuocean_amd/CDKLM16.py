"""CDKLM16 rotating shallow-water simulator on MI355X (gfx950).

Drop-in for the reference's ``SWEsimulators/CDKLM16.py`` class: the same
constructor arguments, ``step``, ``download``, ``upload``, ``updateDt``,
``setSOARModelError``, ``perturbState``, ``attachDrifters`` and friends.  The
device work is done by hand-written HIP kernels in ``libgpuocean_amd.so``
(include/gpuocean_amd.h); torch is used only to own device buffers and the
stream.

Differences in mechanism (results are the same):
  * kernel constants are run-time parameters, so no per-grid JIT;
  * periodic and wall ghost cells are resolved inside the step kernel while
    loading; the global ghost cells are written (as the reference's boundary
    kernels do) whenever something reads them: at the end of ``step``, before
    noise/drifters and on ``upload``;
  * the cell depth Hm is derived from the corner bathymetry inside the kernel.
"""

import ctypes
import logging

import numpy as np
import torch

from . import Common, host
from . import _lib as L
from .forcing import AtmosphericPressure, TimeLevels, WindStress


def _pitch_elems(width):
    """Row pitch in floats: rows start on 256-byte boundaries."""
    return ((width + 63) // 64) * 64


class DeviceArray2D:
    """A (rows, cols) float32 array in HBM with a padded row pitch
    (the role of utils/Common.py:CUDAArray2D)."""

    def __init__(self, rows, cols, device, host_data=None):
        self.rows, self.cols = rows, cols
        self.pitch_elems = _pitch_elems(cols)
        self.pitch = 4 * self.pitch_elems  # bytes, as the reference's kernels take it
        self.buf = torch.zeros((rows, self.pitch_elems), dtype=torch.float32, device=device)
        self.view = self.buf[:, :cols]
        if host_data is not None:
            self.upload(host_data)

    def upload(self, host_data):
        a = np.ascontiguousarray(host_data, dtype=np.float32)
        assert a.shape == (self.rows, self.cols), f"{a.shape} vs {(self.rows, self.cols)}"
        self.view.copy_(torch.from_numpy(a), non_blocking=False)

    def download(self):
        return self.view.cpu().numpy().copy()

    def data_ptr(self):
        return self.buf.data_ptr()

    def copy_from(self, other):
        self.buf.copy_(other.buf)


class CDKLM16:
    """Well-balanced central-upwind scheme of Chertock, Dudzinski, Kurganov and
    Lukacova-Medvidova (2017) for the rotating shallow-water equations."""

    def __init__(self, gpu_ctx, eta0, hu0, hv0, H, nx, ny, dx, dy, dt, g, f, r,
                 subsample_f=10,
                 angle=np.array([[0]], dtype=np.float32),
                 subsample_angle=10,
                 latitude=None,
                 rho_o=1025.0,
                 t=0.0,
                 theta=1.8, rk_order=2,
                 coriolis_beta=0.0,
                 max_wind_direction_perturbation=0,
                 wind=None,
                 wind_stress_factor=1.0,
                 atmospheric_pressure=None,
                 boundary_conditions=None,
                 boundary_conditions_data=None,
                 model_time_step=None,
                 reportGeostrophicEquilibrium=False,
                 write_netcdf=False,
                 comm=None,
                 local_particle_id=0,
                 super_dir_name=None,
                 netcdf_filename=None,
                 ignore_ghostcells=False,
                 courant_number=0.8,
                 offset_x=0, offset_y=0,
                 flux_slope_eps=1.0e-1,
                 desingularization_eps=1.0e-1,
                 depth_cutoff=1.0e-5,
                 one_dimensional=False,
                 flux_balancer=0.8,
                 block_width=12, block_height=32, num_threads_dt=256,
                 use_direct_lookup=False,
                 exact_arithmetic=True,
                 chunk_rows=None,
                 waves_per_simd=None):
        self.logger = logging.getLogger(__name__)
        assert 0 < rk_order < 4, "Only 1st, 2nd and 3rd order Runge Kutta supported"
        if rk_order == 3:
            assert r == 0.0, "3rd order Runge Kutta supported only without friction"
        if latitude is not None:
            raise NotImplementedError("latitude-based coriolis textures are not supported; pass f or f+beta")
        if comm is not None:
            raise NotImplementedError("MPI communicators (MPIOceanModelEnsemble) are not part of gpuocean_amd")
        if reportGeostrophicEquilibrium:
            raise NotImplementedError("reportGeostrophicEquilibrium is not supported")
        if use_direct_lookup:
            raise NotImplementedError("use_direct_lookup is not supported")
        if max_wind_direction_perturbation > 0:
            raise NotImplementedError("max_wind_direction_perturbation is not supported")

        L.lib()  # fail loudly if the HIP library is missing
        self.gpu_ctx = gpu_ctx
        if gpu_ctx is not None and hasattr(gpu_ctx, "device"):
            self.device = gpu_ctx.device
        else:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.gpu_stream = torch.cuda.Stream(device=self.device)

        self.boundary_conditions = Common.BoundaryConditions() if boundary_conditions is None \
            else boundary_conditions
        bc = self.boundary_conditions
        self.bc_data = Common.BoundaryConditionsData() if boundary_conditions_data is None \
            else boundary_conditions_data

        # f at the first interior cell (CDKLM16.py:147-149), then the scalar
        # casts of Simulator.__init__ (Simulator.py:82-97).
        angle = np.asarray(angle)
        f = host.compensate_f(f, coriolis_beta, angle, dx, dy)
        self.nx = np.int32(nx)
        self.ny = np.int32(ny)
        self.ghost_cells_x = np.int32(2)
        self.ghost_cells_y = np.int32(2)
        self.dx = np.float32(dx)
        self.dy = np.float32(dy)
        self.dt = dt
        self.g = np.float32(g)
        self.f = np.float32(f) if np.asarray(f).size == 1 else np.asarray(f, dtype=np.float32)
        self.r = np.float32(r)
        self.coriolis_beta = np.float32(coriolis_beta)
        self.y_zero_reference_cell = np.float32(0)
        self.theta = np.float32(theta)
        self.rk_order = np.int32(rk_order)
        self.t = t
        self.num_iterations = 0
        self.courant_number = courant_number
        self.interior_domain_indices = np.array([-2, -2, 2, 2])
        self.exact_arithmetic = bool(exact_arithmetic)
        self.rho_o = rho_o
        self.wind_stress_factor = wind_stress_factor
        self.one_dimensional = bool(one_dimensional)
        self.desingularization_eps = desingularization_eps
        self.flux_slope_eps = flux_slope_eps
        self.depth_cutoff = depth_cutoff
        self.flux_balancer = flux_balancer

        # Forcing containers (default: 1x1 zero textures).
        self.wind_stress = WindStress() if wind is None else wind
        if self.wind_stress.stress_u is None or self.wind_stress.stress_v is None:
            self.wind_stress.compute_wind_stress_from_wind()
        self.atmospheric_pressure = AtmosphericPressure() if atmospheric_pressure is None \
            else atmospheric_pressure

        eta0 = host.check_state_shape(eta0, nx, ny, "eta0")
        hu0 = host.check_state_shape(hu0, nx, ny, "hu0")
        hv0 = host.check_state_shape(hv0, nx, ny, "hv0")

        # Coriolis texture and the north vector.
        self.coriolis_texture = host.coriolis_texture(self.f, self.coriolis_beta, angle, nx, ny,
                                                      self.dx, self.dy, eta0.shape, subsample_f)
        self.angle = angle
        north_x, north_y, angle_zero = host.north_vector(angle)

        dev = self.device
        with torch.cuda.stream(self.gpu_stream):
            self.f_cells = None
            if self.coriolis_texture is not None:
                self._f_tex = torch.from_numpy(self.coriolis_texture).to(dev)
                self.f_cells = DeviceArray2D(ny + 4, nx + 4, dev)
                L.check(L.lib().gpo_coriolis_cells(int(nx), int(ny), L.ptr(self._f_tex),
                                                   int(self.coriolis_texture.shape[1]),
                                                   int(self.coriolis_texture.shape[0]),
                                                   ctypes.c_void_p(self.f_cells.data_ptr()),
                                                   self.f_cells.pitch, self._stream_handle()),
                        "gpo_coriolis_cells")

            # Bathymetry (utils/Common.py:1414-1491).
            Hi, self.use_mask = host.prepare_bathymetry(H, nx, ny)
            self.mask_value = host.LAND_VALUE
            self.has_land = bool(self.use_mask or np.any(np.abs(Hi - host.LAND_VALUE) < 1e-3))
            self.Hi = DeviceArray2D(ny + 5, nx + 5, dev, Hi)
            self.Bm = DeviceArray2D(ny + 4, nx + 4, dev)
            L.check(L.lib().gpo_bathymetry_init(int(nx), int(ny), bc.packed(), float(self.mask_value),
                                                ctypes.c_void_p(self.Hi.data_ptr()), self.Hi.pitch,
                                                ctypes.c_void_p(self.Bm.data_ptr()), self.Bm.pitch,
                                                self._stream_handle()),
                    "gpo_bathymetry_init")
        self.gpu_stream.synchronize()

        # Adjust eta for possible dry states (CDKLM16.py:320-321).
        Hm = self.Bm.download()
        eta0 = np.maximum(eta0, -Hm)

        with torch.cuda.stream(self.gpu_stream):
            self.h0 = [DeviceArray2D(ny + 4, nx + 4, dev, a) for a in (eta0, hu0, hv0)]
            self.h1 = [DeviceArray2D(ny + 4, nx + 4, dev, a) for a in (eta0, hu0, hv0)]
            self.h_stage = None
            if rk_order == 3:
                self.h_stage = [DeviceArray2D(ny + 4, nx + 4, dev) for _ in range(3)]
            self.dt_buffer = torch.zeros(1, dtype=torch.float32, device=dev)
        self.gpu_stream.synchronize()

        self.params = L.CDKLM16Params()
        p = self.params
        p.nx, p.ny = int(nx), int(ny)
        p.rk_order = int(rk_order)
        p.one_dimensional = int(self.one_dimensional)
        p.has_land = int(self.has_land)
        p.f_is_array = int(self.f_cells is not None)
        p.angle_is_zero = int(angle_zero)
        p.chunk_rows = int(chunk_rows) if chunk_rows else 0
        p.waves_per_simd = int(waves_per_simd) if waves_per_simd else 0
        p.dx = host.kernel_constant(self.dx)
        p.dy = host.kernel_constant(self.dy)
        p.g = host.kernel_constant(self.g)
        p.f = float(self.f) if self.f_cells is None else 0.0
        p.r = host.kernel_constant(self.r)
        p.theta = host.kernel_constant(self.theta)
        p.desing_eps = host.kernel_constant(desingularization_eps)
        p.flux_slope_eps = host.kernel_constant(flux_slope_eps)
        p.depth_cutoff = host.kernel_constant(depth_cutoff)
        p.flux_balancer = host.kernel_constant(flux_balancer)
        p.land_value = float(self.mask_value)
        p.north_x, p.north_y = float(north_x), float(north_y)
        p.wind_stress_factor = host.kernel_constant(wind_stress_factor)
        p.rho_o = host.kernel_constant(rho_o)
        p.divc_ok = 0
        L.check(L.lib().gpo_cdklm16_prepare(ctypes.byref(p), self._stream_handle()), "gpo_cdklm16_prepare")

        self._init_forcing()
        self._bc_times = Common.BCTimeline(self.bc_data)
        self._frs = None
        self.sponge = (ctypes.c_int32 * 4)(int(bc.spongeCells.get('north', 0)), int(bc.spongeCells.get('east', 0)),
                                           int(bc.spongeCells.get('south', 0)), int(bc.spongeCells.get('west', 0)))

        self.model_error = None
        self.hasDrifters = False
        self.drifters = None
        self.drifter_t = self.t
        self.model_time_step = self.dt if model_time_step is None else model_time_step
        self.total_time_steps = 0
        self.local_particle_id = local_particle_id
        self.constant_equilibrium_depth = np.max(H)
        self.comm = None
        self.ignore_ghostcells = ignore_ghostcells
        self.offset_x, self.offset_y = offset_x, offset_y

        # netCDF output, written before dt is computed as the reference does (CDKLM16.py:367-373)
        self.write_netcdf = write_netcdf
        self.sim_writer = None
        if self.write_netcdf:
            from .SimWriter import SimNetCDFWriter
            self.sim_writer = SimNetCDFWriter(self, super_dir_name=super_dir_name, filename=netcdf_filename,
                                              ignore_ghostcells=self.ignore_ghostcells, offset_x=self.offset_x,
                                              offset_y=self.offset_y)

        if self.dt <= 0:
            self.updateDt()

    # ------------------------------------------------------------------ utils
    def _stream_handle(self):
        return ctypes.c_void_p(self.gpu_stream.cuda_stream)

    def _init_forcing(self):
        ws, ap = self.wind_stress, self.atmospheric_pressure
        self._has_wind = not ws.is_zero()
        self._has_pressure = not ap.is_zero()
        self._wind_levels = TimeLevels(ws.t)
        self._pressure_levels = TimeLevels(ap.t)
        self._forcing = L.Forcing()
        self._forcing_bufs = {}

    def _tensor(self, key, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        t = torch.from_numpy(a).to(self.device, non_blocking=False)
        self._forcing_bufs[key] = t
        return t

    def _update_forcing(self):
        """Upload the forcing time levels around self.t and return the ctypes
        struct (or None when there is no forcing)."""
        if not (self._has_wind or self._has_pressure):
            return None
        fo = self._forcing
        with torch.cuda.stream(self.gpu_stream):
            if self._has_wind:
                ws = self.wind_stress
                t0, t1, c0, c1, w = self._wind_levels.update(self.t)
                if c0 or 'wxc' not in self._forcing_bufs:
                    fo.wind_x_cur = self._tensor('wxc', ws.stress_u[t0]).data_ptr()
                    fo.wind_y_cur = self._tensor('wyc', ws.stress_v[t0]).data_ptr()
                if c1 or 'wxn' not in self._forcing_bufs:
                    fo.wind_x_next = self._tensor('wxn', ws.stress_u[t1]).data_ptr()
                    fo.wind_y_next = self._tensor('wyn', ws.stress_v[t1]).data_ptr()
                fo.wind_x_ny, fo.wind_x_nx = ws.stress_u[0].shape
                fo.wind_y_ny, fo.wind_y_nx = ws.stress_v[0].shape
                fo.wind_t = np.float32(w)
            if self._has_pressure:
                ap = self.atmospheric_pressure
                t0, t1, c0, c1, w = self._pressure_levels.update(self.t)
                if c0 or 'pc' not in self._forcing_bufs:
                    fo.pressure_cur = self._tensor('pc', ap.P[t0]).data_ptr()
                if c1 or 'pn' not in self._forcing_bufs:
                    fo.pressure_next = self._tensor('pn', ap.P[t1]).data_ptr()
                fo.pressure_ny, fo.pressure_nx = ap.P[0].shape
                fo.pressure_t = np.float32(w)
        return fo

    def _update_bc_values(self):
        """FRS exterior solution around self.t (utils/Common.py:1193-1308)."""
        bc = self.boundary_conditions
        if not (bc.north == 3 or bc.south == 3 or bc.east == 3 or bc.west == 3):
            return
        c0, c1, i0, i1 = self._bc_times.update(self.t)
        if self._frs is None:
            self._frs = {}
            c0 = c1 = True
        with torch.cuda.stream(self.gpu_stream):
            if c0:
                self._frs['ns_cur'] = torch.from_numpy(Common.pack_bc_data_ns(self.bc_data, i0)).to(self.device)
                self._frs['ew_cur'] = torch.from_numpy(Common.pack_bc_data_ew(self.bc_data, i0)).to(self.device)
            if c1:
                self._frs['ns_next'] = torch.from_numpy(Common.pack_bc_data_ns(self.bc_data, i1)).to(self.device)
                self._frs['ew_next'] = torch.from_numpy(Common.pack_bc_data_ew(self.bc_data, i1)).to(self.device)

    def _apply_bc(self, h):
        """Write ghost cells of a state triple (utils/Common.py:1310-1330)."""
        bc = self.boundary_conditions
        if not (bc.north in (2, 3, 4) or bc.south in (3, 4) or bc.east in (2, 3, 4) or bc.west in (3, 4)):
            return  # walls only: the reference writes nothing
        f = self._frs or {}
        ns_nx = int(self.bc_data.north.h[0].shape[0])
        ew_ny = int(self.bc_data.east.h[0].shape[0])
        L.check(L.lib().gpo_boundary_conditions(
            int(self.nx), int(self.ny), self.boundary_conditions.packed(), self.sponge,
            ctypes.c_void_p(h[0].data_ptr()), ctypes.c_void_p(h[1].data_ptr()), ctypes.c_void_p(h[2].data_ptr()),
            h[0].pitch,
            L.ptr(f.get('ns_cur')), L.ptr(f.get('ns_next')), ns_nx,
            L.ptr(f.get('ew_cur')), L.ptr(f.get('ew_next')), ew_ny,
            float(self._bc_times.bc_t), self._stream_handle()), "gpo_boundary_conditions")

    def _needs_explicit_bc_each_substep(self):
        """Flow relaxation (3) and linear interpolation (4) ghost cells are
        written by their own kernels between the RK stages; walls and periodic
        sides are handled inside the step kernel."""
        bc = self.boundary_conditions
        return any(int(c) in (3, 4) for c in (bc.north, bc.south, bc.east, bc.west))

    def _substep(self, src, dst, qn, dt, rk_step, forcing):
        p = self.params
        L.check(L.lib().gpo_cdklm16_substep(
            ctypes.byref(p), float(dt), int(rk_step),
            ctypes.c_void_p(src[0].data_ptr()), ctypes.c_void_p(src[1].data_ptr()), ctypes.c_void_p(src[2].data_ptr()),
            None if qn is None else ctypes.c_void_p(qn[0].data_ptr()),
            None if qn is None else ctypes.c_void_p(qn[1].data_ptr()),
            None if qn is None else ctypes.c_void_p(qn[2].data_ptr()),
            ctypes.c_void_p(dst[0].data_ptr()), ctypes.c_void_p(dst[1].data_ptr()), ctypes.c_void_p(dst[2].data_ptr()),
            src[0].pitch,
            ctypes.c_void_p(self.Hi.data_ptr()), self.Hi.pitch,
            None if self.f_cells is None else ctypes.c_void_p(self.f_cells.data_ptr()),
            None if forcing is None else ctypes.byref(forcing),
            self.boundary_conditions.packed(), int(self.exact_arithmetic), self._stream_handle()),
            "gpo_cdklm16_substep")

    def _rk_step(self, local_dt, forcing):
        """One RK step h0 -> h0 (CDKLM16.py:517-562)."""
        explicit_bc = self._needs_explicit_bc_each_substep()
        if self.rk_order == 2:
            self._substep(self.h0, self.h1, None, local_dt, 0, forcing)
            if explicit_bc:
                self._apply_bc(self.h1)
            self._substep(self.h1, self.h0, self.h0, local_dt, 1, forcing)
        elif self.rk_order == 1:
            self._substep(self.h0, self.h1, None, local_dt, 0, forcing)
            self.h0, self.h1 = self.h1, self.h0
        else:
            self._substep(self.h0, self.h1, None, local_dt, 0, forcing)
            if explicit_bc:
                self._apply_bc(self.h1)
            with torch.cuda.stream(self.gpu_stream):
                for a, b in zip(self.h_stage, self.h1):
                    a.copy_from(b)  # ghost cells of the stage follow h1
            self._substep(self.h1, self.h_stage, self.h0, local_dt, 1, forcing)
            self.h1, self.h_stage = self.h_stage, self.h1
            if explicit_bc:
                self._apply_bc(self.h1)
            self._substep(self.h1, self.h0, self.h0, local_dt, 2, forcing)

    # ------------------------------------------------------------- public API
    def _batched_eligible(self, update_dt):
        bc = self.boundary_conditions
        return (self.rk_order == 2 and not update_dt and self.model_error is None and not self.hasDrifters
                and not (self._has_wind or self._has_pressure)
                and all(int(c) in (1, 2) for c in (bc.north, bc.east, bc.south, bc.west)))

    def _rk2_steps(self, n, dt):
        p = self.params
        h0, h1 = self.h0, self.h1
        L.check(L.lib().gpo_cdklm16_rk2_steps(
            ctypes.byref(p), int(n), float(dt),
            ctypes.c_void_p(h0[0].data_ptr()), ctypes.c_void_p(h0[1].data_ptr()), ctypes.c_void_p(h0[2].data_ptr()),
            ctypes.c_void_p(h1[0].data_ptr()), ctypes.c_void_p(h1[1].data_ptr()), ctypes.c_void_p(h1[2].data_ptr()),
            h0[0].pitch, ctypes.c_void_p(self.Hi.data_ptr()), self.Hi.pitch,
            None if self.f_cells is None else ctypes.c_void_p(self.f_cells.data_ptr()),
            self.boundary_conditions.packed(), int(self.exact_arithmetic), self._stream_handle()),
            "gpo_cdklm16_rk2_steps")

    def local_dts(self, t_end):
        """The time steps step(t_end) takes: local_dt = f32(min(dt, f32(t_end -
        t_now))) with t_now accumulated in float64 (CDKLM16.py:493-500)."""
        dts = []
        t_now = 0.0
        while t_now < t_end:
            local_dt = np.float32(min(self.dt, np.float32(t_end - t_now)))
            dts.append(local_dt)
            t_now += np.float64(local_dt)
        return dts

    def step(self, t_end=0.0, apply_stochastic_term=True, write_now=True, update_dt=False):
        """Advance by t_end seconds of model time (CDKLM16.py:477-584).

        The host keeps the reference's time bookkeeping (local_dt =
        f32(min(dt, f32(t_end - t_now))), t accumulated in float64).  Runs of
        equal local_dt with nothing but wall/periodic boundaries between steps
        are launched back to back by one library call."""
        if self.t == 0:
            self._update_bc_values()
            self._apply_bc(self.h0)
        if self._batched_eligible(update_dt):
            dts = self.local_dts(t_end)
            i = 0
            while i < len(dts):
                j = i
                while j < len(dts) and dts[j] == dts[i]:
                    j += 1
                self._rk2_steps(j - i, dts[i])
                i = j
            for local_dt in dts:
                self.t += np.float64(local_dt)
                self.num_iterations += 1
            self._apply_bc(self.h0)
            if self.write_netcdf and write_now:
                self.writeState()
            return self.t

        t_now = 0.0
        while t_now < t_end:
            if update_dt:
                self.updateDt()
            local_dt = np.float32(min(self.dt, np.float32(t_end - t_now)))
            self._one_step(local_dt, apply_stochastic_term)
            t_now += np.float64(local_dt)
        if not self._bc_each_step():
            self._apply_bc(self.h0)   # idempotent wall/periodic ghosts, written once at the end
        if self.write_netcdf and write_now:
            self.writeState()
        return self.t

    def step_dts(self, dts, apply_stochastic_term=True):
        """Steps of the given sizes through the per-step path (the tail of an
        ensemble step that the batched launch does not cover)."""
        for local_dt in dts:
            self._one_step(np.float32(local_dt), apply_stochastic_term)
        if not self._bc_each_step():
            self._apply_bc(self.h0)
        return self.t

    def _ghosts_read_each_step(self):
        return (self.model_error is not None) or self.hasDrifters

    def _bc_each_step(self):
        """Whether the boundary conditions are applied after every step, as
        the reference does (CDKLM16.py:571-573): needed when noise or
        drifters read the ghost cells and for flow relaxation, which is not
        idempotent.  Wall and periodic ghosts alone are written once at the
        end of step()."""
        return self._ghosts_read_each_step() or self._needs_explicit_bc_each_substep()

    def _one_step(self, local_dt, apply_stochastic_term):
        forcing = self._update_forcing()
        self._update_bc_values()
        self._rk_step(local_dt, forcing)
        if self.model_error is not None and apply_stochastic_term:
            self.perturbState()
        if self._bc_each_step():
            self._apply_bc(self.h0)
        self.drifterStep(local_dt)
        self.t += np.float64(local_dt)
        self.num_iterations += 1

    def _after_batched_steps(self, n, dt, ghosts_written=False):
        """Host bookkeeping after n batched steps of size dt: time, iteration
        count, the model error's draw index, the drifters' time and the ghost
        cells (unless the batched steps wrote them after every step)."""
        for _ in range(n):
            self.t += np.float64(dt)
            if self.drifters is not None:
                self.drifter_t += dt   # as drifterStep accumulates it
        self.num_iterations += n
        if self.model_error is not None:
            self.model_error.draw += n
            self.model_error._last_scales = (1.0, 1.0)   # the batched model error: q0_scale 1, perturbation_scale 1
        if not ghosts_written:
            self._apply_bc(self.h0)

    def dataAssimilationStep(self, observation_time, otherSim=None, model_error_final_step=True, write_now=True,
                             courant_number=0.8):
        """Steps to observation_time: the leftover step first, then full
        model_time_steps, each followed by model error (CDKLM16.py:742-806)."""
        full_model_time_steps = int(round(observation_time - self.t) / self.model_time_step)
        leftover_step_size = observation_time - self.t - full_model_time_steps * self.model_time_step
        if leftover_step_size / self.model_time_step < 0.1 and full_model_time_steps > 1:
            leftover_step_size += self.model_time_step
            full_model_time_steps -= 1
        if leftover_step_size / self.model_time_step < 0.00001:
            leftover_step_size = 0
        assert full_model_time_steps > 0, "There is less than CDKLM16.model_time_step until the observation"
        self.updateDt(courant_number=courant_number)
        if otherSim is not None:
            otherSim.updateDt(courant_number=courant_number)
        for i in range(full_model_time_steps + 1):
            if i == 0 and leftover_step_size == 0:
                continue
            elif i == 0:
                scale = np.sqrt(leftover_step_size / self.model_time_step)
                self.step(leftover_step_size, apply_stochastic_term=False, write_now=False)
                self.perturbState(perturbation_scale=scale)
                if otherSim is not None:
                    otherSim.step(leftover_step_size, apply_stochastic_term=False, write_now=False)
                    otherSim.perturbSimilarAs(self, perturbation_scale=scale)
            else:
                self.step(self.model_time_step, apply_stochastic_term=False, write_now=False)
                if (i < full_model_time_steps) or model_error_final_step:
                    self.perturbState()
                if otherSim is not None:
                    otherSim.step(self.model_time_step, apply_stochastic_term=False, write_now=False)
                    if (i < full_model_time_steps) or model_error_final_step:
                        otherSim.perturbSimilarAs(self)
            self.total_time_steps += 1
            if self.total_time_steps % 5 == 0:
                self.updateDt(courant_number=courant_number)
        if self.write_netcdf and write_now:
            self.sim_writer.writeTimestep(self)
        assert round(observation_time) == round(self.t), \
            "The simulation time is not the same as observation time after dataAssimilationStep! " \
            f"(self.t, observation_time, diff): {(self.t, observation_time, self.t - observation_time)}"

    def perturbSimilarAs(self, otherSim, perturbation_scale=1.0):
        """Perturbation from otherSim's current random field (CDKLM16.py:735-736)."""
        self.model_error.perturbSimSimilarAs(self, modelError=otherSim.model_error,
                                             perturbation_scale=perturbation_scale)

    def drifterStep(self, dt):
        if self.hasDrifters:
            self.drifters.driftFromSim(self, dt)
            self.drifter_t += dt
            return self.drifter_t

    def _adopt_stream(self, stream):
        """Run this member (and its model error and drifters) on `stream`
        from now on, after the work already queued on its own stream.  Members
        of a batched ensemble share the ensemble's stream, so the ensemble
        needs no per-member stream joins around its launches."""
        if stream is self.gpu_stream:
            return
        stream.wait_stream(self.gpu_stream)
        self.gpu_stream = stream
        if self.model_error is not None:
            self.model_error.gpu_stream = stream
        if self.drifters is not None and hasattr(self.drifters, "setGPUStream"):
            self.drifters.setGPUStream(stream)

    def attachDrifters(self, drifters):
        self.drifters = drifters
        self.hasDrifters = True
        if hasattr(self.drifters, "setGPUStream"):
            self.drifters.setGPUStream(self.gpu_stream)
        self.drifter_t = self.t

    def setSOARModelError(self, small_scale_perturbation_amplitude=None,
                          small_scale_perturbation_interpolation_factor=1,
                          use_lcg=False, xorwow_seed=None,
                          block_width_model_error=16, block_height_model_error=16):
        from .OceanStateNoise import OceanStateNoise
        self.model_error = OceanStateNoise.fromsim(self, soar_q0=small_scale_perturbation_amplitude,
                                                   interpolation_factor=small_scale_perturbation_interpolation_factor,
                                                   use_lcg=use_lcg, xorwow_seed=xorwow_seed)
        if self.write_netcdf:
            self.sim_writer.writeModelError(self)

    def setModelErrorFromFile(self, filename, use_lcg=False, xorwow_seed=None):
        """Model error as recorded in a netCDF file this simulator wrote
        (CDKLM16.py:600-633); SOAR only."""
        from .SimReader import SimNetCDFReader
        reader = SimNetCDFReader(filename, ignore_ghostcells=False)
        sim_name = str(reader.get('simulator_short'))
        assert sim_name == self.__class__.__name__, \
            "Trying to initialize a " + self.__class__.__name__ + " simulator with netCDF file based on " + \
            sim_name + " results."
        args = None
        if reader.has('small_scale_perturbation'):
            if reader.get('small_scale_perturbation') == 'True':
                args = {}
        elif reader.has('has_model_error') and reader.get('has_model_error') == "True":
            if reader.get('model_error_name') != "OceanStateNoise":
                raise NotImplementedError("only OceanStateNoise model error is part of gpuocean_amd")
            args = {}
        if args is not None:
            args['small_scale_perturbation_amplitude'] = reader.get('small_scale_perturbation_amplitude')
            args['small_scale_perturbation_interpolation_factor'] = \
                int(reader.get('small_scale_perturbation_interpolation_factor'))
            self.setSOARModelError(**args, use_lcg=use_lcg, xorwow_seed=xorwow_seed)
        reader.close()

    @classmethod
    def fromfilename(cls, gpu_ctx, filename, cont_write_netcdf=True, new_netcdf_filename=None, time0=None):
        """A simulator hot-started from the last (or the time0) record of a
        netCDF file, with the parameters recorded there (CDKLM16.py:403-473)."""
        from .SimReader import SimNetCDFReader
        reader = SimNetCDFReader(filename, ignore_ghostcells=False)
        sim_name = str(reader.get('simulator_short'))
        assert sim_name == cls.__name__, \
            "Trying to initialize a " + cls.__name__ + " simulator with netCDF file based on " + sim_name + " results."
        H = reader.getH()
        if time0 is None:
            eta0, hu0, hv0, time0 = reader.getLastTimeStep()
        else:
            eta0, hu0, hv0, time0 = reader.getStateAtTime(time0)
        if len(H.shape) == 3:
            H = H[0, :, :]
        params = dict(gpu_ctx=gpu_ctx, eta0=eta0, hu0=hu0, hv0=hv0, H=H, nx=reader.get("nx"), ny=reader.get("ny"),
                      dx=reader.get("dx"), dy=reader.get("dy"), dt=reader.get("dt"), g=reader.get("g"),
                      f=reader.get("coriolis_force"), r=reader.get("bottom_friction_r"), t=float(time0),
                      theta=reader.get("minmod_theta"), rk_order=reader.get("time_integrator"),
                      coriolis_beta=reader.get("coriolis_beta"), write_netcdf=cont_write_netcdf,
                      netcdf_filename=new_netcdf_filename, wind=WindStress(),
                      boundary_conditions=reader.getBC())
        if reader.has('model_time_step'):
            params['model_time_step'] = reader.get('model_time_step')
        reader.close()
        return cls(**params)

    def writeState(self):
        if self.write_netcdf:
            self.sim_writer.writeTimestep(self)

    def perturbState(self, perturbation_scale=1.0, update_random_field=True, q0_scale=1):
        if not q0_scale == 1:
            perturbation_scale = q0_scale
        if self.model_error is not None:
            self.model_error.perturbSim(self, perturbation_scale=perturbation_scale,
                                        update_random_field=update_random_field)

    def applyBoundaryConditions(self):
        self._apply_bc(self.h0)

    def updateDt(self, courant_number=None):
        """dt = courant * CFL maximum (CDKLM16.py:814-841); synchronises."""
        if courant_number is None:
            courant_number = self.courant_number
        with torch.cuda.stream(self.gpu_stream):
            L.check(L.lib().gpo_max_dt(int(self.nx), int(self.ny), float(self.dx), float(self.dy), float(self.g),
                                       float(self.mask_value),
                                       ctypes.c_void_p(self.h0[0].data_ptr()), ctypes.c_void_p(self.h0[1].data_ptr()),
                                       ctypes.c_void_p(self.h0[2].data_ptr()), self.h0[0].pitch,
                                       ctypes.c_void_p(self.Bm.data_ptr()), self.Bm.pitch,
                                       L.ptr(self.dt_buffer), self._stream_handle()), "gpo_max_dt")
            dt_host = self.dt_buffer.cpu().numpy()
        self.dt = courant_number * dt_host[0]

    def _getMaxTimestepHost(self, courant_number=0.8):
        """Host CFL formula kept for tests (CDKLM16.py:843-861)."""
        eta, hu, hv = self.download(interior_domain_only=True)
        Hm = self.downloadBathymetry()[1][2:-2, 2:-2]
        h = eta + Hm
        gravity_waves = np.sqrt(self.g * h)
        u = hu / h
        v = hv / h
        max_dt = 0.25 * min(self.dx / np.max(np.abs(u) + gravity_waves),
                            self.dy / np.max(np.abs(v) + gravity_waves))
        return courant_number * max_dt

    def mass(self):
        """Sum of the interior eta of the latest time step, accumulated in FP64
        on the device (the quantity ConservationOfMass_test.py sums on the host)."""
        with torch.cuda.stream(self.gpu_stream):
            out = torch.empty(1, dtype=torch.float64, device=self.device)
        L.check(L.lib().gpo_mass_sum(int(self.nx), int(self.ny), ctypes.c_void_p(self.h0[0].data_ptr()),
                                     self.h0[0].pitch, ctypes.c_void_p(out.data_ptr()), self._stream_handle()),
                "gpo_mass_sum")
        self.gpu_stream.synchronize()
        return float(out.item())

    def download(self, interior_domain_only=False):
        """(eta, hu, hv) of the latest time step (Simulator.py:348-361)."""
        self.gpu_stream.synchronize()
        eta, hu, hv = (a.download() for a in self.h0)
        if interior_domain_only:
            return eta[2:-2, 2:-2], hu[2:-2, 2:-2], hv[2:-2, 2:-2]
        return eta, hu, hv

    def downloadPrevTimestep(self):
        self.gpu_stream.synchronize()
        return tuple(a.download() for a in self.h1)

    def upload(self, eta0, hu0, hv0, eta1=None, hu1=None, hv1=None):
        """Replace the ocean state (Simulator.py:402-422)."""
        with torch.cuda.stream(self.gpu_stream):
            for dst, src in zip(self.h0, (eta0, hu0, hv0)):
                dst.upload(src)
            for dst, src in zip(self.h1, (eta0, hu0, hv0) if eta1 is None else (eta1, hu1, hv1)):
                dst.upload(src)
        self._update_bc_values()
        self._apply_bc(self.h0)

    def copyState(self, otherSim):
        assert type(otherSim) is type(self)
        assert (self.ny, self.nx) == (otherSim.ny, otherSim.nx)
        otherSim.gpu_stream.synchronize()
        with torch.cuda.stream(self.gpu_stream):
            for a, b in zip(self.h0 + self.h1, otherSim.h0 + otherSim.h1):
                a.copy_from(b)
        self.wind_stress = otherSim.wind_stress
        if otherSim.hasDrifters and self.hasDrifters:
            self.drifters.setDrifterPositions(otherSim.drifters.getDrifterPositions())
            self.drifters.setObservationPosition(otherSim.drifters.getObservationPosition())

    def downloadBathymetry(self, interior_domain_only=False):
        self.gpu_stream.synchronize()
        Bi, Bm = self.Hi.download(), self.Bm.download()
        if self.use_mask:
            Bi = np.ma.array(data=Bi, mask=(Bi == self.mask_value), fill_value=0.0)
            Bm = np.ma.array(data=Bm, mask=(Bm == self.mask_value), fill_value=0.0)
        if interior_domain_only:
            Bi = Bi[2:-2 + 1, 2:-2 + 1]
            Bm = Bm[2:-2, 2:-2]
        return [Bi, Bm]

    def getLandMask(self, interior_domain_only=True):
        if not self.use_mask:
            return None
        Bm = self.Bm.download()
        mask = Bm == self.mask_value
        return mask[2:-2, 2:-2] if interior_domain_only else mask

    def cleanUp(self, do_gc=True):
        self.closeNetCDF()
        self.h0 = self.h1 = self.h_stage = None
        self.Hi = self.Bm = self.f_cells = None
        if self.model_error is not None and hasattr(self.model_error, "cleanUp"):
            self.model_error.cleanUp()
        self.model_error = None
        self.gpu_ctx = None

    def closeNetCDF(self):
        """Closes the netCDF file, if there is one (Simulator.py:318-324)."""
        if self.write_netcdf:
            self.sim_writer.close()
            self.write_netcdf = False

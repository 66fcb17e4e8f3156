"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, and only as the checker or the timed CPU baseline.  It mirrors the
reference's CDKLM16 host loop (SWEsimulators/CDKLM16.py:48-584) on top of the
C++ restatement in cdklm_oracle.cpp, in float32 or float64.
"""

import ctypes
import os
import subprocess

import numpy as np

from gpuocean_amd import Common, host

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("rk_order", ctypes.c_int32),
        ("bc_north", ctypes.c_int32), ("bc_east", ctypes.c_int32), ("bc_south", ctypes.c_int32),
        ("bc_west", ctypes.c_int32),
        ("sponge_north", ctypes.c_int32), ("sponge_east", ctypes.c_int32), ("sponge_south", ctypes.c_int32),
        ("sponge_west", ctypes.c_int32),
        ("one_dimensional", ctypes.c_int32), ("f_nx", ctypes.c_int32), ("f_ny", ctypes.c_int32),
        ("f_arr", ctypes.POINTER(ctypes.c_float)),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("g", ctypes.c_double), ("f", ctypes.c_double),
        ("r", ctypes.c_double), ("theta", ctypes.c_double),
        ("desing_eps", ctypes.c_double), ("flux_slope_eps", ctypes.c_double), ("depth_cutoff", ctypes.c_double),
        ("flux_balancer", ctypes.c_double), ("land_value", ctypes.c_double), ("north_x", ctypes.c_double),
        ("north_y", ctypes.c_double), ("wind_stress_factor", ctypes.c_double), ("rho_o", ctypes.c_double),
    ]


class OracleForcing(ctypes.Structure):
    _fields_ = [
        ("wx_cur", ctypes.c_void_p), ("wx_next", ctypes.c_void_p), ("wy_cur", ctypes.c_void_p),
        ("wy_next", ctypes.c_void_p),
        ("wx_nx", ctypes.c_int32), ("wx_ny", ctypes.c_int32), ("wy_nx", ctypes.c_int32), ("wy_ny", ctypes.c_int32),
        ("wind_t", ctypes.c_float),
        ("p_cur", ctypes.c_void_p), ("p_next", ctypes.c_void_p), ("p_nx", ctypes.c_int32), ("p_ny", ctypes.c_int32),
        ("p_t", ctypes.c_float),
    ]


class OracleFRS(ctypes.Structure):
    _fields_ = [
        ("ns_cur", ctypes.c_void_p), ("ns_next", ctypes.c_void_p), ("ew_cur", ctypes.c_void_p),
        ("ew_next", ctypes.c_void_p), ("ns_nx", ctypes.c_int32), ("ew_ny", ctypes.c_int32), ("t", ctypes.c_float),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(OracleParams)
        vp = ctypes.c_void_p
        for suf in ("f32", "f64"):
            getattr(_lib, f"gpo_oracle_init_bathymetry_{suf}").argtypes = [P, vp, vp]
            getattr(_lib, f"gpo_oracle_run_{suf}").argtypes = [P, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_double,
                                                               ctypes.c_int, ctypes.c_int,
                                                               ctypes.POINTER(OracleForcing),
                                                               ctypes.POINTER(OracleFRS)]
            getattr(_lib, f"gpo_oracle_apply_bc_{suf}").argtypes = [P, vp, vp, vp, ctypes.POINTER(OracleFRS)]
            getattr(_lib, f"gpo_oracle_max_dt_{suf}").argtypes = [P, vp, vp, vp, vp]
        _lib.gpo_oracle_max_dt_f32.restype = ctypes.c_float
        _lib.gpo_oracle_max_dt_f64.restype = ctypes.c_double
        _lib.gpo_oracle_fcell_f32.argtypes = [P, vp]
        _lib.gpo_oracle_substep_f32.argtypes = [P, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_float, ctypes.c_int]
        _lib.gpo_oracle_num_threads.restype = ctypes.c_int
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class OracleCDKLM16:
    """CPU twin of gpuocean_amd.CDKLM16 (same constructor subset, step and
    download), running the oracle restatement in float32 or float64."""

    def __init__(self, eta0, hu0, hv0, H, nx, ny, dx, dy, dt, g, f, r,
                 subsample_f=10, angle=np.array([[0]], dtype=np.float32), rho_o=1025.0, t=0.0, theta=1.8,
                 rk_order=2, coriolis_beta=0.0, wind=None, wind_stress_factor=1.0, atmospheric_pressure=None,
                 boundary_conditions=None, boundary_conditions_data=None, courant_number=0.8,
                 flux_slope_eps=1.0e-1, desingularization_eps=1.0e-1, depth_cutoff=1.0e-5,
                 one_dimensional=False, flux_balancer=0.8, precision="f32", **unused):
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.suf = precision
        self.boundary_conditions = Common.BoundaryConditions() if boundary_conditions is None \
            else boundary_conditions
        self.bc_data = Common.BoundaryConditionsData() if boundary_conditions_data is None \
            else boundary_conditions_data
        bc = self.boundary_conditions
        angle = np.asarray(angle)
        f = host.compensate_f(f, coriolis_beta, angle, dx, dy)
        self.nx, self.ny = int(nx), int(ny)
        self.dx, self.dy = np.float32(dx), np.float32(dy)
        self.g = np.float32(g)
        self.f = np.float32(f) if np.asarray(f).size == 1 else np.asarray(f, dtype=np.float32)
        self.r = np.float32(r)
        self.coriolis_beta = np.float32(coriolis_beta)
        self.dt = dt
        self.t = t
        self.courant_number = courant_number
        self.rk_order = int(rk_order)
        eta0 = host.check_state_shape(eta0, nx, ny, "eta0")
        hu0 = host.check_state_shape(hu0, nx, ny, "hu0")
        hv0 = host.check_state_shape(hv0, nx, ny, "hv0")
        tex = host.coriolis_texture(self.f, self.coriolis_beta, angle, nx, ny, self.dx, self.dy, eta0.shape,
                                    subsample_f)
        self._tex = tex
        north_x, north_y, _ = host.north_vector(angle)

        q = OracleParams()
        q.nx, q.ny, q.rk_order = self.nx, self.ny, self.rk_order
        q.bc_north, q.bc_east, q.bc_south, q.bc_west = int(bc.north), int(bc.east), int(bc.south), int(bc.west)
        q.sponge_north = int(bc.spongeCells.get('north', 0))
        q.sponge_east = int(bc.spongeCells.get('east', 0))
        q.sponge_south = int(bc.spongeCells.get('south', 0))
        q.sponge_west = int(bc.spongeCells.get('west', 0))
        q.one_dimensional = int(bool(one_dimensional))
        if tex is not None:
            q.f_nx, q.f_ny = tex.shape[1], tex.shape[0]
            q.f_arr = tex.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            q.f = 0.0
        else:
            q.f = float(self.f)
        # float32 constants as the kernel sees them; the f64 build widens them.
        kc = host.kernel_constant if precision == "f32" else (lambda x: float(x))
        q.dx, q.dy, q.g = float(kc(self.dx)), float(kc(self.dy)), float(kc(self.g))
        q.r, q.theta = float(kc(self.r)), float(kc(np.float32(theta)))
        q.desing_eps = float(kc(desingularization_eps))
        q.flux_slope_eps = float(kc(flux_slope_eps))
        q.depth_cutoff = float(kc(depth_cutoff))
        q.flux_balancer = float(kc(flux_balancer))
        q.land_value = float(host.LAND_VALUE)
        q.north_x, q.north_y = float(north_x), float(north_y)
        q.wind_stress_factor = float(kc(wind_stress_factor))
        q.rho_o = float(kc(rho_o))
        self.q = q

        Hi, self.use_mask = host.prepare_bathymetry(H, nx, ny)
        self.Hi = np.ascontiguousarray(Hi, dtype=self.dtype)
        self.Bm = np.zeros((ny + 4, nx + 4), dtype=self.dtype)
        getattr(lib(), f"gpo_oracle_init_bathymetry_{self.suf}")(ctypes.byref(q), _p(self.Hi), _p(self.Bm))
        Hm32 = self.Bm.astype(np.float32)
        eta0 = np.maximum(eta0, -Hm32)
        self.h0 = [np.array(a, dtype=self.dtype) for a in (eta0, hu0, hv0)]
        self.h1 = [np.array(a, dtype=self.dtype) for a in (eta0, hu0, hv0)]

        self.wind = wind
        self.pressure = atmospheric_pressure
        self._bc_times = Common.BCTimeline(self.bc_data)
        self._frs = None
        if self.dt <= 0:
            self.updateDt()

    def _frs_struct(self):
        bc = self.boundary_conditions
        if not (bc.north == 3 or bc.south == 3 or bc.east == 3 or bc.west == 3):
            return None
        c0, c1, i0, i1 = self._bc_times.update(self.t)
        if self._frs is None:
            self._frs = {}
        self._frs['ns_cur'] = Common.pack_bc_data_ns(self.bc_data, i0)
        self._frs['ew_cur'] = Common.pack_bc_data_ew(self.bc_data, i0)
        self._frs['ns_next'] = Common.pack_bc_data_ns(self.bc_data, i1)
        self._frs['ew_next'] = Common.pack_bc_data_ew(self.bc_data, i1)
        s = OracleFRS()
        s.ns_cur, s.ns_next = self._frs['ns_cur'].ctypes.data, self._frs['ns_next'].ctypes.data
        s.ew_cur, s.ew_next = self._frs['ew_cur'].ctypes.data, self._frs['ew_next'].ctypes.data
        s.ns_nx = int(self.bc_data.north.h[0].shape[0])
        s.ew_ny = int(self.bc_data.east.h[0].shape[0])
        s.t = float(self._bc_times.bc_t)
        return s

    def _forcing_struct(self):
        if self.wind is None and self.pressure is None:
            return None
        fo = OracleForcing()
        keep = []
        if self.wind is not None:
            ws = self.wind
            if ws.stress_u is None:
                ws.compute_wind_stress_from_wind()
            tt = ws.t
            t0 = max(0, int(np.searchsorted(tt, self.t)) - 1)
            t1 = min(len(tt) - 1, int(np.searchsorted(tt, self.t)))
            arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in
                    (ws.stress_u[t0], ws.stress_u[t1], ws.stress_v[t0], ws.stress_v[t1])]
            keep += arrs
            fo.wx_cur, fo.wx_next, fo.wy_cur, fo.wy_next = [a.ctypes.data for a in arrs]
            fo.wx_ny, fo.wx_nx = arrs[0].shape
            fo.wy_ny, fo.wy_nx = arrs[2].shape
            interval = max(1.0e-10, tt[t1] - tt[t0])
            fo.wind_t = np.float32(max(0.0, min(1.0, (self.t - tt[t0]) / interval)))
        if self.pressure is not None:
            ap = self.pressure
            tt = ap.t
            t0 = max(0, int(np.searchsorted(tt, self.t)) - 1)
            t1 = min(len(tt) - 1, int(np.searchsorted(tt, self.t)))
            arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (ap.P[t0], ap.P[t1])]
            keep += arrs
            fo.p_cur, fo.p_next = [a.ctypes.data for a in arrs]
            fo.p_ny, fo.p_nx = arrs[0].shape
            interval = max(1.0e-10, tt[t1] - tt[t0])
            fo.p_t = np.float32(max(0.0, min(1.0, (self.t - tt[t0]) / interval)))
        self._keep = keep
        return fo

    def _run(self, dt, n, initial_bc):
        fo = self._forcing_struct()
        frs = self._frs_struct()
        getattr(lib(), f"gpo_oracle_run_{self.suf}")(
            ctypes.byref(self.q), *[_p(a) for a in self.h0], *[_p(a) for a in self.h1], _p(self.Hi), _p(self.Bm),
            float(dt), int(n), int(initial_bc), None if fo is None else ctypes.byref(fo),
            None if frs is None else ctypes.byref(frs))

    def step(self, t_end=0.0):
        """Same time-step bookkeeping as CDKLM16.step (CDKLM16.py:477-584)."""
        initial = self.t == 0
        t_now = 0.0
        while t_now < t_end:
            local_dt = np.float32(min(self.dt, np.float32(t_end - t_now)))
            self._run(local_dt, 1, initial)
            initial = False
            self.t += np.float64(local_dt)
            t_now += np.float64(local_dt)
        if initial:
            frs = self._frs_struct()
            getattr(lib(), f"gpo_oracle_apply_bc_{self.suf}")(ctypes.byref(self.q), *[_p(a) for a in self.h0],
                                                             None if frs is None else ctypes.byref(frs))
        return self.t

    def apply_bc(self):
        """The ghost-cell pass on the current state (the FRS data at the
        current time)."""
        frs = self._frs_struct()
        getattr(lib(), f"gpo_oracle_apply_bc_{self.suf}")(ctypes.byref(self.q), *[_p(a) for a in self.h0],
                                                         None if frs is None else ctypes.byref(frs))

    def run_steps(self, n, dt):
        """n steps of a fixed dt in one call (no per-step Python); forcing and
        FRS time levels are held at the current time."""
        self._run(np.float32(dt), n, self.t == 0)
        self.t += n * np.float64(np.float32(dt))

    def rk2_step(self, dt, perturb=None):
        """One RK2 step in the order CDKLM16._one_step takes it (FP32 only):
        substep 0, ghost cells of the stage, substep 1, then perturb(h0) (the
        model error, when given), then the ghost cells of the new state.  The
        ghost cells of h0 are written first when t = 0, as CDKLM16.step does."""
        assert self.suf == "f32"
        L = lib()
        frs = self._frs_struct()
        fr = None if frs is None else ctypes.byref(frs)
        q = ctypes.byref(self.q)
        if self.t == 0:
            L.gpo_oracle_apply_bc_f32(q, *[_p(a) for a in self.h0], fr)
        sub = L.gpo_oracle_substep_f32
        common = (_p(self.Hi), _p(self.Bm), ctypes.c_float(dt))
        sub(q, *[_p(a) for a in self.h0], *[_p(a) for a in self.h1], *common, ctypes.c_int(0))
        L.gpo_oracle_apply_bc_f32(q, *[_p(a) for a in self.h1], fr)
        sub(q, *[_p(a) for a in self.h1], *[_p(a) for a in self.h0], *common, ctypes.c_int(1))
        if perturb is not None:
            perturb(self.h0)
        L.gpo_oracle_apply_bc_f32(q, *[_p(a) for a in self.h0], fr)
        self.t += np.float64(np.float32(dt))

    def download(self, interior_domain_only=False):
        out = [a.copy() for a in self.h0]
        if interior_domain_only:
            out = [a[2:-2, 2:-2] for a in out]
        return tuple(out)

    def max_dt(self):
        return getattr(lib(), f"gpo_oracle_max_dt_{self.suf}")(ctypes.byref(self.q), *[_p(a) for a in self.h0],
                                                               _p(self.Bm))

    def updateDt(self, courant_number=None):
        if courant_number is None:
            courant_number = self.courant_number
        self.dt = courant_number * np.float32(self.max_dt())


def num_threads():
    return lib().gpo_oracle_num_threads()


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def soar(nx, ny, W, scale, per_ns, per_ew, xi, coarse=None, additive=False):
    """SOAR field on the (ny+4) x (nx+4) coarse grid from xi (noise_drift_oracle.cpp)."""
    L = lib()
    xi = np.ascontiguousarray(xi, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float32)
    out = np.zeros((ny + 4, nx + 4), np.float32) if coarse is None else np.ascontiguousarray(coarse, np.float32)
    L.gpo_oracle_soar(ctypes.c_int(nx), ctypes.c_int(ny), _fp(W), ctypes.c_float(scale), ctypes.c_int(int(per_ns)),
                      ctypes.c_int(int(per_ew)), _fp(xi), ctypes.c_int(xi.shape[1]), _fp(out),
                      ctypes.c_int(int(additive)))
    return out


def geostrophic_balance(nx, ny, dx, dy, g, f, beta, north_x, north_y, f_cells, coarse, eta, hu, hv, Hi,
                        land_value=np.float32(1e20)):
    """Adds the balanced perturbation to eta/hu/hv in place (float32 arrays)."""
    L = lib()
    args = [ctypes.c_int(nx), ctypes.c_int(ny)] + [ctypes.c_float(v) for v in (dx, dy, g, f, beta, north_x, north_y)]
    fc = None if f_cells is None else _fp(np.ascontiguousarray(f_cells, np.float32))
    L.gpo_oracle_geostrophic_balance(*args, fc, _fp(np.ascontiguousarray(coarse, np.float32)), _fp(eta), _fp(hu),
                                     _fp(hv), _fp(np.ascontiguousarray(Hi, np.float32)), ctypes.c_float(land_value))


def bicubic_interpolation(nx, ny, dx, dy, cnx, cny, cdx, cdy, oi, oj, g, f, beta, north_x, north_y, f_cells, coarse,
                          eta, hu, hv, Hi, land_value=np.float32(1e20)):
    """Adds the coarse perturbation, bicubically interpolated, and its balanced
    momentum to eta/hu/hv in place (float32 (ny+4) x (nx+4) arrays)."""
    L = lib()
    args = [ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_float(dx), ctypes.c_float(dy), ctypes.c_int(cnx),
            ctypes.c_int(cny), ctypes.c_float(cdx), ctypes.c_float(cdy), ctypes.c_int(oi), ctypes.c_int(oj)]
    args += [ctypes.c_float(v) for v in (g, f, beta, north_x, north_y)]
    fc = None if f_cells is None else _fp(np.ascontiguousarray(f_cells, np.float32))
    L.gpo_oracle_bicubic_interpolation(*args, fc, _fp(np.ascontiguousarray(coarse, np.float32)), _fp(eta), _fp(hu),
                                       _fp(hv), _fp(np.ascontiguousarray(Hi, np.float32)), ctypes.c_float(land_value))


def drift(nx, ny, dx, dy, dt, eta, hu, hv, Hm, per_ns, per_ew, pos, sensitivity=1.0, wind=(0.0, 0.0), factor=0.0):
    """One drifter step in place on pos ((n+1) x 2 float32)."""
    L = lib()
    n = pos.shape[0] - 1
    c = [np.ascontiguousarray(a, np.float32) for a in (eta, hu, hv, Hm)]
    L.gpo_oracle_drift(ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_float(dx), ctypes.c_float(dy),
                       ctypes.c_float(dt), ctypes.c_int(2), ctypes.c_int(2), *[_fp(a) for a in c],
                       ctypes.c_int(int(per_ns)), ctypes.c_int(int(per_ew)), ctypes.c_int(n), _fp(pos),
                       ctypes.c_float(sensitivity), ctypes.c_float(wind[0]), ctypes.c_float(wind[1]),
                       ctypes.c_float(factor))


def philox4x32_10(ctr, key):
    """Reference Philox4x32-10 in Python integers (Salmon et al., SC'11)."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = list(ctr)
    k0, k1 = key
    for r in range(10):
        if r:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF, ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF,
             p0 & 0xFFFFFFFF]
    return c

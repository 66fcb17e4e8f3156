"""CPU checks of the differentiable substep stages (the host build of the
per-cell code the gfx950 kernels run): FP64 forward against the FP64 oracle,
adjoint against the tangent-linear model (dot-product test) and against
central finite differences, on small grids and every boundary combination."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import cases
from gpuocean_amd import Common, host
from gpuocean_amd import _lib as GL
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
LIB = os.path.join(NATIVE, "libadjoint_host.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", NATIVE], check=True)
        _lib = ctypes.CDLL(LIB)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Model:
    """A small CDKLM16 configuration evaluated with the host build."""

    def __init__(self, nx=24, ny=20, bc=Common.BoundaryConditions(1, 1, 1, 1), seed=4, f=1.2e-4, island=False,
                 shelf=False):
        kw = cases.config_c3(n=max(nx, ny), seed=seed)
        n = max(nx, ny)
        kw = dict(kw, nx=nx, ny=ny, f=f, boundary_conditions=bc)
        kw["H"] = kw["H"][:ny + 5, :nx + 5].copy()
        if island:   # land: corner depths of a rectangle in the middle of the domain
            mask = np.zeros(kw["H"].shape, bool)
            mask[ny // 2 - 2:ny // 2 + 3, nx // 2 - 3:nx // 2 + 2] = True
            kw["H"] = np.ma.array(kw["H"], mask=mask)
        if shelf:    # a band of water shallower than the desingularisation depth (0.1 m)
            kw["H"][:, 3:7] = 0.06
        for k in ("eta0", "hu0", "hv0"):
            kw[k] = kw[k][:ny + 4, :nx + 4].copy()
        bg, _ = cases.c3_background_and_truth(dict(kw, eta0=kw["eta0"]), seed=seed)
        self.bg = [np.asarray(a, np.float64) for a in bg]
        if shelf:    # at rest with tiny momenta over the shelf, so that 0.06 + eta stays in (cutoff, 0.1)
            self.bg[0][:, 2:9] = 0.0
            self.bg[1][:, 2:9] *= 1e-3
            self.bg[2][:, 2:9] *= 1e-3
        ora = O.OracleCDKLM16(**kw, flux_balancer=float(np.float32(0.8)), precision="f64")
        self.kw, self.ora = kw, ora
        self.Hi = np.ascontiguousarray(ora.Hi, np.float64)
        self.dt = float(np.float32(ora.dt))
        self.bc = bc.packed()
        q = GL.CDKLM16Params()
        q.nx, q.ny, q.rk_order = nx, ny, 2
        q.one_dimensional = q.f_is_array = 0
        q.has_land = int(island)
        q.angle_is_zero = 1
        q.dx, q.dy = host.kernel_constant(np.float32(200.0)), host.kernel_constant(np.float32(200.0))
        q.g = host.kernel_constant(np.float32(9.81))
        q.f = float(np.float32(host.compensate_f(f, 0.0, np.array([[0]]), 200.0, 200.0)))
        q.r = 0.0
        q.theta = host.kernel_constant(np.float32(1.8))
        q.desing_eps = host.kernel_constant(0.1)
        q.flux_slope_eps = host.kernel_constant(0.1)
        q.depth_cutoff = host.kernel_constant(1e-5)
        q.flux_balancer = host.kernel_constant(0.8)
        q.land_value = float(host.LAND_VALUE)
        q.north_x, q.north_y = 0.0, 1.0
        q.wind_stress_factor, q.rho_o = 1.0, 1025.0
        self.q = q
        self.flag = ctypes.c_int(0)
        self.shape = (ny + 4, nx + 4)

    def zeros(self):
        return [np.zeros(self.shape) for _ in range(3)]

    def substep(self, rk, inp, qn=None):
        out = self.zeros()
        qn = qn if qn is not None else [None] * 3
        rc = lib().gpo_host_substep_f64(ctypes.byref(self.q), ctypes.c_double(self.dt), rk, *[_p(a) for a in inp],
                                        *[_p(a) for a in qn], *[_p(a) for a in out], _p(self.Hi), None,
                                        ctypes.c_uint32(self.bc), ctypes.byref(self.flag))
        assert rc == 0
        return out

    def trajectory(self, q0, n):
        Qs, Q1s = [q0], []
        for _ in range(n):
            q1 = self.substep(0, Qs[-1])
            Q1s.append(q1)
            Qs.append(self.substep(1, q1, Qs[-1]))
        return Qs, Q1s

    def tlm_substep(self, rk, inp, dinp, qn=None, dqn=None):
        out, dout = self.zeros(), self.zeros()
        qn = qn if qn is not None else [None] * 3
        dqn = dqn if dqn is not None else [None] * 3
        rc = lib().gpo_host_substep_tlm_f64(ctypes.byref(self.q), ctypes.c_double(self.dt), rk,
                                            *[_p(a) for a in inp], *[_p(a) for a in dinp], *[_p(a) for a in qn],
                                            *[_p(a) for a in dqn], *[_p(a) for a in out], *[_p(a) for a in dout],
                                            _p(self.Hi), None, ctypes.c_uint32(self.bc), ctypes.byref(self.flag))
        assert rc == 0
        return out, dout

    def adj_substep(self, rk, inp, out, lam_out, lam_in, lam_qn, accumulate):
        rc = lib().gpo_host_substep_adj_f64(ctypes.byref(self.q), ctypes.c_double(self.dt), rk,
                                            *[_p(a) for a in inp], _p(out[0]), *[_p(a) for a in lam_out],
                                            *[_p(a) for a in lam_in],
                                            *[_p(a) for a in (lam_qn if lam_qn is not None else [None] * 3)],
                                            accumulate, _p(self.Hi), None, ctypes.c_uint32(self.bc),
                                            ctypes.byref(self.flag))
        assert rc == 0

    def tangent(self, Qs, Q1s, d0):
        d = d0
        for j in range(len(Q1s)):
            _, d1 = self.tlm_substep(0, Qs[j], d)
            _, d = self.tlm_substep(1, Q1s[j], d1, Qs[j], d)
        return d

    def adjoint(self, Qs, Q1s, lam_final):
        lam = [a.copy() for a in lam_final]
        for j in reversed(range(len(Q1s))):
            L1, X = self.zeros(), self.zeros()
            self.adj_substep(1, Q1s[j], Qs[j + 1], lam, L1, X, 0)
            self.adj_substep(0, Qs[j], Q1s[j], L1, X, None, 1)
            lam = X
        return lam


BCS = [Common.BoundaryConditions(1, 1, 1, 1), Common.BoundaryConditions(1, 2, 1, 2),
       Common.BoundaryConditions(2, 1, 2, 1), Common.BoundaryConditions(2, 2, 2, 2)]
IDS = ["walls", "periodic_x", "periodic_y", "periodic"]


def _interior_noise(shape, rng, scale):
    a = np.zeros(shape)
    a[2:-2, 2:-2] = scale * rng.standard_normal((shape[0] - 4, shape[1] - 4))
    return a


def _dot(a, b):
    return sum(float(np.sum(x[2:-2, 2:-2] * y[2:-2, 2:-2])) for x, y in zip(a, b))


@pytest.mark.parametrize("bc", BCS, ids=IDS)
@pytest.mark.parametrize("island", [False, True], ids=["", "island"])
def test_host_forward_matches_f64_oracle(bc, island):
    m = Model(bc=bc, island=island)
    m.ora.h0 = [a.copy() for a in m.bg]
    m.ora.h1 = [a.copy() for a in m.bg]
    m.ora.run_steps(6, m.dt)
    Qs, _ = m.trajectory([a.copy() for a in m.bg], 6)
    for a, b in zip(Qs[-1], m.ora.download()):
        assert np.max(np.abs(a[2:-2, 2:-2] - b[2:-2, 2:-2])) <= 1e-12 * max(1.0, np.max(np.abs(b)))
    assert m.flag.value == 0


@pytest.mark.parametrize("bc", BCS, ids=IDS)
@pytest.mark.parametrize("island,shelf", [(False, False), (True, False), (False, True)], ids=["", "island", "shelf"])
def test_host_dot_product(bc, island, shelf):
    m = Model(bc=bc, island=island, shelf=shelf)
    Qs, Q1s = m.trajectory([a.copy() for a in m.bg], 4)
    rng = np.random.default_rng(11)
    v = [_interior_noise(m.shape, rng, s) for s in (1e-3, 1e-1, 1e-1)]
    w = [_interior_noise(m.shape, rng, 1.0) for _ in range(3)]
    lhs = _dot(m.tangent(Qs, Q1s, v), w)
    rhs = _dot(v, m.adjoint(Qs, Q1s, w))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs)), (lhs, rhs)


@pytest.mark.parametrize("bc", BCS, ids=IDS)
@pytest.mark.parametrize("island", [False, True], ids=["", "island"])
def test_host_single_substep_transposes(bc, island):
    """Each substep alone, both stages, against the tangent-linear model."""
    m = Model(bc=bc, nx=12, ny=10, island=island)
    Qs, Q1s = m.trajectory([a.copy() for a in m.bg], 1)
    rng = np.random.default_rng(5)
    v = [_interior_noise(m.shape, rng, s) for s in (1e-3, 1e-1, 1e-1)]
    vq = [_interior_noise(m.shape, rng, s) for s in (1e-3, 1e-1, 1e-1)]
    w = [_interior_noise(m.shape, rng, 1.0) for _ in range(3)]
    _, tv0 = m.tlm_substep(0, Qs[0], v)
    li = m.zeros()
    m.adj_substep(0, Qs[0], Q1s[0], w, li, None, 0)
    assert abs(_dot(tv0, w) - _dot(v, li)) <= 1e-11 * abs(_dot(tv0, w))
    _, tv1 = m.tlm_substep(1, Q1s[0], v, Qs[0], vq)
    li, lq = m.zeros(), m.zeros()
    m.adj_substep(1, Q1s[0], Qs[1], w, li, lq, 0)
    lhs, rhs = _dot(tv1, w), _dot(v, li) + _dot(vq, lq)
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("bc", BCS[:2], ids=IDS[:2])
@pytest.mark.parametrize("island,shelf", [(False, False), (True, False), (False, True)], ids=["", "island", "shelf"])
def test_host_adjoint_against_finite_differences(bc, island, shelf):
    """J = <w, Q^n>: the adjoint's <lambda_0, d> against central differences."""
    m = Model(bc=bc, island=island, shelf=shelf)
    n = 5
    rng = np.random.default_rng(3)
    w = [_interior_noise(m.shape, rng, 1.0) for _ in range(3)]
    Qs, Q1s = m.trajectory([a.copy() for a in m.bg], n)
    lam0 = m.adjoint(Qs, Q1s, w)
    # A coast or a shelf adds limiter kinks (adjustSlopes next to near-zero
    # face depths) that a white-noise step can cross; there the best of
    # several step sizes is compared (the function is smooth between kinks).
    steps, tol = ((1e-6, 3e-7, 1e-7, 3e-8), 1e-5) if (island or shelf) else ((1e-6,), 1e-6)
    for _ in range(3):
        d = [_interior_noise(m.shape, rng, s) for s in (1e-2, 1.0, 1.0)]
        if shelf:   # perturbations in proportion to the shelf's tiny depth and momenta
            for a in d:
                a[:, 2:9] *= 1e-3
        ad = _dot(lam0, d)
        errs = []
        for eps in steps:
            Jp = _dot(w, m.trajectory([a + eps * b for a, b in zip(m.bg, d)], n)[0][-1])
            Jm = _dot(w, m.trajectory([a - eps * b for a, b in zip(m.bg, d)], n)[0][-1])
            fd = (Jp - Jm) / (2 * eps)
            errs.append(abs(ad - fd) / abs(fd))
        assert min(errs) <= tol, (ad, errs)

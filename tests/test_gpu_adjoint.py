"""Differentiable CDKLM16: the FP64 forward against the FP64 oracle, the
hand-written adjoint against the tangent-linear model (dot-product test in
FP64) and against central finite differences of the FP64 forward, and the
torch.autograd.Function wrapper.  The adjoint has no counterpart in the
reference; these checks are what pins it."""
import numpy as np
import pytest
import torch

import cases
from gpuocean_amd import CDKLM16, Common
from gpuocean_amd.adjoint import CDKLM16Misfit, DifferentiableCDKLM16
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _frs_data(kw):
    """Flow-relaxation data of one time: the initial state's outermost interior rows and columns."""
    e, u, v = kw["eta0"], kw["hu0"], kw["hv0"]
    ny, nx = e.shape[0] - 4, e.shape[1] - 4

    def side(sl):
        return Common.SingleBoundaryConditionData(h=[e[sl].copy()], hu=[u[sl].copy()], hv=[v[sl].copy()])

    return Common.BoundaryConditionsData(north=side(np.s_[ny + 1, :]), south=side(np.s_[2, :]),
                                         east=side(np.s_[:, nx + 1]), west=side(np.s_[:, 2]))


def _model(gpu_ctx, n=64, bc=None, n_steps=20, k=7, n_obs=30, seed=5, f=1.2e-4, island=False):
    kw = cases.config_c3(n=n, seed=seed)
    kw["f"] = f
    if island:   # land: the corner depths of a rectangle off the centre
        mask = np.zeros(kw["H"].shape, bool)
        mask[n // 3:n // 3 + 6, n // 4:n // 4 + 9] = True
        kw["H"] = np.ma.array(kw["H"], mask=mask)
    if bc is not None:
        kw["boundary_conditions"] = bc
        if any(c == 3 for c in (bc.north, bc.east, bc.south, bc.west)):
            kw["boundary_conditions_data"] = _frs_data(kw)
    bg, truth = cases.c3_background_and_truth(kw, seed=seed)
    sim = CDKLM16(gpu_ctx, **kw)
    pos = cases.c3_drifters(n, n, 200.0, 200.0, n=n_obs, seed=seed)
    # observations: FP32 run of the truth, observed at the positions
    probe = DifferentiableCDKLM16(sim, pos, np.zeros_like(pos), n_steps, checkpoint_every=k)
    obs = _observe(probe, truth, n_steps)
    model = DifferentiableCDKLM16(sim, pos, obs, n_steps, checkpoint_every=k)
    return kw, sim, model, bg


def _observe(model, state, n_steps):
    """(u, v) of the FP64 run at the model's positions after n_steps."""
    model._init_f64()
    Qs, _ = model.trajectory_f64(*[torch.from_numpy(a).double().cuda() for a in state], n_steps=n_steps)
    e, u, v = [a.cpu().numpy() for a in Qs[-1].views()]
    Hi = model._f64["Hi"].cpu().numpy()
    Hm = 0.25 * (Hi[:-1, :-1] + Hi[1:, :-1] + Hi[:-1, 1:] + Hi[1:, 1:])[:e.shape[0], :e.shape[1]]
    out = []
    for x, y in model.pos.cpu().numpy():
        sx, sy = x / model.dx - 0.5, y / model.dy - 0.5
        i0 = min(max(int(np.floor(sx)), 0), model.nx - 2)
        j0 = min(max(int(np.floor(sy)), 0), model.ny - 2)
        ax, ay = sx - i0, sy - j0
        w = [(1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay]
        cells = [(j0 + 2, i0 + 2), (j0 + 2, i0 + 3), (j0 + 3, i0 + 2), (j0 + 3, i0 + 3)]
        uu = sum(wq * u[c] / (e[c] + Hm[c]) for wq, c in zip(w, cells))
        vv = sum(wq * v[c] / (e[c] + Hm[c]) for wq, c in zip(w, cells))
        out.append((uu, vv))
    return np.array(out, np.float32)


def _interior_noise(shape, rng, scale):
    a = np.zeros(shape)
    a[2:-2, 2:-2] = scale * rng.standard_normal((shape[0] - 4, shape[1] - 4))
    return a


BCS = [Common.BoundaryConditions(1, 1, 1, 1), Common.BoundaryConditions(1, 2, 1, 2),
       Common.BoundaryConditions(2, 2, 2, 2)]
_SPONGE = {'north': 8, 'east': 8, 'south': 8, 'west': 8}
BCS_FRS = [Common.BoundaryConditions(3, 3, 3, 3, spongeCells=_SPONGE),
           Common.BoundaryConditions(3, 2, 3, 2, spongeCells=_SPONGE),
           Common.BoundaryConditions(1, 3, 1, 3, spongeCells=_SPONGE)]


@pytest.mark.parametrize("bc", BCS, ids=["walls", "periodic_x", "periodic"])
def test_f64_forward_matches_f64_oracle(gpu_ctx, bc):
    print(f"f64 forward vs oracle, bc={bc.packed():#x}", flush=True)
    kw, sim, model, bg = _model(gpu_ctx, n=48, bc=bc, n_steps=15)
    # the FP64 GPU build widens the float32 parameter block, so the oracle gets
    # the float32 value of the flux balancer (its f64 build would widen 0.8)
    ora = O.OracleCDKLM16(**{**kw, "eta0": bg[0], "hu0": bg[1], "hv0": bg[2], "dt": float(model.dt)},
                          flux_balancer=float(np.float32(0.8)), precision="f64")
    ora.run_steps(15, float(model.dt))
    Qs, _ = model.trajectory_f64(*[torch.from_numpy(a).double().cuda() for a in bg], n_steps=15)
    for a, b in zip(Qs[-1].views(), ora.download()):
        a = a.cpu().numpy()[2:-2, 2:-2]
        b = np.asarray(b, np.float64)[2:-2, 2:-2]
        assert np.max(np.abs(a - b)) <= 1e-11 * max(1.0, np.max(np.abs(b)))
    assert model.flag_value() == 0


@pytest.mark.parametrize("bc", BCS + BCS_FRS, ids=["walls", "periodic_x", "periodic", "frs", "frs_ns_periodic_ew",
                                                   "walls_ns_frs_ew"])
@pytest.mark.parametrize("island", [False, True], ids=["", "island"])
def test_dot_product_f64(gpu_ctx, bc, island):
    """<TLM v, w> == <v, ADJ w> along an FP64 trajectory of 10 steps."""
    print(f"dot product, bc={bc.packed():#x}", flush=True)
    kw, sim, model, bg = _model(gpu_ctx, n=40, bc=bc, n_steps=10, island=island)
    Qs, Q1s = model.trajectory_f64(*[torch.from_numpy(a).double().cuda() for a in bg])
    rng = np.random.default_rng(1)
    shape = bg[0].shape
    v = model._new_state(torch.float64)
    w = model._new_state(torch.float64)
    for a, s in zip(v.views(), (1e-3, 1e-1, 1e-1)):
        a.copy_(torch.from_numpy(_interior_noise(shape, rng, s)))
    for a in w.views():
        a.copy_(torch.from_numpy(_interior_noise(shape, rng, 1.0)))
    tv = model.tangent_f64(Qs, Q1s, v)
    aw = model.adjoint_f64(Qs, Q1s, w)
    lhs = sum(float((x[2:-2, 2:-2] * y[2:-2, 2:-2]).sum()) for x, y in zip(tv.views(), w.views()))
    rhs = sum(float((x[2:-2, 2:-2] * y[2:-2, 2:-2]).sum()) for x, y in zip(v.views(), aw.views()))
    print(f"  <TLM v, w> = {lhs:.15e}, <v, ADJ w> = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(rhs))
    assert model.flag_value() == 0


def _smooth_direction(shape, rng, scale):
    """Seeded smooth random field (Gaussian-filtered, 4 cells) on the interior:
    a perturbation whose own slopes are small, so that a step of 1e-6 along it
    crosses few minmod branch switches."""
    a = np.zeros(shape)
    f = cases.smooth_random_field(shape, 4.0, int(rng.integers(1 << 30)))
    a[2:-2, 2:-2] = scale * (f / np.max(np.abs(f)))[2:-2, 2:-2]
    return a


def _fd_check(model, bg, n_dirs, rng, tol64, tol32):
    """Directional derivatives <g, d> of the FP32 product gradient and of the
    FP64 adjoint build against FP64 central differences (eps 1e-6) along
    smooth random directions; errors relative to the RMS of the FD values."""
    x0 = [torch.from_numpy(a).double().cuda() for a in bg]
    x32 = [torch.from_numpy(a).cuda() for a in bg]
    J32, _ = model.forward(*x32)
    g32 = [a.double().cpu().numpy() for a in model.backward()]
    J64, g64 = model.gradient_f64(*x0)
    g64 = [a.cpu().numpy() for a in g64]
    assert abs(J32 - J64) <= 1e-3 * abs(J64)
    eps = 1e-6
    fds, a32, a64 = [], [], []
    for _ in range(n_dirs):
        d = [_smooth_direction(bg[0].shape, rng, s) for s in (1e-2, 1.0, 1.0)]
        xp = [a + eps * torch.from_numpy(b).cuda() for a, b in zip(x0, d)]
        xm = [a - eps * torch.from_numpy(b).cuda() for a, b in zip(x0, d)]
        fds.append((model.misfit_f64(*xp) - model.misfit_f64(*xm)) / (2 * eps))
        a32.append(sum(float((gg * dd).sum()) for gg, dd in zip(g32, d)))
        a64.append(sum(float((gg * dd).sum()) for gg, dd in zip(g64, d)))
    fds, a32, a64 = np.array(fds), np.array(a32), np.array(a64)
    rms = np.sqrt(np.mean(fds ** 2))
    e32, e64 = np.abs(a32 - fds) / rms, np.abs(a64 - fds) / rms
    for i in range(n_dirs):
        print(f"  FD {fds[i]: .10e}  FP32 adjoint {a32[i]: .10e} ({e32[i]:.2e})  "
              f"FP64 adjoint {a64[i]: .10e} ({e64[i]:.2e})", flush=True)
    print(f"  worst: FP32 {e32.max():.2e}, FP64 {e64.max():.2e} (relative to the RMS directional derivative)",
          flush=True)
    assert e64.max() <= tol64, e64.max()
    assert e32.max() <= tol32, e32.max()
    return g32, g64


@pytest.mark.parametrize("bc,island", [(BCS[0], False), (BCS[1], False), (BCS_FRS[0], False), (BCS_FRS[1], False),
                                       (BCS[1], True), (BCS_FRS[0], True)],
                         ids=["walls", "periodic_x", "frs", "frs_ns_periodic_ew", "periodic_x_island", "frs_island"])
def test_gradient_against_finite_differences_small(gpu_ctx, bc, island):
    print(f"FD check 64x64, bc={bc.packed():#x} island={island}", flush=True)
    kw, sim, model, bg = _model(gpu_ctx, n=64, bc=bc, n_steps=40, k=7, n_obs=30, island=island)
    # next to land, near-zero face depths put adjustSlopes kinks within reach
    # of a 1e-6 step along some directions: a looser FP64 bound there
    _fd_check(model, bg, 4, np.random.default_rng(2), 1e-4 if island else 1e-5, 1e-3)
    assert model.flag_value() == 0


def test_autograd_function_matches_direct_calls(gpu_ctx):
    print("autograd wrapper", flush=True)
    kw, sim, model, bg = _model(gpu_ctx, n=64, n_steps=12, k=5)
    x = [torch.from_numpy(a).cuda().requires_grad_() for a in bg]
    J = CDKLM16Misfit.apply(*x, model)
    (2.5 * J).backward()
    J2, _ = model.forward(*[a.detach() for a in x])
    g = model.backward(2.5)
    assert J.item() == J2
    for a, b in zip(x, g):
        assert torch.equal(a.grad, b)
        assert float(a.grad[:2].abs().sum()) == 0.0   # ghost rows carry no sensitivity
    # the checkpoint interval does not change the gradient
    m1 = DifferentiableCDKLM16(sim, model.pos.cpu().numpy(), model.obs.cpu().numpy(), 12, checkpoint_every=1)
    m1.forward(*[a.detach() for a in x])
    for a, b in zip(m1.backward(2.5), g):
        assert torch.equal(a, b)


def test_c3_gradient_512(gpu_ctx):
    """C3: 512x512, 400 steps, 200 observation points, checkpoints every 20
    steps; 16 smooth random directions against FP64 central differences
    (eps 1e-6).  The FP64 build of the adjoint agrees with the differences to
    1e-3.  The FP32 product gradient linearises the FP32 trajectory, which after
    400 steps differs from the FP64 one by ~2e-4 (rounding growth); the minmod
    limiter switches branches on that scale, so it is held to 5e-2 against the
    differences, and the FP32 adjoint arithmetic itself is pinned separately
    along the FP64 trajectory (1e-3)."""
    print("C3 512x512 FD check", flush=True)
    kw, sim, model, bg = _model(gpu_ctx, n=512, n_steps=400, k=20, n_obs=200, seed=3)
    g32, g64 = _fd_check(model, bg, 16, np.random.default_rng(7), 1e-3, 5e-2)
    for f in range(3):
        rel = np.linalg.norm(g32[f] - g64[f]) / np.linalg.norm(g64[f])
        print(f"  FP32 vs FP64 gradient, field {f}: relative L2 {rel:.2e}", flush=True)
        assert rel <= 5e-2
    # the FP32 adjoint kernels along the FP64 trajectory (cast to float)
    Qs, Q1s = model.trajectory_f64(*[torch.from_numpy(a).double().cuda() for a in bg])
    st = model.stream

    def cast(S):
        T = model._new_state()
        with torch.cuda.stream(st):
            for a, b in zip(T.t, S.t):
                a.copy_(b.float())
        return T

    lam, L1, X = model._new_state(), model._new_state(), model._new_state()
    with torch.cuda.stream(st):
        lam.zero_()
        model.J.zero_()
    model._misfit32(cast(Qs[-1]), lam)
    for j in reversed(range(model.n_steps)):
        q1, qn, qn1 = cast(Q1s[j]), cast(Qs[j]), cast(Qs[j + 1])
        with torch.cuda.stream(st):
            X.zero_()
        model._adj32(1, q1, qn1, lam, L1, X, 0)
        model._adj32(0, qn, q1, L1, X, None, 1)
        lam, X = X, lam
    st.synchronize()
    for f, a in enumerate(lam.views()):
        rel = np.linalg.norm(a.double().cpu().numpy() - g64[f]) / np.linalg.norm(g64[f])
        print(f"  FP32 adjoint on the FP64 trajectory, field {f}: relative L2 {rel:.2e}", flush=True)
        assert rel <= 1e-3
    assert model.flag_value() == 0

"""Derivative of stochastic steps (SOAR model error with fixed normals, whose
balanced momentum depends on the state) and of drifter advection
(driftKernels.cu:74-183) with a drifter-position misfit.

* The differentiated FP32 forward is the product's own stochastic step with
  drifters, bit for bit.
* gpo_balance_adj_f32 against the closed form of the balance's state
  derivative, and the FP64 balance tangent against its adjoint.
* Dot-product tests (tangent-linear model against adjoint) in FP64 on the
  full step with model error and drifters.
* Central finite differences (eps 1e-6, FP64 forward) against the FP64
  adjoint build and the FP32 product gradient.
"""
import ctypes

import numpy as np
import pytest
import torch

import cases
from gpuocean_amd import CDKLM16, Common, GPUDrifterCollection
from gpuocean_amd import _lib as L
from gpuocean_amd.adjoint import DifferentiableCDKLM16

pytestmark = pytest.mark.gpu

Q0 = 1e-4   # SOAR amplitude of these small 200 m grids


def _frs_data(kw):
    e, u, v = kw["eta0"], kw["hu0"], kw["hv0"]
    ny, nx = e.shape[0] - 4, e.shape[1] - 4

    def side(sl):
        return Common.SingleBoundaryConditionData(h=[e[sl].copy()], hu=[u[sl].copy()], hv=[v[sl].copy()])

    return Common.BoundaryConditionsData(north=side(np.s_[ny + 1, :]), south=side(np.s_[2, :]),
                                         east=side(np.s_[:, nx + 1]), west=side(np.s_[:, 2]))


def _sim(gpu_ctx, n, bc, seed, island):
    kw = cases.config_c3(n=n, seed=seed)
    if island:
        mask = np.zeros(kw["H"].shape, bool)
        mask[n // 3:n // 3 + 6, n // 4:n // 4 + 9] = True
        kw["H"] = np.ma.array(kw["H"], mask=mask)
    kw["boundary_conditions"] = bc
    if any(c == 3 for c in (bc.north, bc.east, bc.south, bc.west)):
        kw["boundary_conditions_data"] = _frs_data(kw)
    bg, truth = cases.c3_background_and_truth(kw, seed=seed)
    sim = CDKLM16(gpu_ctx, **{**kw, "eta0": bg[0], "hu0": bg[1], "hv0": bg[2]})
    sim.setSOARModelError(Q0, xorwow_seed=1234 + seed)
    return kw, sim, bg, truth


def _drifters(n, n_d, seed):
    rng = np.random.default_rng(seed + 77)
    return np.stack([rng.uniform(0.3 * n * 200.0, 0.7 * n * 200.0, n_d),
                     rng.uniform(0.3 * n * 200.0, 0.7 * n * 200.0, n_d)], axis=1).astype(np.float32)


def _model(gpu_ctx, n=48, bc=None, n_steps=12, k=5, seed=5, island=False, noise=True, n_d=40, n_obs=0):
    bc = Common.BoundaryConditions(1, 1, 1, 1) if bc is None else bc
    kw, sim, bg, truth = _sim(gpu_ctx, n, bc, seed, island)
    x0 = _drifters(n, n_d, seed)
    pos = cases.c3_drifters(n, n, 200.0, 200.0, n=n_obs, seed=seed) if n_obs else np.zeros((0, 2), np.float32)
    probe = DifferentiableCDKLM16(sim, pos, np.zeros_like(pos), n_steps, checkpoint_every=k, model_error=noise,
                                  drifter_positions=x0, drifter_observed=x0)
    tr = probe._trajectory_generic64(*[torch.from_numpy(a).double().cuda() for a in truth])
    xobs = tr["Xs"][-1].float().cpu().numpy()
    obs = np.zeros_like(pos)
    model = DifferentiableCDKLM16(sim, pos, obs, n_steps, checkpoint_every=k, model_error=noise,
                                  drifter_positions=x0, drifter_observed=xobs)
    return sim, model, bg, x0


BCS = {"walls": Common.BoundaryConditions(1, 1, 1, 1), "periodic_x": Common.BoundaryConditions(1, 2, 1, 2),
       "frs": Common.BoundaryConditions(3, 3, 3, 3, spongeCells={'north': 6, 'east': 6, 'south': 6, 'west': 6})}


@pytest.mark.parametrize("bc", ["walls", "periodic_x"])
def test_differentiated_forward_is_the_product_step(gpu_ctx, bc):
    """The model's FP32 forward (model error of draws 0.., drifters) equals
    CDKLM16.step with the same model error and drifters attached."""
    print(f"differentiated forward vs product step, {bc}", flush=True)
    sim, model, bg, x0 = _model(gpu_ctx, n=64, bc=BCS[bc], n_steps=9, k=4)
    J, A = model.forward(*[torch.from_numpy(a).cuda() for a in bg])
    d = GPUDrifterCollection(gpu_ctx, len(x0) - 1, boundaryConditions=sim.boundary_conditions,
                             domain_size_x=sim.nx * sim.dx, domain_size_y=sim.ny * sim.dy, gpu_stream=sim.gpu_stream)
    d.setDrifterPositions(x0[:-1])
    d.setObservationPosition(x0[-1])
    sim.attachDrifters(d)
    dt = np.float32(model.dt)
    for _ in range(9):
        sim.step(dt)
    for k, (a, b) in enumerate(zip([t.cpu().numpy() for t in A.views()], sim.download())):
        assert np.array_equal(a[2:-2, 2:-2], b[2:-2, 2:-2]), f"field {k}"
    got = np.concatenate([d.getDrifterPositions(), [d.getObservationPosition()]], axis=0)
    assert np.array_equal(got, model._final_x.cpu().numpy())
    assert not np.array_equal(got, x0)


def test_balance_adjoint_f32_closed_form(gpu_ctx):
    """gpo_balance_adj_f32: lambda_eta += -g/f d_y(d) lambda_hu + g/f d_x(d)
    lambda_hv on wet cells (central differences of the coarse field d), all
    three zero on cells with a land corner; lambda_hu, lambda_hv unchanged."""
    n = 40
    kw, sim, bg, truth = _sim(gpu_ctx, n, BCS["walls"], 3, island=True)
    rng = np.random.default_rng(4)
    coarse = rng.standard_normal((n + 4, n + 4)).astype(np.float32)
    lam = [rng.standard_normal((n + 4, n + 4)).astype(np.float32) for _ in range(3)]
    P = sim.h0[0].pitch_elems
    dev = sim.device
    c_t = torch.from_numpy(coarse).to(dev)
    lam_t = [torch.zeros((n + 4, P), dtype=torch.float32, device=dev) for _ in range(3)]
    for a, b in zip(lam_t, lam):
        a[:, :n + 4] = torch.from_numpy(b)
    g, f = np.float32(sim.g), np.float32(sim.f)
    L.check(L.lib().gpo_balance_adj_f32(ctypes.byref(sim.params), float(g), float(f), 0.0, None, 0,
                                         ctypes.c_void_p(c_t.data_ptr()), 4 * (n + 4),
                                         *[ctypes.c_void_p(a.data_ptr()) for a in lam_t], 4 * P,
                                         ctypes.c_void_p(sim.Hi.data_ptr()), sim.Hi.pitch, None),
            "gpo_balance_adj_f32")
    torch.cuda.synchronize()
    out = [a[:, :n + 4].cpu().numpy().astype(np.float64) for a in lam_t]
    Hi = np.asarray(sim.downloadBathymetry()[0].data, np.float64)
    land = np.float64(np.float32(1e20))
    gc = np.float64(g) / np.float64(f)
    wet_seen = dry_seen = 0
    for j in range(2, n + 2):
        for i in range(2, n + 2):
            corners = (Hi[j, i], Hi[j, i + 1], Hi[j + 1, i], Hi[j + 1, i + 1])
            if land in corners:
                dry_seen += 1
                assert out[0][j, i] == 0 and out[1][j, i] == 0 and out[2][j, i] == 0
                continue
            wet_seen += 1
            ddx = (np.float64(coarse[j, i + 1]) - coarse[j, i - 1]) / (2 * 200.0)
            ddy = (np.float64(coarse[j + 1, i]) - coarse[j - 1, i]) / (2 * 200.0)
            ref = lam[0][j, i] - gc * ddy * lam[1][j, i] + gc * ddx * lam[2][j, i]
            assert abs(out[0][j, i] - ref) <= 1e-5 * (abs(ref) + abs(gc * ddy * lam[1][j, i]) + 1e-3)
            assert out[1][j, i] == lam[1][j, i] and out[2][j, i] == lam[2][j, i]
    assert wet_seen > 1000 and dry_seen > 0


def _interior_noise(shape, rng, scale):
    a = np.zeros(shape)
    a[2:-2, 2:-2] = scale * rng.standard_normal((shape[0] - 4, shape[1] - 4))
    return a


@pytest.mark.parametrize("bc,island", [("walls", True), ("periodic_x", False), ("frs", False)])
def test_dot_product_f64_with_model_error_and_drifters(gpu_ctx, bc, island):
    """<TLM (v, vx), (w, wx)> == <(v, vx), ADJ (w, wx)> over 10 stochastic steps
    with drifters: state and drifter positions together."""
    print(f"dot product with model error and drifters, {bc} island={island}", flush=True)
    sim, model, bg, x0 = _model(gpu_ctx, n=40, bc=BCS[bc], n_steps=10, island=island)
    tr = model._trajectory_generic64(*[torch.from_numpy(a).double().cuda() for a in bg])
    rng = np.random.default_rng(1)
    shape = bg[0].shape
    v, w = model._new_state(torch.float64), model._new_state(torch.float64)
    for a, s in zip(v.views(), (1e-3, 1e-1, 1e-1)):
        a.copy_(torch.from_numpy(_interior_noise(shape, rng, s)))
    for a in w.views():
        a.copy_(torch.from_numpy(_interior_noise(shape, rng, 1.0)))
    vx = torch.from_numpy(rng.standard_normal(x0.shape)).cuda()
    wx = torch.from_numpy(rng.standard_normal(x0.shape)).cuda()
    tv, tx = model._tangent_generic64(tr, v, vx)
    aw, ax = model._adjoint_generic64(tr, w, wx)
    model.stream.synchronize()
    lhs = sum(float((x[2:-2, 2:-2] * y[2:-2, 2:-2]).sum()) for x, y in zip(tv.views(), w.views()))
    lhs += float((tx * wx).sum())
    rhs = sum(float((x[2:-2, 2:-2] * y[2:-2, 2:-2]).sum()) for x, y in zip(v.views(), aw.views()))
    rhs += float((vx * ax).sum())
    print(f"  <TLM v, w> = {lhs:.15e}, <v, ADJ w> = {rhs:.15e}, relative {abs(lhs - rhs) / abs(lhs):.2e}",
          flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))
    assert model.flag_value() == 0


def _smooth_direction(shape, rng, scale):
    a = np.zeros(shape)
    f = cases.smooth_random_field(shape, 4.0, int(rng.integers(1 << 30)))
    a[2:-2, 2:-2] = scale * (f / np.max(np.abs(f)))[2:-2, 2:-2]
    return a


@pytest.mark.parametrize("noise,bc", [(True, "walls"), (False, "periodic_x"), (True, "frs")])
def test_drifter_position_gradient_against_finite_differences(gpu_ctx, noise, bc):
    """dJ/dQ0 of the drifter-position misfit (40 drifters, 30 steps, model
    error with fixed normals when `noise`), FP64 adjoint and FP32 product
    against FP64 central differences (eps 1e-6) along 4 smooth directions."""
    print(f"drifter-position FD check, noise={noise}, {bc}", flush=True)
    sim, model, bg, x0 = _model(gpu_ctx, n=64, bc=BCS[bc], n_steps=30, k=7, noise=noise)
    x64 = [torch.from_numpy(a).double().cuda() for a in bg]
    J32, _ = model.forward(*[torch.from_numpy(a).cuda() for a in bg])
    g32 = [a.double().cpu().numpy() for a in model.backward()]
    J64, g64 = model.gradient_f64(*x64)
    g64 = [a.cpu().numpy() for a in g64]
    print(f"  J: FP32 {J32:.9e}, FP64 {J64:.9e}", flush=True)
    assert J64 > 0 and abs(J32 - J64) <= 1e-2 * J64
    rng = np.random.default_rng(2)
    eps = 1e-6
    fds, a32, a64 = [], [], []
    for _ in range(4):
        d = [_smooth_direction(bg[0].shape, rng, s) for s in (1e-2, 1.0, 1.0)]
        xp = [a + eps * torch.from_numpy(b).cuda() for a, b in zip(x64, d)]
        xm = [a - eps * torch.from_numpy(b).cuda() for a, b in zip(x64, d)]
        fds.append((model.misfit_f64(*xp) - model.misfit_f64(*xm)) / (2 * eps))
        a32.append(sum(float((gg * dd).sum()) for gg, dd in zip(g32, d)))
        a64.append(sum(float((gg * dd).sum()) for gg, dd in zip(g64, d)))
    fds, a32, a64 = np.array(fds), np.array(a32), np.array(a64)
    rms = np.sqrt(np.mean(fds ** 2))
    e32, e64 = np.abs(a32 - fds) / rms, np.abs(a64 - fds) / rms
    for i in range(len(fds)):
        print(f"  FD {fds[i]: .10e}  FP32 {a32[i]: .10e} ({e32[i]:.2e})  FP64 {a64[i]: .10e} ({e64[i]:.2e})",
              flush=True)
    assert e64.max() <= 1e-3
    assert e32.max() <= 2e-2
    assert model.flag_value() == 0

"""Batched differentiable ensemble: every member's misfit and gradient, from
single launches over all members, equal bit for bit those of the member's
own DifferentiableCDKLM16 run (whose gradient the adjoint tests pin by
dot-product and finite-difference checks), with and without land."""
import numpy as np
import pytest
import torch

import cases
from gpuocean_amd import CDKLM16, Common
from gpuocean_amd.adjoint import DifferentiableCDKLM16, DifferentiableEnsemble

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (96, 64), (80, 80), (128, 40), (48, 96)]


def _member(gpu_ctx, nx, ny, seed, bc, island=False):
    kw = cases.config_c0(n=64)
    rng = np.random.default_rng(seed)
    jj, ii = np.meshgrid(np.arange(ny + 4), np.arange(nx + 4), indexing="ij")
    eta = 0.3 * np.exp(-((ii - nx / 2 - 3) ** 2 + (jj - ny / 2 + 2) ** 2) / (0.02 * nx * ny))
    p = cases.smooth_random_field((ny + 5, nx + 5), 6.0, seed)
    H = (60.0 + 10.0 * p / np.max(np.abs(p))).astype(np.float32)
    if island:   # land: corner depths of a rectangle in the western half
        mask = np.zeros(H.shape, bool)
        mask[ny // 3:ny // 3 + 5, nx // 5:nx // 5 + 7] = True
        H = np.ma.array(H, mask=mask)
    eta, hu, hv = eta.astype(np.float32), np.zeros((ny + 4, nx + 4), np.float32), np.zeros((ny + 4, nx + 4), np.float32)
    kw.update(eta0=eta, hu0=hu, hv0=hv, nx=nx, ny=ny, H=H, boundary_conditions=bc)
    if any(c == 3 for c in (bc.north, bc.east, bc.south, bc.west)):   # exterior: the initial boundary rows
        def side(sl):
            return Common.SingleBoundaryConditionData(h=[eta[sl].copy()], hu=[hu[sl].copy()], hv=[hv[sl].copy()])
        kw["boundary_conditions_data"] = Common.BoundaryConditionsData(
            north=side(np.s_[ny + 1, :]), south=side(np.s_[2, :]), east=side(np.s_[:, nx + 1]),
            west=side(np.s_[:, 2]))
    sim = CDKLM16(gpu_ctx, **kw)
    x = rng.uniform(2 * sim.dx, (nx - 2) * sim.dx, 12)
    y = rng.uniform(2 * sim.dy, (ny - 2) * sim.dy, 12)
    pos = np.stack([x, y], 1).astype(np.float32)
    obs = (0.01 * rng.standard_normal((12, 2))).astype(np.float32)
    bg = [torch.from_numpy(a).cuda() for a in sim.download()]
    bg[1] = bg[1] + torch.from_numpy((0.02 * rng.standard_normal(bg[1].shape)).astype(np.float32)).cuda()
    return sim, pos, obs, bg


_SP = {'north': 6, 'east': 6, 'south': 6, 'west': 6}
BCS = [Common.BoundaryConditions(1, 1, 1, 1), Common.BoundaryConditions(2, 2, 2, 2),
       Common.BoundaryConditions(1, 2, 1, 2)]
BCS_FRS = [Common.BoundaryConditions(3, 3, 3, 3, spongeCells=_SP), Common.BoundaryConditions(1, 1, 1, 1),
           Common.BoundaryConditions(3, 2, 3, 2, spongeCells=_SP)]


@pytest.mark.parametrize("n_steps,k,island,frs", [(30, 8, False, False), (16, 16, False, False), (9, 4, False, False),
                                                  (12, 5, True, False), (11, 4, True, True)])
def test_batched_gradients_equal_per_member(gpu_ctx, n_steps, k, island, frs):
    print(f"batched adjoint, {n_steps} steps, segments of {k}, land {island}, flow relaxation {frs}", flush=True)
    bcs = BCS_FRS if frs else BCS
    members = [_member(gpu_ctx, nx, ny, 40 + i, bcs[i % 3], island) for i, (nx, ny) in enumerate(SIZES)]
    ens = DifferentiableEnsemble([m[0] for m in members], [m[1] for m in members], [m[2] for m in members],
                                 n_steps, checkpoint_every=k)
    J = ens.forward([m[3] for m in members])
    grads = ens.backward()
    assert ens.flag_value() == 0
    for i, (sim, pos, obs, bg) in enumerate(members):
        one = DifferentiableCDKLM16(sim, pos, obs, n_steps, checkpoint_every=k)
        J1, _ = one.forward(*bg)
        g1 = one.backward()
        assert J[i] == J1, f"member {i}"
        for q in range(3):
            a = grads[i][q][2:-2, 2:-2].cpu().numpy()
            b = g1[q][2:-2, 2:-2].cpu().numpy()
            assert np.array_equal(a, b), f"member {i} field {q}"
        assert grads[i][0].abs().max().item() > 0
    ens.destroy()


@pytest.mark.parametrize("frs,island", [(False, False), (True, True)])
def test_batched_stochastic_drifter_gradients_equal_per_member(gpu_ctx, frs, island):
    """Members with SOAR model error (fixed normals) and drifters, and the
    drifter-position misfit added to the point misfit: the batched misfits,
    final drifter positions and gradients equal each member's own
    step-by-step DifferentiableCDKLM16 run bit for bit."""
    n_steps, k = 13, 5
    print(f"batched stochastic drifter adjoint, flow relaxation {frs}, land {island}", flush=True)
    bcs = BCS_FRS if frs else BCS
    members = [_member(gpu_ctx, nx, ny, 60 + i, bcs[i % 3], island) for i, (nx, ny) in enumerate(SIZES)]
    dpos, dobs = [], []
    for i, (sim, pos, obs, bg) in enumerate(members):
        sim.setSOARModelError(1e-4, xorwow_seed=500 + i)
        rng = np.random.default_rng(90 + i)
        nd = 20 + 7 * i
        x = np.stack([rng.uniform(0.3 * sim.nx * sim.dx, 0.7 * sim.nx * sim.dx, nd),
                      rng.uniform(0.3 * sim.ny * sim.dy, 0.7 * sim.ny * sim.dy, nd)], 1).astype(np.float32)
        dpos.append(x)
        dobs.append((x + rng.normal(0, 30.0, x.shape)).astype(np.float32))
    ens = DifferentiableEnsemble([m[0] for m in members], [m[1] for m in members], [m[2] for m in members],
                                 n_steps, checkpoint_every=k, model_error=True, drifter_positions=dpos,
                                 drifter_observed=dobs)
    J = ens.forward([m[3] for m in members])
    xs = [ens.final_positions(i) for i in range(len(members))]
    grads = ens.backward()
    assert ens.flag_value() == 0
    for i, (sim, pos, obs, bg) in enumerate(members):
        one = DifferentiableCDKLM16(sim, pos, obs, n_steps, checkpoint_every=k, model_error=True,
                                    drifter_positions=dpos[i], drifter_observed=dobs[i])
        J1, _ = one.forward(*bg)
        assert torch.equal(xs[i], one._final_x), f"member {i}: drifter positions"
        g1 = one.backward()
        assert J[i] == J1, f"member {i}"
        for q in range(3):
            a = grads[i][q][2:-2, 2:-2].cpu().numpy()
            b = g1[q][2:-2, 2:-2].cpu().numpy()
            assert np.array_equal(a, b), f"member {i} field {q}"
        assert grads[i][0].abs().max().item() > 0
        assert not torch.equal(xs[i], torch.from_numpy(dpos[i]).cuda())
    ens.destroy()


def test_crowded_drifters_and_points_give_the_same_gradient_every_run(gpu_ctx):
    """Hundreds of drifters and observation points crowded into a few cells
    (many share a cell, across chunks of 256): their sensitivities are added
    in drifter / point order, so two batched runs and each member's own run
    give the same misfits and gradients bit for bit."""
    n_steps, k = 6, 3
    print("crowded drifters and points: repeatability", flush=True)
    sizes, bcs = [(64, 48), (80, 80)], [BCS[0], BCS[1]]
    members, dpos, dobs = [], [], []
    for i, ((nx, ny), bc) in enumerate(zip(sizes, bcs)):
        sim, _, _, bg = _member(gpu_ctx, nx, ny, 120 + i, bc)
        sim.setSOARModelError(1e-4, xorwow_seed=700 + i)
        rng = np.random.default_rng(130 + i)
        cx, cy = 0.5 * nx * sim.dx, 0.5 * ny * sim.dy
        pts = np.stack([rng.uniform(cx - 2 * sim.dx, cx + 2 * sim.dx, 300),
                        rng.uniform(cy - 2 * sim.dy, cy + 2 * sim.dy, 300)], 1).astype(np.float32)
        pobs = (0.01 * rng.standard_normal((300, 2))).astype(np.float32)
        x = np.stack([rng.uniform(cx - 3 * sim.dx, cx + 3 * sim.dx, 600),
                      rng.uniform(cy - 3 * sim.dy, cy + 3 * sim.dy, 600)], 1).astype(np.float32)
        members.append((sim, pts, pobs, bg))
        dpos.append(x)
        dobs.append((x + rng.normal(0, 30.0, x.shape)).astype(np.float32))

    def batched():
        ens = DifferentiableEnsemble([m[0] for m in members], [m[1] for m in members], [m[2] for m in members],
                                     n_steps, checkpoint_every=k, model_error=True, drifter_positions=dpos,
                                     drifter_observed=dobs)
        J = ens.forward([m[3] for m in members])
        g = ens.backward()
        assert ens.flag_value() == 0
        out = ([float(j) for j in J], [[g[i][q].clone() for q in range(3)] for i in range(len(members))])
        ens.destroy()
        return out

    J_a, g_a = batched()
    J_b, g_b = batched()
    assert J_a == J_b
    for i, (sim, pts, pobs, bg) in enumerate(members):
        for q in range(3):
            assert torch.equal(g_a[i][q], g_b[i][q]), f"member {i} field {q}: repeat"
        one = DifferentiableCDKLM16(sim, pts, pobs, n_steps, checkpoint_every=k, model_error=True,
                                    drifter_positions=dpos[i], drifter_observed=dobs[i])
        J1, _ = one.forward(*bg)
        g1 = one.backward()
        assert J_a[i] == J1, f"member {i}"
        for q in range(3):
            assert np.array_equal(g_a[i][q][2:-2, 2:-2].cpu().numpy(), g1[q][2:-2, 2:-2].cpu().numpy()), \
                f"member {i} field {q}"
        assert g_a[i][0].abs().max().item() > 0


@pytest.mark.parametrize("frs,island,noise", [(False, False, False), (False, True, False), (True, True, True)])
def test_fused_reverse_equals_staged(gpu_ctx, frs, island, noise):
    """The fused strip-march reverse and the four staged reverse kernels give
    the same misfits and gradients bit for bit, for chunk heights that do and
    do not divide the grids, over walls, periodic sides, flow relaxation,
    land, model error and drifters, on grids wider than one strip."""
    n_steps, k = 7, 4
    sizes = [(130, 70), (57, 33), (64, 130), (200, 41)]
    print(f"fused reverse against staged, flow relaxation {frs}, land {island}, model error {noise}", flush=True)
    bcs = BCS_FRS if frs else BCS
    members = [_member(gpu_ctx, nx, ny, 80 + i, bcs[i % 3], island) for i, (nx, ny) in enumerate(sizes)]
    dpos = dobs = None
    if noise:
        dpos, dobs = [], []
        for i, (sim, pos, obs, bg) in enumerate(members):
            sim.setSOARModelError(1e-4, xorwow_seed=700 + i)
            rng = np.random.default_rng(110 + i)
            x = np.stack([rng.uniform(0.3 * sim.nx * sim.dx, 0.7 * sim.nx * sim.dx, 25),
                          rng.uniform(0.3 * sim.ny * sim.dy, 0.7 * sim.ny * sim.dy, 25)], 1).astype(np.float32)
            dpos.append(x)
            dobs.append((x + rng.normal(0, 30.0, x.shape)).astype(np.float32))
    ens = DifferentiableEnsemble([m[0] for m in members], [m[1] for m in members], [m[2] for m in members],
                                 n_steps, checkpoint_every=k, model_error=noise, drifter_positions=dpos,
                                 drifter_observed=dobs, fused_reverse=False)
    J0 = ens.forward([m[3] for m in members])
    g0 = [[a.cpu().numpy() for a in g] for g in ens.backward()]
    assert ens.flag_value() == 0
    for rows in (32, 7, 200):
        ens.set_reverse(True, rows)
        J = ens.forward([m[3] for m in members])
        g = ens.backward()
        assert np.array_equal(J, J0), f"rows {rows}"
        for i in range(len(members)):
            for q in range(3):
                a = g[i][q][2:-2, 2:-2].cpu().numpy()
                assert np.array_equal(a, g0[i][q][2:-2, 2:-2]), f"rows {rows} member {i} field {q}"
            assert np.abs(g0[i][0]).max() > 0
    ens.destroy()


def test_member_march_equals_staged(gpu_ctx):
    """A member large enough for the one-member sweep to take the fused march
    (at least 2048 strips x row chunks: 1240 x 2300 cells) gives the same
    gradient as the staged kernels (the batched path with fused_reverse=False)
    bit for bit, with land and periodic east/west sides."""
    nx, ny, n_steps, k = 1240, 2300, 3, 3
    print(f"one-member march against staged, {nx} x {ny}", flush=True)
    sim, pos, obs, bg = _member(gpu_ctx, nx, ny, 140, Common.BoundaryConditions(1, 2, 1, 2), island=True)
    one = DifferentiableCDKLM16(sim, pos, obs, n_steps, checkpoint_every=k)
    J1, _ = one.forward(*bg)
    g1 = one.backward()
    ens = DifferentiableEnsemble([sim], [pos], [obs], n_steps, checkpoint_every=k, fused_reverse=False)
    J = ens.forward([bg])
    g = ens.backward()
    assert J[0] == J1
    for q in range(3):
        assert np.array_equal(g[0][q][2:-2, 2:-2].cpu().numpy(), g1[q][2:-2, 2:-2].cpu().numpy()), f"field {q}"
    assert g1[0].abs().max().item() > 0
    ens.destroy()


@pytest.mark.parametrize("slots", [1, 2, 3])
def test_binomial_checkpointing_equals_every_segment(gpu_ctx, slots):
    """Binomial checkpointing with 1, 2 or 3 state slots (segments recomputed
    more often) gives the misfits, final drifter positions and gradients of
    the default schedule (every segment start kept) bit for bit, with flow
    relaxation, land, model error (its draw index follows the restored
    segment) and drifters."""
    n_steps, k = 23, 4
    print(f"binomial checkpointing, {slots} slots", flush=True)
    members = [_member(gpu_ctx, nx, ny, 160 + i, BCS_FRS[i % 3], True) for i, (nx, ny) in enumerate(SIZES[:3])]
    dpos, dobs = [], []
    for i, (sim, pos, obs, bg) in enumerate(members):
        sim.setSOARModelError(1e-4, xorwow_seed=900 + i)
        rng = np.random.default_rng(170 + i)
        x = np.stack([rng.uniform(0.3 * sim.nx * sim.dx, 0.7 * sim.nx * sim.dx, 30),
                      rng.uniform(0.3 * sim.ny * sim.dy, 0.7 * sim.ny * sim.dy, 30)], 1).astype(np.float32)
        dpos.append(x)
        dobs.append((x + rng.normal(0, 30.0, x.shape)).astype(np.float32))
    args = ([m[0] for m in members], [m[1] for m in members], [m[2] for m in members], n_steps)
    kw = dict(checkpoint_every=k, model_error=True, drifter_positions=dpos, drifter_observed=dobs)
    ref = DifferentiableEnsemble(*args, **kw)
    J0 = ref.forward([m[3] for m in members])
    x0 = [ref.final_positions(i) for i in range(len(members))]
    g0 = [[a[2:-2, 2:-2].cpu().numpy() for a in g] for g in ref.backward()]
    ref.destroy()
    ens = DifferentiableEnsemble(*args, checkpoint_slots=slots, **kw)
    assert ens.schedule is not None and ens.n_slots == slots
    J = ens.forward([m[3] for m in members])
    xs = [ens.final_positions(i) for i in range(len(members))]
    g = ens.backward()
    assert ens.flag_value() == 0
    assert np.array_equal(J, J0)
    for i in range(len(members)):
        assert torch.equal(xs[i], x0[i]), f"member {i}: drifter positions"
        for q in range(3):
            assert np.array_equal(g[i][q][2:-2, 2:-2].cpu().numpy(), g0[i][q]), f"member {i} field {q}"
    ens.destroy()

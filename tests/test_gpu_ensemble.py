"""Batched ensemble: uneven members in one persistent launch per substep (plus
batched SOAR model error and hipGraph replay) give bit for bit the values of
the per-member path, which the other GPU tests pin to the oracle."""
import numpy as np
import pytest
import torch

import cases
from gpuocean_amd import Autotuner, BatchedEnsemble, CDKLM16, Common, OceanModelEnsemble

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (96, 64), (128, 80), (48, 112), (176, 32), (240, 96)]


def _members(gpu_ctx, sizes, noise, seed0=11, bc=None):
    sims = []
    for i, (nx, ny) in enumerate(sizes):
        kw = cases.config_c2_member(nx, ny, seed0 + i)
        if bc is not None:
            kw["boundary_conditions"] = bc
        sim = CDKLM16(gpu_ctx, **kw, local_particle_id=i)
        if noise:
            sim.setSOARModelError(5e-3, xorwow_seed=1000 + i)
        sims.append(sim)
    return sims


def _assert_same(a_sims, b_sims):
    for i, (a, b) in enumerate(zip(a_sims, b_sims)):
        for k, (x, y) in enumerate(zip(a.download(), b.download())):
            assert np.array_equal(x, y), f"member {i} field {k}"
        assert a.t == b.t and a.num_iterations == b.num_iterations
        if a.model_error is not None:
            assert a.model_error.draw == b.model_error.draw


@pytest.mark.parametrize("noise", [False, True])
def test_batched_matches_per_member(gpu_ctx, noise):
    print(f"batched vs per-member, noise={noise}", flush=True)
    plain = _members(gpu_ctx, SIZES, noise)
    batched = _members(gpu_ctx, SIZES, noise)
    ens = BatchedEnsemble(batched)
    n_work, n_blocks = ens.info()
    assert n_work > 0 and n_blocks > 0
    for s in plain:
        for _ in range(7):
            s.step(np.float32(s.dt))
    ens.step(7)
    _assert_same(batched, plain)
    if noise:
        assert ens.device_draws() == [7] * len(SIZES)
        # the batched path keeps no normal/SOAR fields; the getters regenerate them
        for a, b in zip(batched, plain):
            assert np.array_equal(a.model_error.getRandomNumbers(), b.model_error.getRandomNumbers())
            assert np.array_equal(a.model_error.getCoarseBuffer(), b.model_error.getCoarseBuffer())


def test_batched_all_walls_and_all_periodic(gpu_ctx):
    for bc in (Common.BoundaryConditions(1, 1, 1, 1), Common.BoundaryConditions(2, 2, 2, 2)):
        print(f"batched, bc={bc.packed():#x}", flush=True)
        plain = _members(gpu_ctx, SIZES[:4], True, bc=bc)
        batched = _members(gpu_ctx, SIZES[:4], True, bc=bc)
        for s in plain:
            for _ in range(5):
                s.step(np.float32(s.dt))
        BatchedEnsemble(batched, chunk_rows=5, waves_per_simd=1).step(5)
        _assert_same(batched, plain)


def test_small_periodic_members_model_error(gpu_ctx):
    """Members narrower than the model-error tile window (periodic wrap inside
    one tile) and partial tiles."""
    print("small periodic members", flush=True)
    sizes = [(24, 20), (40, 18), (66, 35), (70, 17), (130, 33)]
    bc = Common.BoundaryConditions(2, 2, 2, 2)
    plain = _members(gpu_ctx, sizes, True, bc=bc)
    batched = _members(gpu_ctx, sizes, True, bc=bc)
    for s in plain:
        for _ in range(3):
            s.step(np.float32(s.dt))
    BatchedEnsemble(batched).step(3)
    _assert_same(batched, plain)


def test_graph_replay_matches_direct(gpu_ctx):
    print("graph replay vs direct launches", flush=True)
    a = _members(gpu_ctx, SIZES, True)
    b = _members(gpu_ctx, SIZES, True)
    ea, eb = BatchedEnsemble(a), BatchedEnsemble(b)
    ea.step(9, graph_steps=2)   # 4 replays of a 2-step graph + 1 direct step
    eb.step(9)
    _assert_same(a, b)
    # later replays of the same graph keep drawing fresh random numbers
    ea.step(4, graph_steps=2)
    eb.step(4)
    torch.cuda.synchronize()
    _assert_same(a, b)
    assert ea.device_draws() == [13] * len(SIZES)


def test_changed_dt_and_interleaved_per_member_steps(gpu_ctx):
    print("dt change and per-member steps between batched runs", flush=True)
    a = _members(gpu_ctx, SIZES[:3], True)
    b = _members(gpu_ctx, SIZES[:3], True)
    ea = BatchedEnsemble(a)
    ea.step(3)
    for s in b:
        for _ in range(3):
            s.step(np.float32(s.dt))
    for sa, sb in zip(a, b):   # one member-level step in between, with a new dt
        sa.dt = sb.dt = np.float32(0.5 * sa.dt)
    a[1].step(np.float32(a[1].dt))
    b[1].step(np.float32(b[1].dt))
    ea.step(4)
    for i, s in enumerate(b):
        for _ in range(4):
            s.step(np.float32(s.dt))
    _assert_same(a, b)


def test_autotuner_restores_state(gpu_ctx):
    print("autotuner", flush=True)
    a = _members(gpu_ctx, SIZES[:4], True)
    b = _members(gpu_ctx, SIZES[:4], True)
    ea = BatchedEnsemble(a, autotune=True)
    assert ea.chunk_rows in Autotuner.CHUNK_ROWS and ea.waves_per_simd in Autotuner.WAVES_PER_SIMD
    _assert_same(a, b)                       # tuning left no trace
    ea.step(3)
    for s in b:
        for _ in range(3):
            s.step(np.float32(s.dt))
    _assert_same(a, b)


def test_ocean_model_ensemble_batched_equals_sequential(gpu_ctx):
    print("OceanModelEnsemble.modelStep batched vs sequential", flush=True)
    sizes = SIZES[:5]
    data = [cases.config_c2_member(nx, ny, 40 + i) for i, (nx, ny) in enumerate(sizes)]
    sim_args = {}
    me_args = dict(small_scale_perturbation_amplitude=5e-3, xorwow_seed=77)
    ea = OceanModelEnsemble(gpu_ctx, sim_args, data, me_args, len(sizes), batched=True)
    eb = OceanModelEnsemble(gpu_ctx, sim_args, data, me_args, len(sizes), batched=False)
    for _ in range(3):
        ta = ea.modelStep(600.0)   # members have different dt: common full steps batched, tails per member
        tb = eb.modelStep(600.0)
        assert ta == tb
    assert ea._ensemble is not None
    _assert_same(ea.particles, eb.particles)
    for pa, pb in zip(ea.particles, eb.particles):
        assert pa.dt == pb.dt
    obs = np.array([[5000.0, 7000.0], [20000.0, 15000.0]])
    assert np.array_equal(ea.observeParticles(obs), eb.observeParticles(obs))


COARSE = [((96, 48), 3), ((160, 64), 5), ((112, 96), 7), ((48, 80), 3), ((150, 30), 5)]


def _coarse_members(gpu_ctx, bc=None):
    sims = []
    for i, ((nx, ny), factor) in enumerate(COARSE):
        kw = cases.config_c2_member(nx, ny, 31 + i)
        if bc is not None:
            kw["boundary_conditions"] = bc
        sim = CDKLM16(gpu_ctx, **kw, local_particle_id=i)
        sim.setSOARModelError(5e-3, small_scale_perturbation_interpolation_factor=factor, xorwow_seed=500 + i)
        sims.append(sim)
    return sims


@pytest.mark.parametrize("bc", [None, (1, 1, 1, 1)], ids=["periodic_x", "walls"])
def test_batched_coarse_model_error_matches_per_member(gpu_ctx, bc):
    """Model error generated on grids 3, 5 and 7 times coarser (normal field
    and SOAR on the coarse grid, bicubic interpolation and balance on the
    member's): three batched launches give the per-member values bit for bit,
    directly and through graph replays, and leave the members' normal and
    coarse fields as the per-member path leaves them."""
    print(f"batched coarse model error, bc={bc}", flush=True)
    bcs = None if bc is None else Common.BoundaryConditions(*bc)
    plain = _coarse_members(gpu_ctx, bcs)
    batched = _coarse_members(gpu_ctx, bcs)
    ens = BatchedEnsemble(batched)
    for s in plain:
        for _ in range(4):
            s.step(np.float32(s.dt))
    ens.step(4)
    _assert_same(batched, plain)
    assert ens.device_draws() == [4] * len(COARSE)
    for a, b in zip(batched, plain):
        assert np.array_equal(a.model_error.getRandomNumbers(), b.model_error.getRandomNumbers())
        assert np.array_equal(a.model_error.getCoarseBuffer(), b.model_error.getCoarseBuffer())
    # more steps through a captured graph of three steps, replayed twice
    for s in plain:
        for _ in range(6):
            s.step(np.float32(s.dt))
    ens.step(6, graph_steps=3)
    _assert_same(batched, plain)
    ens.destroy()


def test_coarse_and_fine_model_error_do_not_batch_together(gpu_ctx):
    sims = _coarse_members(gpu_ctx)[:2]
    kw = cases.config_c2_member(64, 48, 3)
    fine = CDKLM16(gpu_ctx, **kw, local_particle_id=9)
    fine.setSOARModelError(5e-3, xorwow_seed=7)
    ok, why = BatchedEnsemble.eligible(sims + [fine])
    assert not ok and "coarser grid" in why
    assert BatchedEnsemble.eligible(sims)[0]

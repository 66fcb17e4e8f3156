"""Batched ensemble with land, flow-relaxation boundaries, SOAR model error
and drifters (the coastal configuration): the batched launches give bit for
bit the states, drifter positions and times of each member's own step path,
which the other GPU tests pin to the oracle."""
import numpy as np
import pytest

import cases
from gpuocean_amd import BatchedEnsemble, CDKLM16, GPUDrifterCollection, OceanModelEnsemble

pytestmark = pytest.mark.gpu

SIZES = [(96, 64), (128, 80), (80, 96), (144, 72), (64, 64), (112, 88)]


def _members(gpu_ctx, codes, noise, n_drifters, seed0=21):
    sims = []
    for i, (nx, ny) in enumerate(SIZES):
        kw = cases.config_c4_member(nx, ny, seed0 + i, codes=codes)
        sim = CDKLM16(gpu_ctx, **kw, local_particle_id=i)
        if noise:
            sim.setSOARModelError(5e-3, xorwow_seed=3000 + i)
        if n_drifters:
            d = GPUDrifterCollection(gpu_ctx, n_drifters, boundaryConditions=sim.boundary_conditions,
                                     domain_size_x=nx * sim.dx, domain_size_y=ny * sim.dy, gpu_stream=sim.gpu_stream)
            d.setDrifterPositions(cases.c4_drifter_positions(nx, ny, sim.dx, sim.dy, n_drifters, seed0 + i))
            sim.attachDrifters(d)
        sims.append(sim)
    return sims


def _assert_same(a_sims, b_sims):
    for i, (a, b) in enumerate(zip(a_sims, b_sims)):
        for k, (x, y) in enumerate(zip(a.download(), b.download())):
            assert np.array_equal(x, y), f"member {i} field {k}"
        assert a.t == b.t and a.num_iterations == b.num_iterations
        if a.model_error is not None:
            assert a.model_error.draw == b.model_error.draw
        if a.drifters is not None:
            assert np.array_equal(a.drifters.getDrifterPositions(), b.drifters.getDrifterPositions()), f"member {i}"
            assert np.array_equal(a.drifters.getObservationPosition(), b.drifters.getObservationPosition())
            assert a.drifter_t == b.drifter_t


@pytest.mark.parametrize("codes,noise,n_drifters", [
    ((3, 3, 3, 1), True, 200),     # coastal: land in the west, sponges elsewhere
    ((3, 2, 3, 2), False, 0),      # sponges north/south, periodic east/west
    ((1, 2, 1, 2), True, 100),     # drifters reading periodic ghost cells
])
def test_batched_coastal_matches_per_member(gpu_ctx, codes, noise, n_drifters):
    print(f"coastal batched vs per-member, codes={codes} noise={noise} drifters={n_drifters}", flush=True)
    plain = _members(gpu_ctx, codes, noise, n_drifters)
    batched = _members(gpu_ctx, codes, noise, n_drifters)
    ok, why = BatchedEnsemble.eligible(batched)
    assert ok, why
    ens = BatchedEnsemble(batched)
    for s in plain:
        for _ in range(6):
            s.step(np.float32(s.dt))
    ens.step(6)
    _assert_same(batched, plain)
    # a second call (no first-step special case), through a captured graph
    for s in plain:
        for _ in range(4):
            s.step(np.float32(s.dt))
    ens.step(4, graph_steps=2)
    _assert_same(batched, plain)


def _time_varying(kw, factors=(1.0, 1.15, 0.9, 1.25), times=(0.0, 12.0, 30.0, 45.0)):
    """The member's boundary data at four times: its jet scaled per level."""
    d = kw["boundary_conditions_data"]
    for side in (d.north, d.south, d.east, d.west):
        side.h = [np.float32(f) * side.h[0] for f in factors]
        side.hu = [np.float32(f) * side.hu[0] for f in factors]
        side.hv = [np.float32(f) * side.hv[0] for f in factors]
        side.numSteps = len(factors)
    d.t, d.numSteps = list(times), len(times)
    return kw


def _frs_members(gpu_ctx, codes):
    sims = []
    for i, (nx, ny) in enumerate(SIZES):
        kw = cases.config_c4_member(nx, ny, 40 + i, codes=codes)
        if i % 2 == 0:
            kw = _time_varying(kw)
        sims.append(CDKLM16(gpu_ctx, **kw, local_particle_id=i))
    return sims


@pytest.mark.parametrize("codes", [(3, 3, 3, 1), (3, 2, 3, 2), (1, 3, 1, 3)])
def test_time_varying_frs_data_batched(gpu_ctx, codes):
    """Flow-relaxation data of several times (members 0, 2, 4; the others
    hold one time): the batched step derives each step's two levels and weight
    on the device from the member's time and gives each member's own values
    bit for bit, across level changes, through graph replays and after
    members took steps of their own."""
    print(f"time-varying flow relaxation, codes={codes}", flush=True)
    plain = _frs_members(gpu_ctx, codes)
    batched = _frs_members(gpu_ctx, codes)
    ok, why = BatchedEnsemble.eligible(batched)
    assert ok, why
    ens = BatchedEnsemble(batched)
    for s in plain:
        for _ in range(8):
            s.step(np.float32(s.dt))
    ens.step(8)
    _assert_same(batched, plain)
    for s in plain:
        for _ in range(8):
            s.step(np.float32(s.dt))
    ens.step(8, graph_steps=4)
    _assert_same(batched, plain)
    assert min(s.t for s in batched) > 45.0   # every level change crossed, and past the last level
    # a step of each member's own moves its time outside the batched path
    for a, b in zip(batched, plain):
        a.step(np.float32(a.dt))
        b.step(np.float32(b.dt))
    for s in plain:
        for _ in range(5):
            s.step(np.float32(s.dt))
    ens.step(5)
    _assert_same(batched, plain)
    ens.destroy()


def test_ocean_model_ensemble_coastal_batched_equals_sequential(gpu_ctx):
    print("OceanModelEnsemble coastal: batched vs sequential, with drifters", flush=True)
    sizes = SIZES[:4]
    data = [cases.config_c4_member(nx, ny, 60 + i) for i, (nx, ny) in enumerate(sizes)]
    me_args = dict(small_scale_perturbation_amplitude=5e-3, xorwow_seed=91)
    ea = OceanModelEnsemble(gpu_ctx, {}, data, me_args, len(sizes), batched=True)
    eb = OceanModelEnsemble(gpu_ctx, {}, data, me_args, len(sizes), batched=False)
    pos = cases.c4_drifter_positions(64, 64, 1000.0, 1000.0, 50, 7)
    ea.attachDrifters(pos)
    eb.attachDrifters(pos)
    for _ in range(3):
        assert ea.modelStep(300.0) == eb.modelStep(300.0)
    assert ea._ensemble is not None
    _assert_same(ea.particles, eb.particles)
    assert np.array_equal(ea.observeParticles(pos[:5]), eb.observeParticles(pos[:5]))


def test_autotuner_restores_coastal_state(gpu_ctx):
    print("autotuner on coastal members with drifters", flush=True)
    a = _members(gpu_ctx, (3, 3, 3, 1), True, 50)
    b = _members(gpu_ctx, (3, 3, 3, 1), True, 50)
    ens = BatchedEnsemble(a, autotune=True)
    _assert_same(a, b)   # tuning left no trace, drifters included
    ens.step(3)
    for s in b:
        for _ in range(3):
            s.step(np.float32(s.dt))
    _assert_same(a, b)

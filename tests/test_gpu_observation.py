"""The library's observation operator (gpo_observe, gpo_observe_underlying_flow)
against the oracle's restatements of the reference's observation loops, bit for
bit, over uneven members, and the ensemble and Observation/ParticleInfo paths
that use it."""
import ctypes

import numpy as np
import pytest
import torch

import cases
from gpuocean_amd import CDKLM16, GPUDrifterCollection, Observation, ObservationType, OceanModelEnsemble, ParticleInfo
from gpuocean_amd import _lib as L
from gpuocean_amd import observe
from oracle import observation as OO

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (96, 64), (48, 112)]


def _random_members(gpu_ctx, sizes=SIZES, seed=7):
    rng = np.random.default_rng(seed)
    sims = []
    for i, (nx, ny) in enumerate(sizes):
        kw = cases.config_c2_member(nx, ny, 100 + i)
        sim = CDKLM16(gpu_ctx, **kw, local_particle_id=i)
        shape = (ny + 4, nx + 4)
        eta = rng.normal(0.0, 0.3, shape).astype(np.float32)
        hu = rng.normal(0.0, 50.0, shape).astype(np.float32)
        hv = rng.normal(0.0, 50.0, shape).astype(np.float32)
        sim.upload(eta, hu, hv)
        sims.append(sim)
    return sims


def _states(sims):
    return [s.download() for s in sims]


def _positions(nx, ny, dx, dy, n, seed, edges=True):
    """Random interior points plus exact cell edges, points just below them
    and negative coordinates numpy wraps."""
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(0, nx * dx, n), rng.uniform(0, ny * dy, n)], axis=1)
    if edges:
        c = rng.integers(0, min(nx, ny), 16)
        e = np.stack([c * dx, c * dy], axis=1).astype(np.float64)
        below = np.nextafter(e, 0)
        neg = np.stack([-rng.uniform(0, nx * dx, 8), -rng.uniform(0, ny * dy, 8)], axis=1)
        p = np.concatenate([p, e, below, neg])
    return p


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True) and a.shape == b.shape


def test_flow_matches_host_loop(gpu_ctx):
    print("observe: FLOW over 3 uneven members", flush=True)
    sims = _random_members(gpu_ctx)
    nx = min(s.nx for s in sims)
    ny = min(s.ny for s in sims)
    pts = _positions(nx, ny, 1000.0, 1000.0, 300, 1)
    got = observe.observe_members(sims, observe.FLOW, positions=pts)
    want = OO.observe_particles(_states(sims), pts, 1000.0, 1000.0)
    assert _same(got, want)
    assert not np.any(np.isnan(got))
    # each member its own points
    per = [_positions(s.nx, s.ny, 1000.0, 1000.0, 200, 10 + k) for k, s in enumerate(sims)]
    got = observe.observe_members(sims, observe.FLOW, positions=per)
    for k, (s, st) in enumerate(zip(sims, _states(sims))):
        assert _same(got[k], OO.observe_particles([st], per[k], 1000.0, 1000.0)[0]), k


def test_points_outside_raise_index_error(gpu_ctx):
    sims = _random_members(gpu_ctx, sizes=[(32, 16)])
    for p in ([[32 * 1000.0, 10.0]], [[10.0, -17 * 1000.0]], [[np.nan, 0.0]]):
        with pytest.raises(IndexError):
            observe.observe_members(sims, observe.FLOW, positions=np.array(p))
    with pytest.raises(IndexError):
        observe.observe_members(sims, observe.STATE, cells=np.array([[0, 16]]))


def test_state_buoy_and_innovation_modes(gpu_ctx):
    print("observe: STATE, BUOY, INNOVATION", flush=True)
    sims = _random_members(gpu_ctx)
    states = _states(sims)
    rng = np.random.default_rng(3)
    cells = np.stack([rng.integers(-40, 48, 64), rng.integers(-40, 48, 64)], axis=1)
    got = observe.observe_members(sims, observe.STATE, cells=cells)
    for k, st in enumerate(states):
        assert _same(got[k], OO.state_samples(*st, cells)), k
    bcells = np.stack([rng.integers(0, 48, 40), rng.integers(0, 48, 40)], axis=1)
    got = observe.observe_members(sims, observe.BUOY, cells=bcells)
    for k, (s, st) in enumerate(zip(sims, states)):
        Bm = s.Bm.download()
        want = OO.buoy_readings(*st, Bm, bcells)
        assert np.array_equal(got[k].astype(np.float32), want) and np.array_equal(got[k], want.astype(np.float64)), k
    obs = np.concatenate([_positions(48, 48, 1000.0, 1000.0, 100, 4, edges=False),
                          rng.normal(0.0, 40.0, (100, 2))], axis=1)
    for md in (None, np.float32(251.7)):
        got = observe.observe_members(sims, observe.INNOVATION, observations=obs, mean_depth=md)
        want = OO.innovation(states, obs, 1000.0, 1000.0, mean_depth=md)
        assert _same(got, want), md


def test_reference_kernel_mirror(gpu_ctx):
    print("observe: observeUnderlyingFlow mirror and the ghosted mode", flush=True)
    sims = _random_members(gpu_ctx, sizes=[(64, 48)])
    s = sims[0]
    eta, hu, hv = s.download()
    rng = np.random.default_rng(9)
    pts = np.stack([rng.uniform(-2500.0, 66500.0, 500), rng.uniform(-2500.0, 50500.0, 500)], axis=1).astype(np.float32)
    pts[:8] = [[0.0, 0.0], [-2000.0, -2000.0], [65999.0, 49999.0], [66000.0, 0.0], [0.1, 0.1],
               [63999.996, 1.0], [1.0, 47999.996], [-2000.1, 0.0]]
    dev = torch.from_numpy(pts).cuda()
    out = torch.zeros((500, 2), dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(L.lib().gpo_observe_underlying_flow(s.nx, s.ny, 1000.0, 1000.0, 2, 2, ctypes.c_void_p(s.h0[0].data_ptr()),
                                                s.h0[0].pitch, ctypes.c_void_p(s.h0[1].data_ptr()), s.h0[1].pitch,
                                                ctypes.c_void_p(s.h0[2].data_ptr()), s.h0[2].pitch, 250.0, 500,
                                                ctypes.c_void_p(dev.data_ptr()), 8, ctypes.c_void_p(out.data_ptr()),
                                                8, stream), "observe_underlying_flow")
    torch.cuda.synchronize()
    want = OO.observe_underlying_flow(hu, hv, pts, 1000.0, 1000.0)
    got = out.cpu().numpy()
    assert _same(got, want)
    assert np.isnan(want).any() and (~np.isnan(want)).sum() > 400
    # the batched operator with the same cell rule
    got2 = observe.observe_members(sims, observe.FLOW | L.OBS_GHOSTED, positions=pts.astype(np.float64))
    assert np.array_equal(got2[0], want.astype(np.float64), equal_nan=True)


def test_ensemble_observation_paths(gpu_ctx):
    print("observe: OceanModelEnsemble.observeParticles / dumpParticleSample / forecasts", flush=True)
    kw = cases.config_c2_member(96, 64, 5)
    data_args = {k: kw[k] for k in ("eta0", "hu0", "hv0", "H", "nx", "ny", "dx", "dy", "boundary_conditions")}
    sim_args = {k: kw[k] for k in ("dt", "g", "f", "r", "courant_number")}
    ens = OceanModelEnsemble(gpu_ctx, sim_args, data_args, {"small_scale_perturbation_amplitude": 5e-3},
                             numParticles=4)
    drifters = _positions(96, 64, 1000.0, 1000.0, 20, 2, edges=False)
    ens.attachDrifters(drifters)
    ens.modelStep(600.0)
    states = _states(ens.particles)
    pts = _positions(96, 64, 1000.0, 1000.0, 50, 8)
    assert _same(ens.observeParticles(pts), OO.observe_particles(states, pts, 1000.0, 1000.0))
    cells = np.array([[3, 4], [95, 63], [-1, 0], [50, 20]])
    ens.dumpParticleSample(cells)
    for i, st in enumerate(states):
        assert np.array_equal(ens.particleInfos[i].get_state_samples(ens.particles[i].t), OO.state_samples(*st, cells))
    # members with their own extra cells take one launch each
    ens.particleInfos[1].setExtraCells(np.array([[7, 7]]))
    ens.modelStep(300.0)
    states = _states(ens.particles)
    ens.dumpParticleSample(cells)
    t = ens.particles[1].t
    assert np.array_equal(ens.particleInfos[1].get_extra_sample(t), OO.state_samples(*states[1], [[7, 7]]))
    assert np.array_equal(ens.particleInfos[2].get_state_samples(t), OO.state_samples(*states[2], cells))
    # drifter forecasts: a row at attach time, then one per dump
    ens.dumpForecastParticleSample()
    for i in range(4):
        f = ens.drifterForecast[i]
        assert f.get_num_observations() == 2
        assert np.array_equal(f.get_drifter_position(ens.particles[i].t), ens.particles[i].drifters.getDrifterPositions())
    ens.cleanUp()


def test_buoy_observations_from_sim(gpu_ctx):
    print("observe: Observation.add_observation_from_sim with buoys", flush=True)
    sim = _random_members(gpu_ctx, sizes=[(96, 64)])[0]
    drifters = GPUDrifterCollection(gpu_ctx, 3, boundaryConditions=sim.boundary_conditions,
                                    domain_size_x=96 * 1000.0, domain_size_y=64 * 1000.0)
    drifters.setDrifterPositions(np.array([[1000.0, 2000.0], [5000.0, 6000.0], [9000.0, 100.0]]))
    sim.attachDrifters(drifters)
    obs = Observation(ObservationType.StaticBuoys, domain_size_x=96 * 1000.0, domain_size_y=64 * 1000.0,
                      nx=96, ny=64, observation_variance=0.0)
    obs.setBuoyCellsByFrequency(10, 10, avoid_boundary=True)
    obs.add_observation_from_sim(sim)
    eta, hu, hv = sim.download()
    want = OO.buoy_readings(eta, hu, hv, sim.Bm.download(), obs.buoy_indices)
    row = obs.obs_df.iloc[0]
    assert row['buoy_observations'].dtype == np.float32 and np.array_equal(row['buoy_observations'], want)
    assert np.array_equal(row['drifter_positions'], sim.drifters.getDrifterPositions())
    # a sample of the same state through ParticleInfo
    pi = ParticleInfo()
    pi.add_state_sample_from_sim(sim, obs.buoy_indices)
    assert np.array_equal(pi.get_state_samples(sim.t), OO.state_samples(eta, hu, hv, obs.buoy_indices))

"""The benchmark configurations and the land/drying path, on the GPU, against
the FP32 oracle directly (not only against another HIP path):

* a coastal member with masked land, a masked island, desingularised and
  drying cells, 240 steps, exact build, bit for bit;
* C1 at its full size (2048 x 1024 double jet), 120 steps, bit for bit, with
  the FP32-vs-FP64 difference stated;
* four C2-sized members (up to 512 x 384) through the batched ensemble with
  SOAR model error, each step against the oracle step + SOAR + balance on the
  same random numbers;
* two C4 members (land, 10-cell flow-relaxation sponges, SOAR, 1000 drifters
  each) through the batched ensemble against the oracle state and drift.
"""
import ctypes

import numpy as np
import pytest

import cases
from gpuocean_amd import BatchedEnsemble, CDKLM16, GPUDrifterCollection, host
from gpuocean_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_coast_island_and_drying_shelf_240_steps(gpu_ctx):
    print("coastal member with island and drying shelf, 240 steps", flush=True)
    kw = cases.config_coast_island()
    sim = CDKLM16(gpu_ctx, **kw)
    ora = O.OracleCDKLM16(**kw)
    assert np.float32(sim.dt) == np.float32(ora.dt)
    dt = np.float32(ora.dt)
    land = np.abs(ora.Bm - host.LAND_VALUE) < 1.0
    desing = dry = 0
    for chunk in range(8):
        sim.step(30 * np.float64(dt))
        ora.run_steps(30, dt)
        for k, (a, b) in enumerate(zip(sim.download(), ora.download())):
            assert np.array_equal(a, b), f"after {30 * (chunk + 1)} steps, field {k}"
        h = ora.h0[0] + ora.Bm
        desing = max(desing, int(np.sum(~land & (h < 0.1) & (h > 2e-5))))
        dry = max(dry, int(np.sum(~land & (h <= 2e-5))))
    print(f"land cells {int(land.sum())}, desingularised up to {desing}, dry up to {dry}", flush=True)
    assert land.sum() > 1000 and desing > 50 and dry > 0


def test_c1_full_size_120_steps(gpu_ctx):
    print("C1 2048x1024, 120 steps against the FP32 oracle", flush=True)
    kw = cases.config_c1()
    sim = CDKLM16(gpu_ctx, **kw)
    ora = O.OracleCDKLM16(**kw)
    assert np.float32(sim.dt) == np.float32(ora.dt)
    dt = np.float32(ora.dt)
    n = 120
    sim.step(n * np.float64(dt))
    ora.run_steps(n, dt)
    f64 = O.OracleCDKLM16(**kw, precision="f64")
    f64.run_steps(n, dt)
    for name, a, b, c in zip(("eta", "hu", "hv"), sim.download(True), ora.download(True), f64.download(True)):
        print(f"C1 {name}: rel L2 vs FP32 oracle {cases.rel_l2(a, b):.3e}, vs FP64 oracle {cases.rel_l2(a, c):.3e}",
              flush=True)
        assert np.array_equal(a, b)


def _oracle_noise(sim, ora):
    """The model error of the step sim just took, added to the oracle state:
    SOAR of the member's latest normal field and its geostrophic balance."""
    me = sim.model_error
    xi = me.getRandomNumbers()
    bc = sim.boundary_conditions
    coarse = O.soar(int(sim.nx), int(sim.ny), me.weights(), 1.0, bc.isPeriodicNorthSouth(), bc.isPeriodicEastWest(),
                    xi)

    def perturb(h):
        O.geostrophic_balance(int(sim.nx), int(sim.ny), sim.dx, sim.dy, sim.g, sim.f, 0.0, 0.0, 1.0, None, coarse,
                              h[0], h[1], h[2], ora.Hi)
    return perturb


def test_c2_sized_members_batched_against_oracle(gpu_ctx):
    sizes = [(512, 384), (128, 128), (320, 256), (496, 144)]
    print(f"C2-sized members {sizes}, batched with SOAR, 20 steps against the oracle", flush=True)
    sims, oras = [], []
    for i, (nx, ny) in enumerate(sizes):
        kw = cases.config_c2_member(nx, ny, 2024 + 1 + i)
        s = CDKLM16(gpu_ctx, **kw, local_particle_id=i)
        s.setSOARModelError(5e-3, xorwow_seed=1_000_003 + i)
        sims.append(s)
        oras.append(O.OracleCDKLM16(**kw))
    ens = BatchedEnsemble(sims)
    for step in range(20):
        ens.step(1)
        for i, (s, o) in enumerate(zip(sims, oras)):
            o.rk2_step(np.float32(s.dt), _oracle_noise(s, o))
            for k, (a, b) in enumerate(zip(s.download(), o.download())):
                assert np.array_equal(a, b), f"step {step}, member {i} {sizes[i]}, field {k}"
    ens.destroy()


def test_c4_members_batched_against_oracle(gpu_ctx):
    sizes = [(256, 256), (384, 320)]
    print(f"C4 members {sizes}: land, 10-cell sponges, SOAR, 1000 drifters, batched, 12 steps", flush=True)
    sims, oras, pos = [], [], []
    for i, (nx, ny) in enumerate(sizes):
        kw = cases.config_c4_member(nx, ny, 4048 + i)
        s = CDKLM16(gpu_ctx, **kw, local_particle_id=i)
        s.setSOARModelError(5e-3, xorwow_seed=77 + i)
        p = cases.c4_drifter_positions(nx, ny, s.dx, s.dy, 1001, 4048 + i)
        d = GPUDrifterCollection(gpu_ctx, 1000, boundaryConditions=s.boundary_conditions, domain_size_x=nx * s.dx,
                                 domain_size_y=ny * s.dy, gpu_stream=s.gpu_stream)
        d.setDrifterPositions(p[:-1])
        d.setObservationPosition(p[-1])
        s.attachDrifters(d)
        sims.append(s)
        oras.append(O.OracleCDKLM16(**kw))
        pos.append(p.copy())
    ens = BatchedEnsemble(sims)
    for step in range(12):
        ens.step(1)
        for i, (s, o) in enumerate(zip(sims, oras)):
            dt = np.float32(s.dt)
            o.rk2_step(dt, _oracle_noise(s, o))
            e, u, v = o.download()
            O.drift(int(s.nx), int(s.ny), s.dx, s.dy, dt, e, u, v, o.Bm.astype(np.float32), False, False, pos[i])
            for k, (a, b) in enumerate(zip(s.download(), o.download())):
                assert np.array_equal(a, b), f"step {step}, member {i}, field {k}"
            got = np.concatenate([s.drifters.getDrifterPositions(), [s.drifters.getObservationPosition()]], axis=0)
            assert np.array_equal(got, pos[i]), f"step {step}, member {i}: drifters"
    land = np.abs(oras[0].Bm - host.LAND_VALUE) < 1.0
    assert land.sum() > 0
    ens.destroy()


def test_ghost_parts_after_set_times(gpu_ctx):
    """Ghost-cell parts launched right after the members' times are set (and
    right after creation) read valid boundary-data levels: the flow-relaxation
    sponges they write equal the members' own ghost-cell pass."""
    from gpuocean_amd import Common
    sims, ref = [], []
    for i, (nx, ny) in enumerate([(96, 80), (64, 72)]):
        kw = cases.config_c4_member(nx, ny, 31 + i)
        data = kw["boundary_conditions_data"]
        # boundary data at two times, so that the device clocks hold several levels
        two = Common.BoundaryConditionsData(
            t=[0.0, 3600.0],
            north=Common.SingleBoundaryConditionData(h=[data.north.h[0], data.north.h[0] + 0.1],
                                                     hu=[data.north.hu[0]] * 2, hv=[data.north.hv[0]] * 2),
            south=Common.SingleBoundaryConditionData(h=[data.south.h[0], data.south.h[0] - 0.1],
                                                     hu=[data.south.hu[0]] * 2, hv=[data.south.hv[0]] * 2),
            east=Common.SingleBoundaryConditionData(h=[data.east.h[0], data.east.h[0] + 0.05],
                                                    hu=[data.east.hu[0]] * 2, hv=[data.east.hv[0]] * 2),
            west=Common.SingleBoundaryConditionData(h=[data.west.h[0]] * 2, hu=[data.west.hu[0]] * 2,
                                                    hv=[data.west.hv[0]] * 2))
        kw["boundary_conditions_data"] = two
        kw["t"] = 900.0 * (i + 1)
        sims.append(CDKLM16(gpu_ctx, **kw))
        ref.append(CDKLM16(gpu_ctx, **kw))
    ens = BatchedEnsemble(sims)
    times = (ctypes.c_double * 2)(1200.0, 2500.0)
    for s, t in zip(sims + ref, [1200.0, 2500.0, 1200.0, 2500.0]):
        s.t = t
    L.check(L.lib().gpo_ensemble_set_times(ens.handle, times, ens._sh()), "gpo_ensemble_set_times")
    L.check(L.lib().gpo_ensemble_part(ens.handle, 3, ens._sh()), "gpo_ensemble_part")
    ens.stream.synchronize()
    for r in ref:
        r._update_bc_values()
        r._apply_bc(r.h0)
    for i, (a, b) in enumerate(zip(sims, ref)):
        for k, (x, y) in enumerate(zip(a.download(), b.download())):
            assert np.array_equal(x, y), f"member {i}, field {k}"
    ens.destroy()


@pytest.mark.parametrize("case", ["coast_island", "c4_member", "periodic_island"])
def test_bathymetry_init_matches_oracle(gpu_ctx, case):
    """gpo_bathymetry_init: the device Hi (land corners, ghost corners) and Bm
    (means of the wet corners, land_value for all-land cells) equal the
    oracle's, bit for bit."""
    from gpuocean_amd import Common
    if case == "coast_island":
        kw = cases.config_coast_island()
    elif case == "c4_member":
        kw = cases.config_c4_member(96, 80, 5)
    else:
        kw = cases.config_coast_island(nx=64, ny=48)
        kw["boundary_conditions"] = Common.BoundaryConditions(2, 2, 2, 2)
    print(f"bathymetry set-up {case}", flush=True)
    sim = CDKLM16(gpu_ctx, **kw)
    ora = O.OracleCDKLM16(**kw)
    Hi, Bm = sim.downloadBathymetry()
    assert np.array_equal(np.ma.getdata(Hi), ora.Hi)
    assert np.array_equal(np.ma.getdata(Bm), ora.Bm)
    assert np.any(ora.Bm == np.float32(host.LAND_VALUE))


def _oracle_da_step(ora, obs_time, mts, courant, perturb, total):
    """dataAssimilationStep's cadence (SWEsimulators/CDKLM16.py:742-806) on the
    oracle: a leftover step first if there is one, then full model time
    steps, each followed by the model error (perturb(scale)), dt from the CFL
    condition at the start and after every fifth model step.  Returns the
    new total_time_steps."""
    full = int(round(obs_time - ora.t) / mts)
    leftover = obs_time - ora.t - full * mts
    if leftover / mts < 0.1 and full > 1:
        leftover += mts
        full -= 1
    if leftover / mts < 0.00001:
        leftover = 0
    ora.updateDt(courant)
    for i in range(full + 1):
        if i == 0 and leftover == 0:
            continue
        ora.step(leftover if i == 0 else mts)
        perturb(np.sqrt(leftover / mts) if i == 0 else 1.0)
        total += 1
        if total % 5 == 0:
            ora.updateDt(courant)
    return total


def test_step_to_observation_against_oracle(gpu_ctx):
    """Data-assimilation stepping (OceanModelEnsemble.stepToObservation, the
    batched path: members with different time steps in shared launches) against
    the oracle: each member's state after every observation time equals the
    oracle's run of dataAssimilationStep's cadence bit for bit, with the
    member's own normal field of each model step (model time step 60 s, six
    observations one model step apart, so dt is also updated after the fifth
    model step), and its dt equals the oracle's."""
    from gpuocean_amd import OceanModelEnsemble
    sizes = [(192, 128), (128, 96), (160, 64)]
    print(f"stepToObservation of {sizes} against the oracle's dataAssimilationStep cadence", flush=True)
    data, oras = [], []
    for i, (nx, ny) in enumerate(sizes):
        kw = cases.config_c2_member(nx, ny, 300 + i)
        data.append({k: kw[k] for k in ("eta0", "hu0", "hv0", "H", "nx", "ny", "dx", "dy", "boundary_conditions")})
        oras.append((kw, O.OracleCDKLM16(**kw)))
    kw0 = oras[0][0]
    sim_args = {k: kw0[k] for k in ("dt", "g", "f", "r", "courant_number")}
    sim_args["model_time_step"] = 60.0
    me = {"small_scale_perturbation_amplitude": 5e-3, "xorwow_seed": 91}
    assert all(kw["f"] == kw0["f"] and kw["g"] == kw0["g"] for kw, _ in oras)
    ens = OceanModelEnsemble(gpu_ctx, sim_args, data, me, numParticles=len(sizes), batched=True)
    totals = [0] * len(sizes)
    for t_obs in (60.0, 120.0, 180.0, 240.0, 300.0, 360.0):
        ens.stepToObservation(t_obs)
        for i, (s, (kw, o)) in enumerate(zip(ens.particles, oras)):
            perturb = _oracle_noise(s, o)
            # dataAssimilationStep's own courant_number default, 0.8; the
            # ghost cells follow the perturbed interior (the step kernel
            # reads wall and periodic neighbours from the interior cells)
            def model_error(scale):
                perturb(o.h0)
                o.apply_bc()
            totals[i] = _oracle_da_step(o, t_obs, 60.0, 0.8, model_error, totals[i])
            assert s.total_time_steps == totals[i]
            assert np.float32(s.dt) == np.float32(o.dt), f"t={t_obs}, member {i}: dt"
            for k, (a, b) in enumerate(zip(s.download(True), o.download(True))):
                assert np.array_equal(a, b), f"t={t_obs}, member {i} {sizes[i]}, field {k}"
    assert ens._ensemble is not None   # the batched path ran
    assert len({round(float(s.dt), 3) for s in ens.particles}) > 1
    ens.cleanUp()

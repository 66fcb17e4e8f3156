"""OceanModelEnsemble.stepToObservation run batched (members with different
time steps sharing launches, the model error once per model time step) gives
every member exactly what its own CDKLM16.dataAssimilationStep gives
(SWEsimulators/CDKLM16.py:742-806): states with ghost cells, time, iteration
and draw counters, time steps and drifter positions."""
import numpy as np
import pytest

import cases
from gpuocean_amd import OceanModelEnsemble

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (96, 64), (48, 80), (128, 32)]
SIZES_3 = [(66, 48), (96, 64), (48, 80), (129, 32)]   # periodic x: nx divisible by 3


def _ensemble(gpu_ctx, batched, drifters=False, noise=True, factor=1):
    data = []
    for i, (nx, ny) in enumerate(SIZES if factor == 1 else SIZES_3):
        kw = cases.config_c2_member(nx, ny, 40 + i)
        data.append({k: kw[k] for k in ("eta0", "hu0", "hv0", "H", "nx", "ny", "dx", "dy", "boundary_conditions")})
    kw = cases.config_c2_member(*SIZES[0], 40)
    sim_args = {k: kw[k] for k in ("dt", "g", "f", "r", "courant_number")}
    sim_args["model_time_step"] = 60.0
    me = {"small_scale_perturbation_amplitude": 5e-3, "xorwow_seed": 77} if noise else {}
    if factor > 1:
        me["small_scale_perturbation_interpolation_factor"] = factor
    ens = OceanModelEnsemble(gpu_ctx, sim_args, data, me, numParticles=len(SIZES), batched=batched)
    if drifters:
        rng = np.random.default_rng(5)
        pos = np.stack([rng.uniform(0, 48e3, 30), rng.uniform(0, 32e3, 30)], axis=1)
        ens.attachDrifters(pos)
    return ens


def _same(a, b, label):
    for i, (x, y) in enumerate(zip(a.particles, b.particles)):
        for k, (u, v) in enumerate(zip(x.download(), y.download())):
            assert np.array_equal(u, v), f"{label}: member {i} field {k}"
        assert x.t == y.t and x.num_iterations == y.num_iterations, f"{label}: member {i} time"
        assert x.total_time_steps == y.total_time_steps and np.float32(x.dt) == np.float32(y.dt), f"{label}: member {i}"
        if x.model_error is not None:
            assert x.model_error.draw == y.model_error.draw, f"{label}: member {i} draw"
        if x.hasDrifters:
            assert np.array_equal(x.drifters.getDrifterPositions(), y.drifters.getDrifterPositions()), \
                f"{label}: member {i} drifters"
            assert x.drifter_t == y.drifter_t


@pytest.mark.parametrize("drifters,noise,factor", [(False, True, 1), (True, True, 1), (False, False, 1),
                                                   (False, True, 3)])
def test_batched_step_to_observation_matches_members(gpu_ctx, drifters, noise, factor):
    print(f"stepToObservation batched vs per member, drifters={drifters}, noise={noise}, factor={factor}",
          flush=True)
    a = _ensemble(gpu_ctx, True, drifters, noise, factor)
    b = _ensemble(gpu_ctx, False, drifters, noise, factor)
    assert len({round(float(s.dt), 3) for s in a.particles}) > 1   # members step with different dt
    # full model steps only, a leftover step first, and the short-leftover merge
    for t_obs in (300.0, 610.0, 1215.0, 1500.0):
        a.stepToObservation(t_obs)
        b.stepToObservation(t_obs)
        _same(a, b, f"t={t_obs}")
    assert a._ensemble is not None   # the batched path ran
    # an inactive member is left alone
    a.particlesActive[2] = b.particlesActive[2] = False
    a.stepToObservation(1800.0)
    b.stepToObservation(1800.0)
    _same(a, b, "one member inactive")
    assert a.particles[2].t == 1500.0
    # and the ensemble still steps with modelStep afterwards
    a.particlesActive[2] = b.particlesActive[2] = True
    a.modelStep(120.0)
    b.modelStep(120.0)
    _same(a, b, "modelStep after")
    a.cleanUp()
    b.cleanUp()


def _coastal_ensemble(gpu_ctx, batched):
    data = []
    for i, (nx, ny) in enumerate([(64, 48), (96, 64), (80, 56)]):
        kw = cases.config_c4_member(nx, ny, 70 + i)
        if i == 1:   # boundary data at three times for one member
            d = kw["boundary_conditions_data"]
            for side in (d.north, d.south, d.east, d.west):
                side.h = [np.float32(f) * side.h[0] for f in (1.0, 1.2, 0.8)]
                side.hu = [np.float32(f) * side.hu[0] for f in (1.0, 1.2, 0.8)]
                side.hv = [np.float32(f) * side.hv[0] for f in (1.0, 1.2, 0.8)]
                side.numSteps = 3
            d.t, d.numSteps = [0.0, 200.0, 500.0], 3
        data.append({k: kw[k] for k in ("eta0", "hu0", "hv0", "H", "nx", "ny", "dx", "dy", "boundary_conditions",
                                        "boundary_conditions_data")})
    sim_args = {"dt": 0.0, "g": 9.81, "f": 1e-4, "r": 0.0, "courant_number": 0.7, "model_time_step": 60.0}
    me = {"small_scale_perturbation_amplitude": 5e-3, "xorwow_seed": 55}
    return OceanModelEnsemble(gpu_ctx, sim_args, data, me, numParticles=3, batched=batched)


def test_batched_step_to_observation_coastal(gpu_ctx):
    """Coastal members (land, flow relaxation, one of them with boundary data
    at three times) through the batched data-assimilation stepping: the device
    boundary-data clocks follow each member's own dataAssimilationStep."""
    print("stepToObservation batched vs per member, coastal", flush=True)
    a = _coastal_ensemble(gpu_ctx, True)
    b = _coastal_ensemble(gpu_ctx, False)
    calls = []
    batched_step = a._batched_da_step
    a._batched_da_step = lambda *args, **kw: (calls.append(1), batched_step(*args, **kw))[1]
    for t_obs in (300.0, 610.0):
        a.stepToObservation(t_obs)
        b.stepToObservation(t_obs)
        _same(a, b, f"coastal t={t_obs}")
    assert a._ensemble is not None and len(calls) == 2   # the batched data-assimilation path ran
    a.cleanUp()
    b.cleanUp()

"""GPU stochastic model error and drifters against the oracle.

The random field is generated on the GPU (Philox4x32-10, pinned by published
known-answer vectors); SOAR, geostrophic balance and drift are then checked
value for value against the oracle given the same random numbers.
"""
import ctypes

import numpy as np
import pytest
import torch

import cases
from gpuocean_amd import CDKLM16, Common, GPUDrifterCollection, OceanStateNoise, WindStress
from gpuocean_amd import _lib as L
from oracle import oracle as O
from test_noise_drift_cpu import PHILOX_KAT, WIND_DRIFT_REF

pytestmark = pytest.mark.gpu


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_philox_known_answers_on_gpu(gpu_ctx):
    ctr = torch.tensor([list(c) for c, _, _ in PHILOX_KAT], dtype=torch.int64).to(torch.uint32)
    key = torch.tensor([list(k) for _, k, _ in PHILOX_KAT], dtype=torch.int64).to(torch.uint32)
    ctr, key = ctr.cuda(), key.cuda()
    out = torch.zeros((len(PHILOX_KAT), 4), dtype=torch.uint32, device="cuda")
    L.check(L.lib().gpo_philox4x32_10(L.ptr(ctr), L.ptr(key), L.ptr(out), len(PHILOX_KAT), _stream()), "philox")
    torch.cuda.synchronize()
    got = out.cpu().to(torch.int64).numpy()
    for row, (_, _, want) in zip(got, PHILOX_KAT):
        assert tuple(int(x) for x in row) == want


def _normals(rows, cols, seed, stream_id, draw):
    out = torch.empty((rows, cols), dtype=torch.float32, device="cuda")
    L.check(L.lib().gpo_random_normal(L.ptr(out), rows, cols, 4 * cols, seed, stream_id, draw, _stream()), "normal")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_random_normal_statistics_and_reproducibility(gpu_ctx):
    a = _normals(1024, 2048, 1234, 7, 0)
    n = a.size
    assert abs(a.mean()) < 5.0 / np.sqrt(n)
    assert abs(a.var() - 1.0) < 5.0 * np.sqrt(2.0 / n)
    assert abs(np.mean(a ** 3)) < 5.0 * np.sqrt(15.0 / n)          # skewness
    assert abs(np.mean(a ** 4) - 3.0) < 5.0 * np.sqrt(96.0 / n)     # kurtosis
    assert np.array_equal(a, _normals(1024, 2048, 1234, 7, 0))       # seeded reproducibility
    assert not np.array_equal(a, _normals(1024, 2048, 1234, 7, 1))   # next draw differs
    assert not np.array_equal(a, _normals(1024, 2048, 1234, 8, 0))   # other stream differs
    # against the generator restated on the host (FP64 Box-Muller), over the
    # first 2048 elements: the hardware log2/sin/cos keep every normal within
    # 1e-5 of it
    flat = a.reshape(-1)[:2048]
    for idx in range(2048):
        words = O.philox4x32_10((idx >> 2, 0, 7, 0x534F4152), (1234, 0))
        x0, x1 = (words[0], words[1]) if (idx & 2) == 0 else (words[2], words[3])
        u1 = (x0 >> 8) * 2.0 ** -24 + 2.0 ** -25
        u2 = (x1 >> 8) * 2.0 ** -24 + 2.0 ** -25
        r = np.sqrt(-2.0 * np.log(u1))
        z = r * (np.sin(2 * np.pi * u2) if idx & 1 else np.cos(2 * np.pi * u2))
        assert abs(flat[idx] - z) < 1e-5 * max(1.0, abs(z))


@pytest.mark.parametrize("periodic", [False, True])
def test_soar_and_balance_match_oracle(gpu_ctx, periodic):
    nx, ny = 70, 45
    bc = Common.BoundaryConditions(2, 2, 2, 2) if periodic else Common.BoundaryConditions()
    rng = np.random.default_rng(3)
    eta = np.zeros((ny + 4, nx + 4), np.float32)
    hu = (rng.standard_normal(eta.shape) * 0.1).astype(np.float32)
    hv = (rng.standard_normal(eta.shape) * 0.1).astype(np.float32)
    H = (80 + 10 * rng.random((ny + 5, nx + 5))).astype(np.float32)
    sim = CDKLM16(gpu_ctx, eta, hu, hv, H, nx, ny, 500.0, 500.0, 1.0, 9.81, 1e-4, 0.0, boundary_conditions=bc)
    sim.setSOARModelError(5e-3, xorwow_seed=42)
    noise = sim.model_error
    Hi_dev, Bm = sim.downloadBathymetry()
    e0, u0, v0 = sim.download()
    sim.perturbState()
    xi = noise.getRandomNumbers()
    W = noise.weights()
    coarse = O.soar(nx, ny, W, 1.0, periodic, periodic, xi)
    assert np.array_equal(noise.getCoarseBuffer(), coarse)
    e, u, v = e0.copy(), u0.copy(), v0.copy()
    O.geostrophic_balance(nx, ny, np.float32(500), np.float32(500), np.float32(9.81), np.float32(1e-4), 0.0, 0.0,
                          1.0, None, coarse, e, u, v, np.asarray(Hi_dev, np.float32))
    ge, gu, gv = sim.download()
    inner = (slice(2, ny + 2), slice(2, nx + 2))
    for a, b in zip((ge, gu, gv), (e, u, v)):
        assert np.array_equal(a[inner], b[inner])


def test_stochastic_steps_match_oracle(gpu_ctx):
    """CDKLM16 steps with SOAR model error (q0=5e-3, L=0.75 dx) against the
    oracle step + SOAR + balance on the same random numbers."""
    nx, ny = 96, 64
    kw = cases.config_c1(nx=nx, ny=ny, seed=3)
    sim = CDKLM16(gpu_ctx, **kw)
    sim.setSOARModelError(5e-3, xorwow_seed=7)
    ora = O.OracleCDKLM16(**kw)
    dt = np.float32(sim.dt)
    assert dt == np.float32(ora.dt)
    gen = OceanStateNoise.fromsim(sim, soar_q0=5e-3, xorwow_seed=7)
    W = gen.weights()
    bc = kw["boundary_conditions"]
    for step in range(8):
        sim.step(dt)
        ora.run_steps(1, dt)
        gen.generateNormalDistribution()
        xi = gen.getRandomNumbers()
        coarse = O.soar(nx, ny, W, 1.0, bc.isPeriodicNorthSouth(), bc.isPeriodicEastWest(), xi)
        e, u, v = ora.h0
        O.geostrophic_balance(nx, ny, sim.dx, sim.dy, sim.g, sim.f, 0.0, 0.0, 1.0, None, coarse, e, u, v, ora.Hi)
        getattr(O.lib(), "gpo_oracle_apply_bc_f32")(ctypes.byref(ora.q), *[O._p(a) for a in ora.h0], None)
        for a, b in zip(sim.download(), ora.download()):
            assert np.array_equal(a, b), f"step {step}"


def test_wind_drift_known_answer_on_gpu(gpu_ctx):
    nx = ny = 100
    dx = dy = 100.0
    z = np.zeros((ny + 4, nx + 4), np.float32)
    Hi = 10 * np.ones((ny + 5, nx + 5), np.float32)
    wind = WindStress(t=np.array([0]), wind_u=np.float32([np.array([[30]])]), wind_v=np.float32([np.array([[0]])]))
    sim = CDKLM16(gpu_ctx, z, z, z, Hi, nx, ny, dx, dy, 0.0, 9.81, 0.0, 0.0,
                  boundary_conditions=Common.BoundaryConditions(2, 2, 2, 2), wind=wind)
    sets = []
    for i in range(12):
        w = WindStress(t=np.array([0]), wind_u=np.float32([np.array([[10 * np.sin(2 * np.pi / 12 * i)]])]),
                       wind_v=np.float32([np.array([[10 * np.cos(2 * np.pi / 12 * i)]])]))
        d = GPUDrifterCollection(gpu_ctx, 1, wind=w, wind_drift_factor=0.02, boundaryConditions=sim.boundary_conditions,
                                 domain_size_x=nx * dx, domain_size_y=ny * dy, gpu_stream=sim.gpu_stream)
        d.setDrifterPositions([[int(0.5 * ny * dy), int(0.5 * ny * dy)]])
        sets.append(d)
    for _ in range(600):
        sim.step(1)
        sim.applyBoundaryConditions()
        for d in sets:
            d.drift(sim.h0[0], sim.h0[1], sim.h0[2], sim.Bm, sim.nx, sim.ny, sim.t, sim.dx, sim.dy, 1, 2, 2)
    got = np.array([list(d.getDrifterPositions()[0]) for d in sets])
    assert np.max(np.abs(got - np.array(WIND_DRIFT_REF))) < 5e-5


@pytest.mark.parametrize("periodic", [False, True])
def test_drift_matches_oracle(gpu_ctx, periodic):
    nx, ny = 90, 70
    dx = dy = 250.0
    rng = np.random.default_rng(9)
    eta = (0.1 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    hu = (5 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    hv = (5 * rng.standard_normal((ny + 4, nx + 4))).astype(np.float32)
    H = (50 + 5 * rng.random((ny + 5, nx + 5))).astype(np.float32)
    bc = Common.BoundaryConditions(2, 2, 2, 2) if periodic else Common.BoundaryConditions()
    sim = CDKLM16(gpu_ctx, eta, hu, hv, H, nx, ny, dx, dy, 1.0, 9.81, 0.0, 0.0, boundary_conditions=bc)
    n = 400
    margin = 0.0 if periodic else 1.5 * dx
    pos = np.stack([rng.uniform(margin, nx * dx - margin, n + 1), rng.uniform(margin, ny * dy - margin, n + 1)],
                   axis=1).astype(np.float32)
    d = GPUDrifterCollection(gpu_ctx, n, boundaryConditions=bc, domain_size_x=nx * dx, domain_size_y=ny * dy,
                             gpu_stream=sim.gpu_stream)
    d.setDrifterPositions(pos[:-1])
    d.setObservationPosition(pos[-1])
    e, u, v = sim.download()
    Bm = np.asarray(sim.downloadBathymetry()[1], np.float32)
    ref = pos.copy()
    for _ in range(20):
        d.driftFromSim(sim, 30.0)
        O.drift(nx, ny, dx, dy, 30.0, e, u, v, Bm, periodic, periodic, ref)
    got = np.concatenate([d.getDrifterPositions(), [d.getObservationPosition()]], axis=0)
    assert np.array_equal(got, ref)


def test_drift_outside_the_arrays_matches_oracle(gpu_ctx):
    """Drifters just outside an open side: the GPU leaves those whose stencil
    leaves the arrays in place, as the oracle does, and moves the rest
    identically."""
    from test_noise_drift_cpu import _drift_guard_case
    nx, ny, dx, dy, eta, hu, hv, Hm, pos, inside = _drift_guard_case()
    Hi = np.full((ny + 5, nx + 5), 10.0, np.float32)
    sim = CDKLM16(gpu_ctx, eta, hu, hv, Hi, nx, ny, dx, dy, 1.0, 9.81, 0.0, 0.0)
    bc = Common.BoundaryConditions()
    d = GPUDrifterCollection(gpu_ctx, len(pos) - 1, boundaryConditions=bc, domain_size_x=nx * dx,
                             domain_size_y=ny * dy, gpu_stream=sim.gpu_stream)
    d.setDrifterPositions(pos[:-1])
    d.setObservationPosition(pos[-1])
    Bm = np.asarray(sim.downloadBathymetry()[1], np.float32)
    e, u, v = sim.download()
    ref = pos.copy()
    for _ in range(3):
        d.driftFromSim(sim, 10.0)
        O.drift(nx, ny, dx, dy, 10.0, e, u, v, Bm, False, False, ref)
    got = np.concatenate([d.getDrifterPositions(), [d.getObservationPosition()]], axis=0)
    assert np.array_equal(got, ref, equal_nan=True)
    assert not np.array_equal(got[inside], pos[inside])

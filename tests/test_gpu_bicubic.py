"""Model error generated on a coarse grid (interpolation_factor 3, 5, 7):
gpo_bicubic_interpolation and the OceanStateNoise / CDKLM16 paths that use it
against the oracle, bit for bit, on the GPU's own random numbers."""
import ctypes

import numpy as np
import pytest

import cases
from gpuocean_amd import CDKLM16, Common, OceanStateNoise
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sim(gpu_ctx, nx, ny, periodic, factor, land=False, beta=0.0, seed=3):
    bc = Common.BoundaryConditions(2, 2, 2, 2) if periodic else Common.BoundaryConditions()
    rng = np.random.default_rng(seed)
    eta = (rng.standard_normal((ny + 4, nx + 4)) * 0.05).astype(np.float32)
    hu = (rng.standard_normal(eta.shape) * 0.1).astype(np.float32)
    hv = (rng.standard_normal(eta.shape) * 0.1).astype(np.float32)
    H = (80 + 10 * rng.random((ny + 5, nx + 5))).astype(np.float32)
    if land:
        mask = np.zeros(H.shape, bool)
        mask[10:20, 5:12] = True
        H = np.ma.array(H, mask=mask)
        eta[11:19, 6:11] = 0.0
        hu[11:19, 6:11] = 0.0
        hv[11:19, 6:11] = 0.0
    sim = CDKLM16(gpu_ctx, eta, hu, hv, H, nx, ny, 500.0, 400.0, 1.0, 9.81, 1e-4, 0.0, boundary_conditions=bc,
                  coriolis_beta=beta)
    sim.setSOARModelError(5e-3, small_scale_perturbation_interpolation_factor=factor, xorwow_seed=42 + factor)
    return sim


def _oracle_perturb(sim, e, u, v, oi=0, oj=0):
    noise = sim.model_error
    xi = noise.getRandomNumbers()
    per_ns, per_ew = bool(noise.periodicNorthSouth), bool(noise.periodicEastWest)
    coarse = O.soar(int(noise.coarse_nx), int(noise.coarse_ny), noise.weights(), 1.0, per_ns, per_ew, xi)
    assert np.array_equal(noise.getCoarseBuffer(), coarse)
    Hi = np.asarray(np.ma.getdata(sim.downloadBathymetry()[0]), np.float32)
    p = sim.params
    O.bicubic_interpolation(int(sim.nx), int(sim.ny), sim.dx, sim.dy, int(noise.coarse_nx), int(noise.coarse_ny),
                            noise.coarse_dx, noise.coarse_dy, oi, oj, sim.g, sim.f, np.float32(sim.coriolis_beta),
                            p.north_x, p.north_y, None, coarse, e, u, v, Hi, land_value=np.float32(sim.mask_value))


@pytest.mark.parametrize("factor,nx,ny,periodic", [(3, 70, 44, False), (3, 75, 45, True), (5, 70, 45, False),
                                                   (5, 75, 45, True), (7, 77, 49, True), (7, 130, 60, False)])
def test_perturbation_matches_oracle(gpu_ctx, factor, nx, ny, periodic):
    print(f"bicubic: factor {factor}, {nx}x{ny}, periodic={periodic}", flush=True)
    sim = _sim(gpu_ctx, nx, ny, periodic, factor, beta=1e-11)
    e, u, v = sim.download()
    sim.perturbState()
    _oracle_perturb(sim, e, u, v)
    inner = (slice(2, ny + 2), slice(2, nx + 2))
    for k, (a, b) in enumerate(zip(sim.download(), (e, u, v))):
        assert np.array_equal(a[inner], b[inner]), k


@pytest.mark.parametrize("align", [(6, 6), (7, 8), (8, 6), (0, 9)])
def test_aligned_coarse_grid_matches_oracle(gpu_ctx, align):
    print(f"bicubic: aligned with cell {align}", flush=True)
    sim = _sim(gpu_ctx, 75, 45, True, 5)
    noise = sim.model_error
    e, u, v = sim.download()
    noise.perturbSim(sim, align_with_cell_i=align[0], align_with_cell_j=align[1])
    oi, oj = noise._obtain_coarse_grid_offset(*align)
    assert (oi, oj) == (-(align[0] % 5 - 2), -(align[1] % 5 - 2))
    _oracle_perturb(sim, e, u, v, oi, oj)
    inner = (slice(2, 47), slice(2, 77))
    for k, (a, b) in enumerate(zip(sim.download(), (e, u, v))):
        assert np.array_equal(a[inner], b[inner]), k


def test_land_cells_keep_their_state(gpu_ctx):
    print("bicubic: land", flush=True)
    sim = _sim(gpu_ctx, 70, 45, False, 3, land=True)
    e, u, v = sim.download()
    sim.perturbState()
    _oracle_perturb(sim, e, u, v)
    g = sim.download()
    inner = (slice(2, 47), slice(2, 72))
    for k, (a, b) in enumerate(zip(g, (e, u, v))):
        assert np.array_equal(a[inner], b[inner]), k
    # cells with a land corner are unchanged
    assert np.array_equal(g[0][12:18, 7:10], e[12:18, 7:10])


def test_stochastic_steps_with_coarse_model_error(gpu_ctx):
    """CDKLM16 steps with model error on a grid 3 times coarser against the
    oracle step + SOAR + bicubic interpolation on the same random numbers."""
    print("bicubic: 6 stochastic steps, factor 3", flush=True)
    nx, ny = 96, 63
    kw = cases.config_c1(nx=nx, ny=ny, seed=5)
    sim = CDKLM16(gpu_ctx, **kw)
    sim.setSOARModelError(5e-3, small_scale_perturbation_interpolation_factor=3, xorwow_seed=9)
    ora = O.OracleCDKLM16(**kw)
    dt = np.float32(sim.dt)
    gen = OceanStateNoise.fromsim(sim, soar_q0=5e-3, interpolation_factor=3, xorwow_seed=9)
    W = gen.weights()
    bc = kw["boundary_conditions"]
    p = sim.params
    for step in range(6):
        sim.step(dt)
        ora.run_steps(1, dt)
        gen.generateNormalDistribution()
        xi = gen.getRandomNumbers()
        coarse = O.soar(int(gen.coarse_nx), int(gen.coarse_ny), W, 1.0, bc.isPeriodicNorthSouth(),
                        bc.isPeriodicEastWest(), xi)
        e, u, v = ora.h0
        O.bicubic_interpolation(nx, ny, sim.dx, sim.dy, int(gen.coarse_nx), int(gen.coarse_ny), gen.coarse_dx,
                                gen.coarse_dy, 0, 0, sim.g, sim.f, 0.0, p.north_x, p.north_y, None, coarse, e, u, v,
                                ora.Hi)
        getattr(O.lib(), "gpo_oracle_apply_bc_f32")(ctypes.byref(ora.q), *[O._p(a) for a in ora.h0], None)
        for a, b in zip(sim.download(), ora.download()):
            assert np.array_equal(a, b), f"step {step}"

"""Flow-relaxation boundaries, wind and pressure forcing against the oracle
(exact build, bit for bit), and conservation of mass summed in FP64 on the
device for the Kelvin-wave set-up of the reference's conservation test."""
import numpy as np
import pytest

from gpuocean_amd import AtmosphericPressure, CDKLM16, Common, WindStress
from oracle.oracle import OracleCDKLM16

pytestmark = pytest.mark.gpu


def _base(nx, ny, seed):
    rng = np.random.default_rng(seed)
    jj, ii = np.meshgrid(np.arange(ny + 4), np.arange(nx + 4), indexing="ij")
    eta = (0.3 * np.exp(-((ii - nx / 2) ** 2 + (jj - ny / 2) ** 2) / 60.0)).astype(np.float32)
    hu = (0.2 * rng.standard_normal(eta.shape)).astype(np.float32)
    hv = (0.2 * rng.standard_normal(eta.shape)).astype(np.float32)
    H = (60.0 + 5.0 * rng.random((ny + 5, nx + 5))).astype(np.float32)
    return dict(eta0=eta, hu0=hu, hv0=hv, H=H, nx=nx, ny=ny, dx=200.0, dy=200.0, dt=0.9, g=9.81, f=1.2e-4, r=0.0)


def _frs_data(nx, ny):
    """Two time levels of a smooth exterior solution on every side."""
    xs = np.linspace(0, 2 * np.pi, nx + 4)
    ys = np.linspace(0, 2 * np.pi, ny + 4)

    def side(n, coord, k):
        return Common.SingleBoundaryConditionData(
            h=[np.float32(0.05 * k * np.sin(coord)) for k in (1.0, 2.0)],
            hu=[np.float32(0.5 * k * np.cos(coord)) for k in (1.0, -1.0)],
            hv=[np.float32(0.2 * k * np.ones(n)) for k in (1.0, 3.0)])

    return Common.BoundaryConditionsData(t=[0.0, 40.0], north=side(nx + 4, xs, 1), south=side(nx + 4, xs, 2),
                                         east=side(ny + 4, ys, 3), west=side(ny + 4, ys, 4))


@pytest.mark.parametrize("codes", [(3, 3, 3, 3), (3, 2, 3, 2), (1, 3, 1, 3)])
def test_frs_boundaries_match_oracle(gpu_ctx, codes):
    print(f"FRS {codes}", flush=True)
    nx, ny = 90, 60
    kw = _base(nx, ny, 1)
    sponge = {'north': 10, 'south': 10, 'east': 10, 'west': 10}
    bc = Common.BoundaryConditions(*codes, spongeCells=sponge)
    kw.update(boundary_conditions=bc, boundary_conditions_data=_frs_data(nx, ny))
    sim = CDKLM16(gpu_ctx, **kw, exact_arithmetic=True)
    ora = OracleCDKLM16(**kw)
    for t_end in (9.0, 27.0):   # second call crosses into the interpolated exterior
        sim.step(t_end)
        ora.step(t_end)
        for a, b in zip(sim.download(), ora.download()):
            assert np.array_equal(a, b), f"t {sim.t}"


@pytest.mark.parametrize("codes", [(4, 1, 4, 1), (1, 4, 1, 4), (4, 4, 4, 4), (4, 2, 4, 2), (2, 4, 2, 4),
                                   (3, 4, 4, 3), (4, 3, 1, 4)])
def test_linear_interpolation_boundaries_match_oracle(gpu_ctx, codes):
    """Open boundaries by linear interpolation (code 4), alone and mixed with
    walls, periodic sides and flow relaxation on the same axis, exact build,
    value for value against the oracle."""
    print(f"linear interpolation {codes}", flush=True)
    nx, ny = 90, 60
    kw = _base(nx, ny, 2)
    sponge = {'north': 8, 'south': 10, 'east': 6, 'west': 9}
    bc = Common.BoundaryConditions(*codes, spongeCells=sponge)
    kw.update(boundary_conditions=bc)
    if 3 in codes:
        kw.update(boundary_conditions_data=_frs_data(nx, ny))
    sim = CDKLM16(gpu_ctx, **kw, exact_arithmetic=True)
    ora = OracleCDKLM16(**kw)
    for a, b in zip(sim.download(), ora.download()):
        assert np.array_equal(a, b), "initial boundary conditions"
    for t_end in (9.0, 27.0):
        sim.step(t_end)
        ora.step(t_end)
        for a, b in zip(sim.download(), ora.download()):
            assert np.array_equal(a, b), f"t {sim.t}"
    sim.cleanUp()


def test_wind_and_pressure_forcing_match_oracle(gpu_ctx):
    print("wind + pressure", flush=True)
    nx, ny = 80, 64

    def forcing():
        yy, xx = np.meshgrid(np.linspace(0, 1, 9), np.linspace(0, 1, 11), indexing="ij")
        wind = WindStress(t=[0.0, 30.0],
                          wind_u=[np.float32(12 * np.cos(3 * yy)), np.float32(8 * np.cos(2 * xx))],
                          wind_v=[np.float32(5 * np.sin(4 * xx)), np.float32(-6 * np.sin(yy))])
        pres = AtmosphericPressure(t=[0.0, 30.0],
                                   P=[np.float32(100000 + 300 * np.sin(5 * xx) * np.cos(3 * yy)),
                                      np.float32(100000 - 200 * np.cos(4 * yy))])
        return wind, pres

    kw = _base(nx, ny, 2)
    kw["boundary_conditions"] = Common.BoundaryConditions(1, 2, 1, 2)
    w1, p1 = forcing()
    w2, p2 = forcing()
    sim = CDKLM16(gpu_ctx, **kw, wind=w1, atmospheric_pressure=p1, exact_arithmetic=True)
    ora = OracleCDKLM16(**kw, wind=w2, atmospheric_pressure=p2)
    for t_end in (10.0, 25.0):
        sim.step(t_end)
        ora.step(t_end)
        for a, b in zip(sim.download(), ora.download()):
            assert np.array_equal(a, b), f"t {sim.t}"


def _kelvin(nx, ny, dx, dy, g, f, depth):
    """Kelvin wave along the south wall (test/schemes/ConservationOfMass_test.py:136-168)."""
    rossby = np.sqrt(g * depth) / f
    geo = g * depth / f
    eta = np.zeros((ny + 4, nx + 4), np.float32)
    hu = np.zeros_like(eta)
    x0, y0 = (nx + 4) / 2, 2 - 0.5
    jj, ii = np.meshgrid(np.arange(ny + 4), np.arange(nx + 4), indexing="ij")
    r_j = np.abs(jj - y0) * dy
    r_i = np.abs(ii - x0) * dx
    e = np.exp(-r_j / rossby) * (1.0 + np.tanh((-r_i + rossby) / (rossby / 3)))
    eta[:] = e
    hu[:] = -geo * (-1.0 / rossby) * np.sign(jj - y0) * e
    return eta, hu


def test_mass_conservation_fp64_sum(gpu_ctx):
    """500x200, walls north/south, periodic east/west, f = 1.2e-4, dt = 50,
    two wave periods sampled 50 times: relative mass change below 5e-6
    (the reference asserts 5 decimal places)."""
    print("mass conservation", flush=True)
    nx, ny, dx, dy, g, f, depth = 500, 200, 10000.0, 10000.0, 9.81, 1.2e-4, 100.0
    eta, hu = _kelvin(nx, ny, dx, dy, g, f, depth)
    sim = CDKLM16(gpu_ctx, eta, hu, np.zeros_like(eta), np.full((ny + 5, nx + 5), depth, np.float32), nx, ny, dx,
                  dy, 50.0, g, f, 0.0, boundary_conditions=Common.BoundaryConditions(1, 2, 1, 2))
    m0 = sim.mass()
    assert abs(m0 - np.sum(eta[2:-2, 2:-2].astype(np.float64))) <= 1e-9 * abs(m0)
    period = dx * nx / np.sqrt(g * depth)
    sub = 2 * period / (50.0 * 49)
    worst = 0.0
    for _ in range(49):
        sim.step(sub * sim.dt)
        m = sim.mass()
        host = np.sum(sim.download(interior_domain_only=True)[0].astype(np.float64))
        assert abs(m - host) <= 1e-9 * abs(host)
        worst = max(worst, abs(m - m0) / abs(m0))
    print(f"  worst relative mass change {worst:.2e}", flush=True)
    assert worst < 5e-6

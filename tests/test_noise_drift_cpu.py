"""Oracle-side checks of the noise and drifter restatements (CPU only)."""
import numpy as np

import cases
from gpuocean_amd import Common, WindStress
from gpuocean_amd.OceanStateNoise import soar_weights
from oracle import oracle as O

# Published known-answer vectors of Philox4x32-10 (Random123 kat_vectors).
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]

# Final drifter positions of the reference's wind-drift test
# (test/windForcing/WindForcing_Test.py:127-140).
WIND_DRIFT_REF = [[5046.37353515625, 5120.1171875], [5106.373046875, 5104.00390625],
                  [5150.29638671875, 5060.05859375], [5166.3740234375, 5000.0],
                  [5150.29638671875, 4939.94140625], [5106.373046875, 4895.99609375],
                  [5046.37353515625, 4879.8828125], [4986.37353515625, 4895.99609375],
                  [4942.4501953125, 4939.94140625], [4926.373046875, 5000.0],
                  [4942.4501953125, 5060.05859375], [4986.37353515625, 5104.00390625]]


def test_python_philox_matches_published_vectors():
    for ctr, key, out in PHILOX_KAT:
        assert tuple(O.philox4x32_10(ctr, key)) == out


def test_soar_weights_closed_form():
    W = soar_weights(np.float32(1000.0), np.float32(1000.0), np.float32(5e-3), np.float32(750.0))
    assert W.shape == (5, 5)
    assert W[2, 2] == np.float32(5e-3)
    d = 1000.0
    assert np.isclose(W[2, 3], 5e-3 * (1 + d / 750.0) * np.exp(-d / 750.0), rtol=1e-7)
    assert np.array_equal(W, W[::-1, :]) and np.array_equal(W, W[:, ::-1]) and np.array_equal(W, W.T)


def test_soar_of_a_delta_is_the_stencil():
    nx, ny = 20, 16
    W = soar_weights(np.float32(200), np.float32(200), np.float32(1.0), np.float32(150))
    xi = np.zeros((ny + 8, nx + 8), np.float32)
    xi[10, 12] = 1.0  # xi row r <-> coarse row r-2
    c = O.soar(nx, ny, W, 1.0, False, False, xi)
    nz = np.argwhere(c != 0)
    assert nz.min(0).tolist() == [6, 8] and nz.max(0).tolist() == [10, 12]
    assert np.array_equal(c[6:11, 8:13], W[::-1, ::-1])
    # periodic wrap in both directions
    xi_p = np.zeros((ny, nx), np.float32)
    xi_p[0, 0] = 1.0
    cp = O.soar(nx, ny, W, 2.0, True, True, xi_p)
    assert np.isclose(cp[2, 2], 2.0 * W[2, 2])           # coarse (2,2) is interior cell (0,0)
    assert np.isclose(cp[2 + ny - 1, 2 + nx - 1], 2.0 * W[3, 3])  # wraps to the far corner


def test_geostrophic_balance_closed_form():
    nx, ny = 12, 10
    dx = dy = np.float32(1000.0)
    g, f = np.float32(9.81), np.float32(1e-4)
    coarse = np.zeros((ny + 4, nx + 4), np.float32)
    jj, ii = np.meshgrid(np.arange(ny + 4), np.arange(nx + 4), indexing="ij")
    coarse[:] = (1e-3 * ii + 2e-3 * jj).astype(np.float32)   # linear d_eta: exact central differences
    z = np.zeros_like(coarse)
    eta, hu, hv = z.copy(), z.copy(), z.copy()
    Hi = np.full((ny + 5, nx + 5), 100.0, np.float32)
    O.geostrophic_balance(nx, ny, dx, dy, g, f, 0.0, 0.0, 1.0, None, coarse, eta, hu, hv, Hi)
    inner = (slice(2, ny + 2), slice(2, nx + 2))
    assert np.array_equal(eta[inner], coarse[inner])
    h = coarse[inner] + 100.0
    assert np.allclose(hu[inner], -(9.81 / 1e-4) * h * 2e-6, rtol=1e-5)
    assert np.allclose(hv[inner], (9.81 / 1e-4) * h * 1e-6, rtol=1e-5)
    assert np.all(eta[:2] == 0) and np.all(hu[:, :2] == 0)


def _wind_drift_known_answer(step_ocean, drift_all):
    """The reference's wind-drift known answer: 100x100, dx=100, H=10, periodic,
    30 m/s wind on the ocean; 12 drifters with 10 m/s winds in 12 directions and
    wind_drift_factor 0.02, 600 steps of 1 s."""
    for _ in range(600):
        step_ocean()
        drift_all()


def test_wind_drift_known_answer_on_the_oracle():
    nx = ny = 100
    dx = dy = 100.0
    z = np.zeros((ny + 4, nx + 4), np.float32)
    Hi = 10 * np.ones((ny + 5, nx + 5), np.float32)
    wind = WindStress(t=np.array([0]), wind_u=np.float32([np.array([[30]])]), wind_v=np.float32([np.array([[0]])]))
    o = O.OracleCDKLM16(z, z, z, Hi, nx, ny, dx, dy, 0.0, 9.81, 0.0, 0.0,
                        boundary_conditions=Common.BoundaryConditions(2, 2, 2, 2), wind=wind)
    pos = [np.array([[5000.0, 5000.0], [0.0, 0.0]], np.float32) for _ in range(12)]
    winds = [(np.float32(10 * np.sin(2 * np.pi / 12 * i)), np.float32(10 * np.cos(2 * np.pi / 12 * i)))
             for i in range(12)]

    def drift_all():
        eta, hu, hv = o.download()
        Bm = o.Bm.astype(np.float32)
        for i in range(12):
            O.drift(nx, ny, dx, dy, 1.0, eta, hu, hv, Bm, True, True, pos[i], wind=winds[i], factor=0.02)

    _wind_drift_known_answer(lambda: o.step(1.0), drift_all)
    got = np.array([p[0] for p in pos], np.float64)
    assert np.max(np.abs(got - np.array(WIND_DRIFT_REF))) < 5e-5   # places=4, as the reference asserts


def _drift_guard_case():
    """A uniform eastward flow (u = 2 m/s) and drifters inside, at the edge of
    and beyond the two ghost cells of an open (non-periodic) domain."""
    nx, ny, dx, dy = 40, 30, 100.0, 100.0
    eta = np.zeros((ny + 4, nx + 4), np.float32)
    Hm = np.full((ny + 4, nx + 4), 10.0, np.float32)
    hu = np.full_like(eta, 20.0)
    hv = np.zeros_like(eta)
    x = np.float32([5.5, nx + 1.2, nx + 1.7, nx + 2.6, -1.2, -1.7, -2.6, 20.5, 20.5, np.nan, 1e30]) * dx
    y = np.float32([15.5, 15.5, 15.5, 15.5, 15.5, 15.5, 15.5, ny + 2.6, -2.6, 15.5, 15.5]) * dy
    pos = np.stack([x, y], axis=1).astype(np.float32)
    # drifters whose four surrounding cell centres lie inside the ghost-inclusive arrays
    inside = np.array([True, True, False, False, True, False, False, False, False, False, False])
    return nx, ny, dx, dy, eta, hu, hv, Hm, pos, inside


def test_drift_outside_the_arrays_is_left_in_place():
    """Drifters whose bilinear stencil would leave the ghost-inclusive arrays
    (the reference reads outside them, driftKernels.cu:118-141) stay put;
    the others move by u dt exactly in a uniform flow."""
    nx, ny, dx, dy, eta, hu, hv, Hm, pos, inside = _drift_guard_case()
    ref = pos.copy()
    O.drift(nx, ny, dx, dy, 10.0, eta, hu, hv, Hm, False, False, ref)
    moved = ref[inside]
    assert np.array_equal(moved[:, 0], pos[inside, 0] + np.float32(2.0) * np.float32(10.0))
    assert np.array_equal(moved[:, 1], pos[inside, 1])
    out = ~inside
    assert np.array_equal(ref[out], pos[out], equal_nan=True)

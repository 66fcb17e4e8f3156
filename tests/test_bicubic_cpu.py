"""The oracle's coarse-grid model error (bicubicInterpolation,
gpu_kernels/ocean_noise.cu:401-659) pinned by what the mathematics fixes:

- at fine cells centred on a coarse centre the interpolant returns the coarse
  value exactly, the reference's own "coarse vs fine" check
  (test/stochastic/OceanStateNoise_test.py:290-297);
- Hermite bicubic interpolation with central-difference slopes reproduces any
  field of degree <= 2 in each coarse coordinate, so a wrong sign, index or
  transposed basis fails;
- the balanced momentum of a perturbation linear in y is -(g/f)(H + d) d'(y);
- cells with a land corner are left unchanged;
- aligning the coarse grid with a cell shifts the perturbation of a periodic
  domain by the offset (OceanStateNoise_test.py:300-355).
"""
import numpy as np
import pytest

from oracle import oracle as O

DX = 7.0   # the reference test's cell size


def _run(nx, ny, factor, coarse, oi=0, oj=0, H=5.0, f=0.02, g=9.81, Hi=None, eta=None):
    cnx, cny = -(-nx // factor), -(-ny // factor)
    assert coarse.shape == (cny + 4, cnx + 4)
    eta = np.zeros((ny + 4, nx + 4), np.float32) if eta is None else eta.copy()
    hu = np.zeros_like(eta)
    hv = np.zeros_like(eta)
    Hi = np.full((ny + 5, nx + 5), H, np.float32) if Hi is None else Hi
    O.bicubic_interpolation(nx, ny, DX, DX, cnx, cny, DX * factor, DX * factor, oi, oj, g, f, 0.0, 0.0, 1.0, None,
                            coarse.astype(np.float32), eta, hu, hv, Hi)
    return eta, hu, hv


@pytest.mark.parametrize("factor", [3, 5, 7])
def test_coarse_centres_reproduced_exactly(factor):
    cnx, cny = 6, 5
    nx, ny = factor * cnx, factor * cny
    coarse = np.random.default_rng(factor).normal(size=(cny + 4, cnx + 4)).astype(np.float32)
    eta, _, _ = _run(nx, ny, factor, coarse)
    c = (factor - 1) // 2
    assert np.array_equal(eta[2:-2, 2:-2][c::factor, c::factor], coarse[2:-2, 2:-2])


@pytest.mark.parametrize("factor,nx,ny", [(3, 24, 21), (5, 40, 25), (7, 35, 28), (3, 23, 20)])
def test_biquadratic_fields_reproduced(factor, nx, ny):
    cnx, cny = -(-nx // factor), -(-ny // factor)
    J, I = np.mgrid[0:cny + 4, 0:cnx + 4].astype(np.float64)

    def q(X, Y):   # degree 2 in each coarse coordinate
        return 0.3 + 0.2 * X - 0.1 * Y + 0.05 * X * Y - 0.02 * X * X + 0.03 * Y * Y + 0.004 * X * X * Y \
            - 0.003 * X * Y * Y + 0.0005 * X * X * Y * Y

    eta, _, _ = _run(nx, ny, factor, q(I, J))
    # fine cell centre (tj, ti) in coarse index coordinates
    tj, ti = np.mgrid[1:ny + 3, 1:nx + 3].astype(np.float64)
    X = ((ti - 2) + 0.5) / factor + 1.5
    Y = ((tj - 2) + 0.5) / factor + 1.5
    want = q(X, Y)[1:-1, 1:-1]
    got = eta[2:-2, 2:-2].astype(np.float64)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) < 2e-6 * scale * 16
    # a dropped cross-derivative term is visible at this size
    assert np.max(np.abs(want - q(X, np.floor(Y))[1:-1, 1:-1])) > 1e3 * 2e-6 * scale


def test_balance_of_a_linear_slope():
    factor, cnx, cny = 3, 8, 8
    nx, ny = factor * cnx, factor * cny
    J, I = np.mgrid[0:cny + 4, 0:cnx + 4].astype(np.float64)
    b = 0.01
    coarse = 0.2 + b * J   # d = 0.2 + b * Y, Y in coarse cells
    g, f, H = 9.81, 0.02, 5.0
    eta, hu, hv = _run(nx, ny, factor, coarse, H=H, f=f, g=g)
    d = eta[2:-2, 2:-2].astype(np.float64)
    slope = b / (DX * factor)                     # d d / d y in metres
    want_hu = -(g / f) * (H + d) * slope
    assert np.allclose(hu[2:-2, 2:-2], want_hu, rtol=2e-5, atol=0)
    assert np.max(np.abs(hv[2:-2, 2:-2])) < 1e-6 * np.max(np.abs(want_hu))


def test_land_cells_unchanged():
    factor, cnx, cny = 5, 4, 4
    nx, ny = factor * cnx, factor * cny
    Hi = np.full((ny + 5, nx + 5), 5.0, np.float32)
    Hi[7, 9] = 1e20   # a land corner: cells (6, 8), (6, 9), (7, 8), (7, 9) keep their state
    eta0 = np.full((ny + 4, nx + 4), 0.125, np.float32)
    coarse = np.random.default_rng(1).normal(size=(cny + 4, cnx + 4)).astype(np.float32)
    eta, hu, hv = _run(nx, ny, factor, coarse, Hi=Hi, eta=eta0)
    for r, c in ((6, 8), (6, 9), (7, 8), (7, 9)):
        assert eta[r, c] == 0.125 and hu[r, c] == 0.0 and hv[r, c] == 0.0
    assert eta[5, 8] != 0.125 and hu[5, 8] != 0.0


@pytest.mark.parametrize("oi,oj", [(1, 1), (0, -1), (-1, 0), (-2, 2)])
def test_alignment_offset_shifts_a_periodic_field(oi, oj):
    factor, cnx, cny = 5, 6, 8
    nx, ny = factor * cnx, factor * cny
    P = np.random.default_rng(2).normal(size=(cny, cnx))
    J, I = np.mgrid[0:cny + 4, 0:cnx + 4]
    coarse = P[(J - 2) % cny, (I - 2) % cnx]      # periodic ghost cells, as SOAR writes them
    base, _, _ = _run(nx, ny, factor, coarse)
    shifted, _, _ = _run(nx, ny, factor, coarse, oi=oi, oj=oj)
    want = np.roll(np.roll(base[2:-2, 2:-2], -oj, axis=0), -oi, axis=1)
    scale = np.max(np.abs(base))
    assert np.max(np.abs(shifted[2:-2, 2:-2] - want)) < 1e-5 * scale

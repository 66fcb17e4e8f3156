"""The CPU oracle against the reference's own golden vectors and invariants.

Golden vectors: test/timestep50 of the reference (packed in
tests/golden/timestep50.npz), checked with the reference's metric
(test/schemes/CDKLM16_test.py:112-143, assertAlmostEqual places=3).
"""
import numpy as np
import pytest

import cases
from oracle.oracle import OracleCDKLM16


@pytest.mark.parametrize("case", cases.REGRESSION_CASES, ids=[c[0] for c in cases.REGRESSION_CASES])
def test_oracle_matches_reference_golden(case):
    kw, key = cases.regression_setup(case)
    if case[0] == "betamodel_central":
        # The golden file predates the Coriolis-texture subsampling now on by
        # default (SWEsimulators/CDKLM16.py:222-223); it matches the
        # full-resolution texture.
        kw["subsample_f"] = 0
    o = OracleCDKLM16(**kw)
    t = o.step(50.0)
    assert abs(t - 50.0) < 1e-4
    metric = cases.check_results_metric(o.download(), cases.golden(key))
    for max_rel, l2_rel in metric:
        assert round(max_rel, 3) == 0.0, metric       # places=3, as the reference asserts
        assert abs(max_rel) < 1e-4, metric            # what this restatement actually achieves
        assert l2_rel < 1e-3, metric


def test_step50_takes_56_steps():
    kw, _ = cases.regression_setup(cases.REGRESSION_CASES[0])
    o = OracleCDKLM16(**kw)
    n = 0
    t_now = 0.0
    while t_now < 50.0:
        local_dt = np.float32(min(o.dt, np.float32(50.0 - t_now)))
        t_now += np.float64(local_dt)
        n += 1
    assert n == 56


@pytest.mark.parametrize("case", cases.REGRESSION_CASES[:3] + cases.REGRESSION_CASES[12:15],
                         ids=[c[0] for c in cases.REGRESSION_CASES[:3] + cases.REGRESSION_CASES[12:15]])
def test_oracle_f32_close_to_f64(case):
    kw, _ = cases.regression_setup(case)
    a = OracleCDKLM16(**kw)
    b = OracleCDKLM16(**kw, precision="f64")
    a.step(50.0)
    b.step(50.0)
    for x, y in zip(a.download(), b.download()):
        assert cases.rel_l2(x[2:-2, 2:-2], y[2:-2, 2:-2]) < 2e-4


def test_betamodel_subsampled_texture_is_transposed():
    """The reference's subsample_texture returns RectBivariateSpline(x, y)(xn, yn),
    shape (new_nx, new_ny), and the kernel reads shape[1] as NX: the beta
    gradient ends up along x.  The restatement keeps that layout."""
    from gpuocean_amd import host
    f = np.float32(0.01)
    beta = np.float32(1e-6)
    tex = host.coriolis_texture(f, beta, np.zeros((1, 1), np.float32), 50, 70, np.float32(200),
                                np.float32(200), (74, 54), 10)
    assert tex.shape == (5, 7)
    assert np.allclose(tex[0, :], tex[-1, :])     # constant along the first axis
    assert np.all(np.diff(tex[0, :]) > 0)         # increasing along the second (y samples)


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("bathy", [False, True])
def test_lake_at_rest_is_exact(precision, bathy):
    """Well-balancing: eta = 0, hu = hv = 0 stays exactly at rest over constant
    depth (RealisticForcing_Test.py:191-194) and at rest to rounding over varying
    bathymetry."""
    nx, ny = 40, 30
    H = np.full((ny + 5, nx + 5), 60.0, np.float32)
    if bathy:
        cases.make_bottom_topography(H, nx, ny, 200.0, 200.0, [2, 2, 2, 2])
        H += 20.0
    z = np.zeros((ny + 4, nx + 4), np.float32)
    o = OracleCDKLM16(z, z, z, H, nx, ny, 200.0, 200.0, 1.0, 9.81, 1.2e-4, 0.0, precision=precision)
    o.step(30.0)
    eta, hu, hv = o.download(interior_domain_only=True)
    if not bathy:
        assert np.all(eta == 0.0)
        assert np.all(hu == 0.0)
        assert np.all(hv == 0.0)
    else:
        # over varying depth the balance holds to rounding of the precision
        tol = 1e-12 if precision == "f64" else 1e-4
        assert np.max(np.abs(eta)) < tol * 0.1
        assert np.max(np.abs(hu)) < tol
        assert np.max(np.abs(hv)) < tol


@pytest.mark.parametrize("bcs", [(1, 1, 1, 1), (2, 2, 2, 2), (1, 2, 1, 2), (2, 1, 2, 1)])
def test_mass_conservation(bcs):
    """Sum of interior eta is conserved for wall/periodic BC
    (ConservationOfMass_test.py:87-144), in FP64 to 1e-12 relative."""
    from gpuocean_amd import Common
    nx, ny = 48, 36
    eta = np.zeros((ny + 4, nx + 4), np.float32)
    cases.add_corner_bump(eta, nx, ny, 200.0, 200.0, [2, 2, 2, 2])
    z = np.zeros_like(eta)
    H = np.full((ny + 5, nx + 5), 60.0, np.float32)
    o = OracleCDKLM16(eta, z, z, H, nx, ny, 200.0, 200.0, 0.9, 9.81, 1.2e-4, 0.0, precision="f64",
                      boundary_conditions=Common.BoundaryConditions(*bcs))
    m0 = o.download(True)[0].sum()
    o.step(40.0)
    m1 = o.download(True)[0].sum()
    assert abs(m1 - m0) <= 1e-10 * max(1.0, abs(m0))


def test_symmetry_central_bump():
    """Walls, f = 0, square domain, symmetric bump.  Swapping x and y is an exact
    symmetry of the scheme (eta = eta^T, hu = hv^T bit for bit).  Mirror symmetry
    holds only approximately: the upwind choice of the transverse flux breaks
    ties towards the plus side (CDKLM16_kernel.cu:147)."""
    nx = ny = 40
    jj, ii = np.meshgrid(np.arange(ny + 4) - 2 + 0.5, np.arange(nx + 4) - 2 + 0.5, indexing="ij")
    eta = (0.5 * np.exp(-((ii - nx / 2) ** 2 + (jj - ny / 2) ** 2) / 50.0)).astype(np.float32)
    z = np.zeros_like(eta)
    H = np.full((ny + 5, nx + 5), 60.0, np.float32)
    o = OracleCDKLM16(eta, z, z, H, nx, ny, 200.0, 200.0, 1.0, 9.81, 0.0, 0.0, precision="f64")
    o.step(20.0)
    e, hu, hv = o.download(True)
    assert np.array_equal(e, e.T)
    assert np.array_equal(hu, hv.T)
    assert np.max(np.abs(e - e[:, ::-1])) < 1e-5 * np.max(np.abs(e))
    assert np.max(np.abs(hu + hu[:, ::-1])) < 1e-4 * np.max(np.abs(hu))


def test_cfl_closed_form():
    """max_dt equals 0.25*min(dx/max(|u|+c), dy/max(|v|+c)), the reference's host
    formula (CDKLM16.py:843-861)."""
    kw = cases.config_c0(n=32)
    rng = np.random.default_rng(3)
    kw["hu0"] = (rng.standard_normal(kw["hu0"].shape) * 5).astype(np.float32)
    kw["hv0"] = (rng.standard_normal(kw["hv0"].shape) * 5).astype(np.float32)
    o = OracleCDKLM16(**kw)
    eta, hu, hv = o.download(True)
    Hm = o.Bm[2:-2, 2:-2].astype(np.float32)
    h = eta + Hm
    c = np.sqrt(np.float32(9.81) * h)
    host = 0.25 * min(np.float32(200.0) / np.max(np.abs(hu / h) + c), np.float32(200.0) / np.max(np.abs(hv / h) + c))
    assert np.float32(o.max_dt()) == np.float32(host)


def test_wall_bc_writes_no_ghost_cells():
    """Walls are handled inside the stencil; global ghost cells keep their
    initial values (CDKLM16_kernel.cu:597-604, Common.py:1310-1330)."""
    kw, _ = cases.regression_setup(cases.REGRESSION_CASES[1])
    o = OracleCDKLM16(**kw)
    e0 = o.download()[0].copy()
    o.step(10.0)
    e1 = o.download()[0]
    assert np.array_equal(e0[:2, :], e1[:2, :])
    assert np.array_equal(e0[:, :2], e1[:, :2])


def test_periodic_ghosts_copy_opposite_interior():
    kw, _ = cases.regression_setup(cases.REGRESSION_CASES[4])
    o = OracleCDKLM16(**kw)
    o.step(5.0)
    e = o.download()[0]
    ny, nx = 70, 50
    assert np.array_equal(e[0:2, :], e[ny:ny + 2, :])
    assert np.array_equal(e[ny + 2:ny + 4, :], e[2:4, :])
    assert np.array_equal(e[:, 0:2], e[:, nx:nx + 2])
    assert np.array_equal(e[:, nx + 2:nx + 4], e[:, 2:4])


def _lin_bc_state(codes, sponge, nx=40, ny=30, seed=9):
    from gpuocean_amd import Common
    rng = np.random.default_rng(seed)
    h0 = [rng.standard_normal((ny + 4, nx + 4)).astype(np.float32) for _ in range(3)]
    H = (50.0 + rng.random((ny + 5, nx + 5))).astype(np.float32)
    o = OracleCDKLM16(*[a.copy() for a in h0], H, nx, ny, 100.0, 100.0, 1.0, 9.81, 0.0, 0.0,
                      boundary_conditions=Common.BoundaryConditions(*codes, spongeCells=sponge))
    o.step(0.0)   # at t = 0 this applies the boundary conditions only
    return h0, o.download()


def _on_line(vals, outer, inner, k_outer, k_inner):
    """vals[k] for k strictly between k_outer and k_inner lie on the straight
    line through (k_outer, outer) and (k_inner, inner), to FP32 rounding."""
    k = np.arange(min(k_outer, k_inner) + 1, max(k_outer, k_inner))
    w = (k - k_outer) / (k_inner - k_outer)
    line = outer[None].astype(np.float64) + w[:, None] * (inner[None].astype(np.float64) - outer[None])
    scale = np.maximum(np.abs(outer), np.abs(inner))[None] + 1e-30
    return np.max(np.abs(vals[k] - line) / scale) < 4e-7


@pytest.mark.parametrize("codes", [(4, 1, 4, 1), (4, 2, 1, 2), (1, 1, 4, 1)])
def test_linear_interpolation_ns_is_a_straight_line(codes):
    """Open boundary by linear interpolation (boundary_kernels.cu:265-325): on
    a code-4 side the rows strictly between the outermost ghost row and the
    row `sponge` cells in become the straight line between those two rows,
    which keep their values; the outermost columns and every other row are
    left alone."""
    sn, ss, nx, ny = 6, 9, 40, 30
    h0, h1 = _lin_bc_state(codes, {'north': sn, 'south': ss, 'east': 0, 'west': 0})
    top = ny + 3
    # periodic east/west overwrites the two ghost columns each side afterwards
    c = slice(2, -2) if codes[1] == 2 else slice(1, -1)
    for a0, a in zip(h0, h1):
        touched = np.zeros(a.shape, bool)
        if codes[2] == 4:
            assert _on_line(a[:, c], a0[0, c], a0[ss, c], 0, ss)
            touched[1:ss, 1:-1] = True
        if codes[0] == 4:
            assert _on_line(a[:, c], a0[top, c], a0[top - sn, c], top, top - sn)
            touched[top - sn + 1:top, 1:-1] = True
        if codes[1] == 2:
            touched[:, :2] = touched[:, -2:] = True
        assert np.array_equal(a[~touched], a0[~touched])
        assert not np.array_equal(a[touched], a0[touched])


@pytest.mark.parametrize("codes", [(1, 4, 1, 4), (2, 4, 2, 4), (1, 1, 1, 4)])
def test_linear_interpolation_ew_is_a_straight_line(codes):
    """The east/west counterpart (boundary_kernels.cu:330-384), by columns."""
    se, sw, nx, ny = 5, 7, 40, 30
    h0, h1 = _lin_bc_state(codes, {'north': 0, 'south': 0, 'east': se, 'west': sw})
    right = nx + 3
    for a0, a in zip(h0, h1):
        if codes[0] == 2:   # periodic north/south first: the columns interpolate the updated rows
            a0 = a0.copy()
            a0[:2], a0[-2:] = a0[ny:ny + 2].copy(), a0[2:4].copy()
        touched = np.zeros(a.shape, bool)
        if codes[3] == 4:
            assert _on_line(a[1:-1].T, a0[1:-1, 0], a0[1:-1, sw], 0, sw)
            touched[1:-1, 1:sw] = True
        if codes[1] == 4:
            assert _on_line(a[1:-1].T, a0[1:-1, right], a0[1:-1, right - se], right, right - se)
            touched[1:-1, right - se + 1:right] = True
        assert np.array_equal(a[~touched], a0[~touched])


def test_linear_interpolation_all_sides_rows_then_columns():
    """With code 4 on every side the rows are interpolated first and the
    columns afterwards from the interpolated rows (utils/Common.py:1310-1330),
    so every band cell away from the corners lies on its column's line and
    every east/west band cell on its row's line."""
    sp = {'north': 5, 'south': 5, 'east': 6, 'west': 6}
    nx, ny = 40, 30
    h0, h1 = _lin_bc_state((4, 4, 4, 4), sp)
    top, right = ny + 3, nx + 3
    for a0, a in zip(h0, h1):
        mid = slice(6, right - 6 + 1)   # columns outside the east/west bands
        assert _on_line(a[:, mid], a0[0, mid], a0[5, mid], 0, 5)
        assert _on_line(a[:, mid], a0[top, mid], a0[top - 5, mid], top, top - 5)
        assert _on_line(a[1:-1].T, a[1:-1, 0], a[1:-1, 6], 0, 6)
        assert _on_line(a[1:-1].T, a[1:-1, right], a[1:-1, right - 6], right, right - 6)
        assert np.array_equal(a[[0, -1]], a0[[0, -1]])   # outermost rows are kept

"""Pins of the oracle's flow-relaxation, land/initBm, drying and
atmospheric-pressure restatements (CPU only).

Each check is a property the reference algorithm fixes independently of how
the oracle evaluates it, chosen so that a plausible restatement mistake (an
index off by one, a wrong data row, a swapped coordinate, a sign or a factor
of the pressure term, a land face carrying flux) fails it:

* flow relaxation (gpu_kernels/boundary_kernels.cu:406-540): ghost rows equal
  the exterior solution, cells beyond the sponges are untouched, the time
  levels are reproduced at their own times, the relaxation weight follows the
  logistic profile 2 / (1 + e^j) of 1 - tanh(j/2), the exterior is the
  reference's texture fetch of the boundary data (half-cell shift included),
  east/west is applied after north/south;
* land (utils/Common.py:1428-1432, gpu_kernels/initHm_kernel.cu:38-106,
  CDKLM16_kernel.cu:583,614-665): Bm is the mean of the wet corners and
  land_value when none is wet; a lake at rest around a masked island stays at
  rest and a wave around it conserves the wet cells' mass;
* drying (CDKLM16_kernel.cu:614-665,1070-1085): a lake at rest over a
  desingularised shelf stays at rest; dry cells sit at -Hm + depth_cutoff with
  zero momentum;
* atmospheric pressure (external_forcing.cu:40-68): the reference's own
  pressure-balanced steady states (test/schemes/RealisticForcing_Test.py:
  77-135, 229-253) hold to its places, and a lake at rest stays exactly at rest.
"""
import ctypes

import numpy as np
import pytest

from gpuocean_amd import Common, host
from gpuocean_amd.forcing import AtmosphericPressure
from oracle import oracle as O


# ----------------------------------------------------------------------------- flow relaxation

def _params(nx, ny, codes, sponge):
    """codes and sponge as (north, east, south, west)."""
    q = O.OracleParams()
    q.nx, q.ny, q.rk_order = nx, ny, 2
    q.bc_north, q.bc_east, q.bc_south, q.bc_west = codes
    q.sponge_north, q.sponge_east, q.sponge_south, q.sponge_west = sponge
    q.land_value = float(host.LAND_VALUE)
    return q


def _frs(ns_cur, ns_next, ew_cur, ew_next, t):
    s = O.OracleFRS()
    s.ns_cur, s.ns_next = ns_cur.ctypes.data, ns_next.ctypes.data
    s.ew_cur, s.ew_next = ew_cur.ctypes.data, ew_next.ctypes.data
    s.ns_nx, s.ew_ny = ns_cur.shape[1], ew_cur.shape[0]
    s.t = float(t)
    return s


def _apply(q, state, frs, prec="f64"):
    dt = np.float64 if prec == "f64" else np.float32
    out = [np.array(a, dtype=dt) for a in state]
    getattr(O.lib(), f"gpo_oracle_apply_bc_{prec}")(ctypes.byref(q), *[ctypes.c_void_p(a.ctypes.data) for a in out],
                                                    ctypes.byref(frs))
    return out


def _data(nx, ny, rng):
    ns = rng.standard_normal((2, nx + 4, 3)).astype(np.float32)
    ew = rng.standard_normal((ny + 4, 2, 3)).astype(np.float32)
    return ns, ew


NX, NY = 36, 28
SP = (6, 5, 7, 4)                     # sponge cells north, east, south, west


def _state(rng):
    return [rng.standard_normal((NY + 4, NX + 4)) for _ in range(3)]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_frs_ghost_rows_equal_the_exterior(prec):
    """alpha = 1 on the two ghost rows (columns) of every code-3 side, so they
    take the exterior value exactly; with constant boundary data that is the
    data itself, on the rows of its own side, and the east/west pass (applied
    last) owns the corners."""
    rng = np.random.default_rng(1)
    q = _params(NX, NY, (3, 3, 3, 3), SP)
    c = {k: rng.standard_normal(3).astype(np.float32) for k in "nesw"}
    ns = np.zeros((2, NX + 4, 3), np.float32)
    ns[0], ns[1] = c["s"], c["n"]
    ew = np.zeros((NY + 4, 2, 3), np.float32)
    ew[:, 0], ew[:, 1] = c["w"], c["e"]
    out = _apply(q, _state(rng), _frs(ns, ns, ew, ew, 0.0), prec)
    cols = slice(SP[3], NX + 4 - SP[1])   # outside the east/west sponges
    for k in range(3):
        for r in (0, 1):
            assert np.all(out[k][r, cols] == np.float64(c["s"][k]))
        for r in (NY + 2, NY + 3):
            assert np.all(out[k][r, cols] == np.float64(c["n"][k]))
        for col in (0, 1):
            assert np.all(out[k][:, col] == np.float64(c["w"][k]))
        for col in (NX + 2, NX + 3):
            assert np.all(out[k][:, col] == np.float64(c["e"][k]))


@pytest.mark.parametrize("codes", [(3, 3, 3, 3), (3, 1, 3, 1), (1, 3, 1, 3), (3, 3, 1, 1)])
def test_frs_cells_beyond_the_sponges_are_untouched(codes):
    rng = np.random.default_rng(2)
    q = _params(NX, NY, codes, SP)
    ns, ew = _data(NX, NY, rng)
    st = _state(rng)
    out = _apply(q, st, _frs(ns, ns, ew, ew, 0.0))
    lo_j = SP[2] if codes[2] == 3 else 0
    hi_j = NY + 4 - (SP[0] if codes[0] == 3 else 0)
    lo_i = SP[3] if codes[3] == 3 else 0
    hi_i = NX + 4 - (SP[1] if codes[1] == 3 else 0)
    changed = np.zeros((NY + 4, NX + 4), bool)
    for a, b in zip(out, st):
        changed |= a != b
    inner = np.zeros_like(changed)
    inner[lo_j:hi_j, lo_i:hi_i] = True
    assert not np.any(changed & inner)
    assert np.all(changed[~inner])   # every sponge cell moves toward random data


def test_frs_time_levels():
    """Weight 0 gives the current level whatever the next one holds, and
    weight 1 the next level whatever the current one holds, bit for bit; in
    between the result is affine in the weight."""
    rng = np.random.default_rng(3)
    q = _params(NX, NY, (3, 3, 3, 3), SP)
    A_ns, A_ew = _data(NX, NY, rng)
    B_ns, B_ew = _data(NX, NY, rng)
    C_ns, C_ew = _data(NX, NY, rng)
    st = _state(rng)
    for prec in ("f32", "f64"):
        r0 = _apply(q, st, _frs(A_ns, B_ns, A_ew, B_ew, 0.0), prec)
        rA = _apply(q, st, _frs(A_ns, C_ns, A_ew, C_ew, 0.0), prec)
        r1 = _apply(q, st, _frs(A_ns, B_ns, A_ew, B_ew, 1.0), prec)
        rB = _apply(q, st, _frs(C_ns, B_ns, C_ew, B_ew, 1.0), prec)
        for a, b in zip(r0, rA):
            assert np.array_equal(a, b)
        for a, b in zip(r1, rB):
            assert np.array_equal(a, b)
        for a, b in zip(r0, r1):
            assert not np.array_equal(a, b)
    w = np.float32(0.375)   # exact in binary: the affine combination is exact up to rounding
    rw = _apply(q, st, _frs(A_ns, B_ns, A_ew, B_ew, w))
    r0 = _apply(q, st, _frs(A_ns, B_ns, A_ew, B_ew, 0.0))
    r1 = _apply(q, st, _frs(A_ns, B_ns, A_ew, B_ew, 1.0))
    for a, b, c in zip(rw, r0, r1):
        assert np.max(np.abs(a - ((1 - w) * b + w * c))) < 1e-13


def test_frs_relaxation_profile():
    """State 0 relaxed toward exterior 1 gives the weight itself: at distance
    d cells inward from the outer ghost row it is 1 - tanh(max(0, d-1)/2) =
    2 / (1 + e^max(0, d-1)), decreasing into the domain; state 1 toward 0
    gives its complement."""
    S = 10
    q = _params(NX, NY, (3, 1, 3, 1), (S, 0, S, 0))
    ns = np.ones((2, NX + 4, 3), np.float32)
    ew = np.ones((NY + 4, 2, 3), np.float32)
    zero = [np.zeros((NY + 4, NX + 4)) for _ in range(3)]
    out = _apply(q, zero, _frs(ns, ns, ew, ew, 0.0))
    d = np.arange(S)
    expect = 2.0 / (1.0 + np.exp(np.maximum(0, d - 1)))
    for k in range(3):
        south = out[k][:S, 5]
        north = out[k][NY + 3 - d, 5]
        assert np.max(np.abs(south - expect)) < 1e-14
        assert np.max(np.abs(north - expect)) < 1e-14
        assert np.all(np.diff(south[1:]) < 0)
    ones = [np.ones((NY + 4, NX + 4)) for _ in range(3)]
    comp = _apply(q, ones, _frs(0 * ns, 0 * ns, 0 * ew, 0 * ew, 0.0))
    assert np.max(np.abs(comp[0][:S, 5] + out[0][:S, 5] - 1.0)) < 1e-14


def test_frs_exterior_is_the_texture_fetch_of_the_data():
    """Boundary data linear along its side, d[i] = a + b i, on a texture as
    wide as the ghost-inclusive grid: the reference samples it at normalised
    coordinate i / (n + 4) with the texture's half-texel offset, which gives
    a + b (i - 1/2) (column 0: d[0]); the south row of the data serves the
    south sponge and the north row the north one.  Likewise down the columns
    for east/west."""
    q = _params(NX, NY, (3, 1, 3, 1), (4, 0, 4, 0))
    i = np.arange(NX + 4)
    ns = np.zeros((2, NX + 4, 3), np.float32)
    ns[0, :, 0], ns[0, :, 1], ns[0, :, 2] = 1.0 + 0.25 * i, -2.0 + 0.5 * i, 3.0
    ns[1, :, 0], ns[1, :, 1], ns[1, :, 2] = -1.0 - 0.125 * i, 4.0, 0.75 * i
    ew = np.zeros((NY + 4, 2, 3), np.float32)
    out = _apply(q, [np.zeros((NY + 4, NX + 4))] * 3, _frs(ns, ns, ew, ew, 0.0))
    x = np.where(i == 0, 0.0, i - 0.5)
    for k in range(3):
        for row, side in ((0, 0), (1, 0), (NY + 2, 1), (NY + 3, 1)):
            ref = ns[side, 0, k] + (ns[side, 1, k] - ns[side, 0, k]) * x
            assert np.max(np.abs(out[k][row] - ref)) < 1e-12, (k, row)
    q = _params(NX, NY, (1, 3, 1, 3), (0, 4, 0, 4))
    j = np.arange(NY + 4)
    ew = np.zeros((NY + 4, 2, 3), np.float32)
    ew[:, 0, 0], ew[:, 1, 0] = 2.0 + 0.5 * j, -1.0 + 0.25 * j
    ew[:, 0, 2], ew[:, 1, 2] = -0.5 * j, 5.0
    out = _apply(q, [np.zeros((NY + 4, NX + 4))] * 3, _frs(ns, ns, ew, ew, 0.0))
    y = np.where(j == 0, 0.0, j - 0.5)
    for k in (0, 2):
        for col, side in ((0, 0), (1, 0), (NX + 2, 1), (NX + 3, 1)):
            ref = ew[0, side, k] + (ew[1, side, k] - ew[0, side, k]) * y
            assert np.max(np.abs(out[k][:, col] - ref)) < 1e-12, (k, col)


def test_frs_interior_at_the_exterior_state_is_unchanged():
    """A state that already equals the exterior solution is left as it is (to
    rounding in FP32, exactly up to 1 ulp in FP64)."""
    rng = np.random.default_rng(4)
    q = _params(NX, NY, (3, 3, 3, 3), SP)
    c = rng.standard_normal(3).astype(np.float32)
    ns = np.broadcast_to(c, (2, NX + 4, 3)).copy()
    ew = np.broadcast_to(c, (NY + 4, 2, 3)).copy()
    st = [np.full((NY + 4, NX + 4), np.float64(c[k])) for k in range(3)]
    for prec, tol in (("f64", 4e-16), ("f32", 2.5e-7)):
        out = _apply(q, st, _frs(ns, ns, ew, ew, 0.5), prec)
        for k in range(3):
            assert np.max(np.abs(out[k] - c[k])) <= tol * abs(c[k])


# ----------------------------------------------------------------------------- land and drying

def _land_case(precision, bc=(1, 1, 1, 1), eta=None, f=1e-4):
    nx, ny = 40, 30
    shape = (ny + 4, nx + 4)
    rng = np.random.default_rng(1)
    H = (20 + 5 * rng.random((ny + 5, nx + 5))).astype(np.float32)
    mask = np.zeros(H.shape, bool)
    mask[12:19, 15:24] = True      # an island of land corners
    mask[14, 25] = True            # one isolated land corner (cells with 3 wet corners)
    mask[5, 5] = mask[6, 5] = mask[5, 6] = True   # three land corners of a cell (1 wet)
    z = np.zeros(shape, np.float32)
    o = O.OracleCDKLM16(z if eta is None else eta, z, z, np.ma.masked_array(H, mask), nx, ny, 100.0, 100.0, 0.0,
                        9.81, f, 0.0, boundary_conditions=Common.BoundaryConditions(*bc), precision=precision)
    return o, H, mask


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_initbm_is_the_mean_of_the_wet_corners(precision):
    o, H, mask = _land_case(precision)
    ny, nx = o.ny, o.nx
    Bm = o.Bm
    n_wet_seen = set()
    for j in range(2, ny + 2):
        for i in range(2, nx + 2):
            c = [(j, i), (j + 1, i), (j, i + 1), (j + 1, i + 1)]
            wet = [np.float64(H[p]) for p in c if not mask[p]]
            n_wet_seen.add(len(wet))
            if not wet:
                assert Bm[j, i] == np.asarray(host.LAND_VALUE, Bm.dtype)
            else:
                ref = np.mean(wet)
                tol = 1e-13 if precision == "f64" else 4e-7
                assert abs(Bm[j, i] - ref) <= tol * ref, (j, i, len(wet))
    assert n_wet_seen == {0, 1, 2, 3, 4}


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (2, 2, 2, 2)])
def test_lake_at_rest_around_a_masked_island(bc):
    """Land faces carry only the wet side's pressure g/2 (eta + H)^2, which the
    bathymetry source balances: a lake at rest around an island, with its
    surface 0.5 m above the datum (so that the two sides of a land face differ),
    stays at rest (FP64 to rounding), and land cells stay land."""
    nx, ny = 40, 30
    eta = np.full((ny + 4, nx + 4), 0.5, np.float32)
    o, H, mask = _land_case("f64", bc, eta=eta)
    o.step(200 * float(o.dt))
    e, u, v = o.download(interior_domain_only=True)
    land = np.abs(o.Bm[2:-2, 2:-2] - host.LAND_VALUE) < 1.0
    assert land.sum() > 40
    assert np.abs(e[~land] - 0.5).max() < 1e-12
    assert max(np.abs(u).max(), np.abs(v).max()) < 1e-11
    assert np.all(u[land] == 0) and np.all(v[land] == 0)
    o32, _, _ = _land_case("f32", bc, eta=eta)
    o32.step(200 * float(o32.dt))
    e32, u32, v32 = o32.download(interior_domain_only=True)
    assert np.abs(e32[~land] - 0.5).max() < 1e-5 and max(np.abs(u32).max(), np.abs(v32).max()) < 1e-3


def test_mass_around_an_island():
    """A bump beside an island, walls, f = 0: the wave reaches and wraps
    around the island and the wet-cell mass is conserved (no mass flux through
    land faces).  (Mirror symmetry is no invariant of the scheme: the upwinded
    tangential flux of CDKLM16_kernel.cu:147 breaks ties toward the plus side.)"""
    nx, ny = 40, 30
    shape = (ny + 4, nx + 4)
    H = np.full((ny + 5, nx + 5), 20.0, np.float32)
    mask = np.zeros(H.shape, bool)
    # corner row K and cell row J (ghost-inclusive) mirror onto 34 - K and 33 - J
    mask[14:21, 17:28] = True
    jj, ii = np.meshgrid(np.arange(ny + 4), np.arange(nx + 4), indexing="ij")
    eta = (0.3 * np.exp(-((ii - 8.5) ** 2 + (jj - 16.5) ** 2) / 18.0)).astype(np.float32)
    z = np.zeros(shape, np.float32)
    o = O.OracleCDKLM16(eta, z, z, np.ma.masked_array(H, mask), nx, ny, 100.0, 100.0, 0.0, 9.81, 0.0, 0.0,
                        precision="f64")
    land = np.abs(o.Bm[2:-2, 2:-2] - host.LAND_VALUE) < 1.0
    m0 = o.download(interior_domain_only=True)[0][~land].sum()
    o.step(150 * float(o.dt))
    e, u, v = o.download(interior_domain_only=True)
    assert abs(e[~land].sum() - m0) < 1e-10 * np.abs(eta).sum()
    assert np.abs(v).max() > 1e-3    # the wave does reach and wrap around the island
    assert np.all(u[land] == 0) and np.all(v[land] == 0)


def test_lake_at_rest_over_a_desingularised_shelf():
    """Depths between depth_cutoff and desingularization_eps (0.02..0.09 m)
    use the desingularised velocities; at rest they stay at rest."""
    nx, ny = 40, 30
    shape = (ny + 4, nx + 4)
    rng = np.random.default_rng(2)
    H = np.full((ny + 5, nx + 5), 20.0, np.float32)
    H[:, :20] = (0.02 + 0.07 * rng.random((ny + 5, 20))).astype(np.float32)
    z = np.zeros(shape, np.float32)
    o = O.OracleCDKLM16(z, z, z, H, nx, ny, 100.0, 100.0, 0.0, 9.81, 0.0, 0.0, precision="f64")
    o.step(100 * float(o.dt))
    e, u, v = o.download(interior_domain_only=True)
    assert np.abs(e).max() < 1e-13 and max(np.abs(u).max(), np.abs(v).max()) < 1e-12


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_dry_cells_sit_at_the_cutoff_with_zero_momentum(precision):
    """A dry patch (eta = -H) carrying momentum, inside a lake: after a step
    its cells farther than the stencil from water hold eta = -Hm +
    depth_cutoff and hu = hv = 0 exactly."""
    nx, ny = 40, 30
    shape = (ny + 4, nx + 4)
    H = np.full((ny + 5, nx + 5), 20.0, np.float32)
    eta = np.zeros(shape, np.float32)
    hu = np.zeros(shape, np.float32)
    hv = np.zeros(shape, np.float32)
    eta[10:24, 12:30] = -20.0
    hu[10:24, 12:30] = 3.0
    hv[10:24, 12:30] = -2.0
    o = O.OracleCDKLM16(eta, hu, hv, H, nx, ny, 100.0, 100.0, 0.5, 9.81, 0.0, 0.0, precision=precision)
    o.step(float(o.dt))
    e, u, v = o.download()
    inner = (slice(13, 21), slice(15, 27))
    dt_ = np.float32 if precision == "f32" else np.float64
    expect = -dt_(20.0) + dt_(np.float32(1e-5) if precision == "f32" else 1e-5)
    assert np.all(e[inner] == expect)
    assert np.all(u[inner] == 0) and np.all(v[inner] == 0)
    assert np.any(e[10:24, 12:30] > expect)    # the water does run in at the edges


# ----------------------------------------------------------------------------- atmospheric pressure

def _pressure_case(kind):
    """RealisticForcing_Test's set-up (test/schemes/RealisticForcing_Test.py:35-55)
    and its balanced pressure fields (:77-135)."""
    nx, ny, dx, dy, g, rho = 500, 400, 800.0, 500.0, 9.81, 1015.0
    shape = (ny + 4, nx + 4)
    p_atm = np.ones(shape, np.float32) * np.float32(100000)
    eta0 = np.zeros(shape, np.float32)
    fullx = np.linspace(-1.5, nx + 1.5, nx + 4)
    fully = np.linspace(-1.5, ny + 1.5, ny + 4)
    X, Y = np.meshgrid(fullx, fully)
    if kind == "linear_x":
        p_atm += (-2000 + 4000 * X / nx).astype(np.float32)
        eta0[:] = -4000 * X / (nx * (g * rho))
    elif kind == "linear_y":
        p_atm += (-2000 + 4000 * Y / ny).astype(np.float32)
        eta0[:] = -4000 * Y / (ny * (g * rho))
    elif kind == "linear_diag":
        p_atm += (-2000 + 4000 * (X + Y) / (nx + ny)).astype(np.float32)
        eta0[:] = -4000 * (X + Y) / ((nx + ny) * (g * rho))
    elif kind == "bumps":
        x = np.linspace(0, 2 * np.pi, nx // 2)
        y = np.linspace(0, 2 * np.pi, ny)
        bx, by = np.meshgrid(x, y)
        a = 0.5 * (1 - np.cos(bx)) * (1 - np.cos(by)) * 2000
        b = 0.5 * (np.cos(bx) - 1) * (1 - np.cos(by)) * 2000
        p = p_atm.astype(np.float64)
        p[:ny, :nx // 2] += a
        p[:ny, nx // 2:nx] += b
        p_atm = p.astype(np.float32)
        eta0[:ny, :nx // 2] = -a / (g * rho)
        eta0[:ny, nx // 2:nx] = -b / (g * rho)
    H = np.ones((shape[0] + 1, shape[1] + 1), np.float32) * 100
    z = np.zeros(shape, np.float32)
    ap = AtmosphericPressure(P=[p_atm]) if kind != "rest" else None
    o = O.OracleCDKLM16(eta0, z, z, H, nx, ny, dx, dy, 0.0, g, 0.0, 0.0, rho_o=rho, atmospheric_pressure=ap)
    return o, eta0


@pytest.mark.parametrize("kind,places", [("rest", None), ("linear_x", 3), ("linear_y", 3), ("linear_diag", 2),
                                         ("bumps", 2)])
def test_pressure_balanced_steady_states(kind, places):
    """check_steady_state of RealisticForcing_Test (:174-193) after step(3600):
    max |hu|, |hv| and |eta - eta0| over the interior without its outer ring
    round to 0 at the reference's places (exactly 0 at rest)."""
    o, eta0 = _pressure_case(kind)
    o.step(3600.0)
    e, u, v = o.download(interior_domain_only=True)
    d = (e - eta0[2:-2, 2:-2])[1:-1, 1:-1]
    worst = [np.abs(u[1:-1, 1:-1]).max(), np.abs(v[1:-1, 1:-1]).max(), np.abs(d).max()]
    if places is None:
        assert worst == [0, 0, 0]
    else:
        for w in worst:
            assert round(float(w), places) == 0, (kind, worst)

"""Initial conditions and configurations used by the tests.

The 50x70 regression cases restate the reference's test set-up
(test/schemes/CDKLM16_test.py:37-419, ICs from test/testUtils.py:99-142):
dx=dy=200, H=60 (or the analytic bathymetry), dt=0.9, step(50.0).
"""

import numpy as np

from gpuocean_amd import Common


def _bump(eta, nx, ny, dx, dy, halo, x_center, y_center):
    """exp(-(x^2+y^2)/size) inside radius size, size = 500*min(dx, dy)
    (test/testUtils.py:111-142), added over ghost cells too."""
    size = 500.0 * min(dx, dy)
    for j in range(-halo[2], ny + halo[0]):
        for i in range(-halo[3], nx + halo[1]):
            x = dx * i - x_center
            y = dy * j - y_center
            if np.sqrt(x ** 2 + y ** 2) < size:
                eta[j + halo[2], i + halo[3]] += np.exp(-(x ** 2 / size + y ** 2 / size))


def add_central_bump(eta, nx, ny, dx, dy, halo):
    _bump(eta, nx, ny, dx, dy, halo, dx * nx / 2.0, dy * ny / 2.0)


def add_corner_bump(eta, nx, ny, dx, dy, halo):
    _bump(eta, nx, ny, dx, dy, halo, 4 * dx, 4 * dy)


def add_upper_corner_bump(eta, nx, ny, dx, dy, halo):
    _bump(eta, nx, ny, dx, dy, halo, (nx - 4) * dx, (ny - 4) * dy)


def make_bottom_topography(Hi, nx, ny, dx, dy, halo):
    """Analytic corner bathymetry (test/testUtils.py:99-108)."""
    for j in range(-halo[2], ny + halo[0] + 1):
        for i in range(-halo[3], nx + halo[1] + 1):
            Hi[j + halo[2], i + halo[3]] = 6 + 2.0 * np.cos(0.3 * (i + i / (np.sin(0.5 * j) + 2.5))) + \
                2.0 * np.sin(2 * np.pi * (j + i) / (2.0 * ny))


BC_SETTINGS = {
    "wall": (1, 1, 1, 1),
    "periodic": (2, 2, 2, 2),
    "periodicNS": (2, 1, 2, 1),
    "periodicEW": (1, 2, 1, 2),
}

# (test name, bc setting, bump, golden key prefix, f, beta, bathymetry)
REGRESSION_CASES = [
    ("wall_central", "wall", "central", "wallBC_central", 0.0, 0.0, False),
    ("wall_corner", "wall", "corner", "wallBC_corner", 0.0, 0.0, False),
    ("wall_upperCorner", "wall", "upperCorner", "wallBC_upperCorner", 0.0, 0.0, False),
    ("periodic_central", "periodic", "central", "wallBC_central", 0.0, 0.0, False),
    ("periodic_corner", "periodic", "corner", "periodic_corner", 0.0, 0.0, False),
    ("periodic_upperCorner", "periodic", "upperCorner", "periodic_upperCorner", 0.0, 0.0, False),
    ("periodicNS_central", "periodicNS", "central", "wallBC_central", 0.0, 0.0, False),
    ("periodicNS_corner", "periodicNS", "corner", "periodicNS_corner", 0.0, 0.0, False),
    ("periodicNS_upperCorner", "periodicNS", "upperCorner", "periodicNS_upperCorner", 0.0, 0.0, False),
    ("periodicEW_central", "periodicEW", "central", "wallBC_central", 0.0, 0.0, False),
    ("periodicEW_corner", "periodicEW", "corner", "periodicEW_corner", 0.0, 0.0, False),
    ("periodicEW_upperCorner", "periodicEW", "upperCorner", "periodicEW_upperCorner", 0.0, 0.0, False),
    ("coriolis_central", "wall", "central", "coriolis_central", 0.01, 0.0, False),
    ("betamodel_central", "wall", "central", "betamodel_central", 0.01 + 2 * 1e-6 * 200.0, 1e-6, False),
    ("bathymetry_central", "wall", "central", "wallBC_central_bathymetry", 0.0, 0.0, True),
]


def regression_setup(case):
    """Constructor kwargs of one 50x70 regression case."""
    name, bcs, bump, key, f, beta, bathy = case
    nx, ny, dx, dy = 50, 70, 200.0, 200.0
    ghosts = [2, 2, 2, 2]
    shape = (ny + 4, nx + 4)
    eta0 = np.zeros(shape, dtype=np.float32)
    hu0 = np.zeros(shape, dtype=np.float32)
    hv0 = np.zeros(shape, dtype=np.float32)
    Hi = np.ones((shape[0] + 1, shape[1] + 1), dtype=np.float32) * 60
    {"central": add_central_bump, "corner": add_corner_bump,
     "upperCorner": add_upper_corner_bump}[bump](eta0, nx, ny, dx, dy, ghosts)
    if bathy:
        make_bottom_topography(Hi, nx, ny, dx, dy, ghosts)
    kw = dict(eta0=eta0, hu0=hu0, hv0=hv0, H=Hi, nx=nx, ny=ny, dx=dx, dy=dy, dt=0.9, g=9.81, f=f, r=0.0,
              boundary_conditions=Common.BoundaryConditions(*BC_SETTINGS[bcs]))
    if beta != 0.0:
        kw["coriolis_beta"] = beta
    return kw, key


def golden(key):
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "timestep50.npz")
    with np.load(path) as z:
        return z[key + "_eta"], z[key + "_u"], z[key + "_v"]


def check_results_metric(sim_arrays, ref_arrays):
    """The reference's checkResults metric (CDKLM16_test.py:112-143): signed
    max(sim-ref)/max|ref| over the interior, and the L2 variant."""
    out = []
    for a, b in zip(sim_arrays, ref_arrays):
        a, b = a[2:-2, 2:-2], b[2:-2, 2:-2]
        scale = np.max(np.abs(b))
        out.append((float(np.max(a - b) / scale), float(np.linalg.norm(a - b) / scale)))
    return out


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


# ------------------------------------------------------------------ benchmark configurations

def config_c0(n=256, bump_amp=0.5, sigma=10.0):
    """256x256, dx=dy=200, H=60 on corners, walls, g=9.81, f=1.2e-4, r=0;
    Gaussian bump eta = 0.5 exp(-r^2/(2*10^2)) cell units, centred at the
    domain centre (cell-centre convention: interior cell (i, j) at i+0.5)."""
    nx = ny = n
    shape = (ny + 4, nx + 4)
    jj, ii = np.meshgrid(np.arange(ny + 4) - 2 + 0.5, np.arange(nx + 4) - 2 + 0.5, indexing="ij")
    eta0 = (bump_amp * np.exp(-((ii - nx / 2) ** 2 + (jj - ny / 2) ** 2) / (2 * sigma ** 2))).astype(np.float32)
    hu0 = np.zeros(shape, np.float32)
    hv0 = np.zeros(shape, np.float32)
    H = np.full((ny + 5, nx + 5), 60.0, np.float32)
    return dict(eta0=eta0, hu0=hu0, hv0=hv0, H=H, nx=nx, ny=ny, dx=200.0, dy=200.0, dt=0.0, g=9.81,
                f=1.2e-4, r=0.0, courant_number=0.7, boundary_conditions=Common.BoundaryConditions(1, 1, 1, 1))


def smooth_random_field(shape, sigma, seed, wrap_x=True):
    """Gaussian-filtered N(0,1) field (sigma in cells): periodic in x, reflecting in y."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(shape)
    ny, nx = shape

    def kernel(n, s):
        k = np.arange(-int(4 * s), int(4 * s) + 1)
        w = np.exp(-0.5 * (k / s) ** 2)
        return k, w / w.sum()

    k, w = kernel(nx, sigma)
    out = np.zeros_like(z)
    for kk, ww in zip(k, w):
        out += ww * np.roll(z, kk, axis=1)
    z = out
    k, w = kernel(ny, sigma)
    pad = len(k) // 2
    zp = np.pad(z, ((pad, pad), (0, 0)), mode="symmetric")
    out = np.zeros_like(z)
    for idx, ww in enumerate(w):
        out += ww * zp[idx:idx + ny, :]
    return out


def config_c1(nx=2048, ny=1024, seed=1):
    """Double jet: 2048x1024, dx=dy=1 km, periodic in x, walls in y, f=1e-4.
    H = 250 + 30*p/max|p| on corners; eta = 0.3(tanh((y-y1)/50km) - tanh((y-y2)/50km)),
    y1 = Ly/4, y2 = 3Ly/4; hu from geostrophic balance, hv = 0."""
    dx = dy = 1000.0
    g, f = 9.81, 1e-4
    p = smooth_random_field((ny + 5, nx + 5), 20.0, seed)
    H = (250.0 + 30.0 * p / np.max(np.abs(p))).astype(np.float32)
    y = (np.arange(ny + 4) - 2 + 0.5) * dy
    Ly = ny * dy
    y1, y2 = Ly / 4, 3 * Ly / 4
    eta_y = 0.3 * (np.tanh((y - y1) / 5e4) - np.tanh((y - y2) / 5e4))
    eta0 = np.repeat(eta_y[:, None], nx + 4, axis=1).astype(np.float32)
    Hm = 0.25 * (H[:-1, :-1] + H[1:, :-1] + H[:-1, 1:] + H[1:, 1:])
    deta_dy = np.gradient(eta_y, dy)
    hu0 = (-(g / f) * (Hm + eta0) * deta_dy[:, None]).astype(np.float32)
    hv0 = np.zeros_like(eta0)
    return dict(eta0=eta0, hu0=hu0, hv0=hv0, H=H, nx=nx, ny=ny, dx=dx, dy=dy, dt=0.0, g=g, f=f, r=0.0,
                courant_number=0.7, boundary_conditions=Common.BoundaryConditions(1, 2, 1, 2))


def config_c1_large(nx=8192, ny=4096):
    """The double jet of config_c1 on a grid larger than the 256 MB last-level
    cache (8192 x 4096: about 0.8 GB of state, bathymetry and stage), with an
    analytic bathymetry 250 + 30 sin(6 pi x / Lx) cos(4 pi y / Ly) instead of
    the filtered random field, whose construction at this size takes minutes."""
    dx = dy = 1000.0
    g, f = 9.81, 1e-4
    xc = (np.arange(nx + 5) - 2.0) / nx
    yc = (np.arange(ny + 5) - 2.0) / ny
    H = (250.0 + 30.0 * np.sin(6 * np.pi * xc)[None, :] * np.cos(4 * np.pi * yc)[:, None]).astype(np.float32)
    y = (np.arange(ny + 4) - 2 + 0.5) * dy
    Ly = ny * dy
    eta_y = 0.3 * (np.tanh((y - Ly / 4) / 5e4) - np.tanh((y - 3 * Ly / 4) / 5e4))
    eta0 = np.repeat(eta_y[:, None], nx + 4, axis=1).astype(np.float32)
    Hm = 0.25 * (H[:-1, :-1] + H[1:, :-1] + H[:-1, 1:] + H[1:, 1:])
    hu0 = (-(g / f) * (Hm + eta0) * np.gradient(eta_y, dy)[:, None]).astype(np.float32)
    return dict(eta0=eta0, hu0=hu0, hv0=np.zeros_like(eta0), H=H, nx=nx, ny=ny, dx=dx, dy=dy, dt=0.0, g=g, f=f,
                r=0.0, courant_number=0.7, boundary_conditions=Common.BoundaryConditions(1, 2, 1, 2))


def c2_member_sizes(n_members=256, seed=2024):
    """Seeded uneven member sizes: nx = 16*round(U{128..512}/16),
    ny = 16*round(U{128..384}/16)."""
    rng = np.random.default_rng(seed)
    nx = 16 * np.round(rng.integers(128, 513, n_members) / 16).astype(int)
    ny = 16 * np.round(rng.integers(128, 385, n_members) / 16).astype(int)
    return [(int(a), int(b)) for a, b in zip(nx, ny)]


def config_c2_member(nx, ny, seed):
    """The double jet of config_c1 on an nx x ny member, with every length
    scaled by the member's Ly / 1024 km: jet width 50 km * s, bathymetry
    smoothing 20 cells * s (at least 2 cells); dx=dy=1 km, periodic in x."""
    dx = dy = 1000.0
    g, f = 9.81, 1e-4
    s = ny / 1024.0
    p = smooth_random_field((ny + 5, nx + 5), max(2.0, 20.0 * s), seed)
    H = (250.0 + 30.0 * p / np.max(np.abs(p))).astype(np.float32)
    y = (np.arange(ny + 4) - 2 + 0.5) * dy
    Ly = ny * dy
    y1, y2 = Ly / 4, 3 * Ly / 4
    w = 5e4 * s
    eta_y = 0.3 * (np.tanh((y - y1) / w) - np.tanh((y - y2) / w))
    eta0 = np.repeat(eta_y[:, None], nx + 4, axis=1).astype(np.float32)
    Hm = 0.25 * (H[:-1, :-1] + H[1:, :-1] + H[:-1, 1:] + H[1:, 1:])
    deta_dy = np.gradient(eta_y, dy)
    hu0 = (-(g / f) * (Hm + eta0) * deta_dy[:, None]).astype(np.float32)
    hv0 = np.zeros_like(eta0)
    return dict(eta0=eta0, hu0=hu0, hv0=hv0, H=H, nx=nx, ny=ny, dx=dx, dy=dy, dt=0.0, g=g, f=f, r=0.0,
                courant_number=0.7, boundary_conditions=Common.BoundaryConditions(1, 2, 1, 2))


def config_c2(n_members=256, seed=2024, sizes=None):
    """Per-member keyword arguments of the batched stochastic ensemble:
    uneven members, SOAR q0 = 5e-3, L = 0.75 dx, seed per member."""
    sizes = c2_member_sizes(n_members, seed) if sizes is None else sizes
    return [config_c2_member(nx, ny, seed + 1 + i) for i, (nx, ny) in enumerate(sizes)]


def config_c3(n=512, seed=3):
    """Configuration-0 physics (dx=dy=200 m, walls, g=9.81, f=1.2e-4, r=0,
    CFL 0.7) over a seeded smooth random bathymetry of 50..150 m on corners,
    Gaussian bump 0.5 m (sigma 10 cells) at the centre, at rest."""
    kw = config_c0(n=n)
    p = smooth_random_field((n + 5, n + 5), 20.0, seed)
    kw["H"] = (100.0 + 50.0 * p / np.max(np.abs(p))).astype(np.float32)
    return kw


def c3_background_and_truth(kw, seed=3):
    """The state the gradient is taken at (the bump plus small smooth random
    perturbations of eta, hu and hv, so that no limiter sits exactly at a
    kink) and the perturbed truth the observations come from."""
    n_y, n_x = kw["eta0"].shape
    r = [smooth_random_field((n_y, n_x), 8.0, seed + 10 + i) for i in range(6)]
    r = [a / np.max(np.abs(a)) for a in r]
    bg = (kw["eta0"] + 0.01 * r[0], 0.5 * r[1], 0.5 * r[2])
    truth = (bg[0] + 0.05 * r[3], bg[1] + 2.0 * r[4], bg[2] + 2.0 * r[5])
    return [a.astype(np.float32) for a in bg], [a.astype(np.float32) for a in truth]


def c3_drifters(nx, ny, dx, dy, n=200, seed=3):
    """Seeded observation positions inside the interior (one cell from the edges)."""
    rng = np.random.default_rng(seed + 100)
    x = rng.uniform(1.0 * dx, (nx - 1.0) * dx, n)
    y = rng.uniform(1.0 * dy, (ny - 1.0) * dy, n)
    return np.stack([x, y], axis=1).astype(np.float32)


def config_c4_member(nx, ny, seed, codes=(3, 3, 3, 1), sponge=10):
    """Coastal member: dx=dy=1 km, f=1e-4; depth ramps linearly from 0 m at
    x = 0.1 Lx to 300 m at the east edge, so the western 10 % (depth <= 0) is
    land (masked corner depths); a seeded zonal jet (width and position
    jittered by the seed) in geostrophic balance over the ramp; flow
    relaxation toward that jet on the sides coded 3 (sponge cells wide),
    boundary data of one time."""
    rng = np.random.default_rng(seed)
    dx = dy = 1000.0
    g, f = 9.81, 1e-4
    Lx, Ly = nx * dx, ny * dy
    xc = (np.arange(nx + 5) - 2) * dx
    Hrow = 300.0 * (xc - 0.1 * Lx) / (0.9 * Lx)
    H = np.repeat(Hrow[None, :], ny + 5, axis=0).astype(np.float32)
    H = np.ma.masked_where(H <= 0.0, H)
    y = (np.arange(ny + 4) - 2 + 0.5) * dy
    yc = Ly * (0.5 + 0.1 * rng.uniform(-1, 1))
    w = Ly * (0.08 + 0.04 * rng.uniform())
    eta_y = 0.2 * np.exp(-((y - yc) / w) ** 2)
    eta0 = np.repeat(eta_y[:, None], nx + 4, axis=1).astype(np.float32)
    Hf = np.ma.filled(H, 0.0)
    Hm = np.maximum(0.25 * (Hf[:-1, :-1] + Hf[1:, :-1] + Hf[:-1, 1:] + Hf[1:, 1:]), 0.0)
    wet = Hm > 1.0
    eta0 = np.where(wet, eta0, 0.0).astype(np.float32)
    deta_dy = np.gradient(eta_y, dy)
    hu0 = np.where(wet, -(g / f) * (Hm + eta0) * deta_dy[:, None], 0.0).astype(np.float32)
    hv0 = np.zeros_like(eta0)

    def ns(row):
        return Common.SingleBoundaryConditionData(h=[eta0[row].copy()], hu=[hu0[row].copy()], hv=[hv0[row].copy()])

    def ew(col):
        return Common.SingleBoundaryConditionData(h=[eta0[:, col].copy()], hu=[hu0[:, col].copy()],
                                                  hv=[hv0[:, col].copy()])

    data = Common.BoundaryConditionsData(north=ns(ny + 1), south=ns(2), east=ew(nx + 1), west=ew(2))
    sp = {'north': sponge, 'east': sponge, 'south': sponge, 'west': sponge}
    return dict(eta0=eta0, hu0=hu0, hv0=hv0, H=H, nx=nx, ny=ny, dx=dx, dy=dy, dt=0.0, g=g, f=f, r=0.0,
                courant_number=0.7, boundary_conditions=Common.BoundaryConditions(*codes, spongeCells=sp),
                boundary_conditions_data=data)


def c4_drifter_positions(nx, ny, dx, dy, n, seed):
    """n seeded drifters in the wet interior, away from the sponges."""
    rng = np.random.default_rng(seed + 500)
    x = rng.uniform(0.3 * nx * dx, 0.8 * nx * dx, n)
    y = rng.uniform(0.25 * ny * dy, 0.75 * ny * dy, n)
    return np.stack([x, y], axis=1).astype(np.float32)


def config_coast_island(nx=160, ny=120, seed=5):
    """Land, a drying shelf and an island in one member: dx=dy=500 m, walls,
    f=1e-4; corner depths ramp from -4 m at the west edge to 60 m at the east
    edge; corners at or above the datum are masked land.  Beside the coast
    lie cells a few centimetres to a metre deep (desingularised cells,
    h < 0.1 m); those shallower than 2 m start dry (eta = -Hm) and the
    waves wet and dry them; a masked island of corners sits offshore; a 0.6 m
    Gaussian bump east of the island sends waves across the shelf and around
    the island."""
    rng = np.random.default_rng(seed)
    dx = dy = 500.0
    xc = (np.arange(nx + 5) - 2.0) / nx
    H = np.repeat((-4.0 + 64.0 * xc)[None, :], ny + 5, axis=0)
    H += 0.3 * rng.standard_normal(H.shape)
    H = H.astype(np.float32)
    mask = H <= 0.0
    cy, cx = ny // 2, int(0.45 * nx)
    jj, ii = np.meshgrid(np.arange(ny + 5), np.arange(nx + 5), indexing="ij")
    mask |= ((jj - cy) / 8.0) ** 2 + ((ii - cx) / 5.0) ** 2 <= 1.0
    Hm_ = np.where(mask, 0.0, H)
    wet_c = (~mask).astype(np.float64)
    n_wet = wet_c[:-1, :-1] + wet_c[1:, :-1] + wet_c[:-1, 1:] + wet_c[1:, 1:]
    Hm = (Hm_[:-1, :-1] + Hm_[1:, :-1] + Hm_[:-1, 1:] + Hm_[1:, 1:]) / np.maximum(n_wet, 1)
    jc, ic = np.meshgrid(np.arange(ny + 4), np.arange(nx + 4), indexing="ij")
    eta = 0.6 * np.exp(-((ic - 0.65 * nx) ** 2 + (jc - cy) ** 2) / (2 * 6.0 ** 2))
    eta = np.where((n_wet > 0) & (Hm < 2.0), -Hm, eta).astype(np.float32)
    z = np.zeros_like(eta)
    return dict(eta0=eta, hu0=z, hv0=z.copy(), H=np.ma.masked_array(H, mask), nx=nx, ny=ny, dx=dx, dy=dy, dt=0.0,
                g=9.81, f=1e-4, r=0.0, courant_number=0.7, boundary_conditions=Common.BoundaryConditions(1, 1, 1, 1))

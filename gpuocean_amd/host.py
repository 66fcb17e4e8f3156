"""Host-side preparation of CDKLM16 inputs, shared by the simulator classes.

Each function restates the numpy set-up the reference does in the CDKLM16
constructor before any kernel runs (SWEsimulators/CDKLM16.py:137-239), so the
device sees the same float32 inputs.
"""

import numpy as np

LAND_VALUE = np.float32(1.0e20)  # Bathymetry.mask_value (utils/Common.py:1424)


def subsample_texture(data, factor):
    """Down-sample a texture by ``factor`` with tensor-product linear
    interpolation between texel centres, as the reference does with a degree-1
    RectBivariateSpline (SWEsimulators/CDKLM16.py:181-192)."""
    ny, nx = data.shape
    dx, dy = 1.0 / nx, 1.0 / ny
    xs = np.linspace(0.5 * dx, 1 - 0.5 * dx, nx)
    ys = np.linspace(0.5 * dy, 1 - 0.5 * dy, ny)
    new_nx, new_ny = max(2, nx // factor), max(2, ny // factor)
    ndx, ndy = 1.0 / new_nx, 1.0 / new_ny
    x_new = np.linspace(0.5 * ndx, 1 - 0.5 * ndx, new_nx)
    y_new = np.linspace(0.5 * ndy, 1 - 0.5 * ndy, new_ny)
    data = np.asarray(data, dtype=np.float64)
    # interpolate along x for every row, then along y for every new column
    tmp = np.stack([np.interp(x_new, xs, data[j, :]) for j in range(ny)], axis=0)
    out = np.stack([np.interp(y_new, ys, tmp[:, i]) for i in range(new_nx)], axis=1)
    # RectBivariateSpline(x, y, data.T)(x_new, y_new) has shape (new_nx, new_ny)
    return out.T


def compensate_f(f, coriolis_beta, angle, dx, dy):
    """f at the first interior cell (SWEsimulators/CDKLM16.py:147-149)."""
    north = np.array([np.sin(angle[0, 0]), np.cos(angle[0, 0])])
    ghost = 2
    return f - coriolis_beta * (ghost * dx * north[0] + ghost * dy * north[1])


def coriolis_texture(f32, beta32, angle, nx, ny, dx32, dy32, eta_shape, subsample_f):
    """The coriolis f texture the kernel samples, or None when f is a single
    constant (SWEsimulators/CDKLM16.py:195-223)."""
    if beta32 != 0.0:
        if angle.size != 1:
            raise RuntimeError("non-constant angle cannot be combined with beta plane model (makes no sense)")
        ghost = 2
        x = np.linspace((-ghost + 0.5) * dx32, (nx + ghost - 0.5) * dx32, nx + 2 * ghost)
        y = np.linspace((-ghost + 0.5) * dy32, (ny + ghost - 0.5) * dy32, ny + 2 * ghost)
        x, y = np.meshgrid(x, y)
        n = x * np.sin(angle[0, 0]) + y * np.cos(angle[0, 0])
        coriolis_f = f32 + beta32 * n
    else:
        farr = np.asarray(f32)
        if farr.size == 1:
            return None
        if farr.shape == tuple(eta_shape):
            coriolis_f = np.array(farr, dtype=np.float32)
        else:
            raise RuntimeError("The shape of f should match up with eta or be scalar.")
    if subsample_f and coriolis_f.size >= int(np.prod(eta_shape)):
        coriolis_f = subsample_texture(coriolis_f, subsample_f)
    return np.ascontiguousarray(coriolis_f, dtype=np.float32)


def north_vector(angle):
    """(sin, cos) of a constant angle; the kernels support a constant angle."""
    angle = np.asarray(angle, dtype=np.float32)
    if angle.size != 1:
        raise NotImplementedError("gpuocean_amd supports a constant angle (1x1 array); "
                                  "per-cell angle textures are not implemented")
    a = angle.reshape(-1)[0]
    return np.float32(np.sin(a)), np.float32(np.cos(a)), bool(a == 0.0)


def prepare_bathymetry(H, nx, ny):
    """Bathymetry on corners as float32 with land (masked) corners set to the
    land value (utils/Common.py:1426-1438).  Returns (Hi, use_mask)."""
    use_mask = False
    if np.ma.is_masked(H):
        H = H.copy().filled(LAND_VALUE).astype(np.float32)
        use_mask = True
    H = np.asarray(H)
    assert H.shape == (ny + 5, nx + 5), \
        "Wrong size of bottom bathymetry, should be defined on cell intersections, not cell centers. " + \
        str((H.shape[1], H.shape[0])) + " vs " + str((nx + 5, ny + 5))
    Hi = np.ascontiguousarray(H, dtype=np.float32)
    return Hi, use_mask


def kernel_constant(x):
    """A kernel constant as the reference's JIT #define carries it:
    "{:.12f}f".format(x) parsed as a float literal (CDKLM16.py:242-269)."""
    return np.float32(float("{:.12f}".format(float(x))))


def check_state_shape(a, nx, ny, name):
    a = np.asarray(a)
    assert a.shape == (ny + 4, nx + 4), f"{name} has shape {a.shape}, expected {(ny + 4, nx + 4)}"
    return np.ascontiguousarray(a, dtype=np.float32)

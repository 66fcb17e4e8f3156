"""Observation operators restated on the CPU -- TEST INFRASTRUCTURE ONLY.

Only tests/ use this module, as the checker of the library's observation
operator.  Each function restates one of the reference's per-member loops over
drifters or cells, on host arrays, in the scalar types the reference computes
in (NumPy 1.x promotion: FP64 quotients for the host paths' cell index, FP32
for the CUDA kernel's and for the buoy factor).
"""

import numpy as np


def _numpy_cell(q, n):
    """Interior index numpy would use for int(np.floor(q)) on an axis of n
    cells, or None where it raises IndexError."""
    if not np.isfinite(q):
        return None
    i = int(np.floor(q))
    if i < -n or i >= n:
        return None
    return i + n if i < 0 else i


def observe_particles(states, positions, dx, dy):
    """(members, D, 2) hu, hv per member at floor(x/dx), floor(y/dy) of the
    interior arrays (ensembles/OceanModelEnsemble.py:184-196).  states: list
    of (eta, hu, hv) ghosted arrays; NaN where numpy would raise."""
    out = np.full((len(states), len(positions), 2), np.nan)
    for p, (eta, hu, hv) in enumerate(states):
        ny, nx = hu.shape[0] - 4, hu.shape[1] - 4
        for d in range(len(positions)):
            i = _numpy_cell(np.float64(positions[d][0]) / np.float64(dx), nx)
            j = _numpy_cell(np.float64(positions[d][1]) / np.float64(dy), ny)
            if i is None or j is None:
                continue
            out[p, d, 0] = hu[2:-2, 2:-2][j, i]
            out[p, d, 1] = hv[2:-2, 2:-2][j, i]
    return out


def innovation(states, observations, dx, dy, mean_depth=None):
    """(members, D, 2) observation*c - (hu, hv) with c = (D + eta)/D in FP32
    when mean_depth is given, else 1 (ensembles/BaseOceanStateEnsemble.py:136-153)."""
    out = np.full((len(states), len(observations), 2), np.nan)
    for p, (eta, hu, hv) in enumerate(states):
        ny, nx = hu.shape[0] - 4, hu.shape[1] - 4
        e, u, v = eta[2:-2, 2:-2], hu[2:-2, 2:-2], hv[2:-2, 2:-2]
        for d in range(len(observations)):
            i = _numpy_cell(np.float64(observations[d][0]) / np.float64(dx), nx)
            j = _numpy_cell(np.float64(observations[d][1]) / np.float64(dy), ny)
            if i is None or j is None:
                continue
            c = 1
            if mean_depth is not None:
                md = np.float32(mean_depth)
                c = (md + e[j, i]) / md
            out[p, d, 0] = np.float64(observations[d][2]) * c - u[j, i]
            out[p, d, 1] = np.float64(observations[d][3]) * c - v[j, i]
    return out


def state_samples(eta, hu, hv, cells):
    """(D, 3) eta, hu, hv at (x, y) interior cells (utils/ParticleInfo.py:96-101)."""
    e, u, v = eta[2:-2, 2:-2], hu[2:-2, 2:-2], hv[2:-2, 2:-2]
    cells = np.asarray(cells)
    out = np.zeros((cells.shape[0], 3))
    out[:, 0] = e[cells[:, 1], cells[:, 0]]
    out[:, 1] = u[cells[:, 1], cells[:, 0]]
    out[:, 2] = v[cells[:, 1], cells[:, 0]]
    return out


def buoy_readings(eta, hu, hv, Bm, cells):
    """(B, 2) FP32 hu*H/(H+eta), hv*H/(H+eta) at (x, y) interior cells, H the
    cell-centre depth (utils/Observation.py:154-166)."""
    e, u, v, H = eta[2:-2, 2:-2], hu[2:-2, 2:-2], hv[2:-2, 2:-2], Bm[2:-2, 2:-2]
    out = np.zeros((len(cells), 2), np.float32)
    for b in range(len(cells)):
        x, y = cells[b][0], cells[b][1]
        f = H[y, x] / (H[y, x] + e[y, x])
        out[b, 0] = u[y, x] * f
        out[b, 1] = v[y, x] * f
    return out


def observe_underlying_flow(hu, hv, positions, dx, dy, x0=2, y0=2):
    """(D, 2) float32 hu, hv at cell floor(x/dx) + x0, floor(y/dy) + y0 of the
    ghosted arrays with FP32 quotients (gpu_kernels/observationKernels.cu:59-78);
    NaN where that cell is outside the arrays."""
    out = np.full((len(positions), 2), np.nan, np.float32)
    rows, cols = hu.shape
    for d in range(len(positions)):
        qx = np.floor(np.float32(positions[d][0]) / np.float32(dx))
        qy = np.floor(np.float32(positions[d][1]) / np.float32(dy))
        if not (np.isfinite(qx) and np.isfinite(qy)):
            continue
        i, j = int(qx) + x0, int(qy) + y0
        if 0 <= i < cols and 0 <= j < rows:
            out[d, 0] = hu[j, i]
            out[d, 1] = hv[j, i]
    return out

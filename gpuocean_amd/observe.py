"""Observation operator over many CDKLM16 members in one launch.

``observe_members`` reads every member's latest state at the same (or each
member's own) positions or cells with the library's ``gpo_observe`` and
returns one host array.  It is what the reference does member by member with a
full-state download and numpy indexing:

- ``OceanModelEnsemble.observeParticles`` (ensembles/OceanModelEnsemble.py:168-198): (hu, hv);
- ``BaseOceanStateEnsemble.observeParticles`` (ensembles/BaseOceanStateEnsemble.py:79-159):
  (hu, hv) or the innovation with the optional eta compensation;
- ``ParticleInfo.add_state_sample_from_sim`` (utils/ParticleInfo.py:81-111): (eta, hu, hv) at cells;
- ``Observation.add_observation_from_sim`` (utils/Observation.py:137-180): buoy readings
  (hu, hv) * H / (H + eta).

Cells are found as numpy would find them on the interior arrays
(``floor(x/dx)`` in FP64, negative indices counting from the end); a point
outside raises ``IndexError`` as numpy does.
"""

import ctypes

import numpy as np
import torch

from . import _lib as L

FLOW, STATE, BUOY, INNOVATION = L.OBS_FLOW, L.OBS_STATE, L.OBS_BUOY, L.OBS_INNOVATION


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class _Table:
    """Device copy of the gpo_obs_member table of a list of simulators, kept
    while the simulators' buffers (and the given points) stay the same."""

    def __init__(self):
        self.key = None
        self.dev = None

    def get(self, sims, point_ptrs, cell_ptrs, dxs, dys, device):
        key = tuple((s.h0[0].data_ptr(), s.h0[1].data_ptr(), s.h0[2].data_ptr(), s.h0[0].pitch, s.Bm.data_ptr(),
                     s.Bm.pitch, int(s.nx), int(s.ny), p, c, dx, dy)
                    for s, p, c, dx, dy in zip(sims, point_ptrs, cell_ptrs, dxs, dys))
        if key == self.key:
            return self.dev
        tab = (L.ObsMember * len(sims))()
        for k, (s, p, c, dx, dy) in enumerate(zip(sims, point_ptrs, cell_ptrs, dxs, dys)):
            m = tab[k]
            m.eta, m.hu, m.hv = s.h0[0].data_ptr(), s.h0[1].data_ptr(), s.h0[2].data_ptr()
            m.pitch = s.h0[0].pitch
            m.Bm, m.Bm_pitch = s.Bm.data_ptr(), s.Bm.pitch
            m.nx, m.ny = int(s.nx), int(s.ny)
            m.dx, m.dy = dx, dy
            m.points, m.cells = p, c
        host = torch.frombuffer(bytearray(bytes(tab)), dtype=torch.uint8)
        self.dev = host.to(device)
        self.key = key
        return self.dev


_default_table = _Table()
_streams = {}


def _stream(dev):
    """One observation stream per device, made once."""
    key = (dev.type, dev.index)
    if key not in _streams:
        _streams[key] = torch.cuda.Stream(device=dev)
    return _streams[key]


def observe_members(sims, mode=FLOW, positions=None, cells=None, observations=None, mean_depth=None,
                    dx=None, dy=None, table=None):
    """Values of each simulator's latest state (h0) at the given points.

    positions: (D, 2) metres, shared by all members, or a list of one (D, 2)
    array per member; cells: (D, 2) interior (x, y) cell indices instead.
    observations (mode INNOVATION): (D, 4) rows (x, y, hu, hv); the cells are
    taken from their first two columns.  mean_depth: eta compensation of the
    innovation (None: none).  dx, dy: the cell sizes the positions are divided
    by (default each member's own).  Returns (members, D, k) float64, k = 3
    for STATE, else 2.  With L.OBS_GHOSTED or-ed into mode, cells follow
    observeUnderlyingFlow's rule and points outside the ghosted arrays give
    NaN instead of raising."""
    sims = list(sims)
    if not sims:
        raise ValueError("observe_members: no members")
    dev = sims[0].device
    if (mode & ~L.OBS_GHOSTED) == INNOVATION:
        if observations is None:
            raise ValueError("observe_members: innovation needs the observations")
        observations = np.ascontiguousarray(observations, dtype=np.float64)
        positions = observations[:, :2]
    per_member = isinstance(positions, (list, tuple))
    keep = []  # device buffers alive until the launch has been read back
    if cells is not None:
        c = np.ascontiguousarray(cells, dtype=np.int64)
        if c.ndim != 2 or c.shape[1] != 2:
            raise ValueError("observe_members: cells must be (D, 2)")
        if np.any(c > np.iinfo(np.int32).max) or np.any(c < np.iinfo(np.int32).min):
            raise IndexError("observe_members: cell index out of range")
        ct = torch.from_numpy(c.astype(np.int32)).to(dev)
        keep.append(ct)
        n_points = c.shape[0]
        point_ptrs, cell_ptrs = [None] * len(sims), [ct.data_ptr()] * len(sims)
    elif positions is not None:
        arrs = [np.ascontiguousarray(p, dtype=np.float64) for p in (positions if per_member else [positions])]
        if per_member and len(arrs) != len(sims):
            raise ValueError("observe_members: one position array per member")
        n_points = arrs[0].shape[0]
        for a in arrs:
            if a.ndim != 2 or a.shape[1] < 2 or a.shape[0] != n_points:
                raise ValueError("observe_members: positions must be (D, 2) with the same D for every member")
        pts = [torch.from_numpy(np.ascontiguousarray(a[:, :2])).to(dev) for a in arrs]
        keep.extend(pts)
        point_ptrs = [p.data_ptr() for p in pts] if per_member else [pts[0].data_ptr()] * len(sims)
        cell_ptrs = [None] * len(sims)
    else:
        raise ValueError("observe_members: positions or cells are required")
    k = 3 if (mode & ~L.OBS_GHOSTED) == STATE else 2
    out = torch.empty((len(sims), n_points, k), dtype=torch.float64, device=dev)
    n_bad = torch.zeros(1, dtype=torch.int32, device=dev)
    obs_dev = None
    if observations is not None:
        obs_dev = torch.from_numpy(observations).to(dev)
        keep.append(obs_dev)
    dxs = [float(s.dx if dx is None else dx) for s in sims]
    dys = [float(s.dy if dy is None else dy) for s in sims]
    stream = _stream(dev)
    seen = set()
    for s in sims:   # members may share a stream: wait on each once
        if id(s.gpu_stream) not in seen:
            seen.add(id(s.gpu_stream))
            stream.wait_stream(s.gpu_stream)
    stream.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(stream):
        tab = (table or _default_table).get(sims, point_ptrs, cell_ptrs, dxs, dys, dev)
        md = 0.0 if mean_depth is None else float(mean_depth)
        if mean_depth is not None and not md > 0.0:
            raise ValueError("observe_members: mean_depth must be positive")
        L.check(L.lib().gpo_observe(_vp(tab), len(sims), n_points, int(mode), _vp(obs_dev), md, _vp(out),
                                    _vp(n_bad), ctypes.c_void_p(stream.cuda_stream)), "gpo_observe")
        host = out.cpu().numpy()
        bad = int(n_bad.item())
    if bad and not (mode & L.OBS_GHOSTED):
        raise IndexError(f"observe_members: {bad} point(s) outside the interior domain")
    return host

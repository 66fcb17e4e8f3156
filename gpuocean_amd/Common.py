"""Host-side containers of the CDKLM16 path: boundary conditions and their data.

Mirrors the reference's ``utils/Common.py`` classes that the CDKLM16 surface
takes as arguments (``BoundaryConditions`` at utils/Common.py:920-1056,
``SingleBoundaryConditionData`` at :1063-1089, ``BoundaryConditionsData`` at
:1091-1116) with the same constructor arguments, codes and assertions, so
scripts written for the reference construct them unchanged.

``CUDAContext`` (utils/Common.py:288) becomes :class:`DeviceContext`: it only
selects the HIP device; kernels are compiled ahead of time for gfx950.
"""

import numpy as np

WALL = 1
PERIODIC = 2
FLOW_RELAXATION = 3
OPEN_LINEAR = 4


class BoundaryConditions:
    """Boundary condition codes per side: 1 wall, 2 periodic, 3 flow relaxation
    scheme (FRS sponge), 4 open linear interpolation (utils/Common.py:920-952)."""

    def __init__(self, north=1, east=1, south=1, west=1,
                 spongeCells=None):
        if spongeCells is None:
            spongeCells = {'north': 0, 'south': 0, 'east': 0, 'west': 0}
        self.north = np.int32(north)
        self.east = np.int32(east)
        self.south = np.int32(south)
        self.west = np.int32(west)
        self.spongeCells = dict(spongeCells)
        for key in self.spongeCells.keys():
            self.spongeCells[key] = np.int32(self.spongeCells[key])

        assert not ((self.north == 2 or self.south == 2) and (self.north != self.south)), \
            'The given periodic boundary conditions are not periodically (north/south)'
        assert not ((self.east == 2 or self.west == 2) and (self.east != self.west)), \
            'The given periodic boundary conditions are not periodically (east/west)'

    def get(self):
        raise NotImplementedError("get() implementation changed! Change calling code")

    def getBCDict(self):
        return {'north': self.north, 'south': self.south, 'east': self.east, 'west': self.west}

    def getSponge(self):
        return self.spongeCells

    def isDefault(self):
        # The reference tests east twice (utils/Common.py:967-971); kept as is.
        return (self.north == 1 and self.east == 1 and self.south == 1 and self.east == 1)

    def isSponge(self):
        return any(v in (3, 4) for v in (self.north, self.east, self.south, self.west))

    def isPeriodicNorthSouth(self):
        return bool(self.north == 2 and self.south == 2)

    def isPeriodicEastWest(self):
        return bool(self.east == 2 and self.west == 2)

    def isPeriodic(self):
        return self.isPeriodicEastWest() and self.isPeriodicNorthSouth()

    @staticmethod
    def _toString(cond):
        return {1: "Wall", 2: "Periodic", 3: "Flow_Relaxation_Scheme",
                4: "Open_Linear_Interpolation"}.get(int(cond), "Invalid_:|")

    def __str__(self):
        # sponge widths as plain integers, as NumPy 1.x printed them, so the
        # text parses back (fromstring) and matches files the reference wrote
        return ("north: " + self._toString(self.north) + ", east: " + self._toString(self.east)
                + ", south: " + self._toString(self.south) + ", west: " + self._toString(self.west)
                + ", spongeCells: " + str({k: int(v) for k, v in self.spongeCells.items()}))

    @classmethod
    def fromstring(cls, bc_string):
        """Inverse of ``__str__`` (utils/Common.py:1012-1044)."""
        def keyword_to_cond(key):
            return {'Wall': 1, 'Periodic': 2, 'Flow_Relaxation_Scheme': 3,
                    'Open_Linear_Interpolation': 4}.get(key, -1)
        clean = bc_string
        for ch in [',', '}', '{', '[', ']', ':', "'"]:
            clean = clean.replace(ch, '')
        arr = clean.split(' ')
        north, east, south, west = (keyword_to_cond(arr[i]) for i in (1, 3, 5, 7))
        if len(arr) > 13:
            sponge = {arr[9]: int(arr[10]), arr[11]: int(arr[12]),
                      arr[13]: int(arr[14]), arr[15]: int(arr[16])}
        else:
            sponge = {'north': int(arr[9]), 'east': int(arr[10]),
                      'south': int(arr[11]), 'west': int(arr[12])}
        return cls(north=north, east=east, south=south, west=west, spongeCells=sponge)

    def packed(self):
        """North<<24 | South<<16 | East<<8 | West, as CDKLM16.callKernel packs
        it (SWEsimulators/CDKLM16.py:695-699)."""
        return int((int(self.north) & 0xFF) << 24 | (int(self.south) & 0xFF) << 16
                   | (int(self.east) & 0xFF) << 8 | (int(self.west) & 0xFF))


class SingleBoundaryConditionData:
    """External solution on one boundary over time (utils/Common.py:1063-1089)."""

    def __init__(self, h=None, hu=None, hv=None):
        self.h = [np.zeros(2, dtype=np.float32, order='C')]
        self.hu = [np.zeros(2, dtype=np.float32, order='C')]
        self.hv = [np.zeros(2, dtype=np.float32, order='C')]
        self.numSteps = 1
        self.shape = self.h[0].shape
        if h is not None:
            self.shape = h[0].shape
            self.numSteps = len(h)
            self.h, self.hu, self.hv = h, hu, hv
            for i in range(len(h)):
                assert h[i].shape == self.shape, str(self.shape) + " vs " + str(h[i].shape)
                assert hu[i].shape == self.shape, str(self.shape) + " vs " + str(hu[i].shape)
                assert hv[i].shape == self.shape, str(self.shape) + " vs " + str(hv[i].shape)
                assert h[i].dtype == 'float32', "h data needs to be of type np.float32"
                assert hu[i].dtype == 'float32', "hu data needs to be of type np.float32"
                assert hv[i].dtype == 'float32', "hv data needs to be of type np.float32"

    def __str__(self):
        return str(self.numSteps) + " steps, each " + str(self.shape)


class BoundaryConditionsData:
    """External solution for all boundaries over time (utils/Common.py:1091-1116)."""

    def __init__(self, t=None, north=None, south=None, east=None, west=None):
        north = SingleBoundaryConditionData() if north is None else north
        south = SingleBoundaryConditionData() if south is None else south
        east = SingleBoundaryConditionData() if east is None else east
        west = SingleBoundaryConditionData() if west is None else west
        self.t = [0]
        self.numSteps = 1
        self.north, self.south, self.east, self.west = north, south, east, west
        if t is not None:
            self.t = t
            self.numSteps = len(t)
        for data in [north, south, east, west]:
            assert data.numSteps == self.numSteps, \
                "Wrong timesteps " + str(data.numSteps) + " vs " + str(self.numSteps)
        assert north.h[0].shape == south.h[0].shape, \
            "Wrong shape of north vs south " + str(north.h[0].shape) + " vs " + str(south.h[0].shape)
        assert east.h[0].shape == west.h[0].shape, \
            "Wrong shape of east vs west " + str(east.h[0].shape) + " vs " + str(west.h[0].shape)

    def __str__(self):
        return ("Steps=" + str(self.numSteps) + ", [north=" + str(self.north) + "]"
                + ", [south=" + str(self.south) + "]" + ", [east=" + str(self.east) + "]"
                + ", [west=" + str(self.west) + "]")


def pack_bc_data_ns(bc_data, t_index):
    """[2, nx, 3] array (south row, north row) of (h, hu, hv), as
    BoundaryConditionsArakawaA.update_bc_values packs it (utils/Common.py:1225-1283)."""
    def row(d):
        h, hu, hv = (np.squeeze(a[t_index]) for a in (d.h, d.hu, d.hv))
        assert len(h.shape) == 1, "NS-data must be one row"
        return np.ascontiguousarray(np.vstack((h, hu, hv)).T.reshape(1, h.shape[0], 3))
    return np.ascontiguousarray(np.vstack((row(bc_data.south), row(bc_data.north))), dtype=np.float32)


def pack_bc_data_ew(bc_data, t_index):
    """[ny, 2, 3] array (west column, east column) of (h, hu, hv)."""
    def col(d):
        h, hu, hv = (np.squeeze(a[t_index]) for a in (d.h, d.hu, d.hv))
        assert len(h.shape) == 1, "EW-data must be one column"
        return np.ascontiguousarray(np.vstack((h, hu, hv)).T.reshape(h.shape[0], 1, 3))
    return np.ascontiguousarray(np.hstack((col(bc_data.west), col(bc_data.east))), dtype=np.float32)


class BCTimeline:
    """Host bookkeeping of the FRS exterior solution in time: which two time
    levels are current and the linear weight between them
    (utils/Common.py:1193-1308)."""

    def __init__(self, bc_data):
        self.bc_data = bc_data
        self.timestamps = [None, None]
        self.bc_t = np.float32(0.0)

    def update(self, t):
        """Returns (changed_current, changed_next, t0_index, t1_index)."""
        tt = self.bc_data.t
        t_max_index = len(tt) - 1
        t0_index = max(0, int(np.searchsorted(tt, t)) - 1)
        t1_index = min(t_max_index, int(np.searchsorted(tt, t)))
        new_t0, new_t1 = tt[t0_index], tt[t1_index]
        old_t0, old_t1 = self.timestamps
        self.timestamps = [new_t0, new_t1]
        elapsed = t - new_t0
        interval = max(1.0e-10, new_t1 - new_t0)
        self.bc_t = np.float32(max(0.0, min(1.0, elapsed / interval)))
        return (new_t0 != old_t0), (new_t1 != old_t1), t0_index, t1_index


class DeviceContext:
    """Stand-in for the reference's ``Common.CUDAContext`` (utils/Common.py:288):
    selects the HIP device that the simulators use."""

    def __init__(self, device=0, context_flags=None, use_cache=True, autotuning=True):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("gpuocean_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device('cuda', device)
        torch.cuda.set_device(self.device)

    def synchronize(self):
        import torch
        torch.cuda.synchronize(self.device)

    def __str__(self):
        import torch
        return "DeviceContext id " + str(torch.cuda.get_device_name(self.device))


# Drop-in alias for scripts that construct ``Common.CUDAContext()``.
CUDAContext = DeviceContext

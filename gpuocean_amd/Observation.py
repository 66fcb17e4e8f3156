"""Drifter and buoy observations of an ocean simulation (utils/Observation.py).

A pandas DataFrame holds, per observation time, the drifter positions and
their observation errors, and for static buoys the buoy readings of hu and hv
(with the depth known and eta unknown) and their errors; the columns and rows
are the reference's, so ``to_pickle`` files interchange.  Buoy readings come
from the library's observation operator on the simulator's state
(``observe.observe_members``) rather than from a full download.

``get_observation`` turns two drifter observations ``observation_interval``
rows apart into the velocity observation ``[x, y, hu, hv]`` of the later one.
Its arithmetic follows the NumPy 1.x promotion rules the reference was written
for (it calls ``np.int``, which NumPy 1.24 removed): the drifter velocity and
its product with the water depth are FP32, the periodic-boundary candidates
and the observation error FP64, the depth lookup index an FP64 quotient.
"""

import numpy as np
import pandas as pd

from .DataAssimilationUtils import ObservationType
from . import observe


def _strong(v):
    """A Python number as NumPy 1.x scalar arithmetic saw it (FP64 or int64);
    NumPy scalars unchanged."""
    if isinstance(v, (np.generic, np.ndarray)):
        return v
    return np.float64(v) if isinstance(v, float) else np.int64(v)


class Observation:
    """Creates, stores and reads observations (utils/Observation.py:37-113)."""

    def __init__(self, observation_type=ObservationType.UnderlyingFlow, domain_size_x=None, domain_size_y=None,
                 nx=None, ny=None, observation_variance=0.0, observation_interval=1, land_mask=None):
        self.observation_type = observation_type
        self._check_observation_type()

        self.time_key = 'time'
        self.drifter_positions_key = 'drifter_positions'
        self.drifter_obs_errors_key = 'drifter_obs_errors'
        self.buoy_observations_key = 'buoy_observations'
        self.buoy_positions_key = 'buoy_positions'
        self.buoy_obs_errors_key = 'buoy_obs_errors'
        self.df_keys = [self.time_key, self.drifter_positions_key, self.drifter_obs_errors_key,
                        self.buoy_observations_key, self.buoy_positions_key, self.buoy_obs_errors_key]
        self.obs_df = pd.DataFrame(columns=self.df_keys)

        self.register_buoys = False
        self.buoy_indices = None
        self.buoy_positions = None
        self.read_buoy = None

        self.drifterSet = None
        self.observationInterval = observation_interval
        self.obs_var = observation_variance
        self.obs_stddev = np.sqrt(observation_variance)
        self.land_mask = land_mask

        if observation_type == ObservationType.StaticBuoys:
            assert nx is not None, 'nx must be provided if observation_type is StaticBuoys'
            assert ny is not None, 'ny must be provided if observation_type is StaticBuoys'
            assert domain_size_x is not None, 'domain_size_x must be provided if observation_type is StaticBuoys'
            assert domain_size_y is not None, 'domain_size_y must be provided if observation_type is StaticBuoys'
        self.domain_size_x = domain_size_x
        self.domain_size_y = domain_size_y
        self.nx = nx
        self.ny = ny

    # ------------------------------------------------------------------ counts

    def get_num_observations(self):
        """Rows (observation times) stored."""
        return self.obs_df[self.time_key].count()

    def get_num_drifters(self, applyDrifterSet=True, ignoreBuoys=False):
        """Buoys read, or drifters in the (selected) set."""
        if self.observation_type == ObservationType.StaticBuoys and not ignoreBuoys:
            return np.sum(self.read_buoy)
        if self.drifterSet is not None and applyDrifterSet:
            return len(self.drifterSet)
        return self.obs_df.iloc[0][self.drifter_positions_key].shape[0]

    # ------------------------------------------------------------------ adding

    def _new_row(self, t, pos, drifter_obs_errors, buoy_observations=None, buoy_positions=None,
                 buoy_obs_errors=None):
        rounded_t = round(t)
        index = self.get_num_observations()
        if index != 0:
            assert self.obs_df[self.obs_df[self.time_key] == rounded_t].time.count() == 0, \
                "Observation for time " + str(rounded_t) + " already exists in DataFrame"
        self.obs_df.loc[index] = {self.time_key: rounded_t, self.drifter_positions_key: pos,
                                  self.buoy_observations_key: buoy_observations,
                                  self.buoy_positions_key: buoy_positions,
                                  self.buoy_obs_errors_key: buoy_obs_errors,
                                  self.drifter_obs_errors_key: drifter_obs_errors}

    def add_observation_from_sim(self, sim):
        """Current drifter positions of `sim`, and with buoys registered their
        readings hu*H/(H+eta), hv*H/(H+eta) at the buoy cells
        (utils/Observation.py:137-180).  Draws the buoy errors, then the
        drifter errors, from np.random."""
        if self.get_num_observations() != 0:
            rounded_t = round(sim.t)
            assert self.obs_df[self.obs_df[self.time_key] == rounded_t].time.count() == 0, \
                "Observation for time " + str(rounded_t) + " already exists in DataFrame"
        buoy_positions = buoy_observations = buoy_obs_errors = None
        if self.register_buoys:
            if self.get_num_observations() == 0:
                buoy_positions = self.buoy_positions
            reading = observe.observe_members([sim], observe.BUOY, cells=self.buoy_indices)[0]
            buoy_observations = reading.astype(self.buoy_positions.dtype)
            buoy_obs_errors = np.random.normal(size=buoy_observations.shape)
        pos = sim.drifters.getDrifterPositions()
        drifter_obs_errors = np.random.normal(size=pos.shape)
        self._new_row(sim.t, pos, drifter_obs_errors, buoy_observations, buoy_positions, buoy_obs_errors)

    def add_observation_from_drifters(self, drifters, t):
        """Positions of a drifter collection at time t, no errors
        (utils/Observation.py:182-203).  Not for buoys."""
        assert not self.register_buoys, "Only floating drifters supported in 'add_observation_from_drifter'"
        self._new_row(t, drifters.getDrifterPositions(), None)

    def add_observations_from_arrays(self, t, x, y):
        """Fills an empty frame from times t (T,) and positions x, y (D, T)
        (utils/Observation.py:205-230)."""
        assert self.get_num_observations() == 0, \
            'This function can only be called when the Observation data frame is empty'
        for i in range(len(t)):
            positions = np.array([x[:, i], y[:, i]]).transpose()
            errors = np.random.normal(size=positions.shape)
            self.obs_df.loc[i] = {self.time_key: t[i], self.drifter_positions_key: positions,
                                  self.buoy_observations_key: None, self.buoy_positions_key: None,
                                  self.buoy_obs_errors_key: None, self.drifter_obs_errors_key: errors}

    def add_observation_from_mldrifters(self, t, mldrifters, ensemble_member):
        """Positions of one ensemble member of a multi-level drifter set
        (anything with getDrifterPositionsForEnsembleMember) at time t
        (utils/Observation.py:232-254)."""
        if self.get_num_observations() != 0:
            rounded_t = round(t)
            assert self.obs_df[self.obs_df[self.time_key] == rounded_t].time.count() == 0, \
                "Observation for time " + str(rounded_t) + " already exists in DataFrame"
        pos = mldrifters.getDrifterPositionsForEnsembleMember(ensemble_member)
        self._new_row(t, pos, np.random.normal(size=pos.shape))

    # ------------------------------------------------------------ configuration

    def setDrifterSet(self, drifterSet):
        """Only these drifter indices are read back (utils/Observation.py:260-272)."""
        assert type(drifterSet) is list, 'drifterSet is required to be a list, but is ' + str(type(drifterSet))
        assert min(drifterSet) >= 0, 'drifterSet contains at least one negative drifter id'
        assert max(drifterSet) < self.get_num_drifters(), 'drifterSet contains indices that are out-of-range'
        self.drifterSet = drifterSet

    def setObservationInterval(self, interval):
        self.observationInterval = interval

    def setBuoyCells(self, buoy_indices):
        """Buoys at the centres of the given (x, y) cells (utils/Observation.py:277-290)."""
        self.buoy_indices = buoy_indices
        pos = buoy_indices.copy().astype(np.float32)
        pos[:, 0] = (pos[:, 0] + np.float32(0.5)) * np.float32(self._cell_size(0))
        pos[:, 1] = (pos[:, 1] + np.float32(0.5)) * np.float32(self._cell_size(1))
        self.buoy_positions = pos
        self.read_buoy = [True] * self.buoy_indices.shape[0]
        self.register_buoys = True

    def setBuoyCellsByFrequency(self, frequency_x, frequency_y, avoid_boundary=False):
        """A buoy every frequency_x, frequency_y cells, skipping land
        (utils/Observation.py:292-314)."""
        x0 = int(frequency_x / 2) if avoid_boundary else 0
        y0 = int(frequency_y / 2) if avoid_boundary else 0
        cells = [[x, y] for y in range(y0, self.ny, frequency_y) for x in range(x0, self.nx, frequency_x)
                 if self.land_mask is None or not self.land_mask[y, x]]
        self.setBuoyCells(np.array(cells, dtype=np.int32))

    def setBuoyReadingArea(self, area='all'):
        """Read only the buoys of an area: 'all', 'south', 'west' or 'sparse'
        (utils/Observation.py:316-331)."""
        if self.observation_type != ObservationType.StaticBuoys:
            return
        idx = self.buoy_indices
        if area == 'south':
            self.read_buoy = [bool(v) for v in idx[:, 1] < self.ny / 2]
        elif area == 'west':
            self.read_buoy = [bool(v) for v in idx[:, 0] < self.nx / 2]
        elif area == 'all':
            self.read_buoy = [True] * idx.shape[0]
        elif area == 'sparse':
            self.read_buoy = [bool(a and b) for a, b in zip(idx[:, 0] % 50 == 0, idx[:, 1] % 50 == 0)]
        else:
            assert area == 'all', 'Invalid area. Must be all, south or west'

    def setBuoySet(self, buoySet):
        assert self.observation_type == ObservationType.StaticBuoys
        self.read_buoy = [False] * self.buoy_indices.shape[0]
        for i in buoySet:
            self.read_buoy[i] = True

    def _cell_size(self, axis):
        """Domain size over cell count in FP64, as NumPy 1.x divided a NumPy
        or Python scalar by an integer."""
        size, n = (self.domain_size_x, self.nx) if axis == 0 else (self.domain_size_y, self.ny)
        return np.float64(size) / np.float64(n)

    # ---------------------------------------------------------------- files

    def to_pickle(self, path):
        """Writes the frame with pandas' pickle format."""
        self.obs_df.to_pickle(path)

    def read_pickle(self, path):
        """Reads a frame written by to_pickle.  pandas unpickles the file, which
        can run code from it: read only files of trusted origin."""
        self.obs_df = pd.read_pickle(path)
        if self.observation_type == ObservationType.StaticBuoys:
            self.buoy_positions = self.obs_df.iloc[0][self.buoy_positions_key].copy()
            idx = self.buoy_positions.copy()
            idx[:, 0] = np.floor(idx[:, 0] / idx.dtype.type(self._cell_size(0)))
            idx[:, 1] = np.floor(idx[:, 1] / idx.dtype.type(self._cell_size(1)))
            self.buoy_indices = idx.astype(np.int32)
            self.read_buoy = [True] * self.buoy_indices.shape[0]

    # ---------------------------------------------------------------- reading

    def _check_observation_type(self):
        assert self.observation_type in (ObservationType.UnderlyingFlow, ObservationType.StaticBuoys), \
            "Only UnderlyingFlow and StaticBuoys ObservationType are supported at the moment."

    def _check_df_at_given_time(self, rounded_t):
        n = self.obs_df[self.obs_df[self.time_key] == rounded_t].time.count()
        assert n > 0, "Observation for time " + str(rounded_t) + " does not exists in DataFrame"
        assert n < 2, "Observation for time " + str(rounded_t) + " has multiple entries in DataFrame"

    def _row(self, t):
        rounded_t = round(t)
        self._check_df_at_given_time(rounded_t)
        return self.obs_df[self.obs_df[self.time_key] == rounded_t].index.values[0]

    def get_observation_times(self):
        """Times with a velocity observation: every observation_interval-th
        row after the first."""
        if self.get_num_observations() < 2:
            return np.array([])
        return self.obs_df.time.values[::self.observationInterval][1:].copy()

    def get_drifter_position(self, t, applyDrifterSet=True, ignoreBuoys=False):
        """Drifter (or read buoy) positions at time t, [[x_1, y_1], ...]."""
        index = self._row(t)
        if self.observation_type == ObservationType.StaticBuoys and not ignoreBuoys:
            return self.buoy_positions.copy()[self.read_buoy, :]
        pos = self.obs_df.iloc[index][self.drifter_positions_key]
        if applyDrifterSet and self.drifterSet is not None:
            return pos[self.drifterSet, :].copy()
        return pos.copy()

    def get_observation(self, t, waterDepth=None, Hm=None):
        """[[x, y, hu, hv], ...] at time t: for buoys their readings, for
        drifters the velocity since the observation observationInterval rows
        earlier times the water depth (waterDepth, or Hm at each drifter's
        cell), the shorter of the direct and the across-the-boundary
        displacement on a domain of known size, plus the stored errors times
        the observation standard deviation (utils/Observation.py:425-536)."""
        assert (waterDepth is not None) or (Hm is not None), \
            'Observation.get_observation() requires either waterDepth or Hm as input argument. Now, neither is provided'
        index = self._row(t)
        self._check_observation_type()
        assert index > self.observationInterval - 1, "Observation can not be made this early in the DataFrame."

        if self.observation_type == ObservationType.StaticBuoys:
            row = self.obs_df.iloc[index]
            observation = np.zeros((self.get_num_drifters(), 4))
            observation[:, :2] = self.buoy_positions.copy()[self.read_buoy, :]
            observation[:, 2:] = row[self.buoy_observations_key][self.read_buoy, :]
            if self.buoy_obs_errors_key in self.obs_df.columns:
                err = row[self.buoy_obs_errors_key]
                if err is not None:
                    observation[:, 2:] += err[self.read_buoy, :] * self.obs_stddev
            return observation

        prev = index - self.observationInterval
        dt = self.obs_df.iloc[index][self.time_key] - self.obs_df.iloc[prev][self.time_key]
        cur = self.obs_df.iloc[index][self.drifter_positions_key]
        old = self.obs_df.iloc[prev][self.drifter_positions_key]
        if self.drifterSet is not None:
            cur, old = cur[self.drifterSet, :], old[self.drifterSet, :]
        D = old.shape[0]
        diff = cur - old
        # array / integer scalar: the array's type (FP32 for simulator drifters)
        u_v = diff / np.asarray(dt).astype(diff.dtype)

        observation = np.zeros((D, 4))
        observation[:, :2] = cur
        if Hm is not None:
            cdx, cdy = self._cell_size(0), self._cell_size(1)
            ix = np.floor(np.asarray(cur[:, 0], dtype=np.float64) / cdx).astype(np.int64)
            iy = np.floor(np.asarray(cur[:, 1], dtype=np.float64) / cdy).astype(np.int64)
            depths = np.asarray(Hm)[iy, ix].astype(np.float64)
        else:
            depths = np.ones(D) * waterDepth
        # array * FP64 scalar: the array's type
        observation[:, 2:] = u_v * depths.astype(u_v.dtype)[:, None]

        dtt = _strong(dt)
        for axis, size in ((0, self.domain_size_x), (1, self.domain_size_y)):
            if not size:
                continue
            size = _strong(size)
            for d in range(D):
                step = cur[d, axis] - old[d, axis]
                H = depths[d]
                vp = (step + size) * H / dtt
                vm = (step - size) * H / dtt
                if abs(vp) < abs(observation[d, 2 + axis]):
                    observation[d, 2 + axis] = vp
                if abs(vm) < abs(observation[d, 2 + axis]):
                    observation[d, 2 + axis] = vm

        if self.drifter_obs_errors_key in self.obs_df.columns:
            err = self.obs_df.iloc[index][self.drifter_obs_errors_key]
            if err is not None:
                if self.drifterSet is not None:
                    err = err[self.drifterSet, :]
                observation[:, 2:] += err * self.obs_stddev
        return observation

    @staticmethod
    def _detect_jump(pos0, pos1, jump_limit=100000):
        return np.sqrt((pos1[0] - pos0[0]) ** 2 + (pos1[1] - pos0[1]) ** 2) > jump_limit

    def get_drifter_path(self, drifter_id, start_t, end_t, in_km=True, keepDomainSize=True,
                         assume_no_boundary_trouble=False):
        """Path pieces of one drifter between start_t and end_t, split where it
        jumps across a periodic boundary (keepDomainSize) or continued past it
        (utils/Observation.py:545-628)."""
        times = self.obs_df.time.values[::self.observationInterval].copy()
        first = np.searchsorted(times, start_t)
        last = min(np.searchsorted(times, end_t) + 1, len(times))
        total = last - first
        positions = np.stack(self.obs_df[self.drifter_positions_key].values[::self.observationInterval].copy(),
                             axis=0)[:, drifter_id, :]
        paths = []
        if assume_no_boundary_trouble:
            paths.append(np.array(positions[first:last, :], dtype=np.float64))
        else:
            path = np.zeros((total, 2))
            n = 0
            shift = np.array([0, 0])
            for i in range(first, last):
                if times[i] < start_t or times[i] > end_t:
                    continue
                p = positions[i, :]
                if n > 0 and self._detect_jump(path[n - 1, :], p + shift):
                    if keepDomainSize:
                        paths.append(path[:n, :])
                        n = 0
                        path = np.zeros((total, 2))
                    else:
                        for a, size in ((0, self.domain_size_x), (1, self.domain_size_y)):
                            delta = p[a] + shift[a] - path[n - 1, a]
                            if min(abs(delta - size), abs(delta + size)) < abs(delta):
                                shift[a] = shift[a] - size if abs(delta - size) < abs(delta + size) else shift[a] + size
                path[n, :] = p + shift
                n += 1
            paths.append(path[:n, :])
        if in_km:
            paths = [p / 1000 for p in paths]
        return paths

"""Ocean state samples of ensemble members at drifter cells (utils/ParticleInfo.py).

A pandas DataFrame keeps, per sample time, (eta, hu, hv) of a member at the
given drifter cells and at an optional fixed set of extra cells, with the
reference's columns.  The samples come from the library's observation
operator, which reads only those cells, instead of a full-state download.
"""

import numpy as np
import pandas as pd

from . import observe


class ParticleInfo:
    """State samples of one ensemble member (utils/ParticleInfo.py:39-53)."""

    def __init__(self):
        self.columns = ('time', 'state_under_drifter', 'extra_states')
        self.state_df = pd.DataFrame(columns=self.columns)
        self.extraCells = None

    def get_num_samples(self):
        return self.state_df[self.columns[0]].count()

    def get_num_drifters(self):
        return self.state_df.iloc[0][self.columns[1]].shape[0]

    def get_num_extra_cells(self):
        return 0 if self.extraCells is None else self.extraCells.shape[0]

    def _add(self, t, state_sample, extra_sample):
        rounded_t = round(t)
        index = self.get_num_samples()
        if index != 0:
            assert self.state_df[self.state_df[self.columns[0]] == rounded_t].time.count() == 0, \
                "State sample for time " + str(rounded_t) + " already exists in DataFrame"
        self.state_df.loc[index] = {self.columns[0]: rounded_t, self.columns[1]: state_sample,
                                    self.columns[2]: extra_sample}

    def add_state_sample_from_sim(self, sim, drifter_cells):
        """(eta, hu, hv) of `sim` at drifter_cells ((D, 2) interior (x, y)
        cells) and at the extra cells (utils/ParticleInfo.py:81-111)."""
        drifter_cells = np.asarray(drifter_cells)
        cells = drifter_cells if self.extraCells is None else np.concatenate(
            [drifter_cells, np.asarray(self.extraCells)], axis=0)
        values = observe.observe_members([sim], observe.STATE, cells=cells)[0]
        D = drifter_cells.shape[0]
        extra = None if self.extraCells is None else values[D:].copy()
        self._add(sim.t, values[:D].copy(), extra)

    def setExtraCells(self, extraCells):
        """A fixed set of (x, y) cells sampled in addition to the drifter cells."""
        self.extraCells = extraCells

    def usePredefinedExtraCellsLovese(self):
        """The reference's fixed cells for its Lofoten case (utils/ParticleInfo.py:124-144)."""
        self.extraCells = np.array([[50, 110], [110, 110], [170, 110], [290, 115], [410, 115], [470, 120],
                                    [50, 155], [200, 165], [290, 165], [350, 155], [410, 155], [470, 155],
                                    [50, 200], [110, 200], [170, 200], [230, 200], [290, 200], [350, 200],
                                    [410, 200], [470, 200]])

    def usePredefinedExtraCells(self):
        """The reference's fixed cells along its April 2019 drifter paths
        (utils/ParticleInfo.py:146-166)."""
        self.extraCells = np.array([[423, 25], [381, 27], [185, 48], [69, 157], [288, 132], [331, 177],
                                    [205, 201], [442, 234], [93, 11], [462, 0], [202, 135], [405, 135],
                                    [315, 229]])

    def to_pickle(self, path):
        """Writes the frame with pandas' pickle format."""
        self.state_df.to_pickle(path)

    def read_pickle(self, path):
        """Reads a frame written by to_pickle.  pandas unpickles the file, which
        can run code from it: read only files of trusted origin."""
        self.state_df = pd.read_pickle(path)

    def _check_df_at_given_time(self, rounded_t):
        n = self.state_df[self.state_df[self.columns[0]] == rounded_t].time.count()
        assert n > 0, "State sample for time " + str(rounded_t) + " does not exists in DataFrame"
        assert n < 2, "State sample for time " + str(rounded_t) + " has multiple entries in DataFrame"

    def _row(self, t):
        rounded_t = round(t)
        self._check_df_at_given_time(rounded_t)
        return self.state_df[self.state_df[self.columns[0]] == rounded_t].index.values[0]

    def get_sample_times(self):
        if self.get_num_samples() < 1:
            return np.array([])
        return self.state_df.time.values.copy()

    def get_state_samples(self, t):
        """[[eta_1, hu_1, hv_1], ...] at the drifter cells at time t."""
        return self.state_df.iloc[self._row(t)][self.columns[1]].copy()

    def get_extra_sample(self, t):
        return self.state_df.iloc[self._row(t)][self.columns[2]].copy()

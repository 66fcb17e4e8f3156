"""Wind stress and atmospheric pressure forcing containers.

Same constructor arguments and fields as the reference's
``utils/WindStress.py:37-123`` and ``utils/AtmosphericPressure.py:25-82``, which
the CDKLM16 constructor takes as ``wind=`` and ``atmospheric_pressure=``.
:class:`DeviceForcing` keeps the two time levels a simulator samples on the
device and the blend weight between them (SWEsimulators/Simulator.py:183-288).
"""

import numpy as np


class WindStress:
    """Wind and/or wind-stress time series on a texture grid."""

    def __init__(self, source_filename=None, t=None, wind_u=None, wind_v=None, stress_u=None, stress_v=None):
        self.source_filename = source_filename
        self.t = [0]
        self.wind_u = [np.zeros((1, 1), dtype=np.float32, order='C')]
        self.wind_v = [np.zeros((1, 1), dtype=np.float32, order='C')]
        self.stress_u = None
        self.stress_v = None
        self.numWindSteps = 1
        if t is not None:
            assert wind_u is not None or stress_u is not None, \
                "any information in x direction missing, provide wind or wind stress"
            assert wind_v is not None or stress_v is not None, \
                "any information in y direction missing, provide wind or wind stress"
            if wind_u is not None or wind_v is not None:
                assert len(t) == len(wind_u), str(len(t)) + " vs " + str(len(wind_u))
                assert len(t) == len(wind_v), str(len(t)) + " vs " + str(len(wind_v))
            if stress_u is not None or stress_v is not None:
                assert len(t) == len(stress_u), str(len(t)) + " vs " + str(len(stress_u))
                assert len(t) == len(stress_v), str(len(t)) + " vs " + str(len(stress_v))
            self.numWindSteps = len(t)
            for i in range(self.numWindSteps):
                if wind_u is not None:
                    assert wind_u[i].dtype == 'float32', "Wind data needs to be of type np.float32"
                    assert wind_v[i].dtype == 'float32', "Wind data needs to be of type np.float32"
                if stress_u is not None:
                    assert stress_u[i].dtype == 'float32', "Wind data needs to be of type np.float32"
                    assert stress_v[i].dtype == 'float32', "Wind data needs to be of type np.float32"
            self.t = t
            self.wind_u = wind_u
            self.wind_v = wind_v
            self.stress_u = stress_u
            self.stress_v = stress_v

    def compute_wind_stress_from_wind(self, wind_u=None, wind_v=None):
        """tau = rho_a * C_drag(|W|) * |W| W / rho_w with the Engedahl (1995) drag
        coefficient (utils/WindStress.py:88-123)."""
        if wind_u is None:
            wind_u = self.wind_u
        if wind_v is None:
            wind_v = self.wind_v
        self.stress_u, self.stress_v = [], []
        for i in range(self.numWindSteps):
            speed = np.sqrt(np.power(wind_u[i], 2) + np.power(wind_v[i], 2))
            c_drag = np.where(speed < 11, 0.0012, 0.00049 + 0.000065 * speed)
            rho_a, rho_w = 1.225, 1025
            stress = c_drag * speed * rho_a / rho_w
            self.stress_u.append(stress * wind_u[i])
            self.stress_v.append(stress * wind_v[i])

    def is_zero(self):
        return all(not np.any(np.asarray(a)) for a in list(self.stress_u) + list(self.stress_v))


class AtmosphericPressure:
    """Atmospheric pressure time series, stored relative to ``shiftValue``."""

    def __init__(self, t=None, P=None, shiftValue=100000):
        self.t = [0]
        self.P = [np.zeros((1, 1), dtype=np.float32, order='C')]
        self.numAtmPressures = 1
        if P is not None:
            if len(P) > 1:
                assert t is not None, "Missing timestamps t"
                assert len(t) == len(P), str(len(t)) + " vs " + str(len(P))
            else:
                t = self.t
            self.numAtmPressures = len(t)
            self.t = t
            self.P = P
        self.shiftValue = shiftValue
        if P is None:
            self.shiftValue = 0
        if not (self.shiftValue == 0):
            for i in range(len(self.P)):
                self.P[i] -= self.shiftValue
        for i in range(len(self.P)):
            if self.P[i].dtype != "float32":
                self.P[i] = self.P[i].astype(np.float32)

    def getOriginalP(self):
        return [p + self.shiftValue for p in self.P]

    def is_zero(self):
        return all(not np.any(np.asarray(a)) for a in self.P)


class TimeLevels:
    """Current/next time indices of a forcing series and the blend weight
    (SWEsimulators/Simulator.py:183-240)."""

    def __init__(self, times):
        self.times = times
        self.stamps = None

    def update(self, t):
        tt = self.times
        t_max_index = len(tt) - 1
        t0 = max(0, int(np.searchsorted(tt, t)) - 1)
        t1 = min(t_max_index, int(np.searchsorted(tt, t)))
        new = (tt[t0], tt[t1])
        old = self.stamps if self.stamps is not None else (None, None)
        self.stamps = new
        interval = max(1.0e-10, new[1] - new[0])
        w = max(0.0, min(1.0, (t - new[0]) / interval))
        return t0, t1, (new[0] != old[0]), (new[1] != old[1]), w

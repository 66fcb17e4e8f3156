"""Observation types of the data-assimilation layer
(dataassimilation/DataAssimilationUtils.py:36-52).  The resampling schemes of
that module belong to the particle filters, which this package does not
provide."""


class ObservationType:
    """Kinds of observation operator, with the reference's values."""
    DrifterPosition = 1
    UnderlyingFlow = 2
    DirectUnderlyingFlow = 3
    StaticBuoys = 4

    @staticmethod
    def _assert_valid(obs_type):
        assert obs_type in (ObservationType.DrifterPosition, ObservationType.UnderlyingFlow,
                            ObservationType.DirectUnderlyingFlow, ObservationType.StaticBuoys), \
            'Provided observation type ' + str(obs_type) + ' is invalid'

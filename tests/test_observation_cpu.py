"""Observation layer on the CPU: the oracle's observation operators pinned by
states that encode their own cell index and by closed forms, and the host
logic of Observation and ParticleInfo (utils/Observation.py,
utils/ParticleInfo.py) against closed-form drifter velocities."""
import numpy as np
import pytest

from gpuocean_amd import Observation, ObservationType, ParticleInfo
from oracle import observation as OO


def _encoded(nx, ny):
    """Ghosted state whose hu holds 1000*row + col of the ghosted array, hv
    its negative and eta 0.5 (exact in FP32)."""
    j, i = np.mgrid[0:ny + 4, 0:nx + 4]
    hu = (1000 * j + i).astype(np.float32)
    return np.full_like(hu, 0.5), hu, -hu


def _decode(v):
    return int(v) // 1000, int(v) % 1000


def test_oracle_cells_follow_numpy_indexing():
    nx, ny, dx, dy = 7, 5, 200.0, 300.0
    states = [_encoded(nx, ny)]
    pts = [((c + 0.5) * dx, (r + 0.5) * dy) for r in range(ny) for c in range(nx)]  # centres
    pts += [(3 * dx, 2 * dy)]                                           # an edge belongs to the cell right of it
    pts += [(np.nextafter(3 * dx, 0), np.nextafter(2 * dy, 0))]        # just below it, to the left
    pts += [(-0.5 * dx, -1.5 * dy)]                                    # numpy's negative indices
    out = OO.observe_particles(states, np.array(pts), dx, dy)[0]
    want = [(r, c) for r in range(ny) for c in range(nx)] + [(2, 3), (1, 2), (ny - 2, nx - 1)]
    for k, (r, c) in enumerate(want):
        assert _decode(out[k, 0]) == (r + 2, c + 2), k
        assert out[k, 1] == -out[k, 0]
    # outside -n..n-1 numpy raises: NaN here
    bad = OO.observe_particles(states, np.array([[nx * dx, 0.0], [0.0, -(ny + 1) * dy], [np.nan, 0.0]]), dx, dy)
    assert np.all(np.isnan(bad))


def test_oracle_ghosted_rule_reaches_ghost_cells():
    nx, ny, dx, dy = 6, 4, 100.0, 100.0
    eta, hu, hv = _encoded(nx, ny)
    pts = np.array([[50.0, 50.0], [-150.0, -50.0], [(nx + 1.5) * dx, (ny + 1.5) * dy], [-250.0, 0.0],
                    [(nx + 2) * dx, 0.0]], np.float32)
    out = OO.observe_underlying_flow(hu, hv, pts, dx, dy)
    assert _decode(out[0, 0]) == (2, 2)
    assert _decode(out[1, 0]) == (1, 0)
    assert _decode(out[2, 0]) == (ny + 3, nx + 3)
    assert np.all(np.isnan(out[3:])) and out.dtype == np.float32


def test_oracle_innovation_and_buoy_closed_forms():
    nx, ny = 4, 3
    rng = np.random.default_rng(0)
    hu = rng.normal(size=(ny + 4, nx + 4)).astype(np.float32)
    hv = rng.normal(size=(ny + 4, nx + 4)).astype(np.float32)
    H = np.full((ny + 4, nx + 4), 80.0, np.float32)
    obs = np.array([[50.0, 50.0, 3.0, -2.0], [150.0, 250.0, 1.5, 4.0]])
    cells = [(0, 0), (1, 2)]
    flat = np.zeros_like(hu)
    # eta = 0: no compensation; eta = mean depth: the observation counts twice
    no = OO.innovation([(flat, hu, hv)], obs, 100.0, 100.0, mean_depth=80.0)[0]
    twice = OO.innovation([(np.full_like(hu, 80.0), hu, hv)], obs, 100.0, 100.0, mean_depth=80.0)[0]
    plain = OO.innovation([(np.full_like(hu, 80.0), hu, hv)], obs, 100.0, 100.0)[0]
    for d, (c, r) in enumerate(cells):
        u, v = hu[r + 2, c + 2], hv[r + 2, c + 2]
        assert no[d, 0] == obs[d, 2] - np.float64(u) and no[d, 1] == obs[d, 3] - np.float64(v)
        assert twice[d, 0] == 2 * obs[d, 2] - np.float64(u)
        assert plain[d, 1] == obs[d, 3] - np.float64(v)
    # buoys: H/(H+eta) is 1 at rest and 1/2 when eta = H
    rest = OO.buoy_readings(flat, hu, hv, H, cells)
    high = OO.buoy_readings(H.copy(), hu, hv, H, cells)
    for b, (c, r) in enumerate(cells):
        assert rest[b, 0] == hu[r + 2, c + 2] and rest[b, 1] == hv[r + 2, c + 2]
        assert high[b, 0] == hu[r + 2, c + 2] / 2 and high[b, 1] == hv[r + 2, c + 2] / 2
    s = OO.state_samples(*_encoded(nx, ny), np.array(cells))
    assert [_decode(v) for v in s[:, 1]] == [(2, 2), (4, 3)] and np.all(s[:, 0] == 0.5)


# ------------------------------------------------------------------ Observation


class _Drifters:
    def __init__(self, pos):
        self.pos = np.asarray(pos, np.float32)

    def getDrifterPositions(self):
        return self.pos.copy()


def _track(obs, times, positions):
    for t, p in zip(times, positions):
        obs.add_observation_from_drifters(_Drifters(p), t)


def test_drifter_velocity_times_depth():
    obs = Observation()
    _track(obs, [0, 60, 120], [[[1000, 2000], [500, 500]], [[1030, 1994], [500, 560]],
                               [[1090, 1994], [440, 560]]])
    assert list(obs.get_observation_times()) == [60, 120]
    o = obs.get_observation(60, waterDepth=100.0)
    assert o.shape == (2, 4)
    assert np.array_equal(o[:, :2], [[1030, 1994], [500, 560]])
    assert np.array_equal(o[:, 2:], [[50.0, -10.0], [0.0, 100.0]])
    # every second row
    obs.setObservationInterval(2)
    assert list(obs.get_observation_times()) == [120]
    o = obs.get_observation(120, waterDepth=100.0)
    assert np.array_equal(o[:, 2:], [[75.0, -5.0], [-50.0, 50.0]])
    obs.setObservationInterval(1)
    obs.setDrifterSet([1])
    assert obs.get_num_drifters() == 1
    assert np.array_equal(obs.get_observation(120, waterDepth=100.0), [[440, 560, -100.0, 0.0]])
    with pytest.raises(AssertionError):
        obs.get_observation(0, waterDepth=100.0)   # nothing before the first row
    with pytest.raises(AssertionError):
        obs.get_observation(60)                    # no depth given


def test_velocity_is_fp32_as_numpy_1_computed_it():
    obs = Observation()
    _track(obs, [0, 7], [[[0.0, 0.0]], [[1.0, 3.0]]])
    o = obs.get_observation(7, waterDepth=0.1)
    # (1/7 in FP32) * (0.1 in FP32), rounded in FP32
    want = np.float32(np.float32(1.0) / np.float32(7)) * np.float32(0.1)
    assert o[0, 2] == np.float64(want) and o.dtype == np.float64
    assert o[0, 2] != (1.0 / 7) * 0.1


def test_periodic_crossing_takes_the_short_way():
    obs = Observation(domain_size_x=10000.0, domain_size_y=8000.0, nx=100, ny=80)
    _track(obs, [0, 60], [[[9990, 10], [5000, 7990]], [[10, 7990], [5060, 20]]])
    o = obs.get_observation(60, waterDepth=100.0)
    want_x0 = (np.float32(10) - np.float32(9990) + np.float64(10000.0)) * np.float64(100.0) / np.int64(60)
    assert o[0, 2] == want_x0 and np.isclose(o[0, 2], 2000.0 / 60)
    assert np.isclose(o[0, 3], -2000.0 / 60)
    assert np.isclose(o[1, 2], 100.0) and np.isclose(o[1, 3], 3000.0 / 60)


def test_depth_from_hm_and_observation_errors():
    np.random.seed(5)
    obs = Observation(domain_size_x=1000.0, domain_size_y=600.0, nx=10, ny=6, observation_variance=4.0)
    sim = type("S", (), {})()
    sim.t = 0.0
    sim.drifters = _Drifters([[150.0, 250.0]])
    obs.add_observation_from_sim(sim)
    sim.t = 30.0
    sim.drifters = _Drifters([[180.0, 250.0]])
    obs.add_observation_from_sim(sim)
    Hm = np.arange(60, dtype=np.float32).reshape(6, 10) + 10.0
    o = obs.get_observation(30, Hm=Hm)
    err = obs.obs_df.iloc[1]['drifter_obs_errors']
    # cell (x, y) = (1, 2): depth Hm[2, 1] = 31
    assert o[0, 2] == np.float64(np.float32(1.0) * np.float32(31.0)) + 2.0 * err[0, 0]
    assert o[0, 3] == 0.0 + 2.0 * err[0, 1]
    with pytest.raises(AssertionError):
        obs.add_observation_from_sim(sim)   # a second row for t = 30


def test_static_buoys_configuration_and_reading():
    mask = np.zeros((100, 200), bool)
    mask[:, :20] = True
    obs = Observation(ObservationType.StaticBuoys, domain_size_x=200 * 50.0, domain_size_y=100 * 50.0,
                      nx=200, ny=100, observation_variance=0.0, land_mask=mask)
    obs.setBuoyCellsByFrequency(25, 25)
    assert len(obs.buoy_indices) == 7 * 4 and obs.buoy_indices[:, 0].min() == 25
    assert np.array_equal(obs.buoy_positions[0], [25.5 * 50, 0.5 * 50])
    obs.setBuoyReadingArea('south')
    assert obs.get_num_drifters() == 2 * 7
    obs.setBuoyReadingArea('west')
    assert obs.get_num_drifters() == 3 * 4   # x = 25, 50, 75 in each of 4 rows
    obs.setBuoyReadingArea('sparse')
    assert obs.get_num_drifters() == int(np.sum((obs.buoy_indices[:, 0] % 50 == 0) & (obs.buoy_indices[:, 1] % 50 == 0)))
    obs.setBuoySet([0, 3])
    assert obs.get_num_drifters() == 2
    readings = np.arange(56, dtype=np.float32).reshape(28, 2)
    obs._new_row(0, np.zeros((0, 2), np.float32), None, readings, obs.buoy_positions, None)
    obs._new_row(60, np.zeros((0, 2), np.float32), None, readings + 1, None, None)
    o = obs.get_observation(60, waterDepth=1.0)
    assert np.array_equal(o[:, 2:], (readings + 1)[[0, 3]]) and np.array_equal(o[:, :2], obs.buoy_positions[[0, 3]])


def test_drifter_path_split_and_continued():
    obs = Observation(domain_size_x=300000.0, domain_size_y=300000.0, nx=300, ny=300)
    xs = [290000.0, 295000.0, 299000.0, 4000.0, 9000.0]
    _track(obs, [0, 60, 120, 180, 240], [[[x, 1000.0]] for x in xs])
    pieces = obs.get_drifter_path(0, 0, 240, in_km=False)
    assert len(pieces) == 2 and pieces[0].shape == (3, 2) and pieces[1].shape == (2, 2)
    whole = obs.get_drifter_path(0, 0, 240, in_km=True, keepDomainSize=False)
    assert len(whole) == 1 and np.allclose(whole[0][:, 0], [290, 295, 299, 304, 309])
    quick = obs.get_drifter_path(0, 60, 180, in_km=False, assume_no_boundary_trouble=True)
    assert np.array_equal(quick[0][:, 0], xs[1:4])


def test_observation_pickle_round_trip(tmp_path):
    obs = Observation(ObservationType.StaticBuoys, domain_size_x=1000.0, domain_size_y=1000.0, nx=10, ny=10)
    obs.setBuoyCells(np.array([[1, 2], [7, 9]], np.int32))
    obs._new_row(0, np.zeros((0, 2), np.float32), None, np.ones((2, 2), np.float32), obs.buoy_positions, None)
    path = str(tmp_path / "obs.bz2")
    obs.to_pickle(path)
    back = Observation(ObservationType.StaticBuoys, domain_size_x=1000.0, domain_size_y=1000.0, nx=10, ny=10)
    back.read_pickle(path)   # a file this test wrote
    assert np.array_equal(back.buoy_indices, [[1, 2], [7, 9]]) and back.get_num_observations() == 1


def test_particle_info_bookkeeping():
    pi = ParticleInfo()
    pi.usePredefinedExtraCells()
    assert pi.get_num_extra_cells() == 13
    pi._add(0.4, np.ones((3, 3)), np.zeros((13, 3)))
    pi._add(60.2, 2 * np.ones((3, 3)), np.zeros((13, 3)))
    assert pi.get_num_samples() == 2 and pi.get_num_drifters() == 3
    assert list(pi.get_sample_times()) == [0, 60]
    assert np.all(pi.get_state_samples(60) == 2) and pi.get_extra_sample(0).shape == (13, 3)
    with pytest.raises(AssertionError):
        pi._add(59.9, np.ones((3, 3)), None)

"""netCDF output of a running CDKLM16 and the hot start from it
(CDKLM16.fromfilename, SWEsimulators/CDKLM16.py:403-473): the restarted
simulator holds the written state and continues exactly as the original."""
import numpy as np
import pytest

import cases
from gpuocean_amd import CDKLM16
from gpuocean_amd.SimReader import SimNetCDFReader

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("land", [False, True])
def test_hot_start_continues_exactly(gpu_ctx, tmp_path, land):
    print(f"netCDF hot start, land={land}", flush=True)
    kw = cases.config_c0(n=96)
    kw["dt"] = 0.5
    if land:
        mask = np.zeros(kw["H"].shape, bool)
        mask[40:50, 10:30] = True
        kw["H"] = np.ma.array(kw["H"], mask=mask)
        for k in ("eta0", "hu0", "hv0"):
            kw[k] = kw[k].copy()
            kw[k][40:49, 10:29] = 0.0
    name = str(tmp_path / "run")
    sim = CDKLM16(gpu_ctx, **kw, write_netcdf=True, netcdf_filename=name, model_time_step=30.0)
    sim.setSOARModelError(5e-3, small_scale_perturbation_interpolation_factor=3, xorwow_seed=3)
    sim.model_error = None   # written to the file; the run itself stays deterministic
    for _ in range(3):
        sim.step(5.0)
    r = SimNetCDFReader(name + ".nc", ignore_ghostcells=False)
    assert r.getNumTimeSteps() == 4
    eta, hu, hv, t = r.getLastTimeStep()
    for a, b in zip((eta, hu, hv), sim.download()):
        assert np.array_equal(a, b)
    assert t == np.float32(sim.t) and r.get('has_model_error') == 'True'
    r.close()
    sim.closeNetCDF()

    hot = CDKLM16.fromfilename(gpu_ctx, name + ".nc", cont_write_netcdf=False)
    assert hot.model_time_step == np.float32(30.0) and np.float32(hot.dt) == np.float32(0.5)
    assert bool(hot.params.has_land) == land
    for a, b in zip(hot.download(), sim.download()):
        assert np.array_equal(a, b)
    for _ in range(4):
        sim.step(5.0)
        hot.step(5.0)
    for k, (a, b) in enumerate(zip(hot.download(), sim.download())):
        assert np.array_equal(a, b), k
    # the model error recorded in the file
    other = CDKLM16(gpu_ctx, **kw)
    other.setModelErrorFromFile(name + ".nc", xorwow_seed=3)
    assert other.model_error is not None and other.model_error.soar_q0 == np.float32(5e-3)
    assert int(other.model_error.interpolation_factor) == 3

"""gpo_cdklm16_adjoint_steps_f32_noalloc against gpo_cdklm16_adjoint_steps_f32
through the C ABI: two calls in a row (the second on a smaller grid, so the
shared scratch holds the first call's values), bit for bit."""
import ctypes

import numpy as np
import pytest
import torch

import cases
from gpuocean_amd import CDKLM16
from gpuocean_amd import _lib as L
from gpuocean_amd.adjoint import _State

pytestmark = pytest.mark.gpu

PP = ctypes.c_void_p


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _sweep_pair(gpu_ctx, n, land, m=6, seed=3):
    """A trajectory of m steps on an n x n C3-like member (walls; with land, an
    island of masked corners), then the reverse sweep of a random lambda by
    both entry points."""
    kw = cases.config_c3(n=n, seed=seed)
    if land:
        H = kw["H"]
        mask = np.zeros(H.shape, bool)
        mask[n // 2 - 4:n // 2 + 4, n // 3 - 3:n // 3 + 3] = True
        kw["H"] = np.ma.masked_array(H, mask)
    sim = CDKLM16(gpu_ctx, **kw)
    dev = sim.device
    p = sim.params
    dt = np.float32(sim.dt)
    bc = sim.boundary_conditions.packed()
    s = torch.cuda.current_stream()
    sh = ctypes.c_void_p(s.cuda_stream)
    qs = [_State(n, n, torch.float32, dev) for _ in range(m + 1)]
    q1s = [_State(n, n, torch.float32, dev) for _ in range(m)]
    torch.cuda.synchronize()
    e, u, v = sim.download()
    qs[0].load(*(torch.from_numpy(a).to(dev) for a in (e, u, v)))
    qs_p = (PP * (3 * (m + 1)))(*[a.data_ptr() for st in qs for a in st.t])
    q1s_p = (PP * (3 * m))(*[a.data_ptr() for st in q1s for a in st.t])
    pitch = qs[0].pitch
    L.check(L.lib().gpo_cdklm16_trajectory_f32(ctypes.byref(p), m, float(dt), qs_p, q1s_p, pitch, _vp(sim.Hi),
                                                sim.Hi.pitch, None, bc, None, sh), "gpo_cdklm16_trajectory_f32")
    g = torch.Generator(device="cpu").manual_seed(seed)
    lam0 = [torch.randn((n + 4, qs[0].P), generator=g).to(dev) for _ in range(3)]
    for a in lam0:
        a[:2].zero_(), a[-2:].zero_(), a[:, :2].zero_(), a[:, n + 2:].zero_()
    lam_a = _State(n, n, torch.float32, dev)
    lam_b = _State(n, n, torch.float32, dev)
    for x, y, z in zip(lam_a.t, lam_b.t, lam0):
        x.copy_(z), y.copy_(z)
    work = [_State(n, n, torch.float32, dev) for _ in range(2)]
    ws = torch.empty(L.lib().gpo_adjoint_workspace_bytes(n, n, 4), dtype=torch.uint8, device=dev)
    flag_a = torch.zeros(1, dtype=torch.int32, device=dev)
    flag_b = torch.zeros(1, dtype=torch.int32, device=dev)
    lam_a_p = (PP * 3)(*[a.data_ptr() for a in lam_a.t])
    lam_b_p = (PP * 3)(*[a.data_ptr() for a in lam_b.t])
    work_p = (PP * 6)(*[a.data_ptr() for st in work for a in st.t])
    L.check(L.lib().gpo_cdklm16_adjoint_steps_f32(ctypes.byref(p), m, float(dt), qs_p, q1s_p, lam_a_p, work_p, pitch,
                                                   _vp(sim.Hi), sim.Hi.pitch, None, bc, None, _vp(ws), _vp(flag_a),
                                                   sh), "gpo_cdklm16_adjoint_steps_f32")
    L.check(L.lib().gpo_cdklm16_adjoint_steps_f32_noalloc(ctypes.byref(p), m, float(dt), qs_p, q1s_p, lam_b_p, pitch,
                                                           _vp(sim.Hi), sim.Hi.pitch, None, bc, None, _vp(flag_b),
                                                           sh), "gpo_cdklm16_adjoint_steps_f32_noalloc")
    torch.cuda.synchronize()
    return lam_a, lam_b, int(flag_a.item()), int(flag_b.item())


@pytest.mark.parametrize("land", [False, True])
def test_noalloc_adjoint_sweep_equals_the_allocating_one(gpu_ctx, land):
    for n in (160, 96):   # the second call on a smaller grid reuses the scratch of the first
        print(f"no-alloc adjoint sweep, {n}x{n}, land={land}", flush=True)
        a, b, fa, fb = _sweep_pair(gpu_ctx, n, land)
        assert fa == fb
        for k, (x, y) in enumerate(zip(a.t, b.t)):
            xi, yi = x[2:n + 2, 2:n + 2], y[2:n + 2, 2:n + 2]
            assert torch.count_nonzero(xi) > 0
            assert torch.equal(xi, yi), f"{n}x{n} field {k}"


def test_noalloc_rejects_grids_beyond_its_scratch(gpu_ctx):
    p = L.CDKLM16Params()
    p.nx, p.ny, p.rk_order = 2000, 2000, 2
    lam = (PP * 3)(None, None, None)
    rc = L.lib().gpo_cdklm16_adjoint_steps_f32_noalloc(ctypes.byref(p), 1, 1.0, lam, lam, lam, 4 * 2048, None,
                                                        4 * 2048, None, 0x01010101, None, None, None)
    assert rc != 0
    assert b"scratch" in L.lib().gpo_last_error()

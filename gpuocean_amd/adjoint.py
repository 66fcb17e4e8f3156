"""Gradient of a point-observation misfit through CDKLM16 time steps.

No counterpart exists in the reference.  ``DifferentiableCDKLM16`` takes the
configuration of a ``CDKLM16`` simulator (RK2, wet or land cells, wall,
periodic or flow-relaxation boundaries with boundary data of one time, no
friction or forcing) and gives

  * the FP32 forward run: the simulator's own step kernel (bit-compatible with
    the host oracle), keeping a checkpoint every ``checkpoint_every`` steps;
  * J = sum over observation points of (u - u_obs)^2 + (v - v_obs)^2, with
    (u, v) = (hu, hv)/h interpolated bilinearly between cell centres;
  * dJ/d(eta0, hu0, hv0) by the hand-written adjoint kernels: each segment
    between checkpoints is recomputed once, storing every stage state, and
    then swept backwards;
  * an FP64 build of the forward, tangent-linear and adjoint runs, used by the
    finite-difference and dot-product checks.

With ``model_error=True`` every step adds the member's SOAR model error for
fixed normals (the Philox draws of the member's own stream), whose balanced
momentum depends on the state, and with ``drifter_positions`` the drifters are
advected after every step as ``CDKLM16`` moves them, and J adds the
drifter-position misfit sum |x_d(T) - x_obs_d|^2.  Both run step by step: the
forward repeats the product's step (substeps, model error, ghost cells,
drift), and the reverse sweep takes drift, sponge, model error and the two
substeps in the opposite order.

``CDKLM16Misfit`` wraps the FP32 path as a ``torch.autograd.Function``;
``DifferentiableEnsemble`` runs the FP32 path for many members at once.
"""

import ctypes

import numpy as np
import torch

from . import _lib as L


def _pitch_elems(width):
    return ((width + 63) // 64) * 64


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class _State:
    """(eta, hu, hv), each (ny+4) x pitch with a padded row pitch."""

    def __init__(self, ny, nx, dtype, device):
        self.P = _pitch_elems(nx + 4)
        self.nx, self.ny = nx, ny
        self.t = [torch.zeros((ny + 4, self.P), dtype=dtype, device=device) for _ in range(3)]
        self.pitch = self.P * self.t[0].element_size()

    def ptrs(self):
        return [_vp(a) for a in self.t]

    def views(self):
        return [a[:, :self.nx + 4] for a in self.t]

    def copy_from(self, other):
        for a, b in zip(self.t, other.t):
            a.copy_(b)

    def load(self, eta, hu, hv):
        for a, b in zip(self.views(), (eta, hu, hv)):
            a.copy_(b)

    def zero_(self):
        for a in self.t:
            a.zero_()


class DifferentiableCDKLM16:
    """Misfit and its gradient for n_steps RK2 steps of fixed dt of the given
    simulator's configuration (its bathymetry, coriolis, boundaries, dt)."""

    def __init__(self, sim, positions, observed, n_steps, dt=None, checkpoint_every=20, _workspace=True,
                 model_error=False, drifter_positions=None, drifter_observed=None):
        p = sim.params
        bc = sim.boundary_conditions
        if int(sim.rk_order) != 2 or float(sim.r) != 0.0:
            raise ValueError("the differentiable step needs RK2 and r = 0")
        if sim._has_wind or sim._has_pressure or sim.one_dimensional:
            raise ValueError("the differentiable step supports no wind, pressure or one-dimensional mode")
        if any(int(c) not in (1, 2, 3) for c in (bc.north, bc.east, bc.south, bc.west)):
            raise ValueError("the differentiable step supports wall, periodic and flow-relaxation boundaries only")
        self.frs = None
        if sim._needs_explicit_bc_each_substep():
            if len(sim.bc_data.t) != 1:
                raise ValueError("the differentiable step needs flow-relaxation data of a single time")
            sim._update_bc_values()
            f = sim._frs
            self.frs = L.Frs()
            for i in range(4):
                self.frs.sponge[i] = int(sim.sponge[i])
            addr = lambda t: None if t is None else t.data_ptr()
            self.frs.ns_cur, self.frs.ns_next = addr(f.get('ns_cur')), addr(f.get('ns_next'))
            self.frs.ns_nx = int(sim.bc_data.north.h[0].shape[0])
            self.frs.ew_cur, self.frs.ew_next = addr(f.get('ew_cur')), addr(f.get('ew_next'))
            self.frs.ew_ny = int(sim.bc_data.east.h[0].shape[0])
            self.frs.t = 1.0
            self._frs_keep = f   # the data tensors stay alive with the model
        if not sim.exact_arithmetic:
            raise ValueError("the differentiable step runs the exact forward build")
        self.sim = sim
        self.params = p
        self.nx, self.ny = int(sim.nx), int(sim.ny)
        self.dx, self.dy = float(sim.dx), float(sim.dy)
        self.bc = bc.packed()
        self.dev = sim.device
        self.stream = sim.gpu_stream
        self.n_steps = int(n_steps)
        self.dt = np.float32(sim.dt if dt is None else dt)
        self.k = max(1, int(checkpoint_every))
        pos = np.asarray(positions, np.float32).reshape(-1, 2)
        obs = np.asarray(observed, np.float32).reshape(-1, 2)
        assert pos.shape == obs.shape
        self.n_obs = pos.shape[0]
        self.pos = torch.from_numpy(pos).to(self.dev)
        self.obs = torch.from_numpy(obs).to(self.dev)
        self.Hi = sim.Hi
        self.fc = sim.f_cells
        lib = L.lib()
        with torch.cuda.stream(self.stream):
            self.ws32 = torch.empty(lib.gpo_adjoint_workspace_bytes(self.nx, self.ny, 4) if _workspace else 0,
                                    dtype=torch.uint8, device=self.dev)
            self.flag = torch.zeros(1, dtype=torch.int32, device=self.dev)
            self.J = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self._f64 = None
        self._seg = None
        # stochastic member (fixed normals) and drifters: the step-by-step path
        self.gen = None
        if model_error:
            me = sim.model_error
            if me is None or int(me.interpolation_factor) != 1:
                raise ValueError("model_error needs the simulator's SOAR model error on its own grid (factor 1)")
            from .OceanStateNoise import OceanStateNoise
            self.gen = OceanStateNoise.fromsim(sim, soar_q0=me.soar_q0, soar_L=me.soar_L, xorwow_seed=me.seed,
                                               stream_id=me.stream_id)
            self.gen.gpu_stream = self.stream
            self.draw0 = int(me.draw)
            self.g32, self.f32, self.beta32 = float(np.float32(sim.g)), float(np.float32(sim.f)), \
                float(np.float32(sim.coriolis_beta))
            if np.asarray(sim.f).size != 1:
                raise ValueError("model_error needs a scalar coriolis parameter")
        self.n_drifters = 0
        if drifter_positions is not None:
            dp = np.asarray(drifter_positions, np.float32).reshape(-1, 2)
            do = np.asarray(drifter_observed, np.float32).reshape(-1, 2)
            assert dp.shape == do.shape and dp.shape[0] >= 1
            self.n_drifters = dp.shape[0]
            self.x0 = torch.from_numpy(dp).to(self.dev)
            self.xobs = torch.from_numpy(do).to(self.dev)
            self.Bm = sim.Bm
            self.per_ns = int(bc.isPeriodicNorthSouth())
            self.per_ew = int(bc.isPeriodicEastWest())
        self.generic = self.gen is not None or self.n_drifters > 0

    # ------------------------------------------------------------------ helpers
    def _sh(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    def _new_state(self, dtype=torch.float32):
        with torch.cuda.stream(self.stream):
            return _State(self.ny, self.nx, dtype, self.dev)

    def _fc_ptr(self):
        return None if self.fc is None else ctypes.c_void_p(self.fc.data_ptr())

    def _substep32(self, src, dst, qn, rk):
        L.check(L.lib().gpo_cdklm16_substep(
            ctypes.byref(self.params), float(self.dt), rk, *src.ptrs(),
            *(qn.ptrs() if qn is not None else [None] * 3), *dst.ptrs(), src.pitch,
            _vp(self.Hi), self.Hi.pitch, self._fc_ptr(), None, self.bc, 1, self._sh()), "gpo_cdklm16_substep")

    def _step32(self, qn, q1, out):
        """q1 = S0(qn); out = S1(q1; qn).  out may be qn."""
        self._substep32(qn, q1, None, 0)
        self._substep32(q1, out, qn, 1)

    def _misfit32(self, st, lam=None):
        L.check(L.lib().gpo_point_misfit_f32(
            self.nx, self.ny, self.dx, self.dy, *st.ptrs(), st.pitch, _vp(self.Hi), self.Hi.pitch, self.n_obs,
            _vp(self.pos), _vp(self.obs), _vp(self.J), *(lam.ptrs() if lam is not None else [None] * 3),
            self._sh()), "gpo_point_misfit_f32")

    def _adj32(self, rk, inp, out, lam_out, lam_in, lam_qn, accumulate):
        L.check(L.lib().gpo_cdklm16_substep_adj_f32(
            ctypes.byref(self.params), float(self.dt), rk, *inp.ptrs(), _vp(out.t[0]), *lam_out.ptrs(),
            *lam_in.ptrs(), *(lam_qn.ptrs() if lam_qn is not None else [None] * 3), int(accumulate), inp.pitch,
            _vp(self.Hi), self.Hi.pitch, self._fc_ptr(), self.bc, _vp(self.ws32), _vp(self.flag), self._sh()),
            "gpo_cdklm16_substep_adj_f32")

    def _frs_ptr(self):
        return None if self.frs is None else ctypes.byref(self.frs)

    def _frs_scale(self, st, f64=False):
        """Transpose (and tangent) of the flow-relaxation blend on st."""
        if self.frs is None:
            return
        fn = L.lib().gpo_frs_adj_f64 if f64 else L.lib().gpo_frs_adj_f32
        L.check(fn(ctypes.byref(self.params), self.bc, self.frs.sponge, *st.ptrs(), st.pitch, self._sh()),
                "gpo_frs_adj")

    def _frs_apply64(self, st):
        if self.frs is not None:
            L.check(L.lib().gpo_frs_apply_f64(ctypes.byref(self.params), self.bc, ctypes.byref(self.frs), *st.ptrs(),
                                              st.pitch, self._sh()), "gpo_frs_apply_f64")

    def _enter(self):
        """Order the work stream after the caller's current stream."""
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))

    def _leave(self):
        """Order the caller's current stream after the work stream."""
        torch.cuda.current_stream(self.dev).wait_stream(self.stream)

    def flag_value(self):
        self.stream.synchronize()
        return int(self.flag.item())

    # -------------------------------------------------------------- FP32 path
    def forward(self, eta0, hu0, hv0):
        """Run n_steps from (eta0, hu0, hv0) ((ny+4) x (nx+4) device float32),
        keeping a checkpoint every checkpoint_every steps.  Returns J (python
        float) and the final state."""
        if self.generic:
            return self._forward_generic32(eta0, hu0, hv0)
        s = self.stream
        self._enter()
        A, B = self._new_state(), self._new_state()
        ckpts = []
        with torch.cuda.stream(s):
            A.load(eta0, hu0, hv0)
        for i0 in range(0, self.n_steps, self.k):
            c = self._new_state()
            with torch.cuda.stream(s):
                c.copy_from(A)
            ckpts.append(c)
            m = min(self.k, self.n_steps - i0)
            if self.frs is None:
                L.check(L.lib().gpo_cdklm16_rk2_steps(
                    ctypes.byref(self.params), m, float(self.dt), *A.ptrs(), *B.ptrs(), A.pitch, _vp(self.Hi),
                    self.Hi.pitch, self._fc_ptr(), self.bc, 1, self._sh()), "gpo_cdklm16_rk2_steps")
            else:   # boundary blend after each substep: steps A -> (B) -> A through the trajectory entry
                PP = ctypes.c_void_p
                qs = (PP * (3 * (m + 1)))(*[a.data_ptr() for _ in range(m + 1) for a in A.t])
                q1s = (PP * (3 * m))(*[a.data_ptr() for _ in range(m) for a in B.t])
                L.check(L.lib().gpo_cdklm16_trajectory_f32(
                    ctypes.byref(self.params), m, float(self.dt), qs, q1s, A.pitch, _vp(self.Hi), self.Hi.pitch,
                    self._fc_ptr(), self.bc, self._frs_ptr(), self._sh()), "gpo_cdklm16_trajectory_f32")
        with torch.cuda.stream(s):
            self.J.zero_()
        self._misfit32(A)
        self._final = A
        self._ckpts = ckpts
        s.synchronize()
        return float(self.J.item()), A

    def observe(self, state=None):
        """(u, v) at the observation positions of a state (default: the final
        state of the last forward run), bilinear between cell centres, as an
        (n, 2) float32 array: the observation operator of the misfit."""
        st = self._final if state is None else state
        self.stream.synchronize()
        e, hu, hv = [a.double().cpu().numpy() for a in st.views()]
        Hi = self.Hi.buf.double().cpu().numpy()
        Hm = 0.25 * (Hi[:-1, :-1] + Hi[1:, :-1] + Hi[:-1, 1:] + Hi[1:, 1:])[:e.shape[0], :e.shape[1]]
        out = np.empty((self.n_obs, 2), np.float32)
        for d, (x, y) in enumerate(self.pos.cpu().numpy()):
            sx, sy = x / self.dx - 0.5, y / self.dy - 0.5
            i0 = min(max(int(np.floor(sx)), 0), self.nx - 2)
            j0 = min(max(int(np.floor(sy)), 0), self.ny - 2)
            ax, ay = sx - i0, sy - j0
            wq = ((1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay)
            cells = ((j0 + 2, i0 + 2), (j0 + 2, i0 + 3), (j0 + 3, i0 + 2), (j0 + 3, i0 + 3))
            out[d, 0] = sum(w * hu[c] / (e[c] + Hm[c]) for w, c in zip(wq, cells))
            out[d, 1] = sum(w * hv[c] / (e[c] + Hm[c]) for w, c in zip(wq, cells))
        return out

    def _segment_buffers(self):
        if self._seg is None:
            Qs = [self._new_state() for _ in range(self.k + 1)]
            Q1s = [self._new_state() for _ in range(self.k)]
            self._lam = [self._new_state() for _ in range(3)]
            PP = ctypes.c_void_p
            self._seg = dict(
                Qs=Qs, Q1s=Q1s,
                qs_p=(PP * (3 * (self.k + 1)))(*[a.data_ptr() for st in Qs for a in st.t]),
                q1s_p=(PP * (3 * self.k))(*[a.data_ptr() for st in Q1s for a in st.t]),
                lam_p=(PP * 3)(*[a.data_ptr() for a in self._lam[0].t]),
                work_p=(PP * 6)(*[a.data_ptr() for st in self._lam[1:] for a in st.t]))
        return self._seg

    def backward(self, grad_J=1.0):
        """dJ/d(eta0, hu0, hv0) of the last forward(), scaled by grad_J, as
        three (ny+4) x (nx+4) float32 tensors (ghost cells zero).  Each
        segment between checkpoints is recomputed once, keeping every state
        and stage, then swept backwards, both by single library calls."""
        if self.generic:
            return self._backward_generic32(grad_J)
        s = self.stream
        self._enter()
        seg = self._segment_buffers()
        lam = self._lam[0]
        lib = L.lib()
        with torch.cuda.stream(s):
            lam.zero_()
            self.J.zero_()
        self._misfit32(self._final, lam)
        with torch.cuda.stream(s):
            if grad_J != 1.0:
                for a in lam.t:
                    a.mul_(float(grad_J))
        n = self.n_steps
        starts = list(range(0, n, self.k))
        for si in reversed(range(len(starts))):
            m = min(self.k, n - starts[si])
            with torch.cuda.stream(s):
                seg["Qs"][0].copy_from(self._ckpts[si])
            L.check(lib.gpo_cdklm16_trajectory_f32(
                ctypes.byref(self.params), m, float(self.dt), seg["qs_p"], seg["q1s_p"], lam.pitch, _vp(self.Hi),
                self.Hi.pitch, self._fc_ptr(), self.bc, self._frs_ptr(), self._sh()), "gpo_cdklm16_trajectory_f32")
            L.check(lib.gpo_cdklm16_adjoint_steps_f32(
                ctypes.byref(self.params), m, float(self.dt), seg["qs_p"], seg["q1s_p"], seg["lam_p"],
                seg["work_p"], lam.pitch, _vp(self.Hi), self.Hi.pitch, self._fc_ptr(), self.bc, self._frs_ptr(),
                _vp(self.ws32), _vp(self.flag), self._sh()), "gpo_cdklm16_adjoint_steps_f32")
        s.synchronize()
        return [a.clone() for a in lam.views()]

    # -------------------------------------------------------------- FP64 path
    def _init_f64(self):
        if self._f64 is not None:
            return
        lib = L.lib()
        with torch.cuda.stream(self.stream):
            Hi64 = self.Hi.buf.double()
            fc64 = None if self.fc is None else self.fc.buf.double()
            ws64 = torch.empty(lib.gpo_adjoint_workspace_bytes(self.nx, self.ny, 8), dtype=torch.uint8,
                               device=self.dev)
            wst = torch.empty(lib.gpo_tlm_workspace_bytes(self.nx, self.ny), dtype=torch.uint8, device=self.dev)
        self._f64 = dict(Hi=Hi64, Hi_pitch=8 * self.Hi.pitch_elems, fc=fc64, ws=ws64, wst=wst)

    def _fc64(self):
        f = self._f64["fc"]
        return None if f is None else _vp(f)

    def _substep64(self, src, dst, qn, rk):
        d = self._f64
        L.check(L.lib().gpo_cdklm16_substep_f64(
            ctypes.byref(self.params), float(self.dt), rk, *src.ptrs(),
            *(qn.ptrs() if qn is not None else [None] * 3), *dst.ptrs(), src.pitch, _vp(d["Hi"]), d["Hi_pitch"],
            self._fc64(), self.bc, _vp(d["ws"]), _vp(self.flag), self._sh()), "gpo_cdklm16_substep_f64")

    def _tlm64(self, src, dsrc, dst, ddst, qn, dqn, rk):
        d = self._f64
        L.check(L.lib().gpo_cdklm16_substep_tlm_f64(
            ctypes.byref(self.params), float(self.dt), rk, *src.ptrs(), *dsrc.ptrs(),
            *(qn.ptrs() if qn is not None else [None] * 3), *(dqn.ptrs() if dqn is not None else [None] * 3),
            *dst.ptrs(), *ddst.ptrs(), src.pitch, _vp(d["Hi"]), d["Hi_pitch"], self._fc64(), self.bc,
            _vp(d["wst"]), _vp(self.flag), self._sh()), "gpo_cdklm16_substep_tlm_f64")

    def _adj64(self, rk, inp, out, lam_out, lam_in, lam_qn, accumulate):
        d = self._f64
        L.check(L.lib().gpo_cdklm16_substep_adj_f64(
            ctypes.byref(self.params), float(self.dt), rk, *inp.ptrs(), _vp(out.t[0]), *lam_out.ptrs(),
            *lam_in.ptrs(), *(lam_qn.ptrs() if lam_qn is not None else [None] * 3), int(accumulate), inp.pitch,
            _vp(d["Hi"]), d["Hi_pitch"], self._fc64(), self.bc, _vp(d["ws"]), _vp(self.flag), self._sh()),
            "gpo_cdklm16_substep_adj_f64")

    def _misfit64(self, st, lam=None):
        d = self._f64
        L.check(L.lib().gpo_point_misfit_f64(
            self.nx, self.ny, self.dx, self.dy, *st.ptrs(), st.pitch, _vp(d["Hi"]), d["Hi_pitch"], self.n_obs,
            _vp(self.pos), _vp(self.obs), _vp(self.J), *(lam.ptrs() if lam is not None else [None] * 3),
            self._sh()), "gpo_point_misfit_f64")

    def trajectory_f64(self, eta0, hu0, hv0, n_steps=None):
        """FP64 states Q^0 .. Q^n and stages Q1^0 .. Q1^(n-1) (kept whole)."""
        if self.generic:
            return self._trajectory_generic64(eta0, hu0, hv0, n_steps)
        self._init_f64()
        self._enter()
        n = self.n_steps if n_steps is None else n_steps
        Qs = [self._new_state(torch.float64) for _ in range(n + 1)]
        Q1s = [self._new_state(torch.float64) for _ in range(n)]
        with torch.cuda.stream(self.stream):
            Qs[0].load(eta0, hu0, hv0)
        for j in range(n):
            self._substep64(Qs[j], Q1s[j], None, 0)
            self._frs_apply64(Q1s[j])
            self._substep64(Q1s[j], Qs[j + 1], Qs[j], 1)
            self._frs_apply64(Qs[j + 1])
        self._leave()
        return Qs, Q1s

    def misfit_f64(self, eta0, hu0, hv0):
        """J of the FP64 forward run from the given (float64) initial state."""
        if self.generic:
            return self._misfit_generic64(eta0, hu0, hv0)
        self._init_f64()
        self._enter()
        A, B = self._new_state(torch.float64), self._new_state(torch.float64)
        with torch.cuda.stream(self.stream):
            A.load(eta0, hu0, hv0)
        for _ in range(self.n_steps):
            self._substep64(A, B, None, 0)
            self._frs_apply64(B)
            self._substep64(B, A, A, 1)
            self._frs_apply64(A)
        with torch.cuda.stream(self.stream):
            self.J.zero_()
        self._misfit64(A)
        self.stream.synchronize()
        return float(self.J.item())

    def adjoint_f64(self, Qs, Q1s, lam_final):
        """lambda_0 = (dQ^n/dQ^0)^T lambda_final along an FP64 trajectory."""
        n = len(Q1s)
        self._enter()
        lam, L1, X = (self._new_state(torch.float64) for _ in range(3))
        with torch.cuda.stream(self.stream):
            lam.copy_from(lam_final)
        for j in reversed(range(n)):
            with torch.cuda.stream(self.stream):
                X.zero_()
            self._frs_scale(lam, f64=True)
            self._adj64(1, Q1s[j], Qs[j + 1], lam, L1, X, 0)
            self._frs_scale(L1, f64=True)
            self._adj64(0, Qs[j], Q1s[j], L1, X, None, 1)
            lam, X = X, lam
        self._leave()
        return lam

    def tangent_f64(self, Qs, Q1s, d0):
        """dQ^n = (dQ^n/dQ^0) d0 along an FP64 trajectory (tangent-linear model)."""
        n = len(Q1s)
        self._enter()
        out = self._new_state(torch.float64)
        d, d1, dn = d0, self._new_state(torch.float64), self._new_state(torch.float64)
        for j in range(n):
            self._tlm64(Qs[j], d, out, d1, None, None, 0)
            self._frs_scale(d1, f64=True)
            self._tlm64(Q1s[j], d1, out, dn, Qs[j], d, 1)
            self._frs_scale(dn, f64=True)
            d = dn
            dn = self._new_state(torch.float64)
        self._leave()
        return d

    def gradient_f64(self, eta0, hu0, hv0):
        """J and dJ/dQ^0 of the FP64 run (trajectory kept whole)."""
        if self.generic:
            tr = self._trajectory_generic64(eta0, hu0, hv0)
            J = self._j_generic64(tr)
            lam, lx = self._final_sens64(tr)
            lam0, _ = self._adjoint_generic64(tr, lam, lx)
            self.stream.synchronize()
            return J, [a.clone() for a in lam0.views()]
        Qs, Q1s = self.trajectory_f64(eta0, hu0, hv0)
        lamN = self._new_state(torch.float64)
        with torch.cuda.stream(self.stream):
            self.J.zero_()
        self._misfit64(Qs[-1], lamN)
        self.stream.synchronize()
        J = float(self.J.item())
        lam0 = self.adjoint_f64(Qs, Q1s, lamN)
        self.stream.synchronize()
        return J, [a.clone() for a in lam0.views()]


    # ------------------------------------------- step-by-step path (model error, drifters)
    def _stage_bc32(self, st):
        """Ghost cells as the product's step writes them: flow-relaxation
        blend and sides, or periodic copies (drifters read ghost cells)."""
        bc = self.sim.boundary_conditions
        if self.frs is None and not (bc.isPeriodicNorthSouth() or bc.isPeriodicEastWest()):
            return
        f = self.frs
        L.check(L.lib().gpo_boundary_conditions(
            self.nx, self.ny, self.bc, self.sim.sponge, *st.ptrs(), st.pitch,
            None if f is None else f.ns_cur, None if f is None else f.ns_next, 0 if f is None else f.ns_nx,
            None if f is None else f.ew_cur, None if f is None else f.ew_next, 0 if f is None else f.ew_ny,
            1.0 if f is None else float(f.t), self._sh()), "gpo_boundary_conditions")

    def _coarse(self, j):
        """The model error's SOAR field of step j (draw draw0 + j)."""
        self.gen._fill(self.draw0 + j)
        self.gen._soar(1.0, 1.0)
        return self.gen.coarse_buffer

    def _noise32(self, st, j):
        c = self._coarse(j)
        sim = self.sim
        L.check(L.lib().gpo_geostrophic_balance(
            self.nx, self.ny, self.dx, self.dy, self.g32, self.f32, self.beta32, float(self.params.north_x),
            float(self.params.north_y), self._fc_ptr(), 0 if self.fc is None else self.fc.pitch, _vp(c),
            4 * (self.nx + 4), *st.ptrs(), st.pitch, _vp(self.Hi), self.Hi.pitch, float(sim.mask_value),
            self._sh()), "gpo_geostrophic_balance")

    def _drift32(self, st, x):
        L.check(L.lib().gpo_drift(
            self.nx, self.ny, self.dx, self.dy, float(self.dt), 2, 2, *st.ptrs(), st.pitch, _vp(self.Bm),
            self.Bm.pitch, self.per_ns, self.per_ew, self.n_drifters - 1, _vp(x), 8, 1.0, None, None, None, None, 0,
            0, 0.0, 0.0, self._sh()), "gpo_drift")

    def _step_generic32(self, src, stage, pre, dst, j, x):
        """CDKLM16._one_step on private buffers: stage = S0(src), ghost cells,
        pre = S1(stage; src), dst = model error(pre), ghost cells, drift of x
        (pre may be dst; dst may be src)."""
        self._substep32(src, stage, None, 0)
        if self.frs is not None:
            self._stage_bc32(stage)
        self._substep32(stage, pre, src, 1)
        if dst is not pre:
            with torch.cuda.stream(self.stream):
                dst.copy_from(pre)
        if self.gen is not None:
            self._noise32(dst, j)
        self._stage_bc32(dst)
        if self.n_drifters:
            self._drift32(dst, x)

    def _drifter_j(self, x):
        if not self.n_drifters:
            return 0.0
        d = x.double() - self.xobs.double()
        return float((d * d).sum().item())

    def _forward_generic32(self, eta0, hu0, hv0):
        s = self.stream
        self._enter()
        A, B = self._new_state(), self._new_state()
        ckpts, xck = [], []
        with torch.cuda.stream(s):
            A.load(eta0, hu0, hv0)
            x = self.x0.clone() if self.n_drifters else None
        for j in range(self.n_steps):
            if j % self.k == 0:
                c = self._new_state()
                with torch.cuda.stream(s):
                    c.copy_from(A)
                    xck.append(x.clone() if x is not None else None)
                ckpts.append(c)
            self._step_generic32(A, B, A, A, j, x)
        with torch.cuda.stream(s):
            self.J.zero_()
        if self.n_obs:
            self._misfit32(A)
        s.synchronize()
        J = float(self.J.item()) + self._drifter_j(x)
        self._final, self._final_x = A, x
        self._ckpts, self._xck = ckpts, xck
        return J, A

    def _backward_generic32(self, grad_J=1.0):
        s = self.stream
        self._enter()
        k, n = self.k, self.n_steps
        if self._seg is None or "Pre" not in self._seg:
            self._seg = dict(Qs=[self._new_state() for _ in range(k + 1)], Q1s=[self._new_state() for _ in range(k)],
                             Pre=[self._new_state() for _ in range(k)])
            self._lam = [self._new_state() for _ in range(3)]
        seg = self._seg
        lam, L1, X = self._lam
        with torch.cuda.stream(s):
            lam.zero_()
            self.J.zero_()
            xs = [torch.empty_like(self.x0) for _ in range(k + 1)] if self.n_drifters else None
            lx = (2.0 * (self._final_x.double() - self.xobs.double())).float() * float(grad_J) \
                if self.n_drifters else None
        if self.n_obs:
            self._misfit32(self._final, lam)
        with torch.cuda.stream(s):
            if grad_J != 1.0:
                for a in lam.t:
                    a.mul_(float(grad_J))
        starts = list(range(0, n, k))
        for si in reversed(range(len(starts))):
            m = min(k, n - starts[si])
            with torch.cuda.stream(s):
                seg["Qs"][0].copy_from(self._ckpts[si])
                if xs is not None:
                    xs[0].copy_(self._xck[si])
            for jj in range(m):
                if xs is not None:
                    with torch.cuda.stream(s):
                        xs[jj + 1].copy_(xs[jj])
                self._step_generic32(seg["Qs"][jj], seg["Q1s"][jj], seg["Pre"][jj], seg["Qs"][jj + 1],
                                     starts[si] + jj, None if xs is None else xs[jj + 1])
            for jj in reversed(range(m)):
                j = starts[si] + jj
                q, q1, pre, q2 = seg["Qs"][jj], seg["Q1s"][jj], seg["Pre"][jj], seg["Qs"][jj + 1]
                if xs is not None:
                    L.check(L.lib().gpo_drift_adj_f32(
                        self.nx, self.ny, self.dx, self.dy, float(self.dt), *q2.ptrs(), q2.pitch, _vp(self.Bm),
                        self.Bm.pitch, self.bc, self.n_drifters, _vp(xs[jj]), _vp(lx), *lam.ptrs(), 1.0,
                        self._sh()), "gpo_drift_adj_f32")
                self._frs_scale(lam)
                if self.gen is not None:
                    c = self._coarse(j)
                    L.check(L.lib().gpo_balance_adj_f32(
                        ctypes.byref(self.params), self.g32, self.f32, self.beta32, self._fc_ptr(),
                        0 if self.fc is None else self.fc.pitch, _vp(c), 4 * (self.nx + 4), *lam.ptrs(), lam.pitch,
                        _vp(self.Hi), self.Hi.pitch, self._sh()), "gpo_balance_adj_f32")
                with torch.cuda.stream(s):
                    X.zero_()
                self._adj32(1, q1, pre, lam, L1, X, 0)
                self._frs_scale(L1)
                self._adj32(0, q, q1, L1, X, None, 1)
                lam, X = X, lam
        self._lam = [lam, L1, X]
        s.synchronize()
        return [a.clone() for a in lam.views()]

    # FP64 counterparts (the verification build)
    def _init_generic64(self):
        self._init_f64()
        if self.n_drifters and "Bm" not in self._f64:
            with torch.cuda.stream(self.stream):
                self._f64["Bm"] = self.Bm.buf.double()

    def _balance64(self, st, j, d=None):
        dd = self._f64
        c = self._coarse(j)
        L.check(L.lib().gpo_balance_f64(
            ctypes.byref(self.params), self.g32, self.f32, self.beta32, self._fc_ptr(),
            0 if self.fc is None else self.fc.pitch, _vp(c), 4 * (self.nx + 4), *st.ptrs(),
            *(d.ptrs() if d is not None else [None] * 3), st.pitch, _vp(dd["Hi"]), dd["Hi_pitch"], self._sh()),
            "gpo_balance_f64")

    def _drift64(self, st, x, dx=None, dst=None):
        Bm = self._f64["Bm"]
        L.check(L.lib().gpo_drift_f64(
            self.nx, self.ny, self.dx, self.dy, float(self.dt), *st.ptrs(), st.pitch, _vp(Bm), 8 * Bm.shape[1],
            self.bc, self.n_drifters, _vp(x), None if dx is None else _vp(dx),
            *(dst.ptrs() if dst is not None else [None] * 3), 1.0, self._sh()), "gpo_drift_f64")

    def _trajectory_generic64(self, eta0, hu0, hv0, n_steps=None):
        """FP64 run kept whole: Qs (states after each step), Q1s (stages),
        Pre (second-substep outputs before model error) and Xs (drifter
        positions before each step, and the final ones)."""
        self._init_generic64()
        self._enter()
        s = self.stream
        n = self.n_steps if n_steps is None else n_steps
        Qs = [self._new_state(torch.float64) for _ in range(n + 1)]
        Q1s = [self._new_state(torch.float64) for _ in range(n)]
        Pre = [self._new_state(torch.float64) for _ in range(n)]
        with torch.cuda.stream(s):
            Qs[0].load(eta0, hu0, hv0)
            Xs = [self.x0.double().clone()] if self.n_drifters else None
        for j in range(n):
            self._substep64(Qs[j], Q1s[j], None, 0)
            self._frs_apply64(Q1s[j])
            self._substep64(Q1s[j], Pre[j], Qs[j], 1)
            with torch.cuda.stream(s):
                Qs[j + 1].copy_from(Pre[j])
            if self.gen is not None:
                self._balance64(Qs[j + 1], j)
            self._frs_apply64(Qs[j + 1])
            if Xs is not None:
                with torch.cuda.stream(s):
                    Xs.append(Xs[-1].clone())
                self._drift64(Qs[j + 1], Xs[-1])
        self._leave()
        return dict(Qs=Qs, Q1s=Q1s, Pre=Pre, Xs=Xs)

    def _j_generic64(self, tr):
        with torch.cuda.stream(self.stream):
            self.J.zero_()
        if self.n_obs:
            self._misfit64(tr["Qs"][-1])
        self.stream.synchronize()
        J = float(self.J.item())
        if self.n_drifters:
            d = tr["Xs"][-1] - self.xobs.double()
            J += float((d * d).sum().item())
        return J

    def _misfit_generic64(self, eta0, hu0, hv0):
        return self._j_generic64(self._trajectory_generic64(eta0, hu0, hv0))

    def _final_sens64(self, tr):
        """dJ/d(final state) and dJ/d(final positions)."""
        lam = self._new_state(torch.float64)
        with torch.cuda.stream(self.stream):
            self.J.zero_()
        if self.n_obs:
            self._misfit64(tr["Qs"][-1], lam)
        lx = None
        if self.n_drifters:
            with torch.cuda.stream(self.stream):
                lx = 2.0 * (tr["Xs"][-1] - self.xobs.double())
        return lam, lx

    def _adjoint_generic64(self, tr, lam_final, lx_final):
        """(lambda_0, lambda_x0) = transposed sensitivities of the FP64 run."""
        Qs, Q1s, Pre, Xs = tr["Qs"], tr["Q1s"], tr["Pre"], tr["Xs"]
        s = self.stream
        self._enter()
        lam, L1, X = (self._new_state(torch.float64) for _ in range(3))
        with torch.cuda.stream(s):
            lam.copy_from(lam_final)
            lx = lx_final.clone() if lx_final is not None else None
        for j in reversed(range(len(Q1s))):
            if lx is not None:
                Bm = self._f64["Bm"]
                L.check(L.lib().gpo_drift_adj_f64(
                    self.nx, self.ny, self.dx, self.dy, float(self.dt), *Qs[j + 1].ptrs(), Qs[j + 1].pitch, _vp(Bm),
                    8 * Bm.shape[1], self.bc, self.n_drifters, _vp(Xs[j]), _vp(lx), *lam.ptrs(), 1.0, self._sh()),
                    "gpo_drift_adj_f64")
            self._frs_scale(lam, f64=True)
            if self.gen is not None:
                c = self._coarse(j)
                dd = self._f64
                L.check(L.lib().gpo_balance_adj_f64(
                    ctypes.byref(self.params), self.g32, self.f32, self.beta32, self._fc_ptr(),
                    0 if self.fc is None else self.fc.pitch, _vp(c), 4 * (self.nx + 4), *lam.ptrs(), lam.pitch,
                    _vp(dd["Hi"]), dd["Hi_pitch"], self._sh()), "gpo_balance_adj_f64")
            with torch.cuda.stream(s):
                X.zero_()
            self._adj64(1, Q1s[j], Pre[j], lam, L1, X, 0)
            self._frs_scale(L1, f64=True)
            self._adj64(0, Qs[j], Q1s[j], L1, X, None, 1)
            lam, X = X, lam
        self._leave()
        return lam, lx

    def _tangent_generic64(self, tr, d0, dx0):
        """(dQ^n, dx^n) for initial perturbations (d0, dx0) along the FP64 run."""
        Qs, Q1s, Pre, Xs = tr["Qs"], tr["Q1s"], tr["Pre"], tr["Xs"]
        s = self.stream
        self._enter()
        out = self._new_state(torch.float64)
        val = self._new_state(torch.float64)
        d = d0
        with torch.cuda.stream(s):
            dx = dx0.clone() if dx0 is not None else None
        for j in range(len(Q1s)):
            d1, dn = self._new_state(torch.float64), self._new_state(torch.float64)
            self._tlm64(Qs[j], d, out, d1, None, None, 0)
            self._frs_scale(d1, f64=True)
            self._tlm64(Q1s[j], d1, out, dn, Qs[j], d, 1)
            if self.gen is not None:
                with torch.cuda.stream(s):
                    val.copy_from(Pre[j])
                self._balance64(val, j, dn)
            self._frs_scale(dn, f64=True)
            if dx is not None:
                with torch.cuda.stream(s):
                    xv = Xs[j].clone()
                self._drift64(Qs[j + 1], xv, dx, dn)
            d = dn
        self._leave()
        return d, dx


class CDKLM16Misfit(torch.autograd.Function):
    """J(eta0, hu0, hv0) of a DifferentiableCDKLM16 run as an autograd node:
    forward = FP32 step kernel with checkpoints, backward = adjoint sweep."""

    @staticmethod
    def forward(ctx, eta0, hu0, hv0, model):
        J, _ = model.forward(eta0.detach(), hu0.detach(), hv0.detach())
        ctx.model = model
        return torch.tensor(J, dtype=torch.float64, device=eta0.device)

    @staticmethod
    def backward(ctx, grad_J):
        g = ctx.model.backward(float(grad_J.item()))
        return g[0], g[1], g[2], None


def binomial_schedule(n_seg, n_slots):
    """Binomial checkpointing (Griewank and Walther's revolve) over n_seg
    segments with n_slots checkpoint slots, slot 0 holding the initial
    state: the actions that reverse every segment, last first, with the
    fewest segment advances.  Actions:

      ("load", slot, a)     the forward run continues from the start of
                            segment a, held in slot
      ("advance", a, m)     the forward run steps through segments a..m-1
      ("save", slot, m)     the start of segment m goes into slot
      ("turn", a)           segment a is recomputed from the forward run's
                            state with every intermediate kept, then swept

    The split points come from the recursion cost(t, s) = min over m of
    m + cost(t - m, s - 1) + cost(m, s) (t segments whose start is held, s
    free slots; cost(t, 0) = t (t - 1) / 2).  Its value is the minimal
    number of advances of binomial checkpointing with S = s + 1 slots,
    r t - C(S + r, S + 1) with r the least repetition number such that
    C(S + r, S) >= t."""
    n_seg, n_slots = int(n_seg), int(n_slots)
    if n_seg < 1 or n_slots < 1:
        raise ValueError("binomial_schedule: need at least one segment and one slot")
    memo = {}

    def cost(t, s):
        if t <= 1:
            return 0
        if s == 0:
            return t * (t - 1) // 2
        if (t, s) not in memo:
            memo[(t, s)] = min(m + cost(t - m, s - 1) + cost(m, s) for m in range(1, t))
        return memo[(t, s)]

    ops = []

    def rev(a, b, slot, free):
        t = b - a
        if t == 1:
            ops.extend([("load", slot, a), ("turn", a)])
        elif not free:
            for i in range(b - 1, a - 1, -1):
                ops.append(("load", slot, a))
                if i > a:
                    ops.append(("advance", a, i))
                ops.append(("turn", i))
        else:
            s = len(free)
            m = a + min(range(1, t), key=lambda q: (q + cost(t - q, s - 1) + cost(q, s), q))
            ops.extend([("load", slot, a), ("advance", a, m), ("save", free[0], m)])
            rev(m, b, free[0], free[1:])
            rev(a, m, slot, free)

    rev(0, n_seg, 0, list(range(1, min(n_slots, n_seg))))
    return ops


def schedule_split(ops):
    """Where the forward run's share of a binomial_schedule ends: before the
    group of the first turn (its load and, without a free slot, its
    advance).  The forward run executes ops[:j] and then steps on to the end
    for the misfit; the sweep executes ops[j:], whose first load brings the
    forward run back."""
    j = next(i for i, op in enumerate(ops) if op[0] == "turn")
    if j > 0 and ops[j - 1][0] == "advance":
        j -= 1
    while j > 0 and ops[j - 1][0] == "load":
        j -= 1
    return j


class DifferentiableEnsemble:
    """The misfit and gradient of DifferentiableCDKLM16 for many members at
    once: every member's forward steps, trajectory recomputation and reverse
    sweep run as single launches over all members (gpo_adj_batch_*), each
    member with its own grid size, dt, observations and checkpoints.  The
    members share the number of steps and the checkpoint spacing.

    checkpoint_slots: states kept per member between the forward run and the
    sweep.  None (default) keeps the start of every segment, so each segment
    is recomputed once; fewer slots follow binomial_schedule (revolve), which
    advances through segments more often to fit memory.  Every choice gives
    the same gradient bit for bit."""

    def __init__(self, sims, positions, observed, n_steps, checkpoint_every=20, dts=None, model_error=False,
                 drifter_positions=None, drifter_observed=None, fused_reverse=True, march_rows=0,
                 checkpoint_slots=None):
        dpos = drifter_positions if drifter_positions is not None else [None] * len(sims)
        dobs = drifter_observed if drifter_observed is not None else [None] * len(sims)
        self.models = [DifferentiableCDKLM16(s, p, o, n_steps, dt=None if dts is None else dts[i],
                                             checkpoint_every=checkpoint_every, _workspace=False,
                                             model_error=model_error, drifter_positions=dpos[i],
                                             drifter_observed=dobs[i])
                       for i, (s, p, o) in enumerate(zip(sims, positions, observed))]
        self.n = len(self.models)
        self.n_steps = int(n_steps)
        self.k = max(1, int(checkpoint_every))
        self.n_seg = max(1, (self.n_steps + self.k - 1) // self.k)
        if checkpoint_slots is None or int(checkpoint_slots) >= self.n_seg:
            self.n_slots, self.schedule = self.n_seg, None
        else:
            self.n_slots = max(1, int(checkpoint_slots))
            self.schedule = binomial_schedule(self.n_seg, self.n_slots)
            self._first_back = schedule_split(self.schedule)
        self.dev = sims[0].device
        self.stream = torch.cuda.Stream(device=self.dev)
        lib = L.lib()
        table = (L.AdjMember * self.n)()
        self._bufs = []
        self._pos = []
        self._coarse_seg = []
        for i, m in enumerate(self.models):
            P = _pitch_elems(m.nx + 4)
            rows = m.ny + 4
            # one allocation per member: state, stage, checkpoints, k+1 states, k stages, 3 lambda states,
            # and with model error k pre-noise eta planes (three to a state)
            n_states = 2 + self.n_slots + (self.k + 1) + self.k + 3 + ((self.k + 2) // 3 if m.gen is not None else 0)
            with torch.cuda.stream(self.stream):
                buf = torch.zeros((n_states, 3, rows, P), dtype=torch.float32, device=self.dev)
                ws = torch.empty(lib.gpo_adjoint_workspace_bytes(m.nx, m.ny, 4), dtype=torch.uint8, device=self.dev)
                # drifter positions: forward, checkpoints, segment, sensitivity
                nd = m.n_drifters
                pos = torch.zeros((1 + self.n_slots + self.k + 1 + 1, max(nd, 1), 2), dtype=torch.float32,
                                  device=self.dev)
            self._bufs.append((buf, ws))
            self._pos.append(pos)
            base = buf.data_ptr()
            st = 3 * rows * P * 4
            e = table[i]
            e.params = m.params
            e.dt = float(m.dt)
            e.boundary_conditions = m.bc
            e.Hi, e.Hi_pitch = m.Hi.data_ptr(), m.Hi.pitch
            e.f_cells = None if m.fc is None else m.fc.data_ptr()
            e.pitch = P * 4
            e.plane = rows * P * 4
            e.state = base
            e.stage = base + st
            e.ckpt = base + 2 * st
            e.qs = base + (2 + self.n_slots) * st
            e.q1s = base + (2 + self.n_slots + self.k + 1) * st
            e.lam = base + (2 + self.n_slots + 2 * self.k + 1) * st
            e.workspace = ws.data_ptr()
            e.has_frs = int(m.frs is not None)
            if m.frs is not None:
                e.frs = m.frs
            if m.gen is not None:
                g = m.gen
                e.has_noise = 1
                e.seed, e.stream_id, e.draw0 = g.seed, g.stream_id, m.draw0
                W = g.weights(1.0).reshape(-1)
                for q in range(25):
                    e.soar_weights[q] = float(W[q])
                e.g, e.f, e.beta = m.g32, m.f32, m.beta32
                e.xi, (e.xi_rows, e.xi_cols) = g.random_numbers.data_ptr(), g.random_numbers.shape
                e.coarse = g.coarse_buffer.data_ptr()
                with torch.cuda.stream(self.stream):   # the segment's SOAR fields, kept for the sweep
                    cs = torch.empty((self.k, m.ny + 4, m.nx + 4), dtype=torch.float32, device=self.dev)
                self._coarse_seg.append(cs)
                e.coarse_seg = cs.data_ptr()
                e.pre = base + (2 + self.n_slots + 2 * self.k + 1 + 3) * st
            if m.n_drifters:
                pb = pos.data_ptr()
                sp = 2 * 4 * m.n_drifters
                e.n_drifters = m.n_drifters
                e.Bm, e.Bm_pitch = m.Bm.data_ptr(), m.Bm.pitch
                e.pos, e.pos_ckpt = pb, pb + sp
                e.pos_seg = pb + (1 + self.n_slots) * sp
                e.lam_pos = pb + (1 + self.n_slots + self.k + 1) * sp
        self._pitch = [_pitch_elems(m.nx + 4) for m in self.models]
        h = ctypes.c_void_p()
        self.stream.synchronize()
        L.check(lib.gpo_adj_batch_create(table, self.n, self.k, self.n_slots, ctypes.byref(h), self._sh()),
                "gpo_adj_batch_create")
        self.handle = h
        self.set_reverse(fused_reverse, march_rows)
        with torch.cuda.stream(self.stream):
            self.J = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
            self.flag = torch.zeros(1, dtype=torch.int32, device=self.dev)

    def _sh(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    def set_reverse(self, fused=True, rows=0):
        """Reverse substeps as one fused strip march per (member, strip,
        chunk of `rows` rows; 0 keeps the current height) or as the four
        staged kernels; both give the same gradient bit for bit."""
        L.check(L.lib().gpo_adj_batch_set_reverse(self.handle, int(bool(fused)), int(rows), self._sh()),
                "gpo_adj_batch_set_reverse")

    def _state(self, i, idx):
        """(eta, hu, hv) views of state idx of member i's buffer."""
        buf = self._bufs[i][0]
        m = self.models[i]
        return [buf[idx, q, :, :m.nx + 4] for q in range(3)]

    def _misfit(self, i, idx, lam_idx=None):
        m = self.models[i]
        buf = self._bufs[i][0]
        P = self._pitch[i]
        st = [buf[idx, q] for q in range(3)]
        lam = [None] * 3 if lam_idx is None else [_vp(buf[lam_idx, q]) for q in range(3)]
        L.check(L.lib().gpo_point_misfit_f32(
            m.nx, m.ny, m.dx, m.dy, *[_vp(a) for a in st], P * 4, _vp(m.Hi), m.Hi.pitch, m.n_obs, _vp(m.pos),
            _vp(m.obs), ctypes.c_void_p(self.J.data_ptr() + 8 * i), *lam, self._sh()), "gpo_point_misfit_f32")

    def forward(self, states):
        """states: one (eta0, hu0, hv0) triple of (ny+4) x (nx+4) device float32
        tensors per member.  Returns the misfits J (numpy float64, one per
        member), keeping a checkpoint every checkpoint_every steps."""
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))
        lib = L.lib()
        with torch.cuda.stream(self.stream):
            for i, st in enumerate(states):
                for a, b in zip(self._state(i, 0), st):
                    a.copy_(b)
                m = self.models[i]
                if m.n_drifters:
                    self._pos[i][0, :m.n_drifters].copy_(m.x0)
        if self.schedule is None:
            for sg in range(self.n_seg):
                m = min(self.k, self.n_steps - sg * self.k)
                L.check(lib.gpo_adj_batch_checkpoint(self.handle, sg, 0, self._sh()), "gpo_adj_batch_checkpoint")
                L.check(lib.gpo_adj_batch_forward(self.handle, m, self._sh()), "gpo_adj_batch_forward")
        else:
            self._slot(0, 0, 0)
            self._at = 0
            self._run_schedule(self.schedule[:self._first_back])
            self._advance(self._at, self.n_seg)
        with torch.cuda.stream(self.stream):
            self.J.zero_()
        for i in range(self.n):
            if self.models[i].n_obs:
                self._misfit(i, 0)
        self.stream.synchronize()
        J = self.J.cpu().numpy().copy()
        for i, m in enumerate(self.models):
            if m.n_drifters:
                J[i] += m._drifter_j(self._pos[i][0, :m.n_drifters])
        return J

    # ---- binomial checkpointing: the forward run's position is self._at (a segment index)
    def _slot(self, slot, mode, segment):
        L.check(L.lib().gpo_adj_batch_slot(self.handle, slot, mode, segment, self._sh()), "gpo_adj_batch_slot")

    def _advance(self, a, m):
        steps = min(m * self.k, self.n_steps) - a * self.k
        if steps > 0:
            L.check(L.lib().gpo_adj_batch_forward(self.handle, steps, self._sh()), "gpo_adj_batch_forward")
        self._at = m

    def _run_schedule(self, ops):
        lib = L.lib()
        for op in ops:
            if op[0] == "load":
                if self._at != op[2]:
                    self._slot(op[1], 2, op[2])
                    self._at = op[2]
            elif op[0] == "advance":
                assert self._at == op[1]
                self._advance(op[1], op[2])
            elif op[0] == "save":
                assert self._at == op[2]
                self._slot(op[1], 0, op[2])
            else:   # turn: recompute segment op[1] from the forward run's state, then sweep it
                sg = op[1]
                assert self._at == sg
                m = min(self.k, self.n_steps - sg * self.k)
                self._slot(0, 3, sg)
                L.check(lib.gpo_adj_batch_trajectory(self.handle, m, self._sh()), "gpo_adj_batch_trajectory")
                L.check(lib.gpo_adj_batch_sweep(self.handle, m, _vp(self.flag), self._sh()), "gpo_adj_batch_sweep")

    def final_positions(self, i):
        """Drifter positions of member i after the last forward()."""
        m = self.models[i]
        self.stream.synchronize()
        return self._pos[i][0, :m.n_drifters].clone()

    def backward(self):
        """dJ_i/d(eta0, hu0, hv0) of every member i of the last forward(), as
        a list of three (ny+4) x (nx+4) float32 tensors per member."""
        lib = L.lib()
        lam_idx = 2 + self.n_slots + 2 * self.k + 1
        with torch.cuda.stream(self.stream):
            for buf, _ in self._bufs:
                buf[lam_idx:lam_idx + 3].zero_()
            self.J.zero_()
            for i, m in enumerate(self.models):
                if m.n_drifters:
                    x = self._pos[i][0, :m.n_drifters]
                    self._pos[i][1 + self.n_slots + self.k + 1, :m.n_drifters].copy_(
                        (2.0 * (x.double() - m.xobs.double())).float())
        for i in range(self.n):
            if self.models[i].n_obs:
                self._misfit(i, 0, lam_idx)
        if self.schedule is None:
            for sg in reversed(range(self.n_seg)):
                m = min(self.k, self.n_steps - sg * self.k)
                L.check(lib.gpo_adj_batch_checkpoint(self.handle, sg, 1, self._sh()), "gpo_adj_batch_checkpoint")
                L.check(lib.gpo_adj_batch_trajectory(self.handle, m, self._sh()), "gpo_adj_batch_trajectory")
                L.check(lib.gpo_adj_batch_sweep(self.handle, m, _vp(self.flag), self._sh()), "gpo_adj_batch_sweep")
        else:
            self._run_schedule(self.schedule[self._first_back:])
        self.stream.synchronize()
        return [[a.clone() for a in self._state(i, lam_idx)] for i in range(self.n)]

    def flag_value(self):
        self.stream.synchronize()
        return int(self.flag.item())

    def destroy(self):
        if getattr(self, "handle", None) is not None:
            self.stream.synchronize()
            L.lib().gpo_adj_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

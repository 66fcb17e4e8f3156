"""Ensembles of CDKLM16 members on one GPU.

``OceanModelEnsemble`` is a drop-in for the reference's
``ensembles/OceanModelEnsemble.py``: the same constructor, ``modelStep``,
``stepToObservation``, ``updateDt``, ``observeParticles`` and
``attachDrifters``.  ``data_args`` may also be a list with one dict per member,
so members can differ in size.

Where the reference steps its members one after another (a Python loop with a
blocking dt read-back per member, OceanModelEnsemble.py:138-150),
``modelStep`` hands every eligible run of full time steps to a
``BatchedEnsemble``: one persistent launch per RK substep over all members
(a table of member/strip/row-chunk work items), then one batched launch each
of the random field, SOAR and geostrophic balance, optionally replayed from a
captured hipGraph.  The batched launches produce bit for bit the values of the
per-member path.
"""

import ctypes
import logging

import numpy as np
import torch

from . import Common
from . import _lib as L
from .CDKLM16 import CDKLM16
from .GPUDrifterCollection import GPUDrifterCollection
from .Observation import Observation
from .ParticleInfo import ParticleInfo
from . import observe


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class BatchedEnsemble:
    """Members stepped together by the library's batched ensemble entry points.

    Every member must use RK2 with wall, periodic or flow-relaxation
    boundaries (flow relaxation toward boundary data of a single time) and no
    wind or pressure forcing; drifters may be attached if they carry no wind
    drag and share the member's periodicity.  All members must agree on exact
    arithmetic, land, coriolis texture, north angle and whether they carry
    SOAR model error.  Each member keeps its own dt, seed and draw index."""

    def __init__(self, sims, chunk_rows=None, waves_per_simd=None, autotune=False):
        ok, why = BatchedEnsemble.eligible(sims)
        if not ok:
            raise ValueError(f"members cannot be batched: {why}")
        self.sims = list(sims)
        self.n = len(self.sims)
        self.device = self.sims[0].device
        self.stream = torch.cuda.Stream(device=self.device)
        for s in self.sims:   # one stream for all members: no joins around each launch
            s._adopt_stream(self.stream)
        self.exact = bool(self.sims[0].exact_arithmetic)
        self.has_noise = self.sims[0].model_error is not None
        self.handle = None
        self._graph_steps = 0
        self._dts = None
        self._draws = None
        if autotune:
            chunk_rows, waves_per_simd = Autotuner.tune(self)
        self.chunk_rows = int(chunk_rows or 0)
        self.waves_per_simd = int(waves_per_simd or 0)
        self._create()

    # ----------------------------------------------------------------- set-up
    @staticmethod
    def eligible(sims):
        if len(sims) == 0:
            return False, "no members"
        s0 = sims[0]
        for s in sims:
            bc = s.boundary_conditions
            if int(s.rk_order) != 2:
                return False, "RK order other than 2"
            if any(int(c) not in (1, 2, 3) for c in (bc.north, bc.east, bc.south, bc.west)):
                return False, "boundaries other than wall, periodic and flow relaxation"
            if s._has_wind or s._has_pressure:
                return False, "wind or pressure forcing"
            if s.write_netcdf:
                return False, "members writing netCDF output after every step"
            if s.hasDrifters:
                d = s.drifters
                if not hasattr(d, "positions") or float(getattr(d, "wind_drift_factor", 0.0)) != 0.0:
                    return False, "drifters without device positions, or with wind drag"
                db = d.boundaryConditions
                if (db.isPeriodicNorthSouth(), db.isPeriodicEastWest()) != \
                        (bc.isPeriodicNorthSouth(), bc.isPeriodicEastWest()):
                    return False, "drifters whose periodicity differs from their member's"
            if (s.model_error is None) != (s0.model_error is None):
                return False, "some members carry model error and some do not"
            if s.model_error is not None and \
                    (int(s.model_error.interpolation_factor) > 1) != (int(s0.model_error.interpolation_factor) > 1):
                return False, "some members generate model error on a coarser grid and some do not"
            if s.device != s0.device or bool(s.exact_arithmetic) != bool(s0.exact_arithmetic):
                return False, "members on different devices or arithmetic modes"
            p, q = s.params, s0.params
            if (p.has_land, p.f_is_array, p.angle_is_zero) != (q.has_land, q.f_is_array, q.angle_is_zero):
                return False, "members differ in land, coriolis texture or north angle"
        return True, ""

    def _member_table(self):
        arr = (L.Member * self.n)()
        for i, s in enumerate(self.sims):
            m = arr[i]
            m.params = s.params
            m.dt = float(np.float32(s.dt))
            m.h0_eta, m.h0_hu, m.h0_hv = (s.h0[k].data_ptr() for k in range(3))
            m.h1_eta, m.h1_hu, m.h1_hv = (s.h1[k].data_ptr() for k in range(3))
            m.pitch = s.h0[0].pitch
            m.Hi, m.Hi_pitch = s.Hi.data_ptr(), s.Hi.pitch
            m.f_cells = None if s.f_cells is None else s.f_cells.data_ptr()
            m.boundary_conditions = s.boundary_conditions.packed()
            me = s.model_error
            m.noise = int(me is not None)
            if me is not None:
                rows, cols = me.random_numbers.shape
                m.xi, m.xi_rows, m.xi_cols, m.xi_pitch = me.random_numbers.data_ptr(), rows, cols, 4 * cols
                m.coarse, m.coarse_pitch = me.coarse_buffer.data_ptr(), 4 * (int(me.coarse_nx) + 4)
                m.seed, m.stream_id, m.draw = me.seed, me.stream_id, me.draw
                W = me.weights(1.0).reshape(-1)
                for k in range(25):
                    m.soar_weights[k] = float(W[k])
                m.perturbation_scale = 1.0
                m.g = float(np.float32(s.g))
                m.f = float(np.float32(s.f)) if np.asarray(s.f).size == 1 else 0.0
                m.beta = float(np.float32(s.coriolis_beta))
                if int(me.interpolation_factor) > 1:
                    # perturbSim's coarse grid, aligned as it aligns it by default
                    m.coarse_nx, m.coarse_ny = int(me.coarse_nx), int(me.coarse_ny)
                    m.coarse_dx, m.coarse_dy = float(me.coarse_dx), float(me.coarse_dy)
                    m.coarse_offset_i, m.coarse_offset_j = me._obtain_coarse_grid_offset(None, None)
            m.bc_each_step = int(self._bc_in_batch(s))
            for k in range(4):
                m.sponge[k] = int(s.sponge[k])
            if s._needs_explicit_bc_each_substep():
                s._update_bc_values()
                f = s._frs
                addr = lambda t: None if t is None else t.data_ptr()
                m.frs_ns_cur, m.frs_ns_next = addr(f.get('ns_cur')), addr(f.get('ns_next'))
                m.frs_ew_cur, m.frs_ew_next = addr(f.get('ew_cur')), addr(f.get('ew_next'))
                m.frs_ns_nx = int(s.bc_data.north.h[0].shape[0])
                m.frs_ew_ny = int(s.bc_data.east.h[0].shape[0])
                m.frs_t = 1.0
                # every level on the device (one or several times); each step's
                # levels and weight are derived there from the member's time
                n_t = len(s.bc_data.t)
                times = torch.tensor(np.asarray(s.bc_data.t, dtype=np.float64), device=s.device)
                ns = torch.from_numpy(np.stack([Common.pack_bc_data_ns(s.bc_data, k) for k in range(n_t)])
                                      .astype(np.float32)).to(s.device)
                ew = torch.from_numpy(np.stack([Common.pack_bc_data_ew(s.bc_data, k) for k in range(n_t)])
                                      .astype(np.float32)).to(s.device)
                self._frs_levels.append((times, ns, ew))
                m.frs_n_times = n_t
                m.frs_times, m.frs_ns_levels, m.frs_ew_levels = times.data_ptr(), ns.data_ptr(), ew.data_ptr()
            m.t = float(s.t)
            if s.hasDrifters:
                d = s.drifters
                m.drifters = 1
                m.n_drifters = int(d.numDrifters)
                m.positions, m.pos_pitch = d.positions.data_ptr(), 8
                m.sensitivity = float(np.float32(d.sensitivity))
                m.Bm, m.Bm_pitch = s.Bm.data_ptr(), s.Bm.pitch
        return arr

    @staticmethod
    def _bc_in_batch(s):
        """Whether the batched step writes this member's ghost cells every
        step (flow relaxation, or drifters that read them)."""
        return s._needs_explicit_bc_each_substep() or s.hasDrifters

    def _create(self):
        self.destroy()
        h = ctypes.c_void_p()
        self._sync_in()
        self._frs_levels = []   # device copies of time-varying boundary data, alive with the ensemble
        L.check(L.lib().gpo_ensemble_create(self._member_table(), self.n, self.chunk_rows, self.waves_per_simd,
                                            int(self.exact), ctypes.byref(h), self._sh()), "gpo_ensemble_create")
        self.handle = h
        self._graph_steps = 0
        self._dts = [np.float32(s.dt) for s in self.sims]
        self._draws = [s.model_error.draw if self.has_noise else 0 for s in self.sims]
        self._times = [s.t for s in self.sims]
        self._has_clocks = bool(self._frs_levels)

    def _sh(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    def _sync_in(self):
        for s in self.sims:
            if s.gpu_stream is not self.stream:   # a member moved to another stream since
                self.stream.wait_stream(s.gpu_stream)

    def _sync_out(self):
        for s in self.sims:
            if s.gpu_stream is not self.stream:
                s.gpu_stream.wait_stream(self.stream)

    def _refresh(self):
        """Push members' dt and draw index to the device tables if they changed
        (and the unit model-error scale of a regular step, if a
        data-assimilation step left other scales)."""
        if getattr(self, "_scales_dirty", False):
            self._select(None, [1.0] * self.n)
            self._scales_dirty = False
        dts = [np.float32(s.dt) for s in self.sims]
        draws = [s.model_error.draw if self.has_noise else 0 for s in self.sims]
        new_dt = dts != self._dts
        new_draw = draws != self._draws
        if new_dt or new_draw:
            dt_arr = (ctypes.c_float * self.n)(*[float(d) for d in dts]) if new_dt else None
            dr_arr = (ctypes.c_uint32 * self.n)(*draws) if new_draw else None
            L.check(L.lib().gpo_ensemble_update(self.handle, dt_arr, dr_arr, self._sh()), "gpo_ensemble_update")
            self._dts, self._draws = dts, draws
        self._sync_times()

    def _sync_times(self):
        """Upload the members' model times to the device boundary-data clocks
        if they moved outside the batched steps."""
        if getattr(self, "_has_clocks", False):
            times = [s.t for s in self.sims]
            if times != self._times:
                arr = (ctypes.c_double * self.n)(*[float(t) for t in times])
                L.check(L.lib().gpo_ensemble_set_times(self.handle, arr, self._sh()), "gpo_ensemble_set_times")
                self._times = times

    def info(self):
        nw, nb = ctypes.c_int32(), ctypes.c_int32()
        L.check(L.lib().gpo_ensemble_info(self.handle, ctypes.byref(nw), ctypes.byref(nb)), "gpo_ensemble_info")
        return nw.value, nb.value

    def device_draws(self):
        out = (ctypes.c_uint32 * self.n)()
        L.check(L.lib().gpo_ensemble_draws(self.handle, out, self._sh()), "gpo_ensemble_draws")
        return list(out)

    # --------------------------------------------------------------- stepping
    def launch(self, n_steps, graph_steps=0):
        """Enqueue n_steps steps on the ensemble stream, with no host
        bookkeeping (members' t, draw and ghost cells are not updated)."""
        if graph_steps and graph_steps > 0:
            if self._graph_steps != graph_steps:
                L.check(L.lib().gpo_ensemble_capture(self.handle, int(graph_steps), self._sh()),
                        "gpo_ensemble_capture")
                self._graph_steps = graph_steps
            reps, rest = divmod(int(n_steps), int(graph_steps))
            if reps:
                L.check(L.lib().gpo_ensemble_replay(self.handle, reps, self._sh()), "gpo_ensemble_replay")
        else:
            rest = int(n_steps)
        if rest:
            L.check(L.lib().gpo_ensemble_step(self.handle, rest, self._sh()), "gpo_ensemble_step")
        self._times = None   # the device clocks moved on; step() records the times its bookkeeping reaches

    def step(self, n_steps, graph_steps=0, initial_bc=True):
        """n_steps steps of every member, each with its own dt; then the host
        bookkeeping of the per-member path (time, iteration count, draw index,
        drifter time, ghost cells).  initial_bc: write the ghost cells of
        members at t = 0 first, as CDKLM16.step does."""
        if n_steps <= 0:
            return
        if initial_bc:
            for s in self.sims:
                if s.t == 0:
                    s._update_bc_values()
                    s._apply_bc(s.h0)
        self._sync_in()
        self._refresh()
        self.launch(n_steps, graph_steps)
        # ghost cells of the members whose steps do not write them, once
        L.check(L.lib().gpo_ensemble_part(self.handle, 5, self._sh()), "gpo_ensemble_part")
        self._draws = [d + n_steps for d in self._draws] if self.has_noise else self._draws
        self._sync_out()
        for s, dt in zip(self.sims, self._dts):
            s._after_batched_steps(int(n_steps), dt, ghosts_written=True)
        self._times = [s.t for s in self.sims]   # the device clocks advanced the same way

    def _select(self, active=None, scales=None):
        a = None if active is None else (ctypes.c_uint8 * self.n)(*[1 if x else 0 for x in active])
        sc = None if scales is None else (ctypes.c_float * self.n)(*[float(np.float32(x)) for x in scales])
        L.check(L.lib().gpo_ensemble_select(self.handle, a, sc, self._sh()), "gpo_ensemble_select")

    def model_step(self, seqs, noise_scales, write_ghosts=True):
        """One model time step of each member as CDKLM16.dataAssimilationStep
        takes it (SWEsimulators/CDKLM16.py:776-795): member p takes the time
        steps seqs[p] (none: it is left alone) with no model error between
        them, then, where noise_scales[p] is not None, the model error once
        with that perturbation scale.  Members with different step counts
        share launches; each launch acts on the members still stepping.  Host
        bookkeeping (time, iteration count, drifter time, draw index) follows
        the per-member path.  write_ghosts=False leaves the ghost cells of
        wall/periodic members without drifters unwritten: nothing reads them
        (the step kernels resolve those boundaries while loading, the model
        error reads only interior cells), and a later model step writes them
        as the per-member path would."""
        sims = self.sims
        if not any(len(q) for q in seqs):
            return
        for s, q in zip(sims, seqs):
            if q and s.t == 0:   # CDKLM16.step writes the ghost cells at t = 0 first
                s._update_bc_values()
                s._apply_bc(s.h0)
        self._sync_in()
        if self._dts is None or self._draws is None:
            self._dts, self._draws = None, None
            self._refresh_all()
        self._sync_times()
        # groups of consecutive inner steps with the same members and time steps
        k_max = max(len(q) for q in seqs)
        groups = []   # [active, dts, count]
        for j in range(k_max):
            act = tuple(j < len(q) for q in seqs)
            dts = tuple(float(np.float32(q[j])) if a else 1.0 for q, a in zip(seqs, act))
            if groups and groups[-1][0] == act and groups[-1][1] == dts:
                groups[-1][2] += 1
            else:
                groups.append([act, dts, 1])
        noisy = [bool(q) and sc is not None for q, sc in zip(seqs, noise_scales)]
        with_noise = self.has_noise and any(noisy)
        sets = [(g[0], g[1]) for g in groups]
        stepped_set = len(sets)   # the members that step: their ghost cells (part 5)
        sets.append((tuple(bool(q) for q in seqs), tuple([1.0] * self.n)))
        if with_noise:
            sets.append((tuple(noisy), tuple([1.0] * self.n)))
        n = self.n
        act_arr = (ctypes.c_uint8 * (len(sets) * n))(*[1 if a else 0 for st in sets for a in st[0]])
        dt_arr = (ctypes.c_float * (len(sets) * n))(*[d for st in sets for d in st[1]])
        scales = [1.0 if sc is None else sc for sc in noise_scales] if with_noise else [1.0] * n
        sc_arr = (ctypes.c_float * n)(*[float(np.float32(x)) for x in scales])
        lib = L.lib()
        sh = self._sh()
        L.check(lib.gpo_ensemble_prepare_sets(self.handle, len(sets), act_arr, dt_arr, sc_arr, sh),
                "gpo_ensemble_prepare_sets")
        for k, g in enumerate(groups):   # parts 0, 1, 3: substeps, ghost cells and drifters
            L.check(lib.gpo_ensemble_run_set(self.handle, k, g[2], 0b1011, sh), "gpo_ensemble_run_set")
        # ghost cells of the new state where the batched parts do not write them
        # (the per-member path writes them after each step, before the model error)
        if write_ghosts:
            L.check(lib.gpo_ensemble_run_set(self.handle, stepped_set, 1, 0b100000, sh), "gpo_ensemble_run_set")
        if with_noise:   # part 4: the model error with each member's scale
            L.check(lib.gpo_ensemble_run_set(self.handle, len(sets) - 1, 1, 0b10000, sh), "gpo_ensemble_run_set")
            self._scales_dirty = True
        self._times = None   # the sets' part 0 moved any device clocks
        self._sync_out()
        for p, (s, q) in enumerate(zip(sims, seqs)):
            for dt in q:
                s.t += np.float64(np.float32(dt))
                s.num_iterations += 1
                if s.drifters is not None:
                    s.drifter_t += np.float32(dt)
            if self.has_noise and noisy[p]:
                s.model_error.draw += 1
                s.model_error._last_scales = (1.0, noise_scales[p])
                self._draws[p] += 1

    def _refresh_all(self):
        dts = [np.float32(s.dt) for s in self.sims]
        draws = [s.model_error.draw if self.has_noise else 0 for s in self.sims]
        dt_arr = (ctypes.c_float * self.n)(*[float(d) for d in dts])
        dr_arr = (ctypes.c_uint32 * self.n)(*draws)
        L.check(L.lib().gpo_ensemble_update(self.handle, dt_arr, dr_arr, self._sh()), "gpo_ensemble_update")
        self._dts, self._draws = dts, draws

    def destroy(self):
        if self.handle is not None:
            torch.cuda.synchronize(self.device)
            L.lib().gpo_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Autotuner:
    """Picks the work-item height (rows per wavefront) and the persistent grid
    size (waves per SIMD) of a batched ensemble by timing a few steps of each
    candidate on the ensemble's own members.  The members' states, drifter
    positions and draw indices are restored afterwards.  Results are cached
    per set of member shapes (and ghost-cell / drifter work)."""

    CHUNK_ROWS = (4, 8, 12, 16, 24, 32)
    WAVES_PER_SIMD = (2, 4, 5, 6, 8)
    _cache = {}

    @staticmethod
    def key(ens):
        shapes = tuple(sorted((int(s.nx), int(s.ny), bool(BatchedEnsemble._bc_in_batch(s)), bool(s.hasDrifters))
                              for s in ens.sims))
        return (ens.exact, ens.has_noise, shapes)

    @classmethod
    def tune(cls, ens, steps=8, repeats=2, verbose=False):
        k = cls.key(ens)
        if k in cls._cache:
            return cls._cache[k]
        ens._sync_in()
        with torch.cuda.stream(ens.stream):
            saved = [[a.buf.clone() for a in s.h0 + s.h1] for s in ens.sims]
            saved_pos = [s.drifters.positions.clone() if s.hasDrifters else None for s in ens.sims]
        draws0 = [s.model_error.draw if ens.has_noise else 0 for s in ens.sims]
        best, best_t = (0, 0), float("inf")
        try:
            for cr in cls.CHUNK_ROWS:
                for w in cls.WAVES_PER_SIMD:
                    ens.chunk_rows, ens.waves_per_simd = cr, w
                    ens._create()
                    ens.launch(1)
                    ms = float("inf")
                    for _ in range(repeats):  # the fastest of a few timings: shapes differ by a few percent
                        t0 = torch.cuda.Event(enable_timing=True)
                        t1 = torch.cuda.Event(enable_timing=True)
                        t0.record(ens.stream)
                        ens.launch(steps)
                        t1.record(ens.stream)
                        t1.synchronize()
                        ms = min(ms, t0.elapsed_time(t1) / steps)
                    if verbose:
                        print(f"autotune chunk_rows={cr} waves_per_simd={w}: {ms:.3f} ms/step", flush=True)
                    if ms < best_t:
                        best, best_t = (cr, w), ms
        finally:
            with torch.cuda.stream(ens.stream):
                for s, bufs, pos in zip(ens.sims, saved, saved_pos):
                    for a, b in zip(s.h0 + s.h1, bufs):
                        a.buf.copy_(b)
                    if pos is not None:
                        s.drifters.positions.copy_(pos)
            ens.stream.synchronize()
            if ens.has_noise:
                for s, d in zip(ens.sims, draws0):
                    s.model_error.draw = d
            ens.destroy()
        cls._cache[k] = best
        return best


class OceanModelEnsemble:
    """numParticles CDKLM16 members on one GPU (OceanModelEnsemble.py:34-218)."""

    def __init__(self, gpu_ctx, sim_args, data_args, model_error_args, numParticles,
                 observation_variance=0.01 ** 2,
                 initialization_variance_factor_ocean_field=0.0,
                 super_dir_name=None, netcdf_filename=None,
                 rank=0, batched=True, autotune=False, graph_steps=0):
        self.logger = logging.getLogger(__name__)
        self.gpu_ctx = gpu_ctx
        self.sim_args = sim_args
        self.data_args = data_args
        self.numParticles = numParticles
        self.observation_variance = observation_variance
        self.initialization_variance_factor_ocean_field = initialization_variance_factor_ocean_field
        if np.isscalar(self.observation_variance):
            self.observation_cov = np.eye(2) * self.observation_variance
            self.observation_cov_inverse = np.eye(2) * (1.0 / self.observation_variance)
        else:
            self.observation_cov = self.observation_variance
            self.observation_cov_inverse = np.linalg.inv(self.observation_cov)

        per_member = isinstance(data_args, (list, tuple))
        if per_member:
            assert len(data_args) == numParticles, "one data_args dict per member"
        self.particles = [None] * numParticles
        self.particleInfos = [None] * numParticles
        self.drifterForecast = [None] * numParticles
        for i in range(numParticles):
            d = data_args[i] if per_member else data_args
            self.particles[i] = CDKLM16(self.gpu_ctx, **self.sim_args, **d, local_particle_id=i,
                                        super_dir_name=super_dir_name, netcdf_filename=netcdf_filename)
            if 'small_scale_perturbation_amplitude' in model_error_args:
                self.particles[i].setSOARModelError(**model_error_args)
            elif 'kl_scaling' in model_error_args:
                raise NotImplementedError("KL model error is not part of gpuocean_amd")
            self.particleInfos[i] = ParticleInfo()
            if self.initialization_variance_factor_ocean_field != 0.0:
                self.particles[i].perturbState(q0_scale=self.initialization_variance_factor_ocean_field)

        self.nx, self.ny = self.particles[0].nx, self.particles[0].ny
        self.dx, self.dy = self.particles[0].dx, self.particles[0].dy
        self.t = self.particles[0].t
        self.particlesActive = [True] * numParticles

        self.batched = batched
        self.autotune = autotune
        self.graph_steps = graph_steps
        self._ensemble = None
        self._obs_table = observe._Table()

    def getNumParticles(self):
        return self.numParticles

    def _batched_ensemble(self):
        if not self.batched:
            return None
        if self._ensemble is None:
            ok, why = BatchedEnsemble.eligible(self.particles)
            if not ok:
                self.logger.debug("stepping members one by one: %s", why)
                return None
            self._ensemble = BatchedEnsemble(self.particles, autotune=self.autotune)
        return self._ensemble

    def attachDrifters(self, drifter_positions):
        if self._ensemble is not None:
            self._ensemble.destroy()
            self._ensemble = None
        for i in range(self.numParticles):
            if len(drifter_positions) > 0:
                p = self.particles[i]
                drifters = GPUDrifterCollection(self.gpu_ctx, len(drifter_positions),
                                                observation_variance=self.observation_variance,
                                                boundaryConditions=p.boundary_conditions,
                                                domain_size_x=p.nx * p.dx, domain_size_y=p.ny * p.dy,
                                                gpu_stream=p.gpu_stream)
                drifters.setDrifterPositions(drifter_positions)
                p.attachDrifters(drifters)
                self.drifterForecast[i] = Observation()
                self.drifterForecast[i].add_observation_from_sim(p)

    def cleanUp(self):
        if self._ensemble is not None:
            self._ensemble.destroy()
            self._ensemble = None
        for p in self.particles:
            if p is not None:
                p.cleanUp(do_gc=False)

    def stepToObservation(self, observation_time, write_now=False):
        """dataAssimilationStep on every active member (OceanModelEnsemble.py:121-134).
        When the active members share their time, model time step and step
        count (total_time_steps), the whole cadence runs batched: the same
        model steps, each member's own dt within them, the model error after
        each, the dt updates every fifth model step, with every member's
        results equal to its own dataAssimilationStep."""
        active = [p for p in range(self.numParticles) if self.particlesActive[p]]
        ens = self._batched_ensemble() if active else None
        if ens is not None and self._da_batchable(active):
            self._batched_da_step(ens, active, observation_time)
        else:
            for p in active:
                self.particles[p].dataAssimilationStep(observation_time, write_now=write_now)
        self.t = observation_time

    def _da_batchable(self, active):
        s0 = self.particles[active[0]]
        for p in active:
            s = self.particles[p]
            if s.t != s0.t or s.model_time_step != s0.model_time_step or s.total_time_steps != s0.total_time_steps:
                return False
        return True

    def _batched_da_step(self, ens, active, observation_time, courant_number=0.8, model_error_final_step=True):
        """CDKLM16.dataAssimilationStep (SWEsimulators/CDKLM16.py:742-806) of
        the given members at once."""
        s0 = self.particles[active[0]]
        mts = s0.model_time_step
        full = int(round(observation_time - s0.t) / mts)
        leftover = observation_time - s0.t - full * mts
        if leftover / mts < 0.1 and full > 1:
            leftover += mts
            full -= 1
        if leftover / mts < 0.00001:
            leftover = 0
        assert full > 0, "There is less than CDKLM16.model_time_step until the observation"
        self.updateDt(courant_number=courant_number, members=active)
        on = set(active)
        for i in range(full + 1):
            if i == 0 and leftover == 0:
                continue
            target = leftover if i == 0 else mts
            scale = np.sqrt(leftover / mts) if i == 0 else 1.0
            noise = i == 0 or i < full or model_error_final_step
            seqs = [self.particles[p].local_dts(target) if p in on else [] for p in range(self.numParticles)]
            scales = [scale if (p in on and noise) else None for p in range(self.numParticles)]
            ens.model_step(seqs, scales, write_ghosts=(i == full))
            for p in active:
                self.particles[p].total_time_steps += 1
            if self.particles[active[0]].total_time_steps % 5 == 0:
                self.updateDt(courant_number=courant_number, members=active)
        for p in active:
            s = self.particles[p]
            assert round(observation_time) == round(s.t), \
                "The simulation time is not the same as observation time after dataAssimilationStep! " \
                f"(self.t, observation_time, diff): {(s.t, observation_time, s.t - observation_time)}"

    def modelStep(self, sub_t, rank="", update_dt=True):
        """Every member steps sub_t seconds, then (update_dt) takes a new dt
        (OceanModelEnsemble.py:138-150).  The leading full-dt steps common to
        all members run batched; each member's remaining steps (a shorter last
        step, or extra steps of a member with a smaller dt) run through its own
        per-step path."""
        ens = self._batched_ensemble()
        if ens is None:
            for p in self.particles:
                self.t = p.step(sub_t)
                if update_dt:
                    p.updateDt()
            return self.t
        seqs = [p.local_dts(sub_t) for p in self.particles]
        n_full = []
        for p, seq in zip(self.particles, seqs):
            full = np.float32(p.dt)
            k = 0
            while k < len(seq) and seq[k] == full:
                k += 1
            n_full.append(k)
        k_common = min(n_full)
        for p in self.particles:
            if p.t == 0:
                p._update_bc_values()
                p._apply_bc(p.h0)
        ens.step(k_common, graph_steps=self.graph_steps, initial_bc=False)
        for p, seq in zip(self.particles, seqs):
            if len(seq) > k_common:
                p.step_dts(seq[k_common:])
            self.t = p.t
        if update_dt:
            self.updateDt()
        return self.t

    def updateDt(self, courant_number=None, members=None):
        """New dt for every member (or the given member indices) from one
        batched set of CFL reductions and a single read-back (replaces the
        per-member blocking updateDt loop)."""
        sims = self.particles if members is None else [self.particles[p] for p in members]
        n = len(sims)
        dev = self.particles[0].device
        out = torch.empty(n, dtype=torch.float32, device=dev)
        stream = torch.cuda.Stream(device=dev)
        sh = ctypes.c_void_p(stream.cuda_stream)
        for p in sims:
            stream.wait_stream(p.gpu_stream)
        lib = L.lib()
        with torch.cuda.stream(stream):
            for i, p in enumerate(sims):
                L.check(lib.gpo_max_dt(int(p.nx), int(p.ny), float(p.dx), float(p.dy), float(p.g),
                                       float(p.mask_value), _vp(p.h0[0]), _vp(p.h0[1]), _vp(p.h0[2]), p.h0[0].pitch,
                                       _vp(p.Bm), p.Bm.pitch, ctypes.c_void_p(out.data_ptr() + 4 * i), sh),
                        "gpo_max_dt")
            host = out.cpu().numpy()
        for i, p in enumerate(sims):
            c = p.courant_number if courant_number is None else courant_number
            p.dt = c * host[i:i + 1][0]

    def _cell_sizes(self):
        """The dx, dy the reference divides positions by: data_args' own."""
        if isinstance(self.data_args, (list, tuple)):
            return None, None
        return self.data_args.get('dx'), self.data_args.get('dy')

    def observeParticles(self, drifter_positions):
        """hu, hv at the cells holding the given positions, per member:
        (numParticles, num_drifters, 2) (OceanModelEnsemble.py:168-198).
        One observation launch over all members and one read-back; the cell
        is np.floor(x/dx) on the interior, as the reference indexes it."""
        drifter_positions = np.asarray(drifter_positions)
        if len(drifter_positions) == 0:
            return np.empty((self.numParticles, 0, 2))
        dx, dy = self._cell_sizes()
        return observe.observe_members(self.particles, observe.FLOW, positions=drifter_positions, dx=dx, dy=dy,
                                       table=self._obs_table)

    def dumpParticleSample(self, drifter_cells):
        """(eta, hu, hv) of every member at drifter_cells (and each member's
        extra cells) into its ParticleInfo (OceanModelEnsemble.py:160-162),
        from one observation launch when the members share their extra cells."""
        drifter_cells = np.asarray(drifter_cells)
        extras = [pi.extraCells for pi in self.particleInfos]
        same = all((e is None and extras[0] is None) or
                   (e is not None and extras[0] is not None and np.array_equal(e, extras[0])) for e in extras)
        if not same:
            for i in range(self.numParticles):
                self.particleInfos[i].add_state_sample_from_sim(self.particles[i], drifter_cells)
            return
        cells = drifter_cells if extras[0] is None else np.concatenate([drifter_cells, np.asarray(extras[0])])
        values = observe.observe_members(self.particles, observe.STATE, cells=cells, table=self._obs_table)
        D = drifter_cells.shape[0]
        for i in range(self.numParticles):
            extra = None if extras[0] is None else values[i, D:].copy()
            self.particleInfos[i]._add(self.particles[i].t, values[i, :D].copy(), extra)

    def dumpForecastParticleSample(self):
        """Each member's drifter positions into its forecast Observation
        (OceanModelEnsemble.py:164-166)."""
        for i in range(self.numParticles):
            self.drifterForecast[i].add_observation_from_sim(self.particles[i])

    def dumpParticleInfosToFiles(self, filename_prefix):
        """Each member's state samples to <prefix>_<member>.bz2 (OceanModelEnsemble.py:200-206)."""
        for p in range(self.getNumParticles()):
            self.particleInfos[p].to_pickle(filename_prefix + "_" + str(p) + ".bz2")

    def dumpDrifterForecastToFiles(self, filename_prefix):
        """Each member's drifter forecast to <prefix>_<member>.bz2 (OceanModelEnsemble.py:208-214)."""
        for p in range(self.getNumParticles()):
            self.drifterForecast[p].to_pickle(filename_prefix + "_" + str(p) + ".bz2")

    def syncGPU(self):
        for p in self.particles:
            p.gpu_stream.synchronize()

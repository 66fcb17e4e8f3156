/*
 * gpuocean_amd -- C ABI of the MI355X (gfx950) CDKLM16 step library.
 *
 * Plain C types only: device pointers come from the caller (torch tensors or
 * hipMalloc), pitches are in BYTES as in the reference's PyCUDA calls, and
 * `stream` is a hipStream_t passed as void*.  Every entry point returns 0 on
 * success or a hipError_t / negative gpo error code; gpo_last_error() gives
 * the message.  Nothing here allocates device memory unless stated.
 *
 * Reference interfaces replaced (paths relative to the reference root
 * src/gpuocean/):
 *   - the JIT-time #defines of CDKLM16_kernel.cu become gpo_cdklm16_params
 *     fields (SWEsimulators/CDKLM16.py:242-269);
 *   - each entry point names the PyCUDA prepared_async_call it replaces.
 */
#ifndef GPUOCEAN_AMD_H
#define GPUOCEAN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes beyond hipError_t. */
#define GPO_ERR_BAD_ARGUMENT (-1)
#define GPO_ERR_UNSUPPORTED (-2)

/* Compile-time constants of the reference kernel, passed at run time.
 * SWEsimulators/CDKLM16.py:242-269 (the `defines` dict). */
typedef struct {
  int32_t nx, ny;                 /* interior cells; arrays carry 2 ghost cells per side */
  int32_t rk_order;               /* 1, 2 or 3 (RK_ORDER) */
  int32_t one_dimensional;        /* ONE_DIMENSIONAL */
  int32_t has_land;               /* Hi holds land_value corners (masked bathymetry) */
  int32_t f_is_array;             /* coriolis f given per cell (see gpo_coriolis_cells) */
  int32_t angle_is_zero;          /* north vector is exactly (0, 1) */
  int32_t chunk_rows;             /* rows marched per wavefront (tuning; 0 = default) */
  float dx, dy, g, f, r;          /* DX, DY, GRAV, constant f, FRIC */
  float theta;                    /* THETA */
  float desing_eps;               /* KPSIMULATOR_DESING_EPS */
  float flux_slope_eps;           /* KPSIMULATOR_FLUX_SLOPE_EPS */
  float depth_cutoff;             /* KPSIMULATOR_DEPTH_CUTOFF */
  float flux_balancer;            /* FLUX_BALANCER */
  float land_value;               /* Bathymetry.mask_value (1e20) */
  float north_x, north_y;         /* sin(angle), cos(angle) for a constant angle */
  float wind_stress_factor, rho_o;/* WIND_STRESS_FACTOR, RHO_O */
  int32_t divc_ok;                /* set by gpo_cdklm16_prepare (0 = IEEE division and sqrt only) */
  int32_t waves_per_simd;         /* persistent substep grid, waves per SIMD (tuning; 0 = default) */
} gpo_cdklm16_params;

/* Optional wind stress / atmospheric pressure textures (time levels `cur`
 * and `next`, dense row-major, blended with weight t).  NULL pointer = none.
 * Replaces the wind_stress_* and atmospheric_pressure_* arguments of
 * cdklm_swe_2D (gpu_kernels/CDKLM16_kernel.cu:502-511). */
typedef struct {
  const float *wind_x_cur, *wind_x_next, *wind_y_cur, *wind_y_next;
  int32_t wind_x_nx, wind_x_ny, wind_y_nx, wind_y_ny;
  float wind_t;
  const float *pressure_cur, *pressure_next;
  int32_t pressure_nx, pressure_ny;
  float pressure_t;
} gpo_forcing;

/* Set-up of a parameter block before its first substep: verifies, on the
 * device and for every numerator in {+0} U +-[2^-96, 2^100], that division by
 * each of dx, dy and 2g equals one Markstein correction step with the
 * correctly rounded reciprocal (bits 0-2 of p->divc_ok), and, for every float
 * in [2^-96, FLT_MAX], that the square root from the hardware reciprocal
 * square root and one correction step equals the IEEE square root (bit 3);
 * the exact build uses these shorter sequences only where verified (results
 * are cached per constant and per process).  Synchronises `stream` and
 * allocates a 4-byte counter.  No reference counterpart: the reference
 * compiles its constants into the kernel (SWEsimulators/CDKLM16.py:242-292). */
int gpo_cdklm16_prepare(gpo_cdklm16_params* p, void* stream);

/* One Runge-Kutta substep of CDKLM16 over the interior.
 * Replaces cdklm_swe_2D (gpu_kernels/CDKLM16_kernel.cu:475-515) as launched by
 * CDKLM16.callKernel (SWEsimulators/CDKLM16.py:688-723).
 *   in_*   : stage input Q (the kernel's eta0/hu0/hv0), ghost cells included
 *   qn_*   : Q^n, read at each interior cell when rk_step > 0 (the kernel
 *            reads it from its output arrays); may alias out_*
 *   out_*  : interior cells written
 *   Hi     : bathymetry on corners, (ny+5) rows, pitch Hi_pitch bytes.  The
 *            cell depth Hm/Bm is derived from Hi inside the kernel.
 *   f_cells: per-cell coriolis f, same pitch as the state (f_is_array), else NULL
 *   boundary_conditions: N<<24 | S<<16 | E<<8 | W (CDKLM16.py:695-699)
 *   exact  : 0 = fast path (reciprocal-based division); 1 = IEEE division and
 *            square root with no FMA contraction (bit-compatible with the
 *            host oracle). */
int gpo_cdklm16_substep(const gpo_cdklm16_params* p, float dt, int32_t rk_step,
                        const float* in_eta, const float* in_hu, const float* in_hv,
                        const float* qn_eta, const float* qn_hu, const float* qn_hv,
                        float* out_eta, float* out_hu, float* out_hv, int32_t pitch,
                        const float* Hi, int32_t Hi_pitch, const float* f_cells,
                        const gpo_forcing* forcing, uint32_t boundary_conditions,
                        int32_t exact, void* stream);

/* n_steps full RK2 steps h0 -> h1 -> h0 with a fixed dt, launched back to back
 * (two substep launches per step) without host round trips.  Only wall and
 * periodic boundaries, which the substep kernel resolves while loading; the
 * global ghost cells are not written (call gpo_boundary_conditions after).
 * Replaces the per-step host loop of CDKLM16.step (SWEsimulators/CDKLM16.py:493-579)
 * for steps with no noise, drifters, forcing updates or FRS boundaries. */
int gpo_cdklm16_rk2_steps(const gpo_cdklm16_params* p, int32_t n_steps, float dt,
                          float* h0_eta, float* h0_hu, float* h0_hv,
                          float* h1_eta, float* h1_hu, float* h1_hv, int32_t pitch,
                          const float* Hi, int32_t Hi_pitch, const float* f_cells,
                          uint32_t boundary_conditions, int32_t exact, void* stream);

/* Ghost-cell boundary conditions on a state triple.
 * Replaces BoundaryConditionsArakawaA.boundaryCondition
 * (utils/Common.py:1310-1330) with periodicBoundary_NS/_EW
 * (gpu_kernels/boundary_kernels.cu:36-110), linearInterpolation_NS/_EW
 * (boundary_kernels.cu:265-384) and flowRelaxationScheme_NS/_EW
 * (boundary_kernels.cu:406-540).  Per axis: periodic, or else flow relaxation
 * on code-3 sides followed by linear interpolation on code-4 sides.  Codes 1
 * (wall) write nothing, as in the reference.  frs_* may be NULL when no side
 * is 3.
 *   frs_ns_* : [2][frs_ns_nx][3] (south, north), frs_ew_* : [frs_ew_ny][2][3]. */
int gpo_boundary_conditions(int32_t nx, int32_t ny, uint32_t boundary_conditions,
                            const int32_t sponge_cells[4] /* n, e, s, w */,
                            float* eta, float* hu, float* hv, int32_t pitch,
                            const float* frs_ns_cur, const float* frs_ns_next, int32_t frs_ns_nx,
                            const float* frs_ew_cur, const float* frs_ew_next, int32_t frs_ew_ny,
                            float frs_t, void* stream);

/* Bathymetry set-up: ghost corners of Hi (periodic or closed) then Bm.
 * Replaces Bathymetry.__init__ (utils/Common.py:1414-1491) with
 * periodic_/closed_boundary_intersections_* (boundary_kernels.cu:122-246) and
 * initBm (gpu_kernels/initHm_kernel.cu:38-106).  Hi is modified in place. */
int gpo_bathymetry_init(int32_t nx, int32_t ny, uint32_t boundary_conditions, float land_value,
                        float* Hi, int32_t Hi_pitch, float* Bm, int32_t Bm_pitch, void* stream);

/* Per-cell coriolis f from a texture, coriolisF evaluated at every cell
 * (gpu_kernels/CDKLM16_kernel.cu:48-57, interpolation.cu:137-190). */
int gpo_coriolis_cells(int32_t nx, int32_t ny, const float* f_tex, int32_t f_nx, int32_t f_ny,
                       float* f_cells, int32_t pitch, void* stream);

/* CFL time step: writes min over wet interior cells of
 * 0.25*min(dx/|u+-c|, dy/|v+-c|) (1e5 if none) to *dt_out (device pointer).
 * Replaces per_block_max_dt + max_dt_reduction (gpu_kernels/max_dt.cu:42-195)
 * called from CDKLM16.updateDt (SWEsimulators/CDKLM16.py:814-841). */
int gpo_max_dt(int32_t nx, int32_t ny, float dx, float dy, float g, float land_value,
               const float* eta, const float* hu, const float* hv, int32_t pitch,
               const float* Bm, int32_t Bm_pitch, float* dt_out, void* stream);

/* Global mass: *sum_out (device double) = sum of the interior eta, summed in
 * FP64.  The conserved quantity of wall/periodic runs
 * (test/schemes/ConservationOfMass_test.py:126-134 sums eta on the host). */
int gpo_mass_sum(int32_t nx, int32_t ny, const float* eta, int32_t pitch, double* sum_out, void* stream);

/* ---- stochastic model error (SOAR) ---------------------------------------- */

/* Raw Philox4x32-10 blocks (Salmon et al., SC'11): out[4i..4i+3] for counter
 * ctr[4i..4i+3] and key key[2i..2i+1], device pointers.  Exposed for the
 * known-answer test of the generator that replaces the reference's curand
 * XORWOW / LCG streams (utils/RandomNumbers.py:36-185). */
int gpo_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int32_t n, void* stream);

/* Standard normal field: out(j, i) from Philox counter (idx/4, draw, stream_id,
 * tag) with idx = j*cols + i, key = seed, Box-Muller on word pairs.  Replaces
 * RandomNumbers.generateNormalDistribution (utils/RandomNumbers.py:168-185). */
int gpo_random_normal(float* out, int32_t rows, int32_t cols, int32_t pitch, uint64_t seed, uint32_t stream_id,
                      uint32_t draw, void* stream);

/* SOAR-correlated field on the (ny+4) x (nx+4) coarse grid from xi with a
 * 5x5 stencil of weights q0 (1 + d/L) exp(-d/L).  Replaces the SOAR kernel
 * (gpu_kernels/ocean_noise.cu:141-241) launched by
 * OceanStateNoise.perturbOceanState (SWEsimulators/OceanStateNoise.py:356-459).
 * xi: (ny+8) x (nx+8) (or ny / nx rows/columns on a periodic axis). */
int gpo_soar(int32_t nx, int32_t ny, const float* weights25, float scale, int32_t periodic_ns,
             int32_t periodic_ew, const float* xi, int32_t xi_pitch, float* coarse, int32_t coarse_pitch,
             int32_t additive, void* stream);

/* Adds the coarse d_eta and its geostrophically balanced hu, hv to the
 * interior state.  Replaces geostrophicBalance (gpu_kernels/ocean_noise.cu:256-390). */
int gpo_geostrophic_balance(int32_t nx, int32_t ny, float dx, float dy, float g, float f, float beta,
                            float north_x, float north_y, const float* f_cells, int32_t f_pitch,
                            const float* coarse, int32_t coarse_pitch, float* eta, float* hu, float* hv,
                            int32_t pitch, const float* Hi, int32_t Hi_pitch, float land_value, void* stream);

/* Adds a perturbation generated on a coarse grid (coarse_nx x coarse_ny cells
 * of coarse_dx x coarse_dy, buffer (coarse_ny+4) x (coarse_nx+4), pitch
 * coarse_pitch bytes) to the interior state: bicubic interpolation from the
 * coarse cell centres (shifted by offset_i, offset_j fine cells) onto the fine
 * cell centres, and the geostrophically balanced hu, hv from the still depth;
 * cells with a land corner are left unchanged.  Two ghost cells on both grids.
 * Replaces bicubicInterpolation (gpu_kernels/ocean_noise.cu:401-659) launched by
 * OceanStateNoise.perturbOceanState when interpolation_factor > 1
 * (SWEsimulators/OceanStateNoise.py:418-441). */
int gpo_bicubic_interpolation(int32_t nx, int32_t ny, float dx, float dy, int32_t coarse_nx, int32_t coarse_ny,
                              float coarse_dx, float coarse_dy, int32_t offset_i, int32_t offset_j, float g, float f,
                              float beta, float north_x, float north_y, const float* f_cells, int32_t f_pitch,
                              const float* coarse, int32_t coarse_pitch, float* eta, float* hu, float* hv,
                              int32_t pitch, const float* Hi, int32_t Hi_pitch, float land_value, void* stream);

/* ---- drifters ----------------------------------------------------------------- */

/* One forward-Euler step of n_drifters + 1 passive drifters (the last one is the
 * observation drifter) with bilinear cell-centre velocity.  Replaces
 * passiveDrifterKernel (gpu_kernels/driftKernels.cu:74-183) launched by
 * GPUDrifterCollection.drift (drifters/GPUDrifterCollection.py:210-232).
 * positions: (n_drifters+1) rows of (x, y), pitch pos_pitch bytes; Hm is the
 * cell-centre depth Bm. */
int gpo_drift(int32_t nx, int32_t ny, float dx, float dy, float dt, int32_t x0ref, int32_t y0ref, const float* eta,
              const float* hu, const float* hv, int32_t pitch, const float* Hm, int32_t Hm_pitch,
              int32_t periodic_ns, int32_t periodic_ew, int32_t n_drifters, float* positions, int32_t pos_pitch,
              float sensitivity, const float* wind_x_cur, const float* wind_x_next, const float* wind_y_cur,
              const float* wind_y_next, int32_t wind_nx, int32_t wind_ny, float wind_t, float wind_drift_factor,
              void* stream);

/* ---- observation operators ---------------------------------------------------- */

/* Mirror of observeUnderlyingFlow (gpu_kernels/observationKernels.cu:29-80),
 * launched by BaseOceanStateEnsemble.observeParticles(gpu=True)
 * (ensembles/BaseOceanStateEnsemble.py:117-131): (hu, hv) of the cell
 * floor(x/dx) + x0ref, floor(y/dy) + y0ref (FP32 quotient) for each of the
 * n_drifters positions (rows of (x, y), pos_pitch bytes), written to rows of
 * (hu, hv) floats, obs_pitch bytes.  eta and H are accepted and unused, as in
 * the reference.  A position whose cell falls outside the (ny+4) x (nx+4)
 * array gives NaN (the reference reads outside the array there). */
int gpo_observe_underlying_flow(int32_t nx, int32_t ny, float dx, float dy, int32_t x0ref, int32_t y0ref,
                                const float* eta, int32_t eta_pitch, const float* hu, int32_t hu_pitch,
                                const float* hv, int32_t hv_pitch, float H, int32_t n_drifters,
                                const float* positions, int32_t pos_pitch, float* observation, int32_t obs_pitch,
                                void* stream);

/* One member observed by gpo_observe: its state (rows of `pitch` bytes,
 * ghost cells included), cell-centre depth Bm (GPO_OBS_BUOY only) and where to
 * observe it: either `points` (n_points (x, y) pairs in metres, FP64) or
 * `cells` (n_points (x, y) interior cell indices); members may share them. */
typedef struct {
  const float* eta;
  const float* hu;
  const float* hv;
  int32_t pitch;
  const float* Bm;
  int32_t Bm_pitch;
  int32_t nx, ny;
  double dx, dy;
  const double* points;
  const int32_t* cells;
} gpo_obs_member;

#define GPO_OBS_FLOW 0       /* (hu, hv) */
#define GPO_OBS_STATE 1      /* (eta, hu, hv) */
#define GPO_OBS_BUOY 2       /* (hu, hv) * H / (H + eta) in FP32 */
#define GPO_OBS_INNOVATION 3 /* (hu_obs, hv_obs) * (D + eta) / D - (hu, hv), D = mean_depth (1 if 0) */
#define GPO_OBS_GHOSTED 16   /* or-ed in: observeUnderlyingFlow's cell rule instead of numpy's */

/* Observation operator over many members in one launch.  members: device
 * array of n_members entries.  out: device doubles, (n_members, n_points, k)
 * with k = 3 for GPO_OBS_STATE, else 2.  Cells follow numpy indexing of the
 * interior arrays, as the reference's host paths do: floor(x/dx) in FP64 (or
 * the cell given), -n..n-1 with negative indices counting from the end.
 * A point outside that range gives NaN and adds one to *n_bad (device int,
 * may be NULL), where numpy raises IndexError.  observed (GPO_OBS_INNOVATION):
 * device doubles, n_points rows of (x, y, hu, hv) as Observation.get_observation
 * returns them.  Replaces the per-member download-and-index loops of
 * OceanModelEnsemble.observeParticles (ensembles/OceanModelEnsemble.py:168-198),
 * BaseOceanStateEnsemble.observeParticles (ensembles/BaseOceanStateEnsemble.py:79-159),
 * ParticleInfo.add_state_sample_from_sim (utils/ParticleInfo.py:81-111) and the
 * buoy readings of Observation.add_observation_from_sim (utils/Observation.py:137-180). */
int gpo_observe(const gpo_obs_member* members, int32_t n_members, int32_t n_points, int32_t mode,
                const double* observed, float mean_depth, double* out, int32_t* n_bad, void* stream);

/* ---- batched ensemble ------------------------------------------------------- */

/* One member of a batched ensemble: its parameter block (after
 * gpo_cdklm16_prepare), time step, double-buffered state and, when the
 * ensemble carries SOAR model error, the buffers and constants that
 * OceanStateNoise.perturbSim would use for it.  Device pointers; pitches in
 * bytes. */
typedef struct {
  gpo_cdklm16_params params;
  float dt;
  float *h0_eta, *h0_hu, *h0_hv;   /* Q^n in, Q^{n+1} out */
  float *h1_eta, *h1_hu, *h1_hv;   /* RK stage */
  int32_t pitch;
  const float* Hi;
  int32_t Hi_pitch;
  const float* f_cells;            /* per-cell f (pitch = pitch) or NULL */
  uint32_t boundary_conditions;    /* wall (1), periodic (2), flow relaxation (3) */
  /* model error (noise != 0) */
  int32_t noise;
  float* xi;                       /* standard normal field, xi_rows x xi_cols */
  int32_t xi_rows, xi_cols, xi_pitch;
  float* coarse;                   /* (ny+4) x (nx+4) SOAR field */
  int32_t coarse_pitch;
  uint64_t seed;
  uint32_t stream_id;
  uint32_t draw;                   /* index of the member's next random field */
  float soar_weights[25];
  float perturbation_scale;
  float g, f, beta;                /* as passed to gpo_geostrophic_balance */
  /* ghost cells (bc_each_step != 0): written after RK substep 0 (on h1, when
   * a side is flow relaxation) and after each step and its model error (on
   * h0), as gpo_boundary_conditions writes them; otherwise never */
  int32_t bc_each_step;
  int32_t sponge[4];               /* N, E, S, W sponge widths */
  const float *frs_ns_cur, *frs_ns_next;  /* [2][frs_ns_nx][3] when N or S is 3 */
  int32_t frs_ns_nx;
  const float *frs_ew_cur, *frs_ew_next;  /* [frs_ew_ny][2][3] when E or W is 3 */
  int32_t frs_ew_ny;
  float frs_t;                     /* weight of next against cur */
  /* passive drifters (drifters != 0): n_drifters + 1 rows of (x, y), moved
   * after each step (after the ghost cells) with the member's dt, no wind */
  int32_t drifters;
  int32_t n_drifters;
  float* positions;
  int32_t pos_pitch;
  float sensitivity;
  const float* Bm;                 /* cell-centre depth, (ny+4) rows */
  int32_t Bm_pitch;
  /* model error on a coarser grid (OceanStateNoise interpolation_factor > 1):
   * coarse_nx > 0 gives its interior size and spacing and the fine-cell offset
   * of gpo_bicubic_interpolation; xi is then (coarse_ny+8) x (coarse_nx+8)
   * (coarse_ny x coarse_nx on periodic axes) and coarse (coarse_ny+4) x
   * (coarse_nx+4); coarse_nx = 0: the model error lives on the grid itself */
  int32_t coarse_nx, coarse_ny;
  float coarse_dx, coarse_dy;
  int32_t coarse_offset_i, coarse_offset_j;
  /* flow-relaxation data as time levels (frs_n_times >= 1): every level on
   * the device; each step takes its two levels and weight from the member's
   * model time t as BoundaryConditionsArakawaA.update_bc_values does
   * (utils/Common.py:1193-1308, in FP64), and t advances by the member's dt.
   * frs_*_cur/next and frs_t are then not used.  frs_n_times = 0: the data
   * of frs_*_cur/next with the fixed weight frs_t. */
  int32_t frs_n_times;
  const double* frs_times;         /* frs_n_times ascending times */
  const float* frs_ns_levels;      /* frs_n_times x [2][frs_ns_nx][3], or NULL */
  const float* frs_ew_levels;      /* frs_n_times x [frs_ew_ny][2][3], or NULL */
  double t;                        /* the member's model time */
} gpo_member;

typedef void* gpo_ensemble;

/* Builds a batched ensemble over n RK2 members of any sizes: one persistent
 * launch per RK substep walks a table of (member, strip, row chunk) items;
 * with model error, one fused launch (normal field, SOAR, geostrophic
 * balance) covers every member; members with bc_each_step get their ghost
 * cells from two batched launches (north/south, then east/west) after
 * substep 0 (flow relaxation only) and after the model error; members with
 * drifters are moved by one batched launch at the end of the step.  Each
 * member's draw index lives on the device and advances once per step, so
 * replays of a captured graph draw fresh random numbers.  Results are bit for
 * bit those of the per-member sequence gpo_cdklm16_substep, (gpo_boundary_conditions),
 * gpo_cdklm16_substep, gpo_random_normal + gpo_soar + gpo_geostrophic_balance
 * (gpo_bicubic_interpolation in place of the balance for coarse-grid model
 * error, run as three batched launches: normal fields, SOAR, interpolation),
 * (gpo_boundary_conditions), (gpo_drift).  chunk_rows (0 = each member's own)
 * and waves_per_simd (0 = 4) set the work-item height and the persistent grid.
 * Allocates the device tables; synchronises `stream`.
 * Replaces the member loop of OceanModelEnsemble.modelStep
 * (ensembles/OceanModelEnsemble.py:138-150, 331-343). */
int gpo_ensemble_create(const gpo_member* members, int32_t n, int32_t chunk_rows, int32_t waves_per_simd,
                        int32_t exact, gpo_ensemble* out, void* stream);

/* New time steps (dts: n host floats) and/or draw indices (draws: n host
 * uint32) of every member; either may be NULL.  Synchronises `stream`.  A
 * captured graph stays valid (it reads both tables at run time). */
int gpo_ensemble_update(gpo_ensemble e, const float* dts, const uint32_t* draws, void* stream);

/* Model times (n host doubles, one per member) of the members whose
 * flow-relaxation data has several times, as they stand before the next
 * step; the others' entries are ignored.  Only the times change: each
 * member's boundary-data levels and weight are derived at its new time on
 * the device, so ghost-cell parts launched before the next step read valid
 * data.  No reference counterpart: the reference keeps each member's time on
 * the host.  Synchronises `stream`. */
int gpo_ensemble_set_times(gpo_ensemble e, const double* t, void* stream);

/* Number of work items and persistent workgroups of an ensemble. */
int gpo_ensemble_info(gpo_ensemble e, int32_t* n_work, int32_t* n_blocks);

/* n_steps steps of every member, launched directly on `stream`.  Ghost cells
 * of wall/periodic members without drifters are not written (part 5 writes
 * them, once after a run of steps). */
int gpo_ensemble_step(gpo_ensemble e, int32_t n_steps, void* stream);

/* One part of a step: part 0 and 1 = RK substep 0 and 1 (with the
 * flow-relaxation ghost cells of the stage), part 2 = the model error (random
 * field, SOAR, balance; nothing without model error), then the ghost cells of
 * the new state and the drifters.  Parts 0, 1, 2 in this order make one
 * gpo_ensemble_step.  Part 3 = part 2 without the model error (a step with
 * apply_stochastic_term=False), part 4 = the model error alone
 * (CDKLM16.perturbState after such steps, as dataAssimilationStep does,
 * SWEsimulators/CDKLM16.py:776-795), part 5 = the ghost cells of the state
 * of members with a periodic side whose steps do not write them (two batched
 * launches, north/south then east/west): the gpo_boundary_conditions call
 * CDKLM16.step ends each step with (SWEsimulators/CDKLM16.py:569-575), made
 * once after a run of steps, since nothing in between reads those cells.
 * Parts act on the selected members. */
int gpo_ensemble_part(gpo_ensemble e, int32_t part, void* stream);

/* Selects the members the following parts act on (active: n host bytes,
 * nonzero = selected; NULL = all members) and/or sets each member's model-error
 * scale (noise_scale: n host floats, the perturbation_scale of
 * OceanStateNoise.perturbSim; NULL = unchanged).  Unselected members are not
 * stepped, perturbed or drifted and keep their draw index.  Graph capture and
 * replay need all members selected.  Synchronises `stream`.  Lets members with
 * different numbers of time steps to an observation time share launches
 * (OceanModelEnsemble.stepToObservation, ensembles/OceanModelEnsemble.py:121-135). */
int gpo_ensemble_select(gpo_ensemble e, const uint8_t* active, const float* noise_scale, void* stream);

/* Prepares n_sets selections at once, for gpo_ensemble_run_set: set k selects
 * the members with active[k*n + m] != 0 and steps them with time steps
 * dts[k*n + m] (NULL: each member's current dt); noise_scale (n floats, may be
 * NULL) sets every member's model-error scale.  One upload and one
 * synchronisation of `stream` for all sets (allocates device memory when the
 * sets outgrow the previous ones).  No reference counterpart: it batches the
 * member loop of OceanModelEnsemble.stepToObservation
 * (ensembles/OceanModelEnsemble.py:121-135). */
int gpo_ensemble_prepare_sets(gpo_ensemble e, int32_t n_sets, const uint8_t* active, const float* dts,
                              const float* noise_scale, void* stream);

/* n_steps steps of prepared set `set`, each made of the parts whose bit is set
 * in `parts` (bit p = gpo_ensemble_part p), launched without host
 * synchronisation.  Afterwards all members are selected again. */
int gpo_ensemble_run_set(gpo_ensemble e, int32_t set, int32_t n_steps, uint32_t parts, void* stream);

/* Captures steps_per_graph steps on `stream` into a hipGraph (replacing any
 * earlier one); gpo_ensemble_replay launches it n_replays times. */
int gpo_ensemble_capture(gpo_ensemble e, int32_t steps_per_graph, void* stream);
int gpo_ensemble_replay(gpo_ensemble e, int32_t n_replays, void* stream);

/* Draw index of every member into draws[n] (host; synchronises `stream`). */
int gpo_ensemble_draws(gpo_ensemble e, uint32_t* draws, void* stream);

/* Frees the device tables and the graph. */
int gpo_ensemble_destroy(gpo_ensemble e);

/* ---- differentiation ------------------------------------------------------------
 * No counterpart in the reference, which has no derivative of the step.  The
 * differentiable substep covers RK2 substeps of wet cells (every cell
 * h >= desing_eps), wall and periodic boundaries, r = 0, no land, wind or
 * pressure.  Unsupported parameters are rejected; state branches outside the
 * scope met at run time set bits of *flag (1: dry or desingularised cell,
 * 2: depth clamp, 8: face depth clamp), flag may be NULL.  Minmod, max and
 * min differentiate through the first minimal/maximal argument.  The stage
 * arithmetic restates gpu_kernels/CDKLM16_kernel.cu:475-1090. */

/* Workspace bytes of one differentiable substep (elem_bytes 4 or 8). */
size_t gpo_adjoint_workspace_bytes(int32_t nx, int32_t ny, int32_t elem_bytes);
/* Workspace bytes of one tangent-linear substep. */
size_t gpo_tlm_workspace_bytes(int32_t nx, int32_t ny);

/* FP64 substep (the verification build): same arguments as
 * gpo_cdklm16_substep with double arrays (pitches in bytes), Hi and f_cells
 * in double. */
int gpo_cdklm16_substep_f64(const gpo_cdklm16_params* p, double dt, int32_t rk_step, const double* in_eta,
                            const double* in_hu, const double* in_hv, const double* qn_eta, const double* qn_hu,
                            const double* qn_hv, double* out_eta, double* out_hu, double* out_hv, int32_t pitch,
                            const double* Hi, int32_t Hi_pitch, const double* f_cells, uint32_t boundary_conditions,
                            void* workspace, int32_t* flag, void* stream);

/* Tangent-linear FP64 substep: out and d_out = (d out / d (in, qn)) . (d_in, d_qn). */
int gpo_cdklm16_substep_tlm_f64(const gpo_cdklm16_params* p, double dt, int32_t rk_step, const double* in_eta,
                                const double* in_hu, const double* in_hv, const double* d_in_eta,
                                const double* d_in_hu, const double* d_in_hv, const double* qn_eta,
                                const double* qn_hu, const double* qn_hv, const double* d_qn_eta,
                                const double* d_qn_hu, const double* d_qn_hv, double* out_eta, double* out_hu,
                                double* out_hv, double* d_out_eta, double* d_out_hu, double* d_out_hv,
                                int32_t pitch, const double* Hi, int32_t Hi_pitch, const double* f_cells,
                                uint32_t boundary_conditions, void* workspace, int32_t* flag, void* stream);

/* Adjoint of one substep out = S(in [, qn]): given lambda_out (interior
 * cells) and the forward's input and output, writes (accumulate = 0) or adds
 * (accumulate = 1) lambda_in = (dS/din)^T lambda_out on the interior cells,
 * ghost-cell sensitivities folded onto the cells they mirror or wrap; for
 * rk_step 1 also adds (dS/dqn)^T lambda_out to lambda_qn (may be NULL).
 * lambda_in may alias lambda_out.  FP32 (product) and FP64 (verification). */
int gpo_cdklm16_substep_adj_f32(const gpo_cdklm16_params* p, float dt, int32_t rk_step, const float* in_eta,
                                const float* in_hu, const float* in_hv, const float* out_eta,
                                const float* lam_out_eta, const float* lam_out_hu, const float* lam_out_hv,
                                float* lam_in_eta, float* lam_in_hu, float* lam_in_hv, float* lam_qn_eta,
                                float* lam_qn_hu, float* lam_qn_hv, int32_t accumulate, int32_t pitch,
                                const float* Hi, int32_t Hi_pitch, const float* f_cells,
                                uint32_t boundary_conditions, void* workspace, int32_t* flag, void* stream);
int gpo_cdklm16_substep_adj_f64(const gpo_cdklm16_params* p, double dt, int32_t rk_step, const double* in_eta,
                                const double* in_hu, const double* in_hv, const double* out_eta,
                                const double* lam_out_eta, const double* lam_out_hu, const double* lam_out_hv,
                                double* lam_in_eta, double* lam_in_hu, double* lam_in_hv, double* lam_qn_eta,
                                double* lam_qn_hu, double* lam_qn_hv, int32_t accumulate, int32_t pitch,
                                const double* Hi, int32_t Hi_pitch, const double* f_cells,
                                uint32_t boundary_conditions, void* workspace, int32_t* flag, void* stream);

/* Flow-relaxation data of a differentiable run (boundary code 3 sides):
 * sponge widths N, E, S, W and the exterior solution, packed as for
 * gpo_boundary_conditions, with the weight t of next against cur. */
typedef struct {
  int32_t sponge[4];
  const float *ns_cur, *ns_next;
  int32_t ns_nx;
  const float *ew_cur, *ew_next;
  int32_t ew_ny;
  float t;
} gpo_frs;

/* n_steps RK2 steps keeping every state: Q^{j+1} = S1(S0(Q^j); Q^j) with the
 * stage Q^(1)_j stored.  qs holds 3(n_steps+1) device pointers (eta, hu, hv of
 * Q^0 .. Q^n; Q^0 is read), q1s 3 n_steps (the stages).  The recomputation a
 * checkpointed gradient makes of one segment; exact build.  With frs (NULL
 * when no side is 3), gpo_boundary_conditions follows each substep, as in
 * the step of a flow-relaxation simulator. */
int gpo_cdklm16_trajectory_f32(const gpo_cdklm16_params* p, int32_t n_steps, float dt, float* const* qs,
                               float* const* q1s, int32_t pitch, const float* Hi, int32_t Hi_pitch,
                               const float* f_cells, uint32_t boundary_conditions, const gpo_frs* frs,
                               void* stream);

/* Reverse sweep over such a trajectory: lam (3 pointers) holds lambda at
 * Q^n on entry and (dQ^n/dQ^0)^T lambda on return (interior cells; the ghost
 * cells of lam are not written); work holds 6 arrays of the state's shape
 * used as scratch. */
int gpo_cdklm16_adjoint_steps_f32(const gpo_cdklm16_params* p, int32_t n_steps, float dt, const float* const* qs,
                                  const float* const* q1s, float* const* lam, float* const* work, int32_t pitch,
                                  const float* Hi, int32_t Hi_pitch, const float* f_cells,
                                  uint32_t boundary_conditions, const gpo_frs* frs, void* workspace,
                                  int32_t* flag, void* stream);

/* gpo_cdklm16_adjoint_steps_f32 without the caller's scratch (`work` and
 * `workspace`): the six lambda arrays and the stage workspace live in a
 * 48 MiB array of library device memory, one per GPU, kept until the process
 * ends, so the call allocates nothing (no hipMalloc/hipFree, no allocating
 * library routine) and can be captured in a hipGraph.  Grids whose scratch,
 * 6 * round_up(pitch * (ny+4), 256) + gpo_adjoint_workspace_bytes(nx, ny, 4)
 * bytes, exceeds 48 MiB (C3's 512 x 512 needs about 40 MiB) are rejected.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race.  Results equal gpo_cdklm16_adjoint_steps_f32's
 * bit for bit.  No counterpart in the reference: it serves callers that cannot
 * allocate (code under graph capture, tight serving loops). */
int gpo_cdklm16_adjoint_steps_f32_noalloc(const gpo_cdklm16_params* p, int32_t n_steps, float dt,
                                          const float* const* qs, const float* const* q1s, float* const* lam,
                                          int32_t pitch, const float* Hi, int32_t Hi_pitch, const float* f_cells,
                                          uint32_t boundary_conditions, const gpo_frs* frs, int32_t* flag,
                                          void* stream);

/* Transpose (equally: tangent) of the flow-relaxation blend: lambda *=
 * (1 - alpha) on the sponge cells of the code-3 sides (the alpha of
 * boundary_kernels.cu:406-540), east/west pass first. */
int gpo_frs_adj_f32(const gpo_cdklm16_params* p, uint32_t boundary_conditions, const int32_t sponge[4],
                    float* lam_eta, float* lam_hu, float* lam_hv, int32_t pitch, void* stream);
int gpo_frs_adj_f64(const gpo_cdklm16_params* p, uint32_t boundary_conditions, const int32_t sponge[4],
                    double* lam_eta, double* lam_hu, double* lam_hv, int32_t pitch, void* stream);
/* The blend itself on an FP64 state (for the FP64 verification runs). */
int gpo_frs_apply_f64(const gpo_cdklm16_params* p, uint32_t boundary_conditions, const gpo_frs* frs, double* eta,
                      double* hu, double* hv, int32_t pitch, void* stream);

/* ---- batched differentiable ensemble: many members' checkpointed forward,
 * trajectory recomputation and reverse sweep, one launch per stage for all
 * members (no counterpart in the reference).  Members: RK2, wall, periodic
 * and flow-relaxation boundaries, land, no friction; exact build.  Every state is three planes
 * (eta, hu, hv) of (ny+4) rows x pitch bytes, `plane` bytes apart; a
 * buffer of several states holds them one after another (3 planes each). */
typedef struct {
  gpo_cdklm16_params params;
  float dt;
  uint32_t boundary_conditions;
  const float* Hi;
  int32_t Hi_pitch;
  const float* f_cells;            /* per-cell f (pitch = pitch) or NULL */
  int32_t pitch;
  int64_t plane;
  float* state;                    /* forward state, stepped in place */
  float* stage;                    /* forward RK stage */
  float* ckpt;                     /* n_segments states: segment starts */
  float* qs;                       /* k+1 states of the segment being swept */
  float* q1s;                      /* k RK stages of that segment */
  float* lam;                      /* 3 states: lambda, then scratch */
  void* workspace;                 /* gpo_adjoint_workspace_bytes(nx, ny, 4) */
  int32_t has_frs;                 /* some side is flow relaxation (3): frs is used */
  gpo_frs frs;
  /* stochastic member: after every step the SOAR model error of fixed normals
   * (draws draw0, draw0 + 1, ... of the member's Philox stream, interpolation
   * factor 1, scalar f), as gpo_random_normal + gpo_soar +
   * gpo_geostrophic_balance add it; all members or none */
  int32_t has_noise;
  uint64_t seed;
  uint32_t stream_id, draw0;
  float soar_weights[25];
  float g, f, beta;
  float* xi;                       /* xi_rows x xi_cols, dense */
  int32_t xi_rows, xi_cols;
  float* coarse;                   /* (ny+4) x (nx+4), dense */
  float* coarse_seg;               /* k such fields: the swept segment's, kept from its recomputation
                                    * (NULL: regenerated in the sweep); all members or none */
  float* pre;                      /* k planes (pitch and plane stride of the states): eta of the
                                    * second-substep outputs before the model error */
  /* drifters advected after every step (gpo_drift without wind); all members
   * or none */
  int32_t n_drifters;
  const float* Bm;
  int32_t Bm_pitch;
  float* pos;                      /* forward positions, n_drifters x 2 */
  float* pos_ckpt;                 /* n_segments x n_drifters x 2 */
  float* pos_seg;                  /* (k+1) x n_drifters x 2: positions within the swept segment */
  float* lam_pos;                  /* n_drifters x 2: sensitivity of the positions */
} gpo_adj_member;

typedef void* gpo_adj_batch;

/* Builds the argument tables of n members with segments of at most k steps
 * and n_segments checkpoint slots (ckpt holds n_segments states per member).
 * Allocates device tables; synchronises. */
int gpo_adj_batch_create(const gpo_adj_member* members, int32_t n, int32_t k, int32_t n_segments,
                         gpo_adj_batch* out, void* stream);
/* n_steps RK2 steps of every member's state (each with its own dt). */
int gpo_adj_batch_forward(gpo_adj_batch b, int32_t n_steps, void* stream);
/* restore = 0: state -> ckpt[segment]; restore = 1: ckpt[segment] -> qs[0]
 * (gpo_adj_batch_slot with slot = segment). */
int gpo_adj_batch_checkpoint(gpo_adj_batch b, int32_t segment, int32_t restore, void* stream);
/* Checkpoint slots for any schedule (binomial checkpointing): the state and
 * drifter positions of every member, with `segment` the segment whose start
 * they hold.  mode 0: state -> ckpt[slot] (the forward run is at the start of
 * segment); 1: ckpt[slot] -> qs[0], the segment start of the next trajectory;
 * 2: ckpt[slot] -> state, the forward run continues from segment; 3: state ->
 * qs[0].  No counterpart in the reference (it has no adjoint). */
int gpo_adj_batch_slot(gpo_adj_batch b, int32_t slot, int32_t mode, int32_t segment, void* stream);
/* qs[0] -> qs[1..m], with the stages in q1s[0..m-1] (m <= k). */
int gpo_adj_batch_trajectory(gpo_adj_batch b, int32_t m, void* stream);
/* lam: lambda at qs[m] in, (dqs[m]/dqs[0])^T lambda out (interior cells),
 * over the trajectory of the last gpo_adj_batch_trajectory(b, m). */
int gpo_adj_batch_sweep(gpo_adj_batch b, int32_t m, int32_t* flag, void* stream);
int gpo_adj_batch_destroy(gpo_adj_batch b);
/* How the sweep reverses each substep: fused = 1 (default) runs one strip
 * march per (member, strip of 58 columns, chunk of `rows` rows) that keeps
 * the stage's intermediates in registers, then folds the ghost-cell
 * sensitivities; fused = 0 runs the four staged kernels (F1, A1, A2, A3)
 * through workspace arrays.  Both give the same gradient bit for bit.
 * rows = 0 keeps the current chunk height.  No counterpart in the reference
 * (its adjoint does not exist); waits for the stream. */
int gpo_adj_batch_set_reverse(gpo_adj_batch b, int32_t fused, int32_t rows, void* stream);

/* Adjoint of the state dependence of gpo_geostrophic_balance for a given
 * coarse field: lambda_eta += -g/f d_y(coarse) lambda_hu + g/f d_x(coarse)
 * lambda_hv, in place (the noise is additive otherwise). */
int gpo_balance_adj_f32(const gpo_cdklm16_params* p, float g, float f, float beta, const float* f_cells,
                        int32_t f_pitch, const float* coarse, int32_t coarse_pitch, float* lam_eta,
                        float* lam_hu, float* lam_hv, int32_t pitch, const float* Hi, int32_t Hi_pitch,
                        void* stream);

/* The same in FP64 (the verification build).  Replaces nothing in the
 * reference (no derivative there); it transposes the balance of
 * gpu_kernels/ocean_noise.cu:256-390. */
int gpo_balance_adj_f64(const gpo_cdklm16_params* p, double g, double f, double beta, const float* f_cells,
                        int32_t f_pitch, const float* coarse, int32_t coarse_pitch, double* lam_eta,
                        double* lam_hu, double* lam_hv, int32_t pitch, const double* Hi, int32_t Hi_pitch,
                        void* stream);

/* The geostrophic balance of a given coarse field added to an FP64 state, the
 * verification counterpart of gpo_geostrophic_balance
 * (gpu_kernels/ocean_noise.cu:256-390, interpolation factor 1), and with
 * d_* (may be NULL) its tangent: d_hu -= g/f d_y(coarse) d_eta,
 * d_hv += g/f d_x(coarse) d_eta; cells with a land corner are set to 0. */
int gpo_balance_f64(const gpo_cdklm16_params* p, double g, double f, double beta, const float* f_cells,
                    int32_t f_pitch, const float* coarse, int32_t coarse_pitch, double* eta, double* hu, double* hv,
                    double* d_eta, double* d_hu, double* d_hv, int32_t pitch, const double* Hi, int32_t Hi_pitch,
                    void* stream);

/* Drifter advection in FP64 (the verification build of gpo_drift, which
 * replaces passiveDrifterKernel, gpu_kernels/driftKernels.cu:74-183; no wind):
 * n positions (x, y pairs, dense) move by dt * sensitivity * (u, v), (u, v) =
 * (hu, hv)/(Bm + eta) bilinear between the cell centres around each position,
 * periodic wrap on periodic axes; a drifter whose stencil leaves the arrays
 * stays put.  Periodic ghost cells are read as the interior cells they copy.
 * With d_positions (may be NULL) the tangent d_positions += dt sensitivity
 * (du/dx d_pos + du/dstate d_state) is formed too (d_eta, d_hu, d_hv: the
 * tangent of the state, or NULL for a fixed state). */
int gpo_drift_f64(int32_t nx, int32_t ny, double dx, double dy, double dt, const double* eta, const double* hu,
                  const double* hv, int32_t pitch, const double* Bm, int32_t Bm_pitch, uint32_t boundary_conditions,
                  int32_t n, double* positions, double* d_positions, const double* d_eta, const double* d_hu,
                  const double* d_hv, double sensitivity, void* stream);

/* Reverse of one drift step from `positions` (before the step): lam_positions
 * (n pairs) holds the sensitivity of the new positions on entry and of the old
 * ones on return; lam_eta/hu/hv accumulate the sensitivity of the state the
 * drift read (periodic ghost cells fold onto the interior cells they copy;
 * other ghost cells are stored constants and receive nothing), added cell by
 * cell in drifter order, so the result repeats bit for bit.  One workgroup
 * takes the drifters 256 at a time. */
int gpo_drift_adj_f32(int32_t nx, int32_t ny, float dx, float dy, float dt, const float* eta, const float* hu,
                      const float* hv, int32_t pitch, const float* Bm, int32_t Bm_pitch, uint32_t boundary_conditions,
                      int32_t n, const float* positions, float* lam_positions, float* lam_eta, float* lam_hu,
                      float* lam_hv, float sensitivity, void* stream);
int gpo_drift_adj_f64(int32_t nx, int32_t ny, double dx, double dy, double dt, const double* eta, const double* hu,
                      const double* hv, int32_t pitch, const double* Bm, int32_t Bm_pitch,
                      uint32_t boundary_conditions, int32_t n, const double* positions, double* lam_positions,
                      double* lam_eta, double* lam_hu, double* lam_hv, double sensitivity, void* stream);

/* Point-observation misfit J += sum_d (u - u_obs)^2 + (v - v_obs)^2 with
 * (u, v) = (hu, hv)/(eta + Hm) interpolated bilinearly between the four cell
 * centres around positions[2d], positions[2d+1] (metres); observed holds
 * (u_obs, v_obs) pairs.  *J (device double) is accumulated; lam_* (may be
 * NULL) accumulate dJ/d(eta, hu, hv), cell by cell in point order (J in a
 * fixed order too), so the results repeat bit for bit.  One workgroup takes
 * the points 256 at a time. */
int gpo_point_misfit_f32(int32_t nx, int32_t ny, float dx, float dy, const float* eta, const float* hu,
                         const float* hv, int32_t pitch, const float* Hi, int32_t Hi_pitch, int32_t n,
                         const float* positions, const float* observed, double* J, float* lam_eta, float* lam_hu,
                         float* lam_hv, void* stream);
int gpo_point_misfit_f64(int32_t nx, int32_t ny, double dx, double dy, const double* eta, const double* hu,
                         const double* hv, int32_t pitch, const double* Hi, int32_t Hi_pitch, int32_t n,
                         const float* positions, const float* observed, double* J, double* lam_eta, double* lam_hu,
                         double* lam_hv, void* stream);

/* Text of the last error on this thread. */
const char* gpo_last_error(void);

/* Library version string. */
const char* gpo_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GPUOCEAN_AMD_H */

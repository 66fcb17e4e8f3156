/* CPU oracle of the CDKLM16 path -- test infrastructure only.
 * Declarations for cdklm_oracle.cpp and noise_drift_oracle.cpp; loaded through
 * ctypes by oracle/oracle.py.  All arrays are dense row-major: cell arrays
 * (ny+4) x (nx+4), corner arrays (ny+5) x (nx+5). */
#ifndef GPO_ORACLE_H
#define GPO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t nx, ny;
  int32_t rk_order;
  int32_t bc_north, bc_east, bc_south, bc_west;
  int32_t sponge_north, sponge_east, sponge_south, sponge_west;
  int32_t one_dimensional;
  int32_t f_nx, f_ny;   /* shape of f_arr when f_arr != NULL */
  const float* f_arr;   /* coriolis f texture, or NULL for constant f */
  double dx, dy, g, f, r, theta;
  double desing_eps, flux_slope_eps, depth_cutoff, flux_balancer;
  double land_value, north_x, north_y, wind_stress_factor, rho_o;
} gpo_oracle_params;

typedef struct {
  const float *wx_cur, *wx_next, *wy_cur, *wy_next;
  int32_t wx_nx, wx_ny, wy_nx, wy_ny;
  float wind_t;
  const float *p_cur, *p_next;
  int32_t p_nx, p_ny;
  float p_t;
} gpo_oracle_forcing;

typedef struct {
  const float *ns_cur, *ns_next; /* [2][ns_nx][3] */
  const float *ew_cur, *ew_next; /* [ew_ny][2][3] */
  int32_t ns_nx, ew_ny;
  float t;
} gpo_oracle_frs;

void gpo_oracle_init_bathymetry_f32(const gpo_oracle_params* q, float* Hi, float* Bm);
void gpo_oracle_init_bathymetry_f64(const gpo_oracle_params* q, double* Hi, double* Bm);
void gpo_oracle_run_f32(const gpo_oracle_params* q, float* h0e, float* h0u, float* h0v, float* h1e, float* h1u,
                        float* h1v, const float* Hi, const float* Bm, double dt, int n_steps, int apply_initial_bc,
                        const gpo_oracle_forcing* forcing, const gpo_oracle_frs* frs);
void gpo_oracle_run_f64(const gpo_oracle_params* q, double* h0e, double* h0u, double* h0v, double* h1e,
                        double* h1u, double* h1v, const double* Hi, const double* Bm, double dt, int n_steps,
                        int apply_initial_bc, const gpo_oracle_forcing* forcing, const gpo_oracle_frs* frs);
void gpo_oracle_substep_f32(const gpo_oracle_params* q, const float* ie, const float* iu, const float* iv,
                            float* oe, float* ou, float* ov, const float* Hi, const float* Bm, float dt,
                            int rk_step);
void gpo_oracle_apply_bc_f32(const gpo_oracle_params* q, float* h, float* u, float* v, const gpo_oracle_frs* frs);
void gpo_oracle_apply_bc_f64(const gpo_oracle_params* q, double* h, double* u, double* v,
                             const gpo_oracle_frs* frs);
float gpo_oracle_max_dt_f32(const gpo_oracle_params* q, const float* eta, const float* hu, const float* hv,
                            const float* Bm);
double gpo_oracle_max_dt_f64(const gpo_oracle_params* q, const double* eta, const double* hu, const double* hv,
                             const double* Bm);
void gpo_oracle_fcell_f32(const gpo_oracle_params* q, float* out);
int gpo_oracle_num_threads(void);

void gpo_oracle_soar(int nx, int ny, const float* W, float scale, int per_ns, int per_ew, const float* xi,
                     int xi_cols, float* coarse, int additive);
void gpo_oracle_geostrophic_balance(int nx, int ny, float dx, float dy, float g, float f, float beta, float north_x,
                                    float north_y, const float* f_cells, const float* coarse, float* eta, float* hu,
                                    float* hv, const float* Hi, float land_value);
void gpo_oracle_bicubic_interpolation(int nx, int ny, float dx, float dy, int cnx, int cny, float cdx, float cdy,
                                      int oi, int oj, float g, float f, float beta, float north_x, float north_y,
                                      const float* f_cells, const float* coarse, float* eta, float* hu, float* hv,
                                      const float* Hi, float land_value);
void gpo_oracle_drift(int nx, int ny, float dx, float dy, float dt, int x0ref, int y0ref, const float* eta,
                      const float* hu, const float* hv, const float* Hm, int per_ns, int per_ew, int n_drifters,
                      float* pos, float sensitivity, float wind_x, float wind_y, float wind_drift_factor);

#ifdef __cplusplus
}
#endif
#endif

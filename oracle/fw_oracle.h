/*
 * fw_oracle.h -- CPU restatement of finitewave's numba kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against (tests/, __graft_entry__.smoke, bench.py cpu_baseline).
 * Nothing in the product package may import, link or call it.
 *
 * Layout follows the reference exactly: C-order arrays of the mesh shape,
 * weights as (*shape, K) float64, an int64 list of active flat indices,
 * one diffusion sweep and one ionic sweep per step (numba's structure).
 */
#ifndef FW_ORACLE_H
#define FW_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int orc_version(void);
int orc_num_threads(void);
void orc_set_num_threads(int n);

/* ---- stencil weights (the @njit compute_weights functions) ---- */
/* isotropic_stencil_2d.py:161-204 (raw, before the numpy scaling) */
void orc_weights_iso2d(double* w, const int8_t* m, const double* dxx,
                       const double* dyy, int64_t ni, int64_t nj);
/* asymmetric_stencil_2d.py:291-433 */
void orc_weights_aniso2d(double* w, const int8_t* m, const double* dxx,
                         const double* dxy, const double* dyx,
                         const double* dyy, int64_t ni, int64_t nj);
/* symmetric_stencil_2d.py:159-250 */
void orc_weights_sym2d(double* w, const int8_t* m, const double* dxx,
                       const double* dxy, const double* dyx, const double* dyy,
                       int64_t ni, int64_t nj);
/* isotropic_stencil_3d.py:112-149 */
void orc_weights_iso3d(double* w, const int8_t* m, const double* dxx,
                       const double* dyy, const double* dzz, int64_t ni,
                       int64_t nj, int64_t nk);
/* asymmetric_stencil_3d.py:163-458; d holds the 9 arrays in the order
 * xx, xy, xz, yx, yy, yz, zx, zy, zz */
void orc_weights_aniso3d(double* w, const int8_t* m, const double* const* d,
                         int64_t ni, int64_t nj, int64_t nk);

/* numpy epilogue of compute_weights: w = ((w*D)*dt)/dr2 ; w[..., c] += 1
 * (isotropic_stencil_2d.py:65-66, asymmetric_stencil_2d.py:81-82) */
void orc_scale_weights(double* w, int64_t n_nodes, int k, int centre,
                       double D, double dt, double dr2);
/* symmetric_stencil_2d.py epilogue: w *= (D*dt)/dr2 ; w[..., 4] += 1 */
void orc_scale_weights_sym(double* w, int64_t n_nodes, double D, double dt,
                           double dr2);

/* ---- diffusion sweeps ---- */
void orc_diffusion_iso2d(double* u_new, const double* u, const double* w,
                         const int64_t* idx, int64_t n, int64_t ni, int64_t nj);
void orc_diffusion_aniso2d(double* u_new, const double* u, const double* w,
                           const int64_t* idx, int64_t n, int64_t ni,
                           int64_t nj);
void orc_diffusion_iso3d(double* u_new, const double* u, const double* w,
                         const int64_t* idx, int64_t n, int64_t ni, int64_t nj,
                         int64_t nk);
void orc_diffusion_aniso3d(double* u_new, const double* u, const double* w,
                           const int64_t* idx, int64_t n, int64_t ni,
                           int64_t nj, int64_t nk);

/* ---- ionic sweeps ----
 * model ids: 0 AlievPanfilov, 1 Barkley, 2 MitchellSchaeffer, 3 FentonKarma,
 * 4 BuenoOrovio, 5 LuoRudy91, 6 TP06, 7 Courtemanche.
 * state: array of pointers, in the order the reference kernel receives them
 * (after u_new, u).  params: the scalars after dt, in the reference's
 * argument order.  Returns 0, or -1 for an unknown model. */
int orc_ionic(int model, double* u_new, const double* u, double* const* state,
              const int64_t* idx, int64_t n, double dt, const double* params);

/* one pass of ActivationTime2DTracker._track over all n cells
 * (activation_time_2d_tracker.py:61-63) */
void orc_act_time(double* act_t, const double* u, int64_t n, double thr,
                  double t);

#ifdef __cplusplus
}
#endif
#endif

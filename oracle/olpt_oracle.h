/*
 * olpt_oracle.h - CPU restatement of the reference IPR hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may call it.
 *
 * Plain C11, FP64 throughout, operations ordered as in the reference so that
 * results agree to the last bit wherever glibc is the common math library.
 * Entry points take the same arguments as the reference harness
 * (ref_harness.cpp) so that tests can run the two side by side.
 * Return value: 0 ok, 1 capacity exceeded, 2 invalid argument,
 * 3 unsupported configuration, 4 allocation failure.
 */
#ifndef OLPT_ORACLE_H
#define OLPT_ORACLE_H

#include <stdint.h>
#include "olpt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int orc_project(const olpt_camera *cam, const double *xyz, int n, double *uv);
int orc_line_of_sight(const olpt_camera *cam, const double *uv, int n,
                      double *out);
int orc_triangulate(const double *lines, int n_line, double *xyz, double *err);
int orc_polyfit(const double *x, const double *y, int n, int order,
                double *coeff);
int orc_trilinear(const double *lim, const double *v, const double *p, double *out);
int orc_is_local_max(const double *img, int H, int W, int row, int col);
int orc_otf_param(const olpt_otf *otf, int cam_id, const double *xyz, int n,
                  double *out);
int orc_find_peaks(const double *img, int H, int W, const olpt_ipr_params *p,
                   int cap, double *xy, int32_t *n_out);
int orc_stereo_match(const olpt_camera *cams, int n_cam, const int32_t *active,
                     const double *xy, const int32_t *n_pts,
                     const olpt_ipr_params *p, int cap_cand, int32_t *cand_ids,
                     double *cand_e, int32_t *sel_idx, int cap_obj,
                     double *obj_xyz, int32_t *obj_ids, int32_t *counts);
int orc_prune(int n_cam, const int32_t *active, const int32_t *n_pts, const int32_t *ids,
              const double *e, int M, uint8_t *sel);
int orc_shake(const olpt_camera *cams, int n_cam, const int32_t *active,
              const double *imgs, int H, int W, const olpt_ipr_params *p,
              const olpt_otf *otf, const double *xyz_in, int n_obj,
              double *xyz_out, uint8_t *flags, double *score, double *res_out);
int orc_residual(const olpt_camera *cams, int n_cam, const double *imgs, int H,
                 int W, const olpt_ipr_params *p, const olpt_otf *otf,
                 const double *xyz, int n_obj, double *res_out);
int orc_run_ipr(const olpt_camera *cams, int n_cam, const double *imgs, int H,
                int W, const olpt_ipr_params *p, const olpt_otf *otf, int cap,
                double *obj_xyz, double *obj_uv, int cap_loop,
                int32_t *loop_total, int32_t *counts);

/* tracking (STB) pieces */
int orc_stb_predict(const double *pos, const int32_t *len, int n_tr, int stride, double *out,
                    int32_t *ok);
int orc_stb_linear_fit(int n_init, const double *pos, const int32_t *len, int n_tr, int stride,
                       int32_t *ok);
int orc_pred_field(const double *lim, int nx, int ny, int nz, double r, const int32_t *nbin,
                   int smooth, const double *sigma, const double *prev, int n_prev,
                   const double *curr, int n_curr, double *field, const double *q, int n_q,
                   double *disp);

#ifdef __cplusplus
}
#endif

#endif

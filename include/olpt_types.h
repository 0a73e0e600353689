/*
 * olpt_types.h - plain-old-data records shared by the HIP library's C-ABI,
 * the CPU oracle and the reference harness.
 *
 * Every field mirrors a member of the reference's C++ objects so that a
 * caller holding reference objects can fill these records field by field:
 *   olpt_camera      <- Camera::_pinhole_param + _max_intensity
 *                      (inc/libMath/Camera.h:23-35, :85)
 *   olpt_otf         <- OTF::_param (inc/libSTB/OTF.h:16-31)
 *   olpt_ipr_params  <- IPRParam / SMParam / ShakeParam / TracerConfig
 *                      (inc/libSTB/Config.h:39-45, :72-80, :83-89, :192-205)
 */
#ifndef OLPT_TYPES_H
#define OLPT_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OLPT_MAX_CAM 8

/* Pinhole camera with OpenCV radial/tangential distortion.
 * Layout of R, Rinv: row-major 3x3.  T maps world -> camera (x_c = R X + T);
 * Tinv = -Rinv T is the camera centre in world coordinates.
 * is_distorted follows the reference loader rule: 1 iff some coefficient is
 * strictly greater than 1e-8 (Camera.cpp:61-64); n_dist is then 4, 5, 8 or
 * 12 (Camera.cpp:71-83). */
typedef struct olpt_camera {
  int32_t n_row;
  int32_t n_col;
  int32_t n_dist;
  int32_t is_distorted;
  double fx, fy, cx, cy;
  double dist[12];
  double R[9];
  double T[3];
  double Rinv[9];
  double Tinv[3];
  double max_intensity;
} olpt_camera;

/* Gaussian OTF grid: per camera, per grid node, (a, b, c, alpha).
 * Grid node (ix, iy, iz) has flat index ix*ny*nz + iy*nz + iz (OTF.h:41-45);
 * value arrays are [n_cam][nx*ny*nz] row-major.  boundary is
 * {xmin, xmax, ymin, ymax, zmin, zmax}. */
typedef struct olpt_otf {
  int32_t n_cam, nx, ny, nz;
  double boundary[6];
  const double *a;
  const double *b;
  const double *c;
  const double *alpha;
} olpt_otf;

/* Tracer IPR parameters (reference defaults in brackets). */
typedef struct olpt_ipr_params {
  /* IPRParam */
  int32_t n_cam_reduced;       /* [1] */
  int32_t n_loop_ipr;          /* [4] */
  int32_t n_loop_ipr_reduced;  /* [2] */
  int32_t n_obj2d_process_max; /* [200000] */
  /* SMParam */
  int32_t match_cam_count;     /* [4] */
  int32_t idmap_cell_px;       /* [4] */
  double tol_2d_px;            /* [1.0] */
  double tol_3d_mm;            /* [2.4e-2] */
  double limit[6];             /* view volume xmin,xmax,ymin,ymax,zmin,zmax */
  /* ShakeParam */
  double shake_width;          /* mm */
  double shakewidth_min;       /* mm, shake_width/20 when read from file */
  int32_t n_shake_loop;        /* [4] */
  int32_t ratio_augimg;        /* [2] */
  double thred_ghost;          /* [0.1] */
  /* TracerConfig */
  int32_t min_obj_int;
  int32_t reserved0;
  double radius_obj;           /* px */
} olpt_ipr_params;

/* Object flags as in Shake.h:19. */
enum { OLPT_FLAG_NONE = 0, OLPT_FLAG_GHOST = 1, OLPT_FLAG_REPEATED = 2 };

#ifdef __cplusplus
}
#endif

#endif /* OLPT_TYPES_H */

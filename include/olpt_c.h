/*
 * olpt_c.h - C-ABI of the MI355X IPR library (libolpt_hip.so).
 *
 * Plain pointers and sizes, no C++ or torch types.  Pointers named d_* are
 * device (HBM) pointers; everything else is host memory.  Entry points never
 * throw: they return an olpt_status and keep a thread-local message for
 * olpt_last_error().  Status codes 1-5 mirror the reference's ErrorCode
 * (inc/error.hpp:41-48); codes >= 100 are GPU-side failures.
 *
 * A context owns one HIP stream (not re-entrant; use one context per
 * thread/stream) and a library-owned device workspace sized by
 * olpt_ctx_reserve().  Arithmetic precision of projection / residual /
 * shake is OLPT_PREC_F32 (default) or OLPT_PREC_F64 (parity mode); line of
 * sight, triangulation and the stereo tolerance tests are always FP64.
 */
#ifndef OLPT_C_H
#define OLPT_C_H

#include <stdint.h>
#include "olpt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct olpt_ctx olpt_ctx;
typedef int32_t olpt_status;

enum {
  OLPT_OK = 0,
  OLPT_E_INVALID_ARGUMENT = 1, /* ErrorCode::InvalidArgument */
  OLPT_E_OUT_OF_RANGE = 2,     /* ErrorCode::OutOfRange */
  OLPT_E_UNSUPPORTED = 3,      /* ErrorCode::UnsupportedType */
  OLPT_E_IO = 4,               /* ErrorCode::IOfailure */
  OLPT_E_NO_ENOUGH_DATA = 5,   /* ErrorCode::NoEnoughData */
  OLPT_E_GPU = 100,            /* HIP runtime error */
  OLPT_E_ALLOC = 101,          /* device allocation failed */
  OLPT_E_CAPACITY = 102,       /* a workspace capacity was exceeded */
  /* pruneMatch (StereoMatch.cpp:631-980) on this input depends on what the
   * library does not replay: eq_eps-equal values that are not transitive
   * (the reference's order is then whatever std::stable_sort produces) or
   * a firing replacement branch.  Outputs are not valid. */
  OLPT_E_PRUNE_TIES = 103
};

enum { OLPT_U8 = 0, OLPT_U16 = 1, OLPT_F32 = 2, OLPT_F64 = 3 };
enum { OLPT_PREC_F32 = 0, OLPT_PREC_F64 = 1 };

/* ---- context -------------------------------------------------------- */
/* stream: a hipStream_t, or NULL for the device's null stream. */
olpt_status olpt_ctx_create(int device, void *stream, olpt_ctx **out);
olpt_status olpt_ctx_destroy(olpt_ctx *ctx);
/* Size the workspace: cameras, image size, per-camera 2D capacity, stereo
 * candidate capacity, per-loop object capacity.  0 selects a default. */
olpt_status olpt_ctx_reserve(olpt_ctx *ctx, int n_cam, int n_row, int n_col,
                             int max_2d_per_cam, int max_candidates,
                             int max_objects);
olpt_status olpt_set_precision(olpt_ctx *ctx, int precision);
/* No counterpart in the reference (its IPR runs one frame per call on host
 * threads).  olpt_ipr_run spreads the frames of one call over n_streams
 * workspaces, each with its own HIP stream, so that one frame's small
 * kernels overlap another frame's work; 1 (default) keeps every frame on
 * the context stream.  Each extra stream holds a full workspace, as sized
 * by olpt_ctx_reserve.  n_streams: 1..8. */
olpt_status olpt_set_concurrency(olpt_ctx *ctx, int n_streams);
/* No counterpart in the reference.  Launch shapes that change how the work
 * is scheduled, never what it computes: every kernel they size walks its
 * work with a grid-stride loop, and every result whose order depends on
 * scheduling is ranked afterwards.  shake_blocks: one-wave workgroups of
 * the shake and proposal kernels (default 4096); build_blocks: 256-thread
 * workgroups of the stereo build kernels (default 1024); prune_blocks: cap
 * on the 256-thread workgroups of the sort/prune/triangulate kernels
 * (default 512), each per frame (a launch over few frames raises them so
 * that it has at least 1024 workgroups, 4096 one-wave ones, in all);
 * n_streams: as olpt_set_concurrency;
 * frames_per_batch (default 1): the frames every kernel launch of
 * olpt_ipr_run covers - its grid takes the frame from blockIdx.z and each
 * workspace holds that many frame slices (memory grows with it; changing it
 * reallocates the workspaces), so one launch sequence serves a whole batch.
 * Each field must lie in its range (blocks 64..65536, streams 1..8, frames
 * 1..1024); a captured graph is dropped.  Python's Context.autotune times candidates on the caller's
 * frames and keeps the fastest. */
typedef struct {
  int32_t n_streams;
  int32_t shake_blocks;
  int32_t build_blocks;
  int32_t prune_blocks;
  int32_t frames_per_batch;
} olpt_tuning;
olpt_status olpt_set_tuning(olpt_ctx *ctx, const olpt_tuning *tuning);
olpt_status olpt_get_tuning(olpt_ctx *ctx, olpt_tuning *tuning);
/* Replaces Camera objects (Camera.h:78-174, pinhole model). */
olpt_status olpt_set_cameras(olpt_ctx *ctx, const olpt_camera *cams, int n_cam);
/* Replaces TracerConfig::_otf (OTF.h:33-78); arrays are copied. */
olpt_status olpt_set_otf(olpt_ctx *ctx, const olpt_otf *otf);
/* OTF::getOTFParam (OTF.cpp:273-338) of the context's OTF, FP64: for camera
 * cam at n world points d_xyz [n][3], d_out [n][5] = (a, b, c, cos alpha,
 * sin alpha) (device buffers; returns when done). */
olpt_status olpt_otf_param(olpt_ctx *ctx, int cam, const double *d_xyz, int n, double *d_out);
olpt_status olpt_sync(olpt_ctx *ctx);
void *olpt_stream(olpt_ctx *ctx);
const char *olpt_last_error(void);
/* Device-side diagnostic counters of the last stereo/shake run:
 * [0] candidate overflow, [1] build list overflow, [2] prune keys with
 * non-transitive eq_eps ties, [3] prune replacement cases, [4] object
 * overflow, [5] prune rounds exhausted, [6] checkMatch triangulation
 * failures, [7] peak overflow. */
olpt_status olpt_diagnostics(olpt_ctx *ctx, int32_t out[8]);
/* Work-list sizes of the last stereo build (frame 0 of the first
 * workspace, as the last stereo call or IPR loop left them): [0] level-1
 * (observation, point) pairs, [1..8] survivors of each build level,
 * [9..16] items enumerated at each level, [17] candidates, [18] selected
 * candidates, [19] objects.  No counterpart in the reference (its DFS keeps
 * no such lists); for profiling the build. */
olpt_status olpt_build_counts(olpt_ctx *ctx, int32_t out[24]);

/* Stage timing with HIP events on the context stream (eager runs only).
 * Slots: 0 peaks, 1 stereo build, 2 sort+prune+triangulate, 3 shake
 * kernel, 4 residual gather, 5 the level-2 band enumeration kernel of the
 * build (k_b_enum, third camera; also inside slot 1).  olpt_stage_times
 * returns the summed milliseconds and launch counts since the last read and
 * resets them.
 * Slot 6 (always recorded): host milliseconds spent issuing the work of
 * the eager olpt_ipr_run calls since the last read, and their number. */
enum { OLPT_STAGE_PEAKS = 0, OLPT_STAGE_BUILD = 1, OLPT_STAGE_PRUNE = 2,
       OLPT_STAGE_SHAKE = 3, OLPT_STAGE_RESIDUAL = 4, OLPT_STAGE_ENUM2 = 5,
       OLPT_STAGE_HOST = 6 };
olpt_status olpt_stage_timing(olpt_ctx *ctx, int enable);
olpt_status olpt_stage_times(olpt_ctx *ctx, double ms[8], int32_t calls[8]);
/* Algorithmic-byte accounting (synchronises once per IPR loop, so only for
 * instrumented runs): returns the bytes and launch counts accumulated since
 * the previous call, resets them, and switches accounting on or off. */
olpt_status olpt_stage_stats(olpt_ctx *ctx, int enable, double bytes[8],
                             int32_t launches[8]);

/* ---- stages ---------------------------------------------------------- */
/* ObjectFinder2D::findObject2D, tracer branch (ObjectFinder.cpp:37-111).
 * One image (dtype OLPT_U8..F64, row-major n_row x n_col).  Writes (x, y)
 * pairs in row-major scan order to d_xy (capacity cap) and the count to
 * *n_out.  Returns OLPT_E_CAPACITY if more than cap peaks exist. */
olpt_status olpt_find_peaks(olpt_ctx *ctx, const void *d_img, int dtype,
                            int n_row, int n_col,
                            const olpt_ipr_params *params, double *d_xy,
                            int cap, int32_t *n_out);

/* No counterpart in the reference: olpt_find_peaks (ObjectFinder2D::
 * findObject2D, tracer branch) for callers that cannot allocate device
 * memory - code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  It takes no context and allocates nothing: its
 * work memory (chunk masks, counts, offsets, scan partials, flags) is a
 * fixed array of the library, one per GPU (about 4.2 MB, held until the
 * process ends), so images are limited to 4096 x 4096 (larger: status
 * OLPT_E_CAPACITY).  Calls on one GPU share that array and must not
 * overlap: issue them on one stream, or order them.  stream: a hipStream_t
 * of the current device (NULL = its null stream).  The work is enqueued and
 * the call returns: *d_n_out (device) receives the total number of peaks,
 * of which the first min(total, cap) are written to d_xy as (x, y) pairs in
 * row-major scan order - the values olpt_find_peaks returns. */
olpt_status olpt_find_peaks_static(void *stream, const void *d_img, int dtype,
                                   int n_row, int n_col,
                                   const olpt_ipr_params *params,
                                   double *d_xy, int cap, int32_t *d_n_out);

/* StereoMatch(cams, obj2d_list, cfg).match() (StereoMatch.cpp:65-178).
 * d_xy: per camera, cap_2d (x, y) pairs ([n_cam][cap_2d][2]); n_pts: host
 * counts; active: host flags (NULL = all active).  Outputs (device):
 * the checked candidates in id-sorted order (d_cand_ids [n_cand][n_cam],
 * d_cand_e), the selection flags d_sel [n_cand] (1 = kept by pruneMatch),
 * the triangulated objects (d_obj_xyz [n_obj][3], d_obj_ids
 * [n_obj][n_cam]).  counts (host) = {n_cand, n_sel, n_obj}.  Any of the
 * candidate outputs may be NULL. */
olpt_status olpt_stereo_match(olpt_ctx *ctx, const int32_t *active,
                              const double *d_xy, int cap_2d,
                              const int32_t *n_pts,
                              const olpt_ipr_params *params,
                              int32_t *d_cand_ids, double *d_cand_e,
                              uint8_t *d_sel, int cap_cand, double *d_obj_xyz,
                              int32_t *d_obj_ids, int cap_obj,
                              int32_t *counts);

/* StereoMatch::pruneMatch (StereoMatch.cpp:631-980) on given candidates
 * (match() calls it on its id-sorted candidates, :153-156): d_cand_ids
 * [n_cand][n_cam] (-1 = camera unused), d_cand_e [n_cand] their e_check,
 * n_pts (host) the 2D point count per camera, active (host, NULL = all
 * active).  d_sel [n_cand] = 1 for the selected candidates (the reference
 * returns them in this index order, :973-978).  OLPT_E_PRUNE_TIES when the
 * selection would depend on non-transitive eq_eps ties or the replacement
 * branch. */
olpt_status olpt_prune_match(olpt_ctx *ctx, const int32_t *active, const int32_t *n_pts,
                             const olpt_ipr_params *params, const int32_t *d_cand_ids,
                             const double *d_cand_e, int n_cand, uint8_t *d_sel);

/* Shake(cams, cfg).runShake(objs, imgs) (Shake.cpp:36-163) on n_obj
 * objects, then calResidualImage of the unflagged objects (IPR mode,
 * Shake.h:40-48).  d_imgs: [n_cam][n_row][n_col] of dtype.  Outputs:
 * shaken positions, flags (Shake.h:19 bits), scores, and (optional)
 * residual images d_res [n_cam][n_row][n_col] in the context precision. */
olpt_status olpt_shake(olpt_ctx *ctx, const int32_t *active,
                       const void *d_imgs, int dtype,
                       const olpt_ipr_params *params, const double *d_xyz_in,
                       int n_obj, double *d_xyz_out, uint8_t *d_flags,
                       double *d_score, void *d_res);

/* IPR-mode residual of given objects (Shake::calResidualImage, all cameras,
 * clamped at 0) into d_res (context precision). */
olpt_status olpt_residual(olpt_ctx *ctx, const void *d_imgs, int dtype,
                          const olpt_ipr_params *params, const double *d_xyz,
                          int n_obj, void *d_res);

/* IPR::runIPR (IPR.cpp:195-306) on one frame, or on n_frames frames stored
 * back to back (d_imgs [n_frames][n_cam][n_row][n_col]).  Per frame f the
 * objects of all loops are written in loop order to d_obj_xyz
 * [f][cap_obj_frame][3] and their final projections to d_obj_uv
 * [f][cap_obj_frame][n_cam][2] (may be NULL); d_n_obj[f] holds the count.
 * loop_total (host, optional, [n_frames][cap_loop]) receives the
 * cumulative count after each IPR loop (0 past the last loop). */
olpt_status olpt_ipr_run(olpt_ctx *ctx, const void *d_imgs, int dtype,
                         int n_frames, const olpt_ipr_params *params,
                         double *d_obj_xyz, double *d_obj_uv,
                         int cap_obj_frame, int32_t *d_n_obj,
                         int32_t *loop_total, int cap_loop);

/* Capture olpt_ipr_run's kernel sequence for these exact arguments into a
 * hipGraph owned by the context, and replay it.  Captured pointers stay
 * fixed; replays overwrite the same outputs. */
olpt_status olpt_ipr_graph_capture(olpt_ctx *ctx, const void *d_imgs,
                                   int dtype, int n_frames,
                                   const olpt_ipr_params *params,
                                   double *d_obj_xyz, double *d_obj_uv,
                                   int cap_obj_frame, int32_t *d_n_obj);
olpt_status olpt_ipr_graph_launch(olpt_ctx *ctx);

/* ---- learned shake proposal ------------------------------------------
 * No counterpart in the reference.  A BF16 MLP on MFMA (FP32 accumulate):
 * 4 cameras x 11x11 augmented residual patches (484 inputs, zero-padded to
 * 512) -> 256 -> 256 -> 256 (GELU) -> (dx, dy, dz in voxels, ghost logit).
 * Parameters are one FP32 array of olpt_mlp_num_params() values: W1
 * [512][256], b1 [256], W2 [256][256], b2, W3 [256][256], b3, W4 [256][32],
 * b4 [32]; rows 484..511 of W1 and columns 4..31 of W4 / b4 are padding and
 * are kept at zero.  A handle owns its stream use and training buffers for
 * batches of up to max_batch rows. */
typedef struct olpt_mlp olpt_mlp;
int32_t olpt_mlp_num_params(void);
olpt_status olpt_mlp_create(int device, void *stream, int max_batch, olpt_mlp **out);
olpt_status olpt_mlp_destroy(olpt_mlp *mlp);
olpt_status olpt_mlp_set_params(olpt_mlp *mlp, const float *params, int32_t n);
olpt_status olpt_mlp_get_params(olpt_mlp *mlp, float *params, int32_t n);
/* d_y [n][4] = MLP(d_x [n][512] bf16), device buffers, on the handle's
 * stream. */
olpt_status olpt_mlp_infer(olpt_mlp *mlp, const void *d_x, int32_t n, float *d_y);
/* One Adam step (beta 0.9 / 0.999) on d_x [n][512] bf16 with targets d_t
 * [n][4] = (dx, dy, dz in voxels, ghost label 0/1): loss = mean over rows
 * of (1 - g) |y_xyz - t_xyz|^2 + ghost_weight * BCE(y_ghost, g).
 * *loss_out (host, may be NULL) receives the batch loss. */
olpt_status olpt_mlp_train_step(olpt_mlp *mlp, const void *d_x, const float *d_t,
                                int32_t n, float lr, float ghost_weight,
                                float *loss_out);
/* Use mlp (NULL: plain shake) in the shake of every IPR loop and olpt_shake
 * call of this context: before the plain iterations each object takes the
 * network's step (clamped to +-max_step_vox voxels per axis, voxel_mm per
 * voxel) when it lowers the shake residue (calShakeResidue, Shake.cpp:
 * 968-1045) on the object's ROI; the plain iterations then run as
 * params->n_shake_loop / shake_width say.  4 cameras, ROI up to 8x8.  The
 * network is not owned; a previously captured graph keeps the old setting. */
olpt_status olpt_set_shake_model(olpt_ctx *ctx, const olpt_mlp *mlp,
                                 double voxel_mm, double max_step_vox);
/* No counterpart in the reference: with the learned proposal in use, an
 * object whose ghost logit (the network's fourth output, at the matched
 * position) exceeds logit_min is flagged a ghost after the shake, beside the
 * score threshold (Shake.cpp:118-136): it leaves the IPR output and the
 * loop's residual image.  INFINITY (the default) turns this off. */
olpt_status olpt_set_ghost_logit(olpt_ctx *ctx, double logit_min);
/* Network inputs of n_obj objects at d_xyz on images d_imgs: residual of
 * the active cameras from all n_obj objects (as at the start of a shake
 * iteration), then per camera the 11x11 window of max(0, residual + own
 * projection) around the rounded projection, divided by the OTF amplitude:
 * d_x [n_obj][512] bf16 (4 x 121 values, then zeros). */
olpt_status olpt_shake_patches(olpt_ctx *ctx, const int32_t *active,
                               const void *d_imgs, int dtype,
                               const olpt_ipr_params *params,
                               const double *d_xyz, int n_obj, void *d_x);
/* Render n_obj particles: max-fusion of their OTF projections over the box
 * of half-size r_px (calROIBound), amplitude d_amp[i] in place of the OTF
 * amplitude when d_amp is not NULL, into d_out [n_cam][n_row][n_col] in the
 * context precision (the training images of the learned proposal). */
olpt_status olpt_render(olpt_ctx *ctx, const double *d_xyz, const double *d_amp,
                        int n_obj, double r_px, void *d_out);

/* Gradient of the last train step's loss (parameter layout); a step with
 * lr = 0 leaves the parameters unchanged. */
olpt_status olpt_mlp_get_grads(olpt_mlp *mlp, float *grads, int32_t n);
const char *olpt_mlp_last_error(void);

/* ---- shake-the-box tracking (STB) ----------------------------------- */
/* STBParam + PFParam + the frame range of BasicSetting (Config.h:14-37,
 * :47-60, :98-101).  Lengths in mm. */
typedef struct olpt_stb_params {
  int32_t frame_start, frame_end;
  int32_t n_initial_frames;        /* [4], at least 4 (Config.cpp:197-199) */
  double radius_search_obj_mm;     /* initial-phase linking radius */
  double radius_search_track_mm;   /* convergence-phase neighbour radius */
  double limit[6];                 /* view volume, as olpt_ipr_params.limit */
  int32_t pf_nx, pf_ny, pf_nz;     /* predictive-field grid */
  double pf_r_mm;                  /* predictive-field search half-width */
  int32_t pf_nbin[3];              /* [11 11 11] displacement bins */
  int32_t pf_smooth;               /* [1] Gaussian smoothing of the field */
  double pf_sigma[3];              /* [1 1 1] grid nodes */
} olpt_stb_params;

typedef struct olpt_stb olpt_stb;
enum { OLPT_TRACK_SHORT_ACTIVE = 0, OLPT_TRACK_LONG_ACTIVE = 1,
       OLPT_TRACK_LONG_INACTIVE = 2, OLPT_TRACK_EXIT = 3 };

/* STB(setting, "Tracer", cfg) (STB.cpp:6-38) on a context whose cameras
 * and OTF are set.  ipr: the IPR / stereo / shake parameters of the object
 * configuration.  output_path (may be NULL or "") receives the files the
 * reference writes: PredField_<f>_<f+1>.csv, InitialTrack/ and
 * ConvergeTrack/ track files (STB.cpp:9-11, :124-157, :66-89). */
olpt_status olpt_stb_create(olpt_ctx *ctx, const olpt_stb_params *stb,
                            const olpt_ipr_params *ipr, const char *output_path,
                            olpt_stb **out);
olpt_status olpt_stb_destroy(olpt_stb *stb);
/* No counterpart in the reference: every IPR call of the tracker (the
 * initial phase's IPR on the frames, STB.cpp:97, and the convergence
 * phase's IPR on the residual images, :454-456) runs with the learned shake
 * proposal mlp (as olpt_set_shake_model) and the IPR parameters ipr (for
 * instance one plain iteration from a small shake width); the shake of the
 * predicted particles (:406-418) keeps the configuration's plain schedule.
 * mlp NULL restores the plain IPR.  The network is not owned. */
olpt_status olpt_stb_set_ipr_model(olpt_stb *stb, const olpt_mlp *mlp, double voxel_mm,
                                   double max_step_vox, const olpt_ipr_params *ipr);
/* STB::processFrame (STB.cpp:40-89): d_imgs [n_cam][n_row][n_col]. */
olpt_status olpt_stb_process_frame(olpt_stb *stb, int frame_id, const void *d_imgs,
                                   int dtype);
/* Size of one track list (OLPT_TRACK_*): tracks and total points. */
olpt_status olpt_stb_track_count(olpt_stb *stb, int which, int32_t *n_tracks,
                                 int32_t *n_points);
/* The list's points, track by track in list order: track index, frame id,
 * world position (host buffers of cap rows). */
olpt_status olpt_stb_get_tracks(olpt_stb *stb, int which, int32_t cap, int32_t *track,
                                int32_t *frame, double *xyz);
/* Replace one list by n_rows points grouped by track index, as
 * STB::loadTracks builds them (STB.cpp:1006-1050). */
olpt_status olpt_stb_set_tracks(olpt_stb *stb, int which, int32_t n_rows,
                                const int32_t *track, const int32_t *frame,
                                const double *xyz);
/* STB::saveTracksAll / loadTracksAll (STB.cpp:926-1005, :1052-1076). */
olpt_status olpt_stb_save_tracks(olpt_stb *stb, const char *folder, int frame);
olpt_status olpt_stb_load_tracks(olpt_stb *stb, const char *folder, int frame);

/* STB::predictNext (STB.cpp:233-316) of n_tr tracks given by their last
 * len_i = min(length, 6) points: d_pos [n_tr][6][3] (oldest first), d_len
 * [n_tr]; d_out [n_tr][3]; d_ok[i] = 0 where the track has fewer than 3
 * points (the reference returns null). */
olpt_status olpt_stb_predict(olpt_ctx *ctx, const double *d_pos, const int32_t *d_len,
                             int n_tr, double *d_out, int32_t *d_ok);
/* STB::checkLinearFit (STB.cpp:874-924): d_pos [n_tr][4][3] holds the last
 * min(length, 4) points, oldest first, d_len the track lengths. */
olpt_status olpt_stb_linear_fit(olpt_ctx *ctx, int n_initial_frames, const double *d_pos,
                                const int32_t *d_len, int n_tr, int32_t *d_ok);
/* PredField::calPredField (PredField.cpp:21-86) between two object sets,
 * then PredField::getDisp (:91-165) at n_q query points.  d_field
 * [nx*ny*nz][3] (may be NULL); d_disp [n_q][3]. */
olpt_status olpt_pred_field(olpt_ctx *ctx, const olpt_stb_params *stb,
                            const double *d_prev, int n_prev, const double *d_curr,
                            int n_curr, double *d_field, const double *d_q, int n_q,
                            double *d_disp);
/* OTF::estimateUniformOTFFromImage (OTF.cpp:16-144) for every camera from
 * n_img frames d_frames [n_img][n_cam][n_row][n_col]: peaks of each image
 * (at most 1000, shuffled as the reference does), a = min(mean + 2 std,
 * max) of the peak pixels, (b, c) by least squares of -ln(I/a) over the
 * peaks' (2r+1)^2 windows, alpha = 0.  abca [n_cam][4] (host); ok[k] = 0
 * where an image of camera k had no peak (its row is left untouched). */
olpt_status olpt_estimate_otf(olpt_ctx *ctx, const void *d_frames, int dtype, int n_img,
                              const olpt_ipr_params *params, double *abca, int32_t *ok);

#ifdef __cplusplus
}
#endif

#endif /* OLPT_C_H */

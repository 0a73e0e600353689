// ref_harness.cpp - extern "C" driver around the reference's own C++ API.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
// reference's sources, where they lie under the reference checkout, into
// oracle/_ref/libolpt_ref.so by oracle/Makefile.  It is used to write golden
// vectors and to time the reference CPU path; nothing in the product links
// it.  It contains no reference code: it fills the reference's public objects
// (Camera, TracerConfig, OTF, Image, Object3D) from the POD records in
// include/olpt_types.h and calls
//   ObjectFinder2D::findObject2D        (ObjectFinder.cpp:6-21)
//   StereoMatch stages via the reference's FRIEND_DEBUG hook
//                                        (StereoMatch.h:120, STBCommons.h:57-61)
//   StereoMatch::match                  (StereoMatch.cpp:65-178)
//   Shake::runShake / calResidualImage  (Shake.cpp:36-163, Shake.h:40-48)
//   IPR::runIPR                         (IPR.cpp:195-306)
//   Camera::project / lineOfSight       (Camera.cpp:648, :996)
//   OTF::getOTFParam                    (OTF.cpp:273-338)
#include <algorithm>
#include <cstring>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <cmath>

#include "Config.h"
#include "IPR.h"
#include "ObjectFinder.h"
#include "OTF.h"
#include "Shake.h"
#include "StereoMatch.h"

#include "olpt_types.h"
#include "ref_sink.hpp"

// ---- reference debug hooks (declared friends when OPENLPT_EXPOSE_PRIVATE) ----
struct DebugAccess_StereoMatch {
  static void build(const StereoMatch &sm, std::vector<int> &seed,
                    std::vector<std::vector<int>> &out) {
    sm.buildMatch(seed, out);
  }
  static bool check(const StereoMatch &sm, const std::vector<int> &ids,
                    double &e) {
    return sm.checkMatch(ids, e);
  }
  static std::vector<std::vector<int>>
  prune(const StereoMatch &sm, const std::vector<std::vector<int>> &c,
        const std::vector<double> &e) {
    return sm.pruneMatch(c, e);
  }
  static std::vector<std::unique_ptr<Object3D>>
  triangulate(const StereoMatch &sm, const std::vector<std::vector<int>> &s) {
    return sm.triangulateMatch(s);
  }
};

struct DebugAccess_Shake {
  static std::vector<double> &score(Shake &s) { return s._score_list; }
};

namespace {


Camera makeCamera(const olpt_camera &c, bool active) {
  Camera cam;
  cam._type = PINHOLE;
  PinholeParam &p = cam._pinhole_param;
  p.n_row = c.n_row;
  p.n_col = c.n_col;
  p.cam_mtx = Matrix<double>(3, 3, 0.0);
  p.cam_mtx(0, 0) = c.fx;
  p.cam_mtx(0, 2) = c.cx;
  p.cam_mtx(1, 1) = c.fy;
  p.cam_mtx(1, 2) = c.cy;
  p.cam_mtx(2, 2) = 1.0;
  p.is_distorted = c.is_distorted != 0;
  p.n_dist_coeff = c.is_distorted ? c.n_dist : 0;
  p.dist_coeff.assign(c.dist, c.dist + (c.n_dist > 0 ? c.n_dist : 0));
  p.r_mtx = Matrix<double>(3, 3, 0.0);
  p.r_mtx_inv = Matrix<double>(3, 3, 0.0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      p.r_mtx(i, j) = c.R[3 * i + j];
      p.r_mtx_inv(i, j) = c.Rinv[3 * i + j];
    }
  p.t_vec = Pt3D(c.T[0], c.T[1], c.T[2]);
  p.t_vec_inv = Pt3D(c.Tinv[0], c.Tinv[1], c.Tinv[2]);
  cam._max_intensity = c.max_intensity;
  cam._is_active = active;
  return cam;
}

std::vector<Camera> makeCameras(const olpt_camera *c, int n,
                                const int32_t *active) {
  std::vector<Camera> cams;
  cams.reserve(n);
  for (int i = 0; i < n; ++i)
    cams.push_back(makeCamera(c[i], active ? active[i] != 0 : true));
  return cams;
}

void fillConfig(TracerConfig &cfg, const olpt_ipr_params &p,
                const olpt_otf *otf, const std::vector<Camera> &cams) {
  cfg._ipr_param.n_cam_reduced = p.n_cam_reduced;
  cfg._ipr_param.n_loop_ipr = p.n_loop_ipr;
  cfg._ipr_param.n_loop_ipr_reduced = p.n_loop_ipr_reduced;
  cfg._ipr_param.n_obj2d_process_max = p.n_obj2d_process_max;
  cfg._sm_param.match_cam_count = p.match_cam_count;
  cfg._sm_param.idmap_cell_px = p.idmap_cell_px;
  cfg._sm_param.tol_2d_px = p.tol_2d_px;
  cfg._sm_param.tol_3d_mm = p.tol_3d_mm;
  cfg._sm_param.limit = AxisLimit(p.limit[0], p.limit[1], p.limit[2],
                                  p.limit[3], p.limit[4], p.limit[5]);
  cfg._shake_param._shake_width = p.shake_width;
  cfg._shake_param._shakewidth_min = p.shakewidth_min;
  cfg._shake_param._n_shake_loop = p.n_shake_loop;
  cfg._shake_param._ratio_augimg = p.ratio_augimg;
  cfg._shake_param._thred_ghost = p.thred_ghost;
  cfg._min_obj_int = p.min_obj_int;
  cfg._radius_obj = p.radius_obj;
  if (otf) {
    AxisLimit lim(otf->boundary[0], otf->boundary[1], otf->boundary[2],
                  otf->boundary[3], otf->boundary[4], otf->boundary[5]);
    cfg._otf.loadParam(otf->n_cam, otf->nx, otf->ny, otf->nz, lim, cams);
    const int ng = otf->nx * otf->ny * otf->nz;
    for (int c = 0; c < otf->n_cam; ++c)
      for (int g = 0; g < ng; ++g) {
        cfg._otf._param.a(c, g) = otf->a[c * ng + g];
        cfg._otf._param.b(c, g) = otf->b[c * ng + g];
        cfg._otf._param.c(c, g) = otf->c[c * ng + g];
        cfg._otf._param.alpha(c, g) = otf->alpha[c * ng + g];
      }
  }
}

Image makeImage(const double *px, int H, int W) {
  Image img(H, W, 0.0);
  img.setData(px, H * W);
  return img;
}

std::vector<Image> makeImages(const double *px, int n_cam, int H, int W) {
  std::vector<Image> imgs;
  imgs.reserve(n_cam);
  for (int c = 0; c < n_cam; ++c)
    imgs.push_back(makeImage(px + (size_t)c * H * W, H, W));
  return imgs;
}

std::vector<std::vector<std::unique_ptr<Object2D>>>
makeObj2d(const double *xy, const int32_t *n_pts, int n_cam) {
  std::vector<std::vector<std::unique_ptr<Object2D>>> out(n_cam);
  size_t off = 0;
  for (int c = 0; c < n_cam; ++c) {
    for (int i = 0; i < n_pts[c]; ++i) {
      auto t = std::make_unique<Tracer2D>();
      t->_pt_center[0] = xy[2 * (off + i)];
      t->_pt_center[1] = xy[2 * (off + i) + 1];
      out[c].push_back(std::move(t));
    }
    off += n_pts[c];
  }
  return out;
}

std::vector<std::unique_ptr<Object3D>>
makeObjects(const TracerConfig &cfg, const double *xyz, int n) {
  std::vector<std::unique_ptr<Object3D>> objs;
  objs.reserve(n);
  for (int i = 0; i < n; ++i) {
    CreateArgs a;
    a._pt_center = Pt3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    objs.push_back(cfg.creatObject3D(std::move(a)));
  }
  return objs;
}

} // namespace

extern "C" {

// Camera::project for n points.
int ref_project(const olpt_camera *cam, const double *xyz, int n,
                double *uv) {
  Camera c = makeCamera(*cam, true);
  for (int i = 0; i < n; ++i) {
    Pt2D p = c.project(Pt3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    uv[2 * i] = p[0];
    uv[2 * i + 1] = p[1];
  }
  return 0;
}

// Camera::lineOfSight for n image points: out = [pt(3), unit(3)] per point.
int ref_line_of_sight(const olpt_camera *cam, const double *uv, int n,
                      double *out) {
  Camera c = makeCamera(*cam, true);
  for (int i = 0; i < n; ++i) {
    Line3D L = c.lineOfSight(Pt2D(uv[2 * i], uv[2 * i + 1]));
    for (int k = 0; k < 3; ++k) {
      out[6 * i + k] = L.pt[k];
      out[6 * i + 3 + k] = L.unit_vector[k];
    }
  }
  return 0;
}

// myMATH::triangulation of n_line lines (pt, unit) -> xyz, err. Returns -1
// if the reference throws.
int ref_triangulate(const double *lines, int n_line, double *xyz,
                    double *err) {
  std::vector<Line3D> los(n_line);
  for (int i = 0; i < n_line; ++i) {
    los[i].pt = Pt3D(lines[6 * i], lines[6 * i + 1], lines[6 * i + 2]);
    los[i].unit_vector =
        Pt3D(lines[6 * i + 3], lines[6 * i + 4], lines[6 * i + 5]);
  }
  Pt3D X;
  try {
    myMATH::triangulation(X, *err, los);
  } catch (...) {
    return -1;
  }
  xyz[0] = X[0];
  xyz[1] = X[1];
  xyz[2] = X[2];
  return 0;
}

// myMATH::polyfit(order).
int ref_polyfit(const double *x, const double *y, int n, int order,
                double *coeff) {
  std::vector<double> xv(x, x + n), yv(y, y + n), cf;
  myMATH::polyfit(cf, xv, yv, order);
  for (size_t i = 0; i < cf.size(); ++i)
    coeff[i] = cf[i];
  return 0;
}

// OTF::getOTFParam -> {a, b, c, cos(alpha), sin(alpha)} for n points.
int ref_otf_param(const olpt_camera *cams, int n_cam, const olpt_otf *otf,
                  int cam_id, const double *xyz, int n, double *out) {
  std::vector<Camera> cv = makeCameras(cams, n_cam, nullptr);
  OTF o;
  AxisLimit lim(otf->boundary[0], otf->boundary[1], otf->boundary[2],
                otf->boundary[3], otf->boundary[4], otf->boundary[5]);
  o.loadParam(otf->n_cam, otf->nx, otf->ny, otf->nz, lim, cv);
  const int ng = otf->nx * otf->ny * otf->nz;
  for (int c = 0; c < otf->n_cam; ++c)
    for (int g = 0; g < ng; ++g) {
      o._param.a(c, g) = otf->a[c * ng + g];
      o._param.b(c, g) = otf->b[c * ng + g];
      o._param.c(c, g) = otf->c[c * ng + g];
      o._param.alpha(c, g) = otf->alpha[c * ng + g];
    }
  for (int i = 0; i < n; ++i) {
    std::vector<double> v =
        o.getOTFParam(cam_id, Pt3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    for (int k = 0; k < 5; ++k)
      out[5 * i + k] = v[k];
  }
  return 0;
}

// ObjectFinder2D::findObject2D (tracer branch) on one image.
int ref_find_peaks(const double *img, int H, int W,
                   const olpt_ipr_params *p, int cap, double *xy,
                   int32_t *n_out) {
  TracerConfig cfg;
  cfg._min_obj_int = p->min_obj_int;
  cfg._radius_obj = p->radius_obj;
  Image im = makeImage(img, H, W);
  ObjectFinder2D finder;
  auto found = finder.findObject2D(im, cfg);
  *n_out = (int32_t)found.size();
  if ((int)found.size() > cap)
    return 1;
  for (size_t i = 0; i < found.size(); ++i) {
    xy[2 * i] = found[i]->_pt_center[0];
    xy[2 * i + 1] = found[i]->_pt_center[1];
  }
  return 0;
}

// StereoMatch, stage by stage (the sequence of StereoMatch::match,
// StereoMatch.cpp:65-178), plus a call of match() itself as a cross-check.
//   cand_ids [cap_cand * n_cam], cand_e [cap_cand]  : checked candidates in
//                                                     id-sorted order
//   sel_idx  [cap_cand]                              : pruned selection, as
//                                                     indices into cand_*
//   obj_xyz  [cap_obj * 3], obj_ids [cap_obj*n_cam]  : triangulated objects
//   counts   [4] = {n_cand, n_sel, n_obj, match_agrees}
int ref_stereo_match(const olpt_camera *cams, int n_cam,
                     const int32_t *active, const double *xy,
                     const int32_t *n_pts, const olpt_ipr_params *p,
                     int cap_cand, int32_t *cand_ids, double *cand_e,
                     int32_t *sel_idx, int cap_obj, double *obj_xyz,
                     int32_t *obj_ids, int32_t *counts) {
  CoutRedirect quiet(false);
  std::vector<Camera> cv = makeCameras(cams, n_cam, active);
  TracerConfig cfg;
  fillConfig(cfg, *p, nullptr, cv);
  auto o2d = makeObj2d(xy, n_pts, n_cam);
  StereoMatch sm(cv, o2d, cfg);

  int ref_cam = -1;
  size_t best = (size_t)-1;
  for (int i = 0; i < n_cam; ++i)
    if (cv[i]._is_active && o2d[i].size() < best) {
      best = o2d[i].size();
      ref_cam = i;
    }
  struct MS {
    std::vector<int> ids;
    double e;
  };
  std::vector<MS> all;
  for (int r = 0; r < (int)o2d[ref_cam].size(); ++r) {
    std::vector<int> seed(n_cam, -1);
    seed[ref_cam] = r;
    std::vector<std::vector<int>> bc;
    DebugAccess_StereoMatch::build(sm, seed, bc);
    for (auto &cand : bc) {
      double e = std::numeric_limits<double>::quiet_NaN();
      if (!DebugAccess_StereoMatch::check(sm, cand, e))
        continue;
      all.push_back(MS{cand, e});
    }
  }
  std::sort(all.begin(), all.end(),
            [](const MS &a, const MS &b) { return a.ids < b.ids; });
  std::vector<std::vector<int>> cands;
  std::vector<double> es;
  for (auto &m : all) {
    cands.push_back(m.ids);
    es.push_back(m.e);
  }
  auto sel = DebugAccess_StereoMatch::prune(sm, cands, es);
  auto objs = DebugAccess_StereoMatch::triangulate(sm, sel);
  auto objs_match = sm.match();

  counts[0] = (int32_t)cands.size();
  counts[1] = (int32_t)sel.size();
  counts[2] = (int32_t)objs.size();
  bool agree = objs.size() == objs_match.size();
  for (size_t i = 0; agree && i < objs.size(); ++i)
    for (int k = 0; k < 3; ++k)
      agree = agree && objs[i]->_pt_center[k] == objs_match[i]->_pt_center[k];
  counts[3] = agree ? 1 : 0;
  if ((int)cands.size() > cap_cand || (int)objs.size() > cap_obj)
    return 1;
  for (size_t j = 0; j < cands.size(); ++j) {
    for (int c = 0; c < n_cam; ++c)
      cand_ids[j * n_cam + c] = cands[j][c];
    cand_e[j] = es[j];
  }
  for (size_t s = 0; s < sel.size(); ++s) {
    auto it = std::lower_bound(cands.begin(), cands.end(), sel[s]);
    sel_idx[s] = (int32_t)(it - cands.begin());
  }
  for (size_t i = 0; i < objs.size(); ++i) {
    for (int k = 0; k < 3; ++k)
      obj_xyz[3 * i + k] = objs[i]->_pt_center[k];
    // recover ids: obj2d_list holds clones of the matched observations
    for (int c = 0; c < n_cam; ++c)
      obj_ids[i * n_cam + c] = -1;
  }
  // objects come from the selected list in order, minus failed gates; map
  // them back by position through the triangulation order.
  {
    size_t oi = 0;
    for (size_t s = 0; s < sel.size() && oi < objs.size(); ++s) {
      bool same = true;
      for (int c = 0; c < n_cam && same; ++c) {
        const int pid = sel[s][c];
        if (pid < 0)
          continue;
        const auto &q = o2d[c][pid]->_pt_center;
        const auto &o = objs[oi]->_obj2d_list[c]->_pt_center;
        same = (q[0] == o[0] && q[1] == o[1]);
      }
      if (same) {
        for (int c = 0; c < n_cam; ++c)
          obj_ids[oi * n_cam + c] = sel[s][c];
        ++oi;
      }
    }
  }
  return 0;
}

// IPR::saveObjInfo (IPR.cpp:307-343) of n tracers: xyz [n][3], uv
// [n][n_cam][2] (a NaN x leaves that camera without an observation).
int ref_save_obj_info(const olpt_camera *cams, int n_cam, const double *xyz, const double *uv,
                      int n, const char *filename) {
  std::vector<Camera> cv = makeCameras(cams, n_cam, nullptr);
  TracerConfig cfg;
  std::vector<std::unique_ptr<Object3D>> objs;
  for (int i = 0; i < n; ++i) {
    auto t = std::make_unique<Tracer3D>();
    t->_pt_center = Pt3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    t->_obj2d_list.resize(n_cam);
    for (int c = 0; c < n_cam; ++c) {
      const double u = uv[((size_t)i * n_cam + c) * 2], v = uv[((size_t)i * n_cam + c) * 2 + 1];
      if (std::isnan(u)) continue;
      auto o = std::make_unique<Tracer2D>();
      o->_pt_center = Pt2D(u, v);
      t->_obj2d_list[c] = std::move(o);
    }
    objs.push_back(std::move(t));
  }
  IPR ipr(cv);
  ipr.saveObjInfo(filename, objs, cfg);
  return 0;
}

// StereoMatch::pruneMatch alone on given candidates: ids [M][n_cam] (-1 =
// camera unused; the id vectors must be distinct), e [M], n_pts[c] 2D
// points per camera (their positions play no part); sel[j] = 1 for the
// selected candidates.
int ref_prune(const olpt_camera *cams, int n_cam, const int32_t *active,
              const int32_t *n_pts, const olpt_ipr_params *p, const int32_t *ids,
              const double *e, int M, uint8_t *sel) {
  CoutRedirect quiet(false);
  std::vector<Camera> cv = makeCameras(cams, n_cam, active);
  TracerConfig cfg;
  fillConfig(cfg, *p, nullptr, cv);
  size_t tot = 0;
  for (int c = 0; c < n_cam; ++c) tot += (size_t)n_pts[c];
  std::vector<double> xy(2 * tot + 2, 0.0);
  auto o2d = makeObj2d(xy.data(), n_pts, n_cam);
  StereoMatch sm(cv, o2d, cfg);
  std::vector<std::vector<int>> cands((size_t)M, std::vector<int>(n_cam, -1));
  std::vector<double> es(e, e + M);
  for (int j = 0; j < M; ++j)
    for (int c = 0; c < n_cam; ++c) cands[j][c] = ids[(size_t)j * n_cam + c];
  auto out = DebugAccess_StereoMatch::prune(sm, cands, es);
  // the output lists the selected candidates in ascending index order
  size_t k = 0;
  for (int j = 0; j < M; ++j) {
    sel[j] = (k < out.size() && out[k] == cands[j]) ? 1 : 0;
    if (sel[j]) ++k;
  }
  return k == out.size() ? 0 : 1;
}

// Shake::runShake on given objects, then the IPR-mode residual of the kept
// objects (IPR.cpp:178-185).
int ref_shake(const olpt_camera *cams, int n_cam, const int32_t *active,
              const double *imgs, int H, int W, const olpt_ipr_params *p,
              const olpt_otf *otf, const double *xyz_in, int n_obj,
              double *xyz_out, uint8_t *flags, double *score,
              double *res_out) {
  CoutRedirect quiet(false);
  std::vector<Camera> cv = makeCameras(cams, n_cam, active);
  TracerConfig cfg;
  fillConfig(cfg, *p, otf, cv);
  auto images = makeImages(imgs, n_cam, H, W);
  auto objs = makeObjects(cfg, xyz_in, n_obj);
  Shake shaker(cv, cfg);
  std::vector<ObjFlag> fl = shaker.runShake(objs, images);
  const auto &sc = DebugAccess_Shake::score(shaker);
  for (int i = 0; i < n_obj; ++i) {
    for (int k = 0; k < 3; ++k)
      xyz_out[3 * i + k] = objs[i]->_pt_center[k];
    flags[i] = static_cast<uint8_t>(fl[i]);
    score[i] = sc[i];
  }
  if (res_out) {
    std::vector<std::unique_ptr<Object3D>> kept;
    for (int i = 0; i < n_obj; ++i)
      if (fl[i] == ObjFlag::None)
        kept.push_back(std::move(objs[i]));
    auto res = shaker.calResidualImage(kept, images);
    for (int c = 0; c < n_cam; ++c)
      std::memcpy(res_out + (size_t)c * H * W, res[c].data(),
                  sizeof(double) * H * W);
  }
  return 0;
}

// IPR::runIPR.  obj_xyz [cap*3], obj_uv [cap*n_cam*2] (final projections),
// loop_total [cap_loop]: cumulative object count after each IPR loop, parsed
// from the reference's own progress lines; counts = {n_obj, n_loop}.
int ref_run_ipr(const olpt_camera *cams, int n_cam, const double *imgs,
                int H, int W, const olpt_ipr_params *p, const olpt_otf *otf,
                int cap, double *obj_xyz, double *obj_uv, int cap_loop,
                int32_t *loop_total, int32_t *counts) {
  std::vector<Camera> cv = makeCameras(cams, n_cam, nullptr);
  TracerConfig cfg;
  fillConfig(cfg, *p, otf, cv);
  auto images = makeImages(imgs, n_cam, H, W);
  std::vector<std::unique_ptr<Object3D>> objs;
  std::string log;
  {
    CoutRedirect cap_out(true);
    IPR ipr(cv);
    objs = ipr.runIPR(cfg, std::move(images));
    log = cap_out.text();
  }
  int n_loop = 0;
  const std::string key = "TOTAL OBJECTS = ";
  for (size_t pos = log.find(key); pos != std::string::npos;
       pos = log.find(key, pos + 1)) {
    if (n_loop < cap_loop)
      loop_total[n_loop] = std::atoi(log.c_str() + pos + key.size());
    ++n_loop;
  }
  counts[0] = (int32_t)objs.size();
  counts[1] = n_loop;
  if ((int)objs.size() > cap)
    return 1;
  for (size_t i = 0; i < objs.size(); ++i) {
    for (int k = 0; k < 3; ++k)
      obj_xyz[3 * i + k] = objs[i]->_pt_center[k];
    for (int c = 0; c < n_cam; ++c) {
      obj_uv[(i * n_cam + c) * 2] = objs[i]->_obj2d_list[c]->_pt_center[0];
      obj_uv[(i * n_cam + c) * 2 + 1] = objs[i]->_obj2d_list[c]->_pt_center[1];
    }
  }
  return 0;
}

} // extern "C"

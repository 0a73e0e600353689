// ref_harness_stb.cpp - extern "C" driver around the reference's STB.
//
// TEST INFRASTRUCTURE ONLY, compiled with the reference's sources into
// oracle/_ref/libolpt_ref.so by oracle/Makefile (never linked by the
// product).  It runs the reference exactly as its command-line program does
// for one object type: BasicSetting::readConfig (Config.cpp:8-158), the STB
// constructor with TracerConfig::readConfig and its OTF estimate from the
// first frames (STB.cpp:6-38, Config.cpp:238-276, OTF.cpp:16-144), then
// ImageIO::loadImg + STB::processFrame per frame (main.cpp:100-119), and
// reads the four track deques back.  Volume self-calibration is switched
// off by marking the cameras calibrated (STB.cpp:631-634), because it may
// rewrite the cameras mid-run.  predictNext / checkLinearFit and the
// predictive field are reached through the reference's FRIEND_DEBUG hook
// (STB.h:107, STBCommons.h:57-61).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "Config.h"
#include "ImageIO.h"
#include "PredField.h"
#include "STB.h"
#include "Track.h"
#include "ref_sink.hpp"

// PredField keeps its field private and has no debug hook.  An explicit
// instantiation may name a private member ([temp.spec.general]: access
// checks do not apply to its template arguments), so a pointer to
// PredField::_disp_field is taken that way: no redefined keyword, one
// definition of the class everywhere.
template <typename Tag, typename Tag::type M> struct MemberPtr {
  friend typename Tag::type member_ptr(Tag) { return M; }
};
struct PredFieldDisp {
  typedef Matrix<double> PredField::*type;
  friend type member_ptr(PredFieldDisp);
};
template struct MemberPtr<PredFieldDisp, &PredField::_disp_field>;

struct DebugAccess_STB {
  static ObjectConfig &cfg(STB &s) { return *s._obj_config; }
  static const BasicSetting &setting(STB &s) { return s._basic_setting; }
  static std::unique_ptr<Object3D> predict(STB &s, const Track &t) { return s.predictNext(t); }
  static bool linfit(STB &s, const Track &t) { return s.checkLinearFit(t); }
};

namespace {

struct RefSTB {
  BasicSetting setting;
  std::unique_ptr<STB> stb;
  std::vector<ImageIO> io;
  std::vector<Image> imgs;
};

struct CoutMute : CoutRedirect {
  CoutMute() : CoutRedirect(false) {}
};

std::deque<Track> *deque_of(STB &s, int which) {
  switch (which) {
  case 0: return &s._short_track_active;
  case 1: return &s._long_track_active;
  case 2: return &s._long_track_inactive;
  case 3: return &s._exit_track;
  default: return nullptr;
  }
}

Track make_track(STB &s, const double *pos, int n, int t0) {
  ObjectConfig &cfg = DebugAccess_STB::cfg(s);
  Track tr;
  for (int i = 0; i < n; ++i) {
    CreateArgs a;
    auto obj = cfg.creatObject3D(std::move(a));
    obj->_pt_center = Pt3D(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    tr.addNext(std::move(obj), t0 + i);
  }
  return tr;
}

} // namespace

extern "C" {

// Read the configuration and build the STB of its first object type.
// otf_abca receives the estimated OTF (a, b, c, alpha of grid node 0 for
// every camera, [n_cam][4]) when not null.  Returns null on failure.
void *ref_stb_open(const char *config_path, double *otf_abca, int quiet) {
  auto *h = new RefSTB();
  std::unique_ptr<CoutMute> mute(quiet ? new CoutMute() : nullptr);
  try {
    if (!h->setting.readConfig(config_path)) {
      delete h;
      return nullptr;
    }
    h->stb.reset(new STB(h->setting, h->setting._object_types.at(0),
                         h->setting._object_config_paths.at(0)));
    ObjectConfig &cfg = DebugAccess_STB::cfg(*h->stb);
    cfg._vsc_param._camera_calibrated = true;
    cfg._vsc_param._otf_calibrated = true;
    for (const auto &p : h->setting._image_file_paths) {
      ImageIO io;
      io.loadImgPath("", p);
      h->io.push_back(io);
    }
    h->imgs.resize(h->io.size());
    if (otf_abca) {
      auto *tc = dynamic_cast<TracerConfig *>(&cfg);
      if (!tc) throw std::runtime_error("not a tracer configuration");
      for (int k = 0; k < h->setting._n_cam; ++k) {
        otf_abca[4 * k + 0] = tc->_otf._param.a(k, 0);
        otf_abca[4 * k + 1] = tc->_otf._param.b(k, 0);
        otf_abca[4 * k + 2] = tc->_otf._param.c(k, 0);
        otf_abca[4 * k + 3] = tc->_otf._param.alpha(k, 0);
      }
    }
  } catch (const std::exception &e) {
    std::cerr << "ref_stb_open: " << e.what() << std::endl;
    delete h;
    return nullptr;
  } catch (...) {
    delete h;
    return nullptr;
  }
  return h;
}

void ref_stb_close(void *hp) { delete static_cast<RefSTB *>(hp); }

// One frame, as the command-line program's loop does.
int ref_stb_process(void *hp, int frame_id, int quiet) {
  auto *h = static_cast<RefSTB *>(hp);
  std::unique_ptr<CoutMute> mute(quiet ? new CoutMute() : nullptr);
  try {
    for (size_t i = 0; i < h->io.size(); ++i) h->imgs[i] = h->io[i].loadImg(frame_id);
    h->stb->processFrame(frame_id, h->imgs);
  } catch (const std::exception &e) {
    std::cerr << "ref_stb_process: " << e.what() << std::endl;
    return 1;
  } catch (...) {
    return 1;
  }
  return 0;
}

// Number of tracks in deque `which` (0 short active, 1 long active, 2 long
// inactive, 3 exited) and their total point count.
int ref_stb_counts(void *hp, int which, int32_t *n_tracks, int32_t *n_points) {
  auto *h = static_cast<RefSTB *>(hp);
  std::deque<Track> *d = deque_of(*h->stb, which);
  if (!d) return 1;
  *n_tracks = (int32_t)d->size();
  long np = 0;
  for (const auto &t : *d) np += (long)t._obj3d_list.size();
  *n_points = (int32_t)np;
  return 0;
}

// Rows (track index in the deque, frame, x, y, z) of deque `which`.
int ref_stb_tracks(void *hp, int which, int cap, int32_t *tid, int32_t *frame, double *xyz) {
  auto *h = static_cast<RefSTB *>(hp);
  std::deque<Track> *d = deque_of(*h->stb, which);
  if (!d) return 1;
  long r = 0;
  for (size_t i = 0; i < d->size(); ++i) {
    const Track &t = (*d)[i];
    for (size_t j = 0; j < t._obj3d_list.size(); ++j) {
      if (r >= cap) return 2;
      tid[r] = (int32_t)i;
      frame[r] = t._t_list[j];
      for (int a = 0; a < 3; ++a) xyz[3 * r + a] = t._obj3d_list[j]->_pt_center[a];
      ++r;
    }
  }
  return 0;
}

// STB::predictNext (STB.cpp:233-316) for tracks given as point lists:
// pos [n_tr][len][3] -> out [n_tr][3]; ok[i] = 0 where it returns null.
int ref_stb_predict(void *hp, const double *pos, int n_tr, int len, double *out, int32_t *ok) {
  auto *h = static_cast<RefSTB *>(hp);
  for (int i = 0; i < n_tr; ++i) {
    Track t = make_track(*h->stb, pos + (long)i * len * 3, len, 0);
    auto p = DebugAccess_STB::predict(*h->stb, t);
    ok[i] = p ? 1 : 0;
    for (int a = 0; a < 3; ++a) out[3 * i + a] = p ? p->_pt_center[a] : 0.0;
  }
  return 0;
}

// STB::checkLinearFit (STB.cpp:874-924) for tracks given as point lists.
int ref_stb_linfit(void *hp, const double *pos, int n_tr, int len, int32_t *ok) {
  auto *h = static_cast<RefSTB *>(hp);
  for (int i = 0; i < n_tr; ++i) {
    Track t = make_track(*h->stb, pos + (long)i * len * 3, len, 0);
    ok[i] = DebugAccess_STB::linfit(*h->stb, t) ? 1 : 0;
  }
  return 0;
}

// PredField::calPredField (PredField.cpp:21-86) with the STB's parameters
// on two point sets -> field [nx*ny*nz][3]; then getDisp at query points.
int ref_pred_field(void *hp, const double *prev, int n_prev, const double *curr, int n_curr,
                   double *field, const double *q, int n_q, double *disp) {
  auto *h = static_cast<RefSTB *>(hp);
  ObjectConfig &cfg = DebugAccess_STB::cfg(*h->stb);
  auto mk = [&](const double *p, int n) {
    std::vector<std::unique_ptr<Object3D>> v;
    for (int i = 0; i < n; ++i) {
      CreateArgs a;
      auto obj = cfg.creatObject3D(std::move(a));
      obj->_pt_center = Pt3D(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
      v.push_back(std::move(obj));
    }
    return v;
  };
  auto a = mk(prev, n_prev), b = mk(curr, n_curr);
  PredField pf(cfg);
  pf.calPredField(a, b);
  const PFParam &pp = cfg._pf_param;
  const int ng = pp.nx * pp.ny * pp.nz;
  const Matrix<double> &f = pf.*member_ptr(PredFieldDisp());
  for (int g = 0; g < ng; ++g)
    for (int k = 0; k < 3; ++k) field[3 * g + k] = f(g, k);
  for (int i = 0; i < n_q; ++i) {
    Pt3D d = pf.getDisp(Pt3D(q[3 * i], q[3 * i + 1], q[3 * i + 2]));
    for (int k = 0; k < 3; ++k) disp[3 * i + k] = d[k];
  }
  return 0;
}

} // extern "C"

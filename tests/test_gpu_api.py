"""The reference-named Python classes (pyopenlpt_amd.lpt: ObjectFinder2D,
StereoMatch, Shake, OTF, IPR) on the golden cases, FP64, against the
reference's outputs: the same bars as the C-ABI tests, through the surface
the reference's notebooks call (pyIPR.cpp, pyShake.cpp, pyStereoMatch.cpp,
pyObjectFinder.cpp, pyOTF.cpp)."""
import ctypes

import numpy as np
import pytest

from common import GOLDEN_CASES, Case

pytestmark = pytest.mark.gpu


def _otf(case):
    from pyopenlpt_amd import config as C
    if "otf_a" in case.z:
        return C.Otf(case.n_cam, tuple(int(v) for v in case.z["otf_shape"]),
                     tuple(float(v) for v in case.z["otf_boundary"]), case.z["otf_a"].copy(),
                     case.z["otf_b"].copy(), case.z["otf_c"].copy(), case.z["otf_alpha"].copy())
    a, b, c, al = case.abca
    n = case.n_cam
    return C.Otf(n, (2, 2, 2), (-20.0, 20.0) * 3, np.full((n, 8), a), np.full((n, 8), b),
                 np.full((n, 8), c), np.full((n, 8), al))


@pytest.fixture(scope="module", params=GOLDEN_CASES)
def case(request, gpu_available):
    from pyopenlpt_amd import config as C
    c = Case(request.param)
    c.cfg = C.TracerConfig.from_ipr_params(c.params, _otf(c))
    return c


def test_object_finder(case):
    from pyopenlpt_amd import lpt
    finder = lpt.ObjectFinder2D()
    for k in range(case.n_cam):
        objs = finder.findObject2D(lpt.Image.from_numpy(case.images[k]), case.cfg)
        got = np.array([o._pt_center for o in objs]).reshape(-1, 2)
        want = case.z[f"peaks_cam{k}"]
        assert got.shape == want.shape
        assert np.abs(got - want).max(initial=0) <= 1e-9


def test_stereo_match_and_prune(case):
    from pyopenlpt_amd import lpt
    obj2d = [[lpt.Tracer2D(p.copy(), case.cfg.radius_obj) for p in case.z[f"peaks_cam{k}"]]
             for k in range(case.n_cam)]
    sm = lpt.StereoMatch(case.cams, obj2d, case.cfg, precision="f64")
    objs = sm.match()
    xyz = np.array([o._pt_center for o in objs]).reshape(-1, 3)
    assert np.abs(xyz - case.z["sm_obj_xyz"]).max(initial=0) <= 1e-12
    for o, ids in zip(objs, case.z["sm_obj_ids"]):
        for k, i in enumerate(ids):
            if i < 0:
                assert o._obj2d_list[k] is None
            else:
                assert np.array_equal(o._obj2d_list[k]._pt_center, case.z[f"peaks_cam{k}"][i])
    sel = sm.pruneMatch(case.z["sm_cand_ids"].tolist(), case.z["sm_cand_e"].tolist())
    assert sel == case.z["sm_cand_ids"][case.z["sm_sel"]].tolist()


def test_shake_run_and_residual(case):
    from pyopenlpt_amd import lpt
    objs = [lpt.Tracer3D(x.copy(), [None] * case.n_cam, case.cfg.radius_obj)
            for x in case.z["sm_obj_xyz"]]
    imgs = [case.images[k] for k in range(case.n_cam)]
    shaker = lpt.Shake(case.cams, case.cfg, precision="f64")
    flags, new = shaker.runShake(objs, imgs)
    assert [int(f) for f in flags] == case.z["sh_flags"].tolist()
    got = np.array([o._pt_center for o in new]).reshape(-1, 3)
    assert np.abs(got - case.z["sh_xyz"]).max(initial=0) <= 1e-9
    assert all(np.array_equal(o._pt_center, x) for o, x in zip(objs, case.z["sm_obj_xyz"]))
    res = shaker.calResidualImage(new, imgs, flags)
    flat = np.stack(res).reshape(case.n_cam, -1)
    assert np.abs(flat[:, case.z["sh_res_idx"]] - case.z["sh_res_sample"]).max() <= 1e-7


def test_otf_get_param(case):
    import oracle as O
    from pyopenlpt_amd import lpt
    otf = lpt.OTF(case.cfg.otf)
    pts = np.random.default_rng(5).uniform(-25, 25, (50, 3))
    o, _keep = case.otf()
    L = O.load_port()
    for k in range(case.n_cam):
        want = np.zeros((len(pts), 5))
        assert L.orc_otf_param(ctypes.byref(o), k, O.dptr(np.ascontiguousarray(pts)), len(pts),
                               O.dptr(want)) == 0
        got = otf.getOTFParams(k, pts)
        assert np.abs(got - want).max() <= 1e-12
        a, b, c, cos_a = otf.getOTFParam(k, pts[0])
        assert np.allclose((a, b, c, cos_a), want[0, :4], rtol=0, atol=1e-12)


def test_ipr_run_and_save(case, tmp_path):
    from pyopenlpt_amd import config as C
    from pyopenlpt_amd import lpt
    ipr = lpt.IPR(case.cams, precision="f64")
    objs = ipr.runIPR(case.cfg, [case.images[k] for k in range(case.n_cam)])
    xyz = np.array([o._pt_center for o in objs]).reshape(-1, 3)
    assert len(objs) == len(case.z["ipr_xyz"])
    assert np.abs(xyz - case.z["ipr_xyz"]).max(initial=0) <= 1e-9
    f1, f2 = tmp_path / "a.csv", tmp_path / "b.csv"
    ipr.saveObjInfo(str(f1), objs, case.cfg)
    uv = np.array([[o2._pt_center for o2 in o._obj2d_list] for o in objs]).reshape(-1, case.n_cam, 2)
    C.save_obj_info(str(f2), xyz, uv)
    assert f1.read_bytes() == f2.read_bytes()

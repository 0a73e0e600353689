"""The oracle's tracking restatements (oracle/olpt_oracle.c: orc_stb_predict,
orc_stb_linear_fit, orc_pred_field) against the compiled reference's golden
vectors (tests/golden/stb_abc_256.npz, written by make_golden_stb.py from
STB::predictNext, STB::checkLinearFit and PredField::calPredField/getDisp),
plus what the mathematics fixes: a straight track passes the linear fit and
is extrapolated one step along the line; a uniformly displaced cloud gives
that displacement back from the predictive field."""
import ctypes
import os

import numpy as np

import oracle as O

GOLD = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                 "stb_abc_256.npz")))
LIM = np.array([-20.0, 20.0, -20.0, 20.0, -20.0, 20.0])


def predict(pos, lens):
    L = O.load_port()
    n, stride = pos.shape[:2]
    out = np.zeros((n, 3))
    ok = np.zeros(n, np.int32)
    assert L.orc_stb_predict(O.dptr(np.ascontiguousarray(pos)), O.iptr(np.ascontiguousarray(lens, np.int32)),
                             n, stride, O.dptr(out), O.iptr(ok)) == 0
    return out, ok


def linfit(pos, lens, n_init=4):
    L = O.load_port()
    n, stride = pos.shape[:2]
    ok = np.zeros(n, np.int32)
    assert L.orc_stb_linear_fit(n_init, O.dptr(np.ascontiguousarray(pos)),
                                O.iptr(np.ascontiguousarray(lens, np.int32)), n, stride,
                                O.iptr(ok)) == 0
    return ok


def pred_field(prev, curr, q, grid=(11, 11, 11), r=None, smooth=1):
    L = O.load_port()
    vox = float(GOLD["voxel_mm"])
    r = 25.0 * vox if r is None else r
    ng = int(np.prod(grid))
    field = np.zeros((ng, 3))
    disp = np.zeros((len(q), 3))
    nbin = np.array([11, 11, 11], np.int32)
    sig = np.ones(3)
    assert L.orc_pred_field(O.dptr(LIM), *grid, ctypes.c_double(r), O.iptr(nbin), smooth, O.dptr(sig),
                            O.dptr(np.ascontiguousarray(prev)), len(prev),
                            O.dptr(np.ascontiguousarray(curr)), len(curr), O.dptr(field),
                            O.dptr(np.ascontiguousarray(q)), len(q), O.dptr(disp)) == 0
    return field, disp


def test_predict_next_matches_reference():
    out, ok = predict(GOLD["pn_pos"], GOLD["pn_len"])
    assert np.array_equal(ok, GOLD["pn_ok"])
    assert np.array_equal(out[ok == 1], GOLD["pn_pred"][ok == 1])


def test_linear_fit_matches_reference():
    ok = linfit(GOLD["lf_pos"], GOLD["lf_len"])
    assert np.array_equal(ok, GOLD["lf_ok"])


def test_pred_field_matches_reference():
    field, disp = pred_field(GOLD["pf_prev"], GOLD["pf_curr"], GOLD["pf_q"])
    assert np.array_equal(field, GOLD["pf_field"])
    assert np.array_equal(disp, GOLD["pf_disp"])


def test_straight_tracks():
    g = np.random.default_rng(3)
    n = 200
    p0 = g.uniform(-15, 15, (n, 3))
    v = g.normal(0, 0.3, (n, 3))
    t = np.arange(7)[None, :, None]
    pos = p0[:, None, :] + v[:, None, :] * t
    lens = np.full(n, 6, np.int32)
    assert linfit(pos, lens).all()
    bent = pos.copy()
    bent[:, 5] += 0.2                       # last point 0.2 mm off the line
    assert not linfit(bent, lens).any()
    # predictNext's normalised LMS from zero converges in one step to the
    # minimum-norm filter f = y_t s / |s|^2 over the previous samples s, so
    # its one-step prediction is y_t (s . s') / (s . s), s' the samples shifted
    # by one (+10 on all samples when |y_t| < 1)
    out, ok = predict(pos, lens)
    assert ok.all()
    for a in range(3):
        ser = pos[:, :6, a].copy()          # 6 points: order 5, all used
        sh = np.abs(ser[:, -1]) < 1.0
        ser[sh] += 10.0
        s0, s1, yt = ser[:, :-1], ser[:, 1:], ser[:, -1]
        want = yt * (s0 * s1).sum(1) / (s0 * s0).sum(1) - np.where(sh, 10.0, 0.0)
        assert np.abs(out[:, a] - want).max() <= 1e-9 * max(1.0, np.abs(want).max())


def test_isolated_displacements_are_recovered_exactly():
    """Points on a 10 mm lattice (no node cube of half-width r = 4.1 mm holds
    two), each displaced by d: a node's histogram is then one Gaussian splat
    exp(-|bin - b(d)|^2), whose per-axis log-parabola fit returns b(d)
    exactly, so with smoothing off the field at every node that sees both
    points is d, and getDisp at a point whose 8 cell corners all see it
    (2 mm grid) is d."""
    ax = np.array([-15.0, -5.0, 5.0, 15.0])
    prev = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3) + 0.3
    d = np.array([0.25, -0.1, 0.05])
    curr = prev + d
    field, disp = pred_field(prev, curr, prev, grid=(21, 21, 21), smooth=0)
    assert np.abs(disp - d).max() < 1e-9
    seen = np.abs(field).sum(1) > 0
    assert seen.sum() > 0 and np.abs(field[seen] - d).max() < 1e-9

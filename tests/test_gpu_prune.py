"""GPU pruneMatch (olpt_prune_match) on hand-built candidate sets against the
compiled reference's selections (tests/golden/make_golden_prune.py).

  random / dense  exact and one-ulp e_check ties, shared points: the GPU's
                  selection equals the reference's, bit for bit
  chain           three candidates on one point whose e_check values form a
                  non-transitive eq_eps chain: the reference's order there is
                  whatever its std::stable_sort produces, so the library
                  refuses with OLPT_E_PRUNE_TIES instead of returning a
                  selection that may differ
"""
import os

import numpy as np
import pytest

from common import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu_available):
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.context import Context
    cams = S.make_rig(256, 256, max_intensity=255)
    c = Context(cams, 256, 256, max_2d=1 << 12, max_cand=1 << 14, max_obj=1 << 12)
    yield c
    c.close()


def _case(name):
    z = np.load(os.path.join(GOLDEN, "prune_cases.npz"))
    return z[f"{name}_ids"], z[f"{name}_e"], z[f"{name}_npts"], z[f"{name}_sel"].astype(bool)


@pytest.mark.parametrize("name", ["random", "dense"])
def test_prune_matches_reference(ctx, name):
    from pyopenlpt_amd import synth as S
    ids, e, npts, want = _case(name)
    params = S.default_params(40.0 / (0.95 * 256))
    got = ctx.prune_match(ids, e, npts, params)
    print(f"prune {name}: {len(e)} candidates, {int(got.sum())} selected (reference {int(want.sum())})")
    assert np.array_equal(got, want)
    assert ctx.diagnostics()[2] == 0


def test_prune_refuses_non_transitive_ties(ctx):
    from pyopenlpt_amd import _lib
    from pyopenlpt_amd import synth as S
    ids, e, npts, _want = _case("chain")
    params = S.default_params(40.0 / (0.95 * 256))
    with pytest.raises(_lib.FatalError) as ei:
        ctx.prune_match(ids, e, npts, params)
    assert ei.value.status == 103
    assert ctx.diagnostics()[2] == 1


def _tol(v):
    return 1e-9 * v + 1e-12


def _with_chain(values, shared=False):
    """The random case plus three candidates with e_check `values` that share
    a point of their own (camera 0, point 300); with shared=True they also
    share their camera-1 point (a second conflict between the same three)."""
    ids, e, npts, _want = _case("random")
    n_pts = [301, 323, 313, 308]
    c1 = [320, 320, 320] if shared else [320, 321, 322]
    extra = np.array([[300, c1[0], 310, 305], [300, c1[1], 311, 306], [300, c1[2], 312, 307]],
                     np.int32)
    ids_c = np.concatenate([ids, extra])
    e_c = np.concatenate([e, np.asarray(values, np.float64)])
    order = np.lexsort(ids_c.T[::-1])
    return ids_c[order], e_c[order], np.array(n_pts, np.int32)


def _chain_values(a):
    """a ~ b, b ~ c, a !~ c under eq_eps (steps of 0.6 of the tolerance)."""
    b = a + 0.6 * _tol(a)
    c = b + 0.6 * _tol(b)
    assert abs(c - a) > _tol(c)
    return [c, b, a]


def _bucket_start_value(slot, size):
    """An e_check value in the middle of the eq_eps check bucket whose home
    slot in a table of `size` slots is `slot` (bucket key: floor(log(v +
    1e-3) / 2e-9) + 2^41 + 2^42 for e_check values)."""
    import math
    k0 = math.floor(math.log(0.01 + 1e-3) / 2e-9)
    off = (1 << 41) + (1 << 42)
    k = k0 + ((slot - (k0 + off)) % size)
    return math.exp((k + 0.5) * 2e-9) - 1e-3


@pytest.mark.parametrize("where", ["spread", "table_wrap"])
def test_prune_refuses_chains_at_any_position(ctx, where):
    """Non-transitive eq_eps chains anywhere in the value range (so across
    and inside the check's bucket boundaries) and at the end of the check
    table (its probe range wraps to slot 0) are all refused."""
    from pyopenlpt_amd import _lib
    from pyopenlpt_amd import synth as S
    params = S.default_params(40.0 / (0.95 * 256))
    if where == "spread":
        rng = np.random.default_rng(77)
        bases = np.exp(rng.uniform(np.log(2e-4), np.log(4e-2), 12))
    else:
        m = len(_case("random")[1]) + 3
        size = 1024
        while size < 8 * m:
            size *= 2
        bases = [_bucket_start_value(size - 1, size), _bucket_start_value(size - 2, size)]
    for a in bases:
        ids, e, npts = _with_chain(_chain_values(float(a)))
        with pytest.raises(_lib.FatalError) as ei:
            ctx.prune_match(ids, e, npts, params)
        assert ei.value.status == 103, a
        assert ctx.diagnostics()[2] == 1


@pytest.mark.parametrize("shared", [False, True])
def test_prune_accepts_transitive_close_values(ctx, shared):
    """Three distinct e_check values within one tolerance of each other (an
    eq_eps clique: the comparator stays a strict weak order) at many
    positions: no refusal, and the selection equals the oracle's pruneMatch."""
    import oracle as O
    from pyopenlpt_amd import synth as S
    params = S.default_params(40.0 / (0.95 * 256))
    rng = np.random.default_rng(78)
    for a in np.exp(rng.uniform(np.log(2e-4), np.log(4e-2), 8)):
        a = float(a)
        vals = [a + 0.45 * _tol(a), a + 0.9 * _tol(a), a]
        ids, e, npts = _with_chain(vals, shared=shared)
        got = ctx.prune_match(ids, e, npts, params)
        assert ctx.diagnostics()[2] == 0
        want = np.zeros(len(e), np.uint8)
        rc = O.load_port().orc_prune(ids.shape[1], None, O.iptr(npts), O.iptr(ids),
                                     O.dptr(e), len(e), O.u8ptr(want))
        assert rc == 0
        assert np.array_equal(got, want.astype(bool)), a

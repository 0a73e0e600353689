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

"""GPU neighbour joining against the oracle: bit-exact against the full-scan
restatement (same q formula and arithmetic), topology-exact against the
restatement of the reference's bin/heap search."""
import numpy as np
import pytest

import mestrado_amd as M
import oracle_bind as O
from test_oracle_nj import _additive, _rand_tree, splits_of_joins, splits_of_tree

pytestmark = pytest.mark.gpu
S = O.blosum62()


def _same(gj, oj, exact=True):
    assert len(gj) == len(oj)
    for g, o in zip(gj, oj):
        assert (g.node, g.left, g.right, g.n) == (o.node, o.left, o.right, o.n)
        if exact:
            assert (g.d, g.Ti, g.Tj, g.q) == (o.d, o.Ti, o.Tj, o.q)
            assert (g.len_left, g.len_right) == (o.lenL, o.lenR)
        else:
            for a, b in ((g.Ti, o.Ti), (g.Tj, o.Tj), (g.len_left, o.lenL), (g.len_right, o.lenR)):
                assert abs(a - b) <= 1e-6 * max(1.0, abs(b))


def test_nj_config0_scores(ctx):
    seqs = M.synth(0, 0)
    D, _ = O.all_pairs(seqs, S)
    D = D.astype(np.float32)
    g = ctx.nj(D)
    _same(g, O.nj_fullscan(D)[0])
    _same(g, O.nj_reference(D)[0], exact=False)


# 230: one workgroup past the first T rebuild (after 200 joins); 450, 700: the
# launches per step
@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (57, 2), (230, 5), (450, 3), (700, 4)])
def test_nj_random_matrices(ctx, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.integers(-300, 1500, (n, n))
    A = np.triu(A) + np.triu(A, 1).T
    D = A.astype(np.float32)
    g = ctx.nj(D)
    _same(g, O.nj_fullscan(D)[0])
    _same(g, O.nj_reference(D)[0], exact=False)


@pytest.mark.parametrize("n", [4, 40, 300])
def test_nj_additive_tree(ctx, n):
    rng = np.random.default_rng(n + 7)
    parent, blen = _rand_tree(rng, n)
    D = _additive(parent, blen, n)
    g = ctx.nj(D)
    assert splits_of_joins(g, n) == splits_of_tree(parent, n)


def test_nj_on_resident_scores(ctx):
    seqs = M.synth(1, 0, n=300)
    st = ctx.upload(seqs)
    sc = st.all_pairs(M.FORM_BATCHED)
    g = st.nj()
    D = sc.astype(np.float32)
    _same(g, O.nj_fullscan(D)[0])
    _same(g, O.nj_reference(D)[0], exact=False)

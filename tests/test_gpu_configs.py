"""The full-size workloads on the device against the oracle: the long-pair set
(256 proteins of 8,000-20,000 residues) through sampled pairs and a subset's
progressive alignment at ~20,000 columns, and the 1,000-protein set's whole
NJ tree and progressive alignment."""
import numpy as np
import pytest

import mestrado_amd as M
import oracle_bind as O

pytestmark = pytest.mark.gpu
S = O.blosum62()


def _gpu_joins(joins):
    out = []
    for J in joins:
        j = M.Join()
        j.node, j.left, j.right = J.node, J.left, J.right
        out.append(j)
    return out


@pytest.fixture(scope="module")
def long_set():
    seqs = M.synth(3, 0)
    assert len(seqs) == 256 and min(map(len, seqs)) >= 8000 and max(map(len, seqs)) <= 20200
    return seqs


def test_long_pairs_sampled(ctx, long_set):
    # every pair here is past the 16-bit bound: the re-based packed kernel in
    # the batched all-pairs matrix, the int32 kernel through explicit pair
    # lists of both forms
    seqs = long_set
    lens = np.array([len(s) for s in seqs])
    longest = list(np.argsort(lens)[-4:])
    pi = [a for a in longest for b in longest if a <= b]
    pj = [b for a in longest for b in longest if a <= b]
    rng = np.random.default_rng(30)
    for _ in range(14):
        a, b = rng.integers(0, len(seqs), 2)
        pi.append(int(a))
        pj.append(int(b))
    codes, offs = O.pack_seqs(seqs)
    want, _ = O.pair_list(codes, offs, pi, pj, S)
    st = ctx.upload(seqs)
    full = st.all_pairs(M.FORM_BATCHED)
    assert (full == full.T).all()
    got = full[np.array(pi), np.array(pj)]
    assert (got == want).all(), [(i, j) for i, j, g, w in zip(pi, pj, got, want) if g != w]
    for form in (M.FORM_PLAIN, M.FORM_BATCHED):
        assert (st.score_pairs(pi, pj, form) == want).all()
    st.free()


def test_long_profiles_msa(ctx, long_set):
    # the four longest sequences (19,000-20,100 residues): three merges whose
    # rows span about 40 strips, on the tiled flag kernel
    seqs = long_set
    idx = np.argsort([len(s) for s in seqs])[-4:]
    sub = [seqs[i] for i in idx]
    st = ctx.upload(sub)
    D = st.all_pairs(M.FORM_BATCHED)
    joins, _ = O.nj_fullscan(D.astype(np.float32))
    rows = st.msa(_gpu_joins(joins))
    st.free()
    assert len(rows[0]) >= max(map(len, sub))
    assert rows == O.progressive_msa_c(sub, joins, S)


def test_config1_tree_and_msa(ctx):
    # the 1,000-protein set: NJ of the resident score matrix (the per-step
    # launches), bit-exact against the full-scan restatement and equal in
    # topology, T and branch lengths to the restatement of the reference's
    # bin and heap search (whose queue-size trigger must not fire), then the
    # progressive alignment along that tree
    seqs = M.synth(1, 0)
    assert len(seqs) == 1000
    st = ctx.upload(seqs)
    D = st.all_pairs(M.FORM_BATCHED)
    codes, offs = O.pack_seqs(seqs)
    rng = np.random.default_rng(31)
    pi, pj = rng.integers(0, 1000, 400), rng.integers(0, 1000, 400)
    assert (D[pi, pj] == O.pair_list(codes, offs, pi, pj, S)[0]).all()
    g = st.nj()
    Df = D.astype(np.float32)
    full, _ = O.nj_fullscan(Df)
    ref, _, trig = O.nj_reference(Df)
    assert trig == 0
    for a, o, r in zip(g, full, ref):
        assert (a.node, a.left, a.right, a.n) == (o.node, o.left, o.right, o.n)
        assert (a.d, a.Ti, a.Tj, a.q, a.len_left, a.len_right) == (o.d, o.Ti, o.Tj, o.q, o.lenL, o.lenR)
        assert (a.node, a.left, a.right, a.n) == (r.node, r.left, r.right, r.n)
        for x, y in ((a.Ti, r.Ti), (a.Tj, r.Tj), (a.len_left, r.lenL), (a.len_right, r.lenR)):
            assert abs(x - y) <= 1e-6 * max(1.0, abs(y))
    rows = st.msa(g)
    st.free()
    assert rows == O.progressive_msa_c(seqs, g, S)

"""GPU progressive alignment against the oracle's profile-profile merges:
identical gapped rows along the same guide tree."""
import numpy as np
import pytest

import mestrado_amd as M
import oracle_bind as O

pytestmark = pytest.mark.gpu
S = O.blosum62()


def _joins(seqs):
    D, _ = O.all_pairs(seqs, S)
    joins, _ = O.nj_fullscan(D.astype(np.float32))
    return joins


def _gpu_joins(joins):
    out = []
    for J in joins:
        j = M.Join()
        j.node, j.left, j.right = J.node, J.left, J.right
        out.append(j)
    return out


def _check(ctx, seqs):
    joins = _joins(seqs)
    st = ctx.upload(seqs)
    rows = st.msa(_gpu_joins(joins))
    ref = O.progressive_msa(seqs, joins, S)
    assert rows == ref
    st.free()
    return rows


def test_two_sequences_is_pairwise(ctx):
    seqs = ["MKTAYIAKQRQISFVKSHFSRQ", "MKTAYAKQRQSFVKSHFSRQAW"]
    rows = _check(ctx, seqs)
    sc, gx, gy, _ = O.nw_align(seqs[0], seqs[1], S)
    assert rows == [gx, gy]


def test_config0_subset(ctx):
    _check(ctx, M.synth(0, 0)[:24])


def test_small_alphabet_ties(ctx):
    rng = np.random.default_rng(6)
    _check(ctx, ["".join(rng.choice(list("AWC"), int(n))) for n in rng.integers(1, 30, 16)])


def test_long_profiles_cross_strips(ctx):
    # profiles longer than one 512-row strip
    rng = np.random.default_rng(7)
    root = "".join(rng.choice(list(O.STANDARD20), 700))
    seqs = []
    for k in range(6):
        s = list(root)
        for t in rng.choice(len(s), 80, replace=False):
            s[t] = rng.choice(list(O.STANDARD20))
        seqs.append("".join(s)[k * 7:])
    _check(ctx, seqs)


def test_long_profiles_tiled(ctx):
    # profiles of more than 8 strips and 2,048 columns take the tiled flag kernel
    rng = np.random.default_rng(8)
    root = "".join(rng.choice(list(O.STANDARD20), 4700))
    seqs = []
    for k in range(4):
        s = list(root)
        for t in rng.choice(len(s), 500, replace=False):
            s[t] = rng.choice(list(O.STANDARD20))
        seqs.append("".join(s)[k * 40: 4700 - k * 30])
    _check(ctx, seqs)

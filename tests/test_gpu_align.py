"""GPU alignment with traceback against the oracle: scores, gapped strings
(path and tie-breaking) bit-exact."""
import numpy as np
import pytest

import mestrado_amd as M
import oracle_bind as O

pytestmark = pytest.mark.gpu
S = O.blosum62()


def _check(ctx, seqs, pi, pj):
    st = ctx.upload(seqs)
    lens = [len(s) for s in seqs]
    sc, gx, gy = st.align_pairs(pi, pj, lens)
    for k, (a, b) in enumerate(zip(pi, pj)):
        osc, ox, oy, _ = O.nw_align(seqs[a], seqs[b], S)
        assert sc[k] == osc, (k, lens[a], lens[b])
        assert (gx[k], gy[k]) == (ox, oy), (k, lens[a], lens[b])
    st.free()


def test_hand_tie_cases(ctx):
    seqs = ["G", "WG", "AAW", "AW", "WCG", "CW", "AA", "A"]
    _check(ctx, seqs, [0, 1, 2, 4, 6, 7], [1, 0, 3, 5, 7, 6])


def test_continuation_tie_cases(ctx):
    # hand-worked cases where extending and opening a gap tie (the reference
    # extends, NeedlemanWunsch.cu:315-316,336): the expected strings are literal
    from test_oracle_nw import CONTINUATION_TIES
    seqs = []
    for x, y, _, _, _ in CONTINUATION_TIES:
        seqs += [x, y]
    st = ctx.upload(seqs)
    k = len(CONTINUATION_TIES)
    sc, gx, gy = st.align_pairs(np.arange(0, 2 * k, 2), np.arange(1, 2 * k, 2),
                                [len(s) for s in seqs])
    for q, (x, y, ext, _, score) in enumerate(CONTINUATION_TIES):
        assert sc[q] == score
        assert (gx[q], gy[q]) == ext, (x, y)
    st.free()


def test_random_small_alphabet_ties(ctx):
    rng = np.random.default_rng(3)
    seqs = ["".join(rng.choice(list("AWCG"), int(n))) for n in rng.integers(1, 40, 60)]
    pi = rng.integers(0, 60, 200)
    pj = rng.integers(0, 60, 200)
    _check(ctx, seqs, pi, pj)


def test_strip_boundaries(ctx):
    rng = np.random.default_rng(4)
    lens = [1, 5, 63, 64, 65, 200, 511, 512, 513, 1024, 1025, 1700]
    seqs = ["".join(rng.choice(list(O.STANDARD20), L)) for L in lens]
    idx = np.arange(len(lens))
    pi, pj = np.meshgrid(idx, idx, indexing="ij")
    _check(ctx, seqs, pi.ravel(), pj.ravel())


def test_packed_rows_per_lane(ctx):
    # the packed flag kernel lays a job's rows out at k = 1..8 rows per lane
    # (one strip of 64k rows up to 512, two strips of 64k each up to 1,024):
    # lengths that give every k in one and in two strips, all pairs, so the
    # traceback steps through each layout's groups
    rng = np.random.default_rng(14)
    lens = [60, 100, 150, 230, 300, 350, 420, 480, 530, 650, 720, 800, 860, 1000]
    seqs = ["".join(rng.choice(list(O.STANDARD20), L)) for L in lens]
    idx = np.arange(len(lens))
    pi, pj = np.meshgrid(idx, idx, indexing="ij")
    _check(ctx, seqs, pi.ravel(), pj.ravel())


def test_long_pairs(ctx):
    # 2,000-6,000 rows: 4 to 12 strips of 512 rows
    rng = np.random.default_rng(12)
    lens = [2049, 3000, 4500, 6000]
    seqs = ["".join(rng.choice(list(O.STANDARD20), L)) for L in lens]
    pi = np.array([0, 1, 2, 3, 3, 1])
    pj = np.array([1, 0, 3, 2, 0, 3])
    _check(ctx, seqs, pi, pj)


def test_tiled_pairs(ctx):
    # more than 8 strips and at least 2,048 columns: tiles of 512 rows x 256
    # columns over anti-diagonal launches; exact and ragged tile edges
    rng = np.random.default_rng(13)
    root = "".join(rng.choice(list(O.STANDARD20), 5200))
    lens = [2048, 4097, 5120, 2300]
    seqs = []
    for L in lens:  # related sequences: long diagonal runs across tile edges
        s = list(root[:L])
        for t in rng.choice(L, L // 6, replace=False):
            s[t] = rng.choice(list(O.STANDARD20))
        seqs.append("".join(s))
    pi = np.array([0, 2, 1, 3, 1])
    pj = np.array([1, 1, 2, 2, 0])
    _check(ctx, seqs, pi, pj)
    kinds = [r[0] for r in ctx.last_launch_stats()]
    assert -3008 in kinds and -3004 in kinds  # tiled and pipelined jobs both ran


def test_config0_pairs(ctx):
    seqs = M.synth(0, 0)
    rng = np.random.default_rng(8)
    pi = rng.integers(0, 64, 120)
    pj = rng.integers(0, 64, 120)
    _check(ctx, seqs, pi, pj)

"""mst_score_pairs_noalloc (no device allocation, a scratch array of the
library) against mst_score_pairs on the same pairs: bit-exact, on two calls
in a row, the second with fewer pairs over shorter sequences."""
import numpy as np
import pytest

import mestrado_amd as M
import oracle_bind as O

pytestmark = pytest.mark.gpu
S = O.blosum62()


def test_noalloc_equals_score_pairs(ctx):
    rng = np.random.default_rng(40)
    lens = list(rng.integers(1, 700, 80)) + [1024, 1025, 1700, 2600]  # one and several strips
    seqs = ["".join(rng.choice(list(O.STANDARD20), int(L))) for L in lens]
    st = ctx.upload(seqs)
    pi, pj = rng.integers(0, len(seqs), 3000), rng.integers(0, len(seqs), 3000)
    want = st.score_pairs(pi, pj, M.FORM_PLAIN)
    assert (st.score_pairs_noalloc(pi, pj) == want).all()
    codes, offs = O.pack_seqs(seqs)
    assert (want[:300] == O.pair_list(codes, offs, pi[:300], pj[:300], S)[0]).all()
    # second call: fewer pairs, all among the short sequences (a smaller launch
    # over the same scratch; nothing of the first call may leak into it)
    qi, qj = rng.integers(0, 40, 77), rng.integers(0, 40, 77)
    assert (st.score_pairs_noalloc(qi, qj) == st.score_pairs(qi, qj, M.FORM_PLAIN)).all()
    st.free()


def test_noalloc_rejects_beyond_scratch(ctx):
    st = ctx.upload(["ACDEFGHIK", "WWW"])
    big = np.zeros(6_000_000, dtype=np.int32)  # 12 B per pair: above 64 MiB
    with pytest.raises(M.MestradoError, match="scratch"):
        st.score_pairs_noalloc(big, big)
    with pytest.raises(M.MestradoError, match="out of range"):
        st.score_pairs_noalloc([0, 2], [1, 1])
    assert (st.score_pairs_noalloc([0, 1], [1, 1]) == st.score_pairs([0, 1], [1, 1])).all()
    st.free()

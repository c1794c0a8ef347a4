"""Pins the NW oracle (oracle/nw.cpp) before it is trusted as the checker.

Independent checks: a textbook three-matrix Gotoh over prefix lengths, brute
force over every alignment of short sequences rescored under the cost model
stated in words (orc_rescore), closed forms, path invariants, and hand-worked
tie cases for the traceback priorities of NeedlemanWunsch.cu:330-336,430-438.
"""
import random

import numpy as np
import pytest

import oracle_bind as O

S = O.blosum62()


def test_blosum_fixture_is_blosum62():
    # spot values of the published BLOSUM62 (letter order ARNDCQEGHILKMFPSTWYVBZX*)
    L = O.LETTERS
    assert S.shape == (24, 24)
    assert (S == S.T).all()
    assert S[L.index("W"), L.index("W")] == 11
    assert S[L.index("C"), L.index("C")] == 9
    assert S[L.index("A"), L.index("A")] == 4
    assert S[L.index("W"), L.index("D")] == -4
    assert S[L.index("*"), L.index("*")] == 1


def test_encode_matches_reference_table():
    # AlignmentScore.cu:10-27 maps '*'=42 -> 23, 'A'=65 -> 0, 'B' -> 20, 'C' -> 4,
    # 'D' -> 3, 'E' -> 6, 'F' -> 13, 'G' -> 7, 'H' -> 8, 'I' -> 9, 'K' -> 11,
    # 'L' -> 10, 'M' -> 12, 'N' -> 2, 'P' -> 14, 'Q' -> 5, 'R' -> 1, 'S' -> 15,
    # 'T' -> 16, 'V' -> 19, 'W' -> 17, 'X' -> 22, 'Y' -> 18, 'Z' -> 21.
    table = {"*": 23, "A": 0, "B": 20, "C": 4, "D": 3, "E": 6, "F": 13, "G": 7, "H": 8,
             "I": 9, "K": 11, "L": 10, "M": 12, "N": 2, "P": 14, "Q": 5, "R": 1, "S": 15,
             "T": 16, "V": 19, "W": 17, "X": 22, "Y": 18, "Z": 21}
    s = "".join(table)
    codes = O.encode_rev(s)
    assert list(codes) == [table[c] for c in reversed(s)]
    # build rule D8: J, O, U and unknown -> X; lower case folds to upper case
    assert list(O.encode_rev("JOU#")) == [22, 22, 22, 22]
    assert list(O.encode_rev("acd")) == list(O.encode_rev("ACD"))


@pytest.mark.parametrize("seed", range(6))
def test_score_equals_gotoh_and_bruteforce(seed):
    rnd = random.Random(seed)
    alpha = O.STANDARD20[: 3 + seed]  # small alphabets force many ties
    for _ in range(60):
        n, m = rnd.randint(1, 6), rnd.randint(1, 6)
        x = "".join(rnd.choice(alpha) for _ in range(n))
        y = "".join(rnd.choice(alpha) for _ in range(m))
        s = O.nw_score(x, y, S)
        assert s == O.gotoh(x, y, S)
        assert s == O.bruteforce(x, y, S)
        sc, gx, gy, _ = O.nw_align(x, y, S)
        assert sc == s
        assert gx.replace("-", "") == x and gy.replace("-", "") == y
        assert len(gx) == len(gy)
        assert not any(a == "-" and b == "-" for a, b in zip(gx, gy))
        assert O.rescore(gx, gy, S) == s  # the traceback path is optimal


def test_score_gotoh_longer_random():
    rng = np.random.default_rng(7)
    for _ in range(30):
        n, m = rng.integers(1, 400, 2)
        x = "".join(rng.choice(list(O.STANDARD20), n))
        y = "".join(rng.choice(list(O.STANDARD20), m))
        s = O.nw_score(x, y, S)
        assert s == O.gotoh(x, y, S)
        sc, gx, gy, _ = O.nw_align(x, y, S)
        assert sc == s and O.rescore(gx, gy, S) == s


def test_closed_forms():
    L = O.LETTERS
    # 1x1: a substitution beats a gap pair (-2 + -10 or -2 + -2 + ... )
    for a in O.STANDARD20:
        for b in O.STANDARD20:
            assert O.nw_score(a, b, S) == max(S[L.index(b), L.index(a)], -2 - 10)
    # identical sequences of standard residues: the gapless path
    rng = np.random.default_rng(3)
    for n in (1, 5, 50, 300):
        x = "".join(rng.choice(list(O.STANDARD20), n))
        assert O.nw_score(x, x, S) == sum(int(S[L.index(c), L.index(c)]) for c in x)
    # symmetry score(x,y) == score(y,x) for the symmetric matrix
    for _ in range(20):
        x = "".join(rng.choice(list(O.STANDARD20), rng.integers(1, 80)))
        y = "".join(rng.choice(list(O.STANDARD20), rng.integers(1, 80)))
        assert O.nw_score(x, y, S) == O.nw_score(y, x, S)
    # one sequence against a single residue it does not contain: every other
    # residue of the long sequence is a gap column
    assert O.nw_score("W", "AAAA", S) == max(
        # W aligned to one A: best placement, gaps around it
        -3 + (-2 * 3),  # leading boundary run of 3 (DP start is the end of the originals)
        -3 - 10 - 2 * 2,
        -2 * 4 - 10,    # W unaligned: boundary run of the 4 A, then W opens a gap
        -2 * 1 - 10 - 2 * 3,
    )


def test_boundary_runs_cost_ge_only_at_dp_start():
    # DP order is the reverse of the forward strings: gaps at the END of the
    # forward sequences are the cheap boundary run; gaps at the start open.
    L = O.LETTERS
    x, y = "WAAAA", "W"
    s = O.nw_score(x, y, S)
    # W/W then 4 trailing A vs gap, lying on the DP boundary: 11 - 2*4
    assert s == 11 - 8
    x, y = "AAAAW", "W"
    # W/W at the end; leading AAAA opens a gap: 11 - 10 - 2*3
    assert O.nw_score(x, y, S) == 11 - 10 - 6
    sc, gx, gy, _ = O.nw_align("AAAAW", "W", S)
    assert (gx, gy) == ("AAAAW", "----W")


def test_traceback_tie_priorities_hand_cases():
    # Each pair has exactly two (three) optimal alignments, found by brute force
    # and worked by hand; the traceback starts at the forward start of both
    # sequences and must take the move of NeedlemanWunsch.cu:430-438: diagonal
    # before up (gap in X) before left (gap in Y).
    # G vs WG: [G/W][-/G] = -2 (boundary) - 2 and [-/W][G/G] = 6 - 10, both -4;
    # the first forward column is diagonal vs up -> diagonal.
    assert O.nw_score("G", "WG", S) == -4
    assert O.nw_align("G", "WG", S)[1:3] == ("G-", "WG")
    # WG vs G: diagonal vs left -> diagonal.
    assert O.nw_align("WG", "G", S)[1:3] == ("WG", "G-")
    # AAW vs AW: [A/A][A/-][W/W] and [A/-][A/A][W/W], both 4 - 10 + 11 = 5.
    assert O.nw_score("AAW", "AW", S) == 5
    assert O.nw_align("AAW", "AW", S)[1:3] == ("AAW", "A-W")
    # WCG vs CW: [W/-][C/C][G/W] = -10 + 9 - 2 and [-/C][W/W][C/-][G/-] =
    # -10 + 11 - 4 (boundary run), both -3; first column left vs up -> up.
    assert O.nw_score("WCG", "CW", S) == -3
    assert O.nw_align("WCG", "CW", S)[1:3] == ("-WCG", "CW--")


# Continuation ties: extending a gap and opening one give the same value, so
# F(r-1,c) - ge == H(r-1,c) - go (or the E analogue). The reference sets the
# continuation bit on that tie (F_up = (F == F_prev - ge), NeedlemanWunsch.cu:
# 315-316; the E bit at :336), and the traceback then keeps the gap going
# (:430-433). Each case below has two optimal alignments (both rescore to the
# brute-force maximum); "open on tie" would give the second string.
CONTINUATION_TIES = [
    # W vs AWW, worked by hand. DP on the reversed inputs: X = [W], Y = [W, W, A].
    # r=0: H = 11 (diagonal). r=1: F = max(-12 - 2, 11 - 10) = 1, H = max(1, -2 + 11) = 9
    # (diagonal). r=2 (A): F = max(F(1) - 2, H(1) - 10) = max(-1, -1) -> tie, contUp set;
    # H = max(F = -1, E = -16, -4 + S[A][W] = -7) = -1 (up). Traceback from r=2: up, then
    # prev was up with contUp -> up again (not r=1's diagonal), then the diagonal W/W.
    #   extend: --W / AWW = -(10 + 2) + 11 = -1;  open: -W- / AWW = -10 + 11 - 2 = -1.
    ("W", "AWW", ("--W", "AWW"), ("-W-", "AWW"), -1),
    # the E analogue (X and Y swapped): left moves instead of up moves
    ("AWW", "W", ("AWW", "--W"), ("AWW", "-W-"), -1),
    # interior ties, both alternatives found by brute force over every alignment
    ("WAGC", "GACGWC", ("WA--GC", "GACGWC"), ("WA-G-C", "GACGWC"), -3),
    ("CGWCWW", "GCWW", ("CGWCWW", "--GCWW"), ("CGWCWW", "-G-CWW"), 17),
    ("CCGW", "WACCW", ("--CCGW", "WACC-W"), ("-CCGW", "WACCW"), 7),
]


@pytest.mark.parametrize("x,y,ext,opn,score", CONTINUATION_TIES)
def test_traceback_continuation_ties_extend(x, y, ext, opn, score):
    assert O.nw_score(x, y, S) == score
    assert O.bruteforce(x, y, S) == score
    # both alternatives are optimal alignments of x and y ...
    for gx, gy in (ext, opn):
        assert gx.replace("-", "") == x and gy.replace("-", "") == y
        assert O.rescore(gx, gy, S) == score
    # ... and the oracle takes the one that extends on the tie
    assert O.nw_align(x, y, S)[1:3] == ext


def test_nibbles_of_reference_layout():
    # First cell of a 1x1 alignment of equal residues: H == diag only, so the
    # direction is CROSSWISE (12) and the continuation bits compare against -inf.
    sc, gx, gy, nib = O.nw_align("A", "A", S)
    assert nib.shape == (1, 1) and nib[0, 0] == 12
    # A/W: H = diag = -3 beats both gaps (-12), crosswise
    sc, gx, gy, nib = O.nw_align("A", "W", S)
    assert nib[0, 0] == 12 and sc == -3


def test_oracle_all_pairs_matches_single_calls():
    rng = np.random.default_rng(11)
    seqs = ["".join(rng.choice(list(O.STANDARD20), rng.integers(1, 60))) for _ in range(12)]
    mat, _ = O.all_pairs(seqs, S, threads=2)
    for i in range(12):
        for j in range(12):
            assert mat[i, j] == O.nw_score(seqs[i], seqs[j], S)

"""Pins the profile-profile oracle: on one-hot profiles (single sequences) it
must reduce exactly to the pairwise recurrence and traceback, the merged rows
must keep every sequence, and tiny profiles are checked by brute force."""
import itertools

import numpy as np

import oracle_bind as O

S = O.blosum62()


def test_one_hot_profiles_reduce_to_pairwise():
    rng = np.random.default_rng(0)
    for _ in range(40):
        x = "".join(rng.choice(list("AWCGSTK"), rng.integers(1, 30)))
        y = "".join(rng.choice(list("AWCGSTK"), rng.integers(1, 30)))
        sc, ops = O.profile_align([x], [y], S)
        psc, gx, gy, _ = O.nw_align(x, y, S)
        assert sc == psc
        rx = "".join("-" if o == "U" else "*" for o in ops)
        assert len(ops) == len(gx)
        assert all((o == "U") == (a == "-") and (o == "L") == (b == "-") for o, a, b in zip(ops, gx, gy))


def _profile_score_bruteforce(rx, ry, go, ge):
    # every path through the two profiles, scored with the rounded column scores
    L = O.LETTERS
    n, m = len(rx[0]), len(ry[0])

    def col(rows, t):
        return [r[t] for r in rows]

    def cs(c, r):
        fx = np.zeros(24, np.float32)
        fy = np.zeros(24, np.float32)
        for ch in col(rx, c):
            if ch != "-":
                fx[L.index(ch)] += 1
        for ch in col(ry, r):
            if ch != "-":
                fy[L.index(ch)] += 1
        fx = (fx / np.float32(len(rx))).astype(np.float32)
        fy = (fy / np.float32(len(ry))).astype(np.float32)
        # same order of fp32 operations as the oracle
        P = np.zeros(24, np.float32)
        for a in range(24):
            acc = np.float32(0)
            for b in range(24):
                acc = np.float32(acc + np.float32(fy[b] * np.float32(S[b, a])))
            P[a] = acc
        acc = np.float32(0)
        for a in range(24):
            acc = np.float32(acc + np.float32(fx[a] * P[a]))
        return int(np.rint(acc))

    best = None
    # DP order is reversed columns; enumerate op strings in DP order
    def rec(i, j, seq):
        nonlocal best
        if i == n and j == m:
            s, prev, first = 0, None, True
            for op, (ci, rj) in seq:
                if prev is not None and op != prev:
                    first = False
                if op == "D":
                    s += cs(n - 1 - ci, m - 1 - rj)
                elif op == prev:
                    s -= ge
                else:
                    s -= ge if first else go
                prev = op
            best = s if best is None else max(best, s)
            return
        if i < n and j < m:
            rec(i + 1, j + 1, seq + [("D", (i, j))])
        if i < n:
            rec(i + 1, j, seq + [("L", (i, j))])
        if j < m:
            rec(i, j + 1, seq + [("U", (i, j))])

    rec(0, 0, [])
    return best


def test_tiny_profiles_bruteforce():
    rng = np.random.default_rng(4)
    for _ in range(12):
        n, m = rng.integers(1, 4, 2)
        kx, ky = rng.integers(1, 4, 2)
        rx = ["".join(rng.choice(list("AWC-"), n)) for _ in range(kx)]
        ry = ["".join(rng.choice(list("AWC-"), m)) for _ in range(ky)]
        sc, ops = O.profile_align(rx, ry, S)
        assert sc == _profile_score_bruteforce(rx, ry, 10, 2)


def test_progressive_msa_keeps_sequences():
    rng = np.random.default_rng(2)
    seqs = ["".join(rng.choice(list(O.STANDARD20), rng.integers(5, 60))) for _ in range(12)]
    D, _ = O.all_pairs(seqs, S)
    joins, _ = O.nj_fullscan(D.astype(np.float32))
    rows = O.progressive_msa(seqs, joins, S)
    assert len(set(len(r) for r in rows)) == 1
    for s, r in zip(seqs, rows):
        assert r.replace("-", "") == s
    assert not any(all(r[t] == "-" for r in rows) for t in range(len(rows[0])))


def test_c_restatement_equals_python():
    # the C++ progressive alignment (used for large sets) equals the Python
    # restatement merge for merge: empty sequences, ties, unknown letters,
    # and profiles longer than its 64-row blocks of column scores
    rng = np.random.default_rng(5)
    cases = [
        ["".join(rng.choice(list("AWC"), int(n))) for n in rng.integers(0, 25, 14)],
        ["".join(rng.choice(list(O.STANDARD20), int(n))) for n in rng.integers(50, 200, 9)],
        ["MKJQ", "mkoq", "", "AUWX*", "W"],
        ["".join(rng.choice(list(O.STANDARD20), 150))] * 3 + ["A"],
    ]
    for seqs in cases:
        D, _ = O.all_pairs(seqs, S)
        joins, _ = O.nj_fullscan(D.astype(np.float32))
        assert O.progressive_msa_c(seqs, joins, S) == O.progressive_msa(seqs, joins, S)

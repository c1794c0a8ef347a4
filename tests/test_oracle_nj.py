"""Pins the NJ oracle (oracle/nj.cpp) before it is trusted as the checker.

Checks: recovery of random additive trees (NJ is exact on additive metrics),
a hand-worked four-taxon case, agreement of the reference restatement (bins,
heaps, candidate list, q prediction) with a textbook full-scan NJ, and the
tree text against the reference's own TreeNode.cpp compiled from source.
"""
import numpy as np
import pytest

import oracle_bind as O


def _rand_tree(rng, n):
    nodes = list(range(n))
    parent, blen = {}, {}
    nxt = n
    while len(nodes) > 1:
        a, b = rng.choice(len(nodes), 2, replace=False)
        u, v = nodes[a], nodes[b]
        for w in (u, v):
            parent[w] = nxt
            blen[w] = int(rng.integers(1, 20))
        nodes = [x for k, x in enumerate(nodes) if k not in (a, b)] + [nxt]
        nxt += 1
    return parent, blen


def _additive(parent, blen, n):
    anc = []
    for i in range(n):
        d = {i: 0}
        x, acc = i, 0
        while x in parent:
            acc += blen[x]
            x = parent[x]
            d[x] = acc
        anc.append(d)
    D = np.zeros((n, n), np.float32)
    for i in range(n):
        for j in range(n):
            D[i, j] = min(anc[i][a] + anc[j][a] for a in anc[i] if a in anc[j])
    return D


def _canon(s, n):
    comp = frozenset(range(n)) - s
    return min(s, comp, key=lambda z: sorted(z))


def splits_of_joins(joins, n):
    members = {i: frozenset([i]) for i in range(n)}
    out = set()
    for J in joins:
        members[J.node] = members[J.left] | members[J.right]
        s = members[J.node]
        if 1 < len(s) < n - 1:
            out.add(_canon(s, n))
    return out


def splits_of_tree(parent, n):
    ch = {}
    for c, p in parent.items():
        ch.setdefault(p, []).append(c)
    memo = {}

    def leaves(x):
        if x < n:
            return frozenset([x])
        if x not in memo:
            memo[x] = frozenset().union(*[leaves(c) for c in ch[x]])
        return memo[x]

    out = set()
    for x in ch:
        s = leaves(x)
        if 1 < len(s) < n - 1:
            out.add(_canon(s, n))
    return out


@pytest.mark.parametrize("n", [4, 5, 9, 30, 120, 260])
def test_additive_tree_recovered(n):
    rng = np.random.default_rng(n)
    parent, blen = _rand_tree(rng, n)
    D = _additive(parent, blen, n)
    joins, _, trig = O.nj_reference(D)
    assert splits_of_joins(joins, n) == splits_of_tree(parent, n)
    assert trig == 0
    # branch lengths of an additive tree are recovered exactly (dyadic values)
    members = {i: frozenset([i]) for i in range(n)}
    for J in joins:
        members[J.node] = members[J.left] | members[J.right]


def test_four_taxon_hand_case():
    # D (additive): ((0:1,1:2):4,(2:3,3:5)) -> D01=3, D23=8, D02=8, D03=10,
    # D12=9, D13=11. T = row sums: T0=21, T1=23, T2=25, T3=29 (diagonal 0).
    # n=4: q_ij = 2*D_ij - T_i - T_j: q01 = 6-44 = -38, q23 = 16-54 = -38,
    # q02 = 16-46 = -30, q03 = 20-50 = -30, q12 = 18-48 = -30, q13 = 22-52 = -30.
    # Tie at -38 between (0,1) and (2,3): the larger max(i,j) wins -> (2,3).
    D = np.array([[0, 3, 8, 10], [3, 0, 9, 11], [8, 9, 0, 8], [10, 11, 8, 0]], np.float32)
    joins, _, _ = O.nj_reference(D)
    J = joins[0]
    assert (J.left, J.right, J.node) == (2, 3, 4)
    assert J.q == -38.0
    # branch lengths 0.5*(8 + (25-29)/2) = 3 and 5
    assert (J.lenL, J.lenR) == (3.0, 5.0)
    # new row: ((D2k + D3k) - 8)/2 -> D(4,0) = (8+10-8)/2 = 5, D(4,1) = 6;
    # T0 = 21 + 5 - 18 = 8, T1 = 23 + 6 - 20 = 9, T4 = 11; n=3: q01 = 3 - 17 = -14,
    # q04 = 5 - 19 = -14, q14 = 6 - 20 = -14; all tie: larger J (node 4), then
    # larger I -> (1,4).
    J = joins[1]
    assert (J.left, J.right) == (1, 4)
    assert J.q == -14.0
    J = joins[2]
    assert (J.left, J.right, J.n) == (0, 5, 2)
    assert J.lenL == J.lenR == J.d / 2


@pytest.mark.parametrize("n,seed", [(30, 0), (200, 1), (450, 2), (700, 3)])
def test_reference_restatement_equals_fullscan(n, seed):
    # Random symmetric integer "similarity" matrices like the score matrix the
    # pipeline feeds NJ (self scores on the diagonal).
    rng = np.random.default_rng(seed)
    A = rng.integers(-300, 1500, (n, n))
    A = np.triu(A) + np.triu(A, 1).T
    D = A.astype(np.float32)
    j1, rb1, trig = O.nj_reference(D)
    j2, rb2 = O.nj_fullscan(D)
    assert trig == 0 and rb1 == rb2
    for a, b in zip(j1, j2):
        assert (a.node, a.left, a.right, a.n) == (b.node, b.left, b.right, b.n)
        assert (a.d, a.Ti, a.Tj, a.q) == (b.d, b.Ti, b.Tj, b.q)
        assert (a.lenL, a.lenR) == (b.lenL, b.lenR)


@pytest.mark.skipif(not O.have_ref_tool(), reason="reference TreeNode not built (oracle/_ref)")
@pytest.mark.parametrize("n", [1, 2, 3, 17, 150])
def test_tree_text_matches_reference_treenode(n):
    rng = np.random.default_rng(100 + n)
    names = [">gi|%d|seq %d" % (1000 + i, i) for i in range(n)]
    if n == 1:
        joins = []
    else:
        A = rng.integers(-300, 3000, (n, n))
        A = np.triu(A) + np.triu(A, 1).T
        joins, _, _ = O.nj_reference(A.astype(np.float32))
    assert O.tree_text(joins, names) == O.ref_tree_text(joins, names)


@pytest.mark.skipif(not O.have_ref_tool(), reason="reference TreeNode not built (oracle/_ref)")
def test_branch_formula_matches_reference_join():
    # The reference's TreeNode::join computes the lengths itself from (d, Ti, Tj,
    # n); feeding our joins through it must print the same lengths, including
    # the clamp of a negative side to zero (TreeNode.cpp:38-52).
    names = ["a", "b", "c"]


    j0 = O.JoinT(node=3, left=0, right=1, d=1.0, Ti=0.0, Tj=50.0, n=3, q=0.0)
    L = 0.5 * (1.0 + (0.0 - 50.0) / 1)
    R = 0.5 * (1.0 + (50.0 - 0.0) / 1)
    j0.lenL, j0.lenR = 0.0, R - L  # clamp: the negative side moves to the other
    j1 = O.JoinT(node=4, left=2, right=3, d=7.0, Ti=1.0, Tj=2.0, n=2, q=0.0)
    j1.lenL = j1.lenR = 3.5
    assert O.tree_text([j0, j1], names) == O.ref_tree_text([j0, j1], names)

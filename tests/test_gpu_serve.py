"""Served form (continuous batching, mst_server_*): every job's matrix is the
oracle's, alone and in a mix; jobs with a tree and an alignment give the
one-job path's (and the oracle's) tree and rows, alone and in a mix;
admission under a small budget, rejection, cancellation and misuse."""
import numpy as np
import pytest

import mestrado_amd as M
import oracle_bind as O

pytestmark = pytest.mark.gpu
S = O.blosum62()


def _job(rng, n, lo, hi):
    return ["".join(rng.choice(list(O.STANDARD20), int(L))) for L in rng.integers(lo, hi + 1, n)]


def _jobs(seed):
    rng = np.random.default_rng(seed)
    spec = [(2, 50, 90), (7, 100, 140), (11, 300, 700), (5, 900, 1300), (3, 1500, 2000),
            (1, 400, 400), (16, 20, 260), (4, 2150, 2400), (9, 60, 1100)]
    return [_job(rng, n, lo, hi) for n, lo, hi in spec]


@pytest.mark.parametrize("k8", [0, 1])
def test_jobs_alone_and_in_mix_match_oracle(ctx, k8):
    # both routes of the rows above 64 k1max: two-wave and multi-sweep
    # classes, or the strips-in-turn kernel
    jobs = _jobs(1)
    want = [O.all_pairs(j, S)[0] for j in jobs]
    with M.Server(ctx, 1 << 30) as srv:
        srv.set("k8", k8)
        for j, w in zip(jobs, want):  # one at a time
            jid = srv.submit(j)
            srv.drain()
            assert srv.job(jid)["state"] == M.JOB_DONE
            assert (srv.result(jid) == w).all()
        srv.set("round_cells", 2e6)  # small rounds: jobs share and split rounds
        ids = []
        for k, j in enumerate(jobs):  # arrivals interleaved with scheduling
            ids.append(srv.submit(j))
            if k % 2:
                srv.step()
        srv.drain()
        for jid, w in zip(ids, want):
            info = srv.job(jid)
            assert info["state"] == M.JOB_DONE
            assert info["t_submit"] <= info["t_admit"] <= info["t_done"]
            assert (srv.result(jid) == w).all()
        st = srv.stats()
        assert st["waiting"] == 0 and st["running"] == 0 and st["arena_used"] == 0


@pytest.mark.parametrize("k8,s6p2", [(1, 0), (0, 1), (0, 0)])
def test_long_rows_multi_sweep_forms(ctx, k8, s6p2):
    # rows above 64 k1max: on the strips-in-turn kernel (score8, the default;
    # each class with its own row buffers and stream), or on the two-wave and
    # multi-sweep classes (3 and 4 rows per lane above 1,024) with the
    # two-pipeline or the one-pipeline kernel; rounds small enough that
    # several are in flight
    rng = np.random.default_rng(9)
    jobs = [_job(rng, n, lo, hi) for n, lo, hi in [(6, 1030, 1500), (5, 1540, 2300), (7, 900, 2000)]]
    want = [O.all_pairs(j, S)[0] for j in jobs]
    with M.Server(ctx, 1 << 30) as srv:
        srv.set("k8", k8)
        srv.set("s6p2", s6p2)
        srv.set("round_cells", 5e6)
        ids = [srv.submit(j) for j in jobs]
        srv.drain()
        for jid, w in zip(ids, want):
            assert srv.job(jid)["state"] == M.JOB_DONE
            assert (srv.result(jid) == w).all()


def test_small_budget_queues_and_rejects(ctx):
    rng = np.random.default_rng(2)
    jobs = [_job(rng, 400, 10, 30) for _ in range(3)]  # 400^2 * 4 B = 625 KiB of scores each
    want = [O.all_pairs(j, S)[0] for j in jobs]
    with M.Server(ctx, 1 << 20) as srv:  # room for one of these jobs at a time
        big = srv.submit(_job(rng, 600, 10, 20))  # 600^2 * 4 B > 1 MiB
        assert srv.job(big)["state"] == M.JOB_REJECTED
        ids = [srv.submit(j) for j in jobs]
        srv.step()
        st = srv.stats()
        assert st["waiting"] >= 1  # the budget holds back the queue
        done = {}
        for _ in range(100000):
            srv.step()
            for jid, w in zip(ids, want):
                if jid not in done and srv.job(jid)["state"] == M.JOB_DONE:
                    done[jid] = srv.result(jid)  # frees the job's memory
            if len(done) == len(ids):
                break
        assert len(done) == len(ids)
        for jid, w in zip(ids, want):
            assert (done[jid] == w).all()
        with pytest.raises(M.MestradoError):
            srv.result(big)


def test_cancel_and_misuse(ctx):
    rng = np.random.default_rng(3)
    jobs = [_job(rng, 20, 200, 400) for _ in range(3)]
    with M.Server(ctx, 1 << 30) as srv:
        a = srv.submit(jobs[0])
        b = srv.submit(jobs[1])
        c = srv.submit(jobs[2])
        srv.cancel(b)  # still waiting
        assert srv.job(b)["state"] == M.JOB_CANCELLED
        with pytest.raises(M.MestradoError):
            srv.result(a)  # not done yet
        srv.drain()
        assert srv.job(a)["state"] == M.JOB_DONE
        assert srv.job(c)["state"] == M.JOB_DONE
        assert (srv.result(c) == O.all_pairs(jobs[2], S)[0]).all()
        srv.release(a)
        with pytest.raises(M.MestradoError):
            srv.job(a)  # forgotten after its result was taken
        with pytest.raises(M.MestradoError):
            srv.set("no-such-key", 1)
        with pytest.raises(M.MestradoError):
            srv.set("est_floor", -1)  # a negative floor on the estimate
        srv.set("est_floor", 0.05)


JOIN_FIELDS = ("node", "left", "right", "d", "Ti", "Tj", "n", "q", "len_left", "len_right")


def _tree_tuple(joins):
    return [tuple(getattr(j, f) for f in JOIN_FIELDS) for j in joins]


def _offline(ctx, seqs):
    """Scores, tree and alignment of the one-job path (all_pairs, nj, msa)."""
    st = ctx.upload(seqs)
    D = st.all_pairs(M.FORM_BATCHED)
    joins = st.nj() if len(seqs) > 1 else []
    rows = st.msa(joins)
    st.free()
    return D, joins, rows


def _msa_jobs(seed):
    rng = np.random.default_rng(seed)
    spec = [(1, 50, 80), (2, 30, 60), (3, 10, 40), (12, 100, 300), (40, 50, 200), (7, 500, 900),
            (6, 0, 3), (9, 600, 1300), (25, 20, 120)]
    return [_job(rng, n, lo, hi) for n, lo, hi in spec]


def _check_job(srv, jid, seqs, want, what):
    D, joins, rows = want
    assert srv.job(jid)["state"] == M.JOB_DONE
    if what == M.SERVE_MSA:
        assert srv.msa(jid) == rows
    if what in (M.SERVE_TREE, M.SERVE_MSA) and len(seqs) > 1:
        assert _tree_tuple(srv.tree(jid)) == _tree_tuple(joins)
    assert (srv.result(jid) == D).all()


@pytest.mark.parametrize("deadlines", [(), (("slack", 3.0), ("est_floor", 0.1))])
def test_msa_jobs_alone_and_in_mix(ctx, deadlines):
    """Scores, NJ tree and alignment of every job equal the one-job path's,
    alone and with jobs of every stage sharing rounds, NJ launches and merge
    batches; the deadline settings change only the order jobs are served in."""
    jobs = _msa_jobs(5)
    want = [_offline(ctx, j) for j in jobs]
    with M.Server(ctx, 1 << 30) as srv:
        for k, v in deadlines:
            srv.set(k, v)
        for j, w in zip(jobs, want):
            jid = srv.submit(j, M.SERVE_MSA)
            srv.drain()
            _check_job(srv, jid, j, w, M.SERVE_MSA)
        srv.set("round_cells", 2e6)
        srv.set("merge_bytes", 1 << 20)  # small merge batches: levels of many jobs interleave
        ids, whats = [], []
        for k, j in enumerate(jobs + jobs[::-1]):
            what = (M.SERVE_MSA, M.SERVE_TREE, M.SERVE_SCORES)[k % 3]
            ids.append(srv.submit(j, what))
            whats.append(what)
            if k % 2:
                srv.step()
        srv.drain()
        for jid, j, w, what in zip(ids, jobs + jobs[::-1], want + want[::-1], whats):
            # stage accounting: a job scored in at least one round after its
            # admission; an alignment job (one sequence too: a batch writes
            # its leaf row) spends its stage with one of its merge batches in
            # flight or none
            tt, js = srv.times(jid), srv.job_stats(jid)
            if any(len(x) for x in j):
                assert js["rounds"] >= 1 and 0 <= js["admit_s"] <= js["first_round_s"] <= tt[2] - tt[0]
            if what == M.SERVE_MSA:
                assert js["batches"] >= 1
                start = tt[3] or tt[2]  # no tree for one sequence: from the scores
                assert abs(js["busy_s"] + js["idle_s"] - (tt[4] - start)) < 1e-3
            else:
                assert js["batches"] == 0
            _check_job(srv, jid, j, w, what)
        st = srv.stats()
        assert st["waiting"] == 0 and st["running"] == 0 and st["arena_used"] == 0


def test_msa_job_matches_oracle(ctx):
    """A served alignment is the oracle's progressive alignment along the
    served tree, and the served tree is the oracle's NJ of the oracle's
    scores."""
    rng = np.random.default_rng(9)
    seqs = _job(rng, 14, 40, 160)
    D, _ = O.all_pairs(seqs, S)
    ref_joins, _ = O.nj_fullscan(D.astype(np.float32))
    with M.Server(ctx, 1 << 28) as srv:
        jid = srv.submit(seqs, M.SERVE_MSA)
        srv.drain()
        joins = srv.tree(jid)
        assert [(j.node, j.left, j.right) for j in joins] == [(j.node, j.left, j.right) for j in ref_joins]
        assert srv.msa(jid) == O.progressive_msa(seqs, joins, S)
        assert (srv.result(jid) == D).all()


def test_msa_job_misuse(ctx):
    rng = np.random.default_rng(4)
    with M.Server(ctx, 1 << 26) as srv:
        jid = srv.submit(_job(rng, 5, 20, 40), M.SERVE_SCORES)
        with pytest.raises(M.MestradoError):
            srv.tree(jid)  # not done
        srv.drain()
        with pytest.raises(M.MestradoError):
            srv.msa(jid)  # submitted without an alignment
        with pytest.raises(M.MestradoError):
            srv.submit(_job(rng, 3, 5, 9), 5)  # not a valid `what`
        srv.release(jid)
        a = srv.submit(_job(rng, 30, 200, 400), M.SERVE_MSA)
        srv.step()
        srv.cancel(a)
        srv.drain()
        assert srv.job(a)["state"] == M.JOB_CANCELLED
        assert srv.stats()["arena_used"] == 0


def test_large_msa_jobs_match_oracle(ctx):
    """Jobs of 210-512 sequences (lengths 100-300) in one mix: their NJ runs
    past the first T rebuild (after max(n/2, 200) joins) in the served
    one-workgroup-per-job NJ launch. Scores, trees and rows against the
    oracle: the full-scan NJ bit-exact, the reference restatement's topology
    (its queue-size trigger not firing), the progressive alignment's rows."""
    rng = np.random.default_rng(21)
    jobs = [_job(rng, n, 100, 300) for n in (210, 347, 512)]
    with M.Server(ctx, 4 << 30) as srv:
        ids = [srv.submit(j, M.SERVE_MSA) for j in jobs]
        srv.drain()
        for jid, seqs in zip(ids, jobs):
            assert srv.job(jid)["state"] == M.JOB_DONE
            D, _ = O.all_pairs(seqs, S)
            Df = D.astype(np.float32)
            full, rebuilds = O.nj_fullscan(Df)
            assert rebuilds >= 1
            ref, _, trig = O.nj_reference(Df)
            assert trig == 0
            joins = srv.tree(jid)
            for a, o, r in zip(joins, full, ref):
                assert (a.node, a.left, a.right, a.n) == (o.node, o.left, o.right, o.n)
                assert (a.d, a.Ti, a.Tj, a.q, a.len_left, a.len_right) == \
                    (o.d, o.Ti, o.Tj, o.q, o.lenL, o.lenR)
                assert (a.left, a.right) == (r.left, r.right)
            assert srv.msa(jid) == O.progressive_msa_c(seqs, joins, S)
            assert (srv.result(jid) == D).all()

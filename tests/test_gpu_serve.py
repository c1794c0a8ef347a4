"""Served form (continuous batching, mst_server_*): every job's matrix is the
oracle's, alone and in a mix; admission under a small budget, rejection,
cancellation and misuse."""
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


def test_jobs_alone_and_in_mix_match_oracle(ctx):
    jobs = _jobs(1)
    want = [O.all_pairs(j, S)[0] for j in jobs]
    with M.Server(ctx, 1 << 30) as srv:
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

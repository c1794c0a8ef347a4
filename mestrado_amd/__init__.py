"""mestrado_amd — MI355X (gfx950) all-pairs Needleman-Wunsch, alignment with
traceback and neighbour joining, behind the C ABI of include/mestrado.h.

The native library (libmestrado.so, built in-tree by mestrado_amd.build) is the
only compute path: importing this package on a machine where the library is
missing raises, and there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmestrado.so")

FORM_PLAIN = 0
FORM_BATCHED = 1


class MestradoError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MestradoError("libmestrado.so is not built (run python -m mestrado_amd.build)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.mst_create.argtypes = [C.c_int, C.c_int, P, C.c_int, C.POINTER(P)]
        L.mst_destroy.argtypes = [P]
        L.mst_last_error.argtypes = [P]
        L.mst_last_error.restype = C.c_char_p
        L.mst_device_info.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.mst_device_arch.argtypes = [P, C.c_char_p, C.c_int]
        L.mst_seqset_upload.argtypes = [P, P, P, C.c_int, C.POINTER(C.c_int)]
        L.mst_seqset_free.argtypes = [P, C.c_int]
        L.mst_seqset_codes.argtypes = [P, C.c_int, P, C.c_int64]
        L.mst_all_pairs.argtypes = [P, C.c_int, C.c_int, P, P]
        L.mst_score_pairs.argtypes = [P, C.c_int, P, P, C.c_int64, C.c_int, P]
        L.mst_score_pairs_noalloc.argtypes = [P, C.c_int, P, P, C.c_int64, P]
        L.mst_all_pairs_resident.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_float)]
        L.mst_all_pairs_launches.argtypes = [P, C.c_int, C.c_int]
        L.mst_last_launch_stats.argtypes = [P, P, P, P, P, C.c_int, C.POINTER(C.c_int)]
        L.mst_valu_peak.argtypes = [P, C.c_int, C.POINTER(C.c_double)]
        L.mst_copy_probe.argtypes = [P, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        L.mst_tune_set.argtypes = [P, C.c_char_p, C.c_int]
        L.mst_server_create.argtypes = [P, C.c_int64, C.POINTER(P)]
        L.mst_server_destroy.argtypes = [P]
        L.mst_server_submit.argtypes = [P, P, P, C.c_int, C.POINTER(C.c_int64)]
        L.mst_server_submit_job.argtypes = [P, P, P, C.c_int, C.c_int, C.POINTER(C.c_int64)]
        L.mst_server_tree.argtypes = [P, C.c_int64, P]
        L.mst_server_job_times.argtypes = [P, C.c_int64, P]
        L.mst_server_job_stats.argtypes = [P, C.c_int64, P]
        L.mst_server_msa.argtypes = [P, C.c_int64, P, C.c_int64, C.POINTER(C.c_int)]
        L.mst_server_step.argtypes = [P, C.POINTER(C.c_int)]
        L.mst_server_drain.argtypes = [P]
        L.mst_server_job.argtypes = [P, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.mst_server_result.argtypes = [P, C.c_int64, P]
        L.mst_server_cancel.argtypes = [P, C.c_int64]
        L.mst_server_completed.argtypes = [P, P, C.c_int, C.POINTER(C.c_int)]
        L.mst_server_stats.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_double)]
        L.mst_server_set.argtypes = [P, C.c_char_p, C.c_double]
        L.mst_tune_get.argtypes = [P, C.c_char_p, C.POINTER(C.c_int)]
        L.mst_align_pairs.argtypes = [P, C.c_int, P, P, C.c_int64, P, P, P, C.c_int64]
        L.mst_msa.argtypes = [P, C.c_int, P, C.POINTER(C.c_int)]
        L.mst_msa_rows.argtypes = [P, P, C.c_int64]
        L.mst_nj.argtypes = [P, P, C.c_int, C.c_float, P]
        L.mst_nj_scores.argtypes = [P, C.c_int, C.c_float, P]
        L.mst_synth.argtypes = [C.c_int, C.c_uint64, C.c_int, P, C.c_int64, P, C.c_int64,
                                C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.mst_synth_job_sizes.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def synth(config, seed=0, n=0):
    """Synthetic set of benchmark configuration `config` -> list of str."""
    L = lib()
    cnt = C.c_int(0)
    tot = C.c_int64(0)
    rc = -2
    if config == 4 and n > 0:  # lengths are at most 2000: one pass with room to spare
        buf = np.zeros(n * 2048, dtype=np.uint8)
        off = np.zeros(n + 1, dtype=np.int64)
        rc = L.mst_synth(config, seed, n, _p(buf), len(buf), _p(off), len(off), C.byref(cnt),
                         C.byref(tot))
    if rc == -2:
        if L.mst_synth(config, seed, n, None, 0, None, 0, C.byref(cnt), C.byref(tot)) != 0:
            raise MestradoError("bad synthetic configuration %r" % config)
        buf = np.zeros(max(1, tot.value), dtype=np.uint8)
        off = np.zeros(cnt.value + 1, dtype=np.int64)
        rc = L.mst_synth(config, seed, n, _p(buf), len(buf), _p(off), len(off), C.byref(cnt),
                         C.byref(tot))
    if rc != 0:
        raise MestradoError("mst_synth failed (%d)" % rc)
    raw = buf.tobytes()
    return [raw[off[i]:off[i + 1]].decode() for i in range(cnt.value)]


def job_sizes(seed, njobs, lo=8, hi=512, exponent=1.5):
    out = np.zeros(njobs, dtype=np.int32)
    lib().mst_synth_job_sizes(seed, njobs, lo, hi, exponent, _p(out))
    return out


def pack(seqs):
    """list of str -> (uint8 ASCII buffer, int64 offsets[n+1])."""
    off = np.zeros(len(seqs) + 1, dtype=np.int64)
    for i, s in enumerate(seqs):
        off[i + 1] = off[i] + len(s)
    buf = np.frombuffer("".join(seqs).encode(), dtype=np.uint8).copy()
    if buf.size == 0:
        buf = np.zeros(1, dtype=np.uint8)
    return buf, off


class Join(C.Structure):
    """mst_join (include/mestrado.h)."""
    _fields_ = [("node", C.c_int), ("left", C.c_int), ("right", C.c_int), ("d", C.c_float),
                ("Ti", C.c_double), ("Tj", C.c_double), ("n", C.c_int), ("q", C.c_double),
                ("len_left", C.c_double), ("len_right", C.c_double)]


class Context:
    """One device context (mst_create/mst_destroy)."""

    def __init__(self, gap_open=10, gap_ext=2, subst=None, device=0, tuned=True):
        """tuned: apply the persisted tuning table of the device
        (mestrado_amd/tuning/<arch>.json, written by mestrado_amd.autotune)."""
        L = lib()
        self._L = L
        h = C.c_void_p()
        s = None
        if subst is not None:
            s = np.ascontiguousarray(subst, dtype=np.int8).reshape(576)
            self._subst = s
        rc = L.mst_create(gap_open, gap_ext, _p(s), device, C.byref(h))
        if rc != 0:
            raise MestradoError("mst_create failed (%d)" % rc)
        self.h = h
        self.tuning = {}
        self.tuned = tuned
        if tuned:
            from . import autotune
            self.tuning = autotune.apply(self)

    def arch(self):
        buf = C.create_string_buffer(64)
        self._ck(self._L.mst_device_arch(self.h, buf, 64), "mst_device_arch")
        return buf.value.decode()

    def _ck(self, rc, what):
        if rc != 0:
            msg = self._L.mst_last_error(self.h)
            raise MestradoError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def close(self):
        if self.h:
            self._L.mst_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        cus = C.c_int(0)
        hbm = C.c_int64(0)
        self._ck(self._L.mst_device_info(self.h, C.byref(cus), C.byref(hbm)), "mst_device_info")
        return cus.value, hbm.value

    def tune_set(self, key, value):
        """Set a tunable of the batched form (drops cached batched plans). An
        explicit setting ends the per-set application of the tuning table's
        buckets at upload, so it is not overridden."""
        self.tuned = False
        self._tune_set(key, value)

    def _tune_set(self, key, value):
        self._ck(self._L.mst_tune_set(self.h, key.encode(), int(value)), "mst_tune_set")

    def tune_get(self, key):
        v = C.c_int(0)
        self._ck(self._L.mst_tune_get(self.h, key.encode(), C.byref(v)), "mst_tune_get")
        return v.value

    def last_launch_stats(self):
        """Per-launch statistics of the last call that records them (all-pairs
        launches, or the flag kernels and traceback of align_pairs):
        [(kernel_class, pairs, cells, ms)]."""
        L = self._L
        cnt = C.c_int(0)
        L.mst_last_launch_stats(self.h, None, None, None, None, 0, C.byref(cnt))
        k = np.zeros(max(1, cnt.value), np.int32)
        p = np.zeros(max(1, cnt.value), np.int64)
        c = np.zeros(max(1, cnt.value), np.int64)
        m = np.zeros(max(1, cnt.value), np.float32)
        self._ck(L.mst_last_launch_stats(self.h, _p(k), _p(p), _p(c), _p(m), cnt.value,
                                         C.byref(cnt)), "mst_last_launch_stats")
        n = cnt.value
        return list(zip(k[:n].tolist(), p[:n].tolist(), c[:n].tolist(), m[:n].tolist()))

    def copy_probe(self, nbytes=1 << 30, reps=3):
        """GB/s (read + written) of a plain device copy of nbytes."""
        g = C.c_double(0)
        self._ck(self._L.mst_copy_probe(self.h, int(nbytes), int(reps), C.byref(g)), "mst_copy_probe")
        return g.value

    def valu_peak(self, packed=False):
        """Measured integer VALU rate of the device in T ops/s (int32, or packed int16)."""
        t = C.c_double(0)
        self._ck(self._L.mst_valu_peak(self.h, 1 if packed else 0, C.byref(t)), "mst_valu_peak")
        return t.value

    def nj(self, D, ratio=0.5):
        """Neighbour joining of an n x n float matrix -> list of Join."""
        D = np.ascontiguousarray(D, dtype=np.float32)
        n = D.shape[0]
        joins = (Join * max(1, n - 1))()
        self._ck(self._L.mst_nj(self.h, _p(D), n, ratio, joins), "mst_nj")
        return list(joins)[: n - 1]

    def upload(self, seqs):
        """Uploads a sequence set. A tuned context (until tune_set is called)
        first applies the tuning table's bucket of the set (by median
        length), when the table has one."""
        if self.tuned:
            from . import autotune
            applied = autotune.apply_for(self, [len(s) for s in seqs])
            if applied:
                self.tuning = applied
        buf, off = pack(seqs)
        sid = C.c_int(0)
        self._ck(self._L.mst_seqset_upload(self.h, _p(buf), _p(off), len(seqs), C.byref(sid)),
                 "mst_seqset_upload")
        return SeqSet(self, sid.value, len(seqs), int(off[-1]))


class SeqSet:
    def __init__(self, ctx, sid, n, total):
        self.ctx, self.id, self.n, self.total = ctx, sid, n, total

    def free(self):
        self.ctx._ck(self.ctx._L.mst_seqset_free(self.ctx.h, self.id), "mst_seqset_free")

    def codes(self):
        out = np.zeros(max(1, self.total), dtype=np.int8)
        self.ctx._ck(self.ctx._L.mst_seqset_codes(self.ctx.h, self.id, _p(out), len(out)),
                     "mst_seqset_codes")
        return out[: self.total]

    def all_pairs(self, form=FORM_PLAIN, want_dist=False):
        sc = np.zeros((self.n, self.n), dtype=np.int32)
        dist = np.zeros((self.n, self.n), dtype=np.float32) if want_dist else None
        self.ctx._ck(self.ctx._L.mst_all_pairs(self.ctx.h, self.id, form, _p(sc), _p(dist)),
                     "mst_all_pairs")
        return (sc, dist) if want_dist else sc

    def score_pairs(self, pi, pj, form=FORM_PLAIN):
        pi = np.ascontiguousarray(pi, dtype=np.int32)
        pj = np.ascontiguousarray(pj, dtype=np.int32)
        out = np.zeros(len(pi), dtype=np.int32)
        self.ctx._ck(self.ctx._L.mst_score_pairs(self.ctx.h, self.id, _p(pi), _p(pj), len(pi), form,
                                                 _p(out)), "mst_score_pairs")
        return out

    def score_pairs_noalloc(self, pi, pj):
        """score_pairs (plain form) through the entry point that allocates no
        device memory (mst_score_pairs_noalloc)."""
        pi = np.ascontiguousarray(pi, dtype=np.int32)
        pj = np.ascontiguousarray(pj, dtype=np.int32)
        out = np.zeros(len(pi), dtype=np.int32)
        self.ctx._ck(self.ctx._L.mst_score_pairs_noalloc(self.ctx.h, self.id, _p(pi), _p(pj), len(pi),
                                                         _p(out)), "mst_score_pairs_noalloc")
        return out

    def align_pairs(self, pi, pj, lengths):
        """Global alignments with traceback -> (scores, [gapped_x], [gapped_y]).
        lengths: the set's sequence lengths (for the output stride)."""
        pi = np.ascontiguousarray(pi, dtype=np.int32)
        pj = np.ascontiguousarray(pj, dtype=np.int32)
        np_ = len(pi)
        stride = int(max([lengths[a] + lengths[b] for a, b in zip(pi, pj)] + [0])) + 1
        sc = np.zeros(np_, dtype=np.int32)
        gx = C.create_string_buffer(max(1, np_ * stride))
        gy = C.create_string_buffer(max(1, np_ * stride))
        self.ctx._ck(self.ctx._L.mst_align_pairs(self.ctx.h, self.id, _p(pi), _p(pj), np_, _p(sc),
                                                 gx, gy, stride), "mst_align_pairs")
        rx, ry = gx.raw, gy.raw
        outx = [rx[k * stride: rx.index(b"\0", k * stride)].decode() for k in range(np_)]
        outy = [ry[k * stride: ry.index(b"\0", k * stride)].decode() for k in range(np_)]
        return sc, outx, outy

    def nj(self, ratio=0.5):
        """NJ on this set's resident score matrix (after all_pairs)."""
        joins = (Join * max(1, self.n - 1))()
        self.ctx._ck(self.ctx._L.mst_nj_scores(self.ctx.h, self.id, ratio, joins), "mst_nj_scores")
        return list(joins)[: self.n - 1]

    def msa(self, joins):
        """Progressive alignment along `joins` -> list of gapped rows."""
        arr = (Join * max(1, len(joins)))(*joins)
        L = C.c_int(0)
        self.ctx._ck(self.ctx._L.mst_msa(self.ctx.h, self.id, arr, C.byref(L)), "mst_msa")
        stride = L.value + 1
        buf = C.create_string_buffer(self.n * stride)
        self.ctx._ck(self.ctx._L.mst_msa_rows(self.ctx.h, buf, stride), "mst_msa_rows")
        raw = buf.raw
        return [raw[r * stride: r * stride + L.value].decode() for r in range(self.n)]

    def all_pairs_resident(self, form=FORM_BATCHED):
        """All pairs with the matrix left on the device; returns device ms."""
        ms = C.c_float(0)
        self.ctx._ck(self.ctx._L.mst_all_pairs_resident(self.ctx.h, self.id, form, C.byref(ms)),
                     "mst_all_pairs_resident")
        return ms.value

    def launch_stats(self, form=FORM_BATCHED):
        """One pass with per-launch events -> [(kernel_class, pairs, cells, ms)].
        kernel_class: K > 0, the int32 kernel at K rows per lane; the packed
        kernels at K rows per lane: -(1000 + 100 W + K) the streamed kernel of
        W waves, -(2000 + K) the multi-sweep kernel, -(3000 + K) stacked row
        pairs, -(4000 + K) score8, -(5000 + K) score8 re-based."""
        L = self.ctx._L
        self.ctx._ck(L.mst_all_pairs_launches(self.ctx.h, self.id, form), "mst_all_pairs_launches")
        cnt = C.c_int(0)
        L.mst_last_launch_stats(self.ctx.h, None, None, None, None, 0, C.byref(cnt))
        k = np.zeros(cnt.value, np.int32)
        p = np.zeros(cnt.value, np.int64)
        c = np.zeros(cnt.value, np.int64)
        m = np.zeros(cnt.value, np.float32)
        self.ctx._ck(L.mst_last_launch_stats(self.ctx.h, _p(k), _p(p), _p(c), _p(m), cnt.value,
                                             C.byref(cnt)), "mst_last_launch_stats")
        return list(zip(k.tolist(), p.tolist(), c.tolist(), m.tolist()))


JOB_WAITING, JOB_RUNNING, JOB_DONE, JOB_CANCELLED, JOB_REJECTED = range(5)
SERVE_SCORES, SERVE_TREE, SERVE_MSA = 1, 3, 7


class Server:
    """Served form (mst_server_* in include/mestrado.h): continuous batching of
    jobs (all-pairs scores, and their NJ tree and alignment when asked) on one
    context."""

    def __init__(self, ctx, budget_bytes=8 << 30, tuned=True):
        """tuned: apply the served settings of the device's tuning table
        (mestrado_amd.autotune --serve), when it has them."""
        self.ctx = ctx
        self._L = ctx._L
        h = C.c_void_p()
        ctx._ck(self._L.mst_server_create(ctx.h, int(budget_bytes), C.byref(h)), "mst_server_create")
        self.h = h
        self.settings = {}
        if tuned:
            from . import autotune
            for k, v in autotune.served_settings(ctx.arch()):
                self.set(k, v)
                self.settings[k] = v

    def close(self):
        if self.h:
            self._L.mst_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, seqs, what=SERVE_SCORES):
        """what: SERVE_SCORES, SERVE_TREE (and the NJ tree) or SERVE_MSA (and
        the alignment along it)."""
        buf, off = pack(seqs)
        j = C.c_int64(0)
        self.ctx._ck(self._L.mst_server_submit_job(self.h, _p(buf), _p(off), len(seqs), what,
                                                   C.byref(j)), "mst_server_submit_job")
        return j.value

    def tree(self, jid):
        """The n-1 joins of a done job submitted with a tree."""
        n = self.job(jid)["n"]
        joins = (Join * max(1, n - 1))()
        self.ctx._ck(self._L.mst_server_tree(self.h, jid, joins), "mst_server_tree")
        return list(joins)[: n - 1]

    def msa(self, jid):
        """Gapped rows (input order) of a done job submitted with an alignment."""
        n = self.job(jid)["n"]
        L = C.c_int(0)
        self.ctx._ck(self._L.mst_server_msa(self.h, jid, None, 0, C.byref(L)), "mst_server_msa")
        stride = L.value + 1
        buf = C.create_string_buffer(n * stride)
        self.ctx._ck(self._L.mst_server_msa(self.h, jid, buf, stride, C.byref(L)), "mst_server_msa")
        raw = buf.raw
        return [raw[r * stride: r * stride + L.value].decode() for r in range(n)]

    def step(self):
        n = C.c_int(0)
        self.ctx._ck(self._L.mst_server_step(self.h, C.byref(n)), "mst_server_step")
        return n.value

    def drain(self):
        self.ctx._ck(self._L.mst_server_drain(self.h), "mst_server_drain")

    def job(self, jid):
        st, n = C.c_int(0), C.c_int(0)
        ts, ta, td = C.c_double(0), C.c_double(0), C.c_double(0)
        self.ctx._ck(self._L.mst_server_job(self.h, jid, C.byref(st), C.byref(n), C.byref(ts),
                                            C.byref(ta), C.byref(td)), "mst_server_job")
        return {"state": st.value, "n": n.value, "t_submit": ts.value, "t_admit": ta.value,
                "t_done": td.value}

    def times(self, jid):
        """Host steady-clock times of a job: submit, admit, scores done, tree
        done, done (0 for a stage not reached)."""
        t = np.zeros(5, dtype=np.float64)
        self.ctx._ck(self._L.mst_server_job_times(self.h, jid, _p(t)), "mst_server_job_times")
        return t

    def job_stats(self, jid):
        """Stage accounting of a job: merge batches that held its merges,
        seconds of the alignment stage with at least one of them in flight
        (busy) or none (idle), score rounds that held its items, and seconds
        from submission to the first of them and to admission."""
        t = np.zeros(6, dtype=np.float64)
        self.ctx._ck(self._L.mst_server_job_stats(self.h, jid, _p(t)), "mst_server_job_stats")
        return {"batches": int(t[0]), "busy_s": float(t[1]), "idle_s": float(t[2]), "rounds": int(t[3]),
                "first_round_s": float(t[4]), "admit_s": float(t[5])}

    def result(self, jid):
        n = self.job(jid)["n"]
        out = np.zeros((n, n), dtype=np.int32)
        self.ctx._ck(self._L.mst_server_result(self.h, jid, _p(out)), "mst_server_result")
        return out

    def release(self, jid):
        self.ctx._ck(self._L.mst_server_result(self.h, jid, None), "mst_server_result")

    def completed(self, cap=4096):
        """Ids of the jobs finished since the previous call."""
        ids = np.zeros(cap, dtype=np.int64)
        n = C.c_int(0)
        self.ctx._ck(self._L.mst_server_completed(self.h, _p(ids), cap, C.byref(n)),
                     "mst_server_completed")
        return ids[: n.value].tolist()

    def cancel(self, jid):
        self.ctx._ck(self._L.mst_server_cancel(self.h, jid), "mst_server_cancel")

    def stats(self):
        w, r, i = C.c_int(0), C.c_int(0), C.c_int(0)
        used, rate = C.c_int64(0), C.c_double(0)
        self.ctx._ck(self._L.mst_server_stats(self.h, C.byref(w), C.byref(r), C.byref(i),
                                              C.byref(used), C.byref(rate)), "mst_server_stats")
        return {"waiting": w.value, "running": r.value, "inflight": i.value,
                "arena_used": used.value, "cells_per_s": rate.value}

    def set(self, key, value):
        self.ctx._ck(self._L.mst_server_set(self.h, key.encode(), float(value)), "mst_server_set")

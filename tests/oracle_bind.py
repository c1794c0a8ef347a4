"""ctypes bindings of the oracle (oracle/liboracle.so) and the reference
driver (oracle/_ref/ref_tool). Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_TOOL = os.path.join(ROOT, "oracle", "_ref", "ref_tool")

LETTERS = "ARNDCQEGHILKMFPSTWYVBZX*"
STANDARD20 = "ACDEFGHIKLMNPQRSTVWY"


class JoinT(C.Structure):
    _fields_ = [("node", C.c_int), ("left", C.c_int), ("right", C.c_int),
                ("d", C.c_float), ("Ti", C.c_double), ("Tj", C.c_double),
                ("n", C.c_int), ("q", C.c_double), ("lenL", C.c_double),
                ("lenR", C.c_double)]


def _i8(a):
    return np.ascontiguousarray(a, dtype=np.int8).ctypes.data_as(C.POINTER(C.c_int8))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
        L = C.CDLL(ORACLE_SO)
        P8 = C.POINTER(C.c_int8)
        PU8 = C.POINTER(C.c_uint8)
        L.orc_encode_reverse.argtypes = [C.c_char_p, C.c_int, P8]
        L.orc_nw_score.argtypes = [P8, C.c_int, P8, C.c_int, P8, C.c_int, C.c_int]
        L.orc_nw_score.restype = C.c_int32
        L.orc_nw_flags.argtypes = [P8, C.c_int, P8, C.c_int, P8, C.c_int, C.c_int, PU8]
        L.orc_nw_flags.restype = C.c_int32
        L.orc_nw_traceback.argtypes = [PU8, P8, C.c_int, P8, C.c_int, C.c_char_p, C.c_char_p]
        L.orc_gotoh_score.argtypes = [P8, C.c_int, P8, C.c_int, P8, C.c_int, C.c_int]
        L.orc_gotoh_score.restype = C.c_int64
        L.orc_rescore.argtypes = [C.c_char_p, C.c_char_p, C.c_int, P8, C.c_int, C.c_int]
        L.orc_rescore.restype = C.c_int64
        L.orc_bruteforce.argtypes = [P8, C.c_int, P8, C.c_int, P8, C.c_int, C.c_int]
        L.orc_bruteforce.restype = C.c_int64
        L.orc_all_pairs.argtypes = [P8, C.POINTER(C.c_int64), C.c_int, P8, C.c_int, C.c_int,
                                    C.POINTER(C.c_int32), C.c_int, C.c_int64]
        L.orc_all_pairs.restype = C.c_double
        L.orc_pair_list.argtypes = [P8, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32), C.c_int64, P8, C.c_int, C.c_int,
                                    C.POINTER(C.c_int32), C.c_int]
        L.orc_pair_list.restype = C.c_double
        L.orc_profile_scores.argtypes = [C.POINTER(C.c_int32), C.c_int, C.c_int,
                                         C.POINTER(C.c_int32), C.c_int, C.c_int, P8,
                                         C.POINTER(C.c_int32)]
        L.orc_nw_flags_sc.argtypes = [C.POINTER(C.c_int32), C.c_int, C.c_int, C.c_int, C.c_int, PU8]
        L.orc_nw_flags_sc.restype = C.c_int32
        L.orc_nw_path.argtypes = [PU8, C.c_int, C.c_int, C.c_char_p]
        L.orc_nj_reference.argtypes = [C.POINTER(C.c_float), C.c_int, C.c_int, C.c_float,
                                       C.c_float, C.c_int, C.c_int, C.POINTER(JoinT),
                                       C.POINTER(C.c_int)]
        L.orc_nj_fullscan.argtypes = [C.POINTER(C.c_float), C.c_int, C.c_float, C.POINTER(JoinT)]
        L.orc_progressive_msa.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.c_int,
                                          C.POINTER(C.c_int32), C.c_int, P8, C.c_int, C.c_int,
                                          C.c_char_p, C.c_int64, C.POINTER(C.c_int), C.c_int]
        L.orc_tree_text.argtypes = [C.POINTER(JoinT), C.c_int, C.POINTER(C.c_char_p),
                                    C.c_char_p, C.c_int]
        _lib = L
    return _lib


def blosum62():
    """BLOSUM62 as stored in the reference's blosum64.txt (24x24, letter order
    ARNDCQEGHILKMFPSTWYVBZX*), committed as a fixture under tests/golden."""
    path = os.path.join(ROOT, "tests", "golden", "blosum62_24.txt")
    rows = [r.strip().rstrip(",") for r in open(path) if r.strip()]
    return np.array([[int(v) for v in r.split(",")] for r in rows], dtype=np.int8)


def encode_rev(s):
    out = np.zeros(len(s), dtype=np.int8)
    lib().orc_encode_reverse(s.encode(), len(s), _i8(out) if False else out.ctypes.data_as(C.POINTER(C.c_int8)))
    return out


def nw_score(x, y, S, go=10, ge=2):
    X, Y = encode_rev(x), encode_rev(y)
    return lib().orc_nw_score(_i8(X), len(X), _i8(Y), len(Y), _i8(S), go, ge)


def nw_score_codes(X, Y, S, go=10, ge=2):
    return lib().orc_nw_score(_i8(X), len(X), _i8(Y), len(Y), _i8(S), go, ge)


def nw_align(x, y, S, go=10, ge=2):
    """Returns (score, gapped_x, gapped_y, nibbles[m][n])."""
    X, Y = encode_rev(x), encode_rev(y)
    n, m = len(X), len(Y)
    nib = np.zeros(max(1, n * m), dtype=np.uint8)
    sc = lib().orc_nw_flags(_i8(X), n, _i8(Y), m, _i8(S), go, ge,
                            nib.ctypes.data_as(C.POINTER(C.c_uint8)))
    bx = C.create_string_buffer(n + m + 1)
    by = C.create_string_buffer(n + m + 1)
    lib().orc_nw_traceback(nib.ctypes.data_as(C.POINTER(C.c_uint8)), _i8(X), n, _i8(Y), m, bx, by)
    return sc, bx.value.decode(), by.value.decode(), nib[: n * m].reshape(m, n) if n * m else nib[:0]


def gotoh(x, y, S, go=10, ge=2):
    X, Y = encode_rev(x), encode_rev(y)
    return lib().orc_gotoh_score(_i8(X), len(X), _i8(Y), len(Y), _i8(S), go, ge)


def rescore(gx, gy, S, go=10, ge=2):
    return lib().orc_rescore(gx.encode(), gy.encode(), len(gx), _i8(S), go, ge)


def bruteforce(x, y, S, go=10, ge=2):
    X, Y = encode_rev(x), encode_rev(y)
    return lib().orc_bruteforce(_i8(X), len(X), _i8(Y), len(Y), _i8(S), go, ge)


def pack_seqs(seqs):
    """ASCII sequences -> (reversed codes concatenated, int64 offsets[n+1])."""
    offs = np.zeros(len(seqs) + 1, dtype=np.int64)
    for i, s in enumerate(seqs):
        offs[i + 1] = offs[i] + len(s)
    codes = np.zeros(max(1, int(offs[-1])), dtype=np.int8)
    for i, s in enumerate(seqs):
        if len(s):
            codes[offs[i]:offs[i + 1]] = encode_rev(s)
    return codes, offs


def all_pairs(seqs, S, go=10, ge=2, threads=0):
    codes, offs = pack_seqs(seqs)
    n = len(seqs)
    out = np.zeros((n, n), dtype=np.int32)
    secs = lib().orc_all_pairs(_i8(codes), offs.ctypes.data_as(C.POINTER(C.c_int64)), n, _i8(S),
                               go, ge, out.ctypes.data_as(C.POINTER(C.c_int32)), threads, 0)
    return out, secs


def pair_list(codes, offs, pi, pj, S, go=10, ge=2, threads=0):
    pi = np.ascontiguousarray(pi, dtype=np.int32)
    pj = np.ascontiguousarray(pj, dtype=np.int32)
    out = np.zeros(len(pi), dtype=np.int32)
    secs = lib().orc_pair_list(_i8(codes), np.ascontiguousarray(offs, dtype=np.int64).ctypes.data_as(C.POINTER(C.c_int64)),
                               pi.ctypes.data_as(C.POINTER(C.c_int32)), pj.ctypes.data_as(C.POINTER(C.c_int32)),
                               len(pi), _i8(S), go, ge, out.ctypes.data_as(C.POINTER(C.c_int32)), threads)
    return out, secs


def nj_reference(D, numBins=30, ratio=0.5, pqRatio=2.0, qda=8, predictQ=True):
    D = np.ascontiguousarray(D, dtype=np.float32)
    n = D.shape[0]
    joins = (JoinT * max(1, n - 1))()
    trig = C.c_int(0)
    rb = lib().orc_nj_reference(D.ctypes.data_as(C.POINTER(C.c_float)), n, numBins, ratio, pqRatio,
                                qda, 1 if predictQ else 0, joins, C.byref(trig))
    return list(joins)[: n - 1], rb, trig.value


def nj_fullscan(D, ratio=0.5):
    D = np.ascontiguousarray(D, dtype=np.float32)
    n = D.shape[0]
    joins = (JoinT * max(1, n - 1))()
    rb = lib().orc_nj_fullscan(D.ctypes.data_as(C.POINTER(C.c_float)), n, ratio, joins)
    return list(joins)[: n - 1], rb


def tree_text(joins, names):
    n0 = len(names)
    arr = (JoinT * max(1, len(joins)))(*joins)
    nm = (C.c_char_p * n0)(*[s.encode() for s in names])
    need = lib().orc_tree_text(arr, n0, nm, None, 0)
    buf = C.create_string_buffer(need)
    lib().orc_tree_text(arr, n0, nm, buf, need)
    return buf.value.decode()


def ref_tree_text(joins, names):
    """Tree text from the reference's own TreeNode (oracle/_ref/ref_tool)."""
    lines = [str(len(names))] + list(names)
    for J in joins:
        lines.append("%d %d %d %s %s %s %d" % (J.node, J.left, J.right, float(J.d).hex(),
                                                float(J.Ti).hex(), float(J.Tj).hex(), J.n))
    out = subprocess.run([REF_TOOL, "tree"], input="\n".join(lines) + "\n", capture_output=True,
                         text=True, check=True)
    return out.stdout


def have_ref_tool():
    return os.path.exists(REF_TOOL)


def profile_counts(rows):
    """rows: list of equal-length strings (letters and '-') -> int32 [len][25]."""
    L = len(rows[0]) if rows else 0
    cnt = np.zeros((L, 25), dtype=np.int32)
    idx = {c: k for k, c in enumerate(LETTERS)}
    for r in rows:
        for t, ch in enumerate(r):
            cnt[t, 24 if ch == "-" else idx.get(ch.upper(), 22)] += 1
    return cnt


def profile_align(rowsX, rowsY, S, go=10, ge=2):
    """Oracle profile-profile alignment (the DP runs on reversed columns).
    Returns (score, ops) with ops in forward order: 'D', 'L' (X column vs gaps),
    'U' (gaps vs Y column)."""
    cx = profile_counts(rowsX)[::-1].copy()
    cy = profile_counts(rowsY)[::-1].copy()
    n, m = len(cx), len(cy)
    sc = np.zeros(max(1, n * m), dtype=np.int32)
    P32 = C.POINTER(C.c_int32)
    lib().orc_profile_scores(cx.ctypes.data_as(P32), n, len(rowsX), cy.ctypes.data_as(P32), m,
                             len(rowsY), _i8(S), sc.ctypes.data_as(P32))
    nib = np.zeros(max(1, n * m), dtype=np.uint8)
    score = lib().orc_nw_flags_sc(sc.ctypes.data_as(P32), n, m, go, ge,
                                  nib.ctypes.data_as(C.POINTER(C.c_uint8)))
    ops = C.create_string_buffer(n + m + 1)
    lib().orc_nw_path(nib.ctypes.data_as(C.POINTER(C.c_uint8)), n, m, ops)
    return score, ops.value.decode()


def progressive_msa(seqs, joins, S, go=10, ge=2):
    """Oracle progressive alignment along `joins` (objects with node/left/right);
    returns the gapped rows in input order."""
    canon = ["".join(LETTERS[c] for c in encode_rev(s)[::-1]) for s in seqs]
    groups = {i: ([i], [canon[i]]) for i in range(len(seqs))}
    for J in joins:
        mx, rx = groups.pop(J.left)
        my, ry = groups.pop(J.right)
        if rx[0] == "" or ry[0] == "":
            ops = ("U" * len(ry[0])) if rx[0] == "" else ("L" * len(rx[0]))
        else:
            _, ops = profile_align(rx, ry, S, go, ge)
        nx = [""] * len(rx)
        ny = [""] * len(ry)
        ix = iy = 0
        for op in ops:
            for k in range(len(rx)):
                nx[k] += "-" if op == "U" else rx[k][ix]
            for k in range(len(ry)):
                ny[k] += "-" if op == "L" else ry[k][iy]
            ix += op != "U"
            iy += op != "L"
        groups[J.node] = (mx + my, nx + ny)
    (members, rows), = groups.values()
    out = [None] * len(seqs)
    for i, r in zip(members, rows):
        out[i] = r
    return out


def progressive_msa_c(seqs, joins, S, go=10, ge=2, threads=0):
    """progressive_msa in C++ (oracle/msa.cpp), for large sets: the same
    merges, rows in input order."""
    n = len(seqs)
    off = np.zeros(n + 1, dtype=np.int64)
    for i, s in enumerate(seqs):
        off[i + 1] = off[i] + len(s)
    j3 = np.array([(J.node, J.left, J.right) for J in joins], dtype=np.int32).reshape(-1)
    if j3.size == 0:
        j3 = np.zeros(3, dtype=np.int32)
    L = C.c_int(0)
    asc = "".join(seqs).encode()
    stride = 2 * max([len(s) for s in seqs] + [0]) + 64  # room for the usual gap columns
    for _ in range(2):
        buf = C.create_string_buffer(n * stride)
        rc = lib().orc_progressive_msa(asc, off.ctypes.data_as(C.POINTER(C.c_int64)), n,
                                       j3.ctypes.data_as(C.POINTER(C.c_int32)), len(joins),
                                       _i8(S), go, ge, buf, stride, C.byref(L), threads)
        if rc != -2:
            break
        stride = L.value + 1
    if rc != 0:
        raise ValueError("orc_progressive_msa: joins do not form one tree")
    raw = buf.raw
    return [raw[r * stride: r * stride + L.value].decode() for r in range(n)]

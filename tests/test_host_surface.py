"""The drop-in host classes against the reference's own FastaFile,
CmdLineOptions and TreeNode, compiled from the reference's sources into
oracle/_ref/ref_tool: same records, same parsed options, same tree text.
Runs on the CPU (no device)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TOOL = os.path.join(ROOT, "mestrado_amd", "bin", "host_tool")
FASTA = os.path.join(ROOT, "tests", "golden", "fasta")
REF_CASES = "/root/reference"

need_ref = pytest.mark.skipif(not O.have_ref_tool(), reason="oracle/_ref not built")


def run(tool, *args, stdin=None):
    p = subprocess.run([tool] + list(args), input=stdin, capture_output=True, text=True)
    return p.returncode, p.stdout


@need_ref
@pytest.mark.parametrize("name", ["basic.fasta", "no_trailing_newline.fasta", "crlf.fasta",
                                  "case05_head.fasta"])
def test_fasta_records_match_reference(name):
    path = os.path.join(FASTA, name)
    assert run(HOST_TOOL, "fasta", path) == run(O.REF_TOOL, "fasta", path)


@need_ref
@pytest.mark.parametrize("name", ["no_header.fasta", "empty_last.fasta"])
def test_fasta_malformed_aborts_like_reference(name):
    path = os.path.join(FASTA, name)
    rc_h, out_h = run(HOST_TOOL, "fasta", path)
    rc_r, out_r = run(O.REF_TOOL, "fasta", path)
    assert rc_h != 0 and rc_r != 0 and out_h == out_r


@need_ref
@pytest.mark.skipif(not os.path.isdir(REF_CASES), reason="reference datasets not present")
@pytest.mark.parametrize("case", ["case01.fasta", "case03.fasta", "case05.fasta"])
def test_fasta_reference_datasets(case):
    path = os.path.join(REF_CASES, case)
    assert run(HOST_TOOL, "fasta", path) == run(O.REF_TOOL, "fasta", path)


@need_ref
@pytest.mark.parametrize("args", [
    ["x.fasta"],
    ["-k", "-R", "-d", "2", "-s", "12", "-r", "0.25", "-m", "3.5", "-q", "0", "dir/sub/file.v1.fasta"],
    ["-i", "m", "-o", "M", "/abs/path/noext"],
    ["-i", "z", "x.fasta"],
    ["-o", "q", "x.fasta"],
    ["-h"],
    [],
    ["missing_file.fasta"],
])
def test_cmdline_matches_reference(args, tmp_path):
    # existing input files are created in a scratch directory
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        for a in args:
            if a.endswith("fasta") and a != "missing_file.fasta" and not a.startswith("/"):
                os.makedirs(os.path.dirname(a) or ".", exist_ok=True)
                open(a, "w").write(">a\nA\n")
        if "/abs/path/noext" in args:
            pytest.skip("absolute path not creatable in the sandbox")
        assert run(HOST_TOOL, "cmd", *args) == run(O.REF_TOOL, "cmd", *args)
    finally:
        os.chdir(cwd)


@need_ref
@pytest.mark.parametrize("n", [1, 2, 5, 64, 300])
def test_tree_text_matches_reference(n):
    rng = np.random.default_rng(n)
    names = [">seq_%d some description" % i for i in range(n)]
    if n == 1:
        joins = []
    else:
        A = rng.integers(-500, 4000, (n, n))
        A = np.triu(A) + np.triu(A, 1).T
        joins, _, _ = O.nj_reference(A.astype(np.float32))
    lines = [str(n)] + names
    for J in joins:
        lines.append("%d %d %d %s %s %s %d" % (J.node, J.left, J.right, float(J.d).hex(),
                                                float(J.Ti).hex(), float(J.Tj).hex(), J.n))
    text = "\n".join(lines) + "\n"
    assert run(HOST_TOOL, "tree", stdin=text) == run(O.REF_TOOL, "tree", stdin=text)


def _matrix_rows(text):
    return [[int(v) for v in line.split()] for line in text.splitlines()]


def test_matrix_file_reader(tmp_path):
    # the -b reader on the blosum64.txt format (the reference ships that file,
    # blosum64.txt:1-24, and reads it nowhere): the fixture parses to BLOSUM62
    path = os.path.join(ROOT, "tests", "golden", "blosum62_24.txt")
    rc, out = run(HOST_TOOL, "matrix", path)
    assert rc == 0 and (np.array(_matrix_rows(out)) == O.blosum62()).all()
    if os.path.exists(os.path.join(REF_CASES, "blosum64.txt")):
        assert run(HOST_TOOL, "matrix", os.path.join(REF_CASES, "blosum64.txt")) == (rc, out)
    rows = open(path).read().splitlines()
    bad = {
        "short.txt": "\n".join(rows[:23]) + "\n",                 # 552 values
        "word.txt": "\n".join(rows[:23] + [rows[23].replace("1", "x")]) + "\n",
        "range.txt": "\n".join(rows[:23] + [rows[23].replace(" 1", " 300")]) + "\n",
    }
    for name, text in bad.items():
        (tmp_path / name).write_text(text)
        assert run(HOST_TOOL, "matrix", str(tmp_path / name))[0] == 1, name
    assert run(HOST_TOOL, "matrix", str(tmp_path / "missing.txt"))[0] == 1
    # whitespace and a trailing comma per row, as blosum64.txt has, are accepted
    (tmp_path / "spaced.txt").write_text("\n".join(" " + r + ", " for r in rows) + "\n")
    assert run(HOST_TOOL, "matrix", str(tmp_path / "spaced.txt")) == (rc, out)


def test_matches_manager_store():
    # getSequence(x, y) is symmetric: the reference swaps so that x >= y
    # (MatchesManager.cpp:37-45); getWindow(wx, wy) addresses the pairs of the
    # window (x / W, y / W). The store returns the two gapped strings as put
    # for the pair, in (larger index, smaller index) order.
    ops = "\n".join([
        "put 5 2 AC-G A-TG",   # x > y: kept as given
        "put 1 7 WW- -WW",     # x < y: kept as (7, 1), strings swapped
        "seq 5 2", "seq 2 5",
        "seq 7 1", "seq 1 7",
        "seq 3 3",
        "win 1 0",             # W = 4: pair (5, 2) lies in window (1, 0)
        "win 0 0", "win 1 1",  # (7, 1) is in window (1, 0) too; nothing in these
        "put 9 8 A A",
        "win 2 2",
    ]) + "\n"
    rc, out = run(HOST_TOOL, "matches", "4", stdin=ops)
    assert rc == 0
    assert out.splitlines() == [
        "AC-G A-TG", "AC-G A-TG",
        "-WW WW-", "-WW WW-",
        "null",
        "AC-G A-TG",
        "null", "null",
        "A A",
    ]

"""The reference's own datasets, case01-06 (README:1-6; the inputs of
exec.sh:11-19), committed gzipped under tests/golden/fasta. Records come from
the FastaFile-compatible reader; all-pairs scores are sampled against the
oracle (J, which the reference maps to a negative index, is read as X on both
sides), including every sequence that holds a J; the NJ tree is compared with
the restatement of the reference's search (and bit for bit with the full-scan
restatement up to 2,000 sequences); the command line's .tree and .msa on
case01 equal the oracle's."""
import gzip
import os
import subprocess

import numpy as np
import pytest

import mestrado_amd as M
import oracle_bind as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FASTA = os.path.join(ROOT, "tests", "golden", "fasta")
HOST_TOOL = os.path.join(ROOT, "mestrado_amd", "bin", "host_tool")
CLUSTAL = os.path.join(ROOT, "mestrado_amd", "bin", "clustal")
S = O.blosum62()
CASES = {"case01": (400, 851), "case02": (1000, 852), "case03": (2000, 240),
         "case04": (4000, 249), "case05": (4000, 71), "case06": (8000, 81)}


def _unpack(case, tmp_path):
    path = tmp_path / (case + ".fasta")
    with gzip.open(os.path.join(FASTA, case + ".fasta.gz"), "rb") as f:
        path.write_bytes(f.read())
    return path


def _records(path):
    out = subprocess.run([HOST_TOOL, "fasta", str(path)], capture_output=True, text=True,
                         check=True).stdout
    names, seqs = [], []
    for line in out.splitlines():
        name, n, seq = line.split("\t")
        assert int(n) == len(seq)
        names.append(name)
        seqs.append(seq)
    return names, seqs


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_dataset(ctx, case, tmp_path):
    names, seqs = _records(_unpack(case, tmp_path))
    n, maxlen = CASES[case]
    assert len(seqs) == n and max(map(len, seqs)) == maxlen
    st = ctx.upload(seqs)
    D = st.all_pairs(M.FORM_BATCHED)
    assert (D == D.T).all()
    rng = np.random.default_rng(int(case[-2:]))
    with_j = [i for i, s in enumerate(seqs) if "J" in s]
    pi = list(rng.integers(0, n, 1500)) + with_j + with_j
    pj = list(rng.integers(0, n, 1500)) + list(rng.integers(0, n, len(with_j))) + with_j
    codes, offs = O.pack_seqs(seqs)
    want, _ = O.pair_list(codes, offs, pi, pj, S)
    assert (D[np.array(pi), np.array(pj)] == want).all()
    g = st.nj()
    st.free()
    Df = D.astype(np.float32)
    ref, _, trig = O.nj_reference(Df)
    assert trig == 0
    for a, r in zip(g, ref):
        assert (a.node, a.left, a.right, a.n) == (r.node, r.left, r.right, r.n)
        for x, y in ((a.Ti, r.Ti), (a.Tj, r.Tj), (a.len_left, r.lenL), (a.len_right, r.lenR)):
            assert abs(x - y) <= 1e-6 * max(1.0, abs(y))
    if n <= 2000:
        full, _ = O.nj_fullscan(Df)
        for a, o in zip(g, full):
            assert (a.left, a.right, a.d, a.Ti, a.Tj, a.q) == (o.left, o.right, o.d, o.Ti, o.Tj, o.q)


def test_clustal_on_case01(ctx, tmp_path):
    path = _unpack("case01", tmp_path)
    p = subprocess.run([CLUSTAL, "-M", path.name], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stderr
    names, seqs = _records(path)
    st = ctx.upload(seqs)
    D = st.all_pairs(M.FORM_BATCHED)
    joins = st.nj()
    st.free()
    ref, _, _ = O.nj_reference(D.astype(np.float32))
    assert [(j.left, j.right) for j in joins] == [(j.left, j.right) for j in ref]
    assert (tmp_path / "case01.tree").read_text() == O.tree_text(ref, names)
    msa = (tmp_path / "case01.msa").read_text().split("\n")
    assert msa[0::2][:400] == names
    assert msa[1::2][:400] == O.progressive_msa_c(seqs, joins, S)

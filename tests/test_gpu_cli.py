"""End-to-end drop-in: the clustal command line on a FASTA file writes the
same <basename>.tree as the oracle pipeline (oracle scores -> restatement of
the reference's NJ -> the reference's own TreeNode text), and the -M
alignment equals the oracle's progressive alignment."""
import os
import re
import subprocess

import numpy as np
import pytest

import mestrado_amd as M
import oracle_bind as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLUSTAL = os.path.join(ROOT, "mestrado_amd", "bin", "clustal")
S = O.blosum62()


def _write_fasta(path, names, seqs, width=80):
    with open(path, "w") as f:
        for n, s in zip(names, seqs):
            f.write(n + "\n")
            for k in range(0, len(s), width):
                f.write(s[k:k + width] + "\n")


@pytest.mark.parametrize("n", [1, 2, 64])
def test_clustal_tree_and_msa(tmp_path, n):
    seqs = M.synth(0, 3)[:n]
    names = [">gi|%d|synthetic protein %d" % (100 + i, i) for i in range(n)]
    fa = tmp_path / "cfg0.fasta"
    _write_fasta(fa, names, seqs)
    p = subprocess.run([CLUSTAL, "-M", str(fa)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    assert re.fullmatch(r"\d+\.\d{4},\d+\.\d{4},\d+\.\d{4},\d+\.\d{4}\n", p.stderr.splitlines(True)[-1])
    tree = (tmp_path / "cfg0.tree").read_text()
    D, _ = O.all_pairs(seqs, S)
    joins, _, _ = O.nj_reference(D.astype(np.float32))
    assert tree == O.tree_text(joins, names)
    if O.have_ref_tool():
        assert tree == O.ref_tree_text(joins, names)
    msa = (tmp_path / "cfg0.msa").read_text().split("\n")
    rows = msa[1::2][:n]
    assert msa[0::2][:n] == names
    if n > 1:
        assert rows == O.progressive_msa(seqs, joins, S)
    else:
        assert rows == seqs


def test_clustal_plain_form_and_matrix_output(tmp_path):
    seqs = M.synth(0, 4)[:20]
    names = [">s%d" % i for i in range(20)]
    fa = tmp_path / "m.fasta"
    _write_fasta(fa, names, seqs)
    p = subprocess.run([CLUSTAL, "-f", "plain", "-o", "m", str(fa)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    lines = (tmp_path / "m.mtx").read_text().splitlines()
    assert lines[0] == "20"
    D, _ = O.all_pairs(seqs, S)
    got = np.array([[float(v) for v in l.split()[1:]] for l in lines[1:]])
    assert (got == D).all()


def test_clustal_bad_arguments():
    assert subprocess.run([CLUSTAL, "-i", "z", "x"], capture_output=True).returncode == 1
    assert subprocess.run([CLUSTAL], capture_output=True).returncode == 0  # help state

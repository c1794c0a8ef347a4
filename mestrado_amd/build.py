"""Builds the native parts of mestrado_amd in-tree.

  libmestrado.so   HIP kernels + C ABI (hipcc, gfx950)
  bin/clustal      the reference-compatible command line (g++, links the above)

Usage: python -m mestrado_amd.build   (or mestrado_amd.build.build()).
"""
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "libmestrado.so")
BIN = os.path.join(PKG, "bin")

ARCH = os.environ.get("MST_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # Integer DP kernels do not need packed fp32; keep it off everywhere.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Bit-exact float paths (NJ, profile scores): no contraction into FMA.
    "-ffp-contract=off",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def _objs(srcs, outdir, jobs=8):
    os.makedirs(outdir, exist_ok=True)
    procs = []
    objs = []
    for src in srcs:
        obj = os.path.join(outdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        if len(procs) >= jobs:
            s, p = procs.pop(0)
            if p.wait():
                raise RuntimeError("compile failed: " + s)
    for s, p in procs:
        if p.wait():
            raise RuntimeError("compile failed: " + s)
    return objs


def build_lib():
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = _objs(srcs, os.path.join(PKG, "_build"))
    for src in sorted(glob.glob(os.path.join(CSRC, "*.cpp"))):
        obj = os.path.join(PKG, "_build", os.path.basename(src) + ".o")
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-c", src, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_cli():
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    if not srcs:
        return None
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "clustal")
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + HOST]
    _run(["g++", "-O2", "-std=c++17"] + inc + ["-o", out] + srcs +
         ["-L" + PKG, "-lmestrado", "-Wl,-rpath,$ORIGIN/.."])
    # device-free driver of the host classes, for the CPU tests
    plain = [os.path.join(HOST, f) for f in ("FastaFile.cpp", "CmdLineOptions.cpp", "TreeNode.cpp",
                                             "MatchesManager.cpp", "MatrixFile.cpp")]
    _run(["g++", "-O2", "-std=c++17"] + inc + ["-o", os.path.join(BIN, "host_tool"),
          os.path.join(HOST, "tools", "host_tool.cpp")] + plain)
    return out


def build():
    build_lib()
    build_cli()


if __name__ == "__main__":
    build()
    sys.exit(0)

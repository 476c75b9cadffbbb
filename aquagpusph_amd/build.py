"""Build the HIP shared library libaquagpusph_amd.so for gfx950 in-tree.

    python -m aquagpusph_amd.build [-j N] [-v]

Every csrc/*.hip and csrc/**/*.cpp source is compiled with hipcc into
aquagpusph_amd/_build/ and linked into aquagpusph_amd/libaquagpusph_amd.so.
Objects are rebuilt when their source or any header is newer.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaquagpusph_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the kernels keep the reference's rounded products and
# sums (no silent FMA contraction), which the bit-exact parity tests rely on.
# -packed-fp32-ops: no packed FP32 (v_pk_*_f32) code generation.
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
          "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
          "-Wno-pass-failed",
          "-Wno-unused-variable", "-I" + os.path.join(ROOT, "include")]


def sources():
    srcs = sorted(glob.glob(os.path.join(CSRC, "**", "*.hip"),
                            recursive=True))
    srcs += sorted(glob.glob(os.path.join(CSRC, "**", "*.cpp"),
                             recursive=True))
    return srcs


def _headers():
    hs = glob.glob(os.path.join(CSRC, "**", "*.hpp"), recursive=True)
    hs += glob.glob(os.path.join(ROOT, "include", "*.h"))
    return hs


def _obj(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "__")
    return os.path.join(BUILD, rel + ".o")


def _compile(src, verbose):
    obj = _obj(src)
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l)
    return src, p.returncode, out


def build(jobs=None, verbose=False, force=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    newest_hdr = max([os.path.getmtime(h) for h in _headers()] + [0])
    todo = []
    for s in srcs:
        o = _obj(s)
        if force or not os.path.exists(o) or \
                os.path.getmtime(o) < max(os.path.getmtime(s), newest_hdr):
            todo.append(s)
    jobs = jobs or min(8, os.cpu_count() or 4)
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, out in ex.map(lambda s: _compile(s, verbose), todo):
            if out.strip():
                print(out, file=sys.stderr)
            if rc != 0:
                failed.append(src)
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    objs = [_obj(s) for s in srcs]
    if force or todo or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        # -z defs: a declaration that matches no definition fails the build
        # instead of the library load
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950",
               "-Wl,-z,defs", "-o", tmp]
        cmd += objs + ["-ldl"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    ap.add_argument("-f", action="store_true", help="rebuild everything")
    a = ap.parse_args()
    print(build(a.j, a.v, a.f))

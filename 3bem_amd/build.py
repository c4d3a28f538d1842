"""Build recipe for the in-tree HIP library 3bem_amd/libtbem_fmm.so (gfx950).

    python 3bem_amd/build.py            # compile (always recompiles)

Every .hip file under csrc/ is compiled by hipcc for gfx950 into one shared
object exporting the C ABI of include/tbem_fmm.h. Packed FP32 instructions are
disabled (-packed-fp32-ops); hipcc's host pass warns that it does not know that
feature, which is harmless.
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtbem_fmm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TBEM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unknown-warning-option",
    "-I", os.path.join(ROOT, "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=True, jobs=None):
    """Compile every csrc/*.hip (in parallel) and link libtbem_fmm.so."""
    t0 = time.time()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    procs = []
    objs = []
    for s in srcs:
        obj = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", s, "-o", obj]
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((s, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for s, text in failed:
            print(f"--- {s}\n{text}", file=sys.stderr)
        raise RuntimeError("HIP build failed: " + ", ".join(os.path.basename(s) for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-shared", *objs, "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        print(r.stdout.decode(errors="replace"), file=sys.stderr)
        raise RuntimeError("HIP link failed")
    os.replace(tmp, LIB)
    build_shim_example()
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


SHIM_EXAMPLE = os.path.join(HERE, "tbem_shim_example")
SPARSE_EXAMPLE = os.path.join(HERE, "tbem_sparse_example")


def build_shim_example():
    """C++ callers of the library through include/tbem/fmm.hpp (host code only, g++)."""
    for name, exe in (("shim_example.cpp", SHIM_EXAMPLE), ("sparse_example.cpp", SPARSE_EXAMPLE)):
        src = os.path.join(ROOT, "tests", "cpp", name)
        cmd = ["g++", "-O2", "-std=c++14", "-I", os.path.join(ROOT, "include"), src, "-L", HERE, "-ltbem_fmm",
               "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            print(r.stdout.decode(errors="replace"), file=sys.stderr)
            raise RuntimeError(f"{name} build failed")


if __name__ == "__main__":
    build()

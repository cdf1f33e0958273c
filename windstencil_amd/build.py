"""Build the HIP library in place: windstencil_amd/libwindstencil_amd.so (gfx950).

    python -m windstencil_amd.build [--verbose] [--src DIR] [--out LIB]

Every source under csrc/ is compiled with hipcc for gfx950 only and linked into
one shared library.  Kernel resource usage (VGPRs, SGPRs, LDS, scratch) is
written to _build/resource_usage.txt.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwindstencil_amd.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(HERE, "..", "include"),
]


def sources(csrc=CSRC):
    return sorted(os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith((".hip", ".cpp")))


def _compile(src, build_dir=BUILD):
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{p.stderr}")
    return obj, p.stderr


def build(verbose=False, jobs=None, csrc=CSRC, out=LIB):
    """Compile every source under `csrc` and link them into `out` (by default the
    in-tree library the package loads)."""
    build_dir = BUILD if out == LIB else out + ".build"
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources(csrc)
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs, usage = [], []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, build_dir), srcs):
            objs.append(obj)
            usage.append(err)
    with open(os.path.join(build_dir, "resource_usage.txt"), "w") as f:
        f.write("\n".join(u for u in usage if u))
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out} from {len(objs)} sources")
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--src", default=CSRC, help="source directory (default: the package's csrc/)")
    ap.add_argument("--out", default=LIB, help="library to write (default: the in-tree library)")
    a = ap.parse_args()
    try:
        build(verbose=True, csrc=os.path.abspath(a.src), out=os.path.abspath(a.out))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)

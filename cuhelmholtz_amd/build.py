"""Build libcuhelmholtz_amd.so in-tree with hipcc for gfx950.

The fixed-shape kernel tables are first generated from the tuning tables
(kernelgen.generate into _build/gen); each source under csrc/ and each
generated source is compiled to an object in _build (in parallel), then
linked into cuhelmholtz_amd/libcuhelmholtz_amd.so; the command-line drivers
(drivers/*.cpp) are linked against it into bin/.
Usage: python -m cuhelmholtz_amd.build  (or build() from Python).
"""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcuhelmholtz_amd.so")
DRIVERS = os.path.join(HERE, "drivers")
BIN = os.path.join(HERE, "bin")
ARCH = os.environ.get("CHZ_OFFLOAD_ARCH", "gfx950")

GEN = os.path.join(BUILD, "gen")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-I", os.path.join(HERE, "..", "include"),
          "-I", CSRC]
# packed FP32 instructions are disabled for all device code; multiply-adds are
# fused within a source expression only (-ffp-contract=on), so a tile body
# rounds the same way in every kernel it is inlined into (a batched problem
# and its single solve give the same bits)
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-ffp-contract=on"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _deps(path, seen=None):
    """Local headers a source includes, transitively (#include "...", beside
    the source or in csrc/)."""
    seen = set() if seen is None else seen
    with open(path) as f:
        for m in re.finditer(r'#include "([^"]+)"', f.read()):
            h = os.path.normpath(os.path.join(os.path.dirname(path), m.group(1)))
            if not os.path.exists(h):
                h = os.path.join(CSRC, m.group(1))
            if h not in seen and os.path.exists(h):
                seen.add(h)
                _deps(h, seen)
    return seen


def _compile(src, incremental=False):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if incremental and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if all(os.path.getmtime(d) < t for d in [src, __file__, *_deps(src)]):
            return obj
    cmd = [hipcc()] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + DEVICE
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def sources():
    """csrc/ sources plus the kernel tables generated from the tuning tables."""
    from cuhelmholtz_amd import kernelgen
    gen = kernelgen.generate(GEN, arch=ARCH)
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp"))) + gen


def build(verbose=True, incremental=False):
    """Compile every HIP source for gfx950 and link the shared library.

    incremental: recompile only sources whose object is older than the source,
    a local header it includes or this script (a development shortcut; the
    default compiles everything)."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    # longest (kernel) files first so the pool stays busy
    srcs.sort(key=lambda s: -os.path.getsize(s))
    jobs = min(len(srcs), max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_drivers()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources, drivers in {BIN}")
    return LIB


def build_drivers():
    """The command-line drivers (drivers/*.cpp, the reference's test programs'
    arguments and output) linked against the library, into bin/."""
    os.makedirs(BIN, exist_ok=True)
    for src in sorted(f for f in os.listdir(DRIVERS) if f.endswith(".cpp")):
        out = os.path.join(BIN, src[:-4])
        cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"), "-I", DRIVERS,
               os.path.join(DRIVERS, src), "-L", HERE, "-lcuhelmholtz_amd", "-Wl,-rpath,$ORIGIN/..", "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)

"""Builds libhjoin.so (the engine: HIP kernels for gfx950 + the C++ host side + the C ABI).

    python hash-join-optimizer_amd/build.py        # or build() from Python

Compiles every source under csrc/ with hipcc for gfx950 only and links the shared library
in-tree, next to this file, so it travels with the repository snapshot.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhjoin.so")
BUILD_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# The reference picks its hash table at configure time: CMake option HASH_ALGO ->
# USE_<NAME>_HASH (CMakeLists.txt:173-187, include/hash_config.h:3-22). This library is the
# value "hip": the GPU tables (bucketized open addressing for left-deep chains, the unchained
# directory for every other join). The reference's CPU tables are its own build, not this one.
HASH_ALGOS = ("hip",)
REFERENCE_ALGOS = ("robinhood", "hopscotch", "cuckoo", "unchained", "standard", "std")


def hash_algo(value=None):
    """the HASH_ALGO this build compiles (environment HASH_ALGO, default "hip"); any other value
    is refused with the reason"""
    v = (value if value is not None else os.environ.get("HASH_ALGO", "hip")).strip().lower()
    if v in HASH_ALGOS:
        return v
    if v in REFERENCE_ALGOS:
        raise ValueError(f"HASH_ALGO={v}: that table is the reference's CPU build "
                         f"(CMakeLists.txt HASH_ALGO); this library provides only HASH_ALGO=hip")
    raise ValueError(f"HASH_ALGO={v}: unknown value; this library provides only HASH_ALGO=hip")
# packed-FP32 instructions are disabled for all device code (no float math here anyway)
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src, algo="hip"):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, f"-DHJ_HASH_ALGO_{algo.upper()}=1", "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    return obj


def build(verbose=False):
    algo = hash_algo()
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda f: _compile(f, algo), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

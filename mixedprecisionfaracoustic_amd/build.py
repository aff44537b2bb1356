"""Build recipe for libporowave.so (hipcc, gfx950) -- in-tree, no JIT cache.

Usage:  python -m mixedprecisionfaracoustic_amd.build   (or build() from Python)
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libporowave.so")
OBJDIR = os.path.join(HERE, "_build")
# standalone HBM bandwidth probe (bench.py --hbm-ceiling), not part of the library
PROBE = os.path.join(HERE, "pw_hbm_probe")
# command-line replacement for the reference program's main(), linked to the library
CLI = os.path.join(HERE, "porowave")
SOURCES = ["pw_api.cpp", "pw_medium.cpp", "pw_server.cpp", "pw_output.cpp", "pw_kernels.hip", "pw_medium_gpu.hip"]
ARCH = os.environ.get("PW_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # keep the compiler from emitting packed FP32 instructions
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # multiply-adds fuse only inside one source expression, so the same
            # formula gives the same bits in every kernel that spells it
            "-ffp-contract=on",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-I", os.path.join(HERE, "..", "include")]


def _compile(src, verbose):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"]
    # extra flags for tuning experiments (e.g. -DPW_WAVES_PER_EU=8)
    extra = os.environ.get("PW_HIPCC_FLAGS", "").split()
    cmd = [hipcc()] + flags() + extra + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    return obj


def _build_probe(verbose):
    cmd = [hipcc()] + flags() + ["-x", "hip", os.path.join(CSRC, "pw_hbm_probe.cpp"), "-o", PROBE + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on pw_hbm_probe.cpp:\n{r.stdout}\n{r.stderr}")
    os.replace(PROBE + ".tmp", PROBE)
    return PROBE


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=7) as ex:
        probe = ex.submit(_build_probe, verbose)
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
        probe.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"), os.path.join(CSRC, "pw_cli.cpp"),
           "-o", CLI + ".tmp", "-L", HERE, "-lporowave", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on pw_cli.cpp:\n{r.stdout}\n{r.stderr}")
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

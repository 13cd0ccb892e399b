"""In-tree build of the engine's shared library for gfx950.

    python -m pricing_amd.build        (from rust-pricing-library_amd/)

Compiles csrc/*.hip (device + host) and csrc/*.cpp (host) with hipcc for
--offload-arch=gfx950, embeds the Joe--Kuo direction asset, and links
pricing_amd/_lib/librpl_amd.so.  The FP64 kernels are built with
-ffp-contract=off (the reference's rounding sequence); the mixed kernels with
contraction on.  Every HIP compile disables packed FP32 instructions.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

from . import sobol_directions

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                       # rust-pricing-library_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB_DIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIB_DIR, "librpl_amd.so")
OBJ_DIR = os.path.join(ROOT, "_build")
ARCH = os.environ.get("PE_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-std=c++17", "-fPIC", "-O3", f"--offload-arch={ARCH}", f"-I{INCLUDE}", f"-I{CSRC}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
          "-Wno-unused-command-line-argument"]

SOURCES = {
    "pe_kernels_f64.hip": ["-ffp-contract=off"],
    "pe_kernels_f32.hip": ["-ffp-contract=fast"],
    "pe_dupire.hip": ["-ffp-contract=off"],
    "pe_noalloc.hip": ["-ffp-contract=off"],
    "pe_plan.cpp": ["-ffp-contract=off"],
    "pe_api.cpp": ["-ffp-contract=off"],
    "pe_surface.cpp": ["-ffp-contract=off"],
    "pe_graph.cpp": ["-ffp-contract=off"],
    "pe_blake3.cpp": [],
}


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout)
    return proc.stdout


def compile_one(src, extra):
    out = os.path.join(OBJ_DIR, src + ".o")
    path = os.path.join(CSRC, src)
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC] + COMMON + extra + lang + ["-c", path, "-o", out])
    return out


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    directions = sobol_directions.default_path()
    os.makedirs(os.path.dirname(directions), exist_ok=True)
    if not os.path.exists(directions):
        sobol_directions.write(directions)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 2)) as pool:
        futures = [pool.submit(compile_one, s, e) for s, e in SOURCES.items()]
        for f in futures:
            objs.append(f.result())
    asm_obj = os.path.join(OBJ_DIR, "pe_directions.o")
    _run(["gcc", "-c", "-x", "assembler-with-cpp", f'-DPE_DIRECTIONS_FILE="{directions}"',
          os.path.join(CSRC, "pe_directions.S"), "-o", asm_obj])
    objs.append(asm_obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print(LIB)
    return LIB


SANITIZE = ["-Xarch_host", "-fsanitize=address,undefined", "-Xarch_host", "-fno-sanitize-recover=all",
            "-Xarch_host", "-fno-omit-frame-pointer", "-g"]
LIB_ASAN = os.path.join(LIB_DIR, "librpl_amd_asan.so")


def build_sanitized(verbose=False):
    """The engine with AddressSanitizer + UndefinedBehaviorSanitizer on its
    host code (plan compiler, graph compiler, C ABI, BLAKE3, surface); the
    device kernels are built as usual.  Load it with the clang ASan runtime
    preloaded (see sanitizer_runtime())."""
    build()   # the ordinary objects, reused for the HIP translation units
    objs = []
    for src, extra in SOURCES.items():
        if src.endswith(".hip"):
            objs.append(os.path.join(OBJ_DIR, src + ".o"))
            continue
        out = os.path.join(OBJ_DIR, src + ".asan.o")
        _run([HIPCC] + COMMON + extra + SANITIZE + ["-c", os.path.join(CSRC, src), "-o", out])
        objs.append(out)
    objs.append(os.path.join(OBJ_DIR, "pe_directions.o"))
    tmp = LIB_ASAN + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-shared-libasan", "-Xarch_host",
          "-fsanitize=address,undefined", "-o", tmp] + objs)
    os.replace(tmp, LIB_ASAN)
    if verbose:
        print(LIB_ASAN)
    return LIB_ASAN


def sanitizer_runtime():
    """The clang ASan runtime the sanitized libraries link against."""
    import glob
    root = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    hits = sorted(glob.glob(os.path.join(root, "llvm", "lib", "clang", "*", "lib", "linux",
                                         "libclang_rt.asan-x86_64.so")))
    if not hits:
        hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build_sanitized(verbose=True)
    else:
        build(verbose=True)
    sys.exit(0)

"""Build script for libdtwj (HIP/gfx950 kernels + C ABI) and the gpu_dtw command-line tool.

Everything is built in-tree with hipcc (no JIT cache):
    <package>/lib/libdtwj.so   shared library exporting include/dtwj.h
    <package>/bin/gpu_dtw      drop-in for the reference's GPU_dtw <data> <time_out> <m> <r>
    <package>/bin/copy_calib   plain 1 GiB copy kernel for calibrating HBM traffic counters
    <package>/bin/fp16_env_check   the packed FP16 envelope decode against plain arithmetic

Usage: python build.py [-j N]
"""
import os
import subprocess
import sys
import concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DTWJ_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for device code (they have produced silent wrong
# results on this hardware); hipcc's host pass warns that it does not know the feature.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
          "-I", CSRC]

LIB_SOURCES = ["dtwj_api.cpp", "series_gen.cpp", "loader.cpp", "prep.hip", "join.hip", "dtw.hip", "proposal.hip",
               "noalloc.hip", "mpx.hip", "mask.hip"]
CLI_SOURCES = ["gpu_dtw_cli.cpp"]
TOOL_SOURCES = ["copy_calib.hip", "fp16_env_check.hip"]   # standalone executables (traffic-counter calibration,
                                                         # check of the packed-width decode)


def _compile(src, force=False):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + ".o")
    deps = [path] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith((".h", ".hpp", ".inc"))]
    deps.append(os.path.join(ROOT, "include", "dtwj.h"))
    if not force and os.path.exists(out) and all(os.path.getmtime(out) >= os.path.getmtime(d) for d in deps):
        return out, None
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip", "--offload-arch=" + ARCH] + DEVICE_FLAGS
    cmd += ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return out, "%s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr)
    return out, None


def build(jobs=8, verbose=False, force=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda src: _compile(src, force), LIB_SOURCES + CLI_SOURCES + TOOL_SOURCES))
    errs = [e for _, e in results if e]
    if errs:
        raise RuntimeError("compilation failed:\n" + "\n".join(errs))
    objs = {os.path.basename(o)[:-2]: o for o, _ in results}
    lib = os.path.join(LIB, "libdtwj.so")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", lib + ".tmp"] + [objs[s] for s in LIB_SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(BIN, "gpu_dtw")
    cmd = [HIPCC, "-o", exe, objs["gpu_dtw_cli.cpp"], "-L", LIB, "-ldtwj", "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("CLI link failed:\n%s\n%s" % (r.stdout, r.stderr))
    for src in TOOL_SOURCES:
        tool = os.path.join(BIN, src.rsplit(".", 1)[0])
        r = subprocess.run([HIPCC, "--offload-arch=" + ARCH, "-o", tool, objs[src]], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s link failed:\n%s\n%s" % (src, r.stdout, r.stderr))
    if verbose:
        print("built", lib, exe)
    return lib


if __name__ == "__main__":
    j = 8
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    build(jobs=j, verbose=True, force="--force" in sys.argv)

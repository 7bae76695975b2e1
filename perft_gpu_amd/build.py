"""Build the engine in-tree: perft_gpu_amd/_build/libperft_gpu_amd.so and the
perft_gpu CLI, compiled by hipcc for gfx950 only.

    python -m perft_gpu_amd.build          # library + CLI
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libperft_gpu_amd.so")
CLI = os.path.join(OUT, "perft_gpu")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instruction selection is disabled for every HIP translation unit.
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
LIB_SOURCES = ["kernels.hip", "stream.hip", "engine.cpp", "engine_tt.cpp", "tables.cpp", "capi.cpp", "probes.hip", "serve.cpp",
               "tt.hip", "tune.cpp", "records.cpp", "tt_file.cpp", "noalloc.hip"]
HEADERS = ["level_dev.h", "wave.h", "board.h", "sliders.h", "movegen.h", "kernels.h", "engine.h", "tables.h",
           "fixed_shift_magics.inc", "serve.h", "tt.h"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), p.stdout))
    return p.stdout


def _obj(src):
    return os.path.join(OUT, os.path.splitext(src)[0] + ".o")


def build(verbose=False, jobs=None):
    os.makedirs(OUT, exist_ok=True)
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmds = []
    for s in srcs:
        lang = [] if s.endswith(".hip") else ["-x", "hip"]
        cmds.append([HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [_obj(s) for s in srcs] + ["-lpthread"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cli.cpp")
    if os.path.exists(cli_src):
        tmpc = CLI + ".tmp"
        _run([HIPCC] + COMMON + ["-x", "hip", cli_src, "-o", tmpc, "-L" + OUT, "-lperft_gpu_amd",
                                  "-Wl,-rpath,$ORIGIN", "-lpthread"])
        os.replace(tmpc, CLI)
    return LIB


def clean():
    shutil.rmtree(OUT, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

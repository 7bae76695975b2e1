import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

HARNESS_SRC = os.path.join(TESTS, "native", "movegen_check.cpp")
HARNESS = os.path.join(TESTS, "native", "_build", "libmovegen_check.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: long-running check")


def build_harness():
    """Test-only build of the product's move generator for the CPU."""
    srcs = [HARNESS_SRC, os.path.join(ROOT, "perft_gpu_amd", "csrc", "tables.cpp")]
    deps = srcs + [os.path.join(ROOT, "perft_gpu_amd", "csrc", f)
                   for f in ("board.h", "sliders.h", "movegen.h", "convert.h", "tables.h", "tt.h")]
    if os.path.exists(HARNESS) and all(os.path.getmtime(HARNESS) >= os.path.getmtime(d) for d in deps):
        return HARNESS
    os.makedirs(os.path.dirname(HARNESS), exist_ok=True)
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-x", "hip"] + srcs + ["-o", HARNESS], check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL)
    return HARNESS


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def pg():
    import perft_gpu_amd
    perft_gpu_amd.load_library()
    return perft_gpu_amd


@pytest.fixture(scope="session")
def engine(pg):
    eng = pg.Engine(backend="ks", arena_bytes=4 << 30)
    yield eng
    eng.close()

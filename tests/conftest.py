"""Shared test set-up: import paths, the gpu marker, and built libraries."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "vibe-splat_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C interface)")


@pytest.fixture(scope="session", autouse=True)
def built_oracle():
    """The CPU oracles are test infrastructure: build them if missing (seconds, g++ only)."""
    lib = ROOT / "oracle" / "liboracle.so"
    if not lib.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    return lib


def pytest_collection_modifyitems(config, items):
    # GPU tests fail loudly (no fallback) when the library is missing; say so up front.
    if any("gpu" in item.keywords for item in items):
        lib = ROOT / "vibe-splat_amd" / "vsplat" / "libvsplat.so"
        if not lib.exists() and os.environ.get("VSPLAT_REQUIRE_LIB", "1") == "1":
            print(f"\nwarning: {lib} is not built; GPU tests will fail", file=sys.stderr)

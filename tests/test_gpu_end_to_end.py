"""The reference's end-to-end training checks on the GPU, through the two
bindings a caller uses:

* through the reference-shaped C++ class (include/vsplat_trainer.hpp), built by
  tests/cpp/Makefile: "Training overfits a single rendered view"
  (tests/test_train.cpp:493-520: 400 steps from a 40x40 grid halve the loss,
  move the positions and the colours) and "A prefetched image of a smaller
  size does not spoil a larger one after it" (:786-810);
* the same prefetch case through the Python binding.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
PROGRAM = ROOT / "tests" / "cpp" / "wrapper_train"


def _program():
    if not PROGRAM.exists():  # normally built by __graft_entry__.build()
        subprocess.run(["make", "-C", str(PROGRAM.parent)], check=True, capture_output=True)
    return PROGRAM


def _run(*args):
    out = subprocess.run([str(_program()), *args], capture_output=True, text=True, timeout=110)
    values = dict(line.split("=", 1) for line in out.stdout.splitlines() if "=" in line)
    assert out.returncode == 0 and "error" not in values, out.stdout + out.stderr
    return values


def test_cpp_wrapper_overfits_one_view(tmp_path):
    # tests/test_train.cpp:493-520, the target a smooth 128x96 picture.
    target = F.smooth_target(128, 96, 23)
    path = tmp_path / "target.rgba"
    path.write_bytes(np.ascontiguousarray(target, np.uint8).tobytes())
    v = _run("overfit", str(path))
    first, last = float(v["first"]), float(v["last"])
    print(f"overfitting one view through the C++ wrapper: loss {first:.4f} -> {last:.4f}")
    assert last < first * 0.5
    assert int(v["count"]) == 1600 and v["positions_moved"] == "1" and v["sh_moved"] == "1"


def test_cpp_wrapper_prefetch_of_a_smaller_image():
    # tests/test_train.cpp:786-810
    v = _run("prefetch")
    expected = float(v["expected"])
    assert abs(float(v["after_small"]) - expected) < 1e-4
    assert abs(float(v["prefetched"]) - expected) < 1e-4
    assert int(v["staging"]) >= int(v["large_bytes"])


def test_prefetch_of_a_smaller_image_through_python():
    # tests/test_train.cpp:786-810
    small, large = F.train_camera(64, 48), F.train_camera(1024, 768)
    large.fx = large.fy = 960.0
    small_target = F.random_target(64, 48, 71)
    large_target, other = F.random_target(1024, 768, 73), F.random_target(1024, 768, 79)
    cloud = F.random_cloud(60, 0, 83)
    cloud.positions[:, 2] = np.abs(cloud.positions[:, 2]) + 2.0
    with vsplat.Trainer(cloud) as reference:
        expected = reference.compute_gradients(large, large_target).loss
    still = vsplat.LearningRates(position=0.0, scale=0.0, rotation=0.0, opacity=0.0, sh=0.0)
    with vsplat.Trainer(cloud) as t:
        t.render(large)
        t.prefetch_target(small_target)
        t.compute_gradients(small, small_target)
        assert abs(t.step(large, large_target, still, next_target=other).loss - expected) < 1e-4
        assert t.staging_bytes() >= large_target.nbytes
        t.step(large, large_target, still, next_target=large_target)
        assert abs(t.step(large, large_target, still).loss - expected) < 1e-4

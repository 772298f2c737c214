"""The C interface without a GPU: the library loads, exports every function
include/vsplat.h declares, and rejects bad arguments with the reference's
error kinds before it touches a device."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import vsplat
from vsplat import fixtures as F

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "vsplat.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(vsplat_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = vsplat.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(vsplat.EXPORTS) == names


def test_defaults_are_the_reference_defaults():
    """TrainerOptions and LearningRates defaults (src/train/trainer.hpp:24-61)."""
    o = vsplat.COptions()
    vsplat.lib().vsplat_default_options(C.byref(o))
    assert abs(o.lambda_ssim - 0.2) < 1e-7 and abs(o.near_plane - 0.01) < 1e-9 and o.selective_adam == 1
    assert o.force_fixed_point == 0 and o.memory_budget == 2 ** 64 - 1 and list(o.background) == [0, 0, 0]
    r = vsplat.CRates()
    vsplat.lib().vsplat_default_rates(C.byref(r))
    want = (1.6e-4, 0.005, 0.001, 0.05, 0.025, 0.05, 0.0)
    got = (r.position, r.scale, r.rotation, r.opacity, r.sh, r.sh_rest_scale, r.trust_region)
    assert np.allclose(got, want, rtol=1e-6)


def test_struct_layouts_match_the_header():
    assert C.sizeof(vsplat.CCamera) == 4 * (16 + 3 + 4 + 2)
    assert C.sizeof(vsplat.CCloud) == 5 * 8 + 8
    assert C.sizeof(vsplat.COptions) == 4 * 12 + 8 + 16  # sh_storage, recompute_transmittance, tile_shape, padding
    assert C.sizeof(vsplat.CStats) == 4 * (3 + 2 + 12 + 6)


def _create(cloud, options=None):
    arrays = [np.ascontiguousarray(a, np.float32) for a in (cloud.positions, cloud.scales, cloud.rotations, cloud.opacities, cloud.sh)]
    cc = vsplat.CCloud(*[a.ctypes.data for a in arrays], cloud.count, cloud.sh_degree)
    handle = C.c_void_p()
    code = vsplat.lib().vsplat_trainer_create(C.byref(cc), None if options is None else C.byref(options), C.byref(handle))
    return code, handle


def test_argument_errors_come_before_any_device_call():
    lib = vsplat.lib()
    handle = C.c_void_p()
    assert lib.vsplat_trainer_create(None, None, C.byref(handle)) == vsplat.INVALID_ARGUMENT
    code, _ = _create(F.Cloud.zeros(0, 0))
    assert code == vsplat.RUNTIME_ERROR and b"empty" in lib.vsplat_last_error()
    code, _ = _create(F.Cloud.zeros(4, 4))
    assert code == vsplat.INVALID_ARGUMENT
    o = vsplat.TrainerOptions(position_noise=-0.1).to_c()
    code, _ = _create(F.random_cloud(4, 0, 1), o)
    assert code == vsplat.INVALID_ARGUMENT
    assert lib.vsplat_render(None, None, None) == vsplat.INVALID_ARGUMENT
    assert lib.vsplat_render_noalloc(None, None, None, 0) == vsplat.INVALID_ARGUMENT
    assert lib.vsplat_gaussian_count(None) == 0 and lib.vsplat_uses_fixed_point(None) == 0
    lib.vsplat_trainer_destroy(None)


def test_python_errors_map_to_reference_exceptions():
    with pytest.raises(RuntimeError, match="empty"):
        vsplat.Trainer(F.Cloud.zeros(0, 1))
    with pytest.raises(ValueError):
        vsplat.Trainer(F.random_cloud(4, 0, 1), vsplat.TrainerOptions(position_noise=-1.0))
    with pytest.raises(ValueError):
        vsplat.math_probe(9, np.zeros(4, np.float32))


def test_last_error_is_per_thread():
    import threading
    lib = vsplat.lib()
    _create(F.Cloud.zeros(0, 0))
    mine = lib.vsplat_last_error()
    seen = []

    def other():
        seen.append(lib.vsplat_last_error())

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert b"empty" in mine and seen == [b""]


WRAPPER_PROGRAM = r"""
#include "vsplat_trainer.hpp"
#include <cstdio>
#include <stdexcept>
#include <type_traits>
int main() {
  vsplat::GaussianCloud empty;
  try { vsplat::Trainer t(empty); return 1; } catch (const std::runtime_error &e) { std::printf("empty:%s\n", e.what()); }
  vsplat::GaussianCloud c;
  c.shDegree = 0;
  c.positions = {0.f, 0.f, 0.f};
  c.scales = {-3.f, -3.f, -3.f};
  c.rotations = {1.f, 0.f, 0.f, 0.f};
  c.opacities = {0.f};
  c.sh = {0.f, 0.f};  // one float short
  try { vsplat::Trainer t(c); return 2; } catch (const std::runtime_error &e) { std::printf("sizes:%s\n", e.what()); }
  c.sh.push_back(0.f);
  vsplat::TrainerOptions o;
  o.positionNoise = -1.f;
  try { vsplat::Trainer t(c, o); return 3; } catch (const std::invalid_argument &e) { std::printf("invalid\n"); }
  static_assert(std::is_base_of<std::logic_error, vsplat::Unsupported>::value, "");
  static_assert(vsplat::Trainer::kBytesPerTileEntry == 16, "");
  vsplat::LearningRates r;
  std::printf("rates:%g %g\n", r.position, r.shRestScale);
  return 0;
}
"""


@pytest.mark.parametrize("std", ["c++17", "c++20"])
def test_cpp_wrapper_maps_statuses_to_reference_exceptions(tmp_path, std):
    """include/vsplat_trainer.hpp compiles as C++17 and C++20, links against the
    library, and turns the statuses the C interface returns before any device
    call into the reference's exception types."""
    import subprocess
    src = tmp_path / "wrapper.cpp"
    src.write_text(WRAPPER_PROGRAM)
    exe = tmp_path / "wrapper"
    libdir = Path(vsplat.__file__).resolve().parent
    subprocess.run(["g++", f"-std={std}", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
                    f"-L{libdir}", "-lvsplat", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert "empty:" in out and "sizes:" in out and "invalid" in out and "rates:0.00016 0.05" in out

"""3DGS .ply files through the library (vsplat_ply_info, vsplat_read_ply,
vsplat_write_ply; src/splat/gaussian_cloud.cpp:95-365), checked against the
oracle's restatement (oracle/ply.py) and the reference's own cases
(tests/test_ply.cpp). Host code only: no GPU."""
import numpy as np
import pytest

from oracle import ply as P
from vsplat import fixtures as F
import vsplat


def _cloud(degree, n=37):
    # tests/test_ply.cpp:20-30
    c = F.Cloud.zeros(n, degree)
    c.positions[:] = (np.arange(c.positions.size, dtype=np.float32) * 0.25).reshape(c.positions.shape)
    c.scales[:] = (-np.arange(c.scales.size, dtype=np.float32) * 0.01).reshape(c.scales.shape)
    c.rotations[:] = (np.arange(c.rotations.size) % 5).astype(np.float32).reshape(c.rotations.shape)
    c.opacities[:] = np.arange(n, dtype=np.float32) - 10.0
    c.sh[:] = np.sin(np.arange(c.sh.size, dtype=np.float32)).reshape(c.sh.shape)
    return c


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_round_trip_every_attribute(tmp_path, degree):
    c = _cloud(degree)
    path = tmp_path / "roundtrip.ply"
    vsplat.save_ply(c, path)
    # The bytes are the reference writer's, as the oracle restates it.
    assert path.read_bytes() == P.write_bytes(c.positions, c.scales, c.rotations, c.opacities, c.sh, degree)
    loaded = vsplat.load_ply(path)
    assert loaded.sh_degree == degree and loaded.count == c.count
    for name in ("positions", "scales", "rotations", "opacities", "sh"):
        np.testing.assert_array_equal(_bits(getattr(loaded, name)), _bits(getattr(c, name)))


def test_f_rest_is_channel_major(tmp_path):
    # tests/test_ply.cpp:47-76, the reference's known answer.
    path = tmp_path / "ascii.ply"
    head = "ply\nformat ascii 1.0\ncomment written by a test\nelement vertex 1\n"
    head += "".join(f"property float {p}\n" for p in ("x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2"))
    head += "".join(f"property float f_rest_{i}\n" for i in range(9))
    head += "".join(f"property float {p}\n" for p in ("opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"))
    head += "element face 0\nproperty list uchar int vertex_indices\nend_header\n"
    path.write_text(head + "1 2 3  0.1 0.2 0.3  10 11 12 20 21 22 30 31 32  0.5  -1 -2 -3  1 0 0 0\n")
    c = vsplat.load_ply(path)
    assert c.count == 1 and c.sh_degree == 1
    np.testing.assert_array_equal(c.positions[0], [1, 2, 3])
    assert c.sh[0, 0, 0] == np.float32(0.1) and c.sh[0, 1, 0] == 10 and c.sh[0, 1, 1] == 20 and c.sh[0, 3, 2] == 32
    assert c.opacities[0] == np.float32(0.5)
    np.testing.assert_array_equal(c.scales[0], [-1, -2, -3])
    want = P.read(path)
    np.testing.assert_array_equal(c.sh, want[4])


def test_other_scalar_types_and_extra_properties(tmp_path):
    # Doubles, bytes and shorts as the reference converts them (gaussian_cloud.cpp:60-81),
    # an unknown property in the middle of the record.
    rng = np.random.default_rng(4)
    n = 50
    names = [("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("red", "u1"), ("f_dc_0", "<f4"), ("f_dc_1", "<f4"), ("f_dc_2", "<f4"),
             ("opacity", "<i2"), ("scale_0", "<f4"), ("scale_1", "<f4"), ("scale_2", "<f4"), ("rot_0", "<f4"), ("rot_1", "<f4"),
             ("rot_2", "<f4"), ("rot_3", "<u4")]
    rec = np.zeros(n, np.dtype(names))
    for name, t in names:
        rec[name] = rng.integers(0, 200, n) if t in ("u1", "<i2", "<u4") else rng.standard_normal(n)
    ptype = {"<f8": "double", "u1": "uchar", "<f4": "float", "<i2": "short", "<u4": "uint"}
    head = f"ply\nformat binary_little_endian 1.0\nelement vertex {n}\n" + "".join(f"property {ptype[t]} {k}\n" for k, t in names)
    path = tmp_path / "types.ply"
    path.write_bytes((head + "end_header\n").encode() + rec.tobytes())
    c = vsplat.load_ply(path)
    want = P.read(path)
    for got, w in zip((c.positions, c.scales, c.rotations, c.opacities, c.sh), want[:5]):
        np.testing.assert_array_equal(_bits(got), _bits(w))
    np.testing.assert_array_equal(c.positions[:, 0], rec["x"].astype(np.float32))


def test_files_that_are_not_splats_are_refused(tmp_path):
    mesh = tmp_path / "mesh.ply"  # tests/test_ply.cpp:78-87
    mesh.write_text("ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nproperty float z\nend_header\n0 0 0\n")
    with pytest.raises(RuntimeError, match="missing property"):
        vsplat.load_ply(mesh)
    other = tmp_path / "other.txt"
    other.write_text("not a ply\n")
    with pytest.raises(RuntimeError, match="not a PLY file"):
        vsplat.load_ply(other)
    c = _cloud(1, 5)
    good = tmp_path / "good.ply"
    vsplat.save_ply(c, good)
    cut = tmp_path / "cut.ply"
    cut.write_bytes(good.read_bytes()[:-7])
    with pytest.raises(RuntimeError, match="truncated"):
        vsplat.load_ply(cut)
    with pytest.raises(RuntimeError, match="Cannot open"):
        vsplat.load_ply(tmp_path / "absent.ply")


_SPLAT_PROPS = "".join(f"property float {p}\n" for p in ("x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1",
                                                            "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"))


@pytest.mark.parametrize("header, message", [
    # The header rules of gaussian_cloud.cpp:96-170, each with its message.
    ("ply\nformat binary_big_endian 1.0\nelement vertex 1\n" + _SPLAT_PROPS, "unsupported PLY format 'binary_big_endian'"),
    ("ply\nformat ascii 1.0\nelement face 2\nproperty list uchar int vertex_indices\nelement vertex 1\n" + _SPLAT_PROPS,
     "element 'face' before the vertex element is not supported"),
    ("ply\nformat ascii 1.0\nelement vertex 1\n" + _SPLAT_PROPS + "element vertex 1\n", "more than one vertex element"),
    ("ply\nformat ascii 1.0\nelement vertex 1\nproperty list uchar int idx\n" + _SPLAT_PROPS,
     "list properties on vertices are not supported"),
    ("ply\nformat ascii 1.0\nelement vertex 1\nproperty half x\n" + _SPLAT_PROPS, "unknown property type 'half'"),
    ("ply\nelement vertex 1\n" + _SPLAT_PROPS, "PLY header has no format or vertex element"),
    ("ply\nformat ascii 1.0\nelement vertex 1\n" + _SPLAT_PROPS + "property float f_rest_0\n",
     "unexpected number of f_rest properties \\(1\\)"),
])
def test_header_errors_carry_the_reference_messages(tmp_path, header, message):
    path = tmp_path / "bad.ply"
    path.write_text(header + "end_header\n" + " ".join(["0"] * 14) + "\n")
    with pytest.raises(RuntimeError, match=message):
        vsplat.load_ply(path)


def test_empty_elements_before_the_vertices_and_ascii_layout_are_accepted(tmp_path):
    # An empty element before the vertices is allowed; ASCII values may wrap
    # lines, carry a '+' sign, and properties after the vertex element are skipped.
    path = tmp_path / "loose.ply"
    head = "ply\r\nformat ascii 1.0\r\nelement camera 0\r\nelement vertex 2\r\n" + _SPLAT_PROPS.replace("\n", "\r\n")
    head += "element face 0\r\nproperty list uchar int vertex_indices\r\nend_header\r\n"
    body = "1 2 3 +0.5 0.25\n-0.125 4 -1 -2 -3 1 0 0 0\n\n  5 6 7 1e-1 2E1 3 0.5 -4 -5 -6 0 1 0 0"
    path.write_bytes((head + body).encode())
    c = vsplat.load_ply(path)
    assert c.count == 2 and c.sh_degree == 0
    np.testing.assert_array_equal(c.positions, [[1, 2, 3], [5, 6, 7]])
    np.testing.assert_array_equal(c.sh[:, 0], np.array([[0.5, 0.25, -0.125], [0.1, 20, 3]], np.float32))
    np.testing.assert_array_equal(c.opacities, [4, 0.5])
    np.testing.assert_array_equal(c.rotations, [[1, 0, 0, 0], [0, 1, 0, 0]])
    cut = tmp_path / "cut.ply"
    cut.write_bytes((head + body.rsplit(" ", 3)[0]).encode())
    with pytest.raises(RuntimeError, match="truncated"):
        vsplat.load_ply(cut)


def test_large_binary_file_decoded_over_threads(tmp_path):
    # More than one decoding thread's share of records (65536 each).
    c = _cloud(3, 200_003)
    path = tmp_path / "large.ply"
    vsplat.save_ply(c, path)
    assert path.read_bytes() == P.write_bytes(c.positions, c.scales, c.rotations, c.opacities, c.sh, 3)
    loaded = vsplat.load_ply(path)
    for name in ("positions", "scales", "rotations", "opacities", "sh"):
        np.testing.assert_array_equal(_bits(getattr(loaded, name)), _bits(getattr(c, name)))

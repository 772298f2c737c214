"""Test infrastructure: a numpy restatement of the reference's 3DGS .ply
reader and writer (src/splat/gaussian_cloud.cpp:95-365), the checker for the
library's vsplat_read_ply / vsplat_write_ply. Pinned by tests/test_ply.py:
the reference's own known answer for the channel-major f_rest order
(tests/test_ply.cpp:47-76), the round trip of every attribute at every SH
degree (tests/test_ply.cpp:20-45), and the rejection of a mesh file
(tests/test_ply.cpp:78-87).
"""
from __future__ import annotations

import numpy as np

# gaussian_cloud.cpp:34-44
TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "<i2", "int16": "<i2", "ushort": "<u2",
         "uint16": "<u2", "int": "<i4", "int32": "<i4", "uint": "<u4", "uint32": "<u4", "float": "<f4", "float32": "<f4",
         "double": "<f8", "float64": "<f8"}
DEGREE_OF_REST = {0: 0, 9: 1, 24: 2, 45: 3}


def _header(data: bytes):
    """(binary, vertex count, [(name, type)], offset of the first record)."""
    if not data.startswith(b"ply"):
        raise ValueError("not a PLY file")
    pos, binary, count, props, in_vertex, saw_vertex, saw_format = 0, True, 0, [], False, False, False
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end].decode("ascii").rstrip("\r")
        pos = end + 1
        words = line.split()
        if not words or words[0] in ("ply", "comment", "obj_info"):
            continue
        if words[0] == "end_header":
            break
        if words[0] == "format":
            if words[1] not in ("binary_little_endian", "ascii"):
                raise ValueError(f"unsupported PLY format '{words[1]}'")
            binary, saw_format = words[1] == "binary_little_endian", True
        elif words[0] == "element":
            if words[1] == "vertex":
                in_vertex, saw_vertex, count = True, True, int(words[2])
            else:
                if not saw_vertex and int(words[2]) > 0:
                    raise ValueError("element before the vertex element")
                in_vertex = False
        elif words[0] == "property" and in_vertex:
            if words[1] == "list":
                raise ValueError("list properties on vertices are not supported")
            props.append((words[2], TYPES[words[1]]))
    if not (saw_format and saw_vertex):
        raise ValueError("PLY header has no format or vertex element")
    return binary, count, props, pos


def read(path):
    """(positions, scales, rotations, opacities, sh, degree) as float32 arrays
    in GaussianCloud's layout, SH (N, coefficients, 3)."""
    data = open(path, "rb").read()
    binary, count, props, start = _header(data)
    names = [n for n, _ in props]
    if binary:
        rec = np.frombuffer(data, np.dtype([(n, t) for n, t in props]), count, start)
        col = {n: rec[n].astype(np.float32) for n in names}
    else:
        values = np.array(data[start:].split()[: count * len(props)], np.float64).reshape(count, len(props))
        col = {n: values[:, i].astype(np.float32) for i, n in enumerate(names)}
    rest = 0
    while f"f_rest_{rest}" in col:
        rest += 1
    if rest not in DEGREE_OF_REST:
        raise ValueError(f"unexpected number of f_rest properties ({rest})")
    for n in ("x", "y", "z", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3", "opacity",
              "f_dc_0", "f_dc_1", "f_dc_2"):
        if n not in col:
            raise ValueError(f"missing property '{n}'")
    degree = DEGREE_OF_REST[rest]
    coeffs = (degree + 1) ** 2
    sh = np.zeros((count, coeffs, 3), np.float32)
    for c in range(3):
        sh[:, 0, c] = col[f"f_dc_{c}"]
        for k in range(1, coeffs):  # f_rest is channel-major in the file
            sh[:, k, c] = col[f"f_rest_{c * (coeffs - 1) + k - 1}"]
    stack = lambda ns: np.stack([col[n] for n in ns], 1)
    return (stack(["x", "y", "z"]), stack(["scale_0", "scale_1", "scale_2"]), stack(["rot_0", "rot_1", "rot_2", "rot_3"]),
            col["opacity"].copy(), sh, degree)


def write_bytes(positions, scales, rotations, opacities, sh, degree) -> bytes:
    """The file save_ply writes (gaussian_cloud.cpp:330-365)."""
    n = len(opacities)
    coeffs = (degree + 1) ** 2
    rest = 3 * (coeffs - 1)
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(rest)] + \
            ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    header = f"ply\nformat binary_little_endian 1.0\nelement vertex {n}\n" + "".join(f"property float {x}\n" for x in names) + \
             "end_header\n"
    sh = np.asarray(sh, np.float32).reshape(n, coeffs, 3)
    cols = [np.asarray(positions, np.float32).reshape(n, 3), np.zeros((n, 3), np.float32), sh[:, 0, :]]
    cols += [sh[:, 1:, c] for c in range(3)]
    cols += [np.asarray(opacities, np.float32).reshape(n, 1), np.asarray(scales, np.float32).reshape(n, 3),
             np.asarray(rotations, np.float32).reshape(n, 4)]
    return header.encode() + np.ascontiguousarray(np.concatenate(cols, 1), "<f4").tobytes()

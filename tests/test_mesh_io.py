"""Mesh file readers (gp_che_read) with the reference's face conventions
(src/che_off.cpp:28-79, src/che_obj.cpp:24-79, src/che_ply.cpp:25-156):
small files written here, the faces each reader must produce derived by hand,
the CPU PTP of a read mesh equal to the PTP of the same arrays, and (GPU)
PTP_GPU of an OFF and a binary PLY file against the FP64 oracle."""
import struct

import numpy as np
import pytest

from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

SQUARE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0], [2, 1, 0]], float)


def _faces(C):
    return C.tables()[0].reshape(-1, 3).tolist()


def test_off_triangles_and_attribute_variants(tmp_path):
    tri = "3 0 1 2\n3 0 2 3\n"
    body = "".join(f"{x} {y} {z}\n" for x, y, z in SQUARE[:4])
    (tmp_path / "a.off").write_text("OFF\n4 2 0\n" + body + tri)
    (tmp_path / "c.off").write_text("COFF\n4 2 0\n" + "".join(f"{x} {y} {z} 255 0 0 255\n" for x, y, z in SQUARE[:4]) + tri)
    (tmp_path / "n.off").write_text("NOFF\n4 2 0\n" + "".join(f"{x} {y} {z} 0 0 1\n" for x, y, z in SQUARE[:4]) + tri)
    for name in ("a.off", "c.off", "n.off"):
        C = P.che.read(tmp_path / name)
        assert _faces(C) == [[0, 1, 2], [0, 2, 3]]
        assert np.array_equal(C.gt(), SQUARE[:4])


def test_off_quads_split_by_index_reuse(tmp_path):
    # (a, b, c, d) -> (a, b, c), (d, a, c)   (che_off.cpp:62-72)
    body = "".join(f"{x} {y} {z}\n" for x, y, z in SQUARE)
    (tmp_path / "q.off").write_text("OFF\n6 2 0\n" + body + "4 0 1 2 3\n4 1 4 5 2\n")
    assert _faces(P.che.read(tmp_path / "q.off")) == [[0, 1, 2], [3, 0, 2], [1, 4, 5], [2, 1, 5]]
    (tmp_path / "mixed.off").write_text("OFF\n6 2 0\n" + body + "4 0 1 2 3\n3 1 4 5\n")
    with pytest.raises(P.GPError) as e:
        P.che.read(tmp_path / "mixed.off")
    assert e.value.status == L.ERR_INVALID


def test_obj_forms_and_quads(tmp_path):
    # "i/t/n" forms read to the first '/'; a quad (a, b, c, d) -> (a, b, d),
    # (b, c, d) (che_obj.cpp:62-72); a pentagon is skipped.
    text = "# comment\n" + "".join(f"v {x} {y} {z}\n" for x, y, z in SQUARE) + \
           "vt 0 0\nvn 0 0 1\nf 1/1/1 2/2/1 3/3/1 4/4/1\nf 2//1 5//1 6//1\nf 1 2 3 4 5\n"
    (tmp_path / "m.obj").write_text(text)
    C = P.che.read(tmp_path / "m.obj")
    assert _faces(C) == [[0, 1, 3], [1, 2, 3], [1, 4, 5]]
    (tmp_path / "bad.obj").write_text("v 0 0 0\nf 1 x 2\n")
    with pytest.raises(P.GPError):
        P.che.read(tmp_path / "bad.obj")


def _ply_binary(path, vtype, ltype, itype, xyz, faces, extra=False):
    fmt = {"float": "f", "double": "d"}[vtype]
    lf = {"uchar": "B", "int": "i"}[ltype]
    fi = {"int": "i", "uint": "I"}[itype]
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {len(xyz)}",
            f"property {vtype} x", f"property {vtype} y", f"property {vtype} z"]
    if extra:
        head.append("property uchar red")
    head += [f"element face {len(faces)}", f"property list {ltype} {itype} vertex_indices", "end_header"]
    data = ("\n".join(head) + "\n").encode()
    for p in xyz:
        data += struct.pack("<" + fmt * 3, *p) + (b"\x07" if extra else b"")
    for f in faces:
        data += struct.pack("<" + lf, 3) + struct.pack("<" + fi * 3, *f)
    path.write_bytes(data)


def test_ply_ascii_and_binary(tmp_path):
    faces = [[0, 1, 2], [0, 2, 3]]
    head = "ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n" \
           "element face 2\nproperty list uchar int vertex_indices\nend_header\n"
    (tmp_path / "a.ply").write_text(head + "".join(f"{x} {y} {z}\n" for x, y, z in SQUARE[:4]) +
                                    "3 0 1 2\n3 0 2 3\n")
    assert _faces(P.che.read(tmp_path / "a.ply")) == faces
    for k, (vt, lt, it, ex) in enumerate([("float", "uchar", "int", False), ("double", "int", "uint", True)]):
        path = tmp_path / f"b{k}.ply"
        _ply_binary(path, vt, lt, it, SQUARE[:4], faces, ex)
        C = P.che.read(path)
        assert _faces(C) == faces
        assert np.allclose(C.gt(), SQUARE[:4])


def test_read_mesh_runs_ptp_like_the_arrays(tmp_path, oracle):
    xyz, f = G.height_field(24, 4)
    f3 = f.reshape(-1, 3)
    (tmp_path / "hf.off").write_text(f"OFF\n{len(xyz)} {len(f3)} 0\n" +
                                     "".join(f"{float(x)!r} {float(y)!r} {float(z)!r}\n" for x, y, z in xyz) +
                                     "".join(f"3 {a} {b} {c}\n" for a, b, c in f3))
    C = P.che.read(tmp_path / "hf.off")
    assert np.array_equal(C.gt(), xyz)
    ref = P.che(xyz, f)
    ts = ref.compute_toplesets([10])
    assert np.array_equal(P.parallel_toplesets_propagation_cpu(C, [10], ts),
                          P.parallel_toplesets_propagation_cpu(ref, [10], ts))


def _write_off(path, xyz, f3):
    with open(path, "w") as fh:
        fh.write(f"OFF\n{len(xyz)} {len(f3)} 0\n")
        np.savetxt(fh, xyz, fmt="%.17g")
        np.savetxt(fh, np.hstack([np.full((len(f3), 1), 3), f3]), fmt="%d")


def _write_ply_binary(path, xyz, f3):
    with open(path, "wb") as fh:
        fh.write((f"ply\nformat binary_little_endian 1.0\nelement vertex {len(xyz)}\nproperty double x\n"
                  f"property double y\nproperty double z\nelement face {len(f3)}\n"
                  f"property list uchar int vertex_indices\nend_header\n").encode())
        xyz.astype("<f8").tofile(fh)
        rec = np.zeros(len(f3), dtype=[("n", "u1"), ("v", "<i4", 3)])
        rec["n"] = 3
        rec["v"] = f3
        rec.tofile(fh)


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["off", "ply"])
def test_read_mesh_ptp_gpu_against_oracle(tmp_path, oracle, fmt):
    # A 128^2 height field written as an OFF text file (17 significant digits)
    # and as a binary little-endian PLY file, read back by gp_che_read, then
    # PTP_GPU of the front end (FP64 by default) against the FP64 oracle run on
    # the arrays that were written: bit-identical, as the reader reproduces
    # positions and faces exactly.
    xyz, f = G.height_field(128, 41)
    f3 = f.reshape(-1, 3)
    src = [4321, 9876]
    path = tmp_path / f"hf.{fmt}"
    (_write_off if fmt == "off" else _write_ply_binary)(path, xyz, f3)
    C = P.che.read(path)
    assert np.array_equal(C.gt(), xyz) and np.array_equal(C.tables()[0].reshape(-1, 3), f3)
    g = P.geodesics(C, src, P.geodesics.PTP_GPU)
    ref, _ = oracle.Mesh(xyz, f).ptp(src, oracle.GPU_PLAIN)
    print(f"\n[read {fmt}] V={len(xyz)} identical={np.array_equal(g.dist, ref)}", flush=True)
    assert np.array_equal(g.dist.view(np.uint64), ref.view(np.uint64))
    g32 = P.geodesics(C, src, P.geodesics.PTP_GPU, precision=L.PREC_F32)
    m = ref > 0
    assert np.max(np.abs(g32.dist[m] - ref[m]) / ref[m]) <= 1e-4


def test_number_forms_and_line_endings(tmp_path):
    # The OFF and ASCII PLY numbers are parsed from memory: each must read as
    # Python's correctly rounded float() of the same text (a leading '+',
    # '.5', '5.', exponents, -0, subnormals, an underflow to 0), CRLF line
    # ends change nothing, and text that is not a finite number is an error.
    words = [["+1.5", ".5", "5."], ["1e-400", "-0", "2.5E+3"],
             ["4.9406564584124654e-324", "1e-310", "-7"], ["0.1", "0.2", "0.30000000000000004"]]
    want = np.array([[float(w) for w in row] for row in words])
    body = "".join(" ".join(row) + "\n" for row in words)
    for name, eol in (("lf.off", "\n"), ("crlf.off", "\r\n")):
        (tmp_path / name).write_bytes(("OFF\n4 2 0\n" + body + "3 0 1 2\n3 0 2 3\n").replace("\n", eol).encode())
        C = P.che.read(tmp_path / name)
        assert np.array_equal(C.gt().view(np.uint64), want.view(np.uint64))
        assert _faces(C) == [[0, 1, 2], [0, 2, 3]]
    head = "ply\nformat ascii 1.0\nelement vertex 4\nproperty double x\nproperty double y\nproperty double z\n" \
           "property uchar red\nelement face 2\nproperty list uchar int vertex_indices\nend_header\n"
    ply = head + "".join(" ".join(row) + " 255\n" for row in words) + "3 0 1 2 9\n3 0 2 3\n"
    (tmp_path / "crlf.ply").write_bytes(ply.replace("\n", "\r\n").encode())
    C = P.che.read(tmp_path / "crlf.ply")
    assert np.array_equal(C.gt().view(np.uint64), want.view(np.uint64))
    assert _faces(C) == [[0, 1, 2], [0, 2, 3]]
    for bad in ("inf 0 0", "nan 0 0", "1e400 0 0", "1.5x 0 0", "0 0"):
        (tmp_path / "bad.off").write_text("OFF\n3 1 0\n" + bad + "\n1 0 0\n0 1 0\n3 0 1 2\n")
        with pytest.raises(P.GPError) as e:
            P.che.read(tmp_path / "bad.off")
        assert e.value.status == L.ERR_INVALID, bad
    # A binary PLY cut inside its faces is an error, not a short read.
    _ply_binary(tmp_path / "cut.ply", "double", "uchar", "int", SQUARE[:4], [[0, 1, 2], [0, 2, 3]])
    data = (tmp_path / "cut.ply").read_bytes()
    (tmp_path / "cut.ply").write_bytes(data[:-5])
    with pytest.raises(P.GPError) as e:
        P.che.read(tmp_path / "cut.ply")
    assert e.value.status == L.ERR_INVALID


def test_unknown_extension_and_missing_file(tmp_path):
    (tmp_path / "m.stl").write_text("solid\n")
    for p in (tmp_path / "m.stl", tmp_path / "none.off"):
        with pytest.raises(P.GPError) as e:
            P.che.read(p)
        assert e.value.status == L.ERR_INVALID

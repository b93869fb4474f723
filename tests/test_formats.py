"""Mesh and texture file formats (SimpleMesh.h:118-160, VectorIO.h, Padding.h) and the apps'
mesh preparation, on the reference's own Julius mesh (tests/golden/julius_mesh.npz)."""
import os

import numpy as np
import pytest

import gradientdomaintextureprocessing_amd as g

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_PLY = "/root/reference/Data/Julius/julius.ply"


def _ascii_ply(path, v, t, uv):
    with open(path, "w") as fh:
        fh.write("ply\nformat ascii 1.0\nelement vertex %d\nproperty float x\nproperty float y\nproperty float z\n" % len(v))
        fh.write("element face %d\nproperty list uchar int vertex_indices\nproperty list uchar float texcoord\nend_header\n" % len(t))
        for p in v:
            fh.write("%r %r %r\n" % tuple(float(np.float32(c)) for c in p))
        for tri, tc in zip(t, uv):
            fh.write("3 %d %d %d 6 %s\n" % (*tri, " ".join(repr(float(np.float32(c))) for c in tc.ravel())))


def _binary_ply(path, v, t, uv):
    with open(path, "wb") as fh:
        fh.write(("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty double x\nproperty double y\nproperty double z\n"
                  "property uchar flags\nelement face %d\nproperty list uchar int vertex_indices\nproperty list uchar double texcoord\n"
                  "end_header\n" % (len(v), len(t))).encode())
        for p in v:
            fh.write(np.asarray(p, "<f8").tobytes() + b"\x07")
        for tri, tc in zip(t, uv):
            fh.write(b"\x03" + np.asarray(tri, "<i4").tobytes() + b"\x06" + np.asarray(tc.ravel(), "<f8").tobytes())


def test_ply_round_trips(tmp_path):
    d = np.load(os.path.join(GOLDEN, "julius_mesh.npz"))
    v, t, uv = d["vertices"][:200], d["triangles"][:50] % 200, d["corner_uv"][:50]
    _binary_ply(tmp_path / "b.ply", v, t, uv)
    vb, tb, ub = g.TexturedMesh.read_ply(tmp_path / "b.ply").arrays()
    assert np.array_equal(vb, v) and np.array_equal(tb, t) and np.array_equal(ub, uv)
    _ascii_ply(tmp_path / "a.ply", v, t, uv)
    va, ta, ua = g.TexturedMesh.read_ply(tmp_path / "a.ply").arrays()
    assert np.array_equal(ta, t) and np.allclose(va, v.astype(np.float32)) and np.allclose(ua, uv.astype(np.float32))


@pytest.mark.skipif(not os.path.exists(REF_PLY), reason="reference data not present")
def test_reference_ply_matches_fixture():
    v, t, uv = g.TexturedMesh.read_ply(REF_PLY).arrays()
    d = np.load(os.path.join(GOLDEN, "julius_mesh.npz"))
    assert np.array_equal(v, d["vertices"]) and np.array_equal(t, d["triangles"]) and np.array_equal(uv, d["corner_uv"])
    assert len(v) == 5076 and len(t) == 9999


def test_bad_ply_fails_loudly(tmp_path):
    (tmp_path / "x.ply").write_text("ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n")
    with pytest.raises(g.api._lib.GdtpError):
        g.TexturedMesh.read_ply(tmp_path / "x.ply")
    with pytest.raises(g.api._lib.GdtpError):
        g.TexturedMesh.read_ply(tmp_path / "missing.ply")


def test_padding_rules():
    # coordinates already inside the texel-centre box: only the multiple-of-8 rounding applies
    p = g.compute_padding(1000, 1000, np.array([[0.3, 0.4], [0.6, 0.7]]))
    assert (p.left, p.bottom, p.right, p.top) == (0, 0, 0, 0)
    p = g.compute_padding(1001, 1003, np.array([[0.3, 0.4], [0.6, 0.7]]))
    assert (1001 + p.left + p.right) % 8 == 0 and (1003 + p.bottom + p.top) % 8 == 0
    # a coordinate 2.3 texels left of the first texel centre needs 3 texels (before rounding)
    p = g.compute_padding(64, 64, np.array([[-1.8 / 64, 0.5], [0.5, 0.5]]))
    assert p.left >= 3 and (64 + p.left + p.right) % 8 == 0
    uv = g.pad_texture_coordinates(p, 64, 64, np.array([[-1.8 / 64, 0.5]]))
    W = 64 + p.left + p.right
    assert uv[0, 0] * W >= 0.5  # inside the texel-centre box of the padded atlas
    img = np.arange(12.0).reshape(3, 4, 1)
    pad = g.Padding(1, 2, 3, 0)
    assert np.array_equal(g.unpad_image(pad, g.pad_image(pad, img)), img)


def test_vector_and_image_files(tmp_path):
    vf = np.random.default_rng(0).standard_normal((17, 2))
    g.write_vector(tmp_path / "a.vf", vf)
    assert np.array_equal(g.read_vector(tmp_path / "a.vf", np.dtype((np.float64, 2))), vf)
    assert os.path.getsize(tmp_path / "a.vf") == 4 + 17 * 16
    im = np.random.default_rng(1).random((5, 7, 3)).astype(np.float32)
    g.write_binary_image(tmp_path / "a.normap", im)
    assert np.array_equal(g.read_binary_image(tmp_path / "a.normap"), im)


def test_julius_hierarchy_and_operator():
    d = np.load(os.path.join(GOLDEN, "julius_mesh.npz"))
    m = g.TexturedMesh.from_arrays(d["vertices"], d["triangles"], d["corner_uv"])
    v, t, uv = m.arrays()
    uv[..., 1] = 1 - uv[..., 1]
    m = g.TexturedMesh.from_arrays(v, t, uv)
    H = g.InitializeHierarchy(m, 256, 256, 3)
    assert H.level_info(0)["charts"] == 28
    from helpers import mass_matrix, oracle_levels, stiffness_matrix
    S = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=0)
    L0 = oracle_levels(H, S)[0]
    M, K = mass_matrix(H, L0), stiffness_matrix(H, L0)
    assert abs(M.sum() - 1.0) < 1e-9
    assert (np.abs(K @ np.ones(L0.N)) / np.abs(K).sum(axis=1).A1).max() < 1e-11

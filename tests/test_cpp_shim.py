"""The reference-signature C++ shim (include/gdtp_reference.hpp) compiled into a TextureFiltering-
shaped program (tests/cpp/shim_app.cpp) that links libgdtp.so: on the CPU its InitializeHierarchy
fills the apps' texel nodes with the library's atlas coordinates; on the GPU its VCycle and
direct solve equal the Python binding's gdtp_vcycle / gdtp_direct_solve bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import gradientdomaintextureprocessing_amd as g
from gradientdomaintextureprocessing_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "shim_app.cpp")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_shim(out_dir):
    exe = os.path.join(out_dir, "shim_app")
    pkg = os.path.dirname(_lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROCM, "include"), SRC, "-o", exe, "-L" + pkg, "-lgdtp", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath," + pkg, "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "llvm", "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def write_mesh(mesh, path):
    v, t, uv = mesh.arrays()
    with open(path, "wb") as fh:
        np.array([len(v), len(t)], np.int32).tofile(fh)
        np.ascontiguousarray(v, np.float64).tofile(fh)
        np.ascontiguousarray(t, np.int32).tofile(fh)
        np.ascontiguousarray(uv, np.float64).tofile(fh)


def run(exe, mode, mesh_path, width, levels, out_path):
    r = subprocess.run([exe, mode, mesh_path, str(width), str(levels), out_path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    with open(out_path, "rb") as fh:
        n = int(np.fromfile(fh, np.int32, 1)[0])
        nodes = np.fromfile(fh, np.int32, 3 * n).reshape(n, 3)
        rest = np.fromfile(fh, np.float64)
    return n, nodes, rest


def test_shim_compiles_and_initialises_hierarchy(tmp_path):
    exe = build_shim(str(tmp_path))
    mesh = g.TexturedMesh.cube(16, 128, 128)
    write_mesh(mesh, str(tmp_path / "mesh.bin"))
    n, nodes, _ = run(exe, "host", str(tmp_path / "mesh.bin"), 128, 3, str(tmp_path / "out.bin"))
    H = g.InitializeHierarchy(mesh, 128, 128, 3)
    ci, cj, chart = H.texel_coordinates(0)
    assert n == len(ci)
    assert np.array_equal(nodes[:, 0], ci) and np.array_equal(nodes[:, 1], cj) and np.array_equal(nodes[:, 2], chart)


@pytest.mark.gpu
def test_shim_vcycle_and_direct_solve_match_binding(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    exe = build_shim(str(tmp_path))
    mesh = g.TexturedMesh.icosphere(4, 256, 256, 0.05, 3)
    write_mesh(mesh, str(tmp_path / "mesh.bin"))
    print("[shim] icosphere 256^2, 4 levels", flush=True)
    n, _, rest = run(exe, "gpu", str(tmp_path / "mesh.bin"), 256, 4, str(tmp_path / "out.bin"))
    mf, xv, exact = (rest.reshape(3, n, 3)[k].T.copy() for k in range(3))
    H = g.InitializeHierarchy(mesh, 256, 256, 4)
    S = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=3)
    i = np.arange(n)
    f = np.stack([((i * 7 + c * 13) % 101) / 100.0 for c in range(3)])
    fd = torch.from_numpy(f).cuda()
    b = torch.empty_like(fd)
    S.apply(g.OPERATOR_MASS, fd, b)
    assert np.array_equal(b.cpu().numpy(), mf)
    x = fd.clone()
    for _ in range(5):
        S.vcycle(x, b)
    assert np.array_equal(x.cpu().numpy(), xv)
    d = torch.empty_like(b)
    S.direct_solve(d, b)
    assert np.array_equal(d.cpu().numpy(), exact)

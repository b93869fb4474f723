"""How much the solution depends on the triangulation of seam pieces.  The reference triangulates
each boundary-cell polygon with Triangle (constrained Delaunay, BoundaryQuadraticElements.inl:
525-642); this library fans each seam piece from its centroid.  Both are quadratic elements on the
same seam nodes, differing only inside the piece, so the difference between two valid
triangulations bounds how far the library's seam rows are from the reference's model.  Here the
centroid fan is compared with a fan from the first ring node (GDTP_SEAM_FAN=vertex, read once per
process, so that model is built in a child process) on the cube at 256^2: the solutions of
(M + 1e-3 S) x = M f differ by < 5e-4 in the mass norm and at deep texels; the vertex fan still
integrates linear functions across seams exactly."""
import os
import subprocess
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import cube_system, mass_matrix, oracle_levels, stiffness_matrix

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
sys.path[:0] = [{root!r}, {root!r} + '/tests', {root!r} + '/oracle']
from helpers import cube_system, mass_matrix, oracle_levels, stiffness_matrix
import test_seams
mesh, H, S = cube_system(256, 2)
L0 = oracle_levels(H, S)[0]
M, K = mass_matrix(H, L0).tocoo(), stiffness_matrix(H, L0).tocoo()
pos = test_seams._texel_positions(mesh, H)
a, c0 = np.array([0.3, -0.7, 0.45]), 0.2
f = pos @ a + c0
energy, m1, m2 = test_seams._surface_integrals(mesh, a, c0)
assert abs(f @ (K.tocsr() @ f) - energy) < 1e-9 * energy
assert abs(f @ (M.tocsr() @ f) - m2) < 1e-11
np.savez({out!r}, mr=M.row, mc=M.col, mv=M.data, sr=K.row, sc=K.col, sv=K.data)
"""


def test_seam_triangulation_changes_solution_little(tmp_path):
    out = str(tmp_path / "vertex_fan.npz")
    env = dict(os.environ, GDTP_SEAM_FAN="vertex")
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=ROOT, out=out)], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    _, H, S = cube_system(256, 2)
    L0 = oracle_levels(H, S)[0]
    Mc, Sc = mass_matrix(H, L0).tocsr(), stiffness_matrix(H, L0).tocsr()
    d = np.load(out)
    n = L0.N
    Mv = sp.csr_matrix((d["mv"], (d["mr"], d["mc"])), shape=(n, n))
    Sv = sp.csr_matrix((d["sv"], (d["sr"], d["sc"])), shape=(n, n))
    assert abs(Sc - Sv).max() > 1e-3 * abs(Sc).max()  # the two models do differ at seam rows
    f = np.random.default_rng(0).random((n, 3))
    Ac, Av = (Mc + 1e-3 * Sc).tocsc(), (Mv + 1e-3 * Sv).tocsc()
    xc, xv = spla.splu(Ac).solve(Mc @ f), spla.splu(Av).solve(Mv @ f)
    e = xc - xv
    mnorm = np.sqrt(np.einsum("ij,ij->j", e, Mc @ e) / np.einsum("ij,ij->j", xc, Mc @ xc))
    deep = np.ones(n, bool)
    deep[L0.bgi] = False
    l2_deep = np.linalg.norm(e[deep], axis=0) / np.linalg.norm(xc[deep], axis=0)
    print(f"[seam model] mass-norm difference {mnorm}, deep-texel difference {l2_deep}")
    assert np.all(mnorm < 5e-4) and np.all(l2_deep < 5e-4)

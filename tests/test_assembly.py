"""Finest-level mass / stiffness (include/Src/MassAndStiffness.h) pinned by closed forms and
invariants: textbook Q1 stencils on flat cube faces, unit total area, S 1 = 0, symmetry, SPD."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from helpers import cube_system, icosphere_system, mass_matrix, oracle_levels, stiffness_matrix


@pytest.mark.parametrize("width", [128, 256])
def test_cube_face_stencils_are_textbook_q1(width):
    _, H, _ = cube_system(width, 2)
    s = 16 * np.floor((min(width / 3, width / 2) - 8) / 16)  # face side in texels
    area = (1.0 / s) ** 2 / 6.0                               # one cell, surface area normalised to 1
    mass = H.array(0, "mass.deep", np.float64).reshape(-1, 10)[:, :9]
    stiff = H.array(0, "stiffness.deep", np.float64).reshape(-1, 10)[:, :9]
    q1m = area / 36.0 * np.array([1, 4, 1, 4, 16, 4, 1, 4, 1], float)
    q1s = np.array([-1, -1, -1, -1, 8, -1, -1, -1, -1], float) / 3.0
    assert len(mass) > 100
    assert np.allclose(mass, q1m, rtol=1e-10, atol=0)
    assert np.allclose(stiff, q1s, rtol=0, atol=1e-12)


@pytest.mark.parametrize("make", [lambda: cube_system(128, 2), lambda: icosphere_system(256, 4, 2)])
def test_total_area_nullspace_symmetry(make):
    _, H, S = make()
    L0 = oracle_levels(H, S)[0]
    M, K = mass_matrix(H, L0), stiffness_matrix(H, L0)
    assert abs(M.sum() - 1.0) < 1e-10                      # metric normalised to unit area (Metric.inl:269-273)
    # constants are in the kernel of S, to rounding relative to each row's magnitude (quadratic
    # elements on thin seam pieces carry large entries that cancel)
    assert (np.abs(K @ np.ones(L0.N)) / np.abs(K).sum(axis=1).A1).max() < 1e-11
    assert abs(M - M.T).max() < 1e-17 and abs(K - K.T).max() < 1e-13
    assert np.all(M.diagonal() > 0) and np.all(K.diagonal() >= -1e-15)
    # M is positive definite: smallest eigenvalue of the (small) matrix is positive
    lam = spla.eigsh(M.tocsc(), k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
    assert lam > 0
    # S is positive semi-definite: x^T S x >= 0 for random x
    x = np.random.default_rng(1).standard_normal((4, L0.N))
    assert np.all(np.einsum("ij,ij->i", x, (K @ x.T).T) >= -1e-12)

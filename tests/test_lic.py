"""Line-integral-convolution metric (include/Src/Metric.inl:12-67) and the anisotropic hierarchy.

The metric is pinned by its eigen-structure: the field direction v and its g-orthogonal
complement w are eigen-directions relative to g, with ratio (across + floor) / floor between the
stretch across and along the field; the embedding metric is the edge Gram matrix with unit total
area."""
import numpy as np

import gradientdomaintextureprocessing_amd as g
from helpers import mass_matrix, oracle_levels, stiffness_matrix


def _torus_field():
    mesh = g.TexturedMesh.torus(64, 32, 4, 2, 256, 256, 1.0, 0.35, 0.0, 5)
    v, t, _ = mesh.arrays()
    f = np.sin(3 * v[:, 0]) + np.cos(2 * v[:, 1]) + v[:, 2]
    return mesh, v, t, g.surface_gradient_field(v, t, f)


def test_embedding_metric_is_the_edge_gram_matrix():
    _, v, t, _ = _torus_field()
    G = g.embedding_metric(v, t, normalize_area=False)
    P = v[t]
    assert np.allclose(G[:, 0, 0], np.sum((P[:, 1] - P[:, 0]) ** 2, 1))
    assert np.allclose(G[:, 0, 1], np.sum((P[:, 1] - P[:, 0]) * (P[:, 2] - P[:, 0]), 1))
    Gn = g.embedding_metric(v, t)
    assert abs(np.sum(np.sqrt(np.linalg.det(Gn)) / 2) - 1.0) < 1e-12


def test_vector_field_metric_eigen_structure():
    _, v, t, vf = _torus_field()
    G = g.embedding_metric(v, t)
    A = g.vector_field_metric(G, vf)
    assert np.allclose(A, np.transpose(A, (0, 2, 1)), rtol=0, atol=1e-12 * np.abs(A).max())
    assert np.all(np.linalg.det(A) > 0) and abs(np.sum(np.sqrt(np.linalg.det(A)) / 2) - 1.0) < 1e-12
    gv = np.einsum("tkl,tl->tk", G, vf)
    w = np.stack([-gv[:, 1], gv[:, 0]], 1)  # g-orthogonal complement of v
    along = np.einsum("tk,tkl,tl->t", vf, A, vf) / np.einsum("tk,tkl,tl->t", vf, G, vf)
    across = np.einsum("tk,tkl,tl->t", w, A, w) / np.einsum("tk,tkl,tl->t", w, G, w)
    ok = np.einsum("tk,tk->t", vf, gv) > 1e-12
    assert np.allclose(across[ok] / along[ok], (1e-1 + 1e-5) / 1e-5, rtol=1e-9)
    # v and w stay eigen-directions: A v is parallel to g v
    Av = np.einsum("tkl,tl->tk", A, vf)
    assert np.abs(Av[ok, 0] * gv[ok, 1] - Av[ok, 1] * gv[ok, 0]).max() <= 1e-9 * np.abs(Av[ok]).max() * np.abs(gv[ok]).max()


def test_anisotropic_hierarchy_assembles():
    mesh, v, t, vf = _torus_field()
    G = g.embedding_metric(v, t)
    H = g.InitializeHierarchy(mesh, 256, 256, 3, triangle_metric=g.vector_field_metric(G, vf))
    S = g.UpdateLinearSystem(H, 1e-4, 1.0, max_cols=0)
    L0 = oracle_levels(H, S)[0]
    M, K = mass_matrix(H, L0), stiffness_matrix(H, L0)
    assert abs(M.sum() - 1.0) < 1e-9          # the metric has unit area
    assert (np.abs(K @ np.ones(L0.N)) / np.abs(K).sum(axis=1).A1).max() < 1e-11

"""Normalised-gradient divergence of the heat-method geodesics (include/Src/VectorFieldIntergration.inl).

Pinned by: the integrated basis gradients of every sample sum to zero (partition of unity), and
on the cube, where a linear function of 3D position is reproduced exactly on every chart and
across every seam, f^T rhs(u) for linear u and f equals the closed form
    -(1/sqrt(A)) sum_faces area * (grad u . grad f) / |grad u|
(tangential gradients; A the total area, which the metric normalises to one)."""
import numpy as np
import pytest

import oracle
from helpers import cube_system, icosphere_system
from test_seams import _texel_positions


def test_samples_integrate_to_zero():
    _, H, _ = icosphere_system(256, 4, 2)
    S = oracle.vf_samples(H)
    v = S["v"]
    scale = np.abs(v).max()
    assert np.abs(v.sum(axis=1)).max() < 1e-6 * scale
    seg = np.repeat(np.arange(len(S["gptr"]) - 1), np.diff(S["gptr"]))
    assert len(seg) > 0
    for d in range(2):
        assert np.abs(np.bincount(seg, S["gV"][:, d])).max() < 1e-6 * scale
    # positions inside the unit cell, tensors positive definite
    assert S["pos"].min() >= -1e-6 and S["pos"].max() <= 1 + 1e-6
    t = S["tensor"]
    assert np.all(t[:, 0] > 0) and np.all(t[:, 0] * t[:, 2] - t[:, 1] ** 2 > 0)


def _q1_grad(x, y):
    """Gradients of the bilinear corner functions (1-x)(1-y), x(1-y), xy, (1-x)y on the unit cell."""
    return np.stack([np.stack([y - 1, x - 1], -1), np.stack([1 - y, -x], -1), np.stack([y, x], -1), np.stack([-y, 1 - x], -1)], 1)


def test_cell_samples_sit_at_the_piece_barycentre():
    # the integral of a linear function over a piece is its area times the value at the barycentre,
    # so v_k = w * grad phi_k(pos) with one weight w > 0 per sample
    _, H, _ = icosphere_system(256, 4, 2)
    S = oracle.vf_samples(H)
    g = _q1_grad(*S["pos"].T) * S["inv"][None, None, :]
    w = np.einsum("skd,skd->s", S["v"], g) / np.einsum("skd,skd->s", g, g)
    assert np.all(w > 0)
    err = np.abs(S["v"] - w[:, None, None] * g).max(axis=(1, 2)) / np.abs(S["v"]).max(axis=(1, 2))
    assert err.max() < 1e-4


def test_cell_gradient_is_the_bilinear_gradient():
    _, H, _ = icosphere_system(256, 4, 2)
    S = oracle.vf_samples(H)
    u = np.random.default_rng(5).standard_normal(H.level_info(0)["texels"])
    g = _q1_grad(*S["pos"].T) * S["inv"][None, None, :]
    expect = np.einsum("sk,skd->sd", u[S["corners"]], g)
    assert np.allclose(oracle.vf_cell_gradients(S, u), expect, rtol=0, atol=1e-9 * np.abs(expect).max())


def _face_closed_form(mesh, a_u, a_f):
    v, t, _ = mesh.arrays()
    P = v[t]
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    area = np.linalg.norm(n, axis=1) / 2
    n = n / (2 * area[:, None])
    gu = a_u - (n @ a_u)[:, None] * n
    gf = a_f - (n @ a_f)[:, None] * n
    val = np.sum(area * np.einsum("ij,ij->i", gu, gf) / np.linalg.norm(gu, axis=1))
    return -val / np.sqrt(area.sum())


@pytest.mark.parametrize("width", [128, 256])
def test_linear_potential_closed_form(width):
    mesh, H, _ = cube_system(width, 2)
    pos = _texel_positions(mesh, H)
    S = oracle.vf_samples(H)
    # u must have a nonzero tangential gradient on every face
    a_u = np.array([0.8, -0.5, 0.33])
    for a_f in [np.array([0.1, 0.7, -0.2]), np.array([-0.6, 0.2, 0.9])]:
        rhs = oracle.vf_divergence(S, pos @ a_u, len(pos))[0]
        f = pos @ a_f + 0.3
        assert abs(f @ rhs - _face_closed_form(mesh, a_u, a_f)) < 1e-6 * np.abs(_face_closed_form(mesh, a_u, a_f))
        # constants integrate to zero against any field
        assert abs(np.ones_like(f) @ rhs) < 1e-6 * np.abs(rhs).sum()

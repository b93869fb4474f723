"""Line integral convolution on the GPU: the diffused noise varies along the field far less than
across it (its anisotropic stiffness energy drops far more than its isotropic one), and both
solves match direct solutions."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

import gradientdomaintextureprocessing_amd as g
from helpers import mass_matrix, oracle_levels, stiffness_matrix

pytestmark = pytest.mark.gpu


def test_lic_smooths_along_the_field():
    mesh = g.TexturedMesh.torus(64, 32, 4, 2, 256, 256, 1.0, 0.35, 0.0, 5)
    v, t, _ = mesh.arrays()
    vf = g.surface_gradient_field(v, t, np.sin(3 * v[:, 0]) + np.cos(2 * v[:, 1]) + v[:, 2])
    lic = g.LineIntegralConvolution(mesh, 256, 256, 4, vf, max_cols=2)
    La = oracle_levels(lic.anisotropic, lic.lic)[0]
    Le = oracle_levels(lic.embedding, lic.sharpen)[0]
    Ma, Sa = mass_matrix(lic.anisotropic, La), stiffness_matrix(lic.anisotropic, La)
    Me, Se = mass_matrix(lic.embedding, Le), stiffness_matrix(lic.embedding, Le)
    noise = np.random.default_rng(4).random((2, La.N))
    x, y = lic.convolve(torch.from_numpy(noise).cuda(), tol=1e-9)
    x, y = x.cpu().numpy(), y.cpu().numpy()
    # direct solutions of both systems
    Aa = (Ma + 1e-4 * Sa).tocsc()
    xd = spla.splu(Aa).solve((Ma @ noise.T)).T
    Ae = (Me + 1e-4 * Se).tocsc()
    yd = spla.splu(Ae).solve(((Me + 100 * 1e-4 * Se) @ xd.T)).T
    for got, want, A in [(x, xd, Aa), (y, yd, Ae)]:
        e = got - want
        assert np.all(np.einsum("ij,ij->i", e, (A @ e.T).T) <= 1e-12 * np.einsum("ij,ij->i", want, (A @ want.T).T))
    # anisotropy: along-field variation (S_a) is removed far more than isotropic variation (S)
    ratio = lambda u: np.einsum("ij,ij->i", u, (Sa @ u.T).T) / np.einsum("ij,ij->i", u, (Se @ u.T).T)
    assert np.all(ratio(x) < 0.1 * ratio(noise))

"""GPU normalised-gradient divergence against the oracle, and the heat-method geodesics recipe
against great-circle distances on a sphere."""
import numpy as np
import pytest
import torch

import gradientdomaintextureprocessing_amd as g
import oracle
from helpers import cube_system, icosphere_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", [lambda: cube_system(128, 3, max_cols=3), lambda: icosphere_system(256, 4, 3, max_cols=3)],
                         ids=["cube128", "icosphere256"])
def test_divergence_matches_oracle(make):
    mesh, H, S = make()
    N = H.level_info(0)["texels"]
    rng = np.random.default_rng(11)
    ci, cj, _ = H.texel_coordinates(0)
    # smooth potentials plus noise, and one constant column (zero gradient: no contribution)
    u = np.stack([np.sin(ci / 17.0) * np.cos(cj / 23.0) + 0.01 * rng.standard_normal(N),
                  (ci - 40.0) ** 2 + 0.5 * (cj - 70.0) ** 2, np.full(N, 2.5)])
    ud = torch.from_numpy(u).cuda()
    rd = torch.empty_like(ud)
    S.geodesic_divergence(ud, rd)
    torch.cuda.synchronize()
    got = rd.cpu().numpy()
    want = oracle.vf_divergence(oracle.vf_samples(H), u, N)
    for k in range(2):
        assert np.abs(got[k] - want[k]).max() <= 1e-12 * np.abs(want[k]).max()
    assert np.abs(got[2]).max() == 0.0


def _texel_points(mesh, H):
    """Surface point of every texel whose centre lies in a triangle of its chart (barycentric)."""
    from scipy.spatial import cKDTree
    v, t, uv = mesh.arrays()
    W = H.width
    ci, cj, ch = H.texel_coordinates(0)
    q = np.stack([(ci + 0.5) / W, (cj + 0.5) / W], 1)
    cen = uv.mean(axis=1)
    tree = cKDTree(cen)
    _, cand = tree.query(q, k=12)
    pts = np.full((len(q), 3), np.nan)
    for k in range(cand.shape[1]):
        tri = cand[:, k]
        a, b, c = uv[tri, 0], uv[tri, 1], uv[tri, 2]
        d = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1])
        l1 = ((q[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (q[:, 1] - a[:, 1])) / d
        l2 = ((b[:, 0] - a[:, 0]) * (q[:, 1] - a[:, 1]) - (q[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1])) / d
        l0 = 1 - l1 - l2
        ok = (l0 >= -1e-9) & (l1 >= -1e-9) & (l2 >= -1e-9) & np.isnan(pts[:, 0])
        P = v[t[tri]]
        pts[ok] = (l0[:, None] * P[:, 0] + l1[:, None] * P[:, 1] + l2[:, None] * P[:, 2])[ok]
    return pts


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_geodesics_on_a_sphere(mode):
    mesh = g.TexturedMesh.icosphere(4, 512, 512, 0.0, 3)
    H = g.InitializeHierarchy(mesh, 512, 512, 5)
    geo = g.Geodesics(H, heat_time=1e-3, max_sources=2)
    pts = _texel_points(mesh, H)
    inside = np.flatnonzero(~np.isnan(pts[:, 0]))
    src = [int(inside[1000]), int(inside[len(inside) // 2 + 777])]
    phi = geo.distances(src, tol=1e-7, mode=mode).cpu().numpy()
    v, t, _ = mesh.arrays()
    P = v[t]
    area = np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1).sum() / 2
    for k, s in enumerate(src):
        p = pts[inside] / np.linalg.norm(pts[inside], axis=1, keepdims=True)
        ps = pts[s] / np.linalg.norm(pts[s])
        d = np.arccos(np.clip(p @ ps, -1, 1)) / np.sqrt(area)  # great-circle distance, unit-area metric
        sel = (d > 0.15 * d.max()) & (d < 0.85 * d.max())
        rel = np.abs(phi[k, inside][sel] - d[sel]) / d[sel]
        assert np.median(rel) < 0.05 and np.percentile(rel, 95) < 0.12
        assert abs(phi[k, s]) == 0.0


@pytest.mark.parametrize("cycles", [0, 2])
def test_distance_system_with_several_cycles_per_iteration(cycles):
    """1e-6 M + S is nearly singular in FP32: the automatic setting (0) runs one V-cycle per FP64
    residual, and a forced two-cycle setting falls back to it once the residual grows; both reach
    the tolerance and agree with the one-cycle solve."""
    mesh = g.TexturedMesh.icosphere(4, 512, 512, 0.0, 3)
    H = g.InitializeHierarchy(mesh, 512, 512, 5)
    geo = g.Geodesics(H, heat_time=1e-3, max_sources=2)
    N = H.level_info(0)["texels"]
    delta = torch.zeros(2, N, dtype=torch.float64, device="cuda")
    delta[0, 10] = delta[1, N // 2] = 1.0
    u = torch.zeros_like(delta)
    geo.heat.solve(u, delta, tol=1e-8, max_iters=80, mode="mixed")
    div = torch.empty_like(u)
    geo.heat.geodesic_divergence(u, div)
    out = []
    for c in (1, cycles):
        geo.distance.configure_mixed(cycles=c)
        phi = torch.zeros_like(u)
        its, rel = geo.distance.solve(phi, div, tol=1e-8, max_iters=80, mode="mixed")
        print(f"[distance] cycles={c}: {its} iterations, rel res {rel.max():.2e}", flush=True)
        assert rel.max() <= 1e-8
        out.append(phi - phi.mean(dim=1, keepdim=True))
    scale = out[0].abs().max().item()
    assert (out[0] - out[1]).abs().max().item() <= 1e-4 * scale


@pytest.mark.parametrize("sharpness", [0.0, 2.0])
def test_texture_filter_matches_direct_solve(sharpness):
    import scipy.sparse.linalg as spla
    from helpers import mass_matrix, oracle_levels, stiffness_matrix
    mesh = g.TexturedMesh.cube(16, 128, 128)
    H = g.InitializeHierarchy(mesh, 128, 128, 3)
    flt = g.TextureFilter(H, stiffness_weight=1e-3, sharpness_weight=sharpness, max_cols=3)
    L0 = oracle_levels(H, flt.system)[0]
    M, K = mass_matrix(H, L0), stiffness_matrix(H, L0)
    x0 = np.random.default_rng(2).random((3, L0.N))
    x, its = flt.filter(torch.from_numpy(x0).cuda(), tol=1e-10, mode="mixed")
    A = (M + 1e-3 * K).tocsc()
    want = spla.splu(A).solve(((M + sharpness * 1e-3 * K) @ x0.T)).T
    e = x.cpu().numpy() - want
    assert np.all(np.einsum("ij,ij->i", e, (A @ e.T).T) <= 1e-14 * np.einsum("ij,ij->i", want, (A @ want.T).T))

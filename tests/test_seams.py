"""Chart coupling along seams (include/Src/BoundaryQuadraticElements.inl:652-773): seam nodes carry
the mean of the charts' bilinear interpolations.

Pinned by closed forms: on the cube every face is flat and its UV map affine, so a linear function
of the 3D position sampled at the texels (exterior texels included, through the face's affine map)
is reproduced exactly by every chart, both charts agree at every seam node, and f^T S f, f^T M f
and 1^T M f equal the surface integrals computed from the triangles alone.  A seam node placed at
the wrong parameter, or read from the wrong chart, breaks the equality.  A jump between charts
costs energy only if the charts are coupled."""
import numpy as np
import pytest

from helpers import cube_system, icosphere_system, mass_matrix, oracle_levels, stiffness_matrix


def _texel_positions(mesh, H):
    """3D position of every level-0 texel through its chart's affine UV->surface map (cube faces)."""
    v, t, uv = mesh.arrays()
    W = H.width
    charts = H.array(0, "charts").reshape(-1, 9)
    ci, cj, ch = H.texel_coordinates(0)
    cen = uv.mean(axis=1) * W - 0.5  # triangle UV centroids in texel coordinates
    tri_chart = np.full(len(t), -1)
    for c, row in enumerate(charts):
        w, h, x0, y0 = row[0], row[1], row[2], row[3]
        inside = (cen[:, 0] >= x0) & (cen[:, 0] <= x0 + w - 1) & (cen[:, 1] >= y0) & (cen[:, 1] <= y0 + h - 1)
        tri_chart[inside] = c
    assert np.all(tri_chart >= 0)
    pos = np.zeros((len(ci), 3))
    u = np.stack([(ci + 0.5) / W, (cj + 0.5) / W], axis=1)
    for c in range(len(charts)):
        k = np.flatnonzero(tri_chart == c)[0]
        P = v[t[k]]                       # 3 x 3
        D = np.stack([uv[k, 1] - uv[k, 0], uv[k, 2] - uv[k, 0]], axis=1)
        E = np.stack([P[1] - P[0], P[2] - P[0]], axis=1)
        A = E @ np.linalg.inv(D)          # 3 x 2
        sel = ch == c
        pos[sel] = P[0] + (u[sel] - uv[k, 0]) @ A.T
    return pos


def _surface_integrals(mesh, a, c0):
    """Exact integrals of g = a.p + c0 over the triangles: Dirichlet energy (scale invariant in
    2D), and the mean of g and of g^2 with the surface area normalised to one."""
    v, t, _ = mesh.arrays()
    P = v[t]
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    area = np.linalg.norm(n, axis=1) / 2
    n = n / (2 * area[:, None])
    energy = np.sum(area * (a @ a - (n @ a) ** 2))
    gv = P @ a + c0
    m1 = np.sum(area * gv.sum(axis=1) / 3)
    m2 = np.sum(area * (np.sum(gv ** 2, axis=1) + gv[:, 0] * gv[:, 1] + gv[:, 1] * gv[:, 2] + gv[:, 0] * gv[:, 2]) / 6)
    A = area.sum()
    return energy, m1 / A, m2 / A


@pytest.mark.parametrize("width", [128, 256])
def test_linear_functions_are_exact_across_seams(width):
    mesh, H, S = cube_system(width, 2)
    L0 = oracle_levels(H, S)[0]
    M, K = mass_matrix(H, L0), stiffness_matrix(H, L0)
    pos = _texel_positions(mesh, H)
    for a, c0 in [(np.array([0.3, -0.7, 0.45]), 0.2), (np.array([1.0, 0.0, 0.0]), 0.0)]:
        f = pos @ a + c0
        energy, m1, m2 = _surface_integrals(mesh, a, c0)
        assert abs(f @ (K @ f) - energy) < 1e-9 * energy
        assert abs(np.ones_like(f) @ (M @ f) - m1) < 1e-11
        assert abs(f @ (M @ f) - m2) < 1e-11


def _chart_jump_energy(H, S):
    L0 = oracle_levels(H, S)[0]
    K = stiffness_matrix(H, L0)
    _, _, ch = H.texel_coordinates(0)
    d = np.random.default_rng(3).standard_normal(ch.max() + 1)[ch]  # one constant per chart
    return d @ (K @ d), K, ch


@pytest.mark.parametrize("make", [lambda: cube_system(128, 2), lambda: icosphere_system(256, 4, 2)])
def test_charts_are_coupled(make):
    _, H, S = make()
    e, K, ch = _chart_jump_energy(H, S)
    # per-chart constants are in the kernel of a decoupled operator; here they are penalised
    assert e > 1.0
    Kc = K.tocoo()
    cross = ch[Kc.row] != ch[Kc.col]
    assert cross.sum() > 0
    # every chart couples to a neighbour, and the coupling graph over charts is connected
    n = ch.max() + 1
    adj = np.zeros((n, n), bool)
    adj[ch[Kc.row[cross]], ch[Kc.col[cross]]] = True
    seen, todo = {0}, [0]
    while todo:
        c = todo.pop()
        for o in np.flatnonzero(adj[c]):
            if o not in seen:
                seen.add(int(o)), todo.append(int(o))
    assert len(seen) == n

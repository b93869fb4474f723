"""Fused four-colour deep sweep (one launch per sweep, tile and halo staged in LDS, ping-pong
correction buffers, optionally adding the prolongation at the start of post-smoothing) and the
colour-pair sweep (one launch per colour pair, a workgroup per sub-grid row) against one launch per
colour and a separate prolongation pass: every tile shape gives bitwise the same iterate of the
mixed solver, on screened operators (FP16 smoother stencil) and on a nearly singular one (FP32
smoother stencil), for column counts that fill one and two column chunks, and for rows of every
width the pair sweep handles.  The FP32 fallback for FP16-range stencils (forced by lowering the
FP16 ratio limit) solves to the direct solution, and so do FP16 Schwarz block inverses forced onto
every level, which flexible CG leaves unread."""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from helpers import noise_texture_rhs, oracle_levels

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import gradientdomaintextureprocessing_amd as g  # noqa: E402


def iterate(S, b, shape, iters, use_graph, fuse_prolongation=False):
    S.tile_shape(shape)
    S.fuse_transfers(prolongation=fuse_prolongation)
    S.configure_mixed(use_graph=use_graph)
    x = torch.zeros_like(b)
    try:
        S.solve(x, b, tol=1e-30, max_iters=iters, mode="mixed")
    except g._lib.GdtpError as e:  # the tolerance is out of reach on purpose: a fixed number of iterations
        assert e.code == 3
    torch.cuda.synchronize()
    return x.cpu().numpy()


@pytest.mark.parametrize("case", ["icosphere512", "cube512-singular"])
@pytest.mark.parametrize("K", [3, 6])
def test_fused_sweep_is_bitwise_four_launches(case, K):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if case == "icosphere512":
        mesh = g.TexturedMesh.icosphere(4, 512, 512, 0.05, 3)
        H = g.InitializeHierarchy(mesh, 512, 512, 5)
        S = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=K)
    else:  # the geodesic distance operator 1e-6 M + S keeps the FP32 smoother stencil
        mesh = g.TexturedMesh.cube(16, 512, 512)
        H = g.InitializeHierarchy(mesh, 512, 512, 5)
        S = g.UpdateLinearSystem(H, 1.0, 1e-6, max_cols=K)
    N = H.level_info(0)["texels"]
    b = torch.from_numpy(np.random.default_rng(K).standard_normal((K, N))).cuda()
    print(f"[fused {case} K={K}] per-colour launches", flush=True)
    want = iterate(S, b, -1, 2, False)
    assert np.all(np.isfinite(want)) and np.abs(want).max() > 0
    for shape in [*range(7), S.TILE_SHAPE_PAIRS]:
        for use_graph in (False, True):
            print(f"[fused {case} K={K}] shape {shape} graph {use_graph}", flush=True)
            got = iterate(S, b, shape, 2, use_graph)
            assert np.array_equal(got, want), f"shape {shape}, graph {use_graph}: max diff {np.abs(got - want).max()}"
            # the prolongation fused into the first post-smoothing sweep changes nothing either
            got = iterate(S, b, shape, 2, use_graph, fuse_prolongation=True)
            assert np.array_equal(got, want), f"shape {shape}, graph {use_graph}, fused prolongation: max diff {np.abs(got - want).max()}"
    # level 0 is swept by tiles (charts too large for one workgroup), so the fused kernel ran
    S.tile_shape(0)
    S.mixed_deep_sweep(0, K)
    torch.cuda.synchronize()


@pytest.mark.parametrize("width,lo,hi", [(2048, 0, 512), (4096, 512, 1024), (8192, 1024, 2048)])
def test_pair_sweep_wide_rows_is_bitwise_four_launches(width, lo, hi):
    """Level-0 rows of about 500, 1000 and 2000 active sub-grid texels (a torus in two charts side
    by side, 256 texels high): the pair sweep packs rows into 512-thread and 1024-thread workgroups
    (one texel per thread) and, for the widest, gives each row a 1024-thread workgroup with two
    texels per thread."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mesh = g.TexturedMesh.torus(48, 24, 1, 2, width, 256)
    H = g.InitializeHierarchy(mesh, width, 256, 4)
    # widest active span of a sub-grid row (texel (i, j) of a chart lies in row j >> 1, column i >> 1)
    ni = H.array(0, "nodeInfo").reshape(-1, 4)
    key = ni[:, 0].astype(np.int64) * 1000000 + (ni[:, 2] >> 1)
    col = ni[:, 1] >> 1
    order = np.argsort(key, kind="stable")
    k, c = key[order], col[order]
    starts = np.flatnonzero(np.r_[True, k[1:] != k[:-1]])
    widest = int((np.maximum.reduceat(c, starts) - np.minimum.reduceat(c, starts) + 1).max())
    assert lo < widest <= hi
    S = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=3)
    N = H.level_info(0)["texels"]
    b = torch.from_numpy(np.random.default_rng(width).standard_normal((3, N))).cuda()
    print(f"[pairs {width}x256 widest span {widest}] per-colour launches", flush=True)
    want = iterate(S, b, -1, 2, False)
    assert np.all(np.isfinite(want)) and np.abs(want).max() > 0
    for use_graph in (False, True):
        print(f"[pairs {width}x256 widest span {widest}] graph {use_graph}", flush=True)
        got = iterate(S, b, S.TILE_SHAPE_PAIRS, 2, use_graph)
        assert np.array_equal(got, want), f"graph {use_graph}: max diff {np.abs(got - want).max()}"


def test_fp32_smoother_fallback_solves_to_direct():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mesh = g.TexturedMesh.icosphere(4, 256, 256, 0.05, 3)
    H = g.InitializeHierarchy(mesh, 256, 256, 4)
    S16 = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=3)
    S16.configure_mixed()
    os.environ["GDTP_FP16_RATIO_LIMIT"] = "0.1"  # every stencil ratio of 1e-3 S + M exceeds it somewhere
    try:
        S32 = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=3)
        S32.configure_mixed()
    finally:
        del os.environ["GDTP_FP16_RATIO_LIMIT"]
    # FP32 smoother stencils: 40 bytes per dense texel instead of 24
    assert S32.memory()[1] > S16.memory()[1]
    L = oracle_levels(H, S32)
    A = oracle.full_matrix(L[0]).tocsc()
    _, b = noise_texture_rhs(H, L[0], K=3, seed=8)
    x = torch.zeros((3, L[0].N), dtype=torch.float64, device="cuda")
    its, rel = S32.solve(x, torch.from_numpy(b).cuda(), tol=1e-9, max_iters=40, mode="mixed")
    xs = x.cpu().numpy()
    lu = spla.splu(A)
    xd = np.stack([lu.solve(b[k]) for k in range(3)])
    e = xs - xd
    en = np.sqrt(np.einsum("ij,ij->i", e, (A @ e.T).T) / np.einsum("ij,ij->i", xd, (A @ xd.T).T))
    print(f"[fp32 fallback] {its} iterations, energy error {en}", flush=True)
    assert np.all(en < 1e-7)


def test_fp16_schwarz_blocks_solve_to_direct():
    """The boundary Schwarz step's block inverses as FP16 quads scaled per core row (forced onto
    every level of a screened system, whose small levels keep FP32 by default): refinement still
    reaches 1e-9 and matches the direct solution, in at most one more iteration than FP32; flexible
    CG on the same system reads the FP32 inverses and matches the system without pairs bitwise."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mesh = g.TexturedMesh.icosphere(4, 256, 256, 0.05, 3)
    H = g.InitializeHierarchy(mesh, 256, 256, 4)
    S32 = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=3)
    S32.configure_mixed()
    os.environ["GDTP_SCHWARZ_HALF_BLOCKS"] = "0"
    try:
        S16 = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=3)
        S16.configure_mixed()
    finally:
        del os.environ["GDTP_SCHWARZ_HALF_BLOCKS"]
    # the FP16 quads are held beside the FP32 inverses
    assert S16.memory()[1] > S32.memory()[1]
    L = oracle_levels(H, S16)
    A = oracle.full_matrix(L[0]).tocsc()
    _, b = noise_texture_rhs(H, L[0], K=3, seed=11)
    lu = spla.splu(A)
    xd = np.stack([lu.solve(b[k]) for k in range(3)])
    its = {}
    for name, S in (("fp32", S32), ("fp16", S16)):
        x = torch.zeros((3, L[0].N), dtype=torch.float64, device="cuda")
        its[name], rel = S.solve(x, torch.from_numpy(b).cuda(), tol=1e-9, max_iters=40, mode="mixed")
        e = x.cpu().numpy() - xd
        en = np.sqrt(np.einsum("ij,ij->i", e, (A @ e.T).T) / np.einsum("ij,ij->i", xd, (A @ xd.T).T))
        print(f"[schwarz {name}] {its[name]} iterations, energy error {en}", flush=True)
        assert np.all(rel <= 1e-9) and np.all(en < 1e-7)
    assert its["fp16"] <= its["fp32"] + 1
    xs = []
    for S in (S32, S16):
        S.configure_mixed(krylov=True)
        try:
            x = torch.zeros((3, L[0].N), dtype=torch.float64, device="cuda")
            S.solve(x, torch.from_numpy(b).cuda(), tol=1e-8, max_iters=60, mode="mixed")
            xs.append(x.cpu().numpy())
        finally:
            S.configure_mixed()
    assert np.array_equal(xs[0], xs[1])

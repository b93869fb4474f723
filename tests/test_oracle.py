"""The oracle (plain-C restatement of the reference V-cycle kernels) pinned by mathematics:
Galerkin identity, true residuals, P^T restriction, a dense-matrix restatement of the band
Gauss-Seidel schedule, energy descent, and convergence to the sparse direct solution."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from helpers import cube_system, icosphere_system, noise_texture_rhs, oracle_levels


@pytest.fixture(scope="module", params=["cube", "icosphere"])
def case(request):
    if request.param == "cube":
        _, H, S = cube_system(128, 3)
    else:
        _, H, S = icosphere_system(256, 4, 3)
    L = oracle_levels(H, S)
    return H, S, L


def test_galerkin_coarse_operators(case):
    """R A P of the fine operator equals the coarse operator (UpdateCoefficients.inl:295-305, 348-371)."""
    H, S, L = case
    for l in range(H.levels - 1):
        P = oracle.prolongation_matrix(L[l], L[l + 1].N)
        G = (P.T @ oracle.full_matrix(L[l]) @ P).tocsr()
        Ac = oracle.full_matrix(L[l + 1])
        assert abs(G - Ac).max() <= 1e-10 * abs(Ac).max()


def test_relaxation_residual_is_true_residual(case):
    H, S, L = case
    x0, b = noise_texture_rhs(H, L[0])
    x, r = x0.copy(), np.zeros_like(x0)
    oracle.relaxation(L[0], spla.splu(L[0].bb.scipy().tocsc()), x, b, r)
    A = oracle.full_matrix(L[0])
    rt = b - (A @ x.T).T
    assert np.abs(r - rt).max() <= 1e-12 * np.abs(b).max()


def test_restriction_is_prolongation_transpose(case):
    H, S, L = case
    rng = np.random.default_rng(3)
    for l in range(H.levels - 1):
        f = rng.standard_normal((2, L[l].N))
        P = oracle.prolongation_matrix(L[l], L[l + 1].N)
        assert np.allclose(oracle.restrict(L[l], L[l + 1], f), (P.T @ f.T).T, atol=1e-12)
        c = rng.standard_normal((2, L[l + 1].N))
        out = np.zeros((2, L[l].N))
        oracle.prolong_add(L[l], L[l + 1], c, out)
        assert np.allclose(out, (P @ c.T).T, atol=1e-13)


def dense_band_relax(L, A, x, b):
    """The band schedule restated with a dense matrix: x_i = (b_i - sum_{j != i} A_ij x_j) / A_ii,
    tasks even then odd, blocks in order, rows lagged by three passes."""
    t = L.tasks
    seg = t["segments"].reshape(-1, 5)
    Ad = A.toarray()
    for color in (0, 1):
        for task in range(color, len(t["taskPtr"]) - 1, 2):
            for blk in range(t["taskPtr"][task], t["taskPtr"][task + 1]):
                r0, nd = t["blockBegin"][blk], t["blockDeepEnd"][blk] - t["blockBegin"][blk]
                for j in range(nd + 3):
                    for g in range(3):
                        row = j - g
                        if 0 <= row < nd:
                            for s in range(t["linePtr"][r0 + row], t["linePtr"][r0 + row + 1]):
                                for i in range(seg[s, 0], seg[s, 1] + 1):
                                    x[:, i] = (b[:, i] - Ad[i] @ x.T + Ad[i, i] * x[:, i]) / Ad[i, i]


def test_band_relax_matches_dense_schedule():
    _, H, S = cube_system(96, 2)
    L = oracle_levels(H, S)[0]
    A = oracle.full_matrix(L)
    rng = np.random.default_rng(5)
    x0, b = rng.standard_normal((2, L.N)), rng.standard_normal((2, L.N))
    x1 = x0.copy()
    oracle.band_relax(L, x1, b)
    x2 = x0.copy()
    dense_band_relax(L, A, x2, b)
    assert np.abs(x1 - x2).max() <= 1e-12 * np.abs(x2).max()


def test_relaxation_decreases_energy_error(case):
    H, S, L = case
    A = oracle.full_matrix(L[0]).tocsc()
    x0, b = noise_texture_rhs(H, L[0], K=1)
    xs = spla.spsolve(A, b[0])
    x = x0.copy()
    lu = spla.splu(L[0].bb.scipy().tocsc())
    e = lambda v: float((v - xs) @ (A @ (v - xs)))
    before = e(x[0])
    oracle.relaxation(L[0], lu, x, b)
    assert e(x[0]) < before


def test_vcycle_converges_to_direct_solution(case):
    H, S, L = case
    A = oracle.full_matrix(L[0]).tocsc()
    x0, b = noise_texture_rhs(H, L[0])
    xd = np.stack([spla.spsolve(A, b[k]) for k in range(3)])
    solv = oracle.Solvers(L, oracle.full_matrix(L[-1]))
    x = x0.copy()
    rel = []
    for _ in range(6):
        oracle.vcycle(L, solv, x, b)
        rel.append(np.linalg.norm(b - (A @ x.T).T) / np.linalg.norm(b))
    assert rel[-1] < 1e-9
    assert all(rel[i + 1] < 0.2 * rel[i] for i in range(4))
    # exterior texels with sliver supports make A ill-conditioned; compare in the energy norm
    e = x - xd
    en = np.sqrt(np.einsum("ij,ij->", e, (A @ e.T).T) / np.einsum("ij,ij->", xd, (A @ xd.T).T))
    assert en < 1e-8


def test_apply_is_full_matrix_product(case):
    """oracle.apply (MultiplyBySystemMatrix_NoReciprocals, IterativeSolvers.inl:881-950) equals the
    product with the sparse matrix assembled row by row from the same stencils and boundary rows,
    for A on every level and for M and S on level 0."""
    from helpers import mass_matrix, stiffness_matrix
    H, S, L = case
    rng = np.random.default_rng(11)
    for l in range(H.levels):
        x = rng.standard_normal((3, L[l].N))
        want = (oracle.full_matrix(L[l]) @ x.T).T
        got = oracle.apply(L[l], L[l].coef_raw, L[l].bd, L[l].bb, x)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    x = rng.standard_normal((2, L[0].N))
    for name, build in (("mass", mass_matrix), ("stiffness", stiffness_matrix)):
        bd = oracle.Csr(H.array(0, name + ".bd.rowptr"), H.array(0, name + ".bd.col"), H.array(0, name + ".bd.val", np.float64), L[0].N)
        bb = oracle.Csr(H.array(0, name + ".bb.rowptr"), H.array(0, name + ".bb.col"), H.array(0, name + ".bb.val", np.float64), L[0].B)
        got = oracle.apply(L[0], H.array(0, name + ".deep", np.float64), bd, bb, x)
        want = (build(H, L[0]) @ x.T).T
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # a planted mistake (a corner and an edge slot exchanged) must fail the comparison
    bad = L[0].coef_raw.reshape(-1, 10).copy()
    bad[:, [0, 1]] = bad[:, [1, 0]]
    x = rng.standard_normal((1, L[0].N))
    got = oracle.apply(L[0], bad.ravel(), L[0].bd, L[0].bb, x)
    want = (oracle.full_matrix(L[0]) @ x.T).T
    assert np.abs(got - want).max() > 1e-6 * np.abs(want).max()


def test_parallel_oracle_equals_one_thread(case):
    """The oracle's OpenMP loops (same-colour band tasks, rows with disjoint outputs) and the
    per-column exact solves give the one-thread V-cycle bit for bit."""
    H, S, L = case
    x0, b = noise_texture_rhs(H, L[0], K=3)
    solv = oracle.Solvers(L, oracle.full_matrix(L[-1]))
    out = []
    for threads in (1, 8):
        oracle.set_threads(threads)
        x = x0.copy()
        oracle.vcycle(L, solv, x, b)
        out.append(x)
    oracle.set_threads(8)
    assert np.array_equal(out[0], out[1])

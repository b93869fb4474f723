"""Edge cases and error paths of the device entry points: vectors of the wrong size, type or
layout are refused before they reach a kernel; zero columns, more columns than the system holds,
invalid weights and a solve that runs out of iterations are reported as errors; the smallest
(one column) and largest (max_cols) batches and a two-level hierarchy solve correctly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import gradientdomaintextureprocessing_amd as g  # noqa: E402
from gradientdomaintextureprocessing_amd import _lib  # noqa: E402


@pytest.fixture(scope="module")
def system():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    H = g.InitializeHierarchy(g.TexturedMesh.cube(16, 128, 128), 128, 128, 3)
    return H, g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=4)


def _vecs(N, K=2):
    return torch.zeros(K, N, dtype=torch.float64, device="cuda"), torch.rand(K, N, dtype=torch.float64, device="cuda")


def test_wrong_vectors_are_refused(system):
    H, S = system
    N = H.level_info(0)["texels"]
    x, b = _vecs(N)
    with pytest.raises(ValueError):
        S.solve(x[:, :-1].contiguous(), b[:, :-1].contiguous())          # wrong texel count
    with pytest.raises(ValueError):
        S.solve(x, b[:1].contiguous())                                     # columns differ
    with pytest.raises(ValueError):
        S.vcycle(x.float(), b.float())                                     # FP32
    with pytest.raises(ValueError):
        S.apply(g.OPERATOR_MASS, torch.zeros(N, 2, dtype=torch.float64, device="cuda").t(), b)  # not contiguous
    with pytest.raises(ValueError):
        S.residual(x.cpu(), b.cpu(), b.cpu())                              # host tensors
    with pytest.raises(ValueError):
        S.restrict(0, x, torch.zeros(2, N, dtype=torch.float64, device="cuda"))  # coarse of the wrong level
    with pytest.raises(ValueError):
        S.relax(7, x, b)                                                   # no such level


def test_column_counts_and_weights(system):
    H, S = system
    N = H.level_info(0)["texels"]
    for K in (0, 5):  # none, and more than max_cols = 4
        x, b = _vecs(N, K)
        with pytest.raises(_lib.GdtpError) as e:
            S.solve(x, b)
        assert e.value.code == 1
    x, b = _vecs(N, 2)
    with pytest.raises(_lib.GdtpError):
        S.solve_weighted(x, b, [[0.0, 1e-3], [1.0, 1e-3]])                 # zero mass weight
    with pytest.raises(_lib.GdtpError):
        S.solve_weighted(x, b, [[1.0, -1e-3], [1.0, 1e-3]])                # negative stiffness
    with pytest.raises(_lib.GdtpError) as e:
        S.solve(x, b, tol=1e-12, max_iters=1)                              # runs out of iterations
    assert e.value.code == 3 and torch.isfinite(x).all()


@pytest.mark.parametrize("K", [1, 4])
def test_one_and_max_columns(system, K):
    H, S = system
    N = H.level_info(0)["texels"]
    x, b = _vecs(N, K)
    its, rel = S.solve(x, b, tol=1e-9, max_iters=40)
    assert rel.max() <= 1e-9
    r = torch.empty_like(b)
    S.residual(x, b, r)
    assert torch.all(r.norm(dim=1) / b.norm(dim=1) <= 1.01e-9)


def test_two_level_hierarchy():
    H = g.InitializeHierarchy(g.TexturedMesh.cube(8, 64, 64), 64, 64, 2)
    S = g.UpdateLinearSystem(H, 1e-3, 1.0, max_cols=2)
    x, b = _vecs(H.level_info(0)["texels"])
    for mode in ("mixed", "fp64"):
        x.zero_()
        its, rel = S.solve(x, b, tol=1e-10, max_iters=60, mode=mode)
        assert rel.max() <= 1e-10, mode

"""Modulated stiffness on the GPU (CellStiffnessToTexelStiffness, IterativeSolvers.inl:1170-1208)
against the oracle's restatement on the same cell products and a random cell mask; a uniform mask
s gives s S f (the stiffness apply of the system), and the filtering right-hand side with that
mask equals the uniform-sharpness one (TextureFiltering.cpp:569-570)."""
import numpy as np
import pytest

import oracle
from helpers import cube_system, icosphere_system

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import gradientdomaintextureprocessing_amd as g  # noqa: E402


@pytest.fixture(scope="module", params=["cube128", "icosphere256"])
def case(request):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _, H, S = cube_system(128, 3, max_cols=3) if request.param == "cube128" else icosphere_system(256, 4, 3, max_cols=3)
    N = H.level_info(0)["texels"]
    f = np.random.default_rng(11).random((3, N))
    mod = g.StiffnessModulation(H, f)
    print(f"[modulation {request.param}] N={N} cells={mod.num_cells} boundary entries={mod.boundary_entries}", flush=True)
    return request.param, H, S, f, mod


def test_kernel_matches_oracle(case):
    name, H, S, f, mod = case
    mask = np.random.default_rng(12).random(mod.num_cells) * 3.0
    t = mod.apply(torch.from_numpy(mask).cuda()).cpu().numpy()
    want = oracle.modulated_stiffness(mod.array("deep"), mod.array("deepCells", np.int32), H.array(0, "deepGlobalIndex"),
                                      mod.array("boundaryRowptr", np.int32), mod.array("boundaryCell", np.int32),
                                      mod.array("boundaryVal"), H.array(0, "boundaryGlobalIndex"), mask, 3, mod.texels)
    assert np.abs(t - want).max() <= 1e-13 * np.abs(want).max()


def test_uniform_mask_is_scaled_stiffness(case):
    name, H, S, f, mod = case
    fg = torch.from_numpy(f).cuda()
    sf = torch.empty_like(fg)
    S.apply(g.OPERATOR_STIFFNESS, fg, sf)
    t = mod.apply(torch.full((mod.num_cells,), 2.5, dtype=torch.float64, device="cuda"))
    assert (t - 2.5 * sf).abs().max().item() <= 1e-12 * sf.abs().max().item()


def test_filter_rhs_with_uniform_mask():
    _, H, _ = cube_system(128, 3)
    flt = g.TextureFilter(H, stiffness_weight=1e-3, sharpness_weight=2.0, max_cols=3)
    x0 = torch.from_numpy(np.random.default_rng(13).random((3, H.level_info(0)["texels"]))).cuda()
    flt.set_signal(x0)
    mask = torch.full((flt.modulation.num_cells,), 2.0, dtype=torch.float64, device="cuda")
    a, b = flt.rhs(x0), flt.rhs(x0, mask)
    assert (a - b).abs().max().item() <= 1e-12 * a.abs().max().item()
    x, its = flt.filter(x0, tol=1e-8, mask=mask)
    y, _ = flt.filter(x0, tol=1e-8)
    assert (x - y).abs().max().item() <= 1e-6 * y.abs().max().item()

"""gdtp_residual_norms_noalloc (scratch in a static device array, no allocation) against
gdtp_residual_norms (scratch allocated by the library), through the C ABI: bitwise equal norms on
two calls in a row, the second on a smaller system with fewer columns; more than 64 columns is
refused with an error."""
import ctypes

import numpy as np
import pytest

from gradientdomaintextureprocessing_amd import _lib
from helpers import cube_system, icosphere_system

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _norms(fn, S, x, b):
    K = x.shape[0]
    rel, norms = np.zeros(K), np.zeros(2 * K)
    rc = fn(S._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(b.data_ptr()), K, rel.ctypes.data, norms.ctypes.data, None)
    return rc, rel, norms


def test_noalloc_matches_allocating_entry_point():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _lib.load()
    rng = np.random.default_rng(21)
    for name, (_, H, S), K in (("icosphere256", icosphere_system(256, 4, 3, max_cols=8), 8),
                               ("cube128", cube_system(128, 3, max_cols=8), 3)):
        N = H.level_info(0)["texels"]
        x = torch.from_numpy(rng.standard_normal((K, N))).cuda()
        b = torch.from_numpy(rng.standard_normal((K, N))).cuda()
        print(f"[no-alloc] {name}: {K} columns, {N} texels", flush=True)
        rc0, rel0, n0 = _norms(lib.gdtp_residual_norms, S, x, b)
        rc1, rel1, n1 = _norms(lib.gdtp_residual_norms_noalloc, S, x, b)
        assert rc0 == 0 and rc1 == 0, lib.gdtp_last_error()
        assert np.array_equal(n0, n1) and np.array_equal(rel0, rel1)
        r = torch.empty_like(b)
        S.residual(x, b, r)
        assert np.allclose(n1[0::2], r.norm(dim=1).cpu().numpy(), rtol=1e-12, atol=0)
        assert np.allclose(n1[1::2], b.norm(dim=1).cpu().numpy(), rtol=1e-12, atol=0)


def test_noalloc_refuses_more_than_64_columns():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _lib.load()
    _, H, S = cube_system(128, 2, max_cols=65)
    N = H.level_info(0)["texels"]
    x = torch.zeros((65, N), dtype=torch.float64, device="cuda")
    print("[no-alloc] 65 columns refused", flush=True)
    rc, _, _ = _norms(lib.gdtp_residual_norms_noalloc, S, x, x)
    assert rc == 1 and b"64 columns" in lib.gdtp_last_error()
    rc, _, _ = _norms(lib.gdtp_residual_norms, S, x, x)
    assert rc == 0

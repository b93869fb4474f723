"""gdtp_residual_norms: per-column ||b - A x|| / ||b|| in FP64 without writing the residual,
against the residual written by gdtp_residual and the oracle's assembled matrix (CPU), and
bitwise repeatable."""
import numpy as np
import pytest

import oracle
from helpers import icosphere_system, oracle_levels

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_residual_norms_match_written_residual_and_oracle():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _, H, S = icosphere_system(256, 4, 3, max_cols=6)
    L = oracle_levels(H, S)
    A = oracle.full_matrix(L[0])
    rng = np.random.default_rng(5)
    x = rng.standard_normal((6, L[0].N))
    b = rng.standard_normal((6, L[0].N))
    xd, bd = torch.from_numpy(x).cuda(), torch.from_numpy(b).cuda()
    print("[residual norms] icosphere256, 6 columns", flush=True)
    rel, rn, bn = S.residual_norms(xd, bd)
    rel2, rn2, bn2 = S.residual_norms(xd, bd)
    assert np.array_equal(rel, rel2) and np.array_equal(rn, rn2) and np.array_equal(bn, bn2)
    r = torch.empty_like(bd)
    S.residual(xd, bd, r)
    want_r = r.norm(dim=1).cpu().numpy()
    assert np.allclose(rn, want_r, rtol=1e-12, atol=0)
    assert np.allclose(bn, np.linalg.norm(b, axis=1), rtol=1e-12, atol=0)
    oracle_r = np.linalg.norm(b - (A @ x.T).T, axis=1)
    assert np.allclose(rn, oracle_r, rtol=1e-11, atol=0)
    assert np.allclose(rel, oracle_r / np.linalg.norm(b, axis=1), rtol=1e-11, atol=0)
    # fewer columns than the system holds: the first three columns alone give the same norms
    rel3, rn3, _ = S.residual_norms(xd[:3].contiguous(), bd[:3].contiguous())
    assert np.array_equal(rn3, rn[:3]) and np.array_equal(rel3, rel[:3])

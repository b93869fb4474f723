"""The reference's own CPU case at its full size: unit cube, 512² atlas, 4 levels, 3-channel
texture, (M + 1e-3 S) x = M f from x = f, five FP64 V-cycles of the reference schedule (band
Gauss-Seidel with 3 lagged passes, exact boundary and coarse solves).  Every cycle's iterate on the
GPU matches the oracle's to 1e-9 relative RMS, and the mixed-precision solve reaches the direct
solution of the same system."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from helpers import cube_system, noise_texture_rhs, oracle_levels

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_c1_five_vcycles_match_oracle_and_mixed_reaches_direct():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _, H, S = cube_system(512, 4, max_cols=3)
    L = oracle_levels(H, S)
    print(f"[c1] levels N={[l.N for l in L]} B={[l.B for l in L]}", flush=True)
    x0, b = noise_texture_rhs(H, L[0], K=3, seed=1)
    solv = oracle.Solvers(L, oracle.full_matrix(L[-1]))
    xg, bg = torch.from_numpy(x0.copy()).cuda(), torch.from_numpy(b).cuda()
    xo = x0.copy()
    for cycle in range(1, 6):
        S.vcycle(xg, bg)
        oracle.vcycle(L, solv, xo, b)
        e = xg.cpu().numpy() - xo
        rms = np.linalg.norm(e) / np.linalg.norm(xo)
        print(f"[c1] cycle {cycle}: rel RMS vs oracle {rms:.2e}", flush=True)
        assert rms < 1e-9
    A = oracle.full_matrix(L[0]).tocsc()
    want = spla.splu(A).solve(b.T).T
    xm = torch.from_numpy(x0.copy()).cuda()
    its, rel = S.solve(xm, bg, tol=1e-10, max_iters=40, mode="mixed")
    err = xm.cpu().numpy() - want
    energy = np.sqrt(np.einsum("ij,ij->i", err, (A @ err.T).T) / np.einsum("ij,ij->i", want, (A @ want.T).T))
    print(f"[c1] mixed: {its} iterations, rel res {rel.max():.1e}, energy error vs direct {energy.max():.1e}", flush=True)
    assert rel.max() <= 1e-10 and energy.max() < 1e-8

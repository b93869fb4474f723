"""The serving workload (C5) under test: one resident mesh per size class (cube 512^2, icosphere
1024^2, 2048^2 and 4096^2, a torus at 1024^2 with line integral convolution), a 60 / 30 / 10 mix
of filtering (stiffness log-uniform in [1e-4, 1e-2], 3 channels), geodesic (1-8 sources) and LIC
requests at rel-res 1e-6, submitted concurrently from two threads; then every request served
alone.  Checks: every filtering and geodesic result meets 1e-6 against its own operator (FP64
residual recomputed on the device) both in the mix and alone, the two agree to the tolerance
(filtering: relative difference <= 1e-4; geodesic distances, which pass the heat solution through
a normalisation: <= 1e-3), and LIC results (the reference's fixed rounds) are identical."""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import gradientdomaintextureprocessing_amd as g  # noqa: E402
from gradientdomaintextureprocessing_amd.serving import SolverServer, latency_percentiles  # noqa: E402

TOL = 1e-6


def _meshes():
    out = {}
    out["cube512"] = (g.TexturedMesh.cube(16, 512, 512), 512, None)
    out["ico1024"] = (g.TexturedMesh.icosphere(5, 1024, 1024, 0.05, 2), 1024, None)
    out["ico2048"] = (g.TexturedMesh.icosphere(6, 2048, 2048, 0.05, 4), 2048, None)
    out["ico4096"] = (g.TexturedMesh.icosphere(7, 4096, 4096, 0.05, 1), 4096, None)
    torus = g.TexturedMesh.torus(256, 128, 8, 4, 1024, 1024, 1.0, 0.35, 0.05, 3)
    v, t, _ = torus.arrays()
    out["torus1024"] = (torus, 1024, g.surface_gradient_field(v, t, np.sin(3 * v[:, 0]) + np.cos(2 * v[:, 1]) + np.sin(4 * v[:, 2])))
    return out


def _filter_residual(m, f, a, x):
    """||M f - (M + a S) x|| / ||M f|| per column on the mesh's level 0 (FP64, on the device)."""
    S = m.systems[m.bucket(a)]
    mf, mx, sx = torch.empty_like(f), torch.empty_like(x), torch.empty_like(x)
    S.apply(g.OPERATOR_MASS, f, mf)
    S.apply(g.OPERATOR_MASS, x, mx)
    S.apply(g.OPERATOR_STIFFNESS, x, sx)
    return ((mf - mx - a * sx).norm(dim=1) / mf.norm(dim=1)).max().item()


def test_c5_mix_alone_and_in_the_mix():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    t0 = time.time()
    server = SolverServer(tol=TOL, max_iters=150)
    try:
        sizes = {}
        for name, (mesh, W, field) in _meshes().items():
            levels = max(4, int(np.log2(W)) - 4)
            lic = g.LineIntegralConvolution(mesh, W, W, levels, field, max_cols=4, assembly="device") if field is not None else None
            H = g.InitializeHierarchy(mesh, W, W, levels, assembly="device")
            server.add_mesh(name, H, max_cols=4, lic=lic)
            sizes[name] = H.level_info(0)["texels"]
            print(f"[c5] {name} resident: {sizes[name]} texels, {time.time() - t0:.0f}s", flush=True)
        rng = np.random.default_rng(5)
        names = list(sizes)
        plan = []
        for i in range(30):
            u = rng.random()
            n = names[rng.integers(len(names))]
            if u < 0.6:
                plan.append(("filter", n, torch.from_numpy(rng.random((3, sizes[n]))).cuda(), float(10 ** rng.uniform(-4, -2))))
            elif u < 0.9:
                plan.append(("geodesic", n, [int(s) for s in rng.integers(0, sizes[n], rng.integers(1, 5))], None))
            else:
                plan.append(("lic", "torus1024", torch.from_numpy(rng.random((2, sizes["torus1024"]))).cuda(), None))

        def submit(kind, n, p, a):
            if kind == "filter":
                return server.submit_filter(n, p, a)
            return server.submit_geodesic(n, p) if kind == "geodesic" else server.submit_lic(n, p)

        handles = [None] * len(plan)

        def feed(idx):
            for i in idx:
                handles[i] = submit(*plan[i])
                time.sleep(rng.exponential(0.01))

        ts = [threading.Thread(target=feed, args=(range(k, len(plan), 2),)) for k in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        mixed = [h.wait(600) for h in handles]
        print(f"[c5] mix of {len(plan)} requests served; latency {latency_percentiles(handles)}", flush=True)
        alone = [submit(*r).wait(600) for r in plan]
        for (kind, n, p, a), xm, xa in zip(plan, mixed, alone):
            m = server.meshes[n]
            if kind == "filter":
                assert _filter_residual(m, p, a, xm) <= 1.01 * TOL and _filter_residual(m, p, a, xa) <= 1.01 * TOL
                rel = ((xm - xa).norm(dim=1) / xa.norm(dim=1)).max().item()
                assert rel <= 1e-4, (n, a, rel)
            elif kind == "geodesic":
                rel = ((xm - xa).norm(dim=1) / xa.norm(dim=1)).max().item()
                print(f"[c5] geodesic {n}: mix vs alone {rel:.1e}", flush=True)
                assert rel <= 1e-3, (n, rel)
            else:
                assert torch.equal(xm, xa)
        print(f"[c5] every result matches alone; total {time.time() - t0:.0f}s", flush=True)
    finally:
        server.close()

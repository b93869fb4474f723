"""Benchmark: screened-Poisson texture filtering solves/s at relative residual 1e-6.

Workload (default): noisy icosphere (327,680 triangles, 20 charts) at a 4096x4096 atlas, 8 levels,
(M + 1e-3 S) x = M f with f a seeded uniform[0,1) RGB texture (float32-stored), three right-hand
sides.  One step = one solve of all three columns from x = 0 until every column's relative
residual ||b - A x|| / ||b|| <= 1e-6; one solve = one column (so a step is 3 solves).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode mixed|mixed-only|fp64] [--config c2|c1|small]
       python bench.py --profile-kernel mixed|band   (calibration copy + the dominant kernel only, for
                                                      rocprofv3 --pmc traffic counts)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: (mesh kind, params, width, FP64 reference-schedule levels, mixed-solver levels, stiffness).
    # Both solvers run on the same hierarchy depth: at 4096^2 the reference's default of 4 levels
    # leaves a 198k-texel coarsest level whose exact solve would dominate the FP64 baseline.
    "c2": ("icosphere", dict(levels=7, amplitude=0.05, seed=1), 4096, 8, 8, 1e-3),
    "c1": ("cube", dict(subdiv=16), 512, 4, 6, 1e-3),
    "small": ("icosphere", dict(levels=5, amplitude=0.05, seed=1), 1024, 7, 7, 1e-3),
}

HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_system(cfg, K, mixed):
    import gradientdomaintextureprocessing_amd as g
    kind, p, W, ref_levels, mix_levels, stiff = CONFIGS[cfg]
    levels = mix_levels if mixed else ref_levels
    t0 = time.time()
    if kind == "icosphere":
        mesh = g.TexturedMesh.icosphere(p["levels"], W, W, p["amplitude"], p["seed"])
    else:
        mesh = g.TexturedMesh.cube(p["subdiv"], W, W)
    H = g.InitializeHierarchy(mesh, W, W, levels)
    t1 = time.time()
    S = g.UpdateLinearSystem(H, stiff, 1.0, max_cols=K)
    t2 = time.time()
    log(f"[bench] setup ({levels} levels): hierarchy+assembly {t1 - t0:.1f}s, coarsening+factorisation+upload {t2 - t1:.1f}s")
    return mesh, H, S


def texture_rhs(H, S, K, seed=1):
    import torch
    tex = np.random.default_rng(seed).random((H.height, H.width, K), dtype=np.float32)
    ci, cj, _ = H.texel_coordinates(0)
    f = torch.from_numpy(np.ascontiguousarray(tex[cj, ci, :].T.astype(np.float64))).cuda()
    b = torch.empty_like(f)
    S.apply(1, f, b)  # b = M f
    return f, b


def solve(S, x, b, tol, max_cycles, mode):
    """One step: from x = 0 until every column reaches tol (reference V-cycles or mixed refinement).
    The solver writes the zeros itself (zero_guess), and the mixed path's first residual is b."""
    its, rel = S.solve(x, b, tol=tol, max_iters=max_cycles, mode=mode, zero_guess=True)
    return its, float(rel.max())


def mixed_roofline(S, H, K, reps=10):
    """Average duration of one launch of the deep Gauss-Seidel sweep (dominant kernel of the mixed
    path) on level 0, and its algorithmic bytes.  With a fused tile shape one launch is a whole
    four-colour sweep (KMixDeepGsFused, for K <= 4); with shape -1 it is one colour plane
    (KMixDeepGs).  B_relax = s*q + 3*K*p per updated deep texel (stencil once, e read, rhs read,
    e written) with the declared FP16-ratio stencil s*q = 8*2 + 2*4 = 24 bytes and p = 4."""
    import torch
    info = H.level_info(0)
    shape = S.tile_shape(level=0)
    fused = shape >= 0
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    S.mixed_deep_sweep(0, K)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        S.mixed_deep_sweep(0, K)
    e1.record(st)
    torch.cuda.synchronize()
    launches = (K + 3) // 4 * reps if fused else 4 * reps
    per_launch_s = e0.elapsed_time(e1) / 1e3 / launches
    updated = info["deep"] if fused else info["deep"] / 4  # deep texels one launch updates
    bytes_per_launch = updated * (24 + 3 * K * 4)
    achieved = bytes_per_launch / per_launch_s
    return dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK, traffic=None,
                kernel="KMixDeepGsFused" if fused else "KMixDeepGs", tile_shape=S.TILE_SHAPES[shape],
                avg_launch_us=per_launch_s * 1e6, algorithmic_bytes_per_launch=bytes_per_launch)


def band_roofline(S, H, x, b, r, reps=5):
    """Average duration of the band Gauss-Seidel kernel (dominant kernel of the FP64 path) and its
    algorithmic bytes: per deep texel one 10-word stencil read, b and x read, x and r written."""
    import torch
    info = H.level_info(0)
    K = x.shape[0]
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    S.band_sweep(0, x, b, r)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        S.band_sweep(0, x, b, r)
    e1.record(st)
    torch.cuda.synchronize()
    launches = 2 * reps
    per_launch_s = e0.elapsed_time(e1) / 1e3 / launches
    bytes_per_launch = info["deep"] / 2 * (10 * 8 + 4 * K * 8)
    achieved = bytes_per_launch / per_launch_s
    return dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK, traffic=None,
                kernel="KBandGs", avg_launch_us=per_launch_s * 1e6, algorithmic_bytes_per_launch=bytes_per_launch)


def error_vs_direct(S, x, b):
    """Relative energy-norm error of the solution x against the FP64 direct-equivalent solution of
    the same system (gdtp_direct_solve: at this size FP64 reference-schedule V-cycles from zero to
    the FP64 residual floor; a sparse factorisation of the 12 M-texel level does not fit the run);
    both relative residuals recomputed in FP64 (gdtp_residual_norms)."""
    import torch
    import gradientdomaintextureprocessing_amd as g
    ref = torch.empty_like(b)
    rel_ref, method = S.direct_solve(ref, b)
    rel_x, _, _ = S.residual_norms(x, b)
    e = x - ref
    ae, aref = torch.empty_like(e), torch.empty_like(ref)
    S.apply(g.OPERATOR_SYSTEM, e, ae)
    S.apply(g.OPERATOR_SYSTEM, ref, aref)
    err = torch.sqrt((e * ae).sum(dim=1) / (ref * aref).sum(dim=1)).cpu().numpy()
    return dict(energy_rel_max=float(err.max()), energy_rel=[float(v) for v in err], rel_res=[float(v) for v in rel_x],
                reference=("LDL^T of the full level-0 matrix" if method == 0 else "FP64 reference-schedule V-cycles to the FP64 floor")
                + f", rel-res {float(rel_ref.max()):.2e}")


COPY_BYTES = 1 << 28


def profile_kernel(cfg, K, which, reps=10, shape=1):
    """Kernel-only run for hardware-counter passes: four copies of a 256 MiB buffer (the FETCH_SIZE /
    WRITE_SIZE calibration: each reads and writes exactly COPY_BYTES), then `reps` launches of the
    dominant kernel of the chosen path on the benchmark workload."""
    import torch
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    mesh, H, S = make_system(cfg, K, mixed=(which != "band"))
    if which != "band":  # "mixed": one launch per colour; "fused": the fused sweep on 64 x 4 tiles
        S.tile_shape(-1 if which == "mixed" else shape)
        f, b = texture_rhs(H, S, K)
        x = torch.zeros_like(b)
        solve(S, x, b, 1e-6, 40, "mixed")  # leaves meaningful FP32 vectors in the solver's buffers
        for _ in range(reps):
            S.mixed_deep_sweep(0, K)
    else:
        f, b = texture_rhs(H, S, K)
        x, r = f.clone(), torch.zeros_like(b)
        for _ in range(reps):
            S.band_sweep(0, x, b, r)
    torch.cuda.synchronize()
    log(f"[bench] profile-kernel {which}: {reps} launches done")


SERVING_MESHES = [  # name, kind, params, width (tori also serve line integral convolution)
    ("cube512", "cube", dict(subdiv=16), 512),
    ("cube1024", "cube", dict(subdiv=16), 1024),
    ("ico1024", "icosphere", dict(levels=5, amplitude=0.05, seed=2), 1024),
    ("ico2048", "icosphere", dict(levels=6, amplitude=0.05, seed=4), 2048),
    ("ico4096", "icosphere", dict(levels=7, amplitude=0.05, seed=1), 4096),
    ("torus512", "torus", dict(nu=128, nv=64, cu=4, cv=2, amplitude=0.05, seed=6), 512),
    ("torus1024", "torus", dict(nu=256, nv=128, cu=8, cv=4, amplitude=0.05, seed=3), 1024),
    ("torus2048", "torus", dict(nu=512, nv=256, cu=8, cv=4, amplitude=0.05, seed=5), 2048),
]


def serving_bench(qps, n_requests, tol, seed=0):
    """Request-loop measurement: the meshes above stay resident; requests arrive as a Poisson
    stream at `qps`: 60 % filtering (stiffness log-uniform in [1e-4, 1e-2], 3 channels, a uniform
    [0,1) texture), 30 % geodesics (1-8 random source texels), 10 % line integral convolution on a
    torus (3 noise channels, the reference's 4 rounds).  Reports p50/p99 latency per class and the
    columns per second."""
    import threading
    import torch
    import gradientdomaintextureprocessing_amd as g
    from gradientdomaintextureprocessing_amd.serving import SolverServer, latency_percentiles
    t0 = time.time()
    server = SolverServer(tol=tol, max_iters=120)
    sizes = {}
    for name, kind, p, W in SERVING_MESHES:
        levels = max(4, int(np.log2(W)) - 4)
        lic = None
        if kind == "cube":
            mesh = g.TexturedMesh.cube(p["subdiv"], W, W)
        elif kind == "icosphere":
            mesh = g.TexturedMesh.icosphere(p["levels"], W, W, p["amplitude"], p["seed"])
        else:
            mesh = g.TexturedMesh.torus(p["nu"], p["nv"], p["cu"], p["cv"], W, W, 1.0, 0.35, p["amplitude"], p["seed"])
            v, t, _ = mesh.arrays()
            ph = np.random.default_rng(p["seed"]).random(3) * 6.28
            field = g.surface_gradient_field(v, t, np.sin(3 * v[:, 0] + ph[0]) + np.cos(2 * v[:, 1] + ph[1]) + np.sin(4 * v[:, 2] + ph[2]))
            lic = g.LineIntegralConvolution(mesh, W, W, levels, field, max_cols=8, assembly="device")
        H = g.InitializeHierarchy(mesh, W, W, levels, assembly="device")
        server.add_mesh(name, H, max_cols=8, lic=lic)
        sizes[name] = H.level_info(0)["texels"]
        log(f"[bench] serving: {name} resident ({sizes[name]} texels, {levels} levels), {time.time() - t0:.1f}s")
    log(f"[bench] serving: {len(SERVING_MESHES)} meshes resident in {time.time() - t0:.1f}s, {server.used / 2**30:.2f} GiB")
    rng = np.random.default_rng(seed)
    names = [m[0] for m in SERVING_MESHES]
    plan = []
    tori = [m[0] for m in SERVING_MESHES if m[1] == "torus"]
    for i in range(n_requests):
        n = names[rng.integers(len(names))]
        u = rng.random()
        if u < 0.6:
            plan.append(("filter", n, float(10 ** rng.uniform(-4, -2))))
        elif u < 0.9:
            plan.append(("geodesic", n, [int(v) for v in rng.integers(0, sizes[n], rng.integers(1, 9))]))
        else:
            plan.append(("lic", tori[rng.integers(len(tori))], None))
    payloads = {n: torch.rand(3, sizes[n], dtype=torch.float64, device="cuda") for n in names}
    for n in names:  # warm every (mesh, kind) path once: graph capture and first-use set-up
        server.submit_filter(n, payloads[n], 1e-3).wait(300)
        server.submit_geodesic(n, [0]).wait(300)
        if n in tori:
            server.submit_lic(n, payloads[n]).wait(300)
    gaps = rng.exponential(1.0 / qps, n_requests)
    handles = []
    torch.cuda.synchronize()
    start = time.perf_counter()
    for (kind, n, a), gap in zip(plan, gaps):
        time.sleep(gap)
        if kind == "filter":
            h = server.submit_filter(n, payloads[n], a)
        elif kind == "geodesic":
            h = server.submit_geodesic(n, a)
        else:
            h = server.submit_lic(n, payloads[n])
        handles.append((kind, n, h))
    for _, _, h in handles:
        h.wait(600)
    el = time.perf_counter() - start
    server.close()
    cols = sum(h.columns for _, _, h in handles)
    out = {"metric": "serving: request latency at rel-residual %g" % tol, "qps_offered": qps, "requests": n_requests,
           "wall_s": el, "columns_per_s": cols / el, "resident_gib": server.used / 2**30,
           "all": latency_percentiles([h for _, _, h in handles])}
    for kind in ("filter", "geodesic", "lic"):
        out[kind] = latency_percentiles([h for k, _, h in handles if k == kind])
        its = [h.iterations for k, _, h in handles if k == kind]
        if its:
            out[kind]["iterations_p50_max"] = [float(np.percentile(its, 50)), int(max(its))]
    out["filter_1024"] = latency_percentiles([h for k, n, h in handles if k == "filter" and n.endswith("1024")])
    out["meshes"] = {n: sizes[n] for n in names}
    print(json.dumps(out))


SERVING_LEG_MESHES = [m for m in SERVING_MESHES if m[0] in ("cube512", "ico1024", "torus1024")]


def serving_leg(qps=20.0, n_requests=150, tol=1e-6, seed=1):
    """Short request-loop measurement for the default run: three resident meshes (cube 512^2,
    noisy icosphere 1024^2, torus 1024^2 with line integral convolution), Poisson arrivals at `qps`
    mixing 60 % filtering, 30 % geodesics and 10 % LIC as in serving_bench.  Returns the 1024^2
    filtering latency percentiles, all-request percentiles and columns per second."""
    import torch
    import gradientdomaintextureprocessing_amd as g
    from gradientdomaintextureprocessing_amd.serving import SolverServer, latency_percentiles
    t0 = time.time()
    server = SolverServer(tol=tol, max_iters=120)
    sizes = {}
    try:
        for name, kind, p, W in SERVING_LEG_MESHES:
            levels = max(4, int(np.log2(W)) - 4)
            lic = None
            if kind == "cube":
                mesh = g.TexturedMesh.cube(p["subdiv"], W, W)
            elif kind == "icosphere":
                mesh = g.TexturedMesh.icosphere(p["levels"], W, W, p["amplitude"], p["seed"])
            else:
                mesh = g.TexturedMesh.torus(p["nu"], p["nv"], p["cu"], p["cv"], W, W, 1.0, 0.35, p["amplitude"], p["seed"])
                v, t, _ = mesh.arrays()
                ph = np.random.default_rng(p["seed"]).random(3) * 6.28
                field = g.surface_gradient_field(v, t, np.sin(3 * v[:, 0] + ph[0]) + np.cos(2 * v[:, 1] + ph[1]) + np.sin(4 * v[:, 2] + ph[2]))
                lic = g.LineIntegralConvolution(mesh, W, W, levels, field, max_cols=8, assembly="device")
            H = g.InitializeHierarchy(mesh, W, W, levels, assembly="device")
            server.add_mesh(name, H, max_cols=8, lic=lic)
            sizes[name] = H.level_info(0)["texels"]
        admit = time.time() - t0
        log(f"[bench] serving leg: {len(sizes)} meshes resident in {admit:.1f}s")
        rng = np.random.default_rng(seed)
        names = list(sizes)
        payloads = {n: torch.rand(3, sizes[n], dtype=torch.float64, device="cuda") for n in names}
        plan = []
        for _ in range(n_requests):
            n = names[rng.integers(len(names))]
            u = rng.random()
            if u < 0.6:
                plan.append(("filter", n, float(10 ** rng.uniform(-4, -2))))
            elif u < 0.9:
                plan.append(("geodesic", n, [int(v) for v in rng.integers(0, sizes[n], rng.integers(1, 9))]))
            else:
                plan.append(("lic", "torus1024", None))
        gaps = rng.exponential(1.0 / qps, n_requests)
        handles = []
        torch.cuda.synchronize()
        start = time.perf_counter()
        for (kind, n, a), gap in zip(plan, gaps):
            time.sleep(gap)
            if kind == "filter":
                handles.append((kind, n, server.submit_filter(n, payloads[n], a)))
            elif kind == "geodesic":
                handles.append((kind, n, server.submit_geodesic(n, a)))
            else:
                handles.append((kind, n, server.submit_lic(n, payloads[n])))
        for _, _, h in handles:
            h.wait(600)
        el = time.perf_counter() - start
    finally:
        server.close()
    cols = sum(h.columns for _, _, h in handles)
    return dict(qps_offered=qps, requests=n_requests, wall_s=el, columns_per_s=cols / el, admission_s=admit,
                meshes={n: sizes[n] for n in names}, tol=tol,
                filter_1024=latency_percentiles([h for k, n, h in handles if k == "filter" and n.endswith("1024")]),
                all=latency_percentiles([h for _, _, h in handles]))


def load_traffic(kernel, workload_key):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE and WRITE_SIZE passes, scaled
    by the calibration copy) and recorded in profiles/traffic.json; None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh).get(kernel)
    if not rec or rec.get("workload") != workload_key:
        return None
    return rec["bytes_per_launch"]


def host_cpu():
    """Host CPU model, sockets, cores and threads (lscpu)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return dict(model=info.get("Model name"), sockets=info.get("Socket(s)"), cores_per_socket=info.get("Core(s) per socket"),
                threads_per_core=info.get("Thread(s) per core"), cpus=info.get("CPU(s)"))


def cpu_baseline(H, S, K, tol, runs=5):
    """Oracle (plain-C restatement of the reference V-cycle, OpenMP over same-colour band tasks and
    rows, exact sparse solves per column in parallel) on the host's cores: one solve of the same
    workload from x = 0 to `tol` counts the V-cycles n (and warms up), then `runs` single V-cycles
    are timed with the wall clock; value = K / (n x median V-cycle time)."""
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(threads)
    L = [oracle.level_from_product(H, S, l) for l in range(H.levels)]
    t0 = time.perf_counter()
    solv = oracle.Solvers(L, oracle.full_matrix(L[-1]))
    A = oracle.full_matrix(L[0])
    setup = time.perf_counter() - t0
    tex = np.random.default_rng(1).random((H.height, H.width, K), dtype=np.float32)
    ci, cj, _ = H.texel_coordinates(0)
    f = np.ascontiguousarray(tex[cj, ci, :].T.astype(np.float64))
    bd = oracle.Csr(H.array(0, "mass.bd.rowptr"), H.array(0, "mass.bd.col"), H.array(0, "mass.bd.val", np.float64), L[0].N)
    bb = oracle.Csr(H.array(0, "mass.bb.rowptr"), H.array(0, "mass.bb.col"), H.array(0, "mass.bb.val", np.float64), L[0].B)
    b = oracle.apply(L[0], H.array(0, "mass.deep", np.float64), bd, bb, f)
    x = np.zeros_like(b)
    bn = np.linalg.norm(b, axis=1)
    cycles, rel = 0, 1.0
    while rel > tol and cycles < 40:
        oracle.vcycle(L, solv, x, b)
        cycles += 1
        rel = float((np.linalg.norm(b - (A @ x.T).T, axis=1) / bn).max())
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.vcycle(L, solv, x, b)
        times.append(time.perf_counter() - t0)
    per = float(np.median(times))
    return dict(value=K / (cycles * per), unit="solves/s", cores=threads, kind="port",
                sample=f"oracle V-cycle on the same {K}-column system: {cycles} cycles reach rel-res {rel:.1e}; "
                       f"median of {runs} timed V-cycles {per:.3f} s (all {threads} threads, OMP_PROC_BIND="
                       f"{os.environ['OMP_PROC_BIND']}, OMP_PLACES={os.environ['OMP_PLACES']}); setup {setup:.1f} s not counted",
                vcycle_s=[float(t) for t in times], host=host_cpu())


def geodesics_bench(steps, warmup, tol, sources=64, seed=7):
    """Heat-method geodesics batched as columns: noisy icosphere (327,680 triangles, amplitude
    0.05) on a 2048^2 atlas, 64 seeded source texels; (M + t S) u = delta with the reference's
    t = 1e-3 of the unit-area surface (Geodesics.cpp:19; with t = h^2 the heat decays below any
    solve tolerance within a few hundred texels and the far-field gradient direction is noise),
    normalised-gradient divergence, (1e-6 M + S) phi = div, phi -= phi[source]; every solve to
    `tol`.  One step = all 64 sources."""
    import torch
    import gradientdomaintextureprocessing_amd as g
    t0 = time.time()
    mesh = g.TexturedMesh.icosphere(7, 2048, 2048, 0.05, 3)
    H = g.InitializeHierarchy(mesh, 2048, 2048, 7)
    N = H.level_info(0)["texels"]
    geo = g.Geodesics(H, heat_time=1e-3, max_sources=sources)
    log(f"[bench] geodesics setup {time.time() - t0:.1f}s, {N} texels")
    src = [int(v) for v in np.random.default_rng(seed).choice(N, sources, replace=False)]
    lat = []
    for step in range(warmup + steps):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        phi = geo.distances(src, tol=tol, mode="mixed", max_iters=200)
        torch.cuda.synchronize()
        if step >= warmup:
            lat.append(time.perf_counter() - ts)
    ms = float(np.mean(lat)) * 1e3
    return dict(metric=f"geodesic sources/s at rel-residual {tol:g} (2048² noisy icosphere)",
                value=sources / (ms / 1e3), unit="sources/s", n_gpus=1, steps=steps, warmup=warmup, ms_per_step=ms,
                higher_is_better=True, dtype="f32 relax / f64 refine", data="synthetic",
                config=dict(workload=f"c3: heat-method geodesics, {N} texels, {sources} sources as columns, t=1e-3",
                            max_abs_phi=float(phi.abs().max())),
                latency_ms=dict(p50=float(np.percentile(lat, 50)) * 1e3, p99=float(np.percentile(lat, 99)) * 1e3))


def lic_bench(steps, warmup, columns=16, rounds=4, seed=9):
    """Line integral convolution as anisotropic diffusion: noisy torus (R 1, r 0.35, 262,144
    triangles, 32 charts) on a 2048^2 atlas, field = surface gradient of seeded smooth noise,
    float32 white-noise input, 16 diffusion lengths as columns, each with its own operator
    (M_a + l_c S_a) x_c = M_a noise_c, l_c log-spaced over [1e-5, 1e-3]
    (LineIntegralConvolution.cpp:503-527, 611-619).  Geometric multigrid does not converge on this
    anisotropy (the reference runs a fixed number of rounds), so one step is `rounds` iterations of
    the weighted solve (flexible CG, one V-cycle of the system's own l = 1e-4 hierarchy per
    iteration), with each column's relative residual after them reported, and column 3 solved alone
    compared with the batch."""
    import torch
    import gradientdomaintextureprocessing_amd as g
    t0 = time.time()
    mesh = g.TexturedMesh.torus(512, 256, 8, 4, 2048, 2048, 1.0, 0.35, 0.05, 5)
    v, t, _ = mesh.arrays()
    rng = np.random.default_rng(seed)
    ph = rng.random(4) * 6.28
    f = np.sin(3 * v[:, 0] + ph[0]) + np.cos(2 * v[:, 1] + ph[1]) + np.sin(4 * v[:, 2] + ph[2]) + 0.5 * np.sin(5 * v[:, 0] * v[:, 1] + ph[3])
    lic = g.LineIntegralConvolution(mesh, 2048, 2048, 7, g.surface_gradient_field(v, t, f), max_cols=columns)
    S = lic.lic
    N = lic.anisotropic.level_info(0)["texels"]
    log(f"[bench] lic setup {time.time() - t0:.1f}s, {N} texels")
    lengths = np.logspace(-5, -3, columns)
    noise = torch.from_numpy(rng.random((columns, N), dtype=np.float32).astype(np.float64)).cuda()
    lat, x, rel = [], None, None
    for step in range(warmup + steps):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        x, rel = lic.convolve_lengths(noise, lengths, rounds=rounds)
        torch.cuda.synchronize()
        if step >= warmup:
            lat.append(time.perf_counter() - ts)
    xa, _ = lic.convolve_lengths(noise[3:4].contiguous(), lengths[3:4], rounds=rounds)
    alone = float((x[3] - xa[0]).abs().max() / x[3].abs().max())
    ms = float(np.mean(lat)) * 1e3
    return dict(metric="LIC columns/s, 16 diffusion lengths, reference round count (2048² noisy torus)",
                value=columns / (ms / 1e3), unit="columns/s", n_gpus=1, steps=steps, warmup=warmup, ms_per_step=ms,
                higher_is_better=True, dtype="f64 operator / f32 V-cycle", data="synthetic",
                config=dict(workload=f"c4: anisotropic diffusion, {N} texels, {columns} columns, one length each, {rounds} rounds",
                            lengths=[float(l) for l in lengths], rel_res_after=[float(r) for r in rel],
                            column3_alone_vs_batch=alone),
                latency_ms=dict(p50=float(np.percentile(lat, 50)) * 1e3, p99=float(np.percentile(lat, 99)) * 1e3))


def assembly_bench():
    """Mesh admission: hierarchy + finest-level assembly of the benchmark mesh (noisy icosphere,
    327,680 triangles, 4096^2 atlas, 8 levels), interior cells integrated on the host (OpenMP over
    charts) or on the GPU; best of two runs each."""
    import gradientdomaintextureprocessing_amd as g
    mesh = g.TexturedMesh.icosphere(7, 4096, 4096, 0.05, 1)
    times = {}
    for assembly in ("host", "device", "host", "device"):
        t0 = time.perf_counter()
        H = g.InitializeHierarchy(mesh, 4096, 4096, 8, assembly=assembly)
        el = time.perf_counter() - t0
        times[assembly] = min(times.get(assembly, 1e9), el)
        log(f"[bench] hierarchy + assembly ({assembly}): {el:.2f}s")
        del H
    return dict(metric="mesh admission: hierarchy + assembly seconds (4096² icosphere)", host_s=times["host"],
                device_s=times["device"], speedup=times["host"] / times["device"], cores=os.cpu_count(),
                config=dict(workload="icosphere 327,680 triangles, 4096^2 atlas, 8 levels"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="mixed", choices=["fp64", "mixed", "mixed-only"])
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--serving-leg", default="on", choices=["on", "off"],
                    help="short request-loop leg (three meshes, 20 QPS) reported beside the solve throughput")
    ap.add_argument("--profile-kernel", default=None, choices=["mixed", "fused", "band"])
    ap.add_argument("--tile-shape", type=int, default=1, help="fused tile shape index for --profile-kernel fused")
    ap.add_argument("--tile-autotune", default="on", choices=["on", "off"],
                    help="time every deep-sweep tile shape per level before the timed steps (off: the library's automatic shapes, "
                         "e.g. under a profiler, whose overhead would skew the choice)")
    ap.add_argument("--serving", action="store_true", help="request-loop benchmark (resident meshes, Poisson arrivals)")
    ap.add_argument("--workload", default=None, choices=["c3", "c4", "assembly"],
                    help="other workloads: c3 batched geodesics, c4 line integral convolution, assembly = mesh admission "
                         "time (one JSON line each)")
    ap.add_argument("--qps", type=float, default=40.0)
    ap.add_argument("--requests", type=int, default=300)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); across GPUs the solver would shard by "
              "requests / RHS columns with every mesh resident on each GPU, which needs no collective", file=sys.stderr)
        sys.exit(2)
    import torch
    K = 3
    max_cycles = 40
    if args.profile_kernel:
        profile_kernel(args.config, K, args.profile_kernel, shape=args.tile_shape)
        return
    if args.serving:
        serving_bench(args.qps, args.requests, args.tol)
        return
    if args.workload == "assembly":
        print(json.dumps(assembly_bench()), flush=True)
        return
    if args.workload:
        steps = min(args.steps, 10)
        res = geodesics_bench(steps, args.warmup, args.tol) if args.workload == "c3" else lic_bench(steps, args.warmup)
        print(json.dumps(res), flush=True)
        return
    results = {}
    modes = {"mixed": ["mixed", "fp64"], "mixed-only": ["mixed"], "fp64": ["fp64"]}[args.mode]
    for mode in modes:
        mesh, H, S = make_system(args.config, K, mixed=(mode == "mixed"))
        f, b = texture_rhs(H, S, K)
        x, r = torch.zeros_like(b), torch.zeros_like(b)
        tiles = None
        if mode == "mixed" and args.tile_autotune == "on":  # pick each level's deep-sweep tile shape on this system (K columns)
            solve(S, x, b, args.tol, max_cycles, mode)
            tiles = {f"level {l}": {str(S.TILE_SHAPES[k]): round(v, 2) for k, v in t.items()} for l, t in S.autotune_tiles(K).items()}
            log(f"[bench] tile shapes per level (us per sweep, fastest first): {tiles}")
        for _ in range(args.warmup):
            cycles, rel = solve(S, x, b, args.tol, max_cycles, mode)
            log(f"[bench] {mode} warmup: {cycles} iterations, rel res {rel:.2e}")
        ref_bytes, mix_bytes = S.memory()
        log(f"[bench] {mode} device memory: reference system {ref_bytes / 1e9:.2f} GB, mixed system {mix_bytes / 1e9:.2f} GB")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cyc, lat = [], []
        for _ in range(args.steps):
            ts = time.perf_counter()
            c, rel = solve(S, x, b, args.tol, max_cycles, mode)  # returns after the device finished
            lat.append((time.perf_counter() - ts) * 1e3)
            cyc.append(c)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res = dict(value=K * args.steps / el, ms_per_step=el / args.steps * 1e3, iterations=cyc, rel_res=rel, levels=H.levels,
                   latency_ms=dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99))), tiles=tiles)
        log(f"[bench] {mode}: {args.steps} steps in {el:.3f}s ({res['value']:.1f} solves/s), iterations/step {cyc}, "
            f"final rel res {rel:.2e}")
        if mode == "mixed":
            res["error_vs_direct"] = error_vs_direct(S, x, b)
            log(f"[bench] mixed error vs FP64 direct-equivalent: {res['error_vs_direct']}")
            res["roofline"] = mixed_roofline(S, H, K)
        else:
            res["roofline"] = band_roofline(S, H, x, b, r)
        res["roofline"]["traffic"] = load_traffic(res["roofline"]["kernel"], args.config)
        log(f"[bench] {mode} roofline: {res['roofline']}")
        if mode == "fp64" and args.cpu_baseline == "on":
            res["cpu"] = cpu_baseline(H, S, K, args.tol)
            log(f"[bench] cpu baseline: {res['cpu']}")
        results[mode] = (res, H)
        del S, H, mesh
    serving = None
    if args.serving_leg == "on" and args.config == "c2":
        serving = serving_leg()
        log(f"[bench] serving leg: {serving}")
    main_mode = modes[0]
    res, H = results[main_mode]
    value, ms, cyc, roof = res["value"], res["ms_per_step"], res["iterations"], res["roofline"]
    fp = results.get("fp64", (None, None))[0]
    cpu = fp.get("cpu") if fp else None
    info = H.level_info(0)
    out = {
        "metric": "solves/s at rel-residual 1e-6 (4096² atlas) vs FP64 GPU; p50/p99 req latency",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 relax (f16 stencil ratios) / f64 refine" if main_mode == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: {CONFIGS[args.config][0]} {CONFIGS[args.config][2]}^2 atlas, "
                               f"{info['texels']} texels, 3 RHS, (M+1e-3 S)x=Mf from x=0 to rel-res {args.tol:g}",
                   "mode": main_mode, "levels": H.levels, "iterations_per_step": cyc, "tile_autotune_us": res.get("tiles")},
        "latency_ms": res["latency_ms"],
        "serving": serving,
        "fp64_reference_gpu": {k: v for k, v in fp.items() if k in ("value", "ms_per_step", "iterations", "levels")} if fp else None,
        "speedup_vs_fp64_gpu": value / fp["value"] if fp else None,
        "roofline": roof, "cpu_baseline": cpu,
        "error_vs_direct": res.get("error_vs_direct"),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

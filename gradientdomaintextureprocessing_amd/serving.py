"""Request loop over resident meshes with continuous batching.

Each resident mesh keeps its hierarchy, its multigrid systems and their factorisations on the GPU
(one system per stiffness bucket, built once).  Requests name a mesh and carry their own weights:

  - "filter":   (M + a S) x = M f + s a S f for the request's a (TextureFiltering.cpp:562-718),
                x starts at f, one column per channel;
  - "geodesic": heat-method distances from source texels (Geodesics.cpp:508-550), one column per
                source;
  - "lic":      line integral convolution of a noise texture (LineIntegralConvolution.cpp:503-527)
                on meshes admitted with a LineIntegralConvolution object: the reference's schedule
                of `rounds` V-cycle rounds (anisotropic diffusion does not converge under geometric
                multigrid, so no tolerance applies), yielding to short requests after every V-cycle
                (its FP64 band kernels run for milliseconds and hold up the hardware queue they share
                with short requests); one column per channel.

Each mesh has a worker lane per kind (filtering, geodesics, LIC) with its own systems, buffers and
HIP streams, so that short filtering requests never queue behind geodesic or LIC batches.  A lane forms
batches from its queue as slots free up: requests are taken in arrival order while their columns
fit the system's capacity, every column keeping its own operator weights (parametric solves share
the bucket's V-cycle as preconditioner), so a request that arrives while others run joins the
next batch.  Short work — filtering on meshes of at most `short_texels` texels — runs concurrently
on all meshes; everything else (geodesics, LIC, filtering on large meshes) is heavy: at most
`heavy_concurrency` heavy batches run at a time over all meshes, and each holds back while any
short batch runs (at most `yield_ms`): host-side priorities for the short requests (HIP stream
priorities are not used).  Heavy batches also yield inside their solves, between refinement
iterations (gdtp_solve_yielding: the solve continues afterwards, flexible CG keeping its search
direction) or LIC rounds: while short requests are queued or running, the next iteration waits
(at most `yield_ms` each time), so a short request shares the GPU with at most one iteration.  Meshes are admitted against an HBM budget.  Results are the same alone
and in a batch up to the requested tolerance: every column is solved to it.
"""
import math
import queue
import threading
import time

import numpy as np

from . import _lib, api


class Request:
    def __init__(self, mesh, kind, payload, weight, sharpness=0.0):
        self.mesh, self.kind, self.payload, self.weight, self.sharpness = mesh, kind, payload, weight, sharpness
        self.submitted = time.perf_counter()
        self.done = threading.Event()
        self.result = None
        self.error = None
        self.iterations = 0
        self.latency = None

    @property
    def columns(self):
        return len(self.payload) if self.kind == "geodesic" else int(self.payload.shape[0])

    def wait(self, timeout=None):
        if not self.done.wait(timeout):
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise self.error
        return self.result


class ResidentMesh:
    """Hierarchy and systems of one mesh: filtering buckets (M + a_b S) and, for geodesics, a heat
    system (M + t S) and the distance system (1e-6 M + S)."""

    def __init__(self, name, hierarchy, buckets=(1e-4, 1e-3, 1e-2), max_cols=8, geodesics=True, heat_time=1e-3, lic=None, lic_rounds=4,
                 lic_mode="fp64"):
        self.name, self.hierarchy, self.max_cols, self.heat_time = name, hierarchy, max_cols, heat_time
        self.lic, self.lic_rounds = lic, lic_rounds  # api.LineIntegralConvolution (its own hierarchies and systems) or None
        self.lic_mode = lic_mode  # the reference's FP64 V-cycle rounds, or mixed-precision refinement steps
        self.buckets = sorted(buckets)
        self.systems = {a: api.UpdateLinearSystem(hierarchy, a, 1.0, max_cols=max_cols) for a in self.buckets}
        self.heat = api.UpdateLinearSystem(hierarchy, heat_time, 1.0, max_cols=max_cols) if geodesics else None
        self.distance = api.UpdateLinearSystem(hierarchy, 1.0, 1e-6, max_cols=max_cols) if geodesics else None
        self.N = hierarchy.level_info(0)["texels"]

    def all_systems(self):
        extra = [self.lic.lic, self.lic.sharpen] if self.lic is not None else []
        return list(self.systems.values()) + [s for s in (self.heat, self.distance) if s is not None] + extra

    def lanes(self):
        return ["filter", "geodesic"] + (["lic"] if self.lic is not None else [])

    def bucket(self, a):
        return min(self.buckets, key=lambda b: abs(math.log(a) - math.log(b)))

    def memory(self):
        total = 0
        for s in self.all_systems():
            r, m = s.memory()
            total += r + m
        return total


class SolverServer:
    """Continuous-batching solve server (see the module docstring).  tol: relative residual per
    column; hbm_budget: bytes of device memory the resident systems may hold (None: no limit)."""

    def __init__(self, tol=1e-6, hbm_budget=None, max_iters=80, yield_ms=100.0, heavy_concurrency=1, short_texels=1_500_000):
        import torch
        self.torch = torch
        self.tol, self.hbm_budget, self.max_iters, self.yield_ms = tol, hbm_budget, max_iters, yield_ms
        self.heavy = threading.BoundedSemaphore(heavy_concurrency) if heavy_concurrency else None
        self.short_texels = short_texels
        self.short_pending = 0  # short requests submitted and not finished (queued or running)
        self.running_lock = threading.Condition()
        self.meshes = {}
        self.queues = {}
        self.workers = {}
        self.used = 0
        self.stop = threading.Event()

    # -- residency
    def add_mesh(self, name, hierarchy, warm=True, **kw):
        """Admit a mesh.  Unless `buckets` is given, meshes of at most `short_texels` texels get five
        stiffness buckets log-spaced over [1e-4, 1e-2] (their filtering is the latency-critical
        class: a weight at most a quarter decade from its bucket keeps the weighted solve short),
        larger ones three."""
        if "buckets" not in kw:
            small = hierarchy.level_info(0)["texels"] <= self.short_texels
            kw["buckets"] = tuple(np.logspace(-4, -2, 5 if small else 3))
        m = ResidentMesh(name, hierarchy, **kw)
        for s in m.all_systems():  # builds the mixed solver now: first requests should not pay for it
            s.configure_mixed(krylov=s is m.heat or s is m.distance)
        mem = m.memory()
        if self.hbm_budget is not None and self.used + mem > self.hbm_budget:
            del m
            raise MemoryError(f"mesh {name} needs {mem / 2**30:.2f} GiB; {(self.hbm_budget - self.used) / 2**30:.2f} GiB left")
        self.used += mem
        if warm:
            self._warm(m)
        self.meshes[name] = m
        for lane in m.lanes():
            self.queues[(name, lane)] = queue.Queue()
            t = threading.Thread(target=self._work, args=(name, lane), daemon=True)
            self.workers[(name, lane)] = t
            t.start()
        return m

    def _warm(self, m):
        """Capture the refinement graphs of every batch width each lane will hand its systems (the
        per-lane buffers make the pointers of a width fixed), so that no request pays a capture."""
        torch = self.torch
        g = torch.Generator(device="cuda").manual_seed(0)
        rand = lambda K: torch.rand(K, m.N, dtype=torch.float64, device="cuda", generator=g)
        fb = self._buffers(m, "filter")
        for K in range(1, m.max_cols + 1):
            x, b = fb["x"][:K], fb["b"][:K]
            for a, system in m.systems.items():
                b.copy_(rand(K))
                x.zero_()
                system.solve_weighted(x, b, [[1.0, a]] * K, tol=1e-2, max_iters=self.max_iters)
            if m.heat is not None:
                gb = self._buffers(m, "geodesic")
                x, b, t0, t1 = gb["x"][:K], gb["b"][:K], gb["t0"][:K], gb["t1"][:K]
                b.copy_(rand(K))
                x.zero_()
                m.heat.solve(x, b, tol=1e-2, max_iters=self.max_iters, mode="mixed")
                # like a divergence: orthogonal to constants (the near-null mode of 1e-6 M + S)
                t0.copy_(rand(K))
                t0 -= t0.mean(dim=1, keepdim=True)
                t1.zero_()
                m.distance.solve(t1, t0, tol=1e-2, max_iters=self.max_iters, mode="mixed")
        if m.lic is not None:  # every batch width's buffers and captured steps
            for K in range(1, m.max_cols + 1):
                m.lic.convolve(rand(K), passes=1, mode=m.lic_mode)
        torch.cuda.synchronize()

    # -- requests
    def submit_filter(self, mesh, texels, stiffness, sharpness=0.0):
        """texels: [channels, N] FP64 GPU tensor."""
        r = Request(mesh, "filter", texels, stiffness, sharpness)
        self._enqueue(r)
        return r

    def submit_geodesic(self, mesh, sources):
        r = Request(mesh, "geodesic", list(sources), self.meshes[mesh].heat_time)
        self._enqueue(r)
        return r

    def submit_lic(self, mesh, noise):
        """noise: [channels, N] FP64 GPU tensor (texels of the mesh's anisotropic hierarchy, which
        has the same atlas and texel order)."""
        r = Request(mesh, "lic", noise, 0.0)
        self._enqueue(r)
        return r

    def _enqueue(self, r):
        m = self.meshes.get(r.mesh)
        if m is None:
            raise KeyError(f"mesh {r.mesh} is not resident")
        if r.columns < 1 or r.columns > m.max_cols:
            raise ValueError(f"a request needs 1..{m.max_cols} columns")
        if r.kind == "geodesic" and m.distance is None:
            raise ValueError("mesh was added without geodesics")
        if r.kind == "lic" and m.lic is None:
            raise ValueError("mesh was added without a line integral convolution")
        if self._short(m, r.kind):
            with self.running_lock:
                self.short_pending += 1
        self.queues[(r.mesh, r.kind)].put(r)

    def _short(self, m, lane):
        return lane == "filter" and m.N <= self.short_texels

    def close(self):
        self.stop.set()
        for name, q in self.queues.items():
            q.put(None)
        for t in self.workers.values():
            t.join(timeout=60)

    # -- the per-mesh loop
    def _work(self, name, lane):
        torch = self.torch
        m = self.meshes[name]
        q = self.queues[(name, lane)]
        stream = torch.cuda.Stream()
        pending = []
        while not self.stop.is_set():
            if not pending:
                r = q.get()
                if r is None:
                    return
                pending.append(r)
            while True:  # everything that arrived meanwhile joins the candidates
                try:
                    r = q.get_nowait()
                except queue.Empty:
                    break
                if r is None:
                    self.stop.set()
                    break
                pending.append(r)
            batch, rest = self._form_batch(m, pending)
            pending = rest
            short = self._short(m, lane)
            heavy = not short and self.heavy is not None
            if heavy:
                self.heavy.acquire()
            if not short:  # let submitted short requests finish first (bounded wait)
                self._yield_to_filters()
            try:
                with torch.cuda.stream(stream):
                    try:
                        self._run(m, batch, lane)
                    except Exception as e:  # a failed batch fails its requests, not the server
                        for r in batch:
                            r.error = e
                            r.done.set()
            finally:
                if heavy:
                    self.heavy.release()
            if short:
                with self.running_lock:
                    self.short_pending -= len(batch)
                    self.running_lock.notify_all()

    @staticmethod
    def _form_batch(m, pending):
        """Requests in arrival order of the first request's kind (and, for filtering, bucket) while
        their columns fit; the others stay pending."""
        first = pending[0]
        key = (first.kind, m.bucket(first.weight) if first.kind == "filter" else None)
        batch, rest, cols = [], [], 0
        for r in pending:
            rkey = (r.kind, m.bucket(r.weight) if r.kind == "filter" else None)
            if rkey == key and cols + r.columns <= m.max_cols:
                batch.append(r)
                cols += r.columns
            else:
                rest.append(r)
        return batch, rest

    def _buffers(self, m, lane):
        """Per-lane device vectors of max_cols columns, reused by every batch: the first K
        columns of a buffer are contiguous, so a batch of K columns always hands the solver the
        same pointers and replays its captured refinement step."""
        if not hasattr(m, "buf"):
            m.buf = {}
        if lane not in m.buf:
            torch = self.torch
            z = lambda: torch.zeros(m.max_cols, m.N, dtype=torch.float64, device="cuda")
            m.buf[lane] = dict(x=z(), b=z(), t0=z(), t1=z())
        return m.buf[lane]

    def _yield_to_filters(self):
        with self.running_lock:
            self.running_lock.wait_for(lambda: self.short_pending == 0, timeout=self.yield_ms / 1e3)

    def _solve_yielding(self, system, x, b):
        """Mixed solve to self.tol from x = 0 (whatever x holds) that yields to running filtering
        batches between its iterations (flexible CG keeps its search direction across the yields)."""
        its, _ = system.solve(x, b, tol=self.tol, max_iters=self.max_iters, mode="mixed", between=self._yield_to_filters,
                              zero_guess=True)
        return its

    def _run(self, m, batch, lane):
        torch = self.torch
        K = sum(r.columns for r in batch)
        if lane != "lic":
            buf = self._buffers(m, lane)
            x, b, t0, t1 = (buf[k][:K] for k in ("x", "b", "t0", "t1"))
        if batch[0].kind == "filter":
            system = m.systems[m.bucket(batch[0].weight)]
            torch.cat([r.payload for r in batch], dim=0, out=t0)  # f
            w = np.concatenate([[[1.0, r.weight]] * r.columns for r in batch])
            system.apply(api.OPERATOR_MASS, t0, b)
            if any(r.sharpness for r in batch):
                system.apply(api.OPERATOR_STIFFNESS, t0, t1)
                scale = torch.tensor(np.concatenate([[r.sharpness * r.weight] * r.columns for r in batch]), dtype=b.dtype, device=b.device)
                b += scale[:, None] * t1
            x.copy_(t0)
            its, _ = system.solve_weighted(x, b, w, tol=self.tol, max_iters=self.max_iters)
            out = x
        elif batch[0].kind == "lic":
            noise = torch.cat([r.payload for r in batch], dim=0)
            def between():  # V-cycles launch asynchronously: finish this one before yielding
                torch.cuda.current_stream().synchronize()
                self._yield_to_filters()

            _, out = m.lic.convolve(noise, passes=m.lic_rounds, mode=m.lic_mode, between_rounds=between)
            its = m.lic_rounds
        else:
            system = m.heat
            src = torch.as_tensor([s for r in batch for s in r.payload], device="cuda")
            rows = torch.arange(K, device="cuda")
            b.zero_()
            b[rows, src] = 1.0
            its = self._solve_yielding(system, x, b)
            system.geodesic_divergence(x, t0)
            its2 = self._solve_yielding(m.distance, t1, t0)
            its += its2
            t1 -= t1[rows, src][:, None]
            out = t1
        results = out.clone()
        torch.cuda.current_stream().synchronize()
        now = time.perf_counter()
        c = 0
        for r in batch:
            r.result = results[c:c + r.columns]
            r.iterations = its
            r.latency = now - r.submitted
            c += r.columns
            r.done.set()


def latency_percentiles(requests):
    lat = np.array([r.latency for r in requests if r.latency is not None]) * 1e3
    if len(lat) == 0:
        return dict(p50=None, p99=None, n=0)
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)), n=int(len(lat)))

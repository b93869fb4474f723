"""Host-side mirror of the reference's texel-solver call surface, over the libgdtp C ABI.

Reference names are kept where they exist (include/Src/*.inl, TextureFiltering/*.cpp):
InitializeHierarchy, UpdateLinearSystem, VCycle, Relaxation, RelaxationAndResidual,
MultiplyByRestriction, AddProlongation, MultiplyBySystemMatrix_NoReciprocals, solve.
Device vectors are torch tensors on the GPU, float64, shape [ncols, N_level], contiguous.
"""
import ctypes

import numpy as np

from . import _lib

OPERATOR_SYSTEM, OPERATOR_MASS, OPERATOR_STIFFNESS = 0, 1, 2
METRIC_EMBEDDING, METRIC_UNIFORM = 0, 1


def _ptr(t):
    if t is None:
        return None
    if not (t.is_cuda and t.is_contiguous()):
        raise ValueError("device vectors must be contiguous GPU tensors")
    return ctypes.c_void_p(t.data_ptr())


def _vec(t, n, K=None, name="vector"):
    """Device pointer of a [K, n] float64 GPU tensor, checked: the C ABI trusts sizes it is given."""
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.dim() == 2 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous 2-D float64 GPU tensor [columns, texels]")
    if t.shape[1] != n or (K is not None and t.shape[0] != K):
        raise ValueError(f"{name} has shape {tuple(t.shape)}; expected [{K if K is not None else 'columns'}, {n}]")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


_YIELD = ctypes.CFUNCTYPE(None, ctypes.c_void_p)  # void (*between)(void*) of gdtp_solve_yielding


class TexturedMesh:
    """Triangle mesh with per-corner texture coordinates (include/Src/SimpleMesh.h:45-49)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_arrays(cls, vertices, triangles, corner_uv):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        t = np.ascontiguousarray(triangles, dtype=np.int32)
        uv = np.ascontiguousarray(corner_uv, dtype=np.float64)
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.gdtp_mesh_create(v.ctypes.data, v.shape[0], t.ctypes.data, t.shape[0], uv.ctypes.data, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def read_ply(cls, path):
        """PLY with per-face texcoord lists (ReadTexturedMesh, SimpleMesh.h:118-160)."""
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.gdtp_mesh_read_ply(str(path).encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def cube(cls, subdiv, width, height):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.gdtp_mesh_create_cube(subdiv, width, height, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def icosphere(cls, levels, width, height, amplitude=0.05, seed=0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.gdtp_mesh_create_icosphere(levels, width, height, amplitude, seed, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def torus(cls, nu, nv, charts_u, charts_v, width, height, major_radius=1.0, minor_radius=0.35, amplitude=0.05, seed=0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.gdtp_mesh_create_torus(nu, nv, charts_u, charts_v, major_radius, minor_radius, amplitude, seed, width, height,
                                              ctypes.byref(h)))
        return cls(h)

    def arrays(self):
        lib = _lib.load()
        nv, nt = ctypes.c_int(), ctypes.c_int()
        _lib.check(lib.gdtp_mesh_counts(self._h, ctypes.byref(nv), ctypes.byref(nt)))
        v = np.zeros((nv.value, 3)); t = np.zeros((nt.value, 3), np.int32); uv = np.zeros((nt.value, 3, 2))
        _lib.check(lib.gdtp_mesh_read(self._h, v.ctypes.data, t.ctypes.data, uv.ctypes.data))
        return v, t, uv

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.gdtp_mesh_destroy(self._h)
            self._h = None


class HierarchicalSystem:
    """Texel hierarchy plus finest-level mass and stiffness (Hierarchy.h:265-275)."""

    def __init__(self, handle, levels, width, height):
        self._h, self.levels, self.width, self.height = handle, levels, width, height

    def level_info(self, level):
        info = np.zeros(8, np.int32)
        _lib.check(_lib.load().gdtp_hierarchy_level_info(self._h, level, info.ctypes.data))
        keys = ["texels", "deep", "boundary", "charts", "raster_lines", "tasks", "blocks", "segments"]
        return dict(zip(keys, (int(v) for v in info)))

    def array(self, level, name, dtype=np.int32):
        lib = _lib.load()
        fn = lib.gdtp_hierarchy_array_i32 if dtype == np.int32 else lib.gdtp_hierarchy_array_f64
        n = ctypes.c_size_t(0)
        _lib.check(fn(self._h, level, name.encode(), None, ctypes.byref(n)))
        out = np.zeros(n.value, dtype)
        if n.value:
            _lib.check(fn(self._h, level, name.encode(), out.ctypes.data, ctypes.byref(n)))
        return out

    def texel_coordinates(self, level=0):
        """(atlas column, atlas row, chart) of every texel of `level` in global order
        (TextureNodeInfo, HierarchyConstruction.inl:534-555)."""
        ni = self.array(level, "nodeInfo").reshape(-1, 4)
        charts = self.array(level, "charts").reshape(-1, 9)
        ci = charts[ni[:, 0], 2] + ni[:, 1]
        cj = charts[ni[:, 0], 3] + ni[:, 2]
        return ci, cj, ni[:, 0]

    def cell_coordinates(self, level=0):
        """(atlas column, atlas row, chart, type) of every active cell of `level` in global cell
        order (chart by chart, rows then columns; type 1 interior, 0 boundary); a cell spans the
        texels (column, row) .. (column + 1, row + 1)."""
        info = self.array(level, "cellInfo").reshape(-1, 4)
        charts = self.array(level, "charts").reshape(-1, 9)
        return charts[info[:, 0], 2] + info[:, 1], charts[info[:, 0], 3] + info[:, 2], info[:, 0], info[:, 3]

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.gdtp_hierarchy_destroy(self._h)
            self._h = None


class StiffnessModulation:
    """Per-cell stiffness of fixed signal columns, weighted per cell on demand: the sharpening
    brush of TextureFiltering.  Set-up applies every active cell's stiffness to the signal
    (InitializeMassAndStiffness with computeCellBasedStiffness, MassAndStiffness.h:296-395,
    490-549); `apply(mask)` is CellStiffnessToTexelStiffness (IterativeSolvers.inl:1170-1208):
    t[c] = sum over cells k of mask[k] * S_k signal[c], so a mask of ones gives S signal."""

    def __init__(self, hierarchy, signal):
        f = signal.detach().cpu().numpy() if hasattr(signal, "detach") else np.asarray(signal)
        f = np.ascontiguousarray(np.atleast_2d(f), dtype=np.float64)
        self.hierarchy = hierarchy
        h = ctypes.c_void_p()
        _lib.check(_lib.load().gdtp_modulation_create(hierarchy._h, f.ctypes.data, f.shape[0], ctypes.byref(h)))
        self._h = h
        info = np.zeros(4, np.int32)
        _lib.check(_lib.load().gdtp_modulation_info(self._h, info.ctypes.data))
        self.columns, self.texels, self.num_cells, self.boundary_entries = (int(v) for v in info)

    def apply(self, mask, out=None):
        """mask: [num_cells] FP64 GPU tensor; returns out ([columns, texels] FP64)."""
        import torch
        if mask.shape != (self.num_cells,) or mask.dtype != torch.float64:
            raise ValueError(f"mask must be a float64 vector of {self.num_cells} cells")
        if out is None:
            out = torch.empty(self.columns, self.texels, dtype=torch.float64, device=mask.device)
        po = _vec(out, self.texels, self.columns, "out")
        _lib.check(_lib.load().gdtp_modulated_stiffness(self._h, _ptr(mask), po, _stream()))
        return out

    def array(self, name, dtype=np.float64):
        lib = _lib.load()
        fn = lib.gdtp_modulation_array_f64 if dtype == np.float64 else lib.gdtp_modulation_array_i32
        n = ctypes.c_size_t(0)
        _lib.check(fn(self._h, name.encode(), None, ctypes.byref(n)))
        out = np.zeros(n.value, dtype)
        if n.value:
            _lib.check(fn(self._h, name.encode(), out.ctypes.data, ctypes.byref(n)))
        return out

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.gdtp_modulation_destroy(self._h)
            self._h = None


def InitializeHierarchy(mesh, width, height, levels=4, metric=METRIC_EMBEDDING, triangle_metric=None, assembly="host"):
    """InitializeHierarchy + InitializeMetric + InitializeMassAndStiffness
    (HierarchyConstruction.inl:799-849, Metric.inl:448-466, MassAndStiffness.h:490-549).
    triangle_metric ([T, 2, 2], the surface metric in each triangle's edge basis) replaces the
    embedding / uniform metric (Metric.inl:468-486).  assembly="device" integrates the interior
    cells on the GPU (boundary cells and seams on the host)."""
    if assembly not in ("host", "device"):
        raise ValueError("assembly is 'host' or 'device'")
    h = ctypes.c_void_p()
    m = None
    if triangle_metric is not None:
        g = np.asarray(triangle_metric, dtype=np.float64)
        m = np.ascontiguousarray(np.stack([g[:, 0, 0], 0.5 * (g[:, 0, 1] + g[:, 1, 0]), g[:, 1, 1]], axis=1))
    lib = _lib.load()
    if assembly == "device":
        _lib.check(lib.gdtp_hierarchy_create_ex(mesh._h, width, height, levels, metric, None if m is None else m.ctypes.data, 1, ctypes.byref(h)))
    elif m is None:
        _lib.check(lib.gdtp_hierarchy_create(mesh._h, width, height, levels, metric, ctypes.byref(h)))
    else:
        _lib.check(lib.gdtp_hierarchy_create_with_metric(mesh._h, width, height, levels, m.ctypes.data, ctypes.byref(h)))
    return HierarchicalSystem(h, levels, width, height)


def embedding_metric(vertices, triangles, normalize_area=True):
    """Per-triangle Gram matrix of the edges (p1 - p0, p2 - p0), scaled to unit total area
    (InitializeEmbeddingMetric, Metric.inl:248-276).  Returns [T, 2, 2]."""
    P = np.asarray(vertices, np.float64)[np.asarray(triangles)]
    d = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]], axis=1)
    g = np.einsum("tkc,tlc->tkl", d, d)
    if normalize_area:
        g /= np.sum(np.sqrt(np.linalg.det(g)) / 2)
    return g


def vector_field_metric(g, vf, normalize_area=True, along=0.0, across=1e-1, floor=1e-5):
    """Anisotropic metric of a tangent field (InitializeVectorFieldMetric, Metric.inl:12-67):
    with v the field (per triangle, edge-basis coordinates) and w its g-orthogonal complement,
    both g-normalised and B = [v w], the metric is  g B diag(along, across) B^-1 + floor * g,
    zero-field triangles get floor * g; scaled to unit total area."""
    g = np.asarray(g, np.float64)
    v = np.asarray(vf, np.float64)
    out = floor * g.copy()
    gv = np.einsum("tkl,tl->tk", g, v)
    nrm = np.einsum("tk,tk->t", v, gv)
    ok = nrm > 0
    w = np.stack([-gv[:, 1], gv[:, 0]], axis=1)
    v0 = v / np.sqrt(np.where(ok, nrm, 1.0))[:, None]
    nw = np.einsum("tk,tkl,tl->t", w, g, w)
    v1 = w / np.sqrt(np.where(nw > 0, nw, 1.0))[:, None]
    B = np.stack([v0, v1], axis=2)  # columns v0, v1
    D = np.diag([along, across])
    u = np.einsum("tkl,tlm,mn,tnp->tkp", g, B, D, np.linalg.inv(np.where(ok[:, None, None], B, np.eye(2))))
    out[ok] += u[ok]
    if normalize_area:
        out /= np.sum(np.sqrt(np.linalg.det(out)) / 2)
    return out


class MultigridSystem:
    """A = screen*M + stiffness*S over the hierarchy, with device buffers for max_cols columns."""

    def __init__(self, handle, hierarchy, max_cols):
        self._h, self.hierarchy, self.max_cols = handle, hierarchy, max_cols
        info = [hierarchy.level_info(l) for l in range(hierarchy.levels)]
        self._texels, self._boundary = [i["texels"] for i in info], [i["boundary"] for i in info]

    def _level(self, level):
        if not 0 <= level < len(self._texels):
            raise ValueError(f"level {level} outside [0, {len(self._texels) - 1}]")
        return self._texels[level]

    def update(self, stiffness_weight, screen_weight=1.0):
        _lib.check(_lib.load().gdtp_system_update(self._h, screen_weight, stiffness_weight))

    def array(self, level, name, dtype=np.float64):
        lib = _lib.load()
        fn = lib.gdtp_system_array_i32 if dtype == np.int32 else lib.gdtp_system_array_f64
        n = ctypes.c_size_t(0)
        _lib.check(fn(self._h, level, name.encode(), None, ctypes.byref(n)))
        out = np.zeros(n.value, dtype)
        if n.value:
            _lib.check(fn(self._h, level, name.encode(), out.ctypes.data, ctypes.byref(n)))
        return out

    # -- device operations
    def _pair(self, x, b, n, *more):
        K = x.shape[0] if hasattr(x, "shape") else None
        ptrs = [_vec(x, n, None, "x"), _vec(b, n, K, "b")] + [None if m is None else _vec(m, n, K, "output") for m in more]
        return ptrs, K

    def vcycle(self, x, b):
        (px, pb), K = self._pair(x, b, self._texels[0])
        _lib.check(_lib.load().gdtp_vcycle(self._h, px, pb, K, _stream()))

    def relax(self, level, x, b, residual=None):
        (px, pb, pr), K = self._pair(x, b, self._level(level), residual)
        _lib.check(_lib.load().gdtp_relax(self._h, level, px, pb, pr, K, _stream()))

    def band_sweep(self, level, x, b, residual=None):
        (px, pb, pr), K = self._pair(x, b, self._level(level), residual)
        _lib.check(_lib.load().gdtp_band_sweep(self._h, level, px, pb, pr, K, _stream()))

    def restrict(self, level, fine, coarse):
        n = self._level(level)
        pf = _vec(fine, n, None, "fine")
        pc = _vec(coarse, self._level(level + 1), fine.shape[0], "coarse")
        _lib.check(_lib.load().gdtp_restrict(self._h, level, pf, pc, fine.shape[0], _stream()))

    def prolong_add(self, level, coarse, fine):
        n = self._level(level)
        pf = _vec(fine, n, None, "fine")
        pc = _vec(coarse, self._level(level + 1), fine.shape[0], "coarse")
        _lib.check(_lib.load().gdtp_prolong_add(self._h, level, pc, pf, fine.shape[0], _stream()))

    def boundary_solve(self, level, rhs, x):
        self._level(level)
        (pr, px), K = self._pair(rhs, x, self._boundary[level])
        _lib.check(_lib.load().gdtp_boundary_solve(self._h, level, pr, px, K, _stream()))

    def coarse_solve(self, rhs, x):
        (pr, px), K = self._pair(rhs, x, self._texels[-1])
        _lib.check(_lib.load().gdtp_coarse_solve(self._h, pr, px, K, _stream()))

    def apply(self, which, x, y):
        (px, py), K = self._pair(x, y, self._texels[0])
        _lib.check(_lib.load().gdtp_apply(self._h, which, px, py, K, _stream()))

    def residual(self, x, b, r):
        (px, pb, pr), K = self._pair(x, b, self._texels[0], r)
        _lib.check(_lib.load().gdtp_residual(self._h, px, pb, pr, K, _stream()))

    def residual_norms(self, x, b):
        """(rel_res, residual norms, rhs norms) per column: ||b - A x|| / ||b|| in FP64 on level 0
        (ComputeSystemResdiual, IterativeSolvers.inl:953-1007), without writing the residual."""
        (px, pb), K = self._pair(x, b, self._texels[0])
        rel, norms = np.zeros(K), np.zeros(2 * K)
        _lib.check(_lib.load().gdtp_residual_norms(self._h, px, pb, K, rel.ctypes.data, norms.ctypes.data, _stream()))
        return rel, norms[0::2].copy(), norms[1::2].copy()

    def direct_solve(self, x, b):
        """Exact FP64 solve of A x = b on level 0 (the apps' --useDirectSolver path): LDL^T of the
        full matrix up to 2^20 texels, else FP64 V-cycles to the FP64 residual floor.  Returns
        (rel_res per column, method: 0 factorisation / 1 V-cycles)."""
        (px, pb), K = self._pair(x, b, self._texels[0])
        rel = np.zeros(K)
        m = ctypes.c_int(-1)
        _lib.check(_lib.load().gdtp_direct_solve(self._h, px, pb, K, rel.ctypes.data, ctypes.byref(m), _stream()))
        return rel, m.value

    def geodesic_divergence(self, potential, rhs):
        """rhs = integral < -grad u / |grad u|_g , grad phi_t >_g per column of `potential`."""
        (pu, pr), K = self._pair(potential, rhs, self._texels[0])
        _lib.check(_lib.load().gdtp_geodesic_divergence(self._h, pu, pr, K, _stream()))

    def solve(self, x, b, tol=1e-6, max_iters=50, mode="mixed", between=None, zero_guess=False):
        """Solve A x = b to relative residual `tol` per column; returns (iterations, rel_res).
        between: optional callable run after every unconverged iteration (the device is idle
        then; the solve continues afterwards without restarting).  zero_guess: start from x = 0
        whatever x holds (the solve writes the zeros; the mixed path's first residual is b)."""
        (px, pb), K = self._pair(x, b, self._texels[0])
        rel = np.zeros(K)
        it = ctypes.c_int(0)
        m = (1 if mode == "mixed" else 0) | (0x100 if zero_guess else 0)
        if between is None:
            _lib.check(_lib.load().gdtp_solve(self._h, px, pb, K, tol, max_iters, m, rel.ctypes.data,
                                              ctypes.byref(it), _stream()))
        else:
            cb = _YIELD(lambda _user: between())
            _lib.check(_lib.load().gdtp_solve_yielding(self._h, px, pb, K, tol, max_iters, m, rel.ctypes.data,
                                                       ctypes.byref(it), _stream(), ctypes.cast(cb, ctypes.c_void_p), None))
        return it.value, rel

    def solve_weighted(self, x, b, weights, tol=1e-6, max_iters=60):
        """Column c solves (weights[c][0] M + weights[c][1] S) x = b (mixed precision; this system's
        V-cycle preconditions every column).  Returns (iterations, rel_res)."""
        (px, pb), K = self._pair(x, b, self._texels[0])
        w = np.ascontiguousarray(np.asarray(weights, np.float64).reshape(K, 2))
        rel = np.zeros(K)
        it = ctypes.c_int(0)
        _lib.check(_lib.load().gdtp_solve_weighted(self._h, px, pb, K, w.ctypes.data, tol, max_iters, rel.ctypes.data,
                                                   ctypes.byref(it), _stream()))
        return it.value, rel

    def configure_mixed(self, nu_pre=1, nu_post=2, boundary_sweeps=1, omega=1.0, use_graph=True, krylov=False, cycles=0):
        """Smoothing schedule of the FP32 V-cycle (nu_pre steps of boundary Schwarz then the four
        deep colours; nu_post sweeps of the deep colours in reverse order, then one boundary Schwarz
        step; boundary_sweeps Schwarz sweeps per Schwarz step), graph capture, the outer iteration
        (flexible CG with krylov=True, else plain iterative refinement) and the V-cycles per outer
        iteration (0: automatic, two unless the operator is nearly singular in FP32)."""
        lib = _lib.load()
        _lib.check(lib.gdtp_mixed_acceleration(self._h, 1 if krylov else 0))
        _lib.check(lib.gdtp_mixed_cycles(self._h, int(cycles)))
        _lib.check(lib.gdtp_mixed_configure(self._h, nu_pre, nu_post, boundary_sweeps, omega, 1 if use_graph else 0))

    # (pre-smoothing, post-smoothing, Schwarz sweeps, flexible CG, V-cycles per residual)
    TUNE_CANDIDATES = [(1, 2, 1, False, 2), (1, 2, 1, False, 1), (1, 1, 1, False, 2), (2, 2, 1, False, 2), (2, 1, 1, False, 2),
                       (1, 2, 2, False, 2), (1, 2, 1, True, 1), (1, 1, 1, True, 1), (2, 2, 1, True, 1)]

    def autotune_mixed(self, columns=None, tol=1e-6, candidates=None, repeats=2, max_iters=60, seed=0):
        """Time mixed-solver schedules on this system and keep the fastest.  Each candidate
        (pre, post, Schwarz sweeps, flexible CG, V-cycles per residual) solves seeded right-hand
        sides b = M f (f uniform [0,1), `columns` columns, default max_cols) from x = 0 to `tol`,
        `repeats` times after one untimed run (graph capture); candidates that do not converge are
        skipped.  Returns the measurements, fastest first; the system keeps the fastest schedule."""
        import time
        import torch
        K = columns or self.max_cols
        N = self._texels[0]
        f = torch.from_numpy(np.random.default_rng(seed).random((K, N))).cuda()
        b = torch.empty_like(f)
        self.apply(OPERATOR_MASS, f, b)
        x = torch.zeros_like(b)
        results = []
        for cand in candidates or self.TUNE_CANDIDATES:
            pre, post, sweeps, krylov, cycles = cand
            self.configure_mixed(pre, post, sweeps, 1.0, True, krylov, cycles)
            try:
                x.zero_()
                self.solve(x, b, tol=tol, max_iters=max_iters)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(repeats):
                    x.zero_()
                    its, rel = self.solve(x, b, tol=tol, max_iters=max_iters)
                torch.cuda.synchronize()
                results.append(dict(schedule=cand, ms=(time.perf_counter() - t0) / repeats * 1e3, iterations=its))
            except _lib.GdtpError as e:
                if e.code != 3:
                    raise
        if not results:
            raise RuntimeError("no candidate schedule reached the tolerance")
        results.sort(key=lambda r: r["ms"])
        pre, post, sweeps, krylov, cycles = results[0]["schedule"]
        self.configure_mixed(pre, post, sweeps, 1.0, True, krylov, cycles)
        return results

    def mixed_deep_sweep(self, level, ncols):
        _lib.check(_lib.load().gdtp_mixed_deep_sweep(self._h, level, ncols, _stream()))

    @property
    def TILE_SHAPES(self):
        """Deep-sweep tile shapes: index -> label (fused shapes 0..6, -1 one launch per colour)."""
        return tile_shapes()

    TILE_SHAPE_AUTO = -3
    TILE_SHAPE_PAIRS = -4

    def tile_shape(self, shape=None, level=None):
        """Set (0..6 fused, -1 one launch per colour, TILE_SHAPE_PAIRS one launch per colour pair,
        TILE_SHAPE_AUTO per level by size, the default) and/or return the deep-sweep tile shape
        setting of every level, or the shape one level uses."""
        cur = ctypes.c_int(0)
        arg = -2 if shape is None else int(shape)
        if level is None:
            _lib.check(_lib.load().gdtp_mixed_tile_shape(self._h, arg, ctypes.byref(cur)))
        else:
            _lib.check(_lib.load().gdtp_mixed_level_tile_shape(self._h, int(level), arg, ctypes.byref(cur)))
        return cur.value

    def fuse_transfers(self, prolongation=True):
        """Fuse the prolongation into the first post-smoothing sweep of fused levels, or keep it a
        separate pass (the default); the iterate is bitwise the same."""
        _lib.check(_lib.load().gdtp_mixed_fuse_transfers(self._h, 1 if prolongation else 0))

    def autotune_tiles(self, columns=None, sweeps=20, shapes=None):
        """Time `sweeps` deep sweeps of every level swept by tiles for each tile shape and keep
        each level's fastest (every shape gives the same iterate).  Returns
        {level: {shape: microseconds per sweep}}, fastest first."""
        import torch
        K = columns or self.max_cols
        shapes = list(tile_shapes()) if shapes is None else shapes
        st = torch.cuda.current_stream()
        out = {}
        for level in range(self.hierarchy.levels - 1):
            times = {}
            try:
                for sh in shapes:
                    self.tile_shape(sh, level)
                    self.mixed_deep_sweep(level, K)
                    torch.cuda.synchronize()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(st)
                    for _ in range(sweeps):
                        self.mixed_deep_sweep(level, K)
                    e1.record(st)
                    torch.cuda.synchronize()
                    times[sh] = e0.elapsed_time(e1) * 1e3 / sweeps
            except _lib.GdtpError:  # a level swept one chart per workgroup has no tiles
                continue
            best = min(times, key=times.get)
            self.tile_shape(best, level)
            out[level] = dict(sorted(times.items(), key=lambda kv: kv[1]))
        return out

    def memory(self):
        r, m = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(_lib.load().gdtp_system_memory(self._h, ctypes.byref(r), ctypes.byref(m)))
        return r.value, m.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.gdtp_system_destroy(self._h)
            self._h = None


_PREFETCH = {0: "fetch per colour", 1: "one colour ahead", 2: "all colours ahead"}


def tile_shapes():
    """The fused deep-sweep tile shapes the library offers, index -> label ("64x4 tile, 256
    threads, one colour ahead, XCD-contiguous"), -1 -> "per-colour" (one launch per colour) and
    -4 -> "colour pairs" (one launch per colour pair, a workgroup per row)."""
    out, lib = {}, _lib.load()
    w, h, t, v = (ctypes.c_int() for _ in range(4))
    k = 0
    while lib.gdtp_mixed_tile_shape_info(k, ctypes.byref(w), ctypes.byref(h), ctypes.byref(t), ctypes.byref(v)) == 0:
        out[k] = f"{w.value}x{h.value} tile, {t.value} threads, {_PREFETCH[v.value & 3]}" + (", XCD-contiguous" if v.value & 4 else "")
        k += 1
    out[-1] = "per-colour"
    out[MultigridSystem.TILE_SHAPE_PAIRS] = "colour pairs"
    return out


def UpdateLinearSystem(hierarchy, stiffness_weight, screen_weight=1.0, max_cols=3):
    """UpdateLinearSystem (UpdateCoefficients.inl:463-513); max_cols=0 builds the host operator only."""
    h = ctypes.c_void_p()
    _lib.check(_lib.load().gdtp_system_create(hierarchy._h, screen_weight, stiffness_weight, max_cols, ctypes.byref(h)))
    return MultigridSystem(h, hierarchy, max_cols)


def VCycle(system, x, b):
    """VCycle (IterativeSolvers.inl:1211-1333) on level-0 device vectors."""
    system.vcycle(x, b)


def RelaxationAndResidual(system, level, x, b, residual):
    system.relax(level, x, b, residual)


def Relaxation(system, level, x, b):
    system.relax(level, x, b, None)


def MultiplyByRestriction(system, level, fine, coarse):
    system.restrict(level, fine, coarse)


def AddProlongation(system, level, coarse, fine):
    system.prolong_add(level, coarse, fine)


def MultiplyBySystemMatrix_NoReciprocals(system, which, x, y):
    system.apply(which, x, y)


# ---------------------------------------------------------------------------- application recipes
def _solve_mode(mode):
    if mode not in ("mixed", "fp64"):
        raise ValueError("mode must be 'mixed' or 'fp64'")
    return mode


class TextureFilter:
    """Screened-Poisson texture filtering (TextureFiltering.cpp:581-718, 562-579):
        (M + a S) x = M x0 + s a S x0,   x starts at x0,
    with a the stiffness (smoothing) weight and s the sharpness weight (s = 0: pure smoothing,
    s > 1: sharpening).  Columns are channels.  A per-cell sharpness (the app's brush) replaces
    s S x0 by sum_k s_k S_k x0 after `set_signal(x0)` (TextureFiltering.cpp:569-570)."""

    def __init__(self, hierarchy, stiffness_weight=1e-3, sharpness_weight=1.0, max_cols=3):
        self.hierarchy, self.stiffness, self.sharpness = hierarchy, stiffness_weight, sharpness_weight
        self.system = UpdateLinearSystem(hierarchy, stiffness_weight, 1.0, max_cols=max_cols)
        self.modulation = None

    def set_signal(self, x0):
        """Cell-based stiffness of the input texels, for per-cell sharpness masks."""
        self.modulation = StiffnessModulation(self.hierarchy, x0)

    def set_stiffness(self, stiffness_weight):
        """UpdateLinearSystem on a new weight (the app's slider, TextureFiltering.cpp:515-545)."""
        self.stiffness = stiffness_weight
        self.system.update(stiffness_weight, 1.0)

    def rhs(self, x0, mask=None):
        """M x0 + s a S x0, or with a per-cell mask (after set_signal(x0)) M x0 + a t(mask)."""
        import torch
        m, s = torch.empty_like(x0), torch.empty_like(x0)
        self.system.apply(OPERATOR_MASS, x0, m)
        if mask is not None:
            if self.modulation is None:
                raise ValueError("set_signal(x0) builds the cell-based stiffness a mask weights")
            return m + self.stiffness * self.modulation.apply(mask)
        if self.sharpness == 0:
            return m
        self.system.apply(OPERATOR_STIFFNESS, x0, s)
        return m + (self.sharpness * self.stiffness) * s

    def filter(self, x0, tol=1e-6, mode="mixed", max_iters=60, mask=None):
        """Filtered texels for input texels x0 ([K, N] FP64 on the GPU); returns (x, iterations)."""
        b = self.rhs(x0, mask)
        x = x0.clone()
        its, _ = self.system.solve(x, b, tol=tol, max_iters=max_iters, mode=_solve_mode(mode))
        return x, its


class Geodesics:
    """Heat-method geodesic distances from source texels (Geodesics.cpp:508-550, 600-716):
        (M + t S) u = delta_source,   rhs = div(-grad u / |grad u|),   (1e-6 M + S) phi = rhs,
    and phi -= phi[source].  One column per source.  t defaults to the reference's 1e-3
    (Geodesics.cpp:19), in the area-normalised metric."""

    def __init__(self, hierarchy, heat_time=1e-3, max_sources=8):
        self.hierarchy, self.t = hierarchy, heat_time
        self.heat = UpdateLinearSystem(hierarchy, heat_time, 1.0, max_cols=max_sources)
        self.distance = UpdateLinearSystem(hierarchy, 1.0, 1e-6, max_cols=max_sources)
        self.max_sources = max_sources
        if max_sources > 0:  # flexible CG: fewer passes than refinement on both (measured at 2048², 64 sources)
            for s in (self.heat, self.distance):
                s.configure_mixed(krylov=True)

    def distances(self, sources, tol=1e-6, mode="mixed", max_iters=80):
        """Distances ([len(sources), N] FP64 on the GPU) from each source texel (global index)."""
        import torch
        K = len(sources)
        if not 1 <= K <= self.max_sources:
            raise ValueError(f"between 1 and {self.max_sources} sources")
        N = self.hierarchy.level_info(0)["texels"]
        src = torch.as_tensor(np.asarray(sources, dtype=np.int64), device="cuda")
        if int(src.min()) < 0 or int(src.max()) >= N:
            raise ValueError("source texel out of range")
        delta = torch.zeros(K, N, dtype=torch.float64, device="cuda")
        delta[torch.arange(K, device="cuda"), src] = 1.0
        u = torch.empty_like(delta)  # both solves start from zero (zero_guess writes the zeros)
        self.heat.solve(u, delta, tol=tol, max_iters=max_iters, mode=_solve_mode(mode), zero_guess=True)
        div = torch.empty_like(u)
        self.heat.geodesic_divergence(u, div)
        phi = torch.empty_like(u)
        self.distance.solve(phi, div, tol=tol, max_iters=max_iters, mode=mode, zero_guess=True)
        phi -= phi[torch.arange(K, device="cuda"), src][:, None]
        return phi


def surface_gradient_field(vertices, triangles, f):
    """Per-triangle gradient of the per-vertex scalar f, as contravariant components in the edge
    basis (p1 - p0, p2 - p0): g^-1 (f1 - f0, f2 - f0) with g the edge Gram matrix."""
    t = np.asarray(triangles)
    g = embedding_metric(vertices, t, normalize_area=False)
    f = np.asarray(f, np.float64)[t]
    return np.linalg.solve(g, np.stack([f[:, 1] - f[:, 0], f[:, 2] - f[:, 0]], axis=1)[..., None])[..., 0]


class LineIntegralConvolution:
    """Line integral convolution as anisotropic diffusion (LineIntegralConvolution.cpp:503-527,
    560-710), on two hierarchies over the same atlas:
        (M_a + l S_a) x = M_a noise        with the vector-field metric (vector_field_metric),
        (M + a S) y = M x + s a S x        with the embedding metric (the sharpening pass),
    defaults as the reference's: l = 1e-4 (licStiffness), a = 1e-4 (stiffnessWeight),
    s = 100 (sharpnessWeight).  vector_field: per mesh triangle, edge-basis components."""

    def __init__(self, mesh, width, height, levels, vector_field, convolution_weight=1e-4, stiffness_weight=1e-4,
                 sharpening_weight=100.0, max_cols=3, assembly="host"):
        v, t, _ = mesh.arrays()
        metric = vector_field_metric(embedding_metric(v, t), vector_field)
        self.anisotropic = InitializeHierarchy(mesh, width, height, levels, triangle_metric=metric, assembly=assembly)
        self.embedding = InitializeHierarchy(mesh, width, height, levels, assembly=assembly)
        self.convolution_weight, self.stiffness, self.sharpening = convolution_weight, stiffness_weight, sharpening_weight
        self.lic = UpdateLinearSystem(self.anisotropic, convolution_weight, 1.0, max_cols=max_cols)
        self.sharpen = UpdateLinearSystem(self.embedding, stiffness_weight, 1.0, max_cols=max_cols)

    def convolve_lengths(self, noise, lengths, rounds=4):
        """Diffusion with one length per column (LineIntegralConvolution.cpp:503-527, 611-619 with
        the licStiffness slider, batched): column c solves (M_a + lengths[c] S_a) x_c = M_a noise_c
        by `rounds` iterations of the weighted solve (flexible CG) from x = 0, as the reference runs
        a fixed number of rounds.  Each column is preconditioned by the V-cycle of the anisotropic
        system at the power of ten nearest its length (built on first use, kept): one V-cycle serves
        lengths within a factor of about three of its own, and a column's iterate does not depend on
        the other columns of the batch.  Returns (x [K, N], relative residual per column against
        its own operator)."""
        import torch
        lengths = np.asarray(lengths, np.float64)
        b = torch.empty_like(noise)
        self.lic.apply(OPERATOR_MASS, noise, b)
        x = torch.zeros_like(b)
        decade = np.round(np.log10(lengths)).astype(int)
        centres = 10.0 ** np.unique(decade)
        group = np.searchsorted(np.unique(decade), decade)
        if not hasattr(self, "_length_systems"):
            self._length_systems = {}
        for gi, lam in enumerate(centres):
            cols = np.nonzero(group == gi)[0]
            if len(cols) == 0:
                continue
            key = float(lam)
            if key not in self._length_systems:
                self._length_systems[key] = UpdateLinearSystem(self.anisotropic, key, 1.0, max_cols=max(len(lengths), self.lic.max_cols))
            system = self._length_systems[key]
            idx = torch.as_tensor(cols, device=b.device)
            xg, bg = torch.zeros((len(cols), b.shape[1]), dtype=b.dtype, device=b.device), b[idx].contiguous()
            try:  # a fixed number of rounds: "not converged" is the expected outcome
                system.solve_weighted(xg, bg, [[1.0, float(l)] for l in lengths[cols]], tol=1e-300, max_iters=rounds)
            except _lib.GdtpError as e:
                if e.code != 3:
                    raise
            x[idx] = xg
        return x, self._weighted_residuals(x, b, lengths)

    def _weighted_residuals(self, x, b, lengths):
        import torch
        m, s = torch.empty_like(x), torch.empty_like(x)
        self.lic.apply(OPERATOR_MASS, x, m)
        self.lic.apply(OPERATOR_STIFFNESS, x, s)
        lam = torch.tensor(np.asarray(lengths, np.float64), device=x.device)[:, None]
        return ((b - m - lam * s).norm(dim=1) / b.norm(dim=1)).cpu().numpy()

    def convolve(self, noise, passes=4, tol=None, mode="fp64", max_iters=200, between_rounds=None):
        """(x, y): the anisotropically diffused noise and its sharpened version ([K, N] FP64).

        Default (tol=None): the reference's schedule (LineIntegralConvolution.cpp:503-527, 707),
        `passes` rounds of one V-cycle on the LIC system, the sharpening rhs from the current x,
        one V-cycle on the sharpening system, from x = y = 0.  Strongly anisotropic diffusion
        converges slowly under geometric multigrid, so the app runs a fixed number of rounds.
        mode="fp64" runs the reference-schedule V-cycle; mode="mixed" one mixed-precision
        refinement step instead (x += V(b - A x) with the FP32 V-cycle: short kernels, but a
        different smoother, so after a few unconverged rounds its iterate is not the reference's).
        `between_rounds` is called after each half-round (each V-cycle).  With tol: both systems solved to that
        relative residual (flexible CG around the mixed V-cycle)."""
        import torch
        K = noise.shape[0]
        buf = self._buffers(K, noise)
        x, y, b, r = buf["x"], buf["y"], buf["b"], buf["r"]
        self.lic.apply(OPERATOR_MASS, noise, b)
        x.zero_()
        y.zero_()

        def sharpen_rhs():
            m, s = buf["m"], buf["s"]
            self.sharpen.apply(OPERATOR_MASS, x, m)
            self.sharpen.apply(OPERATOR_STIFFNESS, x, s)
            torch.add(m, s, alpha=self.sharpening * self.stiffness, out=r)
            return r

        def one_round(system, u, rhs):
            if mode == "fp64":
                system.vcycle(u, rhs)
                return
            try:  # one refinement step; "not converged" is the expected outcome
                system.solve(u, rhs, tol=1e-300, max_iters=1, mode="mixed")
            except _lib.GdtpError as e:
                if e.code != 3:
                    raise

        if tol is None:
            if mode != "fp64" and getattr(self, "_round_schedule", None) != "mixed":
                for system in (self.lic, self.sharpen):
                    system.configure_mixed(cycles=1)  # one V-cycle per round, as the reference's
                self._round_schedule = "mixed"
            for _ in range(passes):
                one_round(self.lic, x, b)
                if between_rounds is not None:  # e.g. a serving loop letting short requests run
                    between_rounds()
                one_round(self.sharpen, y, sharpen_rhs())
                if between_rounds is not None:
                    between_rounds()
            return x.clone(), y.clone()
        for system in (self.lic, self.sharpen):
            system.configure_mixed(krylov=True)
        self._round_schedule = None
        self.lic.solve(x, b, tol=tol, max_iters=max_iters, mode=_solve_mode(mode))
        y.copy_(x)
        self.sharpen.solve(y, sharpen_rhs(), tol=tol, max_iters=max_iters, mode=mode)
        return x.clone(), y.clone()

    def _buffers(self, K, like):
        """Device vectors per column count, reused across calls (fixed pointers let the mixed
        solver replay its captured steps)."""
        import torch
        if not hasattr(self, "_buf"):
            self._buf = {}
        if K not in self._buf:
            self._buf[K] = {k: torch.zeros(K, like.shape[1], dtype=torch.float64, device=like.device) for k in ("x", "y", "b", "r", "m", "s")}
        return self._buf[K]


# ---------------------------------------------------------------------------- file formats and the apps' preprocessing
class Padding:
    """Texels added on each side of the atlas so that every texture coordinate lies inside the
    texel-centre box and both sides are multiples of 8 (Padding.h:4-45)."""

    def __init__(self, left=0, bottom=0, right=0, top=0):
        self.left, self.bottom, self.right, self.top = left, bottom, right, top

    @property
    def non_trivial(self):
        return bool(self.left or self.bottom or self.right or self.top)


def compute_padding(width, height, uv):
    """ComputePadding (Padding.h:16-45); uv: [..., 2] texture coordinates."""
    import math
    uv = np.asarray(uv, np.float64).reshape(-1, 2)
    lo, hi = uv.min(axis=0), uv.max(axis=0)
    pmin = (0.5 / width, 0.5 / height)
    pmax = ((width - 0.5) / width, (height - 0.5) / height)
    left = math.ceil((pmin[0] - lo[0]) * width) if lo[0] < pmin[0] else 0
    bottom = math.ceil((pmin[1] - lo[1]) * height) if lo[1] < pmin[1] else 0
    right = math.ceil((hi[0] - pmax[0]) * width) if hi[0] > pmax[0] else 0
    top = math.ceil((hi[1] - pmax[1]) * height) if hi[1] > pmax[1] else 0
    nw, nh = width + left + right, height + bottom + top
    left += 8 * ((nw - 1) // 8 + 1) - nw
    bottom += 8 * ((nh - 1) // 8 + 1) - nh
    return Padding(left, bottom, right, top)


def pad_texture_coordinates(padding, width, height, uv):
    """PaddTextureCoordinates (Padding.h:70-77): returns the new coordinates."""
    uv = np.array(uv, np.float64)
    nw, nh = width + padding.left + padding.right, height + padding.bottom + padding.top
    uv[..., 0] = (uv[..., 0] * width + padding.left) / nw
    uv[..., 1] = (uv[..., 1] * height + padding.bottom) / nh
    return uv


def pad_image(padding, image):
    """PaddImage (Padding.h:46-57): image [H, W, C] (row j = texture row j) with clamped borders."""
    h, w = image.shape[:2]
    jj = np.clip(np.arange(h + padding.bottom + padding.top) - padding.bottom, 0, h - 1)
    ii = np.clip(np.arange(w + padding.left + padding.right) - padding.left, 0, w - 1)
    return image[jj][:, ii]


def unpad_image(padding, image):
    """UnpaddImage (Padding.h:58-66)."""
    h, w = image.shape[:2]
    return image[padding.bottom:h - padding.top, padding.left:w - padding.right]


def load_textured_mesh(path, width, height, flip_v=True, jitter=False, seed=0):
    """The apps' mesh preparation (TextureFiltering.cpp:807-837): read the PLY, centre the
    vertices and scale them to unit radius, flip v, optionally jitter texture coordinates by
    1e-3 texel (seeded here; the reference draws from rand(), AtlasMesh.inl:36-55), and pad the
    atlas.  Returns (mesh, width, height, padding) with the padded resolution."""
    m = TexturedMesh.read_ply(path)
    v, t, uv = m.arrays()
    v = v - v.mean(axis=0)
    v /= np.linalg.norm(v, axis=1).max()
    if flip_v:
        uv[..., 1] = 1.0 - uv[..., 1]
    if jitter:
        uv = uv + (np.random.default_rng(seed).random(uv.shape) * 2 - 1) * (1e-3 / max(width, height))
    pad = compute_padding(width, height, uv)
    if pad.non_trivial:
        uv = pad_texture_coordinates(pad, width, height, uv)
        width, height = width + pad.left + pad.right, height + pad.bottom + pad.top
    return TexturedMesh.from_arrays(v, t, uv), width, height, pad


def read_vector(path, dtype):
    """ReadVector (VectorIO.h:8-18): int32 count, then count records of `dtype` (e.g. the .vf
    files: np.dtype((np.float64, 2)))."""
    with open(path, "rb") as fh:
        n = int(np.frombuffer(fh.read(4), np.int32)[0])
        return np.frombuffer(fh.read(), dtype, count=n).copy()


def write_vector(path, vec):
    """WriteVector (VectorIO.h:19-27)."""
    vec = np.ascontiguousarray(vec)
    with open(path, "wb") as fh:
        fh.write(np.int32(len(vec)).tobytes())
        fh.write(vec.tobytes())


def read_binary_image(path, channels=3, dtype=np.float32):
    """ReadBinaryImage (VectorIO.h:39-50), e.g. .normap: int32 width, height, then width*height
    records of `channels` values; returns [height, width, channels]."""
    with open(path, "rb") as fh:
        w, h = (int(x) for x in np.frombuffer(fh.read(8), np.int32))
        return np.frombuffer(fh.read(), dtype, count=w * h * channels).reshape(h, w, channels).copy()


def write_binary_image(path, image):
    """WriteBinaryImage (VectorIO.h:28-38)."""
    image = np.ascontiguousarray(image)
    with open(path, "wb") as fh:
        fh.write(np.array([image.shape[1], image.shape[0]], np.int32).tobytes())
        fh.write(image.tobytes())

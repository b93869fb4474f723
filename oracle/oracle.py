"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

CPU restatement of the reference V-cycle (include/Src/IterativeSolvers.inl:1211-1333): the
arithmetic kernels are the plain-C restatement in oracle.c (built into oracle/_build/liboracle.so
by oracle/Makefile); exact boundary / coarse solves use SciPy's sparse LU, i.e. the exact solution
that the reference's SimplicialLDLT (include/Src/Solver.h:145-259) also computes, up to rounding.
Inputs are plain arrays (hierarchy index structures and level coefficients).

Parity pins (tests/test_oracle.py): residuals equal b - A x of an independently built sparse A;
restriction equals P^T of the prolongation matrix; prolongation reproduces bilinear fields;
one relaxation matches a dense-matrix restatement of the band schedule on small atlases; the
V-cycle converges to the sparse direct solution.
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")


class _Tasks(ctypes.Structure):
    _fields_ = [("n_tasks", ctypes.c_int), ("task_ptr", ctypes.c_void_p), ("block_begin", ctypes.c_void_p),
                ("block_deep_end", ctypes.c_void_p), ("block_pad_end", ctypes.c_void_p), ("line_ptr", ctypes.c_void_p),
                ("segments", ctypes.c_void_p)]


class _Csr(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int), ("rowptr", ctypes.c_void_p), ("col", ctypes.c_void_p), ("val", ctypes.c_void_p)]


_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "oracle"], check=True, stdout=subprocess.DEVNULL)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Csr:
    def __init__(self, rowptr, col, val, ncols):
        self.rowptr, self.col, self.val, self.ncols = _i(rowptr), _i(col), _d(val), ncols

    @property
    def rows(self):
        return len(self.rowptr) - 1

    def c(self):
        return _Csr(self.rows, self.rowptr.ctypes.data, self.col.ctypes.data, self.val.ctypes.data)

    def scipy(self):
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.rows, self.ncols))


class LevelData:
    """Index structures and coefficients of one level (all plain arrays)."""

    def __init__(self, N, D, B, bgi, raster, tasks, coef, coef_raw, bd, bb, restr_lines=None, prol_lines=None, brestr=None):
        self.N, self.D, self.B = N, D, B
        self.bgi = _i(bgi)
        self.raster = _i(raster)
        self.tasks = {k: _i(v) for k, v in tasks.items()}
        self.coef, self.coef_raw = _d(coef), _d(coef_raw)
        self.bd, self.bb = bd, bb
        self.restr_lines = _i(restr_lines) if restr_lines is not None else None
        self.prol_lines = _i(prol_lines) if prol_lines is not None else None
        self.brestr = brestr

    def ctasks(self):
        t = self.tasks
        return _Tasks(len(t["taskPtr"]) - 1, t["taskPtr"].ctypes.data, t["blockBegin"].ctypes.data, t["blockDeepEnd"].ctypes.data,
                      t["blockPadEnd"].ctypes.data, t["linePtr"].ctypes.data, t["segments"].ctypes.data)


def level_from_product(H, S, level):
    """Collect a LevelData from product handles (hierarchy arrays are index data, not results)."""
    info = H.level_info(level)
    N, D, B = info["texels"], info["deep"], info["boundary"]
    tasks = {k: H.array(level, "tasks." + k) for k in ["taskPtr", "blockBegin", "blockDeepEnd", "blockPadEnd", "linePtr", "segments"]}
    bd = Csr(S.array(level, "bd.rowptr", np.int32), S.array(level, "bd.col", np.int32), S.array(level, "bd.val"), N)
    bb = Csr(S.array(level, "bb.rowptr", np.int32), S.array(level, "bb.col", np.int32), S.array(level, "bb.val"), B)
    kw = {}
    if level > 0:
        kw["restr_lines"] = H.array(level, "restrictionLines")
    if level + 1 < H.levels:
        kw["prol_lines"] = H.array(level, "prolongationLines")
        kw["brestr"] = Csr(H.array(level, "boundaryRestriction.rowptr"), H.array(level, "boundaryRestriction.col"),
                           H.array(level, "boundaryRestriction.val", np.float64), N)
    return LevelData(N, D, B, H.array(level, "boundaryGlobalIndex"), H.array(level, "rasterLines"), tasks,
                     S.array(level, "deepRecip"), S.array(level, "deepRaw"), bd, bb, **kw)


# ------------------------------------------------------------------ kernels
def band_relax(L, x, b, r=None):
    """Deep Gauss-Seidel sweep of the band tasks, in place on x ([K, N]); residual into r."""
    lib = load()
    K = x.shape[0]
    t = L.ctasks()
    lib.oracle_band_relax(ctypes.byref(t), ctypes.c_void_p(L.coef.ctypes.data), ctypes.c_void_p(x.ctypes.data),
                          ctypes.c_void_p(b.ctypes.data), ctypes.c_void_p(r.ctypes.data) if r is not None else None, L.N, K)


def boundary_rhs(L, x, b):
    vb = np.zeros((x.shape[0], L.B))
    bd = L.bd.c()
    load().oracle_boundary_rhs(L.B, ctypes.c_void_p(L.bgi.ctypes.data), ctypes.byref(bd), ctypes.c_void_p(x.ctypes.data),
                               ctypes.c_void_p(b.ctypes.data), ctypes.c_void_p(vb.ctypes.data), L.N, x.shape[0])
    return vb


def boundary_residual(L, x, b, r):
    bd, bb = L.bd.c(), L.bb.c()
    load().oracle_boundary_residual(L.B, ctypes.c_void_p(L.bgi.ctypes.data), ctypes.byref(bd), ctypes.byref(bb),
                                    ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(b.ctypes.data), ctypes.c_void_p(r.ctypes.data),
                                    L.N, x.shape[0])


def restrict(Lf, Lc, fine):
    K = fine.shape[0]
    out = np.zeros((K, Lc.N))
    rs = Lf.brestr.c()
    load().oracle_restrict(Lc.B, ctypes.c_void_p(Lc.bgi.ctypes.data), ctypes.byref(rs), len(Lc.restr_lines) // 5,
                           ctypes.c_void_p(Lc.restr_lines.ctypes.data), ctypes.c_void_p(fine.ctypes.data), Lf.N,
                           ctypes.c_void_p(out.ctypes.data), Lc.N, K)
    return out


def prolong_add(Lf, Lc, coarse, fine):
    load().oracle_prolong_add(len(Lf.prol_lines) // 6, ctypes.c_void_p(Lf.prol_lines.ctypes.data), ctypes.c_void_p(coarse.ctypes.data),
                              Lc.N, ctypes.c_void_p(fine.ctypes.data), Lf.N, coarse.shape[0])


def apply(L, coef_raw, bd, bb, x):
    y = np.zeros_like(x)
    cbd, cbb = bd.c(), bb.c()
    coef_raw = _d(coef_raw)
    load().oracle_apply(len(L.raster) // 5, ctypes.c_void_p(L.raster.ctypes.data), ctypes.c_void_p(coef_raw.ctypes.data), L.B,
                        ctypes.c_void_p(L.bgi.ctypes.data), ctypes.byref(cbd), ctypes.byref(cbb), ctypes.c_void_p(x.ctypes.data),
                        ctypes.c_void_p(y.ctypes.data), L.N, x.shape[0])
    return y


# ------------------------------------------------------------------ matrices (independent assembly)
def full_matrix(L, coef_raw=None, bd=None, bb=None):
    """Sparse N x N matrix of a level from its deep rows (raster lines, 10-slot stencils) and
    boundary rows (FullMatrixConstruction, UpdateCoefficients.inl:137-219)."""
    coef_raw = L.coef_raw if coef_raw is None else coef_raw
    bd = L.bd if bd is None else bd
    bb = L.bb if bb is None else bb
    rows, cols, vals = [], [], []
    c = coef_raw.reshape(-1, 10)
    for s, e, p, n, d in L.raster.reshape(-1, 5):
        ln = e - s + 1
        i = np.arange(ln)
        for dy, base in enumerate((p, s, n)):
            for dx in range(3):
                rows.append(s + i); cols.append(base + i + dx - 1); vals.append(c[d + i, 3 * dy + dx])
    bdm = bd.scipy().tocoo()
    rows.append(L.bgi[bdm.row]); cols.append(bdm.col); vals.append(bdm.data)
    bbm = bb.scipy().tocoo()
    rows.append(L.bgi[bbm.row]); cols.append(L.bgi[bbm.col]); vals.append(bbm.data)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(L.N, L.N))
    return A


def prolongation_matrix(Lf, Nc):
    """P (Nf x Nc) from the prolongation lines (ProlongationAndRestriction.inl:173-200)."""
    rows, cols, vals = [], [], []
    for start, length, center, prev, nxt, aligned in Lf.prol_lines.reshape(-1, 6):
        off = 0 if aligned else 1
        i = np.arange(length)
        a, b = (i + off) // 2, (i + off + 1) // 2
        if center != -1:
            for q in (a, b):
                rows.append(start + i); cols.append(center + q); vals.append(np.full(length, 0.5))
        else:
            for base in (prev, nxt):
                for q in (a, b):
                    rows.append(start + i); cols.append(base + q); vals.append(np.full(length, 0.25))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(Lf.N, Nc))


# ------------------------------------------------------------------ V-cycle
class Solvers:
    """Exact solvers of the boundary blocks (levels < L-1) and of the coarsest level."""

    def __init__(self, levels, coarse_matrix):
        self.boundary = [spla.splu(L.bb.scipy().tocsc()) for L in levels[:-1]]
        self.coarse = spla.splu(sp.csc_matrix(coarse_matrix))


_pool = None


def set_threads(n):
    """Threads of the C kernels' parallel loops and of the per-column exact solves."""
    global _pool
    load().oracle_set_threads(int(n))
    _pool = None


def _solve(lu, rhs):
    """Exact solve per column, the columns in parallel (the reference solves its three channels in
    an OpenMP loop, Solver.h:197-198)."""
    global _pool
    K = rhs.shape[0]
    if K == 1:
        return np.ascontiguousarray(lu.solve(rhs[0])[None])
    if _pool is None:
        from concurrent.futures import ThreadPoolExecutor
        _pool = ThreadPoolExecutor(max(1, min(16, load().oracle_max_threads())))
    return np.ascontiguousarray(np.stack(list(_pool.map(lambda k: lu.solve(rhs[k]), range(K)))))


def relaxation(L, lu, x, b, r=None):
    """Relaxation (boundary first) / RelaxationAndResidual (IterativeSolvers.inl:174-558)."""
    vb = boundary_rhs(L, x, b)
    x[:, L.bgi] = _solve(lu, vb)
    band_relax(L, x, b, r)
    if r is not None:
        boundary_residual(L, x, b, r)


def vcycle(levels, solvers, x, b):
    """One reference V-cycle on level-0 arrays x (in/out, [K, N0]) and b."""
    nl = len(levels)
    K = x.shape[0]
    X = [x] + [np.zeros((K, L.N)) for L in levels[1:]]
    Bv = [b] + [None] * (nl - 1)
    for l in range(nl - 1):
        r = np.zeros((K, levels[l].N))
        relaxation(levels[l], solvers.boundary[l], X[l], Bv[l], r)
        Bv[l + 1] = restrict(levels[l], levels[l + 1], r)
    X[nl - 1][:] = _solve(solvers.coarse, Bv[nl - 1])
    for l in range(nl - 2, -1, -1):
        prolong_add(levels[l], levels[l + 1], X[l + 1], X[l])
        relaxation(levels[l], solvers.boundary[l], X[l], Bv[l], None)
    return x


# ---------------------------------------------------------------------------- normalised gradient
def vf_samples(H):
    """The library's normalised-gradient integration samples (exported arrays, FP32 values)."""
    cell = H.array(0, "vf.cellData", np.float64).reshape(-1, 13)
    return dict(corners=H.array(0, "vf.cellCorners").reshape(-1, 4), pos=cell[:, 0:2], tensor=cell[:, 2:5],
                v=cell[:, 5:13].reshape(-1, 4, 2), inv=H.array(0, "vf.invCell", np.float64),
                gptr=H.array(0, "vf.genPtr"), gtex=H.array(0, "vf.genTexel"),
                gten=H.array(0, "vf.genTensor", np.float64).reshape(-1, 3),
                gG=H.array(0, "vf.genG", np.float64).reshape(-1, 2), gV=H.array(0, "vf.genV", np.float64).reshape(-1, 2))


def _normalise(n, t):
    """X = g^-1 n / sqrt(n^T g^-1 n), zero where the gradient vanishes
    (VectorFieldIntergration.inl:508-518)."""
    ax = t[:, 0] * n[:, 0] + t[:, 1] * n[:, 1]
    ay = t[:, 1] * n[:, 0] + t[:, 2] * n[:, 1]
    l2 = n[:, 0] * ax + n[:, 1] * ay
    il = np.where(l2 > 0, 1.0 / np.sqrt(np.where(l2 > 0, l2, 1.0)), 0.0)
    return ax * il, ay * il


def vf_cell_gradients(S, u):
    """Gradient (UV units) of the bilinear interpolant of the 4 cell corners at each cell sample
    (VectorFieldIntergration.inl:500-503, corner order (0,0),(1,0),(1,1),(0,1)), written in the
    algebraically identical difference form that is exactly zero for a constant."""
    cv, (x, y) = u[S["corners"]], S["pos"].T
    gx = ((1 - y) * (cv[:, 1] - cv[:, 0]) + y * (cv[:, 2] - cv[:, 3])) * S["inv"][0]
    gy = ((1 - x) * (cv[:, 3] - cv[:, 0]) + x * (cv[:, 2] - cv[:, 1])) * S["inv"][1]
    return np.stack([gx, gy], 1)


def vf_divergence(S, u, N):
    """rhs_t = sum over samples of < X, v(t) >, X the metric-normalised negative gradient of the
    potential u at the sample (IntergrateNormalizedVectorFieldFromSamples,
    VectorFieldIntergration.inl:452-582: bilinear cell gradient at the sample position, scaled by
    the inverse cell size; the quadratic seam samples expand their node values into texel
    weights).  u: [N] or [K, N] FP64."""
    u = np.atleast_2d(u)
    out = np.zeros_like(u)
    c = S["corners"]
    for k in range(u.shape[0]):
        X, Y = _normalise(-vf_cell_gradients(S, u[k]), S["tensor"])
        for j in range(4):
            np.add.at(out[k], c[:, j], X * S["v"][:, j, 0] + Y * S["v"][:, j, 1])
        # general samples
        ns = len(S["gptr"]) - 1
        if ns:
            seg = np.repeat(np.arange(ns), np.diff(S["gptr"]))
            ut = u[k][S["gtex"]] - u[k][S["gtex"][S["gptr"][:-1]]][seg]  # differences to the first entry
            n = -np.stack([np.bincount(seg, S["gG"][:, 0] * ut, ns), np.bincount(seg, S["gG"][:, 1] * ut, ns)], 1)
            X, Y = _normalise(n, S["gten"])
            np.add.at(out[k], S["gtex"], X[seg] * S["gV"][:, 0] + Y[seg] * S["gV"][:, 1])
    return out


def modulated_stiffness(deep, deep_cells, deep_global, b_rowptr, b_cell, b_val, b_global, mask, K, N):
    """CellStiffnessToTexelStiffness (include/Src/IterativeSolvers.inl:1170-1208): boundary texels
    take their sparse row over cells times the cell mask (:1177-1181); deep texel d takes its four
    cell products weighted by the masks of the cells (i-1,j-1), (i,j-1), (i-1,j), (i,j), summed in
    that order (:1184-1192).  deep: [K][4D]; deep_cells: [2D] (the two left cells; the right ones
    follow); b_val: [K][nnz].  Returns [K][N]."""
    D = len(deep_global)
    out = np.zeros((K, N))
    cp, cn = deep_cells[0::2], deep_cells[1::2]
    m = np.stack([mask[cp], mask[cp + 1], mask[cn], mask[cn + 1]], axis=1)
    c = deep.reshape(K, D, 4)
    for k in range(K):
        out[k, deep_global] = ((c[k, :, 0] * m[:, 0] + c[k, :, 1] * m[:, 1]) + c[k, :, 2] * m[:, 2]) + c[k, :, 3] * m[:, 3]
    nnz = len(b_cell)
    v = b_val.reshape(K, nnz)
    rows = np.repeat(np.arange(len(b_global)), np.diff(b_rowptr))
    for k in range(K):
        out[k, b_global] = np.bincount(rows, weights=v[k] * mask[b_cell], minlength=len(b_global))
    return out

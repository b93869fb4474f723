"""ORACLE — test infrastructure only (tests/).  Independent numpy restatement of the texel
hierarchy's index work, from the mesh's per-corner texture coordinates:

  charts       triangles joined across edges whose two texture coordinates agree
               (ChartDecomposition.h:30-72), numbered by their first triangle;
  grid charts  origin = the texel centre below-left of the chart's UV box centre, grid reaching
               past the box by whole cells (AtlasCharts.inl:49-62, HierarchyConstruction.inl:506-528);
  node types   -1 inactive, 1 inside a chart triangle, 0 a corner of a boundary cell, 2 deep (all
               four cells around the node interior); cell types -1, 0 boundary (crossed by a chart
               boundary edge, or holding a whole one), 1 interior (four interior corners)
               (HierarchyConstruction.inl:60-226);
  numbering    chart-major raster order of active texels, deep and boundary sub-numberings
               (:230-290), raster lines of deep runs (:372-404);
  band tasks   blocks of 128 x 16 texels advancing 126 columns and 16 rows, deep segments of the
               block interior and padded segments of the whole block, one task per block row
               (:406-490).

Written as array operations over all (triangle, node) and (edge, cell side) pairs rather than the
reference's per-triangle loops; it shares no code with the library.  Compared bit for bit with
the library's exported arrays by tests/test_hierarchy_oracle.py."""
import numpy as np

BLOCK_W, BLOCK_H, PAD_W, PAD_H = 128, 16, 2, 0


def charts_of(tri, uv):
    """Chart index per triangle: union of triangles that share an edge with equal texture
    coordinates at both ends (opposite orientation), numbered by first triangle."""
    nt = len(tri)
    a = tri.reshape(-1)
    b = tri[:, [1, 2, 0]].reshape(-1)
    he_uv0 = uv.reshape(nt, 3, 2).reshape(-1, 2)
    he_uv1 = uv.reshape(nt, 3, 2)[:, [1, 2, 0]].reshape(-1, 2)
    key = {}
    for h in range(3 * nt):
        key[(int(a[h]), int(b[h]))] = h
    parent = np.arange(nt)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for h in range(3 * nt):
        o = key.get((int(b[h]), int(a[h])))
        if o is None:
            continue
        if np.array_equal(he_uv0[h], he_uv1[o]) and np.array_equal(he_uv1[h], he_uv0[o]):
            ra, rb = find(h // 3), find(o // 3)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    roots = np.array([find(t) for t in range(nt)])
    _, first = np.unique(roots, return_index=True)
    order = np.argsort(first)
    label = np.empty(nt, np.int64)
    for c, r in enumerate(roots[first[order]]):
        label[roots == r] = c
    return label


def boundary_half_edges(tri, uv, chart):
    """Half-edges (triangle, k) with no opposite half-edge in the same chart carrying the same
    texture coordinates (chart boundary: mesh boundary or texture seam)."""
    nt = len(tri)
    uvt = uv.reshape(nt, 3, 2)
    key = {(int(tri[t, k]), int(tri[t, (k + 1) % 3])): (t, k) for t in range(nt) for k in range(3)}
    out = []
    for t in range(nt):
        for k in range(3):
            o = key.get((int(tri[t, (k + 1) % 3]), int(tri[t, k])))
            inner = False
            if o is not None:
                to, ko = o
                inner = (chart[to] == chart[t] and np.array_equal(uvt[t, k], uvt[to, (ko + 1) % 3])
                         and np.array_equal(uvt[t, (k + 1) % 3], uvt[to, ko]))
            if not inner:
                out.append((t, k))
    return out


def grid_chart(uvc, W, H, cw, ch):
    """Grid geometry of one chart from its corner UVs: (width, height, corner UV, origin texel,
    offset of the origin in the grid)."""
    lo, hi = uvc.min(axis=0), uvc.max(axis=0)
    mid = (lo + hi) / 2.0
    mid = np.array([mid[0] * W, mid[1] * H]) - 0.5
    mid = np.floor(mid)
    origin_coords = np.round(mid).astype(np.int64)
    origin = (mid + 0.5) / np.array([W, H], float)
    cs = np.array([cw, ch])
    half0 = np.ceil((origin - lo) / cs).astype(np.int64)
    half1 = np.ceil((hi - origin) / cs).astype(np.int64)
    corner = origin - cs * half0
    return int(half0[0] + half1[0] + 1), int(half0[1] + half1[1] + 1), corner, origin_coords - half0, half0


def classify(tpos, edges, w, h, cw, ch):
    """Node and cell types of one chart grid.  tpos: [nt, 3, 2] triangle corners relative to the
    grid corner; edges: [ne, 2, 2] boundary edges (start, end) relative to the grid corner."""
    node = np.full((h, w), -1, np.int64)
    cell = np.full((h - 1, w - 1), -1, np.int64)
    inv = np.array([1.0 / cw, 1.0 / ch])
    # (1) nodes inside a triangle (closed barycentric test), over every (triangle, bbox node) pair
    lo = np.floor(np.maximum(tpos.min(axis=1), 0.0) * inv).astype(np.int64)
    hi = np.ceil(np.minimum(tpos.max(axis=1), 1.0) * inv).astype(np.int64)
    ext = hi - lo + 1
    cnt = ext[:, 0] * ext[:, 1]
    t_of = np.repeat(np.arange(len(tpos)), cnt)
    k = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    ii = lo[t_of, 0] + k % ext[t_of, 0]
    jj = lo[t_of, 1] + k // ext[t_of, 0]
    p0, d1, d2 = tpos[:, 0], tpos[:, 1] - tpos[:, 0], tpos[:, 2] - tpos[:, 0]
    det = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
    # inverse of [[d1x, d2x], [d1y, d2y]]
    m00, m01, m10, m11 = d2[:, 1] / det, -d2[:, 0] / det, -d1[:, 1] / det, d1[:, 0] / det
    px = ii * cw - p0[t_of, 0]
    py = jj * ch - p0[t_of, 1]
    b0 = m00[t_of] * px + m01[t_of] * py
    b1 = m10[t_of] * px + m11[t_of] * py
    inside = (b0 >= 0) & (b1 >= 0) & (b0 + b1 <= 1)
    node[jj[inside], ii[inside]] = 1
    # (2) boundary edges: cell sides they cross, and cells that hold a whole edge
    touched = np.zeros((h, w), bool)
    e0, e1 = edges[:, 0], edges[:, 1]
    elo = np.floor(np.maximum(np.minimum(e0, e1), 0.0) * inv).astype(np.int64)
    ehi = np.ceil(np.minimum(np.maximum(e0, e1), 1.0) * inv).astype(np.int64)
    dirv = e1 - e0
    en = np.stack([dirv[:, 1], -dirv[:, 0]], 1)
    en = en / np.sqrt((en ** 2).sum(1))[:, None]
    elev = ((en * e0).sum(1) + (en * e1).sum(1)) / 2.0
    ext = ehi - elo + 1
    cnt = ext[:, 0] * ext[:, 1]
    e_of = np.repeat(np.arange(len(edges)), cnt)
    k = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    ci = elo[e_of, 0] + k % ext[e_of, 0]
    cj = elo[e_of, 1] + k // ext[e_of, 0]
    for c in (0, 1):
        a = np.stack([ci * cw, cj * ch], 1)
        b = a.copy()
        b[:, c] += cw if c == 0 else ch
        side = b - a
        sn = np.stack([side[:, 1], -side[:, 0]], 1)
        sn = sn / np.sqrt((sn ** 2).sum(1))[:, None]
        slev = ((sn * a).sum(1) + (sn * b).sum(1)) / 2.0
        n_e, l_e = en[e_of], elev[e_of]
        opp_e = ((n_e * a).sum(1) - l_e) * ((n_e * b).sum(1) - l_e) < 0.0
        opp_s = ((sn * e1[e_of]).sum(1) - slev) * ((sn * e0[e_of]).sum(1) - slev) < 0.0
        hit = opp_e & opp_s
        hi_, hj = ci[hit], cj[hit]
        if c == 0:  # horizontal side from (i, j): the cells above and below it
            cell[hj - 1, hi_] = 0
            cell[hj, hi_] = 0
        else:  # vertical side: the cells left and right of it
            cell[hj, hi_ - 1] = 0
            cell[hj, hi_] = 0
        for dn in (-1, 0, 1):
            for dc in (0, 1):
                ni, nj = hi_.copy(), hj.copy()
                if c == 0:
                    nj, ni = nj + dn, ni + dc
                else:
                    ni, nj = ni + dn, nj + dc
                touched[np.clip(nj, 0, h - 1), np.clip(ni, 0, w - 1)] = True
    whole = (elo[:, 0] + 1 == ehi[:, 0]) & (elo[:, 1] + 1 == ehi[:, 1])
    for i0, j0 in elo[whole]:
        cell[j0, i0] = 0
        for di in (0, 1):
            for dj in (0, 1):
                touched[min(max(j0 + dj, 0), h - 1), min(max(i0 + di, 0), w - 1)] = True
    node[touched & (node != 1)] = 0
    # (3) interior cells
    four = (node[:-1, :-1] == 1) & (node[:-1, 1:] == 1) & (node[1:, :-1] == 1) & (node[1:, 1:] == 1)
    cell[four & (cell == -1)] = 1
    # (4) deep nodes: the four cells around the node are interior
    deep = np.zeros((h, w), bool)
    deep[1:-1, 1:-1] = (cell[:-1, :-1] == 1) & (cell[:-1, 1:] == 1) & (cell[1:, :-1] == 1) & (cell[1:, 1:] == 1)
    node[deep] = 2
    return node, cell


def band_tasks(node, gidx, didx):
    """Band tasks of one chart: list of tasks, each a list of blocks (deep rows, padded rows), each
    row a list of segments (currentStart, currentEnd, previousStart, nextStart, deepStart)."""
    h, w = node.shape
    isdeep = didx >= 0
    hoff, voff = BLOCK_W - PAD_W, BLOCK_H - PAD_H
    nbx = (w - PAD_W - 1) // hoff + 1
    nby = (h - PAD_H - 1) // voff + 1

    def runs(j, i0, i1):
        segs = []
        row = isdeep[j, i0:i1 + 1]
        if not row.any():
            return segs
        d = np.diff(np.concatenate([[0], row.astype(np.int8), [0]]))
        for s, e in zip(np.nonzero(d == 1)[0], np.nonzero(d == -1)[0] - 1):
            a, b = i0 + s, i0 + e
            segs.append((int(gidx[j, a]), int(gidx[j, b]), int(gidx[j - 1, a]), int(gidx[j + 1, a]), int(didx[j, a])))
        return segs

    tasks = []
    for bj in range(nby):
        vs, ve = bj * voff, min((bj + 1) * voff + PAD_H + 1, h - 1)
        blocks = []
        for bi in range(nbx):
            hs, he = bi * hoff, min((bi + 1) * hoff + PAD_W + 1, w - 1)
            deep_rows = [r for r in (runs(j, hs + 1, he - 1) for j in range(vs + 1, ve)) if r]
            if not deep_rows:
                continue
            pad_rows = [r for r in (runs(j, hs, he) for j in range(vs, ve + 1)) if r]
            blocks.append((deep_rows, pad_rows))
        if blocks:
            tasks.append(blocks)
    return tasks


def hierarchy_level(tri, uv, W, H, level):
    """Every index array of one level that the library exports, restated: dict of numpy arrays
    named as the library's (charts, nodeType, cellType, nodeInfo, rasterLines, tasks.*)."""
    cw, ch = 2.0 ** level / W, 2.0 ** level / H
    chart = charts_of(tri, uv)
    bhe = boundary_half_edges(tri, uv, chart)
    uvt = uv.reshape(len(tri), 3, 2)
    charts, ntypes, ctypes_, nodeinfo, raster = [], [], [], [], []
    seg_all, line_ptr, blk_begin, blk_deep, blk_pad, task_ptr = [], [0], [], [], [], [0]
    gbase = dbase = bbase = 0
    for c in range(chart.max() + 1):
        ts = np.nonzero(chart == c)[0]
        w, h, corner, cc, off = grid_chart(uvt[ts].reshape(-1, 2), W, H, cw, ch)
        tpos = uvt[ts] - corner
        e = np.array([[uvt[t, k], uvt[t, (k + 1) % 3]] for t, k in bhe if chart[t] == c]).reshape(-1, 2, 2) - corner
        node, cell = classify(tpos, e, w, h, cw, ch)
        active = node != -1
        gidx = np.full((h, w), -1, np.int64)
        gidx[active] = gbase + np.arange(active.sum())
        didx = np.full((h, w), -1, np.int64)
        dm = node == 2
        didx[dm] = dbase + np.arange(dm.sum())
        jj, ii = np.nonzero(active)
        nodeinfo.append(np.stack([np.full(len(ii), c), ii, jj, node[jj, ii]], 1))
        for j in range(h):
            row = dm[j]
            if not row.any():
                continue
            d = np.diff(np.concatenate([[0], row.astype(np.int8), [0]]))
            for s, t in zip(np.nonzero(d == 1)[0], np.nonzero(d == -1)[0] - 1):
                raster.append((gidx[j, s], gidx[j, t], gidx[j - 1, s], gidx[j + 1, s], didx[j, s]))
        for blocks in band_tasks(node, gidx, didx):
            for deep_rows, pad_rows in blocks:
                blk_begin.append(len(line_ptr) - 1)
                for r in deep_rows:
                    seg_all.extend(r)
                    line_ptr.append(len(seg_all))
                blk_deep.append(len(line_ptr) - 1)
                for r in pad_rows:
                    seg_all.extend(r)
                    line_ptr.append(len(seg_all))
                blk_pad.append(len(line_ptr) - 1)
            task_ptr.append(len(blk_begin))
        charts.append((w, h, cc[0], cc[1], off[0], off[1], gbase, dbase, bbase))
        gbase += int(active.sum())
        dbase += int(dm.sum())
        bbase += int(((node == 0) | (node == 1)).sum())
        ntypes.append(node.reshape(-1))
        ctypes_.append(cell.reshape(-1))
    i32 = lambda a: np.asarray(a, np.int64).reshape(-1)  # noqa: E731
    return {"charts": i32(charts), "nodeType": np.concatenate(ntypes), "cellType": np.concatenate(ctypes_),
            "nodeInfo": np.concatenate(nodeinfo).reshape(-1), "rasterLines": i32(raster), "tasks.taskPtr": i32(task_ptr),
            "tasks.blockBegin": i32(blk_begin), "tasks.blockDeepEnd": i32(blk_deep), "tasks.blockPadEnd": i32(blk_pad),
            "tasks.linePtr": i32(line_ptr), "tasks.segments": i32(seg_all)}

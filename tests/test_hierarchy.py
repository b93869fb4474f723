"""Texel hierarchy invariants (include/Src/HierarchyConstruction.inl, ProlongationAndRestriction.inl).

The reference builder cannot be compiled without stand-ins for headers this image lacks (its
image reader needs the pre-1.5 libpng layout), so the hierarchy is pinned by the rules it must
satisfy, checked from independent recomputations on the exported grids."""
import numpy as np
import pytest

import oracle
from helpers import cube_system, icosphere_system, oracle_levels


@pytest.fixture(scope="module", params=["cube", "icosphere"])
def sys_(request):
    if request.param == "cube":
        return cube_system(128, 3)
    return icosphere_system(256, 4, 3)


def chart_grids(H, level):
    charts = H.array(level, "charts").reshape(-1, 9)
    nt = H.array(level, "nodeType")
    ct = H.array(level, "cellType")
    out, a, b = [], 0, 0
    for w, h, *_ in charts:
        out.append((nt[a:a + w * h].reshape(h, w), ct[b:b + (w - 1) * (h - 1)].reshape(h - 1, w - 1)))
        a += w * h
        b += (w - 1) * (h - 1)
    return out


def test_partition_and_raster_order(sys_):
    _, H, _ = sys_
    for l in range(H.levels):
        info = H.level_info(l)
        assert info["texels"] == info["deep"] + info["boundary"]
        ni = H.array(l, "nodeInfo").reshape(-1, 4)
        key = ni[:, 0].astype(np.int64) * 10**8 + ni[:, 2].astype(np.int64) * 10**4 + ni[:, 1]
        assert np.all(np.diff(key) > 0), "global index is not chart-major raster order"
        bad = H.array(l, "boundaryAndDeepIndex")
        bgi, dgi = H.array(l, "boundaryGlobalIndex"), H.array(l, "deepGlobalIndex")
        assert np.array_equal(bad[bgi], np.arange(1, len(bgi) + 1))
        assert np.array_equal(bad[dgi], -np.arange(1, len(dgi) + 1))
        assert np.all((ni[bgi, 3] == 0) | (ni[bgi, 3] == 1)) and np.all(ni[dgi, 3] == 2)


def test_deep_rule(sys_):
    """A texel is deep iff its four incident cells are interior (HierarchyConstruction.inl:219-226)."""
    _, H, _ = sys_
    for l in range(H.levels):
        for nt, ct in chart_grids(H, l):
            h, w = nt.shape
            interior = ct == 1
            deep = np.zeros_like(nt, dtype=bool)
            deep[1:h - 1, 1:w - 1] = interior[:-1, :-1] & interior[:-1, 1:] & interior[1:, :-1] & interior[1:, 1:]
            assert np.array_equal(deep, nt == 2)
            # interior cells have four interior (type >= 1) corners
            ok = (nt[:-1, :-1] >= 1) & (nt[:-1, 1:] >= 1) & (nt[1:, :-1] >= 1) & (nt[1:, 1:] >= 1)
            assert np.all(ok[interior])


def test_raster_lines_cover_deep_texels(sys_):
    _, H, _ = sys_
    for l in range(H.levels):
        ni = H.array(l, "nodeInfo").reshape(-1, 4)
        rl = H.array(l, "rasterLines").reshape(-1, 5)
        seen = np.zeros(len(ni), int)
        for s, e, p, n, d in rl:
            seen[s:e + 1] += 1
            assert ni[p, 1] == ni[s, 1] and ni[p, 2] == ni[s, 2] - 1
            assert ni[n, 1] == ni[s, 1] and ni[n, 2] == ni[s, 2] + 1
        assert np.array_equal(seen, (ni[:, 3] == 2).astype(int))


def test_band_tasks_cover_and_are_race_free(sys_):
    """Every deep texel is relaxed; tasks of one colour never touch rows within one of each other
    (IterativeSolvers.inl:511-521, HierarchyConstruction.inl:397-490)."""
    _, H, _ = sys_
    for l in range(H.levels):
        ni = H.array(l, "nodeInfo").reshape(-1, 4)
        t = {k: H.array(l, "tasks." + k) for k in ["taskPtr", "blockBegin", "blockDeepEnd", "blockPadEnd", "linePtr", "segments"]}
        seg = t["segments"].reshape(-1, 5)
        relaxed = np.zeros(len(ni), bool)
        rows_of_task = []
        for task in range(len(t["taskPtr"]) - 1):
            rows = set()
            for blk in range(t["taskPtr"][task], t["taskPtr"][task + 1]):
                for r in range(t["blockBegin"][blk], t["blockDeepEnd"][blk]):
                    for s in range(t["linePtr"][r], t["linePtr"][r + 1]):
                        cs, ce = seg[s, 0], seg[s, 1]
                        relaxed[cs:ce + 1] = True
                        assert np.all(ni[cs:ce + 1, 3] == 2)
                        rows.add((ni[cs, 0], ni[cs, 2]))
            rows_of_task.append(rows)
        assert np.array_equal(relaxed, ni[:, 3] == 2)
        for color in range(2):
            touched = {}
            for task in range(color, len(rows_of_task), 2):
                for (c, j) in rows_of_task[task]:
                    for dj in (-1, 0, 1):
                        other = touched.get((c, j + dj))
                        assert other is None or other == task, "same-colour tasks share a stencil row"
                for (c, j) in rows_of_task[task]:
                    touched[(c, j)] = task


def test_prolongation_rows_and_linear_precision(sys_):
    """P rows sum to one and P reproduces fields linear in chart-grid coordinates
    (ProlongationAndRestriction.inl:99-216)."""
    _, H, S = sys_
    L = oracle_levels(H, S)
    for l in range(H.levels - 1):
        P = oracle.prolongation_matrix(L[l], L[l + 1].N)
        assert np.allclose(P @ np.ones(L[l + 1].N), 1.0, atol=1e-14)
        nf, nc = H.array(l, "nodeInfo").reshape(-1, 4), H.array(l + 1, "nodeInfo").reshape(-1, 4)
        cf, cc = H.array(l, "charts").reshape(-1, 9), H.array(l + 1, "charts").reshape(-1, 9)
        # fine coordinates relative to the chart centre; coarse ones scaled by two
        fu = nf[:, 1] - cf[nf[:, 0], 4]; fv = nf[:, 2] - cf[nf[:, 0], 5]
        cu = 2 * (nc[:, 1] - cc[nc[:, 0], 4]); cv = 2 * (nc[:, 2] - cc[nc[:, 0], 5])
        for a, b in [(1.0, 0.0), (0.0, 1.0), (0.3, -0.7)]:
            coarse = a * cu + b * cv + 5.0 * nc[:, 0]
            fine = a * fu + b * fv + 5.0 * nf[:, 0]
            assert np.allclose(P @ coarse, fine, atol=1e-12)


def test_boundary_restriction_is_prolongation_transpose(sys_):
    _, H, S = sys_
    L = oracle_levels(H, S)
    for l in range(H.levels - 1):
        P = oracle.prolongation_matrix(L[l], L[l + 1].N).tocsc()
        R = L[l].brestr.scipy()
        PT = P.T.tocsr()[L[l + 1].bgi, :]
        assert abs(R - PT).max() < 1e-15

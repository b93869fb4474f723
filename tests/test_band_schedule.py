"""The GPU band kernel runs the blocks of a band task overlapped in time (KBandGs: block b's local
steps at task steps blockStart[b] + t).  This checks the schedule the library computes against the
reference's order (blocks of a task one after another, IterativeSolvers.inl:511-515) texel by texel:
every pair of accesses to one texel by consecutive blocks (x written then read, read then written,
written twice; residual written twice) keeps its order, blocks further apart touch disjoint texels,
and no wave group holds two blocks at once.  A schedule that starts a block after half its lag fails."""
import numpy as np
import pytest

import gradientdomaintextureprocessing_amd as g

GROUPS = 3  # wave groups of the band kernel (kBandGroups)


def block_accesses(H, level, b):
    """Per access kind, (texel, local step) pairs of block b: x written, x read, residual written."""
    seg = H.array(level, "tasks.segments").reshape(-1, 5).astype(np.int64)
    lp = H.array(level, "tasks.linePtr").astype(np.int64)
    bb, bd, bp = (H.array(level, n).astype(np.int64) for n in ("tasks.blockBegin", "tasks.blockDeepEnd", "tasks.blockPadEnd"))
    W, R, RW = [], [], []

    def line(l):
        gs, ps, ns = [], [], []
        for s in seg[lp[l]:lp[l + 1]]:
            k = np.arange(s[1] - s[0] + 1)
            gs.append(s[0] + k), ps.append(s[2] + k), ns.append(s[3] + k)
        return np.concatenate(gs), np.concatenate(ps), np.concatenate(ns)

    def reads(gv, pv, nv, own):
        parts = [pv - 1, pv, pv + 1, gv - 1, gv + 1, nv - 1, nv, nv + 1] + ([gv] if own else [])
        return np.concatenate(parts)

    for i in range(bd[b] - bb[b]):
        gv, pv, nv = line(bb[b] + i)
        for w in range(3):  # passes at local steps i, i + 2, i + 4
            t = i + 2 * w
            W.append((gv, t)), R.append((reads(gv, pv, nv, False), t))
    for i in range(bp[b] - bd[b]):  # residual of padded row i at local step i + 6
        gv, pv, nv = line(bd[b] + i)
        R.append((reads(gv, pv, nv, True), i + 6)), RW.append((gv, i + 6))
    cat = lambda acc: (np.concatenate([a for a, _ in acc]) if acc else np.zeros(0, np.int64),
                       np.concatenate([np.full(len(a), t) for a, t in acc]) if acc else np.zeros(0, np.int64))
    return cat(W), cat(R), cat(RW)


def extremes(tex, step, start, N, last):
    out = np.full(N, -1 if last else np.iinfo(np.int64).max, np.int64)
    (np.maximum if last else np.minimum).at(out, tex, step + start)
    return out


def check_schedule(H, level, start):
    N = H.level_info(level)["texels"]
    tp = H.array(level, "tasks.taskPtr").astype(np.int64)
    bb, bd, bp = (H.array(level, n).astype(np.int64) for n in ("tasks.blockBegin", "tasks.blockDeepEnd", "tasks.blockPadEnd"))
    steps = np.maximum(bd - bb + 4, bp - bd + 6)
    task_steps = H.array(level, "tasks.taskSteps").astype(np.int64)
    pairs = 0
    for t in range(len(tp) - 1):
        blocks = range(tp[t], tp[t + 1])
        assert task_steps[t] >= max(start[b] + steps[b] for b in blocks)
        acc = {b: block_accesses(H, level, b) for b in blocks}
        for b in blocks:
            if b + GROUPS < tp[t + 1]:  # a wave group holds one block at a time
                assert start[b + GROUPS] >= start[b] + steps[b]
            if b + 1 >= tp[t + 1]:
                continue
            (w0, tw0), (r0, tr0), (q0, tq0) = acc[b]
            (w1, tw1), (r1, tr1), (q1, tq1) = acc[b + 1]
            lastW, lastR, lastQ = (extremes(a, s, start[b], N, True) for a, s in ((w0, tw0), (r0, tr0), (q0, tq0)))
            firstW, firstR, firstQ = (extremes(a, s, start[b + 1], N, False) for a, s in ((w1, tw1), (r1, tr1), (q1, tq1)))
            for before, after in ((lastW, firstR), (lastW, firstW), (lastR, firstW), (lastQ, firstQ)):
                both = (before >= 0) & (after < np.iinfo(np.int64).max)
                if not np.all(before[both] < after[both]):
                    return False
            pairs += 1
            if b + 2 < tp[t + 1]:  # blocks two apart share no texel an access of either writes
                (w2, _), (r2, _), (q2, _) = acc[b + 2]
                assert not (np.intersect1d(w0, np.concatenate([w2, r2])).size or np.intersect1d(r0, w2).size
                            or np.intersect1d(q0, q2).size)
    assert pairs > 0
    return True


@pytest.mark.parametrize("case", ["icosphere1024", "julius1024"])
def test_band_pipeline_keeps_reference_order(case):
    if case == "icosphere1024":
        mesh = g.TexturedMesh.icosphere(4, 1024, 1024, 0.05, 7)
    else:
        import os
        d = np.load(os.path.join(os.path.dirname(__file__), "golden", "julius_mesh.npz"))
        mesh = g.TexturedMesh.from_arrays(d["vertices"], d["triangles"], d["corner_uv"])
    H = g.InitializeHierarchy(mesh, 1024, 1024, 2)
    for level in range(H.levels - 1):
        start = H.array(level, "tasks.blockStart").astype(np.int64)
        assert check_schedule(H, level, start)
        tp = H.array(level, "tasks.taskPtr")
        bb, bd, bp = (H.array(level, n).astype(np.int64) for n in ("tasks.blockBegin", "tasks.blockDeepEnd", "tasks.blockPadEnd"))
        seq = np.maximum(bd - bb + 4, bp - bd + 6)
        sequential = sum(seq[tp[t]:tp[t + 1]].sum() for t in range(len(tp) - 1))
        piped = H.array(level, "tasks.taskSteps").astype(np.int64).sum()
        print(f"[{case} level {level}] task steps {piped} pipelined against {sequential} block after block")
        assert piped < sequential


def test_early_block_breaks_order():
    mesh = g.TexturedMesh.icosphere(4, 1024, 1024, 0.05, 7)
    H = g.InitializeHierarchy(mesh, 1024, 1024, 1)
    start = H.array(0, "tasks.blockStart").astype(np.int64)
    tp = H.array(0, "tasks.taskPtr")
    # the second block of the first task with several blocks, started after half its lag
    t = next(t for t in range(len(tp) - 1) if tp[t + 1] - tp[t] >= 3)
    b = tp[t] + 1
    early = start.copy()
    early[b:tp[t + 1]] -= (start[b] - start[b - 1] + 1) // 2
    assert not check_schedule(H, 0, early)


@pytest.mark.parametrize("case", ["icosphere512", "julius512"])
def test_band_row_elements(case):
    """The band kernel's row element table restates every line's segments texel by texel: global
    index, indices above and below, deep index (-d-1 where the element starts a segment)."""
    if case == "icosphere512":
        mesh = g.TexturedMesh.icosphere(4, 512, 512, 0.05, 7)
    else:
        import os
        d = np.load(os.path.join(os.path.dirname(__file__), "golden", "julius_mesh.npz"))
        mesh = g.TexturedMesh.from_arrays(d["vertices"], d["triangles"], d["corner_uv"])
    H = g.InitializeHierarchy(mesh, 512, 512, 2)
    A = lambda n: H.array(0, n).astype(np.int64)
    seg, lp, le = A("tasks.segments").reshape(-1, 5), A("tasks.linePtr"), A("tasks.lineElem")
    el = A("tasks.elems").reshape(-1, 4)
    assert le[-1] == len(el)
    for l in range(len(lp) - 1):
        want = [(s[0] + o, s[2] + o, s[3] + o, -(s[4] + o) - 1 if o == 0 else s[4] + o) for s in seg[lp[l]:lp[l + 1]]
                for o in range(s[1] - s[0] + 1)]
        assert [tuple(e) for e in el[le[l]:le[l + 1]]] == want, l

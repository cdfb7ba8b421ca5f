"""A candidate grid for gat_edr_filter_grid built in NumPy with the float32
operations the host builder and the kernel use, the box rule of the
candidate filter, and a restatement of the kernel's cell scan — used to check
the scan's coverage and its one-cell-per-pair reporting on the CPU, and to
drive the kernel through its C ABI on the GPU."""
import numpy as np

F = np.float32


def keep(q, c, u):
    """The filter's rule for boxes q (n, 4) and c (n, 4): no side a float
    difference of u or more beyond the other (NaN compares keep the pair)."""
    u = F(u)
    far = ((c[..., 0] - q[..., 2]) >= u) | ((q[..., 0] - c[..., 2]) >= u) | \
          ((c[..., 1] - q[..., 3]) >= u) | ((q[..., 1] - c[..., 3]) >= u)
    return ~far


def cell(v, o, inv, g):
    t = np.floor((F(v) - F(o)) * F(inv)).astype(F)
    return np.minimum(np.maximum(t, F(0)), F(g - 1)).astype(np.int64)


def finite_box(b):
    return np.isfinite(b).all(axis=-1) & (b[..., 0] <= b[..., 2]) & (b[..., 1] <= b[..., 3])


def build(boxes, g, max_cells=64):
    """CSR over g x g cells of the finite boxes' extent; others, and boxes
    touching more than max_cells cells, go to the always-tested list."""
    boxes = boxes.astype(F)
    ok = finite_box(boxes)
    if ok.any():
        x0, y0 = F(boxes[ok, 0].min()), F(boxes[ok, 1].min())
        ex, ey = F(boxes[ok, 2].max() - x0), F(boxes[ok, 3].max() - y0)
    else:
        x0 = y0 = ex = ey = F(0)
    invx = F(F(g) / ex) if ex > 0 else F(0)
    invy = F(F(g) / ey) if ey > 0 else F(0)
    lists = [[] for _ in range(g * g)]
    always = []
    for t, b in enumerate(boxes):
        if not ok[t]:
            always.append(t)
            continue
        c0, c1 = cell(b[0], x0, invx, g), cell(b[2], x0, invx, g)
        r0, r1 = cell(b[1], y0, invy, g), cell(b[3], y0, invy, g)
        if (c1 - c0 + 1) * (r1 - r0 + 1) > max_cells:
            always.append(t)
            continue
        for r in range(r0, r1 + 1):
            for c in range(c0, c1 + 1):
                lists[r * g + c].append(t)
    start = np.zeros(g * g + 1, np.uint32)
    start[1:] = np.cumsum([len(x) for x in lists])
    ids = np.array([t for x in lists for t in x] or [0], np.int32)
    return {"g": g, "x0": x0, "y0": y0, "invx": invx, "invy": invy, "start": start, "id": ids,
            "box": boxes[ids].astype(F), "always": np.array(always or [-1], np.int32), "nalways": len(always)}


def scan(grid, qboxes, cboxes, u):
    """The kernel's pairs, restated: cells of each query's widened box, the
    reporting cell of each candidate, the always list, the every-candidate
    path for queries without a finite box."""
    u = F(u)
    g, out = grid["g"], []
    for qi, q in enumerate(qboxes.astype(F)):
        if not finite_box(q):
            out += [(qi, c) for c in np.nonzero(keep(q, cboxes, u))[0]]
            continue
        for c in grid["always"][:grid["nalways"]]:
            if keep(q, cboxes[c], u):
                out.append((qi, int(c)))
        cx0 = max(cell(q[0] - u, grid["x0"], grid["invx"], g) - 1, 0)
        cx1 = min(cell(q[2] + u, grid["x0"], grid["invx"], g) + 1, g - 1)
        cy0 = max(cell(q[1] - u, grid["y0"], grid["invy"], g) - 1, 0)
        cy1 = min(cell(q[3] + u, grid["y0"], grid["invy"], g) + 1, g - 1)
        for cy in range(cy0, cy1 + 1):
            for cx in range(cx0, cx1 + 1):
                k = cy * g + cx
                for e in range(grid["start"][k], grid["start"][k + 1]):
                    b = grid["box"][e]
                    rx = max(cell(b[0], grid["x0"], grid["invx"], g), cx0)
                    ry = max(cell(b[1], grid["y0"], grid["invy"], g), cy0)
                    if rx == cx and ry == cy and keep(q, b, u):
                        out.append((qi, int(grid["id"][e])))
    return out


def cases(seed, nq, nc):
    """Small boxes, boxes u apart (to the float), NaN, inverted and wide ones."""
    rng = np.random.default_rng(seed)

    def boxes(n):
        lo = rng.random((n, 2)).astype(F)
        side = (rng.random((n, 2)) * 0.01).astype(F)
        b = np.concatenate([lo, lo + side], axis=1).astype(F)
        k = max(1, n // 20)
        b[rng.integers(0, n, k)] = F(np.nan)
        b[rng.integers(0, n, k), 0] = F(0.9)  # inverted when xmax < 0.9
        b[rng.integers(0, n, k)] = np.array([0.1, 0.1, 0.8, 0.9], F)  # wide
        return b

    return boxes(nq), boxes(nc)

"""CPU build of the LBVH from the same Morton codes (test helper).

Restates libigl_amd/csrc/lbvh.hip step by step in numpy/Python so the GPU
node array can be compared bit for bit: float32 box centres and scaling (no
fused multiply-add on either side), 30-bit Morton codes, a stable sort of the
codes, the Karras binary radix tree, the binary nodes' boxes, and the
top-down collapse into 4-wide nodes: each starts from its binary node's two
children and opens its largest member (half area dx dy + dy dz + dz dx in
float32, first on ties) until it has four or only leaves; nodes are numbered
level by level.
"""
import numpy as np

EMPTY = 0xFFFFFFFF
INTERNAL = 0x80000000


def tri_boxes(V, F):
    U = np.asarray(V, dtype=np.float32)
    p = U[F]  # nf x 3 x 3
    lo = p[:, 0, :].copy()
    hi = p[:, 0, :].copy()
    for k in (1, 2):
        v = p[:, k, :]
        lo = np.where(lo > v, v, lo)  # h_min: keep the first on NaN/equality
        hi = np.where(hi < v, v, hi)
    return lo.astype(np.float32), hi.astype(np.float32)


def _expand10(v):
    v = v.astype(np.uint64)
    v = (v * 0x00010001) & 0xFF0000FF
    v = (v * 0x00000101) & 0x0F00F00F
    v = (v * 0x00000011) & 0xC30C30C3
    v = (v * 0x00000005) & 0x49249249
    return v.astype(np.uint32)


def morton_codes(V, F):
    U = np.asarray(V, dtype=np.float32)
    fin = np.where(np.isfinite(U), U, np.nan)
    lo_s = np.nan_to_num(np.nanmin(fin, axis=0), nan=0.0).astype(np.float32)
    hi_s = np.nan_to_num(np.nanmax(fin, axis=0), nan=0.0).astype(np.float32)
    ext = (hi_s - lo_s).astype(np.float32)
    scale = np.where(ext > 0, np.float32(1024.0) / np.where(ext > 0, ext, np.float32(1)), np.float32(0)).astype(np.float32)
    lo, hi = tri_boxes(V, F)
    keep = np.all(np.isfinite(lo), axis=1) & np.all(np.isfinite(hi), axis=1)
    c = (np.float32(0.5) * (lo + hi)).astype(np.float32)
    t = ((c - lo_s).astype(np.float32) * scale).astype(np.float32)
    t = np.minimum(np.maximum(t, np.float32(0)), np.float32(1023))
    ti = np.where(np.isfinite(t), t, 0).astype(np.uint32)
    code = (_expand10(ti[:, 0]) << 2) | (_expand10(ti[:, 1]) << 1) | _expand10(ti[:, 2])
    idx = np.nonzero(keep)[0].astype(np.uint32)
    return code[idx].astype(np.uint32), idx


def _clz32(x):
    return 32 - int(x).bit_length()


def lbvh_nodes(V, F):
    codes, idx = morton_codes(V, F)
    order = np.argsort(codes, kind="stable")
    keys = codes[order]
    tri = idx[order]
    n = keys.size
    if n == 0:
        return np.zeros((0, 4), np.uint32)
    if n == 1:
        return np.array([[tri[0], EMPTY, EMPTY, EMPTY]], np.uint32)
    K = [int(k) for k in keys]

    def delta(i, j):
        if j < 0 or j >= n:
            return -1
        a, b = K[i], K[j]
        if a != b:
            return _clz32(a ^ b)
        return 32 + _clz32(i ^ j)

    m = n - 1
    LEAF = 1 << 31
    left = [0] * m
    right = [0] * m
    parent = [-1] * m
    for i in range(m):
        d = 1 if delta(i, i + 1) - delta(i, i - 1) >= 0 else -1
        dmin = delta(i, i - d)
        lmax = 2
        while delta(i, i + lmax * d) > dmin:
            lmax *= 2
        l = 0
        t = lmax // 2
        while t >= 1:
            if delta(i, i + (l + t) * d) > dmin:
                l += t
            t //= 2
        j = i + l * d
        dn = delta(i, j)
        s = 0
        t = l
        while True:
            t = (t + 1) // 2
            if delta(i, i + (s + t) * d) > dn:
                s += t
            if t <= 1:
                break
        g = i + s * d + (-1 if d < 0 else 0)
        lo, hi = min(i, j), max(i, j)
        L = (LEAF | g) if lo == g else g
        R = (LEAF | (g + 1)) if hi == g + 1 else g + 1
        left[i], right[i] = L, R
        if not L & LEAF:
            parent[L] = i
        if not R & LEAF:
            parent[R] = i
    # binary node boxes (children before parents: a post-order walk)
    lo_t, hi_t = tri_boxes(V, F)
    box = {}

    def leaf_box(k):
        t = int(tri[k])
        return (lo_t[t].copy(), hi_t[t].copy())

    stack = [(0, False)]
    while stack:
        i, ready = stack.pop()
        if ready:
            bs = [leaf_box(c & ~LEAF) if c & LEAF else box[c] for c in (left[i], right[i])]
            box[i] = (np.minimum(bs[0][0], bs[1][0]).astype(np.float32), np.maximum(bs[0][1], bs[1][1]).astype(np.float32))
            continue
        stack.append((i, True))
        for c in (left[i], right[i]):
            if not c & LEAF:
                stack.append((c, False))

    def area(i):
        lo, hi = box[i]
        d = (hi - lo).astype(np.float32)
        return np.float32(np.float32(d[0] * d[1]) + np.float32(d[1] * d[2])) + np.float32(d[2] * d[0])

    out_nodes = []
    frontier = [0]
    base = 0
    while frontier:
        nxt = []
        level = []
        for b in frontier:
            mem = [left[b], right[b]]
            while len(mem) < 4:
                best, ba = -1, np.float32(-1.0)
                for t, x in enumerate(mem):
                    if not x & LEAF:
                        a = area(x)
                        if a > ba:
                            best, ba = t, a
                if best < 0:
                    break
                x = mem[best]
                mem[best:best + 1] = [left[x], right[x]]
            row = []
            for x in mem:
                if x & LEAF:
                    row.append(int(tri[x & ~LEAF]))
                else:
                    row.append(INTERNAL | (base + len(frontier) + len(nxt)))
                    nxt.append(x)
            level.append(row + [EMPTY] * (4 - len(row)))
        out_nodes.extend(level)
        base += len(frontier)
        frontier = nxt
    return np.array(out_nodes, np.uint32).reshape(-1, 4)

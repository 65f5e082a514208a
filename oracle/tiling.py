"""CPU oracle of the tiled detection path (strips, global Otsu, ownership band, dedupe).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates reference subpx/_gpu/centers.py:1778-1862 (strips of tile_h rows stepping
tile_h - overlap, global Otsu on img[::ds, ::ds] times thr_scale, per-strip fixed
threshold, centres kept in [y0 + overlap//2, y1 - overlap//2) except at the first / last
strip) and dedupe_centers (reference subpx/centers.py:335-378: O(N^2) radius union-find
with union by rank and path halving, clusters replaced by their means, ordered by root).
Pinned by: a strip size covering the image reduces to the untiled oracle; three strips
over a page whose holes lie inside the strips that own them equal the untiled oracle, and
holes placed exactly on the band boundaries are kept once each (before the dedupe); dedupe
of well-separated points is the identity and of exact duplicates their mean
(tests/test_oracle_tiling.py).
"""

from __future__ import annotations

import numpy as np

from . import pipeline, segment


def dedupe_centers(centers_xy, eps: float = 1.5) -> np.ndarray:
    pts = np.asarray(centers_xy, dtype=np.float64)
    n = pts.shape[0]
    if n == 0:
        return pts.reshape(0, 2)
    parent = np.arange(n, dtype=np.int32)
    rank = np.zeros(n, dtype=np.int8)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return int(a)

    def union(a, b):
        ra, rb = find(a), find(b)
        if ra == rb:
            return
        if rank[ra] < rank[rb]:
            parent[ra] = rb
        elif rank[ra] > rank[rb]:
            parent[rb] = ra
        else:
            parent[rb] = ra
            rank[ra] += 1

    eps2 = float(eps) ** 2
    for i in range(n):
        d2 = np.sum((pts[i + 1:] - pts[i]) ** 2, axis=1)
        for j in np.where(d2 <= eps2)[0] + (i + 1):
            union(i, int(j))
    roots = np.array([find(i) for i in range(n)], dtype=np.int32)
    unique_roots, inv = np.unique(roots, return_inverse=True)
    sums = np.zeros((unique_roots.size, 2))
    counts = np.zeros(unique_roots.size, dtype=np.int32)
    np.add.at(sums, inv, pts)
    np.add.at(counts, inv, 1)
    return sums / np.maximum(counts[:, None], 1)


def detect_centers_tiled_logquad(img, *, tile_h, overlap=128, otsu_downsample=4, thr_scale=0.8,
                                 area_min=1, area_max=50, pad=3, connectivity=8, invert=False,
                                 dedupe_eps=1.5, refine="logquad", logquad_variant="cpu", raw=False):
    """Centres of the tiled path after the dedupe; ``raw=True`` returns the centres the strips
    kept, before the dedupe."""
    g = np.asarray(img)
    H = g.shape[0]
    ds = max(1, int(otsu_downsample))
    pre = float(segment.otsu_u16(g, ds)) * float(thr_scale)
    t = int(np.floor(pre))
    step = tile_h - overlap
    out = []
    y0 = 0
    while y0 < H:
        y1 = min(H, y0 + tile_h)
        c, _ = pipeline.detect_centers(g[y0:y1], threshold_value=t, invert=invert, area_min=area_min,
                                       area_max=area_max, pad=pad, connectivity=connectivity, refine=refine,
                                       logquad_variant=logquad_variant)
        if c.shape[0]:
            c = c.copy()
            c[:, 1] += y0
            half = overlap // 2
            lo = y0 if y0 == 0 else y0 + half
            hi = y1 if y1 == H else y1 - half
            m = (c[:, 1] >= lo) & (c[:, 1] < hi)
            out.append(c[m])
        if y1 == H:
            break
        y0 += step
    if not out:
        return np.zeros((0, 2))
    kept = np.vstack(out)
    return kept if raw else dedupe_centers(kept, eps=dedupe_eps)

"""Writes tests/golden/tiny_scenes.npz: known-answer vectors for small scenes.

The expected outputs come from tests/_bruteforce.py (pure Python / float64,
per-pixel traversal of Python-int-sorted tile lists), not from the oracle or
the HIP kernels, so they check both.  The reference repository ships no test
vectors for this path; these are the project's own.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from sv_raster_amd import synthetic as S  # noqa: E402
from tests import _bruteforce as BF  # noqa: E402

CASES = [  # (seed, W, H, camera direction or None)
    (0, 32, 32, None),
    (1, 40, 24, (0.02, 0.03, 1.0)),
    (2, 32, 32, (1.0, -0.01, 0.02)),
]


def preprocess_py(cam, centers, lengths):
    """Float32 restatement of the reference preprocess in numpy scalars (same operation order)."""
    f = np.float32
    W, H = cam["W"], cam["H"]
    w2c, c2w = cam["w2c"].astype(f).ravel(), cam["c2w"].astype(f).ravel()
    fx = f(f(0.5) * f(W)) / f(cam["tanfovx"])
    fy = f(f(0.5) * f(H)) / f(cam["tanfovy"])
    ro = c2w[[3, 7, 11]]
    gx, gy = (W + 15) // 16, (H + 15) // 16
    N = centers.shape[0]
    n_dup = np.zeros(N, np.uint32); bbox = np.zeros((N, 2), np.uint32); qb = np.zeros(N, np.uint32)
    near = f(cam["near"])
    for v in range(N):
        c = centers[v].astype(f); r = f(0.5) * f(lengths[v, 0])
        rel = c - ro
        if (rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2] < near * near:
            continue
        umin = vmin = f(1e9); umax = vmax = f(-1e9); bits = 0
        for i in range(8):
            s = np.array([((i >> 2) & 1) * 2 - 1, ((i >> 1) & 1) * 2 - 1, (i & 1) * 2 - 1], f)
            p = c + r * s
            cz = ((w2c[8] * p[0] + w2c[9] * p[1]) + w2c[10] * p[2]) + w2c[11]
            if cz < near:
                continue
            cxv = ((w2c[0] * p[0] + w2c[1] * p[1]) + w2c[2] * p[2]) + w2c[3]
            cyv = ((w2c[4] * p[0] + w2c[5] * p[1]) + w2c[6] * p[2]) + w2c[7]
            iz = f(1.0) / cz
            u, vv = cxv * iz, cyv * iz
            q = (int(p[0] < ro[0]) << 2) | (int(p[1] < ro[1]) << 1) | int(p[2] < ro[2])
            umin = umin if umin < u else u; vmin = vmin if vmin < vv else vv
            umax = umax if umax > u else u; vmax = vmax if vmax > vv else vv
            bits |= 1 << q
        cxh, cyh = f(cam["cx"]) - f(0.5), f(cam["cy"]) - f(0.5)
        bx0 = max(fx * umin + cxh, f(0)); by0 = max(fy * vmin + cyh, f(0))
        bx1 = min(fx * umax + cxh, f(W)); by1 = min(fy * vmax + cyh, f(H))
        if bx0 > bx1 or by0 > by1:
            continue
        r0 = [int(np.rint(z)) for z in (bx0, by0, bx1, by1)]
        bbox[v] = ((r0[0] << 16) | r0[1], (r0[2] << 16) | r0[3])
        tiles = (min(r0[3] // 16, gy - 1) - min(r0[1] // 16, gy - 1) + 1) * \
                (min(r0[2] // 16, gx - 1) - min(r0[0] // 16, gx - 1) + 1)
        n_dup[v] = tiles * bin(bits).count("1"); qb[v] = bits
    return n_dup, bbox, qb


def main():
    out = {}
    for k, (seed, W, H, d) in enumerate(CASES):
        cfg = dict(n=300, lv_min=2, lv_max=5, outside=0, views=[(W, H)])
        sc = S.make_scene(cfg, seed=seed)
        cam = S.make_camera(W, H, seed=seed, direction=d)
        rgbs = np.random.default_rng(seed).uniform(0, 1, (sc["octpath"].shape[0], 3)).astype(np.float32)
        n_dup, bbox, qbits = preprocess_py(cam, sc["vox_centers"], sc["vox_lengths"])
        lists = BF.tile_lists(cam, sc["octpath"].reshape(-1), n_dup, bbox, qbits)
        ent = sorted(((t << 63) | r, v, q) for t, items in lists.items() for r, v, q in items)
        color, T, nc, amb = BF.render(cam, sc["octpath"].reshape(-1), n_dup, bbox, qbits, sc["vox_centers"],
                                      sc["vox_lengths"].reshape(-1), sc["geos"], rgbs)
        p = f"c{k}_"
        out.update({p + "octpath": sc["octpath"], p + "centers": sc["vox_centers"], p + "lengths": sc["vox_lengths"],
                    p + "geos": sc["geos"], p + "rgbs": rgbs, p + "w2c": cam["w2c"], p + "c2w": cam["c2w"],
                    p + "cam": np.array([W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], cam["near"]],
                                        np.float64),
                    p + "n_dup": n_dup, p + "bbox": bbox, p + "qbits": qbits,
                    p + "keys_hi": np.array([e[0] >> 64 for e in ent], np.uint64),
                    p + "keys_lo": np.array([e[0] & ((1 << 64) - 1) for e in ent], np.uint64),
                    p + "vox_list": np.array([(e[2] << 29) | e[1] for e in ent], np.uint32),
                    p + "color": color.astype(np.float32), p + "T": T.astype(np.float32),
                    p + "n_contrib": nc.astype(np.uint32), p + "ambiguous": amb})
    out["n_cases"] = np.array(len(CASES))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny_scenes.npz"), **out)


if __name__ == "__main__":
    main()

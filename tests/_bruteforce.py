"""Independent pure-Python/numpy restatement used to pin the oracle.

It avoids every structure the oracle and the kernels share: no tile lists, no
sorting network, no packed keys.  Per pixel it gathers the voxels whose
quadrant set contains the pixel's ray quadrant and whose bbox covers the
pixel, orders them by octpath ^ repeat(q) with Python integers, and composites
in float64 with numpy's exp.  The tile-list position (n_contrib) is computed
from a separately built per-tile list of (rank, voxel, quadrant) triples.
"""
import math

import numpy as np

MAX_ALPHA, MIN_ALPHA, EARLY_STOP_T = 0.99999, 0.00001, 0.0001


def rep(q):
    v = 0
    for _ in range(21):
        v = (v << 3) | q
    return v


def ray(px, py, cam):
    W, H = cam["W"], cam["H"]
    c2w = cam["c2w"].astype(np.float64)
    d = np.array([(px + 0.5 - cam["cx"]) * 2 * cam["tanfovx"] / W, (py + 0.5 - cam["cy"]) * 2 * cam["tanfovy"] / H, 1.0])
    rd = c2w[:3, :3] @ d
    rd /= np.linalg.norm(d)
    return c2w[:3, 3], rd


def quad_of(rd):
    return (int(rd[0] < 0) << 2) | (int(rd[1] < 0) << 1) | int(rd[2] < 0)


def slab(c, l, ro, rd):
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / rd
        t0 = (c - l / 2 - ro) * inv
        t1 = (c + l / 2 - ro) * inv
    lo, hi = np.minimum(t0, t1), np.maximum(t0, t1)
    return lo.max(), hi.min()


def el11(x):
    return x if x > 1.1 else math.exp(0.909090909091 * x - 0.904689820196)


def density(geo, q):
    # trilinear interpolation, corner index = (x << 2) | (y << 1) | z
    return sum(geo[i] * (q[0] if i & 4 else 1 - q[0]) * (q[1] if i & 2 else 1 - q[1]) * (q[2] if i & 1 else 1 - q[2])
               for i in range(8))


def tile_lists(cam, octpaths, n_dup, bbox, qbits):
    gx = (cam["W"] + 15) // 16
    gy = (cam["H"] + 15) // 16
    lists = {}
    for v in np.nonzero(n_dup)[0]:
        bx, by = int(bbox[v, 0]), int(bbox[v, 1])
        x0, y0, x1, y1 = bx >> 16, bx & 0xFFFF, by >> 16, by & 0xFFFF
        for q in range(8):
            if not (int(qbits[v]) >> q) & 1:
                continue
            rank = int(np.asarray(octpaths[v]).reshape(-1)[0]) ^ rep(q)
            for ty in range(min(y0 // 16, gy - 1), min(y1 // 16, gy - 1) + 1):
                for tx in range(min(x0 // 16, gx - 1), min(x1 // 16, gx - 1) + 1):
                    lists.setdefault(ty * gx + tx, []).append((rank, int(v), q))
    for k in lists:
        lists[k].sort()
    return lists


def render(cam, octpaths, n_dup, bbox, qbits, centers, lengths, geos, rgbs, bg=0.0, pixels=None):
    """float64 compositor; returns color [3,H,W], T [H,W], n_contrib [H,W], ambiguous [H,W] (a threshold
    test was within 1e-4 relative, so float32 may legitimately decide it the other way)."""
    W, H = cam["W"], cam["H"]
    gx = (W + 15) // 16
    lists = tile_lists(cam, octpaths, n_dup, bbox, qbits)
    color = np.zeros((3, H, W))
    T_img = np.zeros((H, W))
    nc = np.zeros((H, W), np.int64)
    amb = np.zeros((H, W), bool)
    pix = pixels if pixels is not None else [(x, y) for y in range(H) for x in range(W)]
    for px, py in pix:
        ro, rd = ray(px, py, cam)
        pq = quad_of(rd)
        lst = lists.get((py // 16) * gx + (px // 16), [])
        T, C, last = 1.0, np.zeros(3), 0
        for pos, (rank, v, q) in enumerate(lst):
            if q != pq:
                continue
            bx, by = int(bbox[v, 0]), int(bbox[v, 1])
            if not ((bx >> 16) <= px <= (by >> 16) and (bx & 0xFFFF) <= py <= (by & 0xFFFF)):
                continue
            c, l = centers[v].astype(np.float64), float(np.asarray(lengths[v]).reshape(-1)[0])
            a, b = slab(c, l, ro, rd)
            if a > b:
                continue
            if abs(a - b) < 1e-6 * max(1.0, abs(a)):
                amb[py, px] = True
            t = a + 0.5 * (b - a)
            qt = (ro + t * rd - (c - l / 2)) / l
            vol = 100.0 * (b - a) * el11(density(geos[v].astype(np.float64), qt))
            alpha = min(MAX_ALPHA, 1.0 - math.exp(-vol))
            if abs(alpha - MIN_ALPHA) < 1e-4 * MIN_ALPHA:
                amb[py, px] = True
            if alpha < MIN_ALPHA:
                continue
            C += T * alpha * rgbs[v].astype(np.float64)
            T *= 1.0 - alpha
            last = pos + 1
            if abs(T - EARLY_STOP_T) < 1e-4 * EARLY_STOP_T:
                amb[py, px] = True
            if T < EARLY_STOP_T:
                break
        color[:, py, px] = C + T * bg
        T_img[py, px] = T
        nc[py, px] = last
    return color, T_img, nc, amb


SH_C0 = 0.5 / math.sqrt(math.pi)
SH_C1 = math.sqrt(3.0) / (2.0 * math.sqrt(math.pi))


def sh_basis(d):
    """Real SH up to degree 3 in the 3DGS sign convention, from the textbook normalisations."""
    x, y, z = d
    c2 = [math.sqrt(15 / math.pi) / 2, -math.sqrt(15 / math.pi) / 2, math.sqrt(5 / math.pi) / 4,
          -math.sqrt(15 / math.pi) / 2, math.sqrt(15 / math.pi) / 4]
    c3 = [-math.sqrt(35 / (2 * math.pi)) / 4, math.sqrt(105 / math.pi) / 2, -math.sqrt(21 / (2 * math.pi)) / 4,
          math.sqrt(7 / math.pi) / 4, -math.sqrt(21 / (2 * math.pi)) / 4, math.sqrt(105 / math.pi) / 4,
          -math.sqrt(35 / (2 * math.pi)) / 4]
    xx, yy, zz = x * x, y * y, z * z
    return np.array([
        SH_C0, -SH_C1 * y, SH_C1 * z, -SH_C1 * x,
        c2[0] * x * y, c2[1] * y * z, c2[2] * (2 * zz - xx - yy), c2[3] * x * z, c2[4] * (xx - yy),
        c3[0] * y * (3 * xx - yy), c3[1] * x * y * z, c3[2] * y * (4 * zz - xx - yy),
        c3[3] * z * (2 * zz - 3 * xx - 3 * yy), c3[4] * x * (4 * zz - xx - yy), c3[5] * z * (xx - yy),
        c3[6] * x * (xx - 3 * yy)])


def sh_rgb(center, cam_pos, sh0, shs):
    d = np.asarray(center, np.float64) - np.asarray(cam_pos, np.float64)
    d /= np.linalg.norm(d)
    b = sh_basis(d)
    coef = np.concatenate([np.asarray(sh0, np.float64)[None], np.asarray(shs, np.float64)], 0)
    return np.maximum(b @ coef + 0.5, 0.0), b


def contracted(x):
    return x if x < 1.0 else 2.0 - 1.0 / x


def lin_normal(g):
    g = np.asarray(g, np.float64)
    nx = g[[4, 5, 6, 7]].sum() - g[[0, 1, 2, 3]].sum()
    ny = g[[2, 3, 6, 7]].sum() - g[[0, 1, 4, 5]].sum()
    nz = g[[1, 3, 5, 7]].sum() - g[[0, 2, 4, 6]].sum()
    return np.array([nx, ny, nz])


def render_ext(cam, octpaths, n_dup, bbox, qbits, centers, lengths, geos, rgbs, pixels):
    """float64, n_samp 1: per pixel the mean depth, the median depth (16-sample refinement inside the voxel
    where T first drops to 0.5), the distortion sum, the normal, and the contributing entries
    (voxel, weight pt_w, a, b); plus max_w per voxel over the given pixels."""
    W, H = cam["W"], cam["H"]
    gx = (W + 15) // 16
    lists = tile_lists(cam, octpaths, n_dup, bbox, qbits)
    out = {}
    max_w = np.zeros(len(centers))
    for px, py in pixels:
        ro, rd = ray(px, py, cam)
        d_cam = np.array([(px + 0.5 - cam["cx"]) * 2 * cam["tanfovx"] / W,
                          (py + 0.5 - cam["cy"]) * 2 * cam["tanfovy"] / H, 1.0])
        rn = np.linalg.norm(d_cam)
        pq = quad_of(rd)
        T, D, Dd, Nn = 1.0, 0.0, 0.0, np.zeros(3)
        med = None
        entries = []
        for rank, v, q in lists.get((py // 16) * gx + (px // 16), []):
            if q != pq:
                continue
            bx, by = int(bbox[v, 0]), int(bbox[v, 1])
            if not ((bx >> 16) <= px <= (by >> 16) and (bx & 0xFFFF) <= py <= (by & 0xFFFF)):
                continue
            c, l = centers[v].astype(np.float64), float(np.asarray(lengths[v]).reshape(-1)[0])
            a, b = slab(c, l, ro, rd)
            if a > b:
                continue
            t = a + 0.5 * (b - a)
            g = geos[v].astype(np.float64)
            vol = 100.0 * (b - a) * el11(density(g, (ro + t * rd - (c - l / 2)) / l))
            alpha = min(MAX_ALPHA, 1.0 - math.exp(-vol))
            if alpha < MIN_ALPHA:
                continue
            w = T * alpha
            D += T * alpha * 0.5 * (a + b)
            if T > 0.5:
                med = (v, T)
            Dd += w * 0.5 * (contracted(a) + contracted(b))
            n = lin_normal(g)
            Nn += w * n / math.sqrt(n @ n + 1e-15)
            entries.append((v, w, a, b))
            max_w[v] = max(max_w[v], w)
            T *= 1.0 - alpha
            if T < EARLY_STOP_T:
                break
        Dmed = 0.0
        if med is not None:
            v, Tm = med
            c, l = centers[v].astype(np.float64), float(np.asarray(lengths[v]).reshape(-1)[0])
            a, b = slab(c, l, ro, rd)
            st = (b - a) / 16
            Dmed = a - 0.5 * st
            g = geos[v].astype(np.float64)
            for k in range(16):
                if Tm <= 0.5:
                    break
                Dmed += st
                qt = (ro + (a + (k + 0.5) * st) * rd - (c - l / 2)) / l
                Tm *= math.exp(-100.0 * st * el11(density(g, qt)))
        out[(px, py)] = dict(depth=D / rn, median=Dmed / rn, dist=Dd, normal=Nn, entries=entries, T=T)
    return out, max_w

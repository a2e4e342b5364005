"""Known-answer vectors (tests/golden/tiny_scenes.npz, written by
tests/golden/make_golden.py from the independent brute force) checked against
the oracle on the CPU and against the HIP path on the GPU."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_scenes.npz")


def _cases():
    z = np.load(GOLDEN)
    for k in range(int(z["n_cases"])):
        p = f"c{k}_"
        c = {name[len(p):]: z[name] for name in z.files if name.startswith(p)}
        W, H, tx, ty, cx, cy, near = c["cam"]
        c["camd"] = dict(W=int(W), H=int(H), tanfovx=float(tx), tanfovy=float(ty), cx=float(cx), cy=float(cy),
                         near=float(near), w2c=c["w2c"], c2w=c["c2w"])
        yield k, c


def _check(k, c, n_dup, bbox, qbits, hi, lo, vl, color, T, nc):
    assert np.array_equal(np.asarray(n_dup, np.uint32), c["n_dup"]), k
    live = c["n_dup"] > 0
    assert np.array_equal(np.asarray(bbox, np.uint32)[live], c["bbox"][live]), k
    assert np.array_equal(np.asarray(qbits, np.uint32)[live], c["qbits"][live]), k
    assert np.array_equal(hi, c["keys_hi"]) and np.array_equal(lo, c["keys_lo"]), k
    assert np.array_equal(vl, c["vox_list"]), k
    ok = ~c["ambiguous"]
    assert np.abs(color[:, ok] - c["color"][:, ok]).max() < 2e-5, k
    assert np.abs(T.reshape(ok.shape)[ok] - c["T"][ok]).max() < 2e-5, k
    assert np.array_equal(np.asarray(nc, np.uint32).reshape(ok.shape)[ok], c["n_contrib"][ok]), k


def test_oracle_matches_golden():
    n = 0
    for k, c in _cases():
        cam = c["camd"]
        n_dup, bbox, qbits = O.preprocess(cam, c["centers"], c["lengths"])
        hi, lo, vl, ranges = O.binning(cam, c["octpath"], n_dup, bbox, qbits)
        color, T, nc, _ = O.render_fwd(cam, ranges, vl, bbox, c["centers"], c["lengths"], c["geos"], c["rgbs"])
        _check(k, c, n_dup, bbox, qbits, hi, lo, vl, color, T, nc)
        n += 1
    assert n == 3


@pytest.mark.gpu
def test_hip_matches_golden():
    import torch
    from sv_raster_amd import _C
    for k, c in _cases():
        cam = c["camd"]
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        paths, centers, lengths = t(c["octpath"]), t(c["centers"]), t(c["lengths"])
        w2c, c2w = t(c["w2c"]), t(c["c2w"])
        W, H = cam["W"], cam["H"]
        n_dup, geom = _C.rasterize_preprocess(W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w,
                                              cam["near"], paths, centers, lengths)
        R, binning, img, color, _, _, T, _ = _C.rasterize_voxels(
            1, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, 0.0, False, False, False,
            False, paths, centers, lengths, t(c["geos"]), t(c["rgbs"]), geom)
        _, _, nc = _C.unpack_ImageState(W, H, img)
        hi, lo, vals = _C.export_sort_keys(W, H, R, binning, paths)
        gnd, gbb, gqb = _C.unpack_geom_state(paths.shape[0], geom)
        torch.cuda.synchronize()
        _check(k, c, gnd.cpu().numpy().view(np.uint32), gbb.cpu().numpy().view(np.uint32),
               gqb.cpu().numpy().view(np.uint32), hi.cpu().numpy().view(np.uint64),
               lo.cpu().numpy().view(np.uint64), vals.cpu().numpy().view(np.uint32), color.cpu().numpy(),
               T.cpu().numpy(), nc.cpu().numpy().view(np.uint32))

"""SH colour and its backward on the GPU (through the C ABI) against the oracle
(sh_compute.cu restated), for index lists that exercise every staging path of
the degree-3 kernels: runs of 64 consecutive ids at aligned and misaligned
starts (16-byte span copies), runs ending at the last voxel, gaps, unsorted
ids and duplicates (row gathers), lists of most voxels (the dense pass over
marked rows, sorted and unsorted with duplicates), and all voxels (no index
list)."""
import numpy as np
import pytest

RTOL = 1e-5


def _idx_cases(N, rng):
    perm = rng.permutation(N)
    dup = np.concatenate([np.arange(0, 40), np.arange(30, 100)])
    return {
        "run_aligned": np.arange(128, 128 + 640),
        "run_misaligned": np.arange(5, 5 + 64 * 7 + 3),
        "run_to_end": np.arange(N - 200, N),
        "gaps": np.sort(rng.choice(N, N // 3, replace=False)),
        "unsorted": perm[: N // 2],
        "duplicates": dup,
        # lists covering at least half of the voxels take the dense path (listed rows marked)
        "most_sorted": np.sort(rng.choice(N, (4 * N) // 5, replace=False)),
        "most_unsorted_duplicates": np.concatenate([perm[: (3 * N) // 4], perm[:50]]),
        "all": None,
    }


@pytest.mark.gpu
@pytest.mark.parametrize("deg", [0, 1, 3])
def test_sh_forward_backward_match_oracle(deg):
    import torch
    from oracle import oracle as O
    from sv_raster_amd import _C
    rng = np.random.default_rng(11 + deg)
    N = 3001
    centers = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    sh0 = rng.uniform(0, 1, (N, 3)).astype(np.float32)
    shs = rng.normal(0, 0.05, (N, 15, 3)).astype(np.float32)
    cam_pos = np.array([0.3, -2.5, 1.7], np.float32)
    dL = rng.normal(0, 1, (N, 3)).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    for name, idx in _idx_cases(N, rng).items():
        print(f"sh deg {deg} idx {name}", flush=True)
        ti = t(idx.astype(np.int64)) if idx is not None else None
        rgbs = _C.sh_compute(deg, ti, t(centers), t(cam_pos), t(sh0), t(shs))
        ref = O.sh_fwd(deg, idx, centers, cam_pos, sh0, shs)
        got = rgbs.cpu().numpy()
        assert np.abs(got - ref).max() <= RTOL * max(1.0, np.abs(ref).max()), name
        d0, ds = _C.sh_compute_bw(deg, 16, ti, t(centers), t(cam_pos), rgbs, t(dL))
        r0, rs = O.sh_bwd(deg, 16, idx, centers, cam_pos, dL)
        assert np.abs(d0.cpu().numpy() - r0).max() <= RTOL * max(1.0, np.abs(r0).max()), name
        assert np.abs(ds.cpu().numpy() - rs.reshape(ds.shape)).max() <= RTOL * max(1.0, np.abs(rs).max()), name

"""Multi-view batch: one pass per stage over V cameras of one octree.

No reference counterpart as a whole.  A reference training loop calls
`renderer.rasterize_voxels` once per camera (backends/new_cuda/
new_svraster_cuda/renderer.py:39-127) and its `vox_fn` evaluates SH colours
per camera (renderer.py:251-316).  `rasterize_voxels_batch` gives, for every
view, the image and transmittance that `renderer.rasterize_voxels` gives for
that camera with SH colours of `sh0`/`shs`, and back-propagates the sum over
views of the per-view gradients into the shared `geos`, `sh0`, `shs` and
`subdiv_p`, with all views handled by one launch per stage:

* preprocess, key duplication and one sort over all (view, voxel) pairs;
* SH colours of all views from one read of the coefficients;
* the compositor forward and backward as persistent launches over one queue
  of all views' tiles (longest lists first);
* per-voxel gradient sums over views in the order a per-view loop would
  accumulate them.

The camera set is fixed in a `ViewBatch` (host and device copies of the
`svr_view` table of include/svraster_hip.h).
"""
import ctypes

import numpy as np
import torch

from . import _C, _lib

_L = _lib.load()


# binning tiles of a batch (svr_view.tile_shape): name -> (code, width, height)
TILE_SHAPES = {"16x16": (0, 16, 16), "8x8": (1, 8, 8), "32x8": (2, 32, 8)}
# backward depth of a batch (svr_view.bwd_window, bwd_sync): LDS window in 64-entry chunks, and the number of a
# wave's pixels from which an entry's gradient terms are summed across the wave at once
BWD_WINDOWS = (2, 4, 6, 8)
BWD_SYNCS = (4, 8, 16)
BWD_DEFAULT = (4, 8)


def _mat12(m):
    a = m.detach().float().cpu().numpy() if torch.is_tensor(m) else np.asarray(m, np.float32)
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    if a.size < 12:
        raise Exception("camera matrices need at least 12 values (3 rows of a 4x4)")
    return a[:12]


class ViewBatch:
    """A fixed set of cameras rendered together.

    `settings` is a sequence of `renderer.RasterSettings` (color_mode "sh";
    the same n_samp_per_vox for all views; depth, normal and max_w tracking
    off).  `tile_shape` is the binning tile ("16x16", the reference's, or
    "8x8" / "32x8"; images and gradients do not depend on it, see `autotune`).
    Reading the camera matrices synchronises once, here.
    """

    def __init__(self, settings, device=None, tile_shape="16x16", bwd_window=BWD_DEFAULT[0], bwd_sync=BWD_DEFAULT[1]):
        if tile_shape not in TILE_SHAPES:
            raise Exception(f"tile_shape must be one of {sorted(TILE_SHAPES)}")
        if bwd_window not in BWD_WINDOWS or bwd_sync not in BWD_SYNCS:
            raise Exception(f"bwd_window must be one of {BWD_WINDOWS} and bwd_sync one of {BWD_SYNCS}")
        self.bwd_window, self.bwd_sync = int(bwd_window), int(bwd_sync)
        code, tw, th = TILE_SHAPES[tile_shape]
        self.tile_shape = tile_shape
        self.settings = list(settings)
        settings = list(settings)
        if not settings:
            raise Exception("ViewBatch needs at least one view")
        self.n_samp = int(settings[0].n_samp_per_vox)
        if not 1 <= self.n_samp <= _C.MAX_N_SAMP:
            raise Exception(f"n_samp_per_vox should be in range [1, {_C.MAX_N_SAMP}].")
        V = len(settings)
        table = (_lib.SvrView * V)()
        self.sizes = []
        tile_base, pix_base = 0, 0
        for u, s in enumerate(settings):
            if int(s.n_samp_per_vox) != self.n_samp:
                raise Exception("all views of a batch need the same n_samp_per_vox")
            opts = [n for n in ("need_depth", "need_normal", "track_max_w", "lambda_R_concen", "lambda_ascending",
                                "lambda_dist") if getattr(s, n, 0)]
            if opts:
                raise NotImplementedError(f"the multi-view batch renders colour and T only (got {', '.join(opts)}); "
                                          "use renderer.rasterize_voxels per view for these outputs")
            W, H = int(s.image_width), int(s.image_height)
            v = table[u]
            v.width, v.height = W, H
            v.tanfovx, v.tanfovy = float(s.tanfovx), float(s.tanfovy)
            v.cx, v.cy = float(s.cx), float(s.cy)
            v.near_plane = float(getattr(s, "near", 0.02))
            v.bg_color = float(getattr(s, "bg_color", 0.0))
            v.w2c[:] = _mat12(s.w2c_matrix).tolist()
            v.c2w[:] = _mat12(s.c2w_matrix).tolist()
            v.tile_base, v.tile_shape, v.pix_base = tile_base, code, pix_base
            v.bwd_window, v.bwd_sync = self.bwd_window, self.bwd_sync
            self.sizes.append((W, H))
            tile_base += ((W + tw - 1) // tw) * ((H + th - 1) // th)
            pix_base += W * H
        if device is None:
            m = settings[0].c2w_matrix
            device = m.device if torch.is_tensor(m) and m.is_cuda else torch.device("cuda")
        self.device = torch.device(device)
        self.host = table
        raw = np.frombuffer(bytes(memoryview(table)), dtype=np.uint8).copy()
        self.dev = torch.from_numpy(raw).to(self.device)
        self.n_views = V
        self.total_tiles = tile_base
        self.total_pixels = pix_base
        self.pix_base = [table[u].pix_base for u in range(V)]

    def host_ptr(self):
        return ctypes.cast(self.host, ctypes.c_void_p)

    def set_backward(self, bwd_window, bwd_sync):
        """Selects the backward window depth and sync threshold (read on the host at each backward call; images
        and gradients do not depend on them beyond float summation order)."""
        if bwd_window not in BWD_WINDOWS or bwd_sync not in BWD_SYNCS:
            raise Exception(f"bwd_window must be one of {BWD_WINDOWS} and bwd_sync one of {BWD_SYNCS}")
        self.bwd_window, self.bwd_sync = int(bwd_window), int(bwd_sync)
        for u in range(self.n_views):
            self.host[u].bwd_window, self.host[u].bwd_sync = self.bwd_window, self.bwd_sync
        raw = np.frombuffer(bytes(memoryview(self.host)), dtype=np.uint8).copy()
        self.dev.copy_(torch.from_numpy(raw).to(self.device))
        return self

    def split(self, color_flat, T_flat):
        """Per-view [3, H, W] colour and [1, H, W] transmittance views of the flat buffers."""
        colors, Ts = [], []
        for (W, H), pb in zip(self.sizes, self.pix_base):
            colors.append(color_flat.narrow(0, 3 * pb, 3 * W * H).view(3, H, W))
            Ts.append(T_flat.narrow(0, pb, W * H).view(1, H, W))
        return colors, Ts


def _state(views, vox_centers, vox_lengths, sh0, shs, active_sh_degree, capacity=None):
    """Preprocess of all pairs, R, and the SH colours of all views.  With a capacity the entry count stays
    on the device (no synchronisation) and R is the capacity."""
    dev = vox_centers.device
    V, N = views.n_views, vox_centers.size(0)
    st = _C._stream(dev)
    centers, lengths = _C._f32(vox_centers), _C._f32(vox_lengths)
    n_dup = torch.empty(V * N, dtype=torch.int32, device=dev)
    geom = torch.empty(int(_L.svr_batch_geom_state_size(V, N)), dtype=torch.uint8, device=dev)
    R = ctypes.c_int64(0 if capacity is None else int(capacity))
    if N > 0:
        _lib.check(_L.svr_batch_preprocess(V, views.host_ptr(), _C._ptr(views.dev), N, _C._ptr(centers),
                                           _C._ptr(lengths), _C._ptr(n_dup), _C._ptr(geom), st),
                   "batch_preprocess")
        if capacity is None:
            _lib.check(_L.svr_num_rendered(V * N, _C._ptr(geom), st, ctypes.byref(R)), "num_rendered")
        else:
            _lib.check(_L.svr_scan_entries(V * N, _C._ptr(geom), st), "scan_entries")
    rgbs = torch.empty((V * N, 3), dtype=torch.float32, device=dev)
    if N > 0:
        _lib.check(_L.svr_batch_sh_compute(int(active_sh_degree), V, _C._ptr(views.dev), N, _C._ptr(centers),
                                           _C._ptr(geom), _C._ptr(_C._f32(sh0)), _C._ptr(_C._f32(shs)),
                                           _C._ptr(rgbs), st), "batch_sh_compute")
    return n_dup, geom, int(R.value), rgbs


def batch_forward(views, active_sh_degree, octree_paths, vox_centers, vox_lengths, geos, sh0, shs, capacity=None,
                  overflow=None, table=None):
    """Forward of all views; returns the state the backward needs (also used by the tests):
    color [3 * P_total], T [P_total], R, rgbs [V * N, 3], n_dup [V * N] and the opaque buffers.

    capacity / overflow (an int32 [2] device tensor): the capture-safe form, with buffers sized for
    `capacity` entries and no host synchronisation; a frame over the capacity sets overflow[0] (sticky) and
    overflow[1] (this frame) to 1 and renders no entries.  table: the octree's rank-order table
    (_C.build_order_table); by default the cached one of octree_paths."""
    dev = vox_centers.device
    V, N = views.n_views, vox_centers.size(0)
    if (capacity is None) != (overflow is None):
        raise Exception("capacity and overflow go together")
    n_dup, geom, R, rgbs = _state(views, vox_centers, vox_lengths, sh0, shs, active_sh_degree, capacity)
    binning = torch.empty(int(_L.svr_binning_state_size(R)) if R > 0 else 0, dtype=torch.uint8, device=dev)
    img = torch.empty(int(_L.svr_batch_image_state_size(views.total_tiles, views.total_pixels, R)), dtype=torch.uint8,
                      device=dev)
    color = torch.empty(3 * views.total_pixels, dtype=torch.float32, device=dev)
    T = torch.empty(views.total_pixels, dtype=torch.float32, device=dev)
    if table is None and N > 0:
        table = _C.order_table_cache.get(octree_paths)
    if overflow is not None and overflow.numel() < 2:
        raise Exception("overflow needs 2 int32 words")
    centers, lengths, geos_c = _C._f32(vox_centers), _C._f32(vox_lengths), _C._f32_a16(geos)
    if capacity is None:
        _lib.check(_L.svr_batch_rasterize_forward(views.n_samp, V, views.host_ptr(), _C._ptr(views.dev), N,
                                                  _C._ptr(centers), _C._ptr(lengths), _C._ptr(geos_c),
                                                  _C._ptr(rgbs), _C._ptr(geom), _C._ptr(table), R,
                                                  _C._ptr(binning), _C._ptr(img), _C._ptr(color), _C._ptr(T),
                                                  _C._stream(dev)), "batch_rasterize_forward")
    else:
        _lib.check(_L.svr_batch_rasterize_forward_capacity(views.n_samp, V, views.host_ptr(), _C._ptr(views.dev),
                                                           N, _C._ptr(centers), _C._ptr(lengths), _C._ptr(geos_c),
                                                           _C._ptr(rgbs), _C._ptr(geom), _C._ptr(table), R,
                                                           _C._ptr(binning), _C._ptr(img), _C._ptr(color),
                                                           _C._ptr(T), _C._ptr(overflow), _C._stream(dev)),
                   "batch_rasterize_forward_capacity")
    return dict(color=color, T=T, R=R, rgbs=rgbs, n_dup=n_dup, geom=geom, binning=binning, img=img,
                centers=centers, lengths=lengths, geos=geos_c, deg=int(active_sh_degree),
                capacity_mode=capacity is not None)


def batch_backward(views, state, dL_dcolor, dL_dT, out=None):
    """Gradients (dgeos [N, 8], dsh0 [N, 3], dshs [N, 15, 3], dsub [N, 1], drgbs [V * N, 3]) of a batch forward.
    `out` may hold preallocated (dgeos, dsh0, dshs, dsub) tensors to write into."""
    centers, lengths, T = state["centers"], state["lengths"], state["T"]
    R = state["R"]
    dev = centers.device
    V, N = views.n_views, centers.size(0)
    st = _C._stream(dev)
    dL_dcolor = torch.zeros(3 * T.numel(), dtype=torch.float32, device=dev) if dL_dcolor is None \
        else _C._f32(dL_dcolor)
    dL_dT = torch.zeros_like(T) if dL_dT is None else _C._f32(dL_dT)
    ws = torch.empty(int(_L.svr_batch_backward_workspace_size(V, N, R)), dtype=torch.uint8, device=dev)
    if out is None:
        dgeos = torch.empty((N, 8), dtype=torch.float32, device=dev)
        dsh0 = torch.empty((N, 3), dtype=torch.float32, device=dev)
        dshs = torch.empty((N, 15, 3), dtype=torch.float32, device=dev)
        dsub = torch.empty((N, 1), dtype=torch.float32, device=dev)
    else:
        dgeos, dsh0, dshs, dsub = out
    drgbs = torch.empty((V * N, 3), dtype=torch.float32, device=dev)
    backward = _L.svr_batch_rasterize_backward_capacity if state.get("capacity_mode") else \
        _L.svr_batch_rasterize_backward
    if N > 0:
        _lib.check(backward(views.n_samp, V, views.host_ptr(), _C._ptr(views.dev), N, _C._ptr(centers),
                            _C._ptr(lengths), _C._ptr(state["geos"]), _C._ptr(state["rgbs"]), _C._ptr(state["geom"]),
                            R, _C._ptr(state["binning"]), _C._ptr(state["img"]), _C._ptr(T), _C._ptr(dL_dcolor),
                            _C._ptr(dL_dT), _C._ptr(ws), _C._ptr(dgeos), _C._ptr(drgbs), _C._ptr(dsub), st),
                   "batch_rasterize_backward")
        _lib.check(_L.svr_batch_sh_compute_bw(state["deg"], V, _C._ptr(views.dev), N, _C._ptr(centers),
                                              _C._ptr(state["geom"]), _C._ptr(drgbs), _C._ptr(dsh0), _C._ptr(dshs),
                                              st), "batch_sh_compute_bw")
    return dgeos, dsh0, dshs, dsub, drgbs


class _RasterizeBatch(torch.autograd.Function):
    """Outputs: the V per-view colour images, then the V per-view T images (views of the flat buffers).  The
    backward gathers their gradients into the flat layout with one concatenation: had the function returned
    the flat buffers and left the per-view split to autograd, every view's gradient would be widened to a
    full-size zero tensor and accumulated, O(V^2) traffic over the batch's pixels."""

    @staticmethod
    def forward(ctx, views, active_sh_degree, octree_paths, vox_centers, vox_lengths, geos, sh0, shs, subdiv_p):
        state = batch_forward(views, active_sh_degree, octree_paths, vox_centers, vox_lengths, geos, sh0, shs)
        ctx.views, ctx.state = views, state
        colors, Ts = views.split(state["color"], state["T"])
        return (*colors, *Ts)

    @staticmethod
    def backward(ctx, *grads):
        views, state = ctx.views, ctx.state
        V = views.n_views
        dev = state["color"].device

        def flat(gs, chans):
            if all(g is None for g in gs):
                return None
            return torch.cat([(g.reshape(-1) if g is not None else
                               torch.zeros(chans * W * H, dtype=torch.float32, device=dev))
                              for g, (W, H) in zip(gs, views.sizes)])

        dL_dcolor, dL_dT = flat(grads[:V], 3), flat(grads[V:], 1)
        dgeos, dsh0, dshs, dsub, _ = batch_backward(views, state, dL_dcolor, dL_dT)
        ctx.state = None
        return None, None, None, None, None, dgeos, dsh0, dshs, dsub


def rasterize_voxels_batch(views, octree_paths, vox_centers, vox_lengths, geos, sh0, shs, subdiv_p=None,
                           active_sh_degree=3):
    """Render every camera of `views` (a ViewBatch).

    geos [N, 8], sh0 [N, 3] and shs [N, 15, 3] are shared by all views.
    Returns (colors, Ts): lists with one [3, H, W] and one [1, H, W] tensor per
    view.  Gradients flow to geos, sh0, shs and subdiv_p ([N, 1]; its gradient
    is the summed subdivision priority of the backward).
    """
    N = octree_paths.shape[0]
    if vox_centers.shape[0] != N or vox_lengths.shape[0] != N or geos.shape[0] != N or sh0.shape[0] != N \
            or shs.shape[0] != N:
        raise Exception("Size mismatched.")
    if len(vox_centers.shape) != 2 or vox_centers.shape[1] != 3:
        raise Exception("Expect vox_centers in shape [N, 3].")
    if len(geos.shape) != 2 or geos.shape[1] != 8:
        raise Exception("Expect geos in shape [N, 8].")
    if tuple(shs.shape[1:]) != (15, 3) or tuple(sh0.shape[1:]) != (3,):
        raise Exception("Expect sh0 in shape [N, 3] and shs in shape [N, 15, 3].")
    if not 0 <= int(active_sh_degree) <= 3:
        raise Exception("active_sh_degree should be in range [0, 3].")
    _C._require_cuda(octree_paths, vox_centers, vox_lengths, geos, sh0, shs)
    if subdiv_p is None:
        subdiv_p = torch.zeros((N, 1), dtype=torch.float32, device=vox_centers.device)
    out = _RasterizeBatch.apply(views, int(active_sh_degree), octree_paths, vox_centers, vox_lengths, geos, sh0, shs,
                                subdiv_p)
    return list(out[:views.n_views]), list(out[views.n_views:])


def unpack_image_state(views, imageBuffer):
    """ranges [T_total, 2], tile_last [T_total], n_contrib [P_total] (int32) of a batch image state."""
    dev = imageBuffer.device
    T, P = views.total_tiles, views.total_pixels
    ranges = torch.empty((T, 2), dtype=torch.int32, device=dev)
    tile_last = torch.empty(T, dtype=torch.int32, device=dev)
    n_contrib = torch.empty(P, dtype=torch.int32, device=dev)
    _lib.check(_L.svr_batch_unpack_image_state(T, P, _C._ptr(imageBuffer), _C._ptr(ranges), _C._ptr(tile_last),
                                               _C._ptr(n_contrib), _C._stream(dev)), "batch_unpack_image_state")
    return ranges, tile_last, n_contrib


class GraphTrainStep:
    """A batch training step replayed from a captured HIP graph (no reference counterpart).

    One replay = forward of all views of `views`, the gradient of the summed per-view MSE to `targets`
    (one [3, H, W] tensor per view; dL/dcolor = 2 (color - target) / (3 H W) per view, the backward of
    torch's mse_loss), the batch backward and `optimizer.step()` (a SparseAdam switched to device step
    counts, whose bias terms equal the host-step ones bit for bit).  Nothing in a replay synchronises with the
    host: the list buffers are sized for `capacity` entries (by default 1.25 x the count of the current
    parameters plus a margin).  A frame over the capacity renders no entries and its Adam update is skipped
    (parameters, moments and step counts unchanged, with sparse or dense Adam); `overflowed()` reads the
    sticky flag and `recapture()` re-sizes.  Parameters are updated in place; `color` holds the last replay's
    images (flat, per-view blocks as in ViewBatch.split).

    The graph reads the tensors it was captured with by address, so the trainer holds them: the octree's
    rank-order table, the gradient tensors it installed as geos.grad / sh0.grad / shs.grad (the optimizer
    reads them), the targets and the views.  Replacing or freeing those gradients (for example
    zero_grad(set_to_none=True)) does not reach the graph, which keeps using the held tensors.
    """

    def __init__(self, views, octree_paths, vox_centers, vox_lengths, geos, sh0, shs, optimizer, targets,
                 active_sh_degree=3, capacity=None, headroom=1.25):
        self.views, self.octree_paths = views, octree_paths
        self.centers, self.lengths = _C._f32(vox_centers), _C._f32(vox_lengths)
        self.geos, self.sh0, self.shs = geos, sh0, shs
        self.optimizer, self.deg, self.headroom = optimizer, int(active_sh_degree), headroom
        dev = self.centers.device
        if len(targets) != views.n_views:
            raise Exception("one target image per view")
        self.target = torch.cat([_C._f32(t).reshape(-1) for t in targets])
        self.weight = torch.cat([torch.full((3 * W * H,), 2.0 / (3 * W * H), dtype=torch.float32, device=dev)
                                 for W, H in views.sizes])
        if self.target.numel() != 3 * views.total_pixels:
            raise Exception("target sizes do not match the views")
        self.zero_dT = torch.zeros(views.total_pixels, dtype=torch.float32, device=dev)
        # [0]: sticky overflow flag, [1]: this frame's (the Adam update's skip flag)
        self.overflow = torch.zeros(2, dtype=torch.int32, device=dev)
        for p in (geos, sh0, shs):
            p.grad = torch.zeros_like(p)
        self.grads = (geos.grad, sh0.grad, shs.grad)
        self.dsub = torch.zeros((self.centers.size(0), 1), dtype=torch.float32, device=dev)
        optimizer.use_device_steps(skip=self.overflow[1:2])
        self.table = _C.order_table_cache.get(octree_paths) if self.centers.size(0) > 0 else None
        self.recapture(capacity)

    def _measure(self):
        with torch.no_grad():
            st = batch_forward(self.views, self.deg, self.octree_paths, self.centers, self.lengths, self.geos,
                               self.sh0, self.shs, table=self.table)
        return st["R"]

    def _fwd_bwd(self):
        st = batch_forward(self.views, self.deg, self.octree_paths, self.centers, self.lengths, self.geos, self.sh0,
                           self.shs, capacity=self.capacity, overflow=self.overflow, table=self.table)
        dcol = (st["color"] - self.target) * self.weight
        batch_backward(self.views, st, dcol, self.zero_dT, out=(self.grads[0], self.grads[1], self.grads[2], self.dsub))
        self.color = st["color"]

    def recapture(self, capacity=None):
        """(Re)size the entry buffers (measuring the current count when capacity is None) and capture."""
        if capacity is None:
            capacity = int(self._measure() * self.headroom) + 65536
        self.capacity = int(capacity)
        self.overflow.zero_()
        with torch.no_grad():
            # the captured optimizer step reads the held gradient tensors
            for p, g in zip((self.geos, self.sh0, self.shs), self.grads):
                p.grad = g
            self._fwd_bwd()  # outside the graph: allocator and first-use warm-up
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._fwd_bwd()
                self.optimizer.step()
        torch.cuda.synchronize()

    def step(self):
        self.graph.replay()
        self.optimizer.note_device_steps(1)

    def overflowed(self):
        """True if a replay since the last capture exceeded the capacity (synchronises)."""
        return bool(int(self.overflow[0].item()))


_TUNED = {}


def _spread(ts):
    ts = sorted(ts)
    return {"median_ms": ts[len(ts) // 2], "min_ms": ts[0], "max_ms": ts[-1], "runs": len(ts)}


def _separated(times):
    """True when the fastest candidate's slowest run beats every other candidate's fastest run."""
    med = {k: sorted(v)[len(v) // 2] for k, v in times.items()}
    best = min(med, key=med.get)
    return all(max(times[best]) < min(v) for k, v in times.items() if k != best), best


def _tune_rounds(cands, run, min_rounds, max_rounds, setup=None):
    """Times run(c) for every candidate in blocks of consecutive runs (setup(c), one untimed warm-up, then
    min_rounds timed runs), then further blocks for the candidates whose fastest run is not beaten by the
    leader's slowest, until the fastest candidate separates from the rest or max_rounds runs are reached.
    Consecutive blocks keep each candidate's buffers in the caching allocator between its runs (a batch can
    hold a large part of the device memory, and interleaving candidates of different sizes would reallocate
    it between runs)."""
    times = {c: [] for c in cands}

    def block(c, n):
        if setup is not None:
            setup(c)
        run(c)
        for _ in range(n):
            times[c].append(run(c))

    for c in cands:
        block(c, min_rounds)
    while True:
        sep, best = _separated(times)
        if sep:
            break
        contenders = [c for c in cands if c != best and min(times[c]) <= max(times[best])]
        todo = [c for c in [best] + contenders if len(times[c]) < max_rounds]
        if not todo:
            break
        for c in todo:
            block(c, min(min_rounds, max_rounds - len(times[c])))
    sep, best = _separated(times)
    return best, sep, times


def autotune(settings, octree_paths, vox_centers, vox_lengths, geos, sh0, shs, candidates=tuple(TILE_SHAPES),
             windows=BWD_WINDOWS, syncs=BWD_SYNCS, min_rounds=3, max_rounds=10, active_sh_degree=3):
    """Picks the binning tile and the backward depth (LDS window, sync threshold) of a multi-view workload by
    timing them on the actual scene and cameras (the first call for a workload; later calls with the same voxel
    count, image sizes and n_samp return the cached choice).  Returns (ViewBatch with the choice, report).

    Stage 1 times a forward + backward per tile shape (default backward depth); stage 2 times the backward
    alone for every (window, sync) pair on the chosen tile, all from one forward state.  Each stage times its
    candidates in blocks of consecutive runs and repeats the blocks of the candidates still in contention until
    the fastest one's slowest run beats every other candidate's fastest run, or max_rounds runs; the report
    gives every candidate's median, min and max and whether that separation was reached.

    The tile shape trades sort size against list relevance: small tiles give more list entries (more keys to
    sort) but each wave walks only the entries of its own 8x8 pixels; large tiles sort fewer entries but every
    wave tests the entries of its neighbours' pixels too.  The window sets how many 64-entry chunks the
    backward stages in LDS at a time (deeper windows share each staging barrier among more entries, shallower
    ones leave LDS for more workgroups per CU); the sync threshold sets which entries have their gradient terms
    summed across the wave at once instead of per pixel."""
    settings = list(settings)
    key = (int(vox_centers.shape[0]), tuple((int(s.image_width), int(s.image_height)) for s in settings),
           int(settings[0].n_samp_per_vox), tuple(candidates), tuple(windows), tuple(syncs))
    if key in _TUNED:
        (best, kw, sync), report = _TUNED[key]
        return ViewBatch(settings, tile_shape=best, bwd_window=kw, bwd_sync=sync), report
    import time

    def fwd_bwd(name):
        views = ViewBatch(settings, tile_shape=name)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = batch_forward(views, active_sh_degree, octree_paths, vox_centers, vox_lengths, geos, sh0, shs)
        batch_backward(views, st, st["color"] - 0.5, None)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    with torch.no_grad():
        best, sep1, t1 = _tune_rounds(list(candidates), fwd_bwd, min_rounds, max_rounds,
                                      setup=lambda c: torch.cuda.empty_cache())
        torch.cuda.empty_cache()
        views = ViewBatch(settings, tile_shape=best)
        st = batch_forward(views, active_sh_degree, octree_paths, vox_centers, vox_lengths, geos, sh0, shs)
        dcol = st["color"] - 0.5
        if views.n_samp != 1:  # n_samp 2 and 3 always run the default backward depth
            pairs = [BWD_DEFAULT]
        else:
            pairs = [(w, s) for w in windows for s in syncs]

        def bwd(pair):
            views.set_backward(*pair)
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            start.record()
            batch_backward(views, st, dcol, None)
            end.record()
            end.synchronize()
            return start.elapsed_time(end)

        (kw, sync), sep2, t2 = _tune_rounds(pairs, bwd, min_rounds, max_rounds)
        del st, dcol
    report = {"tile_shape": best, "bwd_window": kw, "bwd_sync": sync,
              "fwd_bwd_ms": {k: _spread(v) for k, v in t1.items()}, "tile_separated": sep1,
              "bwd_ms": {f"{w}x{s}": _spread(v) for (w, s), v in t2.items()}, "bwd_separated": sep2}
    _TUNED[key] = ((best, kw, sync), report)
    return ViewBatch(settings, tile_shape=best, bwd_window=kw, bwd_sync=sync), report



# ---------------------------------------------------------------------------
# Per-voxel statistics over many cameras (the loops of SRC/sparse_voxel_gears/
# adaptive.py:190-221 and NC renderer.py:408-438, as batch passes)

def _pix_size(views):
    """2 * tanfovx / width per view (cameras.py:54-55), in float64 on the host, then float32 on the device."""
    ps = np.array([2.0 * float(s.tanfovx) / int(s.image_width) for s in views.settings], np.float64)
    return torch.from_numpy(ps.astype(np.float32)).to(views.device)


def _pair_preprocess(views, vox_centers, vox_lengths):
    dev = vox_centers.device
    V, N = views.n_views, vox_centers.size(0)
    centers, lengths = _C._f32(vox_centers), _C._f32(vox_lengths)
    n_dup = torch.empty(V * N, dtype=torch.int32, device=dev)
    geom = torch.empty(int(_L.svr_batch_geom_state_size(V, N)), dtype=torch.uint8, device=dev)
    _lib.check(_L.svr_batch_preprocess(V, views.host_ptr(), _C._ptr(views.dev), N, _C._ptr(centers),
                                       _C._ptr(lengths), _C._ptr(n_dup), _C._ptr(geom), _C._stream(dev)),
               "batch_preprocess")
    return centers, lengths, n_dup, geom


@torch.no_grad()
def compute_training_stat(views, octree_paths, vox_centers, vox_lengths, geos):
    """`SparseVoxelModel.compute_training_stat` (adaptive.py:190-221) over the cameras of a ViewBatch in one
    pass: per voxel, max_w (largest blending weight over the views), min_samp_interval (minimum over the
    views where the voxel has a positive weight of its camera-axis distance times the pixel size; 1e30 if
    none) and view_cnt (number of such views).  Each [N, 1] float32; the same values as rendering every camera
    with track_max_w and folding the results in camera order."""
    dev = vox_centers.device
    V, N = views.n_views, vox_centers.size(0)
    max_w = torch.zeros((N, 1), dtype=torch.float32, device=dev)
    msi = torch.full((N, 1), 1e30, dtype=torch.float32, device=dev)
    cnt = torch.zeros((N, 1), dtype=torch.float32, device=dev)
    if N == 0:
        return {"max_w": max_w, "min_samp_interval": msi, "view_cnt": cnt}
    st = _C._stream(dev)
    centers, lengths, _, geom = _pair_preprocess(views, vox_centers, vox_lengths)
    R = ctypes.c_int64(0)
    _lib.check(_L.svr_num_rendered(V * N, _C._ptr(geom), st, ctypes.byref(R)), "num_rendered")
    R = int(R.value)
    binning = torch.empty(int(_L.svr_binning_state_size(R)) if R > 0 else 0, dtype=torch.uint8, device=dev)
    img = torch.empty(int(_L.svr_batch_image_state_size(views.total_tiles, views.total_pixels, R)), dtype=torch.uint8,
                      device=dev)
    T = torch.empty(views.total_pixels, dtype=torch.float32, device=dev)
    pair_w = torch.empty(V * N, dtype=torch.float32, device=dev)
    table = _C.order_table_cache.get(octree_paths)
    geos_c = _C._f32_a16(geos)
    _lib.check(_L.svr_batch_rasterize_max_w(views.n_samp, V, views.host_ptr(), _C._ptr(views.dev), N,
                                            _C._ptr(centers), _C._ptr(lengths), _C._ptr(geos_c), _C._ptr(geom),
                                            _C._ptr(table), R, _C._ptr(binning), _C._ptr(img), _C._ptr(T),
                                            _C._ptr(pair_w), st), "batch_rasterize_max_w")
    pix = _pix_size(views)
    _lib.check(_L.svr_batch_training_stat(V, _C._ptr(views.dev), N, _C._ptr(centers), _C._ptr(pix), _C._ptr(pair_w),
                                          _C._ptr(max_w), _C._ptr(msi), _C._ptr(cnt), st), "batch_training_stat")
    return {"max_w": max_w, "min_samp_interval": msi, "view_cnt": cnt}


def _camera_views(cameras, near):
    from .renderer import RasterSettings
    settings = [RasterSettings(color_mode="sh", n_samp_per_vox=1, image_width=cam.image_width,
                               image_height=cam.image_height, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy, cx=cam.cx,
                               cy=cam.cy, w2c_matrix=cam.w2c, c2w_matrix=cam.c2w, near=near) for cam in cameras]
    return ViewBatch(settings)


def _mark(cameras, octree_paths, vox_centers, vox_lengths, near, rate):
    dev = vox_centers.device
    N = vox_centers.size(0)
    out = torch.zeros(N, dtype=torch.float32 if rate else torch.bool, device=dev)
    cameras = list(cameras)
    if N == 0 or not cameras:
        return out
    views = _camera_views(cameras, near)
    centers, lengths, n_dup, _ = _pair_preprocess(views, vox_centers, vox_lengths)
    # the cameras' own pixel size where they carry one (cameras.py:54-55), rounded to float once
    ps = [float(getattr(c, "pix_size", 2.0 * float(c.tanfovx) / int(c.image_width))) for c in cameras]
    pix = torch.tensor(ps, dtype=torch.float32).to(dev)
    _lib.check(_L.svr_batch_mark_samp(views.n_views, _C._ptr(views.dev), N, _C._ptr(centers), _C._ptr(lengths),
                                      _C._ptr(pix), _C._ptr(n_dup), _C._ptr(out) if rate else None,
                                      None if rate else _C._ptr(out), _C._stream(dev)), "batch_mark_samp")
    return out


@torch.no_grad()
def mark_max_samp_rate(cameras, octree_paths, vox_centers, vox_lengths, near=0.02):
    """`renderer.mark_max_samp_rate` (NC renderer.py:408-423) for all cameras in one pass: per voxel the
    largest length / (camera-axis distance x pixel size) over the cameras that see it beyond `near` ([N])."""
    return _mark(cameras, octree_paths, vox_centers, vox_lengths, near, True)


@torch.no_grad()
def mark_near(cameras, octree_paths, vox_centers, vox_lengths, near=0.2):
    """`renderer.mark_near` (NC renderer.py:426-438) for all cameras in one pass: True for voxels some camera
    sees within `near` + the voxel's length along its axis ([N] bool)."""
    return _mark(cameras, octree_paths, vox_centers, vox_lengths, near, False)

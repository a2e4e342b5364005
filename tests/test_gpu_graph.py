"""Capture-safe batch training step (GraphTrainStep, hipGraph replay) against
the eager batch step: after every replay, with eager work and host
synchronisation in between, the images match the eager run within 1e-5 and
the parameters within 1e-4 (relative, with an absolute floor of 1e-3 x the
largest parameter: the per-entry gradient rows are summed with float atomics
in a run-dependent order, and Adam's normalisation with eps 1e-15 magnifies
that rounding noise for gradients that nearly cancel); a frame over the entry
capacity is flagged and changes nothing."""
import numpy as np
import pytest

from tests._harness import rel_err
from tests.test_gpu_batch import _settings, _views


def _setup(seed=4):
    import torch
    from sv_raster_amd import synthetic as S
    cases = _views("C", [((64, 64), None), ((48, 40), None), ((80, 56), None)], seed=seed)
    sc = cases[0]["scene"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    geom = (t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"]))
    targets = [t(S.make_target(c["cam"]["H"], c["cam"]["W"], seed + u)) for u, c in enumerate(cases)]
    return cases, sc, geom, targets, t


def _params(sc, t):
    from sv_raster_amd import sparse_adam
    p = [t(sc["geos"]).requires_grad_(), t(sc["sh0"]).requires_grad_(), t(sc["shs"]).requires_grad_()]
    opt = sparse_adam.SparseAdam([{"params": [p[0]], "lr": 0.025}, {"params": [p[1]], "lr": 0.01},
                                  {"params": [p[2]], "lr": 0.01}], betas=(0.1, 0.99), eps=1e-15, sparse=True)
    return p, opt


@pytest.mark.gpu
def test_graph_step_matches_eager_steps():
    import torch
    from sv_raster_amd import batch
    cases, sc, (octpath, centers, lengths), targets, t = _setup()
    views = batch.ViewBatch([_settings(c["cam"]) for c in cases])

    pe, opt_e = _params(sc, t)
    pg, opt_g = _params(sc, t)
    trainer = batch.GraphTrainStep(views, octpath, centers, lengths, pg[0], pg[1], pg[2], opt_g, targets)
    print("captured, capacity", trainer.capacity, flush=True)
    for k in range(4):
        colors, _ = batch.rasterize_voxels_batch(views, octpath, centers, lengths, *pe)
        loss = sum(torch.nn.functional.mse_loss(c, tg) for c, tg in zip(colors, targets))
        opt_e.zero_grad(set_to_none=True)
        loss.backward()
        opt_e.step()
        eager_color = torch.cat([c.detach().reshape(-1) for c in colors])

        trainer.step()
        # eager work and a host synchronisation between replays
        _ = (pg[0] * 2.0).sum().item()
        torch.cuda.synchronize()
        e = rel_err(trainer.color.cpu().numpy(), eager_color.cpu().numpy(), 0)
        assert e <= 1e-5, f"replay {k}: image rel err {e:.3g}"
        for a, b, name in zip(pg, pe, ("geos", "sh0", "shs")):
            e = rel_err(a.detach().cpu().numpy(), b.detach().cpu().numpy())
            assert e <= 1e-4, f"replay {k}: {name} rel err {e:.3g}"
    assert not trainer.overflowed()


@pytest.mark.gpu
def test_graph_adam_matches_oracle_adam():
    """Every replay's Adam update equals the oracle's restatement of adam_step.cu applied to the gradients that
    replay computed, bit for bit (parameters and both moments), so the device step counts and the host-filled
    bias terms give exactly the host-step update."""
    import torch
    from oracle import oracle as O
    from sv_raster_amd import batch
    cases, sc, (octpath, centers, lengths), targets, t = _setup(seed=6)
    views = batch.ViewBatch([_settings(c["cam"]) for c in cases])
    pg, opt_g = _params(sc, t)
    host = [p.detach().cpu().numpy().copy() for p in pg]
    moments = [(np.zeros_like(h), np.zeros_like(h)) for h in host]
    lrs = (0.025, 0.01, 0.01)
    trainer = batch.GraphTrainStep(views, octpath, centers, lengths, pg[0], pg[1], pg[2], opt_g, targets)
    for k in range(1, 6):
        trainer.step()
        torch.cuda.synchronize()
        for i, (p, g) in enumerate(zip(pg, trainer.grads)):
            gh = g.detach().cpu().numpy().astype(np.float32).copy()
            m, v = moments[i]
            O.adam_step(host[i], gh, m, v, k, lrs[i], 0.1, 0.99, 1e-15, sparse=True)
            st = opt_g.state[p]
            assert np.array_equal(p.detach().cpu().numpy().view(np.uint32), host[i].view(np.uint32)), (k, i)
            assert np.array_equal(st["exp_avg"].cpu().numpy().view(np.uint32), m.view(np.uint32)), (k, i)
            assert np.array_equal(st["exp_avg_sq"].cpu().numpy().view(np.uint32), v.view(np.uint32)), (k, i)
    assert not trainer.overflowed()


@pytest.mark.gpu
def test_graph_survives_order_table_cache_eviction():
    """The trainer holds the octree's rank-order table: after the table cache has moved on to other octrees
    (and their allocations may reuse freed memory), a replay still matches the eager step."""
    import torch
    from sv_raster_amd import _C, batch
    from sv_raster_amd import synthetic as S
    cases, sc, (octpath, centers, lengths), targets, t = _setup(seed=7)
    views = batch.ViewBatch([_settings(c["cam"]) for c in cases])
    pe, opt_e = _params(sc, t)
    pg, opt_g = _params(sc, t)
    trainer = batch.GraphTrainStep(views, octpath, centers, lengths, pg[0], pg[1], pg[2], opt_g, targets)
    others = []
    for j in range(_C.order_table_cache.capacity + 2):
        other = t(S.make_scene("C", seed=50 + j)["octpath"])
        _C.order_table_cache.get(other)
        others.append(other)
    torch.cuda.empty_cache()
    junk = [torch.full((1 << 16,), -1, dtype=torch.int64, device="cuda") for _ in range(8)]
    torch.cuda.synchronize()
    trainer.step()
    colors, _ = batch.rasterize_voxels_batch(views, octpath, centers, lengths, *pe)
    loss = sum(torch.nn.functional.mse_loss(c, tg) for c, tg in zip(colors, targets))
    opt_e.zero_grad(set_to_none=True)
    loss.backward()
    opt_e.step()
    torch.cuda.synchronize()
    eager_color = torch.cat([c.detach().reshape(-1) for c in colors])
    assert rel_err(trainer.color.cpu().numpy(), eager_color.cpu().numpy(), 0) <= 1e-5
    for a, b in zip(pg, pe):
        assert rel_err(a.detach().cpu().numpy(), b.detach().cpu().numpy()) <= 1e-4
    del junk, others


@pytest.mark.gpu
@pytest.mark.parametrize("sparse", [True, False])
def test_graph_overflow_is_flagged_and_harmless(sparse):
    """A frame over the entry capacity is flagged and changes nothing: parameters, moments and step counts stay
    as they were, with sparse and with dense Adam."""
    import torch
    from sv_raster_amd import batch, sparse_adam
    cases, sc, (octpath, centers, lengths), targets, t = _setup(seed=5)
    views = batch.ViewBatch([_settings(c["cam"]) for c in cases])
    pg = [t(sc["geos"]).requires_grad_(), t(sc["sh0"]).requires_grad_(), t(sc["shs"]).requires_grad_()]
    opt_g = sparse_adam.SparseAdam([{"params": [pg[0]], "lr": 0.025}, {"params": [pg[1]], "lr": 0.01},
                                    {"params": [pg[2]], "lr": 0.01}], betas=(0.1, 0.99), eps=1e-15, sparse=sparse)
    before = [p.detach().clone() for p in pg]
    trainer = batch.GraphTrainStep(views, octpath, centers, lengths, pg[0], pg[1], pg[2], opt_g, targets,
                                   capacity=64)
    trainer.step()
    trainer.step()
    torch.cuda.synchronize()
    assert trainer.overflowed()
    for a, b in zip(pg, before):
        assert torch.equal(a.detach(), b)
        st = opt_g.state[a]
        assert int(st["step_dev"].item()) == 0
        assert float(st["exp_avg"].abs().max()) == 0.0 and float(st["exp_avg_sq"].abs().max()) == 0.0
    assert float(trainer.color.abs().max()) == 0.0  # background colour 0, no entries rendered

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
def test_graph_overflow_is_flagged_and_harmless():
    import torch
    from sv_raster_amd import batch
    cases, sc, (octpath, centers, lengths), targets, t = _setup(seed=5)
    views = batch.ViewBatch([_settings(c["cam"]) for c in cases])
    pg, opt_g = _params(sc, t)
    before = [p.detach().clone() for p in pg]
    trainer = batch.GraphTrainStep(views, octpath, centers, lengths, pg[0], pg[1], pg[2], opt_g, targets,
                                   capacity=64)
    trainer.step()
    torch.cuda.synchronize()
    assert trainer.overflowed()
    for a, b in zip(pg, before):
        assert torch.equal(a.detach(), b)
    assert float(trainer.color.abs().max()) == 0.0  # background colour 0, no entries rendered

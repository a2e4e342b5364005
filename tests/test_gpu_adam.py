"""Sparse Adam on the GPU (through the C ABI) against the oracle's restatement of
adam_step.cu, bit for bit: the same fma / sqrt / division sequence per element.

Sizes on and off multiples of 4 and views at a 4-byte offset cover both the
16-byte vector kernel (with its scalar tail) and the scalar kernel; sparse
gradients have exact zeros both in whole 4-element vectors and mixed inside
one vector."""
import numpy as np
import pytest


def _case(n, seed, density):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 1, n).astype(np.float32)
    g = rng.normal(0, 1e-2, n).astype(np.float32)
    g[rng.random(n) >= density] = 0.0
    if n >= 16:
        g[8:12] = 0.0  # one whole zero vector
    m = rng.normal(0, 1e-3, n).astype(np.float32)
    v = np.abs(rng.normal(0, 1e-4, n)).astype(np.float32)
    return p, g, m, v


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 7, 4096, 100003])
@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("sparse,biased", [(True, False), (False, False), (True, True), (False, True)])
def test_adam_matches_oracle(n, offset, sparse, biased):
    import torch
    from oracle import oracle as O
    from sv_raster_amd import _C
    print(f"adam n={n} offset={offset} sparse={sparse} biased={biased}", flush=True)
    p, g, m, v = _case(n, seed=n + 7 * offset, density=0.4)

    def dev(a):  # a view starting `offset` floats into its allocation (offset 1: not 16-byte aligned)
        buf = torch.zeros(a.size + offset, dtype=torch.float32, device="cuda")
        buf[offset:] = torch.from_numpy(a).cuda()
        return buf[offset:]

    tp, tg, tm, tv = dev(p), dev(g), dev(m), dev(v)
    for step in (1, 2, 3):
        if biased:
            _C.biased_adam_step(sparse, tp, tg, tm, tv, 0.01, 0.1, 0.99, 1e-15)
        else:
            _C.unbiased_adam_step(sparse, tp, tg, tm, tv, step, 0.025, 0.1, 0.99, 1e-15)
        O.adam_step(p, g, m, v, step, 0.01 if biased else 0.025, 0.1, 0.99, 1e-15, biased=biased, sparse=sparse)
    torch.cuda.synchronize()
    for name, a, b in (("param", tp, p), ("exp_avg", tm, m), ("exp_avg_sq", tv, v)):
        assert np.array_equal(a.cpu().numpy().view(np.uint32), b.view(np.uint32)), name
    if sparse:
        assert np.array_equal(tp.cpu().numpy()[g == 0], _case(n, seed=n + 7 * offset, density=0.4)[0][g == 0])


@pytest.mark.gpu
@pytest.mark.parametrize("sparse", [True, False])
def test_adam_device_step_matches_host_step(sparse):
    """The capture-safe form (step count on the device, bias terms from the host-filled table) equals the
    host-step form and the oracle bit for bit, also across table refills (a 3-step table over 10 steps)."""
    import torch
    from oracle import oracle as O
    from sv_raster_amd import sparse_adam
    print(f"adam device steps sparse={sparse}", flush=True)
    p, g, m, v = _case(10001, seed=3, density=0.5)
    ph, pd = torch.from_numpy(p.copy()).cuda(), torch.from_numpy(p.copy()).cuda()
    oh = sparse_adam.SparseAdam([ph], lr=0.025, betas=(0.1, 0.99), eps=1e-15, sparse=sparse)
    od = sparse_adam.SparseAdam([pd], lr=0.025, betas=(0.1, 0.99), eps=1e-15, sparse=sparse)
    od.BIAS_TABLE_STEPS = 3
    od.use_device_steps()
    po, mo, vo = p.copy(), np.zeros_like(p), np.zeros_like(p)
    rng = np.random.default_rng(11)
    for step in range(1, 11):
        gs = (g * rng.normal(1.0, 0.5, g.size)).astype(np.float32)
        ph.grad = torch.from_numpy(gs).cuda()
        pd.grad = torch.from_numpy(gs).cuda()
        oh.step()
        od.step()
        O.adam_step(po, gs, mo, vo, step, 0.025, 0.1, 0.99, 1e-15, sparse=sparse)
    torch.cuda.synchronize()
    assert int(od.state[pd]["step_dev"].item()) == 10
    for a, b in ((ph, pd), (oh.state[ph]["exp_avg"], od.state[pd]["exp_avg"]),
                 (oh.state[ph]["exp_avg_sq"], od.state[pd]["exp_avg_sq"])):
        assert torch.equal(a, b)
    assert np.array_equal(pd.cpu().numpy().view(np.uint32), po.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("sparse", [True, False])
def test_adam_device_step_skip_changes_nothing(sparse):
    """A nonzero skip word (a frame over its entry capacity) leaves parameters, moments and the step count as
    they were, for sparse and dense Adam; the next step without skip is the host form's next step."""
    import torch
    from sv_raster_amd import sparse_adam
    print(f"adam skip sparse={sparse}", flush=True)
    p, g, m, v = _case(4099, seed=5, density=0.5)
    skip = torch.zeros(1, dtype=torch.int32, device="cuda")
    pd, ph = torch.from_numpy(p.copy()).cuda(), torch.from_numpy(p.copy()).cuda()
    od = sparse_adam.SparseAdam([pd], lr=0.025, betas=(0.1, 0.99), eps=1e-15, sparse=sparse)
    oh = sparse_adam.SparseAdam([ph], lr=0.025, betas=(0.1, 0.99), eps=1e-15, sparse=sparse)
    od.use_device_steps(skip=skip)
    pd.grad = torch.from_numpy(g).cuda()
    ph.grad = torch.from_numpy(g).cuda()
    od.step()
    oh.step()
    before = [t.clone() for t in (pd, od.state[pd]["exp_avg"], od.state[pd]["exp_avg_sq"])]
    skip.fill_(1)
    pd.grad = torch.zeros_like(pd)  # what an overflowing frame produces
    od.step()
    torch.cuda.synchronize()
    assert int(od.state[pd]["step_dev"].item()) == 1
    for a, b in zip((pd, od.state[pd]["exp_avg"], od.state[pd]["exp_avg_sq"]), before):
        assert torch.equal(a, b)
    skip.fill_(0)
    pd.grad = torch.from_numpy(g).cuda()
    od.step()
    oh.step()
    torch.cuda.synchronize()
    assert torch.equal(pd, ph)


@pytest.mark.gpu
@pytest.mark.parametrize("sparse,biased", [(True, False), (False, False), (True, True)])
def test_adam_multi_matches_oracle(sparse, biased):
    """One launch over several tensors (SparseAdam.step's path): each tensor, with its own step count, lr,
    betas and eps, equals the oracle bit for bit.  Eleven tensors (two launches of at most eight), sizes on
    and off multiples of 4, one view at a 4-byte offset (sent through the per-tensor kernel)."""
    import torch
    from oracle import oracle as O
    from sv_raster_amd import _C
    sizes = [1, 7, 4096, 100003, 64, 5, 12, 333, 8, 100000, 3]
    print(f"adam multi sparse={sparse} biased={biased}", flush=True)
    host, dev, hyper = [], [], []
    for i, n in enumerate(sizes):
        p, g, m, v = _case(n, seed=100 + i, density=0.4)
        offset = 1 if i == 3 else 0
        ts = []
        for a in (p, g, m, v):
            buf = torch.zeros(a.size + offset, dtype=torch.float32, device="cuda")
            buf[offset:] = torch.from_numpy(a).cuda()
            ts.append(buf[offset:])
        host.append([p, g, m, v])
        dev.append(ts)
        hyper.append((0.01 * (1 + i), 0.1 + 0.02 * i, 0.99 - 0.001 * i, 1e-15 * (1 + i)))
    for k in range(1, 4):
        steps = [k + i for i in range(len(sizes))]
        _C.adam_step_multi(biased, sparse, [(t[0], t[1], t[2], t[3], s, *h) for t, s, h in zip(dev, steps, hyper)])
        for (p, g, m, v), s, (lr, b1, b2, eps) in zip(host, steps, hyper):
            O.adam_step(p, g, m, v, s, lr, b1, b2, eps, biased=biased, sparse=sparse)
    torch.cuda.synchronize()
    for i, (ts, hs) in enumerate(zip(dev, host)):
        for name, a, b in zip(("param", "grad", "exp_avg", "exp_avg_sq"), ts, hs):
            assert np.array_equal(a.cpu().numpy().view(np.uint32), b.view(np.uint32)), (i, name)


@pytest.mark.gpu
def test_sparse_adam_step_matches_oracle():
    """SparseAdam.step (three parameter groups with their own lr, sparse) against the oracle, bit for bit."""
    import torch
    from oracle import oracle as O
    from sv_raster_amd import sparse_adam
    sizes, lrs = [8000, 3000, 45000], [0.025, 0.01, 0.01]
    params, host = [], []
    for i, n in enumerate(sizes):
        p, g, _, _ = _case(n, seed=200 + i, density=0.6)
        t = torch.from_numpy(p).cuda().requires_grad_()
        params.append(t)
        host.append([p, g, np.zeros_like(p), np.zeros_like(p)])
    opt = sparse_adam.SparseAdam([{"params": [t], "lr": lr} for t, lr in zip(params, lrs)], betas=(0.1, 0.99),
                                 eps=1e-15, sparse=True)
    for step in (1, 2, 3):
        for t, h in zip(params, host):
            t.grad = torch.from_numpy(h[1]).cuda()
        opt.step()
        for h, lr in zip(host, lrs):
            O.adam_step(h[0], h[1], h[2], h[3], step, lr, 0.1, 0.99, 1e-15, sparse=True)
    torch.cuda.synchronize()
    for t, h in zip(params, host):
        assert np.array_equal(t.detach().cpu().numpy().view(np.uint32), h[0].view(np.uint32))
        st = opt.state[t]
        assert np.array_equal(st["exp_avg"].cpu().numpy().view(np.uint32), h[2].view(np.uint32))
        assert np.array_equal(st["exp_avg_sq"].cpu().numpy().view(np.uint32), h[3].view(np.uint32))

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
def test_adam_device_step_matches_host_step():
    """The capture-safe form (step count and bias terms on the device) equals the host-step form."""
    import torch
    from sv_raster_amd import _C
    p, g, m, v = _case(10001, seed=3, density=0.5)
    t = lambda a: torch.from_numpy(a.copy()).cuda()  # noqa: E731
    a = [t(x) for x in (p, g, m, v)]
    b = [t(x) for x in (p, g, m, v)]
    step_dev = torch.zeros(1, dtype=torch.int32, device="cuda")
    coef = torch.zeros(2, dtype=torch.float32, device="cuda")
    for step in (1, 2, 3, 4):
        _C.unbiased_adam_step(True, a[0], a[1], a[2], a[3], step, 0.025, 0.1, 0.99, 1e-15)
        _C.unbiased_adam_step_device(True, b[0], b[1], b[2], b[3], step_dev, coef, 0.025, 0.1, 0.99, 1e-15)
    torch.cuda.synchronize()
    assert int(step_dev.item()) == 4
    assert torch.equal(a[2], b[2]) and torch.equal(a[3], b[3])  # moments do not use the bias terms
    # the bias terms come from pow() on the device instead of the host: equal up to float rounding
    assert torch.allclose(a[0], b[0], rtol=1e-6, atol=0.0)

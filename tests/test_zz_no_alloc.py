"""The allocation-free entry point wrapd_mlp_grad_noalloc against wrapd_mlp_grad.

Both run the same kernels in the same order on the same tuples, so loss and gradient must
be identical bit for bit.  Two calls in a row, the second with fewer tuples (the static
scratch array keeps the first call's contents), then one at the stated maximum, and the
size check beyond it.  A third context on the same GPU shares the array.
"""
import numpy as np
import pytest

from wrapd_amd.solver import Device, WrapdError

pytestmark = pytest.mark.gpu


def _tuples(rng, n):
    s = np.sort(rng.uniform(0.5, 1.6, (n, 3)), axis=1)[:, ::-1]
    mu = 10 ** rng.uniform(4.5, 6.5, n)
    lam = mu * rng.uniform(2.0, 12.0, n)
    wov = mu * 10 ** rng.uniform(-1, 1.5, n)
    x = s + rng.normal(0, 0.05, (n, 3))
    return np.column_stack([s, wov, mu, lam, x]).astype(np.float32)


def _same(d, t):
    loss_a, g_a = d.mlp_grad(t)
    loss_b, g_b = d.mlp_grad_noalloc(t)
    print(f"n {len(t)}: loss {loss_a:.7e} / {loss_b:.7e}, max |dg| {np.max(np.abs(g_a - g_b)):.1e}")
    assert np.isfinite(loss_a) and np.all(np.isfinite(g_a))
    assert np.array_equal(np.float32(loss_a), np.float32(loss_b))
    assert np.array_equal(g_a, g_b)
    return g_b


def test_noalloc_grad_matches_allocating_entry_point():
    rng = np.random.default_rng(23)
    d = Device()
    d.mlp_init(seed=3)
    g1 = _same(d, _tuples(rng, 20000))
    g2 = _same(d, _tuples(rng, 777))      # smaller: stale scratch contents beyond 777 must not leak in
    assert not np.array_equal(g1, g2)
    _same(d, _tuples(rng, 65536))         # the stated maximum
    with pytest.raises(WrapdError) as e:
        d.mlp_grad_noalloc(_tuples(rng, 65537))
    assert e.value.status == 1            # WRAPD_ERR_ARG
    other = Device()                      # another context on the same GPU shares the array
    other.mlp_init(seed=4)
    _same(other, _tuples(rng, 1000))
    _same(d, _tuples(rng, 500))

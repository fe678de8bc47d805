"""GPU: gsk_reduce_eta_noalloc, the no-allocation variant of gsk_reduce's real-space reductions
(mean eta_v and the sum of eta^2) on order parameters held in the caller's device memory.

Bar: bit-exact against gsk_reduce on a handle holding the same fields (same kernel, block count
and summation order), on two calls in a row, the second with a smaller input; a ragged size
against an FP64 numpy sum; sizes beyond the scratch are rejected."""
import sys

import numpy as np
import pytest

import gsk

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def test_reduce_eta_noalloc_matches_reduce(gpu):
    import torch

    for n, V, seed in [(32, 3, 5), (16, 2, 9)]:  # the second call has the smaller input
        say(f"noalloc reduction n={n} V={V}")
        cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=V)
        with gsk.Sim(cfg) as s:
            s.build_btensor(gsk.cubic_modulus(0.7, 0.3, 0.2), gsk.mart_eigenstrains(V))
            s.init_uniform("eta", -0.3, 0.5, seed)
            want_mean = s.mean_eta()
            want_s2 = s.reduce(gsk.RED_SUM_ETA2)[0]
            eta = s.get_field("eta")
        d = torch.from_numpy(eta).to("cuda:0")
        torch.cuda.synchronize()
        got_mean = gsk.reduce_eta_noalloc(d.data_ptr(), V, n**3, gsk.RED_MEAN_ETA)
        got_s2 = gsk.reduce_eta_noalloc(d.data_ptr(), V, n**3, gsk.RED_SUM_ETA2)
        assert np.array_equal(got_mean, want_mean), (got_mean, want_mean)
        assert got_s2[0] == want_s2


def test_reduce_eta_noalloc_ragged_and_limits(gpu):
    import torch

    say("noalloc reduction ragged size and limits")
    rng = np.random.default_rng(4)
    V, nvox = 5, 1001
    eta = rng.uniform(-1.0, 1.0, (V, nvox)).astype(np.float32)
    d = torch.from_numpy(eta).to("cuda:0")
    torch.cuda.synchronize()
    got = gsk.reduce_eta_noalloc(d.data_ptr(), V, nvox, gsk.RED_MEAN_ETA)
    want = eta.astype(np.float64).mean(axis=1)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-16)
    s2 = gsk.reduce_eta_noalloc(d.data_ptr(), V, nvox, gsk.RED_SUM_ETA2)[0]
    assert abs(s2 - float((eta.astype(np.float64) ** 2).sum())) < 1e-12 * s2
    with pytest.raises(gsk.GskError):
        gsk.reduce_eta_noalloc(d.data_ptr(), 9, nvox // 2, gsk.RED_MEAN_ETA)
    with pytest.raises(gsk.GskError):
        gsk.reduce_eta_noalloc(d.data_ptr(), V, nvox, gsk.RED_ENERGY)

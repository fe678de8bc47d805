"""GPU: the checkpointed adjoint of the FP32 martensite step (gsk_adjoint_*) and its torch binding.

Bars: against the FP64 store-all adjoint of the oracle (itself pinned by central
differences in test_oracle.py) within FP32 tolerances written in each test; bit-exact
between checkpoint intervals (the recomputed segments reproduce the recorded trajectory
exactly and the reductions are deterministic).
"""
import sys

import numpy as np
import pytest

import gsk
import oracle as O

pytestmark = pytest.mark.gpu

C = gsk.cubic_modulus(0.7, 0.3, 0.2)
V = 3


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def make(n, noise=0.01):
    cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, A=-0.3, B=3.0, C=3.0, kappa=1.2, L=0.8, dt=0.05,
                             w_el=20.0, noise=noise, seed=9)
    s = gsk.Sim(cfg)
    s.build_btensor(C, gsk.mart_eigenstrains(V))
    s.init_uniform("eta", -0.1, 0.1, 21)
    return s, cfg


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("n,steps,K", [(16, 10, 4), (32, 12, 5)])
def test_adjoint_matches_oracle(gpu, n, steps, K):
    say(f"adjoint n={n} steps={steps} K={K}")
    s, cfg = make(n)
    eta0 = s.get_field("eta").astype(np.float64)
    rng = np.random.default_rng(3)
    gT = rng.standard_normal(eta0.shape).astype(np.float32)
    s.adjoint_forward(steps, K)
    ge, gp = s.adjoint_backward(gT)
    s.close()
    p = O.mart_params(A=cfg.A, B=cfg.B, C=cfg.C, kappa=cfg.kappa, L=cfg.L, dt=cfg.dt, w_el=cfg.w_el, noise=cfg.noise,
                      seed=cfg.seed)
    we, wp = O.mart_adjoint(eta0, O.btensor(n, C, gsk.mart_eigenstrains(V)), p, steps, gT.astype(np.float64))
    say(f"  rel err eta0 {rel(ge, we):.2e}  params {[f'{abs(a - b) / abs(b):.1e}' for a, b in zip(gp, wp)]}")
    assert rel(ge, we) < 1e-4
    # L and kappa: sums over every step and voxel; the s_p gradients are differences of
    # elastic inner products and carry FP32 cancellation (relative 1e-3 allowed)
    assert abs(gp[V] - wp[V]) < 1e-4 * abs(wp[V])
    assert abs(gp[V + 1] - wp[V + 1]) < 1e-4 * abs(wp[V + 1])
    for q in range(V):
        assert abs(gp[q] - wp[q]) < 1e-3 * abs(wp[q]), (q, gp[q], wp[q])


def test_adjoint_checkpoint_interval_is_exact(gpu):
    say("adjoint K=1 (store-all) vs K=3 vs K=7 vs K=20")
    out = []
    for K in (1, 3, 7, 20):
        s, _ = make(32)
        gT = np.random.default_rng(5).standard_normal((V, 32, 32, 32)).astype(np.float32)
        s.adjoint_forward(14, K)
        etaT = s.get_field("eta")
        ge, gp = s.adjoint_backward(gT)
        info = s.adjoint_info()
        # the handle is left at the final state, ready to continue
        assert np.array_equal(s.get_field("eta"), etaT) and s.step_index == 14
        assert info["stored_states"] == (14 + K - 1) // K + 1
        out.append((ge, gp))
        s.close()
    for ge, gp in out[1:]:
        assert np.array_equal(ge.view(np.uint32), out[0][0].view(np.uint32))
        assert np.array_equal(gp, out[0][1])


def test_adjoint_forward_equals_step(gpu):
    say("adjoint forward reproduces gsk_step")
    a, _ = make(32)
    b, _ = make(32)
    a.step(9)
    b.adjoint_forward(9, 4)
    assert np.array_equal(a.get_field("eta").view(np.uint32), b.get_field("eta").view(np.uint32))
    a.close()
    b.close()


def test_torch_rollout_gradients(gpu):
    torch = pytest.importorskip("torch")
    from gsk.autograd import rollout

    say("torch autograd rollout")
    n, steps = 16, 6
    s, cfg = make(n, noise=0.0)
    eta0_np = s.get_field("eta")
    eps = gsk.mart_eigenstrains(V)
    target = torch.from_numpy(np.random.default_rng(2).uniform(-0.1, 0.1, (V, n, n, n)).astype(np.float32))
    eta0 = torch.from_numpy(eta0_np.copy()).requires_grad_(True)
    params = torch.tensor([1.1, 0.9, 1.0, cfg.L, cfg.kappa], dtype=torch.float64, requires_grad=True)
    etaT = rollout(s, eta0, params, C, eps, steps, K=4)
    loss = 0.5 * ((etaT - target) ** 2).sum()
    loss.backward()
    # same gradients from the oracle at the same parameters (eigenstrains scaled by s)
    sc = np.array([1.1, 0.9, 1.0])
    p = O.mart_params(A=cfg.A, B=cfg.B, C=cfg.C, kappa=cfg.kappa, L=cfg.L, dt=cfg.dt, w_el=cfg.w_el)
    B = O.btensor(n, C, eps * sc[:, None, None])
    wT = O.mart_run(eta0_np.astype(np.float64), B, p, steps)
    we, wp = O.mart_adjoint(eta0_np.astype(np.float64), B, p, steps, wT - target.numpy().astype(np.float64))
    wp[:V] /= sc
    assert rel(eta0.grad.numpy(), we) < 1e-4
    g = params.grad.numpy()
    assert np.all(np.abs(g - wp) < 1e-3 * np.abs(wp)), (g, wp)
    s.close()

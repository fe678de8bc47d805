"""GPU: the reference rules of both plugins (rule = GSK_RULE_REFERENCE) against the FP64 oracle.

The oracle restates src/dynamics_mart.cu:181-417 and src/dynamics_precipitate.cu:118-289
(or_mart_ref_run / or_prec_ref_run, pinned in test_oracle.py). The GPU computes in FP32 like
the reference; tolerances (relative L2 after the stated steps) are written in each test.
"""
import sys

import numpy as np
import pytest

import gsk
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("chem_form,noise", [(0, 1e-4), (1, 0.0)])
def test_mart_reference_rule_matches_oracle(gpu, chem_form, noise):
    say(f"mart reference rule chem_form={chem_form} noise={noise}")
    n, V, steps = 32, 3, 40
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.01)
    cfg = gsk.reference_config(gsk.PHYS_MART, n=n, nvar=V, temperature=300.0, chem_form=chem_form, noise=noise,
                               w_el=1e3, seed=6, dt=0.005)
    rng = np.random.default_rng(1)
    defect = (rng.random((n, n, n)) < 0.01).astype(np.float32)
    stress = rng.uniform(-0.01, 0.01, (6, n, n, n)).astype(np.float32)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", 0.0, 0.2, 3)
        s.set_field("defect", defect)
        s.set_field("dislocationstress", stress)
        eta0 = s.get_field("eta").astype(np.float64)
        s.step(steps)
        got = s.get_field("eta", np.float64)
    a1 = 0.02 * (cfg.temperature - cfg.T0)
    p = O.mart_ref_params(a1=a1, a2=cfg.B, a3=cfg.C, w_grad=cfg.w_grad, w_chem=cfg.w_chem, w_el=cfg.w_el,
                          w_dis=cfg.w_dis, w_noise=noise, dt=cfg.dt, dx=cfg.dx, kconv=cfg.kconv, chem_form=chem_form,
                          seed=cfg.seed)
    comp = [0, 1, 2, 1, 3, 4, 2, 4, 5]
    sig = stress.astype(np.float64)[comp].reshape(3, 3, n, n, n)
    K = 25.0 * np.einsum("ijxyz,pij->pxyz", sig, eps)
    want = O.mart_ref_run(eta0, O.btensor(n, C, eps, kconv=cfg.kconv), p, steps, defect=defect.astype(np.float64),
                          K=K)
    err = rel(got, want)
    say(f"  rel L2 {err:.2e}, |eta| max {np.abs(want).max():.3f}")
    assert np.isfinite(got).all()
    assert np.all(got[:, defect > 0] == 0.0)
    assert err < 2e-5


@pytest.mark.parametrize("n,V", [(16, 6), (32, 2)])
def test_mart_reference_rule_variant_counts(gpu, n, V):
    """Variant counts beyond the 16-byte update kernel's 1..4 (V = 6: the general kernel,
    `k_ref_mart`) and another one inside it (V = 2), with noise, defects and dislocation stress,
    against the FP64 oracle. Eigenstrains of variants p and p + 3 share an axis."""
    say(f"mart reference rule n={n} V={V}")
    steps = 20
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    rng = np.random.default_rng(3)
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.01) + 0.002 * rng.standard_normal((V, 3, 3))
    eps = 0.5 * (eps + eps.transpose(0, 2, 1))
    cfg = gsk.reference_config(gsk.PHYS_MART, n=n, nvar=V, temperature=300.0, noise=1e-4, w_el=1e3, seed=9,
                               dt=0.005)
    defect = (rng.random((n, n, n)) < 0.02).astype(np.float32)
    stress = rng.uniform(-0.01, 0.01, (6, n, n, n)).astype(np.float32)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", 0.0, 0.2, 5)
        s.set_field("defect", defect)
        s.set_field("dislocationstress", stress)
        eta0 = s.get_field("eta").astype(np.float64)
        s.step(steps)
        got = s.get_field("eta", np.float64)
        forces = s.mart_forces()
    a1 = 0.02 * (cfg.temperature - cfg.T0)
    p = O.mart_ref_params(a1=a1, a2=cfg.B, a3=cfg.C, w_grad=cfg.w_grad, w_chem=cfg.w_chem, w_el=cfg.w_el,
                          w_dis=cfg.w_dis, w_noise=cfg.noise, dt=cfg.dt, dx=cfg.dx, kconv=cfg.kconv,
                          chem_form=cfg.chem_form, seed=cfg.seed)
    comp = [0, 1, 2, 1, 3, 4, 2, 4, 5]
    sig = stress.astype(np.float64)[comp].reshape(3, 3, n, n, n)
    K = 25.0 * np.einsum("ijxyz,pij->pxyz", sig, eps)
    want = O.mart_ref_run(eta0, O.btensor(n, C, eps, kconv=cfg.kconv), p, steps, defect=defect.astype(np.float64),
                          K=K)
    err = rel(got, want)
    say(f"  rel L2 {err:.2e}")
    assert np.isfinite(got).all() and np.isfinite(forces).all()
    assert np.all(got[:, defect > 0] == 0.0)
    assert err < 2e-5


def _mart_ref_256(noise, n=256):
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    eps = gsk.mart_eigenstrains(3, base=0.0, axis=0.01)
    cfg = gsk.reference_config(gsk.PHYS_MART, n=n, nvar=3, temperature=300.0, noise=noise, w_el=1e3, seed=6,
                               dt=0.005)
    return C, eps, cfg


def test_mart_reference_rule_matches_oracle_at_256(gpu):
    """n = 256, where the elastic R of the reference rule runs through the sub-plane kernels and
    the split-order contraction (three launches) instead of the chain of z, y and x kernels:
    three steps with noise, defects and dislocation stress against the FP64 oracle."""
    say("mart reference rule n=256 (split R) vs oracle")
    n, V, steps = 256, 3, 3
    C, eps, cfg = _mart_ref_256(1e-4)
    rng = np.random.default_rng(2)
    defect = (rng.random((n, n, n)) < 0.01).astype(np.float32)
    stress = rng.uniform(-0.01, 0.01, (6, n, n, n)).astype(np.float32)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", 0.0, 0.2, 3)
        s.set_field("defect", defect)
        s.set_field("dislocationstress", stress)
        eta0 = s.get_field("eta").astype(np.float64)
        s.step(steps)
        got = s.get_field("eta", np.float64)
        names, _ = s.kernel_times(1)
    assert "sub_inv_R" in names, names
    a1 = 0.02 * (cfg.temperature - cfg.T0)
    p = O.mart_ref_params(a1=a1, a2=cfg.B, a3=cfg.C, w_grad=cfg.w_grad, w_chem=cfg.w_chem, w_el=cfg.w_el,
                          w_dis=cfg.w_dis, w_noise=cfg.noise, dt=cfg.dt, dx=cfg.dx, kconv=cfg.kconv,
                          chem_form=cfg.chem_form, seed=cfg.seed)
    comp = [0, 1, 2, 1, 3, 4, 2, 4, 5]
    sig = stress.astype(np.float64)[comp].reshape(3, 3, n, n, n)
    K = 25.0 * np.einsum("ijxyz,pij->pxyz", sig, eps)
    del sig
    want = O.mart_ref_run(eta0, O.btensor(n, C, eps, kconv=cfg.kconv), p, steps, defect=defect.astype(np.float64),
                          K=K)
    err = rel(got, want)
    say(f"  rel L2 {err:.2e}")
    assert np.isfinite(got).all()
    assert np.all(got[:, defect > 0] == 0.0)
    assert err < 1e-5


@pytest.mark.parametrize("n", [256, 192, 128])
def test_mart_reference_rule_split_matches_chain_at_256(gpu, monkeypatch, n):
    """The reference rule's elastic R at n = 256 and 192 through the sub-plane kernels (default)
    and through the chain of z, y and x kernels (GSK_SPLIT=0 when the handle is created): two
    FP32 orderings of one computation, so 10 steps with noise agree to FP32 rounding."""
    say(f"mart reference rule n={n}: split R vs chain")
    C, eps, cfg = _mart_ref_256(1e-3, n)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("GSK_SPLIT", mode)
        with gsk.Sim(cfg) as s:
            s.build_btensor(C, eps)
            s.init_uniform("eta", 0.0, 0.2, 8)
            s.step(10)
            out[mode] = (s.get_field("eta", np.float64), s.mart_forces())
            names, _ = s.kernel_times(1)
        assert ("sub_inv_R" in names) == (mode == "1"), names
        assert ("zinv_R" in names) == (mode == "0"), names
    (a, fa), (b, fb) = out["1"], out["0"]
    assert np.isfinite(a).all()
    assert float(np.abs(a - b).max()) < 2e-6 * float(np.abs(b).max())
    assert rel(a, b) < 1e-6
    # the four force means of the last step (kernel_times ran one more step on each handle)
    say(f"  force means split {fa}, chain {fb}")
    assert np.all(np.abs(fa - fb) <= 1e-4 * np.abs(fb).max())


def test_prec_reference_rule_matches_oracle(gpu):
    say("precipitate reference rule")
    n, V, steps = 32, 3, 30
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.02)
    cfg = gsk.reference_config(gsk.PHYS_PREC, n=n, nvar=V, M=1.0, L=1.0, kappa=1.5, kappa_c=1.5, dt=0.002, seed=8,
                               xi=50.0)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("concentration", 0.245, 0.255, 11)
        s.init_uniform("eta", 0.0, 0.1, 12)
        c0 = s.get_field("concentration", np.float64)
        h0 = s.get_field("eta", np.float64)
        m0 = s.mass()
        s.step(steps)
        c1 = s.get_field("concentration", np.float64)
        h1 = s.get_field("eta", np.float64)
        m1 = s.mass()
    p = O.prec_params(A=list(cfg.a), c1=cfg.c1, c2=cfg.c2, M=cfg.M, L=cfg.L, kappa_c=cfg.kappa_c,
                      kappa_eta=cfg.kappa, xi=cfg.xi, S=0.0, dt=cfg.dt, noise=cfg.noise, kconv=cfg.kconv,
                      seed=cfg.seed)
    cw, hw = O.prec_ref_run(c0, h0, O.btensor(n, C, eps, kconv=cfg.kconv), p, steps)
    say(f"  rel L2 eta {rel(h1, hw):.2e}  c-cbar {rel(c1 - c0.mean(), cw - c0.mean()):.2e}")
    assert rel(h1, hw) < 1e-4
    assert rel(c1 - c0.mean(), cw - c0.mean()) < 1e-4
    assert m1 == m0


@pytest.mark.parametrize("n", [256, 192, 128])
def test_prec_reference_rule_split_matches_chain_at_256(gpu, monkeypatch, n):
    """The precipitate reference rule at n = 256 through the split-y kernels (R of the eta^2
    source in three launches, local forces with it in the sub-plane kernel, x update, sub-plane
    inverse) and through the chain of z, y and x kernels (GSK_SPLIT=0): 10 steps with noise agree
    to FP32 rounding, and sum c is conserved by both."""
    say(f"prec reference rule n={n}: split vs chain")
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    eps = gsk.mart_eigenstrains(3, base=0.0, axis=0.02)
    cfg = gsk.reference_config(gsk.PHYS_PREC, n=n, nvar=3, M=1.0, L=1.0, kappa=1.5, kappa_c=1.5, dt=0.002,
                               seed=8, xi=50.0, noise=0.01)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("GSK_SPLIT", mode)
        with gsk.Sim(cfg) as s:
            s.build_btensor(C, eps)
            s.init_uniform("concentration", 0.245, 0.255, 11)
            s.init_uniform("eta", 0.0, 0.1, 12)
            m0 = s.mass()
            s.step(10)
            out[mode] = (s.get_field("concentration", np.float64), s.get_field("eta", np.float64), m0, s.mass())
            names, _ = s.kernel_times(1)
        assert ("sub_inv_R" in names) == (mode == "1"), names
    (ca, ea, m0a, m1a), (cb, eb, m0b, m1b) = out["1"], out["0"]
    say(f"  max|d eta| {np.abs(ea - eb).max():.2e}  max|d c| {np.abs(ca - cb).max():.2e}")
    assert np.isfinite(ea).all() and np.isfinite(ca).all()
    assert m1a == m0a and m1b == m0b
    assert float(np.abs(ca - cb).max()) < 1e-6
    assert float(np.abs(ea - eb).max()) < 2e-6 * float(np.abs(eb).max())
    assert rel(ea, eb) < 1e-6


def test_reference_rule_errors(gpu):
    cfg = gsk.reference_config(gsk.PHYS_MART, n=16, nvar=2)
    with gsk.Sim(cfg) as s:
        s.build_btensor(gsk.cubic_modulus(3.5, 1.5, 1.0), gsk.mart_eigenstrains(2))
        s.init_uniform("eta", 0.0, 0.1, 1)
        with pytest.raises(gsk.GskError):
            s.energy()
        with pytest.raises(gsk.GskError):
            s.adjoint_forward(2, 1)
    spec = gsk.default_config(gsk.PHYS_MART, n=16, nvar=2)
    with gsk.Sim(spec) as s:
        with pytest.raises(gsk.GskError):
            s.set_field("defect", np.zeros((16, 16, 16), np.float32))


def test_mart_reference_force_means(gpu):
    """The Get columns gradient / chemical / elastic / dislocation (src/dynamics_mart.cu:446-451):
    means over voxels and variants of the four force terms of the last step.
    (a) Uniform eta, uniform dislocation stress: Grad = 0, Chem from the HEAD branches,
        El_p = -w_el (B(0) eta^2)_p eta_p with B(0) = C eps eps, Dis_p = -w_dis K_p eta_p with
        K_p = 25 sigma : eps_p -- each mean is a closed form.
    (b) A random state without noise or defects: the four means add up to the mean increment
        (eta1 - eta0)/dt of the FP64 oracle's step.
    Before the first step every mean is zero."""
    say("mart reference force means")
    n, V = 16, 3
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.01)
    cfg = gsk.reference_config(gsk.PHYS_MART, n=n, nvar=V, temperature=300.0, noise=0.0, w_el=1e3, w_dis=2.0, dt=0.005)
    vals = np.array([0.3, -0.2, 0.5])
    sig6 = np.array([0.01, 0.002, -0.003, -0.02, 0.001, 0.015])
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.set_field("eta", (vals[:, None, None, None] * np.ones((V, n, n, n))).astype(np.float32))
        s.set_field("dislocationstress", (sig6[:, None, None, None] * np.ones((6, n, n, n))).astype(np.float32))
        assert np.all(s.mart_forces() == 0.0)
        s.step(1)
        got = s.mart_forces()
    a1, a2, a3 = 0.02 * (cfg.temperature - cfg.T0), cfg.B, cfg.C
    sc = np.sqrt(abs(a1 / (a2 - a3)))
    S2 = ((sc * vals) ** 2).sum()
    chem = np.where(vals >= 0, -cfg.w_chem * (sc * vals) * (a1 - a2 * (sc * vals) ** 2 + a3 * S2),
                    1000.0 * cfg.w_chem * a2 * vals**2)
    B0 = np.einsum("ijkl,pij,qkl->pq", C, eps, eps)
    el = -cfg.w_el * (B0 @ vals**2) * vals
    comp = [0, 1, 2, 1, 3, 4, 2, 4, 5]
    K = 25.0 * np.einsum("ij,pij->p", sig6[comp].reshape(3, 3), eps)
    dis = -cfg.w_dis * K * vals
    want = np.array([0.0, chem.mean(), el.mean(), dis.mean()])
    say(f"  uniform: got {got}, want {want}")
    assert abs(got[0]) < 1e-6
    np.testing.assert_allclose(got[1:], want[1:], rtol=2e-5)

    cfg = gsk.reference_config(gsk.PHYS_MART, n=32, nvar=V, temperature=300.0, noise=0.0, w_el=1e3, dt=0.005)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", 0.0, 0.2, 5)
        eta0 = s.get_field("eta").astype(np.float64)
        s.step(1)
        got = s.mart_forces()
    p = O.mart_ref_params(a1=a1, a2=a2, a3=a3, w_grad=cfg.w_grad, w_chem=cfg.w_chem, w_el=cfg.w_el, w_dis=cfg.w_dis,
                          w_noise=0.0, dt=cfg.dt, dx=cfg.dx, kconv=cfg.kconv)
    eta1 = O.mart_ref_run(eta0, O.btensor(32, C, eps, kconv=cfg.kconv), p, 1)
    inc = ((eta1 - eta0) / cfg.dt).mean()
    say(f"  random: sum of means {got.sum():.6e}, oracle mean increment {inc:.6e}, terms {got}")
    assert got[3] == 0.0 and abs(got[0]) < 1e-6 * np.abs(got).max()
    assert abs(got.sum() - inc) < 1e-4 * np.abs(got[1:3]).max()

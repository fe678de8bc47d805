"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Bars: bit-exact for the Philox noise and the initial fields; 1e-12 relative for
the FP64 paths (B tensor, plain step); FP32 paths within the tolerances written
in each test (single-precision FFT rounding, ~1e-6 relative per transform).
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


def cubic(c11, c12, c44):
    return gsk.cubic_modulus(c11, c12, c44)


C_CFG0 = None


def c_cfg0():
    # C11/C12/C44 = 350/150/100 scaled by 1/500
    return cubic(350 / 500, 150 / 500, 100 / 500)


def mart_sim(n, V=3, path=gsk.PATH_FUSED32, **kw):
    cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, path=path, **kw)
    s = gsk.Sim(cfg)
    s.build_btensor(c_cfg0(), gsk.mart_eigenstrains(V))
    return s


def oracle_mart_params(cfg):
    return O.mart_params(A=cfg.A, B=cfg.B, C=cfg.C, kappa=cfg.kappa, L=cfg.L, dt=cfg.dt, w_el=cfg.w_el,
                         noise=cfg.noise, dx=cfg.dx, kconv=cfg.kconv, seed=cfg.seed)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# ------------------------------------------------------------------ noise and initial fields
@pytest.mark.parametrize("seed,step,field", [(0, 0, 0), (7, 3, 1), (2**40 + 5, 123456, 2), (0xFFFFFFFF, 1, 7)])
def test_noise_bit_exact(gpu, seed, step, field):
    say(f"noise seed={seed} step={step} field={field}")
    nv = 1 << 20
    g = gsk.noise_field(seed, step, field, nv)
    c = O.noise_field(seed, step, field, nv)
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32))


def test_uniform_init_bit_exact(gpu):
    say("init_uniform mart n=32")
    n, V = 32, 3
    s = mart_sim(n, V)
    s.init_uniform("eta", -0.01, 0.01, 0)
    got = s.get_field("eta")
    for v in range(V):
        want = O.uniform_field(0, v, n**3, -0.01, 0.01).reshape(n, n, n)
        assert np.array_equal(got[v], want)
    s.close()


# ------------------------------------------------------------------ B tensor (FP64 on the device)
@pytest.mark.parametrize("kconv,dx", [(0, 1.0), (1, 1.0), (0, 0.5)])
def test_btensor_fp64_matches_oracle(gpu, kconv, dx):
    say(f"btensor kconv={kconv} dx={dx}")
    n = 16
    C = cubic(3.5, 1.5, 1.0)
    e = gsk.mart_eigenstrains(4)
    g = gsk.btensor_full(n, dx, kconv, C, e)
    o = O.btensor(n, C, e, dx=dx, kconv=kconv)
    np.testing.assert_allclose(g, o, rtol=0, atol=1e-14)
    # non-square (6 unit base strains x V), as the stress plugin uses
    base = np.zeros((6, 3, 3))
    for i, (a, b) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        base[i, a, b] = base[i, b, a] = 1.0
    g2 = gsk.btensor_full(n, dx, kconv, C, base, e)
    o2 = O.btensor(n, C, base, e, dx=dx, kconv=kconv)
    np.testing.assert_allclose(g2, o2, rtol=0, atol=1e-13)


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [16, 32, 64, 96, 128, 192, 256])
def test_fft_r2c_c2r_vs_numpy(gpu, n):
    say(f"fft n={n}")
    rng = np.random.default_rng(n)
    cnt = 2 if n <= 128 else 1
    x = rng.standard_normal((cnt, n, n, n)).astype(np.float32)
    X = gsk.fft_r2c(x)
    want = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3))
    err = np.abs(X - want).max() / np.abs(want).max()
    assert err < 2e-6, err
    y = gsk.fft_c2r(X) / n**3
    assert np.abs(y - x).max() < 2e-5
    # C2R of an exact spectrum (Hermitian-consistent input)
    z = gsk.fft_c2r(want.astype(np.complex64)) / n**3
    assert np.abs(z - x).max() < 2e-5


def test_fft_512_roundtrip_and_impulse(gpu):
    say("fft n=512")
    n = 512
    x = np.zeros((1, n, n, n), np.float32)
    x[0, 3, 5, 7] = 1.0
    X = gsk.fft_r2c(x)
    k = np.arange(n)
    # impulse -> plane wave, checked on a few slices
    for kx in (0, 1, 255, 256, 511):
        ph = np.exp(-2j * np.pi * (3 * kx + 5 * k[:, None] + 7 * np.arange(n // 2 + 1)[None, :]) / n)
        assert np.abs(X[0, kx] - ph).max() < 5e-6
    y = gsk.fft_c2r(X) / n**3
    assert np.abs(y - x).max() < 1e-6


# ------------------------------------------------------------------ martensite
def run_mart_pair(n, V, steps, path, **kw):
    s = mart_sim(n, V, path=path, **kw)
    s.init_uniform("eta", -0.01, 0.01, 0)
    eta0 = s.get_field("eta").astype(np.float64)
    s.step(steps)
    got = s.get_field("eta", np.float64)
    E = s.energy()
    cfg = s.cfg
    s.close()
    p = oracle_mart_params(cfg)
    B = O.btensor(n, c_cfg0(), gsk.mart_eigenstrains(V))
    want = O.mart_run(eta0, B, p, steps)
    Ew = O.mart_energy(want, B, p)
    return got, want, E, Ew


@pytest.mark.parametrize("n,steps", [(16, 50), (32, 200)])
def test_mart_fused32_matches_oracle(gpu, n, steps):
    say(f"mart fused n={n} steps={steps}")
    got, want, E, Ew = run_mart_pair(n, 3, steps, gsk.PATH_FUSED32)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() / scale < 2e-5
    assert rel_l2(got, want) < 2e-5
    np.testing.assert_allclose(E, Ew, rtol=1e-4, atol=1e-9 * abs(Ew[3]))


@pytest.mark.parametrize("n,steps", [(16, 50), (32, 100)])
def test_mart_plain64_matches_oracle(gpu, n, steps):
    say(f"mart plain64 n={n} steps={steps}")
    got, want, E, Ew = run_mart_pair(n, 3, steps, gsk.PATH_PLAIN64)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-11
    np.testing.assert_allclose(E, Ew, rtol=1e-10)


def test_mart_strong_dynamics_with_noise(gpu):
    """Martensite below the transition (A < 0) with noise: nonlinear and elastic terms matter."""
    say("mart fused n=32 A=-0.5 noise")
    n, V, steps = 32, 3, 100
    kw = dict(A=-0.5, B=3.0, C=3.0, noise=0.002, seed=9, w_el=20.0)
    for path, tol in ((gsk.PATH_FUSED32, 5e-5), (gsk.PATH_PLAIN64, 1e-10)):
        got, want, E, Ew = run_mart_pair(n, V, steps, path, **kw)
        assert np.abs(got - want).max() / np.abs(want).max() < tol, path
        np.testing.assert_allclose(E, Ew, rtol=max(tol * 10, 1e-10), atol=1e-9 * abs(Ew[3]))


@pytest.mark.parametrize("n,V", [(32, 1), (32, 2), (32, 4), (64, 2), (128, 1)])
def test_mart_fused32_variant_counts(gpu, n, V):
    """The fused step with 1, 2 and 4 variants (at 32^3, 64^3 and 128^3 the plane-resident
    passes, instantiated per variant count) against the FP64 oracle, below the transition,
    with elasticity and noise."""
    say(f"mart fused n={n} V={V} A=-0.5 noise")
    kw = dict(A=-0.5, B=3.0, C=3.0, noise=0.002, seed=13, w_el=20.0)
    got, want, E, Ew = run_mart_pair(n, V, 30, gsk.PATH_FUSED32, **kw)
    assert np.abs(got - want).max() / np.abs(want).max() < 5e-5
    assert rel_l2(got, want) < 5e-5
    np.testing.assert_allclose(E, Ew, rtol=5e-4, atol=1e-9 * abs(Ew[3]))


@pytest.mark.parametrize("n", [64, 96, 256, 512])
def test_mart_noise_enters_exactly(gpu, n):
    """With every other force zero (A = B = C = 0, no elasticity, kappa = 0), eta = 0 and
    dt = L = 1, one step gives eta = -a xi with xi the Philox normals of (seed, step 0,
    variant, voxel), which noise_field reproduces bit for bit. Checks that each voxel of the
    middle z pass gets its own normal (its lanes share Philox blocks at 64, 256, 512)."""
    say(f"mart noise placement n={n}")
    V, seed = 3, 21
    s = mart_sim(n, V, A=0.0, B=0.0, C=0.0, w_el=0.0, kappa=0.0, L=1.0, dt=1.0, noise=1.0, seed=seed)
    s.set_field("eta", np.zeros((V, n, n, n), np.float32))
    s.step(1)
    got = s.get_field("eta")
    s.close()
    for v in range(V):
        want = -gsk.noise_field(seed, 0, v, n**3).reshape(n, n, n)
        assert float(np.abs(got[v] - want).max()) < 1e-5 * float(np.abs(want).max()), v


def test_mart_linear_mode_at_256(gpu):
    """Full-size property: with B = C = 0 and no elasticity, a single mode decays by
    ((1 - dt L A)/(1 + dt L kappa k^2))^steps exactly."""
    say("mart linear mode n=256")
    n, steps = 256, 20
    s = mart_sim(n, 1, w_el=0.0, B=0.0, C=0.0, A=0.3, kappa=1.5, L=0.7, dt=0.05)
    x = np.arange(n, dtype=np.float64)
    m = (5, 17, 100)
    wave = 1e-3 * np.cos(2 * np.pi * (m[0] * x[:, None, None] + m[1] * x[None, :, None] + m[2] * x[None, None, :]) / n)
    s.set_field("eta", wave[None].astype(np.float32))
    s.step(steps)
    got = s.get_field("eta", np.float64)[0]
    s.close()
    k2 = sum((2 * np.pi * mi / n) ** 2 for mi in m)
    g = (1 - 0.05 * 0.7 * 0.3) / (1 + 0.05 * 0.7 * 1.5 * k2)
    assert np.abs(got - wave * g**steps).max() < 1e-3 * 1e-4


def test_mart_translation_equivariance_at_256(gpu):
    say("mart translation n=256")
    n, V = 256, 3
    s = mart_sim(n, V, A=-0.5, w_el=20.0)
    s.init_uniform("eta", -0.05, 0.05, 3)
    e0 = s.get_field("eta")
    s.step(5)
    a = s.get_field("eta")
    sh = (17, 3, 101)
    s.set_field("eta", np.roll(e0, sh, axis=(1, 2, 3)))
    s.set_step(0)
    s.step(5)
    b = s.get_field("eta")
    s.close()
    assert np.abs(np.roll(a, sh, axis=(1, 2, 3)) - b).max() < 1e-5 * np.abs(a).max()


@pytest.mark.parametrize("n,V", [(192, 3), (192, 1), (192, 4), (192, 2), (128, 3), (128, 1), (128, 4)])
def test_split_schedule_matches_chain_at_192(gpu, monkeypatch, n, V):
    """The split-y schedule at n = 192 (sub-planes of 48 rows, 384-thread sub-plane groups, x
    transforms 12 * 16 whose last pass leaves values in 384 of 512 threads) and n = 128 (32 rows,
    256-thread groups, rows of 32 voxel quads) against what GSK_SPLIT=0 runs (the chain of y and
    z kernels; the plane-resident kernels at 128): 20 steps of strong dynamics with noise agree
    to FP32 rounding, and so do the energies."""
    say(f"split-y schedule vs y/z chain n={n} V={V}")
    kw = dict(A=-0.5, w_el=20.0, noise=0.002, seed=9)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("GSK_SPLIT", mode)
        s = mart_sim(n, V, **kw)
        s.init_uniform("eta", -0.05, 0.05, 6)
        s.step(20)
        out[mode] = (s.get_field("eta", np.float64), s.energy())
        names, _ = s.kernel_times(1)
        s.close()
        assert ("sub_mid" in names) == (mode == "1"), names
    (a, Ea), (b, Eb) = out["1"], out["0"]
    assert np.isfinite(a).all()
    assert float(np.abs(a - b).max()) < 2e-6 * float(np.abs(b).max())
    assert rel_l2(a, b) < 1e-6
    np.testing.assert_allclose(Ea, Eb, rtol=1e-5, atol=1e-9 * abs(Eb[3]))


@pytest.mark.parametrize("V", [3, 1, 2, 4])
def test_split_schedule_matches_chain_at_256(gpu, monkeypatch, V):
    """At 256^3 the fused martensite step runs the split-y schedule (four launches; sub-plane
    kernels keep 64 y-rows in LDS through y, z and y) unless GSK_SPLIT=0 is set when the handle
    is created, which keeps the chain of y and z kernels (eight launches). The two FP32
    schedules order their arithmetic differently but follow one trajectory: 20 steps of strong
    dynamics with noise agree to FP32 rounding, and so do the energies. Every variant count of
    the split kernels (1..4; the contraction keeps V - 1 fields in registers and one in LDS)."""
    say(f"split-y schedule vs y/z chain n=256 V={V}")
    kw = dict(A=-0.5, w_el=20.0, noise=0.002, seed=9)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("GSK_SPLIT", mode)
        s = mart_sim(256, V, **kw)
        s.init_uniform("eta", -0.05, 0.05, 6)
        s.step(20)
        out[mode] = (s.get_field("eta", np.float64), s.energy())
        names, _ = s.kernel_times(1)  # which schedule ran
        s.close()
        assert ("sub_mid" in names) == (mode == "1"), names
    (a, Ea), (b, Eb) = out["1"], out["0"]
    assert np.isfinite(a).all()
    assert float(np.abs(a - b).max()) < 2e-6 * float(np.abs(b).max())
    assert rel_l2(a, b) < 1e-6
    np.testing.assert_allclose(Ea, Eb, rtol=1e-5, atol=1e-9 * abs(Eb[3]))


# ------------------------------------------------------------------ precipitate
def test_mart_fused_vs_plain_at_512(gpu):
    """512^3 with the 512^3 configuration's noise (0.002), strong dynamics: the FP32 fused step
    against the FP64 plain path, an independent implementation (C2C transforms, full-layout
    FP64 B tensor). Catches any wrong B entry or transform at the largest grid, where the oracle
    is too slow. FP32 tolerance: a few steps keep the relative error near 1e-6."""
    say("mart fused vs plain n=512")
    n, V, steps = 512, 3, 4
    kw = dict(A=-0.5, w_el=20.0, noise=0.002, seed=11)
    out = []
    for path in (gsk.PATH_FUSED32, gsk.PATH_PLAIN64):
        s = mart_sim(n, V, path=path, **kw)
        s.init_uniform("eta", -0.05, 0.05, 4)
        s.step(steps)
        out.append(s.get_field("eta"))
        s.close()
    got, want = out
    assert np.isfinite(got).all()
    assert rel_l2(got.astype(np.float64), want.astype(np.float64)) < 2e-6
    assert float(np.abs(got - want).max()) < 2e-6 * float(np.abs(want).max())


@pytest.mark.parametrize("n", [64, 96, 128, 192, 256])
def test_mart_square_carry_bit_identical(gpu, n):
    """A step leaves the transformed eta^2 of its result for the next step (along z and y at
    n = 64, where the plane-resident passes run; along z at n = 96 and 256, or along z and y at
    256 in the sub-plane kernels of the split-y schedule, GSK_SPLIT=1). Steps that run back to back must
    give bit for bit the trajectory of steps separated by energy reductions and a set_step to
    the current index (which drops the carried transform, so the next step transforms eta^2
    afresh)."""
    say(f"mart eta^2 carry-over n={n}")
    kw = dict(A=-0.5, w_el=20.0, noise=0.002, seed=5)
    a = mart_sim(n, 3, **kw)
    b = mart_sim(n, 3, **kw)
    for s in (a, b):
        s.init_uniform("eta", -0.05, 0.05, 2)
    a.step(6)
    for k in (2, 1, 3):
        b.step(k)
        b.energy()
        b.set_step(b.step_index)
    ea, eb = a.get_field("eta"), b.get_field("eta")
    a.close()
    b.close()
    assert np.array_equal(ea.view(np.uint32), eb.view(np.uint32))


@pytest.mark.parametrize("n", [64, 128, 192, 256])
def test_prec_local_force_carry_bit_identical(gpu, n):
    """The precipitate step's last pass also forms the next step's local forces (with the next
    step's noise): the last z pass at n = 64, the sub-plane local-force kernel of the split-y
    schedule at 256. Steps back to back must give bit for bit the trajectory of steps whose
    carried-over forces are dropped in between (a parameter set to its own value)."""
    say(f"prec local-force carry-over n={n}")
    a, b = prec_sim(n), prec_sim(n)
    for s in (a, b):
        s.init_uniform("concentration", 0.245, 0.255, 11)
        s.init_uniform("eta", 0.0, 0.01, 12)
    a.step(5)
    for k in (2, 1, 2):
        b.step(k)
        b.set_param("M", 1.0)
    fa = [a.get_field(nm) for nm in ("concentration", "eta")]
    fb = [b.get_field(nm) for nm in ("concentration", "eta")]
    a.close()
    b.close()
    for x, y in zip(fa, fb):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("n", [96, 128, 192])
def test_mart_fused_vs_plain_mixed_sizes(gpu, n):
    """The fused step at the radix-3 sizes (96 = 3*32, 192 = 3*64) and at 128 against the FP64
    plain path (pinned by the oracle), strong dynamics with noise, a few steps."""
    say(f"mart fused vs plain n={n}")
    kw = dict(A=-0.5, w_el=20.0, noise=0.002, seed=13)
    out = []
    for path in (gsk.PATH_FUSED32, gsk.PATH_PLAIN64):
        s = mart_sim(n, 3, path=path, **kw)
        s.init_uniform("eta", -0.05, 0.05, 6)
        s.step(6)
        out.append(s.get_field("eta").astype(np.float64))
        s.close()
    got, want = out
    assert rel_l2(got, want) < 2e-6
    assert float(np.abs(got - want).max()) < 2e-6 * float(np.abs(want).max())


def test_mart_two_handles_bit_identical_at_512(gpu):
    """Two handles created one after the other give bit-identical trajectories (every buffer
    is ordered on the handle's own stream, including its zero fill at creation)."""
    say("mart determinism n=512")
    res = []
    for _ in range(2):
        s = mart_sim(512, 3, A=-0.5, w_el=20.0)
        s.init_uniform("eta", -0.05, 0.05, 4)
        s.step(2)
        res.append(s.get_field("eta"))
        s.close()
    assert np.array_equal(res[0].view(np.uint32), res[1].view(np.uint32))


def prec_sim(n, V=4, **kw):
    base = dict(n=n, nvar=V, M=1.0, L=1.0, kappa=1.5, kappa_c=1.5, dt=0.005, noise=0.001, seed=5, S=54.0)
    base.update(kw)
    cfg = gsk.default_config(gsk.PHYS_PREC, **base)
    s = gsk.Sim(cfg)
    s.build_btensor(cubic(3.5, 1.5, 1.0), 0.02 * np.eye(3)[None])
    return s


def oracle_prec_params(cfg):
    return O.prec_params(A=list(cfg.a), c1=cfg.c1, c2=cfg.c2, M=cfg.M, L=cfg.L, kappa_c=cfg.kappa_c,
                         kappa_eta=cfg.kappa, xi=cfg.xi, S=cfg.S, dt=cfg.dt, noise=cfg.noise, dx=cfg.dx,
                         kconv=cfg.kconv, seed=cfg.seed)


def prec_init(s, seed=11):
    s.init_uniform("concentration", 0.245, 0.255, seed)
    s.init_uniform("eta", 0.0, 0.01, seed + 1)


@pytest.mark.parametrize("n,steps", [(16, 50), (32, 200), (96, 20), (256, 3)])
def test_prec_fused32_matches_oracle(gpu, n, steps):
    say(f"prec fused n={n} steps={steps}")
    s = prec_sim(n)
    prec_init(s)
    c0 = s.get_field("concentration", np.float64)
    e0 = s.get_field("eta", np.float64)
    m0, r0 = s.mass(), s.mass_real()
    s.step(steps)
    c1 = s.get_field("concentration", np.float64)
    e1 = s.get_field("eta", np.float64)
    E = s.energy()
    m1, r1 = s.mass(), s.mass_real()
    cfg = s.cfg
    s.close()
    p = oracle_prec_params(cfg)
    Bc = O.btensor(n, cubic(3.5, 1.5, 1.0), 0.02 * np.eye(3)[None])[0, 0]
    cw, ew = O.prec_run(c0, e0, Bc, p, steps)
    assert np.abs(c1 - cw).max() < 2e-6
    assert np.abs(e1 - ew).max() / np.abs(ew).max() < 2e-5
    # FP64 mean + exact k = 0 state: conserved to the last bit
    assert m1 == m0
    assert abs(m0 - c0.sum()) / c0.sum() < 1e-12
    # FP64 sum of the FP32 field: one IFFT's rounding (bias ~1e-2 ulp of dc per voxel), not accumulated
    assert abs(r1 - m0) / m0 < 2e-10 and abs(r0 - m0) / m0 < 2e-10
    assert abs(cw.sum() - c0.sum()) / c0.sum() < 1e-13
    Ew = O.prec_energy(cw, ew, Bc, p)
    np.testing.assert_allclose(E, Ew, rtol=1e-5, atol=1e-9 * abs(Ew[3]))


@pytest.mark.parametrize("n,V", [(256, 4), (256, 1), (256, 2), (192, 4), (192, 1), (128, 4)])
def test_prec_split_schedule_matches_chain(gpu, monkeypatch, n, V):
    """At 256^3 and 192^3 the precipitate step runs the split-y schedule (x update, sub-plane inverse,
    sub-plane local forces: three launches) unless GSK_SPLIT=0 is set when the handle is
    created (the chain of y and z kernels, four launches). Two FP32 orderings of one step: 20
    steps with noise agree to FP32 rounding in c, eta and the energies; both conserve sum c to the
    last bit. V order parameters: 2..5 fields of the local-force kernel."""
    say(f"prec split-y schedule vs y/z chain n={n} V={V}")
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("GSK_SPLIT", mode)
        s = prec_sim(n, V=V, noise=0.01)
        prec_init(s)
        m0 = s.mass()
        s.step(20)
        out[mode] = (s.get_field("concentration", np.float64), s.get_field("eta", np.float64), s.energy(), m0,
                     s.mass())
        names, _ = s.kernel_times(1)
        s.close()
        assert ("sub_local" in names) == (mode == "1"), names
    (ca, ea, Ea, m0a, m1a), (cb, eb, Eb, m0b, m1b) = out["1"], out["0"]
    assert np.isfinite(ea).all() and np.isfinite(ca).all()
    assert m1a == m0a and m1b == m0b and m0a == m0b
    assert float(np.abs(ca - cb).max()) < 1e-6
    assert float(np.abs(ea - eb).max()) < 2e-6 * float(np.abs(eb).max())
    assert rel_l2(ea, eb) < 1e-6
    np.testing.assert_allclose(Ea, Eb, rtol=1e-5, atol=1e-9 * abs(Eb[3]))


def test_prec_mass_conservation_at_256(gpu):
    say("prec mass n=256")
    s = prec_sim(256)
    prec_init(s)
    m0 = s.mass()
    s.step(40)
    m1, r1 = s.mass(), s.mass_real()
    e = s.get_field("eta")
    s.close()
    assert np.isfinite(e).all()
    assert m1 == m0
    assert abs(r1 - m0) / m0 < 1e-10


def test_resume_is_bit_exact(gpu):
    """Counter-based noise: stopping, copying the fields out and back, and resuming at the
    same step index reproduces an uninterrupted run bit for bit (FP64 plain path)."""
    say("mart plain64 resume n=32")
    n, V = 32, 3
    a = mart_sim(n, V, path=gsk.PATH_PLAIN64, noise=0.002, seed=4, A=-0.5)
    a.init_uniform("eta", -0.01, 0.01, 0)
    e0 = a.get_field("eta", np.float64)
    a.step(10)
    full = a.get_field("eta", np.float64)
    b = mart_sim(n, V, path=gsk.PATH_PLAIN64, noise=0.002, seed=4, A=-0.5)
    b.set_field("eta", e0)
    b.step(6)
    mid = b.get_field("eta", np.float64)
    c = mart_sim(n, V, path=gsk.PATH_PLAIN64, noise=0.002, seed=4, A=-0.5)
    c.set_field("eta", mid)
    c.set_step(6)
    c.step(4)
    assert np.array_equal(c.get_field("eta", np.float64), full)
    for x in (a, b, c):
        x.close()


@pytest.mark.parametrize("n", [64, 256])
def test_fft_agrees_with_rocfft(gpu, n):
    """The library's R2C matches rocFFT's on the same data (rocFFT is only a comparator)."""
    say(f"fft vs rocfft n={n}")
    r = gsk.fft_compare(n, count=1, reps=3)
    assert r["max_rel_diff"] < 2e-6, r
    assert r["hand_r2c_ms"] > 0 and r["rocfft_r2c_ms"] > 0


def test_mart_fused32_general_eigenstrains(gpu):
    """Eigenstrains with off-diagonal entries make B(k) lose its reflection symmetries, so the
    Nyquist planes of the half spectra must carry the average of B at +-n/2 to reproduce the
    reference's real part of a full complex product (the oracle's convention)."""
    say("mart fused32 with general eigenstrains")
    n, V, steps = 32, 3, 40
    rng = np.random.default_rng(17)
    e = rng.uniform(-0.03, 0.03, (V, 3, 3))
    eps = 0.5 * (e + e.transpose(0, 2, 1))
    C = gsk.cubic_modulus(0.7, 0.3, 0.2)
    cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, w_el=50.0, A=-0.2)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", -0.1, 0.1, 4)
        eta0 = s.get_field("eta").astype(np.float64)
        s.step(steps)
        got = s.get_field("eta", np.float64)
    want = O.mart_run(eta0, O.btensor(n, C, eps), oracle_mart_params(cfg), steps)
    assert rel_l2(got, want) < 2e-5, rel_l2(got, want)


@pytest.mark.parametrize("n,steps", [(16, 40), (32, 50)])
def test_prec_plain64_matches_oracle(gpu, n, steps):
    """The precipitate spectral rule on the FP64 plain path (one kernel per operation, C2C
    transforms, the full-layout FP64 B(k)) against the FP64 oracle, with eta noise: 1e-11."""
    say(f"prec plain64 n={n} steps={steps}")
    s = prec_sim(n, path=gsk.PATH_PLAIN64)
    prec_init(s)
    c0 = s.get_field("concentration", np.float64)
    e0 = s.get_field("eta", np.float64)
    m0 = s.mass()
    s.step(steps)
    c1 = s.get_field("concentration", np.float64)
    e1 = s.get_field("eta", np.float64)
    E = s.energy()
    m1 = s.mass()
    cfg = s.cfg
    s.close()
    p = oracle_prec_params(cfg)
    Bc = O.btensor(n, cubic(3.5, 1.5, 1.0), 0.02 * np.eye(3)[None])[0, 0]
    cw, ew = O.prec_run(c0, e0, Bc, p, steps)
    assert np.abs(c1 - cw).max() < 1e-11 * np.abs(cw).max()
    assert np.abs(e1 - ew).max() < 1e-11 * np.abs(ew).max()
    assert abs(m0 - c0.sum()) / c0.sum() < 1e-14
    assert abs(m1 - m0) / m0 < 1e-13
    Ew = O.prec_energy(cw, ew, Bc, p)
    np.testing.assert_allclose(E, Ew, rtol=1e-10, atol=1e-12 * abs(Ew[3]))


@pytest.mark.parametrize("physics,V", [("mart", 3), ("mart", 2), ("prec", 4), ("prec", 3)])
def test_plain_kernel_times_follow_the_paired_step(gpu, physics, V):
    """Per-kernel timing on the FP64 plain path names one slot per launch of the paired step: a
    slot list that missed a launch (or kept a removed one) fails with a slot count mismatch."""
    say(f"plain64 kernel_times physics={physics} V={V}")
    n = 32
    s = mart_sim(n, V, path=gsk.PATH_PLAIN64) if physics == "mart" else prec_sim(n, V, path=gsk.PATH_PLAIN64)
    with s:
        names, ms = s.kernel_times(2)
        nbytes = s.kernel_bytes()
    P = (V + 1) // 2
    if physics == "mart":
        want = (["p_square_pair"] + ["p_z_fwdS", "p_y_fwdS", "p_x_fwdS"] * P + ["p_contract_pair"]
                + ["p_z_invR", "p_y_invR", "p_x_invR"] * P + ["p_force_pair"] + ["p_z_fwdG", "p_y_fwdG", "p_x_fwdG"] * V
                + ["p_update_pair"] + ["p_z_invE", "p_y_invE", "p_x_invE"] * P + ["p_extract_pair"])
    else:
        F = V + 1
        want = (["p_prec_local"] + ["p_z_fwdG", "p_y_fwdG", "p_x_fwdG"] * F + ["p_prec_update_pair"]
                + ["p_z_invE", "p_y_invE", "p_x_invE"] * ((F + 1) // 2) + ["p_extract_pair"])
    assert names == want
    assert len(ms) == len(want) and (ms > 0).all()
    assert len(nbytes) == len(want) and (nbytes > 0).all()

"""Pins the FP64 oracle (CPU, no GPU needed).

The reference (General-Simulator) ships no tests, fixtures or golden vectors for
this path, so each oracle routine is pinned here by an independent source:
numpy.fft (pocketfft) vectors, brute-force DFTs, published Philox KATs plus
rocRAND's implementation, closed forms of the Khachaturyan tensor evaluated with
numpy.linalg, finite differences of the energy, and exact linear-mode decay.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def fftg():
    return np.load(os.path.join(GOLD, "fft_golden.npz"))


# ------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 3, 4, 8, 16, 32, 48, 64, 96, 128, 192, 256, 512])
def test_fft1d_matches_pocketfft(fftg, n):
    x = fftg[f"c1d_{n}_in"]
    np.testing.assert_allclose(O.fft1d(x, -1), fftg[f"c1d_{n}_fwd"], rtol=0, atol=1e-11 * n)
    np.testing.assert_allclose(O.fft1d(x, +1), fftg[f"c1d_{n}_bwd"], rtol=0, atol=1e-11 * n)


@pytest.mark.parametrize("n", [3, 5, 6, 8, 12, 16])
def test_fft1d_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    D = np.exp(-2j * np.pi * np.outer(j, j) / n)
    np.testing.assert_allclose(O.fft1d(x, -1), D @ x, atol=1e-12 * n)


def test_fft3d_golden(fftg):
    np.testing.assert_allclose(O.fft3d(fftg["c3d_6_in"]), fftg["c3d_6_fftn"], atol=1e-11)
    for n in (8, 16):
        r = fftg[f"r3d_{n}_in"]
        want = fftg[f"r3d_{n}_rfftn"]
        for f in range(r.shape[0]):
            got = O.fft3d(r[f].astype(np.complex128))[:, :, : n // 2 + 1]
            np.testing.assert_allclose(got, want[f], atol=1e-10)


def test_fft3d_impulse_plane_wave_parseval():
    n = 16
    a = np.zeros((n, n, n), complex)
    a[3, 5, 7] = 1.0
    F = O.fft3d(a)
    k = np.arange(n)
    expect = np.exp(-2j * np.pi * (3 * k[:, None, None] + 5 * k[None, :, None] + 7 * k[None, None, :]) / n)
    np.testing.assert_allclose(F, expect, atol=1e-12)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    X = O.fft3d(x)
    assert abs(np.sum(abs(X) ** 2) / n**3 - np.sum(abs(x) ** 2)) < 1e-9 * np.sum(abs(x) ** 2)
    np.testing.assert_allclose(O.fft3d(X, +1) / n**3, x, atol=1e-12)


# ------------------------------------------------------------------------- Philox + normals
PUBLISHED_KAT = [  # Random123 kat_vectors, philox4x32 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def test_philox_published_kat():
    for ctr, key, out in PUBLISHED_KAT:
        assert list(O.philox(ctr, key)) == out


def test_philox_matches_rocrand():
    with open(os.path.join(GOLD, "philox_rocrand.json")) as f:
        vecs = json.load(f)
    assert len(vecs) == 64
    for v in vecs:
        assert list(O.philox(v["ctr"], v["key"])) == v["out"]


def test_normal_components_accuracy():
    us = np.concatenate([np.linspace(2.0**-24, 1.0, 20001, dtype=np.float32), np.float32(2.0) ** -np.arange(0, 25)])
    err = max(abs(O.log_f(float(u)) - math.log(float(u))) for u in us)
    assert err < 2e-6
    xs = np.linspace(1e-6, 34.0, 5001, dtype=np.float32)
    rel = max(abs(O.sqrt_f(float(x)) - math.sqrt(float(x))) / math.sqrt(float(x)) for x in xs)
    assert rel < 1e-6
    assert O.sqrt_f(0.0) == 0.0
    for t in np.linspace(0.0, math.pi / 2 * (1 - 2**-22), 4001, dtype=np.float32):
        s, c = O.sincos_f(float(t))
        assert abs(s - math.sin(float(t))) < 1e-6 and abs(c - math.cos(float(t))) < 1e-6


def test_normals_statistics():
    from scipy import stats
    z = O.noise_field(seed=7, step=3, field=1, nvox=1 << 21).astype(np.float64)
    assert abs(z.mean()) < 5e-3
    assert abs(z.var() - 1.0) < 5e-3
    assert abs(stats.skew(z)) < 1e-2 and abs(stats.kurtosis(z)) < 2e-2
    ks = stats.kstest(z[:200000], "norm")
    assert ks.pvalue > 1e-3
    # different (step, field, seed) give different streams
    z2 = O.noise_field(seed=7, step=4, field=1, nvox=1024)
    assert not np.array_equal(z[:1024].astype(np.float32), z2)


def test_uniform_field():
    u = O.uniform_field(seed=0, field=2, nvox=1 << 16, lo=-0.01, hi=0.01)
    assert u.dtype == np.float32 and u.min() >= -0.01 and u.max() < 0.01
    assert abs(u.mean()) < 1e-4
    bits = O.philox([5, 2, 0, 1], [0, 0])  # group 5 of field 2, purpose 1
    want = (-0.01 + 0.02 * ((bits >> 8).astype(np.float64) * 2.0**-24)).astype(np.float32)
    np.testing.assert_array_equal(u[20:24], want)


# ------------------------------------------------------------------------- B tensor
def cubic(c11, c12, c44):
    C = np.zeros((3, 3, 3, 3))
    for i in range(3):
        C[i, i, i, i] = c11
        for j in range(3):
            if i != j:
                C[i, i, j, j] = c12
                C[i, j, i, j] = C[i, j, j, i] = c44
    return C


def mart_eps(V=3):
    e = np.zeros((V, 3, 3))
    for p in range(V):
        e[p] = -0.04 * np.eye(3)
        e[p, p % 3, p % 3] += 0.12
    return e


def b_numpy(C, e1, e2, nvec):
    """Khachaturyan B at one unit vector using einsum and numpy.linalg.inv."""
    cee = np.einsum("ijkl,pij,qkl->pq", C, e1, e2)
    if np.linalg.norm(nvec) == 0:
        return cee
    s1 = np.einsum("ijkl,pkl->pij", C, e1)
    s2 = np.einsum("ijkl,qkl->qij", C, e2)
    G = np.linalg.inv(np.einsum("ikjl,k,l->ij", C, nvec, nvec))
    return cee - np.einsum("i,pij,jk,qkl,l->pq", nvec, s1, G, s2, nvec)


def test_btensor_closed_forms():
    C = cubic(0.7, 0.3, 0.2)
    e = mart_eps()
    n = 8
    B = O.btensor(n, C, e)
    # k = 0: C eps eps
    np.testing.assert_allclose(B[:, :, 0, 0, 0], np.einsum("ijkl,pij,qkl->pq", C, e, e), atol=1e-14)
    # [100]: Cee - s_xx^p s_xx^q / C11 (diagonal eigenstrains)
    s = np.einsum("ijkl,pkl->pij", C, e)
    want = np.einsum("ijkl,pij,qkl->pq", C, e, e) - np.outer(s[:, 0, 0], s[:, 0, 0]) / 0.7
    np.testing.assert_allclose(B[:, :, 1, 0, 0], want, atol=1e-14)
    # general directions against numpy.linalg
    for (x, y, z) in [(1, 1, 0), (1, 1, 1), (2, 1, 3), (7, 1, 5), (4, 4, 4), (3, 6, 2)]:
        m = np.array([x if x < n // 2 else x - n, y if y < n // 2 else y - n, z if z < n // 2 else z - n], float)
        u = m / np.linalg.norm(m)
        np.testing.assert_allclose(B[:, :, x, y, z], b_numpy(C, e, e, u), atol=1e-13)


def test_btensor_symmetries():
    C = cubic(3.5, 1.5, 1.0)
    e = mart_eps()
    n = 8
    B = O.btensor(n, C, e)
    np.testing.assert_allclose(B, np.transpose(B, (1, 0, 2, 3, 4)), atol=1e-14)
    idx = (-np.arange(n)) % n
    np.testing.assert_allclose(B, B[:, :, idx][:, :, :, idx][:, :, :, :, idx], atol=1e-14)
    # cubic symmetry: swapping the x and y axes swaps variants 0 and 1
    Bs = np.transpose(B, (0, 1, 3, 2, 4))[[1, 0, 2]][:, [1, 0, 2]]
    np.testing.assert_allclose(B, Bs, atol=1e-14)


def test_btensor_isotropic_dilatational_is_direction_free():
    c11, c12 = 3.0, 1.0
    C = cubic(c11, c12, (c11 - c12) / 2)
    e = 0.02 * np.eye(3)[None]
    B = O.btensor(8, C, e)[0, 0]
    nz = B.reshape(-1)[1:]
    s0 = (c11 + 2 * c12) * 0.02
    want = 3 * s0 * 0.02 - s0**2 / c11
    np.testing.assert_allclose(nz, want, rtol=1e-12)


def test_btensor_reference_wavevector_convention():
    # reference form 2*3.14*dx*m/n: direction only changes when dx differs per axis, so
    # check the k-component helper directly
    assert O.kcomp(3, 16, 0.5, 1) == pytest.approx(2 * 3.14 * 0.5 / 16 * 3, rel=1e-15)
    assert O.kcomp(12, 16, 1.0, 0) == pytest.approx(2 * math.pi * (12 - 16) / 16, rel=1e-15)


# ------------------------------------------------------------------------- update rules
def test_mart_force_is_energy_derivative():
    n, V = 8, 3
    rng = np.random.default_rng(3)
    eta = rng.uniform(-0.3, 0.5, (V, n, n, n))
    B = O.btensor(n, cubic(0.7, 0.3, 0.2), mart_eps())
    p = O.mart_params(w_el=5.0)
    F = O.mart_force(eta, B, p)
    h = 1e-5
    for (v, x, y, z) in [(0, 1, 2, 3), (1, 7, 0, 4), (2, 5, 5, 5)]:
        ep, em = eta.copy(), eta.copy()
        ep[v, x, y, z] += h
        em[v, x, y, z] -= h
        Ep, Em = O.mart_energy(ep, B, p), O.mart_energy(em, B, p)
        fd = ((Ep[0] + Ep[2]) - (Em[0] + Em[2])) / (2 * h)
        assert fd == pytest.approx(F[v, x, y, z], rel=1e-6, abs=1e-9)


def test_prec_local_is_energy_derivative():
    p = O.prec_params()
    rng = np.random.default_rng(4)
    for V in (1, 2, 3, 4):
        eta = rng.uniform(-0.8, 0.9, V)
        c = 0.3
        mu, F, f = O.prec_local(c, eta, p)
        h = 1e-6
        assert (O.prec_local(c + h, eta, p)[2] - O.prec_local(c - h, eta, p)[2]) / (2 * h) == pytest.approx(mu, rel=1e-7)
        for v in range(V):
            ep, em = eta.copy(), eta.copy()
            ep[v] += h
            em[v] -= h
            fd = (O.prec_local(c, ep, p)[2] - O.prec_local(c, em, p)[2]) / (2 * h)
            assert fd == pytest.approx(F[v], rel=1e-6, abs=1e-9)


def test_mart_linear_mode_exact_decay():
    n, V = 16, 1
    p = O.mart_params(A=0.3, B=0.0, C=0.0, w_el=0.0, kappa=1.5, L=0.7, dt=0.05)
    x = np.arange(n)
    m = (2, 1, 3)
    eps = 1e-3
    eta = eps * np.cos(2 * np.pi * (m[0] * x[:, None, None] + m[1] * x[None, :, None] + m[2] * x[None, None, :]) / n)
    B = np.zeros((1, 1, n, n, n))
    out = O.mart_run(eta[None], B, p, 40)
    k2 = sum((2 * np.pi * mi / n) ** 2 for mi in m)
    g = (1 - p.dt * p.L * p.A) / (1 + p.dt * p.L * p.kappa * k2)
    np.testing.assert_allclose(out[0], eta * g**40, atol=1e-15)


def test_prec_mass_conserved_and_energy_decreases():
    n, V = 16, 2
    p = O.prec_params(M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5, dt=0.005, xi=400.0, S=54.0)
    rng = np.random.default_rng(5)
    c = 0.25 + rng.uniform(-0.005, 0.005, (n, n, n))
    eta = rng.uniform(0, 0.01, (V, n, n, n))
    Bc = O.btensor(n, cubic(3.5, 1.5, 1.0), 0.02 * np.eye(3)[None])[0, 0]
    m0 = c.sum()
    E0 = O.prec_energy(c, eta, Bc, p)[3]
    c1, e1 = O.prec_run(c, eta, Bc, p, 20)
    assert abs(c1.sum() - m0) / m0 < 1e-13
    E1 = O.prec_energy(c1, e1, Bc, p)[3]
    assert E1 < E0


def test_mart_energy_nonincreasing():
    n, V = 16, 3
    rng = np.random.default_rng(6)
    eta = rng.uniform(-0.3, 0.3, (V, n, n, n))
    B = O.btensor(n, cubic(0.7, 0.3, 0.2), mart_eps())
    p = O.mart_params(A=-0.5, B=3.0, C=3.0, w_el=1.0, dt=0.01)
    E = [O.mart_energy(eta, B, p)[3]]
    for _ in range(5):
        eta = O.mart_run(eta, B, p, 4)
        E.append(O.mart_energy(eta, B, p)[3])
    assert all(b <= a + 1e-12 for a, b in zip(E, E[1:]))


def test_noise_enters_force_linearly():
    n, V = 8, 2
    rng = np.random.default_rng(8)
    eta = rng.uniform(-0.1, 0.1, (V, n, n, n))
    B = np.zeros((V, V, n, n, n))
    p0 = O.mart_params(noise=0.0, seed=11)
    p1 = O.mart_params(noise=0.01, seed=11)
    d = O.mart_run(eta, B, p1, 1, step0=5) - O.mart_run(eta, B, p0, 1, step0=5)
    # d_hat = -dt L a xi_hat / (1 + dt L kappa k^2)
    xi = np.stack([O.noise_field(11, 5, v, n**3).reshape(n, n, n) for v in range(V)]).astype(np.float64)
    k = 2 * np.pi * np.fft.fftfreq(n)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    want = np.real(np.fft.ifftn(-p1.dt * p1.L * 0.01 * np.fft.fftn(xi, axes=(1, 2, 3)) / (1 + p1.dt * p1.L * k2),
                                axes=(1, 2, 3)))
    np.testing.assert_allclose(d, want, atol=1e-15)


def test_prec_reference_rule_unstable_without_stabilisation():
    """The reference CH rule (S = 0) amplifies modes near k^2 ~ 16 at dt=0.005, M=1,
    kappa_c=1.5, A1=54 by (dt M A1 k^2 - 1)/(1 + dt M kappa_c k^4) > 1; S = A1 removes it."""
    n = 16
    p0 = O.prec_params(M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5, dt=0.005, xi=0.0, S=0.0)
    pS = O.prec_params(M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5, dt=0.005, xi=0.0, S=54.0)
    x = np.arange(n)
    m = (6, 6, 5)  # k^2 = 97 (2 pi/16)^2 ~ 15.0
    wave = 1e-4 * np.cos(2 * np.pi * (m[0] * x[:, None, None] + m[1] * x[None, :, None] + m[2] * x[None, None, :]) / n)
    c = 0.25 + wave
    eta = np.zeros((1, n, n, n))
    Bc = np.zeros((n, n, n))
    k2 = sum((2 * np.pi * mi / n) ** 2 for mi in m)
    g0 = (1 - p0.dt * p0.M * p0.A[1] * k2) / (1 + p0.dt * p0.M * p0.kappa_c * k2**2)
    gS = 1 / (1 + pS.dt * pS.M * (pS.S * k2 + pS.kappa_c * k2**2))
    assert abs(g0) > 1.0
    c0, _ = O.prec_run(c, eta, Bc, p0, 10)
    cS, _ = O.prec_run(c, eta, Bc, pS, 10)
    # mu = A1 (c - c1) + A2/2 * 0 is linear: each mode evolves by its exact factor
    np.testing.assert_allclose(c0 - 0.25, wave * g0**10, atol=1e-14)
    np.testing.assert_allclose(cS - 0.25, wave * gS**10, atol=1e-14)


# ------------------------------------------------------------------------- adjoint
def _cubic(c11, c12, c44):
    C = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    if i == j and k == l:
                        C[i, j, k, l] = c11 if i == k else c12
                    elif (i == k and j == l) or (i == l and j == k):
                        C[i, j, k, l] = c44
    return C


def _eps(V, scale=None):
    e = np.zeros((V, 3, 3))
    for p in range(V):
        e[p] = np.eye(3) * -0.04
        e[p, p % 3, p % 3] += 0.12
        if scale is not None:
            e[p] *= scale[p]
    return e


def test_mart_adjoint_matches_central_differences():
    """The store-all adjoint against central differences of the FP64 forward model for every
    gradient it returns: eta0 (random direction), eigenstrain magnitudes s_p, L, kappa.
    The target sits 1e-3 from eta(T) so the loss is small next to its gradients and the
    differences resolve them (h = 1e-6; h = 1e-4 for the s_p gradients, ~1e-7 in size,
    where rounding of the loss would otherwise dominate)."""
    n, V, steps = 16, 3, 8
    rng = np.random.default_rng(4)
    C = _cubic(0.7, 0.3, 0.2)
    p = O.mart_params(A=-0.3, B=3.0, C=3.0, kappa=1.2, L=0.8, dt=0.05, w_el=20.0, noise=0.01, seed=9)
    eta0 = rng.uniform(-0.1, 0.1, (V, n, n, n))
    target = O.mart_run(eta0, O.btensor(n, C, _eps(V)), p, steps) + 1e-3 * rng.standard_normal(eta0.shape)

    def loss(e0=eta0, s=np.ones(V), **kw):
        q = O.mart_params(**{f: getattr(p, f) for f, _ in O.MartParams._fields_})
        for k, v in kw.items():
            setattr(q, k, v)
        eT = O.mart_run(e0, O.btensor(n, C, _eps(V, s)), q, steps)
        return 0.5 * float(((eT - target) ** 2).sum()), eT

    _, eT = loss()
    g_eta, g_par = O.mart_adjoint(eta0, O.btensor(n, C, _eps(V)), p, steps, eT - target)
    h = 1e-6
    d = rng.standard_normal(eta0.shape)
    fd = (loss(e0=eta0 + h * d)[0] - loss(e0=eta0 - h * d)[0]) / (2 * h)
    assert abs(fd - float((g_eta * d).sum())) < 1e-7 * abs(fd), (fd, float((g_eta * d).sum()))
    fd = (loss(L=p.L + h)[0] - loss(L=p.L - h)[0]) / (2 * h)
    assert abs(fd - g_par[V]) < 1e-7 * abs(fd), (fd, g_par[V])
    fd = (loss(kappa=p.kappa + h)[0] - loss(kappa=p.kappa - h)[0]) / (2 * h)
    assert abs(fd - g_par[V + 1]) < 1e-7 * abs(fd), (fd, g_par[V + 1])
    hs = 1e-4
    for q in range(V):
        sp, sm = np.ones(V), np.ones(V)
        sp[q] += hs
        sm[q] -= hs
        fd = (loss(s=sp)[0] - loss(s=sm)[0]) / (2 * hs)
        assert abs(fd - g_par[q]) < 1e-5 * abs(fd), (q, fd, g_par[q])
    # every gradient is non-trivial, so a dropped term cannot hide behind a zero
    assert np.all(np.abs(g_par) > 1e-9) and np.abs(g_eta).max() > 1e-6


# ------------------------------------------------------------------------- reference rules
def test_mart_ref_plane_waves_decay_by_exact_factor():
    """Linear regime of the reference martensite rule (master chemical form with a2 = a3 = 0,
    no elasticity, no noise): a cosine along each axis is an eigenvector of the Grad stencil
    of src/dynamics_mart.cu:181-188 with eigenvalue (2 cos(theta) - 2)/(2 dx)/3, so one step
    multiplies it by 1 + dt (w_grad (2 cos(theta) - 2)/(6 dx) - w_chem a1) exactly."""
    n, dx, dt, wG, wC, a1 = 16, 0.7, 0.01, 2.5, 1.3, 0.4
    x = np.arange(n)
    modes = [(0, 1, 0.3), (1, 2, -0.2), (2, 5, 0.1)]
    eta = np.zeros((1, n, n, n))
    want = np.zeros_like(eta)
    for axis, m, amp in modes:
        th = 2 * np.pi * m / n
        shape = [1, 1, 1]
        shape[axis] = n
        wave = amp * np.cos(th * x).reshape(shape) * np.ones((n, n, n))
        eta[0] += wave
        want[0] += wave * (1 + dt * (wG * (2 * np.cos(th) - 2) / (6 * dx) - wC * a1))
    p = O.mart_ref_params(a1=a1, a2=0.0, a3=0.0, w_grad=wG, w_chem=wC, w_el=0.0, w_noise=0.0, dt=dt, dx=dx,
                          chem_form=1)
    got = O.mart_ref_run(eta, np.zeros((1, 1, n, n, n)), p, 1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)


def test_mart_ref_uniform_states_defect_and_dislocation():
    """Uniform fields (Grad = 0, no elasticity): the HEAD chemical force on each branch,
    defect blocking to zero and linear dislocation decay."""
    n, dt = 8, 0.01
    a1, a2, a3, wC = -1.5, 32.05, 37.5, 1.0
    sc = math.sqrt(abs(a1 / (a2 - a3)))
    e1, e2 = 0.3, -0.2
    eta = np.stack([np.full((n, n, n), e1), np.full((n, n, n), e2)])
    S2 = (sc * e1) ** 2 + (sc * e2) ** 2
    f1 = -wC * (sc * e1) * (a1 - a2 * (sc * e1) ** 2 + a3 * S2)    # eta >= 0 branch
    f2 = 1000.0 * wC * a2 * e2 * e2                                 # eta < 0 branch
    p = O.mart_ref_params(a1=a1, a2=a2, a3=a3, w_grad=2.5, w_chem=wC, w_el=0.0, w_noise=0.0, dt=dt, chem_form=0)
    B0 = np.zeros((2, 2, n, n, n))
    got = O.mart_ref_run(eta, B0, p, 1)
    np.testing.assert_allclose(got[0], e1 + dt * f1, rtol=1e-14)
    np.testing.assert_allclose(got[1], e2 + dt * f2, rtol=1e-14)
    defect = np.zeros((n, n, n))
    defect[1, 2, 3] = 1.0
    got = O.mart_ref_run(eta, B0, p, 1, defect=defect)
    assert got[0, 1, 2, 3] == 0.0 and got[1, 1, 2, 3] == 0.0 and got[0, 0, 0, 0] != 0.0
    q = O.mart_ref_params(a1=0.0, a2=1.0, a3=0.0, w_grad=0.0, w_chem=0.0, w_el=0.0, w_dis=2.0, w_noise=0.0, dt=dt,
                          chem_form=1)
    K = np.full((2, n, n, n), 3.0)
    got = O.mart_ref_run(eta, B0, q, 1, K=K)
    np.testing.assert_allclose(got[:, 0, 0, 0], [e1 * (1 - dt * 2.0 * 3.0), e2 * (1 - dt * 2.0 * 3.0)], rtol=1e-14)


def test_prec_ref_elastic_matches_martensite_rule():
    """With every Landau coefficient zero the reference precipitate eta update is the
    semi-implicit step with F = xi * 2 R eta, which is the martensite rule with w_el = xi."""
    n, V, steps = 16, 2, 3
    rng = np.random.default_rng(8)
    C = _cubic(0.7, 0.3, 0.2)
    B = O.btensor(n, C, _eps(V))
    eta = rng.uniform(-0.1, 0.1, (V, n, n, n))
    c = rng.uniform(0.2, 0.3, (n, n, n))
    pp = O.prec_params(A=[0.0] * 8, xi=5.0, L=0.7, kappa_eta=1.1, dt=0.02, M=1.0, kappa_c=1.0)
    _, got = O.prec_ref_run(c, eta, B, pp, steps)
    pm = O.mart_params(A=0.0, B=0.0, C=0.0, w_el=5.0, L=0.7, kappa=1.1, dt=0.02)
    np.testing.assert_allclose(got, O.mart_run(eta, B, pm, steps), rtol=0, atol=1e-13)


def test_prec_ref_without_order_parameters_is_the_plain_ch_step():
    """eta = 0 removes the eta^2 source: the reference rule equals the spectral rule with no
    elasticity and no stabilisation (noise off)."""
    n, V = 16, 2
    rng = np.random.default_rng(9)
    c = rng.uniform(0.2, 0.3, (n, n, n))
    eta = np.zeros((V, n, n, n))
    pp = O.prec_params(xi=400.0, S=0.0, dt=0.001, M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5)
    cr, er = O.prec_ref_run(c, eta, O.btensor(n, _cubic(3.5, 1.5, 1.0), _eps(V)), pp, 4)
    q = O.prec_params(xi=0.0, S=0.0, dt=0.001, M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5)
    cs, es = O.prec_run(c, eta, np.zeros((n, n, n)), q, 4)
    np.testing.assert_allclose(cr, cs, rtol=0, atol=1e-13)
    assert np.abs(er).max() == 0.0


def test_prec_ref_conserves_mass_with_noise():
    n, V = 16, 3
    rng = np.random.default_rng(10)
    c = rng.uniform(0.2, 0.3, (n, n, n))
    eta = rng.uniform(0.0, 0.05, (V, n, n, n))
    pp = O.prec_params(noise=1e-4, seed=4, dt=0.001, M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5)
    c1, _ = O.prec_ref_run(c, eta, O.btensor(n, _cubic(3.5, 1.5, 1.0), _eps(V)), pp, 5)
    assert abs(c1.sum() - c.sum()) < 1e-11 * c.sum()
    assert np.abs(c1 - c).max() > 0


def test_stress_uniform_state_is_minus_xi_sigma():
    """A uniform eta has only k = 0, where B_bv(0) = C_ijkl e^b_ij eps^v_kl = sigma^v_ab for the
    unit base strain b = (a, b): the stress plugin must return -xi sum_v sigma^v_ab eta_v^2
    (HEAD source) or -xi sum_v sigma^v_ab eta_v (master source), component order 00,01,02,11,12,22."""
    n, V, xi = 8, 3, 4000.0
    C = _cubic(5.39, 3.39, 0.6)
    rng = np.random.default_rng(12)
    eps = rng.uniform(-0.02, 0.02, (V, 3, 3))
    vals = np.array([0.3, -0.5, 0.95])
    eta = vals[:, None, None, None] * np.ones((V, n, n, n))
    sig = np.einsum("ijkl,vkl->vij", C, eps)
    for source, s_v in ((2, vals**2), (1, vals)):
        st, d = O.stress(eta, C, eps, xi=xi, source=source)
        want = [-xi * np.dot(sig[:, a, b], s_v) for a, b in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]]
        np.testing.assert_allclose(st[:, 0, 0, 0], want, rtol=1e-12)
        np.testing.assert_allclose(st.std(axis=(1, 2, 3)), 0.0, atol=1e-9 * xi)
        assert np.all(d == 1.0)  # |0.95| > 0.9 everywhere
    st, d = O.stress(eta * 0.5, C, eps, xi=xi)
    assert np.all(d == 0.0)

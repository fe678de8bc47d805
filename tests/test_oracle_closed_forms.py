"""Closed-form pins of the FP64 oracle's elastic and gradient terms (CPU, no GPU needed).

Each check evaluates a case the mathematics fixes exactly (a single Fourier mode, a
uniform state, a bilinear identity) from quantities pinned elsewhere (the k = 0 and
[100] values of the Khachaturyan tensor, numpy.linalg's 3x3 inverse), so that a
plausible mistake in the oracle -- a doubled or dropped factor, a transposed operand --
fails here even where the other oracle checks are blind to it:

* the elastic term xi B(k) c_hat of the Cahn-Hilliard update
  (src/dynamics_precipitate.cu:197-206 with the dilatational c source): one cosine mode
  of c must be multiplied by an exact factor per step;
* the gradient, chemical and elastic parts of or_mart_energy / or_prec_energy;
* the martensite reference-rule elastic force at a uniform state
  (src/dynamics_mart.cu:283-301 HEAD: El_p = -w_el R_p eta_p, no factor 2);
* the rectangular tensor B(eps1, eps2) used by the stress plugin
  (src/gtensorb.cu:160-181 with V1 != V2).
"""
import numpy as np
import pytest

import oracle as O


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
    """Khachaturyan B at one unit vector from einsum and numpy.linalg.inv."""
    cee = np.einsum("ijkl,pij,qkl->pq", C, e1, e2)
    if np.linalg.norm(nvec) == 0:
        return cee
    s1 = np.einsum("ijkl,pkl->pij", C, e1)
    s2 = np.einsum("ijkl,qkl->qij", C, e2)
    G = np.linalg.inv(np.einsum("ikjl,k,l->ij", C, nvec, nvec))
    return cee - np.einsum("i,pij,jk,qkl,l->pq", nvec, s1, G, s2, nvec)


def cosine(n, m, amp):
    x = np.arange(n)
    th = 2 * np.pi * (m[0] * x[:, None, None] + m[1] * x[None, :, None] + m[2] * x[None, None, :]) / n
    return amp * np.cos(th)


def k2_of(n, m):
    return sum((2 * np.pi * mi / n) ** 2 for mi in m)


# ------------------------------------------------------------------------- (a) CH elastic mode
@pytest.mark.parametrize("S", [0.0, 54.0])
def test_prec_ch_single_mode_with_isotropic_elasticity(S):
    """eta = 0 makes mu_c = A1 (c - c1) linear, and an isotropic C with the dilatational
    eigenstrain e I gives the direction-free B(k != 0) = 3 s0 e - s0^2 / C11, s0 = (C11 + 2 C12) e.
    A cosine mode of c is then multiplied each step by
        g = (1 - dt M k^2 (A1 + xi B - S)) / (1 + dt M (S k^2 + kappa_c k^4))."""
    n, e0 = 16, 0.02
    c11, c12 = 3.0, 1.0
    C = cubic(c11, c12, (c11 - c12) / 2)
    Bc = O.btensor(n, C, e0 * np.eye(3)[None])[0, 0]
    s0 = (c11 + 2 * c12) * e0
    Biso = 3 * s0 * e0 - s0**2 / c11
    dt = 0.005 if S else 0.0005
    p = O.prec_params(M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5, dt=dt, xi=4000.0, S=S)
    A1 = p.A[1]
    modes = [((1, 1, 0), 0.003), ((0, 0, 2), -0.002), ((3, 1, 2), 0.001)]
    c0 = 0.25
    c = np.full((n, n, n), c0)
    for m, a in modes:
        c = c + cosine(n, m, a)
    steps = 30
    c1, eta1 = O.prec_run(c, np.zeros((1, n, n, n)), Bc, p, steps)
    want = np.full((n, n, n), c0)
    for m, a in modes:
        kk = k2_of(n, m)
        g = (1 - dt * p.M * kk * (A1 + p.xi * Biso - S)) / (1 + dt * p.M * (S * kk + p.kappa_c * kk * kk))
        want = want + cosine(n, m, a * g**steps)
    np.testing.assert_allclose(c1, want, rtol=0, atol=1e-14)
    assert np.abs(eta1).max() == 0.0


def test_prec_ch_single_mode_with_cubic_elasticity_along_axes():
    """Cubic C with the dilatational eigenstrain: along [100] (and [010], [001] by cubic
    symmetry) the acoustic tensor is diag(C11, C44, C44), so B = 3 s0 e - s0^2 / C11."""
    n, e0 = 16, 0.02
    c11, c12, c44 = 3.5, 1.5, 1.0
    Bc = O.btensor(n, cubic(c11, c12, c44), e0 * np.eye(3)[None])[0, 0]
    s0 = (c11 + 2 * c12) * e0
    Bax = 3 * s0 * e0 - s0**2 / c11
    p = O.prec_params(M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5, dt=0.005, xi=4000.0, S=54.0)
    modes = [((2, 0, 0), 0.003), ((0, 3, 0), 0.002), ((0, 0, 1), -0.001)]
    c = np.full((n, n, n), 0.3)
    for m, a in modes:
        c = c + cosine(n, m, a)
    c1, _ = O.prec_run(c, np.zeros((2, n, n, n)), Bc, p, 25)
    want = np.full((n, n, n), 0.3)
    for m, a in modes:
        kk = k2_of(n, m)
        g = (1 - p.dt * p.M * kk * (p.A[1] + p.xi * Bax - p.S)) / (1 + p.dt * p.M * (p.S * kk + p.kappa_c * kk**2))
        want = want + cosine(n, m, a * g**25)
    np.testing.assert_allclose(c1, want, rtol=0, atol=1e-14)


# ------------------------------------------------------------------------- (b) energies
def test_mart_energy_of_a_single_mode():
    """eta_0 = a cos(k0.x), eta_1 = b (uniform), eta_2 = 0 on n^3 = N voxels (3 m, 4 m not
    aliased to 0):
        E_grad = kappa k0^2 a^2 N / 4
        E_chem = A/2 N (a^2/2 + b^2) - B/3 N b^3 + C/4 N (3 a^4/8 + a^2 b^2 + b^4)
        E_el   = w/2 N [a^4/4 B00(0) + a^4/8 B00(2k0) + a^2 b^2/2 (B01(0) + B10(0)) + b^4 B11(0)]
    with B(0) = C eps eps and B(2k0) from numpy.linalg along k0."""
    n, V = 16, 3
    m = (1, 2, 0)
    a, b = 0.4, -0.3
    C = cubic(0.7, 0.3, 0.2)
    eps = mart_eps(V)
    B = O.btensor(n, C, eps)
    eta = np.zeros((V, n, n, n))
    eta[0] = cosine(n, m, a)
    eta[1] = b
    p = O.mart_params(A=0.2, B=3.0, C=3.0, kappa=1.3, w_el=7.0)
    E = O.mart_energy(eta, B, p)
    N = n**3
    k0 = k2_of(n, m)
    B0 = np.einsum("ijkl,pij,qkl->pq", C, eps, eps)
    u = np.array(m, float) / np.linalg.norm(m)
    B2 = b_numpy(C, eps, eps, u)
    e_grad = p.kappa * k0 * a * a * N / 4
    e_chem = p.A / 2 * N * (a * a / 2 + b * b) - p.B / 3 * N * b**3 + p.C / 4 * N * (3 * a**4 / 8 + a * a * b * b + b**4)
    e_el = p.w_el / 2 * N * (a**4 / 4 * B0[0, 0] + a**4 / 8 * B2[0, 0] + a * a * b * b / 2 * (B0[0, 1] + B0[1, 0])
                             + b**4 * B0[1, 1])
    np.testing.assert_allclose(E[:3], [e_chem, e_grad, e_el], rtol=1e-12)
    assert E[3] == pytest.approx(e_chem + e_grad + e_el, rel=1e-13)


def test_prec_energy_of_single_modes():
    """Concentration mode c = c0 + a cos(k0.x), eta = 0:
        E_chem = A1/2 N ((c0 - c1)^2 + a^2/2),  E_grad = kappa_c k0^2 a^2 N / 4,
        E_el   = xi/2 N (B(0) c0^2 + B(k0) a^2 / 2),  B(0) = 3 s0 e, B(k0) = 3 s0 e - s0^2/C11 (isotropic).
    Order-parameter mode eta_0 = b cos(k1.x), c = c0 (6 m not aliased to 0):
        E_chem = N (A1/2 (c0 - c1)^2 + A2/4 (c0 - c2) b^2 - 3 A3/32 b^4 + 5 A4/96 b^6),
        E_grad = kappa_eta k1^2 b^2 N / 4,  E_el = xi/2 N B(0) c0^2."""
    n, e0 = 16, 0.02
    c11, c12 = 3.0, 1.0
    Bc = O.btensor(n, cubic(c11, c12, (c11 - c12) / 2), e0 * np.eye(3)[None])[0, 0]
    s0 = (c11 + 2 * c12) * e0
    B0, Bk = 3 * s0 * e0, 3 * s0 * e0 - s0**2 / c11
    p = O.prec_params(kappa_c=1.5, kappa_eta=2.5, xi=400.0)
    A = p.A
    N = n**3
    c0, a = 0.25, 0.004
    m = (2, 1, 1)
    c = c0 + cosine(n, m, a)
    E = O.prec_energy(c, np.zeros((1, n, n, n)), Bc, p)
    want = [A[1] / 2 * N * ((c0 - p.c1) ** 2 + a * a / 2), p.kappa_c * k2_of(n, m) * a * a * N / 4,
            p.xi / 2 * N * (B0 * c0**2 + Bk * a * a / 2)]
    np.testing.assert_allclose(E[:3], want, rtol=1e-12)
    b, m1 = 0.7, (1, 0, 0)
    eta = np.zeros((2, n, n, n))
    eta[0] = cosine(n, m1, b)
    E = O.prec_energy(np.full((n, n, n), c0), eta, Bc, p)
    want = [N * (A[1] / 2 * (c0 - p.c1) ** 2 + A[2] / 4 * (c0 - p.c2) * b * b - 3 * A[3] / 32 * b**4
                 + 5 * A[4] / 96 * b**6),
            p.kappa_eta * k2_of(n, m1) * b * b * N / 4, p.xi / 2 * N * B0 * c0**2]
    np.testing.assert_allclose(E[:3], want, rtol=1e-12)


# ------------------------------------------------------------------------- (c) mart reference El
def test_mart_ref_elastic_force_at_uniform_state():
    """A uniform eta has only k = 0, where B(0) = C eps eps, so R_p = sum_q B_pq(0) eta_q^2 and
    one explicit step is eta_p + dt (-w_el R_p eta_p) (src/dynamics_mart.cu:283-301, HEAD)."""
    n, V = 8, 3
    C = cubic(3.5, 1.5, 1.0)
    eps = mart_eps(V)
    B = O.btensor(n, C, eps)
    vals = np.array([0.3, -0.2, 0.5])
    eta = vals[:, None, None, None] * np.ones((V, n, n, n))
    dt, w = 0.01, 1000.0
    p = O.mart_ref_params(a1=0.0, a2=1.0, a3=0.0, w_grad=2.5, w_chem=0.0, w_el=w, w_dis=0.0, w_noise=0.0, dt=dt,
                          chem_form=1)
    got = O.mart_ref_run(eta, B, p, 1)
    B0 = np.einsum("ijkl,pij,qkl->pq", C, eps, eps)
    want = vals + dt * (-w * (B0 @ vals**2) * vals)
    for v in range(V):
        np.testing.assert_allclose(got[v], want[v], rtol=1e-13)
    assert np.abs(want - vals).min() > 1e-4  # the elastic step is not negligible


# ------------------------------------------------------------------------- (d) rectangular B
BASE = np.zeros((6, 3, 3))
for _b, (_i, _j) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
    BASE[_b, _i, _j] = 1.0


def test_rectangular_btensor_closed_forms():
    """B(eps1, eps2) with V1 = 6 unit base strains and V2 = 3 general eigenstrains: k = 0 is
    C eps1 eps2; [100] with diagonal-free shear strains against numpy.linalg along several
    directions; swapping the roles transposes p and q."""
    n = 8
    C = cubic(5.39, 3.39, 0.6)
    rng = np.random.default_rng(21)
    e2 = rng.uniform(-0.02, 0.02, (3, 3, 3))
    e2 = 0.5 * (e2 + np.transpose(e2, (0, 2, 1)))
    Br = O.btensor(n, C, BASE, e2)
    assert Br.shape == (6, 3, n, n, n)
    np.testing.assert_allclose(Br[:, :, 0, 0, 0], np.einsum("ijkl,pij,qkl->pq", C, BASE, e2), atol=1e-14)
    for (x, y, z) in [(1, 0, 0), (0, 3, 0), (1, 1, 0), (1, 1, 1), (2, 1, 3), (7, 1, 5), (3, 6, 2)]:
        mvec = np.array([x if x < n // 2 else x - n, y if y < n // 2 else y - n, z if z < n // 2 else z - n], float)
        u = mvec / np.linalg.norm(mvec)
        np.testing.assert_allclose(Br[:, :, x, y, z], b_numpy(C, BASE, e2, u), atol=1e-13)
    Bt = O.btensor(n, C, e2, BASE)
    np.testing.assert_allclose(Bt, np.transpose(Br, (1, 0, 2, 3, 4)), atol=1e-14)


def test_rectangular_btensor_is_bilinear_in_the_strains():
    """B is bilinear in (eps1, eps2). For a symmetric eps = sum_b w_b eps[b] E_b over the six unit
    base strains E_b (w_b = 1 on the diagonal, 2 off it, by the minor symmetry of C), the
    square tensor must equal the rectangular one contracted with those weights:
        B(eps, eps)[p, q] = sum_b w_b eps_q[b] B(eps, E)[p, b]."""
    n = 8
    C = cubic(3.5, 1.5, 1.0)
    rng = np.random.default_rng(22)
    eps = rng.uniform(-0.03, 0.03, (3, 3, 3))
    eps = 0.5 * (eps + np.transpose(eps, (0, 2, 1)))
    Bsq = O.btensor(n, C, eps)
    Br = O.btensor(n, C, eps, BASE)
    wts = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])
    coef = np.stack([[wts[b] * eps[q][i, j] for b, (i, j) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)])]
                     for q in range(3)])  # [q][b]
    want = np.einsum("qb,pbxyz->pqxyz", coef, Br)
    np.testing.assert_allclose(Bsq, want, atol=1e-14)

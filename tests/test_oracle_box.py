"""Pins of the FP64 oracle on box grids (nx x ny x nz, spacings dx, dy, dz; CPU, no GPU needed).

The reference reads `gridsize nx ny nz dx dy dz` (src/dynamics_mart.cu:18), builds the wave
numbers of each axis from its own n and spacing (src/gtensorb.cu:14-24) and divides the
difference quotients of its gradient term by each axis's spacing (src/dynamics_mart.cu:181-188).
Each box routine of the oracle is checked against something independent of it:

* the 3D transform against numpy.fft.fftn;
* the Khachaturyan tensor at every voxel against einsum + numpy.linalg.inv at the per-axis wave
  vector, in both wave-number conventions;
* every step rule (martensite and precipitate, spectral and reference rule) by tiling: a field
  on an (nx, ny, nz) box, repeated to fill a cube, evolves on the cube (whose oracle is pinned
  by the other oracle tests) exactly as the box field repeated -- the repeated field holds only
  the cube's wave vectors that are the box's, and its difference quotients see the same
  periodic neighbours;
* the per-axis spacings by single modes: the semi-implicit Allen-Cahn and Cahn-Hilliard
  factors at k_d = 2 pi m / (n_d Delta_d), the reference rule's difference-quotient factor
  (2 cos(2 pi m / n_d) - 2) / (6 Delta_d), and the gradient energy kappa k^2 a^2 N / 4.
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
    cee = np.einsum("ijkl,pij,qkl->pq", C, e1, e2)
    if np.linalg.norm(nvec) <= 5e-7:
        return cee
    nvec = nvec / np.linalg.norm(nvec)
    s1 = np.einsum("ijkl,pkl->pij", C, e1)
    s2 = np.einsum("ijkl,qkl->qij", C, e2)
    G = np.linalg.inv(np.einsum("ikjl,k,l->ij", C, nvec, nvec))
    return cee - np.einsum("i,pij,jk,qkl,l->pq", nvec, s1, G, s2, nvec)


def kval(m, n, d, kconv):
    t = m if m < n // 2 else m - n
    return 2 * 3.14 * d / n * t if kconv == 1 else 2 * np.pi * t / (n * d)


def tile(a, g, n):
    """Repeat the last three axes of a (box g) to fill an n^3 cube."""
    reps = (1,) * (a.ndim - 3) + (n // g.nx, n // g.ny, n // g.nz)
    return np.tile(a, reps)


def test_fft3d_box_matches_numpy():
    rng = np.random.default_rng(1)
    for dims in [(16, 8, 32), (12, 16, 4), (8, 32, 16)]:
        g = O.grid(*dims)
        x = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
        f = O.fft3d_box(x, g, -1)
        assert np.allclose(f, np.fft.fftn(x), rtol=0, atol=1e-10 * np.abs(f).max())
        b = O.fft3d_box(x, g, +1)
        assert np.allclose(b, np.fft.ifftn(x) * x.size, rtol=0, atol=1e-10 * np.abs(b).max())


@pytest.mark.parametrize("kconv", [0, 1])
def test_btensor_box_matches_numpy(kconv):
    g = O.grid(8, 4, 16, 1.0, 2.0, 0.5)
    C = cubic(3.5, 1.5, 1.0)
    e = mart_eps(3)
    B = O.btensor_box(g, C, e, kconv=kconv)
    assert B.shape == (3, 3, 8, 4, 16)
    for x in range(8):
        for y in range(4):
            for z in range(16):
                k = np.array([kval(x, 8, 1.0, kconv), kval(y, 4, 2.0, kconv), kval(z, 16, 0.5, kconv)])
                np.testing.assert_allclose(B[:, :, x, y, z], b_numpy(C, e, e, k), rtol=1e-12, atol=1e-13)


BOXES = [(16, 8, 32), (32, 16, 16), (8, 32, 16)]


@pytest.mark.parametrize("dims", BOXES)
@pytest.mark.parametrize("kconv", [0, 1])
def test_mart_box_equals_tiled_cube(dims, kconv):
    n, V = 32, 3
    g = O.grid(*dims)
    C = cubic(0.7, 0.3, 0.2)
    e = mart_eps(V)
    rng = np.random.default_rng(2)
    eta = rng.uniform(-0.3, 0.3, (V,) + dims)
    p = O.mart_params(A=-0.5, w_el=20.0, kconv=kconv)
    got = O.mart_run_box(eta, O.btensor_box(g, C, e, kconv=kconv), p, g, 4)
    want = O.mart_run(tile(eta, g, n), O.btensor(n, C, e, kconv=kconv), p, 4)
    np.testing.assert_allclose(tile(got, g, n), want, rtol=0, atol=1e-12 * np.abs(want).max())
    Eb = O.mart_energy_box(got, O.btensor_box(g, C, e, kconv=kconv), p, g)
    Ec = O.mart_energy(want, O.btensor(n, C, e, kconv=kconv), p)
    np.testing.assert_allclose(Eb * (n**3 / np.prod(dims)), Ec, rtol=1e-11)


@pytest.mark.parametrize("dims", BOXES[:2])
def test_mart_ref_box_equals_tiled_cube(dims):
    n, V = 32, 3
    g = O.grid(*dims)
    C = cubic(3.5, 1.5, 1.0)
    e = mart_eps(V)
    rng = np.random.default_rng(3)
    eta = rng.uniform(-0.2, 0.6, (V,) + dims)
    defect = (rng.uniform(0, 1, dims) > 0.9).astype(np.float64)
    K = rng.uniform(-1, 1, (V,) + dims)
    p = O.mart_ref_params(a1=-1.0, w_el=50.0, w_noise=0.0, dt=0.002)
    got = O.mart_ref_run_box(eta, O.btensor_box(g, C, e, kconv=1), p, g, 3, defect=defect, K=K)
    want = O.mart_ref_run(tile(eta, g, n), O.btensor(n, C, e, kconv=1), p, 3, defect=tile(defect, g, n),
                          K=tile(K, g, n))
    np.testing.assert_allclose(tile(got, g, n), want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("dims", BOXES[:2])
def test_prec_box_equals_tiled_cube(dims):
    n, V = 32, 3
    g = O.grid(*dims)
    C = cubic(3.0, 1.0, 1.0)
    rng = np.random.default_rng(4)
    c = 0.25 + rng.uniform(-0.05, 0.05, dims)
    eta = rng.uniform(0, 0.3, (V,) + dims)
    p = O.prec_params(M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5, dt=0.005, S=54.0, xi=400.0)
    Bb = O.btensor_box(g, C, 0.02 * np.eye(3)[None])[0, 0]
    Bn = O.btensor(n, C, 0.02 * np.eye(3)[None])[0, 0]
    cb, eb = O.prec_run_box(c, eta, Bb, p, g, 3)
    cc, ec = O.prec_run(tile(c, g, n), tile(eta, g, n), Bn, p, 3)
    np.testing.assert_allclose(tile(cb, g, n), cc, rtol=0, atol=1e-13)
    np.testing.assert_allclose(tile(eb, g, n), ec, rtol=0, atol=1e-12 * np.abs(ec).max())
    Eb = O.prec_energy_box(cb, eb, Bb, p, g)
    Ec = O.prec_energy(cc, ec, Bn, p)
    np.testing.assert_allclose(Eb * (n**3 / np.prod(dims)), Ec, rtol=1e-11)


def test_prec_ref_box_equals_tiled_cube():
    n, V, dims = 32, 2, (16, 32, 8)
    g = O.grid(*dims)
    C = cubic(3.0, 1.0, 1.0)
    e = mart_eps(V)
    rng = np.random.default_rng(5)
    c = 0.25 + rng.uniform(-0.05, 0.05, dims)
    eta = rng.uniform(0, 0.3, (V,) + dims)
    p = O.prec_params(dt=0.001, noise=0.0, kconv=1)
    cb, eb = O.prec_ref_run_box(c, eta, O.btensor_box(g, C, e, kconv=1), p, g, 3)
    cc, ec = O.prec_ref_run(tile(c, g, n), tile(eta, g, n), O.btensor(n, C, e, kconv=1), p, 3)
    np.testing.assert_allclose(tile(cb, g, n), cc, rtol=0, atol=1e-13)
    np.testing.assert_allclose(tile(eb, g, n), ec, rtol=0, atol=1e-12 * np.abs(ec).max())


def mode(dims, axis, m, amp):
    n = dims[axis]
    t = np.arange(n)
    shape = [1, 1, 1]
    shape[axis] = n
    return np.broadcast_to(amp * np.cos(2 * np.pi * m * t / n).reshape(shape), dims).copy()


SP = (1.0, 0.5, 2.0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_mart_box_spacing_single_mode_factor(axis):
    """B = C = 0, no elasticity: a cosine mode along one axis is multiplied per step by
    (1 - dt L A) / (1 + dt L kappa k^2), k = 2 pi m / (n_axis Delta_axis)."""
    dims = (8, 16, 32)
    g = O.grid(*dims, *SP)
    m, amp, steps = 3, 0.01, 5
    eta = mode(dims, axis, m, amp)[None]
    p = O.mart_params(A=0.3, B=0.0, C=0.0, w_el=0.0, kappa=1.3, L=0.7, dt=0.05)
    got = O.mart_run_box(eta, np.zeros((1, 1) + dims), p, g, steps)
    k2 = (2 * np.pi * m / (dims[axis] * SP[axis])) ** 2
    f = (1 - p.dt * p.L * p.A) / (1 + p.dt * p.L * p.kappa * k2)
    np.testing.assert_allclose(got, eta * f**steps, rtol=0, atol=1e-15)
    E = O.mart_energy_box(eta[0][None], np.zeros((1, 1) + dims), p, g)
    N = np.prod(dims)
    np.testing.assert_allclose(E[1], p.kappa * k2 * amp**2 * N / 4, rtol=1e-12)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_prec_box_spacing_single_mode_factor(axis):
    """eta = 0, xi = 0: a cosine mode of c is multiplied per step by
    (1 - dt M k^2 (A1 - S)) / (1 + dt M (S k^2 + kappa_c k^4)); eta = 0 stays 0."""
    dims = (16, 8, 32)
    g = O.grid(*dims, *SP)
    m, amp, steps = 2, 0.003, 4
    c = 0.3 + mode(dims, axis, m, amp)
    eta = np.zeros((2,) + dims)
    p = O.prec_params(M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5, dt=0.005, S=54.0, xi=0.0)
    cb, eb = O.prec_run_box(c, eta, np.zeros(dims), p, g, steps)
    k2 = (2 * np.pi * m / (dims[axis] * SP[axis])) ** 2
    A1 = p.A[1]
    f = (1 - p.dt * p.M * k2 * (A1 - p.S)) / (1 + p.dt * p.M * (p.S * k2 + p.kappa_c * k2 * k2))
    np.testing.assert_allclose(cb - 0.3, (c - 0.3) * f**steps, rtol=0, atol=1e-14)
    assert np.abs(eb).max() == 0.0
    E = O.prec_energy_box(c, eta, np.zeros(dims), p, g)
    np.testing.assert_allclose(E[1], p.kappa_c * k2 * amp**2 * np.prod(dims) / 4, rtol=1e-12)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_mart_ref_box_difference_quotient_spacing(axis):
    """Reference rule with only the gradient term: a cosine mode along one axis is multiplied
    per step by 1 + dt w_grad (2 cos(2 pi m / n) - 2) / (2 Delta) / 3 (src/dynamics_mart.cu:181-188)."""
    dims = (16, 32, 8)
    g = O.grid(*dims, *SP)
    m, amp, steps = 1, 0.2, 3
    eta = mode(dims, axis, m, amp)[None]
    p = O.mart_ref_params(a1=0.0, a2=1.0, a3=0.0, w_chem=0.0, w_el=0.0, w_dis=0.0, w_noise=0.0, w_grad=2.5, dt=0.01)
    got = O.mart_ref_run_box(eta, np.zeros((1, 1) + dims), p, g, steps)
    th = 2 * np.pi * m / dims[axis]
    f = 1 + p.dt * p.w_grad * (2 * np.cos(th) - 2) / (2 * SP[axis]) / 3
    np.testing.assert_allclose(got, eta * f**steps, rtol=0, atol=1e-15)

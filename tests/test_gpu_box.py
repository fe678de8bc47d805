"""GPU: box grids (nx x ny x nz, spacings dx, dy, dz; the reference's `gridsize nx ny nz dx dy dz`,
src/dynamics_mart.cu:18) through the C ABI against the FP64 oracle's box routines (pinned in
tests/test_oracle_box.py by numpy, closed forms and tiling onto the cube oracle).

Bars: FP64 box path (GSK_PATH_PLAIN64) 1e-11 relative; FP32 box path (GSK_PATH_FUSED32) within the
tolerance written in each test; the Philox noise enters by voxel index as on cubes.
"""
import sys

import numpy as np
import pytest

import gsk
import oracle as O

pytestmark = pytest.mark.gpu

BOXES = [((32, 16, 64), (1.0, 0.5, 2.0)), ((16, 64, 32), (1.0, 1.0, 1.0)), ((96, 32, 16), (0.7, 0.7, 1.3))]


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def box_cfg(physics, dims, sp, ref=False, **kw):
    mk = gsk.reference_config if ref else gsk.default_config
    return mk(physics, n=dims[0], ny=dims[1], nz=dims[2], dx=sp[0], dy=sp[1], dz=sp[2], **kw)


def grid(dims, sp):
    return O.grid(*dims, *sp)


# the box steps share transforms between pairs of real fields: odd and even V, one variant
# (a single pair with an empty partner) and the largest count
@pytest.mark.parametrize("dims,sp,V", [(d, sp, 3) for d, sp in BOXES] + [(*BOXES[0], V) for V in (1, 2, 4, 8)])
@pytest.mark.parametrize("path", [gsk.PATH_PLAIN64, gsk.PATH_FUSED32])
def test_box_mart_spectral_matches_oracle(gpu, dims, sp, V, path):
    say(f"box mart spectral dims={dims} sp={sp} V={V} path={path}")
    steps = 12
    C = gsk.cubic_modulus(0.7, 0.3, 0.2)
    eps = gsk.mart_eigenstrains(V)
    cfg = box_cfg(gsk.PHYS_MART, dims, sp, nvar=V, path=path, A=-0.5, w_el=20.0, noise=0.002, seed=9)
    with gsk.Sim(cfg) as s:
        assert s.dims == dims
        s.build_btensor(C, eps)
        s.init_uniform("eta", -0.2, 0.2, 4)
        eta0 = s.get_field("eta", np.float64)
        E0 = s.energy()
        s.step(steps)
        got = s.get_field("eta", np.float64)
        E1 = s.energy()
    g = grid(dims, sp)
    p = O.mart_params(A=cfg.A, B=cfg.B, C=cfg.C, kappa=cfg.kappa, L=cfg.L, dt=cfg.dt, w_el=cfg.w_el, noise=cfg.noise,
                      seed=cfg.seed)
    B = O.btensor_box(g, C, eps)
    want = O.mart_run_box(eta0, B, p, g, steps)
    Ew0, Ew1 = O.mart_energy_box(eta0, B, p, g), O.mart_energy_box(want, B, p, g)
    err = rel(got, want)
    eE = abs(E1[3] - Ew1[3]) / abs(Ew1[3])
    say(f"  rel L2 {err:.2e}, energy {eE:.2e}, |eta| max {np.abs(want).max():.3f}")
    tol = 1e-11 if path == gsk.PATH_PLAIN64 else 2e-5
    assert err < tol
    assert abs(E0[3] - Ew0[3]) / abs(Ew0[3]) < (1e-11 if path == gsk.PATH_PLAIN64 else 1e-5)
    assert eE < (1e-11 if path == gsk.PATH_PLAIN64 else 1e-4)


# F = V + 1 fields: even (V = 3, 7) and odd (V = 2, 4) counts of fields to pair
@pytest.mark.parametrize("dims,sp,V", [(d, sp, 3) for d, sp in BOXES[:2]] + [(*BOXES[0], 2), (*BOXES[2], 4),
                                                                           (*BOXES[1], 7)])
@pytest.mark.parametrize("path", [gsk.PATH_PLAIN64, gsk.PATH_FUSED32])
def test_box_prec_spectral_matches_oracle(gpu, dims, sp, V, path):
    say(f"box precipitate spectral dims={dims} sp={sp} V={V} path={path}")
    steps = 20
    C = gsk.cubic_modulus(3.0, 1.0, 1.0)
    cfg = box_cfg(gsk.PHYS_PREC, dims, sp, nvar=V, path=path, M=1.0, L=1.0, kappa=1.5, kappa_c=1.5, dt=0.005, S=54.0,
                  noise=0.001, seed=3)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, 0.02 * np.eye(3)[None])
        s.init_uniform("concentration", 0.245, 0.255, 1)
        s.init_uniform("eta", 0.0, 0.01, 2)
        c0 = s.get_field("concentration", np.float64)
        h0 = s.get_field("eta", np.float64)
        m0 = s.mass()
        s.step(steps)
        c1 = s.get_field("concentration", np.float64)
        h1 = s.get_field("eta", np.float64)
        m1 = s.mass()
        E1 = s.energy()
    g = grid(dims, sp)
    p = O.prec_params(A=list(cfg.a), c1=cfg.c1, c2=cfg.c2, M=cfg.M, L=cfg.L, kappa_c=cfg.kappa_c, kappa_eta=cfg.kappa,
                      xi=cfg.xi, S=cfg.S, dt=cfg.dt, noise=cfg.noise, seed=cfg.seed)
    Bc = O.btensor_box(g, C, 0.02 * np.eye(3)[None])[0, 0]
    cw, hw = O.prec_run_box(c0, h0, Bc, p, g, steps)
    Ew = O.prec_energy_box(cw, hw, Bc, p, g)
    eh, ec, dc = rel(h1, hw), rel(c1 - c0.mean(), cw - c0.mean()), float(np.abs(c1 - cw).max())
    dm = abs(m1 - m0) / abs(m0)
    dE = abs(E1[3] - Ew[3]) / abs(Ew[3])
    say(f"  rel L2 eta {eh:.2e}, c - cbar {ec:.2e}, max|d c| {dc:.1e}, sum c drift {dm:.1e}, energy {dE:.2e}")
    if path == gsk.PATH_PLAIN64:
        assert eh < 1e-11 and ec < 1e-11 and dm < 1e-13 and dE < 1e-11
    else:
        # FP32: c = cbar (FP64 mean rounded to binary32) + an FP32 fluctuation, as on the cube path
        assert eh < 2e-5 and dc < 2e-6 and dm < 1e-10 and dE < 1e-5


# (16, 32, 96): z rows of 24 quads, which do not divide a wave, so the update kernel loads the z
# neighbours across quad boundaries instead of taking them from the neighbouring lane
# V = 4: R of two full pairs
@pytest.mark.parametrize("dims,sp,V", [(d, sp, 3) for d, sp in BOXES[:2] + [((16, 32, 96), (1.0, 0.8, 1.2))]]
                         + [(*BOXES[1], 4)])
def test_box_mart_reference_rule_matches_oracle(gpu, dims, sp, V):
    say(f"box mart reference rule dims={dims} sp={sp} V={V}")
    steps = 40
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.01)
    cfg = box_cfg(gsk.PHYS_MART, dims, sp, ref=True, nvar=V, temperature=300.0, noise=1e-4, w_el=1e3, seed=6, dt=0.005)
    rng = np.random.default_rng(1)
    defect = (rng.random(dims) < 0.01).astype(np.float32)
    stress = rng.uniform(-0.01, 0.01, (6,) + dims).astype(np.float32)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", 0.0, 0.2, 3)
        s.set_field("defect", defect)
        s.set_field("dislocationstress", stress)
        eta0 = s.get_field("eta", np.float64)
        s.step(steps)
        got = s.get_field("eta", np.float64)
        fm = s.mart_forces()
    g = grid(dims, sp)
    a1 = 0.02 * (cfg.temperature - cfg.T0)
    p = O.mart_ref_params(a1=a1, a2=cfg.B, a3=cfg.C, w_grad=cfg.w_grad, w_chem=cfg.w_chem, w_el=cfg.w_el,
                          w_dis=cfg.w_dis, w_noise=cfg.noise, dt=cfg.dt, kconv=cfg.kconv, chem_form=0, seed=cfg.seed)
    comp = [0, 1, 2, 1, 3, 4, 2, 4, 5]
    K = 25.0 * np.einsum("ijxyz,pij->pxyz", stress.astype(np.float64)[comp].reshape((3, 3) + dims), eps)
    want = O.mart_ref_run_box(eta0, O.btensor_box(g, C, eps, kconv=cfg.kconv), p, g, steps,
                              defect=defect.astype(np.float64), K=K)
    err = rel(got, want)
    say(f"  rel L2 {err:.2e}, |eta| max {np.abs(want).max():.3f}, force means {fm}")
    assert np.isfinite(got).all() and np.all(got[:, defect > 0] == 0.0)
    assert err < 2e-5
    assert np.isfinite(fm).all() and fm[3] != 0.0


@pytest.mark.parametrize("V", [3, 2])
def test_box_prec_reference_rule_matches_oracle(gpu, V):
    dims, sp = (32, 64, 16), (1.0, 0.5, 1.0)
    say(f"box precipitate reference rule dims={dims} sp={sp} V={V}")
    steps = 30
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.02)
    cfg = box_cfg(gsk.PHYS_PREC, dims, sp, ref=True, nvar=V, M=1.0, L=1.0, kappa=1.5, kappa_c=1.5, dt=0.002, seed=8,
                  xi=50.0)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("concentration", 0.245, 0.255, 11)
        s.init_uniform("eta", 0.0, 0.1, 12)
        c0 = s.get_field("concentration", np.float64)
        h0 = s.get_field("eta", np.float64)
        s.step(steps)
        c1 = s.get_field("concentration", np.float64)
        h1 = s.get_field("eta", np.float64)
    g = grid(dims, sp)
    p = O.prec_params(A=list(cfg.a), c1=cfg.c1, c2=cfg.c2, M=cfg.M, L=cfg.L, kappa_c=cfg.kappa_c,
                      kappa_eta=cfg.kappa, xi=cfg.xi, S=0.0, dt=cfg.dt, noise=cfg.noise, kconv=cfg.kconv,
                      seed=cfg.seed)
    cw, hw = O.prec_ref_run_box(c0, h0, O.btensor_box(g, C, eps, kconv=cfg.kconv), p, g, steps)
    say(f"  rel L2 eta {rel(h1, hw):.2e}  c-cbar {rel(c1 - c0.mean(), cw - c0.mean()):.2e}")
    assert rel(h1, hw) < 1e-4
    assert rel(c1 - c0.mean(), cw - c0.mean()) < 1e-4


def test_box_equal_edges_unequal_spacing_and_errors(gpu):
    """A cube with unequal spacings is a box too; handles of box grids refuse the cube-only
    entry points (adjoint, per-kernel timing) and ensembles refuse box members."""
    say("box: cube edges with unequal spacings; cube-only entry points")
    dims, sp = (32, 32, 32), (1.0, 1.0, 0.5)
    V = 2
    C = gsk.cubic_modulus(0.7, 0.3, 0.2)
    eps = gsk.mart_eigenstrains(V)
    cfg = box_cfg(gsk.PHYS_MART, dims, sp, nvar=V, path=gsk.PATH_PLAIN64, A=-0.5, w_el=20.0)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", -0.2, 0.2, 4)
        eta0 = s.get_field("eta", np.float64)
        s.step(5)
        got = s.get_field("eta", np.float64)
        with pytest.raises(gsk.GskError):
            s.kernel_times(1)
    g = grid(dims, sp)
    p = O.mart_params(A=cfg.A, B=cfg.B, C=cfg.C, kappa=cfg.kappa, L=cfg.L, dt=cfg.dt, w_el=cfg.w_el)
    want = O.mart_run_box(eta0, O.btensor_box(g, C, eps), p, g, 5)
    assert rel(got, want) < 1e-11
    # the same edges with one spacing is a cube: the oracle's cube and box runs agree
    np.testing.assert_allclose(O.mart_run_box(eta0, O.btensor_box(grid(dims, (1.0, 1.0, 1.0)), C, eps), p,
                                              grid(dims, (1.0, 1.0, 1.0)), 2),
                               O.mart_run(eta0, O.btensor(32, C, eps), p, 2), rtol=0, atol=1e-15)
    f32 = box_cfg(gsk.PHYS_MART, (32, 16, 32), (1, 1, 1), nvar=V)
    with gsk.Sim(f32) as s:
        s.build_btensor(C, eps)
        with pytest.raises(gsk.GskError):
            s.adjoint_forward(2, 1)
    with pytest.raises(gsk.GskError):
        gsk.Ensemble([f32])

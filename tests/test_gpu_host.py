"""GPU: the gsi script host driving the plugins end to end (src/gs.cu:168-212 run loop,
src/dynamics.h:9-25 plugin interface) against the same steps taken through the C ABI.

Bar: bit-exact. The host only forwards parameters and fields; the dumped binary fields
(float32, include/datamain.th:356-418) must equal the library's output for the same
configuration, including the per-step temperature of `set fix temperature`.
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import gsk

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GSI = os.path.join(ROOT, "general-simulator_amd", "gsi")


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def run_gsi(tmp_path, script):
    p = tmp_path / "in.test"
    p.write_text(script)
    r = subprocess.run([GSI, str(p)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def read_binary(path):
    raw = open(path, "rb").read()
    rank = struct.unpack("i", raw[:4])[0]
    dims = struct.unpack("i" * rank, raw[4:4 + 4 * rank])
    return np.frombuffer(raw[4 + 4 * rank:], np.float32).reshape(dims)


def tensor_lines(eps):
    V = eps.shape[0]
    return f"3 {V} 3 3\n" + " ".join(f"{x:.17g}" for x in eps.ravel()) + "\n"


def test_gsi_spectral_mart_equals_library(gpu, tmp_path):
    say("gsi spectral martensite vs library")
    n, V, steps = 32, 3, 12
    eps = gsk.mart_eigenstrains(V)
    script = (f"sys mart\nset rule spectral\nset gridsize {n} {n} {n} 1 1 1\nset deltatime 0.01\nset coefficient -0.5 3 3\n"
              f"set weightelastic 20\nset noiseamp 0.002\nset seed 3\nset cubicmodulus 0.7 0.3 0.2\n"
              f"readhere varianttensor\n{tensor_lines(eps)}set initeta uniform -0.05 0.05 7\n"
              f"set binary on\ndump out 1 eta\ninfo 3 eta step\nrun {steps}\n")
    out = run_gsi(tmp_path, script)
    got = read_binary(tmp_path / "out" / f"eta.{steps}.data")
    cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, dt=0.01, A=-0.5, B=3.0, C=3.0, w_el=20.0, noise=0.002,
                             seed=3)
    with gsk.Sim(cfg) as s:
        s.build_btensor(gsk.cubic_modulus(0.7, 0.3, 0.2), eps)
        s.init_uniform("eta", -0.05, 0.05, 7)
        s.step(steps)
        want = s.get_field("eta")
    assert "Info" in out
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_gsi_reference_mart_with_temperature_ramp(gpu, tmp_path):
    say("gsi reference martensite with a temperature ramp")
    n, V, steps = 32, 3, 10
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.01)
    script = (f"sys mart\nset gridsize {n} {n} {n} 1 1 1\nset rule reference\nset deltatime 0.005\n"
              f"set coefficient 0 32.05 37.5\nset weightelastic 1000\nset weightnoise 0.0001\nset seed 6\n"
              f"readhere varianttensor\n{tensor_lines(eps)}set initeta uniform 0 0.2 3\n"
              f"set fix temperature 300 250\nset binary on\ndump out 1 eta\ninfo 2 eta temperature\nrun {steps}\n")
    run_gsi(tmp_path, script)
    got = read_binary(tmp_path / "out" / f"eta.{steps}.data")
    cfg = gsk.reference_config(gsk.PHYS_MART, n=n, nvar=V, dt=0.005, B=32.05, C=37.5, w_el=1000.0, noise=1e-4, seed=6)
    with gsk.Sim(cfg) as s:
        s.build_btensor(gsk.cubic_modulus(3.5, 1.5, 1.0), eps)
        s.init_uniform("eta", 0.0, 0.2, 3)
        for i in range(1, steps + 1):
            # GS::Run calls Fix(i / total) before step i; the temperature travels as text (6 digits)
            s.set_param("temperature", float(f"{300 + (i / steps) * (250 - 300):.6g}"))
            s.step(1)
        want = s.get_field("eta")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_stress_matches_oracle(gpu):
    """gsk_stress (the stress plugin on the GPU) against the FP64 oracle, both sources."""
    import oracle as O

    say("stress vs oracle")
    n, V = 32, 4
    rng = np.random.default_rng(4)
    eta = rng.uniform(-1.0, 1.0, (V, n, n, n)).astype(np.float32)
    eps = rng.uniform(-0.02, 0.02, (V, 3, 3))
    C = gsk.cubic_modulus(5.39, 3.39, 0.6)
    for source in (2, 1):
        st, d = gsk.stress(eta, C, eps, xi=4000.0, source=source, dx=0.2, kconv=gsk.K_REFERENCE)
        wst, wd = O.stress(eta.astype(np.float64), C, eps, xi=4000.0, source=source, dx=0.2, kconv=gsk.K_REFERENCE)
        err = float(np.linalg.norm((st - wst).ravel()) / np.linalg.norm(wst.ravel()))
        say(f"  source {source}: rel L2 {err:.2e}")
        assert err < 1e-5
        assert np.array_equal(d, wd.astype(np.float32))


def test_gsi_pipeline_cores_precipitate_stress_mart(gpu, tmp_path):
    """The reference's script chain in one gsi run: cores seeds c and eta, the precipitate
    (reference rule) evolves them, stress turns eta into the dislocation stress and the defect
    map, and the martensite reference rule consumes both. Checks the hand-offs: stress equals
    gsk_stress of the dumped eta, and the martensite eta is zero on every defect voxel."""
    say("gsi pipeline")
    n, V = 32, 3
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.01)
    vt = tensor_lines(eps)
    script = (f"set gridsize {n} {n} {n} 1 1 1\nset binary on\n"
              f"sys cores\nset variantn {V}\nset coresn 2\nset radius 5\nset method regular\nrun\n"
              f"sys precipitate\nset rule reference\nset deltatime 0.002\nset meta 1\nset lpp 1\nset arfi 1.5\n"
              f"set beta 1.5\nset xi 50\nreadhere varianttensor\n{vt}run 20\n"
              f"dump pre 1 eta\nrun 1\n"
              f"sys stress\nset xi 4000\ndump st 1 stress defect\nrun\n"
              f"shell cp st/stress.1.data ds.data\nread dislocationstress ds.data\n"
              f"sys mart\nset rule reference\nset deltatime 0.005\nset weightelastic 1000\nset temperature 300\n"
              f"dump mt 1 eta\nrun 5\n")
    out = run_gsi(tmp_path, script)
    pre = read_binary(tmp_path / "pre" / "eta.21.data")
    st = read_binary(tmp_path / "st" / "stress.1.data")
    df = read_binary(tmp_path / "st" / "defect.1.data")
    mt = read_binary(tmp_path / "mt" / "eta.5.data")
    assert np.isfinite(pre).all() and np.isfinite(st).all() and np.isfinite(mt).all(), out
    want, wd = gsk.stress(pre, _stress_default_modulus(), eps, xi=4000.0)
    assert np.array_equal(df, wd)
    assert float(np.abs(st - want).max()) <= 1e-6 * float(np.abs(want).max())
    assert df.sum() > 0 and np.all(mt[:, df > 0] == 0.0)


def _stress_default_modulus():
    C = np.zeros((3, 3, 3, 3))
    for (i, j, k, l), v in {(0, 0, 0, 0): 5.39, (1, 1, 1, 1): 5.39, (2, 2, 2, 2): 5.22, (0, 0, 1, 1): 3.39,
                            (1, 1, 0, 0): 3.39, (1, 1, 2, 2): 3.56, (2, 2, 1, 1): 3.56, (2, 2, 0, 0): 3.56,
                            (0, 0, 2, 2): 3.56}.items():
        C[i, j, k, l] = v
    for (a, b), v in {(0, 1): 0.6, (1, 2): 0.77, (0, 2): 0.77}.items():
        for p in ((a, b), (b, a)):
            for q in ((a, b), (b, a)):
                C[p + q] = v
    return C


def test_gsi_restart_with_ahead_steps(gpu, tmp_path):
    """Restart through a dump (src/gs.cu:183-184): `read eta` of the binary dump at step 5, then
    `set ahead_steps 5` and `run 5`, must reproduce the uninterrupted run's step-10 dump.

    The fused step's state is the FP32 k-space field, of which the dumped real field is the
    rounded inverse transform, so the restart agrees to FP32 rounding (a few ulp), not bit for
    bit. The noise counter must resume at step 5: a restart without ahead_steps draws step 0's
    normals again and is off by the noise increment, far above the rounding."""
    say("gsi restart through ahead_steps")
    n, V = 32, 3
    eps = gsk.mart_eigenstrains(V)
    head = (f"sys mart\nset rule spectral\nset gridsize {n} {n} {n} 1 1 1\nset deltatime 0.01\nset coefficient -0.5 3 3\n"
            f"set weightelastic 20\nset noiseamp 0.002\nset seed 9\nset cubicmodulus 0.7 0.3 0.2\n"
            f"readhere varianttensor\n{tensor_lines(eps)}set binary on\n")
    run_gsi(tmp_path, head + "set initeta uniform -0.05 0.05 4\ndump full 2 eta\nrun 10\n")
    full5 = read_binary(tmp_path / "full" / "eta.5.data")
    full10 = read_binary(tmp_path / "full" / "eta.10.data")
    run_gsi(tmp_path, head + "read eta full/eta.5.data\nset ahead_steps 5\ndump again 1 eta\nrun 5\n")
    again = read_binary(tmp_path / "again" / "eta.10.data")
    run_gsi(tmp_path, head + "read eta full/eta.5.data\ndump wrong 1 eta\nrun 5\n")
    wrong = read_binary(tmp_path / "wrong" / "eta.5.data")
    scale = float(np.abs(full10).max())
    d_again = float(np.abs(again - full10).max())
    d_wrong = float(np.abs(wrong - full10).max())
    say(f"  restart max|d| {d_again:.3e}, without ahead_steps {d_wrong:.3e}, max|eta| {scale:.3e}")
    assert not np.array_equal(full5, full10)
    assert d_again <= 2e-6 * scale
    assert d_wrong > 1e-4 * scale


def test_gsi_reference_mart_info_force_columns(gpu, tmp_path):
    """An in.mart-style script (no `set rule`: the plugin's own update rule) with the reference's
    info columns temperature eta gradient chemical elastic dislocation
    (scripts/in.mart:26, src/dynamics_mart.cu:443-456): every column is a number, and the force
    columns equal the library's force means after the same steps."""
    say("gsi reference martensite info force columns")
    n, V, steps = 32, 3, 8
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.01)
    rng = np.random.default_rng(3)
    stress = rng.uniform(-0.01, 0.01, (6, n, n, n)).astype(np.float32)
    with open(tmp_path / "ds.data", "wb") as f:
        f.write(struct.pack("5i", 4, 6, n, n, n))
        f.write(stress.tobytes())
    script = (f"sys mart\nset gridsize {n} {n} {n} 1 1 1\nset deltatime 0.005\nset transitiontemperature 450\n"
              f"set weightdislocation .1\nset weightelastic 1000\nset weightnoise 0\nset binary on\n"
              f"read dislocationstress ds.data\nreadhere varianttensor\n{tensor_lines(eps)}"
              f"set initeta uniform 0 0.2 3\ninfo 1 temperature eta gradient chemical elastic dislocation\n"
              f"set fix temperature 460 300\nrun {steps}\n")
    out = run_gsi(tmp_path, script)
    assert "reference update rule" in out
    rows = [l.split() for l in out.splitlines() if l.startswith("Step")]
    last = rows[-1]
    assert int(last[1]) == steps
    vals = [float(x) for x in last[2:]]
    assert len(vals) == 6 and all(np.isfinite(vals)), last
    cfg = gsk.reference_config(gsk.PHYS_MART, n=n, nvar=V, dt=0.005, w_el=1000.0, w_dis=0.1, noise=0.0)
    with gsk.Sim(cfg) as s:
        s.build_btensor(gsk.cubic_modulus(3.5, 1.5, 1.0), eps)
        s.set_field("dislocationstress", stress)
        s.init_uniform("eta", 0.0, 0.2, 3)
        for i in range(1, steps + 1):
            s.set_param("temperature", float(f"{460 + (i / steps) * (300 - 460):.6g}"))
            s.step(1)
        want = s.mart_forces()
    np.testing.assert_allclose(vals[2:], want, rtol=2e-5, atol=1e-12)


def test_gsi_box_grid_reference_mart_equals_library(gpu, tmp_path):
    """`set gridsize nx ny nz dx dy dz` with unequal edges and spacings (src/dynamics_mart.cu:18):
    gsi hands the box to the library's box path; the dump has the box's shape and equals the
    library's own run bit for bit."""
    say("gsi box grid reference martensite vs library")
    dims, sp, V, steps = (32, 16, 64), (1.0, 0.5, 2.0), 3, 8
    eps = gsk.mart_eigenstrains(V, base=0.0, axis=0.01)
    script = (f"sys mart\nset gridsize {dims[0]} {dims[1]} {dims[2]} {sp[0]} {sp[1]} {sp[2]}\nset deltatime 0.005\n"
              f"set coefficient 0 32.05 37.5\nset weightelastic 1000\nset weightnoise 0.0001\nset seed 6\n"
              f"readhere varianttensor\n{tensor_lines(eps)}set initeta uniform 0 0.2 3\n"
              f"set binary on\ndump out 1 eta\nrun {steps}\n")
    run_gsi(tmp_path, script)
    got = read_binary(tmp_path / "out" / f"eta.{steps}.data")
    assert got.shape == (V,) + dims
    cfg = gsk.reference_config(gsk.PHYS_MART, n=dims[0], ny=dims[1], nz=dims[2], dx=sp[0], dy=sp[1], dz=sp[2],
                               nvar=V, dt=0.005, B=32.05, C=37.5, w_el=1000.0, noise=1e-4, seed=6)
    with gsk.Sim(cfg) as s:
        s.build_btensor(gsk.cubic_modulus(3.5, 1.5, 1.0), eps)
        s.init_uniform("eta", 0.0, 0.2, 3)
        s.step(steps)
        want = s.get_field("eta")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))

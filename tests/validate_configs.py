"""Accuracy report for the BASELINE configurations against the FP64 oracle.

Not collected by pytest (long oracle runs). Usage on a GPU machine:
    python tests/validate_configs.py [--only 0,1,2] [--out report.json]
Prints one JSON object per configuration: max|d eta|, relative L2, energy error,
and for the precipitate the FP64 drift of sum(c).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "general-simulator_amd"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import gsk  # noqa: E402
import oracle as O  # noqa: E402


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def chunked(run, state, steps, chunk=100, label=""):
    """Advance an oracle run in chunks, printing progress between them."""
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        state = run(state, done, k)
        done += k
        print(f"[validate] {label} oracle {done}/{steps}", flush=True)
    return state


def mart_case(n, steps, path, noise=0.0, seed=0, V=3):
    C = gsk.cubic_modulus(350 / 500, 150 / 500, 100 / 500)
    eps = gsk.mart_eigenstrains(V)
    cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, path=path, noise=noise, seed=seed)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", -0.01, 0.01, 0)
        eta0 = s.get_field("eta").astype(np.float64)
        t0 = time.perf_counter()
        s.step(steps)
        tg = time.perf_counter() - t0
        got = s.get_field("eta", np.float64)
        E = s.energy()
    B = O.btensor(n, C, eps)
    p = O.mart_params(noise=noise, seed=seed)
    t0 = time.perf_counter()
    want = chunked(lambda e, s0, k: O.mart_run(e, B, p, k, step0=s0), eta0, steps, label=f"mart {n}^3")
    tc = time.perf_counter() - t0
    Ew = O.mart_energy(want, B, p)
    return {
        "n": n, "V": V, "steps": steps, "path": "fp32_fused" if path == gsk.PATH_FUSED32 else "fp64_plain",
        "noise": noise, "max_abs_d_eta": float(np.abs(got - want).max()),
        "max_abs_eta": float(np.abs(want).max()), "rel_l2": rel_l2(got, want),
        "energy_gpu": E.tolist(), "energy_oracle": Ew.tolist(),
        "rel_energy_err": float(abs(E[3] - Ew[3]) / abs(Ew[3])), "gpu_s": tg, "oracle_s": tc,
        "oracle_threads": O.threads(),
    }


def prec_case(n, steps, V=4):
    C = gsk.cubic_modulus(3.5, 1.5, 1.0)
    e0 = 0.02 * np.eye(3)[None]
    cfg = gsk.default_config(gsk.PHYS_PREC, n=n, nvar=V, M=1.0, L=1.0, kappa=1.5, kappa_c=1.5, dt=0.005,
                             noise=0.001, seed=5, S=54.0)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, e0)
        s.init_uniform("concentration", 0.245, 0.255, 11)
        s.init_uniform("eta", 0.0, 0.01, 12)
        c0 = s.get_field("concentration", np.float64)
        h0 = s.get_field("eta", np.float64)
        m0 = s.mass()
        s.step(steps)
        c1 = s.get_field("concentration", np.float64)
        h1 = s.get_field("eta", np.float64)
        m1 = s.mass()
        E = s.energy()
    p = O.prec_params(A=list(cfg.a), c1=cfg.c1, c2=cfg.c2, M=1.0, L=1.0, kappa_c=1.5, kappa_eta=1.5, xi=cfg.xi,
                      S=54.0, dt=0.005, noise=0.001, seed=5)
    Bc = O.btensor(n, C, e0)[0, 0]
    cw, hw = chunked(lambda st, s0, k: O.prec_run(st[0], st[1], Bc, p, k, step0=s0), (c0, h0), steps,
                     label=f"prec {n}^3")
    Ew = O.prec_energy(cw, hw, Bc, p)
    return {
        "n": n, "V": V, "steps": steps, "path": "fp32_fused",
        "max_abs_d_eta": float(np.abs(h1 - hw).max()), "max_abs_eta": float(np.abs(hw).max()),
        "rel_l2_eta": rel_l2(h1, hw), "max_abs_d_c": float(np.abs(c1 - cw).max()), "rel_l2_c": rel_l2(c1, cw),
        "mass_drift_rel": float(abs(m1 - m0) / m0), "energy_gpu": E.tolist(), "energy_oracle": Ew.tolist(),
        "rel_energy_err": float(abs(E[3] - Ew[3]) / abs(Ew[3])),
    }


def mart_512_property(steps=500):
    """Config 2 at full size: finite, bounded, energy non-increasing without noise is not
    expected (noise on), so report the state statistics and the mean-field drift."""
    cfg = gsk.default_config(gsk.PHYS_MART, n=512, nvar=3, noise=0.002, seed=2)
    with gsk.Sim(cfg) as s:
        s.build_btensor(gsk.cubic_modulus(0.7, 0.3, 0.2), gsk.mart_eigenstrains(3))
        s.init_uniform("eta", -0.01, 0.01, 0)
        t0 = time.perf_counter()
        s.step(steps)
        dt = time.perf_counter() - t0
        e = s.get_field("eta")
        E = s.energy()
    return {"n": 512, "steps": steps, "finite": bool(np.isfinite(e).all()), "max_abs_eta": float(np.abs(e).max()),
            "energy": E.tolist(), "wall_s": dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="0,1,2")
    ap.add_argument("--out", default="")
    ap.add_argument("--threads", type=int, default=min(os.cpu_count() or 1, 16))
    a = ap.parse_args()
    O.set_threads(a.threads)
    only = set(int(x) for x in a.only.split(","))
    rep = {}
    if 0 in only:
        print("[validate] config 0: 64^3, 1000 steps, FP64 plain and FP32 fused", flush=True)
        rep["config0_fp64_plain"] = mart_case(64, 1000, gsk.PATH_PLAIN64)
        print(json.dumps(rep["config0_fp64_plain"]), flush=True)
        rep["config0_fp32_fused"] = mart_case(64, 1000, gsk.PATH_FUSED32)
        print(json.dumps(rep["config0_fp32_fused"]), flush=True)
    if 1 in only:
        print("[validate] config 1 physics at 64^3, 2000 steps", flush=True)
        rep["config1_prec_64"] = prec_case(64, 2000)
        print(json.dumps(rep["config1_prec_64"]), flush=True)
    if 2 in only:
        print("[validate] config 2 physics (noise 0.002) at 64^3 vs oracle, and 512^3 run", flush=True)
        rep["config2_noise_64"] = mart_case(64, 500, gsk.PATH_FUSED32, noise=0.002, seed=2)
        print(json.dumps(rep["config2_noise_64"]), flush=True)
        rep["config2_512"] = mart_512_property()
        print(json.dumps(rep["config2_512"]), flush=True)
        for n in (256, 512):
            r = gsk.fft_compare(n, count=1, reps=10)
            r3 = gsk.fft_compare(n, count=3, reps=5)
            alg = 3 * 8 * n**3  # three axis passes, 8 B/voxel each, per field
            rep[f"fft_compare_{n}"] = {
                "one_field": r, "three_fields": r3,
                "hand_r2c_GBps_algorithmic": alg / (r["hand_r2c_ms"] / 1e3) / 1e9,
                "rocfft_r2c_GBps_algorithmic": alg / (r["rocfft_r2c_ms"] / 1e3) / 1e9}
            print(json.dumps({f"fft_compare_{n}": rep[f"fft_compare_{n}"]}), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()

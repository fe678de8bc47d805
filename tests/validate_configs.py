"""Accuracy report of the benchmark workloads against the FP64 oracle: the 64^3 martensite
step (FP64 plain and FP32 fused, 1000 steps), the precipitate physics at 64^3 (2000 steps),
the martensite step with noise (64^3 against the oracle, 512^3 run, hand FFT vs rocFFT) and
the adjoint (32^3 against the oracle and finite differences, memory and time at 128^3).

Not collected by pytest (long oracle runs). Usage on a GPU machine:
    python tests/validate_configs.py [--only mart64,prec64,noise,adjoint] [--out report.json]
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
    """The 512^3 martensite step with noise 0.002: finite, bounded; energy non-increasing is not
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


def adjoint_case(n=32, steps=200, K=20, fd=True):
    """Adjoint of the martensite step: 3 variants, noise off; loss = 1/2 |eta(T) - eta_target|^2 with
    eta_target from the forward model (initial field seed 7) at perturbed parameters
    (s = 1.05/0.95/1.0, L x1.1, kappa x0.9). GPU checkpointed adjoint vs the oracle's FP64
    store-all adjoint, and the oracle adjoint vs central differences (h = 1e-6)."""
    V = 3
    C = gsk.cubic_modulus(0.7, 0.3, 0.2)
    eps = gsk.mart_eigenstrains(V)
    cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, w_el=20.0)
    pert = dict(s=np.array([1.05, 0.95, 1.0]), L=cfg.L * 1.1, kappa=cfg.kappa * 0.9)
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.init_uniform("eta", -0.05, 0.05, 7)
        eta0 = s.get_field("eta")
    p = O.mart_params(A=cfg.A, B=cfg.B, C=cfg.C, kappa=cfg.kappa, L=cfg.L, dt=cfg.dt, w_el=cfg.w_el)
    pp = O.mart_params(A=cfg.A, B=cfg.B, C=cfg.C, kappa=pert["kappa"], L=pert["L"], dt=cfg.dt, w_el=cfg.w_el)
    target = O.mart_run(eta0.astype(np.float64), O.btensor(n, C, eps * pert["s"][:, None, None]), pp, steps)
    B = O.btensor(n, C, eps)
    eT = O.mart_run(eta0.astype(np.float64), B, p, steps)
    t0 = time.perf_counter()
    we, wp = O.mart_adjoint(eta0.astype(np.float64), B, p, steps, eT - target)
    t_or = time.perf_counter() - t0
    with gsk.Sim(cfg) as s:
        s.build_btensor(C, eps)
        s.set_field("eta", eta0)
        s.adjoint_forward(steps, K)
        gT = (s.get_field("eta", np.float64) - target).astype(np.float32)
        ge, gp = s.adjoint_backward(gT)
        info = s.adjoint_info()
    out = {"n": n, "steps": steps, "K": K, "loss": 0.5 * float(((eT - target) ** 2).sum()),
           "rel_l2_grad_eta0": rel_l2(ge, we), "grad_params_gpu": gp.tolist(), "grad_params_oracle": wp.tolist(),
           "rel_err_params": [abs(a - b) / abs(b) for a, b in zip(gp, wp)], "oracle_adjoint_s": t_or,
           "gpu_info": info}
    if fd:
        h = 1e-6

        def loss(**kw):
            q = O.mart_params(A=cfg.A, B=cfg.B, C=cfg.C, kappa=kw.get("kappa", cfg.kappa), L=kw.get("L", cfg.L),
                              dt=cfg.dt, w_el=cfg.w_el)
            sc = kw.get("s", np.ones(V))
            e = O.mart_run(kw.get("e0", eta0.astype(np.float64)), O.btensor(n, C, eps * sc[:, None, None]), q, steps)
            return 0.5 * float(((e - target) ** 2).sum())

        fdv = {}
        fdv["L"] = (loss(L=cfg.L + h) - loss(L=cfg.L - h)) / (2 * h)
        fdv["kappa"] = (loss(kappa=cfg.kappa + h) - loss(kappa=cfg.kappa - h)) / (2 * h)
        for q in range(V):
            sp, sm = np.ones(V), np.ones(V)
            sp[q] += h
            sm[q] -= h
            fdv[f"s{q + 1}"] = (loss(s=sp) - loss(s=sm)) / (2 * h)
        d = np.random.default_rng(1).standard_normal(eta0.shape)
        fdv["eta0_dir"] = (loss(e0=eta0 + h * d) - loss(e0=eta0 - h * d)) / (2 * h)
        adj = {"L": wp[V], "kappa": wp[V + 1], "eta0_dir": float((we * d).sum())}
        adj.update({f"s{q + 1}": wp[q] for q in range(V)})
        out["fd_vs_oracle_adjoint_rel"] = {k: abs(fdv[k] - adj[k]) / abs(fdv[k]) for k in fdv}
    return out


def adjoint_memory_case(n=128, steps=200):
    """Adjoint at 128^3: checkpointed (K = 20) vs store-all (K = 1) memory and time."""
    V = 3
    C = gsk.cubic_modulus(0.7, 0.3, 0.2)
    cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=V, w_el=20.0)
    res = {}
    grads = {}
    for K in (20, 1):
        with gsk.Sim(cfg) as s:
            s.build_btensor(C, gsk.mart_eigenstrains(V))
            s.init_uniform("eta", -0.05, 0.05, 7)
            s.adjoint_forward(steps, K)
            gT = s.get_field("eta")  # d/deta(T) of 1/2 |eta(T)|^2
            ge, gp = s.adjoint_backward(gT)
            res[f"K{K}"] = s.adjoint_info()
            grads[K] = (ge, gp)
        print(f"[validate] adjoint {n}^3 K={K}: {res[f'K{K}']}", flush=True)
    res["bit_identical_K20_vs_K1"] = bool(np.array_equal(grads[20][0], grads[1][0]) and
                                          np.array_equal(grads[20][1], grads[1][1]))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="mart64,prec64,noise,adjoint")
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--out", default="")
    ap.add_argument("--threads", type=int, default=min(os.cpu_count() or 1, 16))
    a = ap.parse_args()
    O.set_threads(a.threads)
    only = set(a.only.split(","))
    rep = {}
    if "mart64" in only:
        print("[validate] martensite 64^3, 1000 steps, FP64 plain and FP32 fused", flush=True)
        rep["mart64_fp64_plain"] = mart_case(64, 1000, gsk.PATH_PLAIN64)
        print(json.dumps(rep["mart64_fp64_plain"]), flush=True)
        rep["mart64_fp32_fused"] = mart_case(64, 1000, gsk.PATH_FUSED32)
        print(json.dumps(rep["mart64_fp32_fused"]), flush=True)
    if "prec64" in only:
        print("[validate] precipitate physics at 64^3, 2000 steps", flush=True)
        rep["prec64"] = prec_case(64, 2000)
        print(json.dumps(rep["prec64"]), flush=True)
    if "noise" in only:
        print("[validate] martensite with noise 0.002: 64^3 vs oracle, and a 512^3 run", flush=True)
        rep["mart64_noise"] = mart_case(64, 500, gsk.PATH_FUSED32, noise=0.002, seed=2)
        print(json.dumps(rep["mart64_noise"]), flush=True)
        rep["mart512_noise"] = mart_512_property()
        print(json.dumps(rep["mart512_noise"]), flush=True)
        for n in (256, 512):
            r = gsk.fft_compare(n, count=1, reps=10)
            r3 = gsk.fft_compare(n, count=3, reps=5)
            alg = 3 * 8 * n**3  # three axis passes, 8 B/voxel each, per field
            rep[f"fft_compare_{n}"] = {
                "one_field": r, "three_fields": r3,
                "hand_r2c_GBps_algorithmic": alg / (r["hand_r2c_ms"] / 1e3) / 1e9,
                "rocfft_r2c_GBps_algorithmic": alg / (r["rocfft_r2c_ms"] / 1e3) / 1e9}
            print(json.dumps({f"fft_compare_{n}": rep[f"fft_compare_{n}"]}), flush=True)
    if "adjoint" in only:
        print("[validate] adjoint at 32^3 vs oracle and finite differences; memory/time at 128^3", flush=True)
        rep["adjoint32"] = adjoint_case(32, a.adj_steps)
        print(json.dumps(rep["adjoint32"]), flush=True)
        rep["adjoint128_memory"] = adjoint_memory_case(128, 200)
        print(json.dumps(rep["adjoint128_memory"]), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()

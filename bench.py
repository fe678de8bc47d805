"""Benchmark: Mvoxel-steps/s of the phase-field time step on one MI355X.

Workload: 256^3 periodic grid, 3-variant martensite (semi-implicit Allen-Cahn with
Khachaturyan elasticity, Landau A=0.2 B=3 C=3, kappa=L=1, dt=0.01, cubic
C11/C12/C44 = 0.7/0.3/0.2, tetragonal eigenstrains), FP32 fused spectral path,
synthetic initial fields (Philox uniform in [-0.01, 0.01], seed 0), inputs resident
in HBM. The FP64 plain path is timed beside it, and the FP64 CPU oracle is timed on
a bounded sample of the same workload on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "general-simulator_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0

# kernel symbol (a regular expression over the demangled name) behind each per-kernel timing
# slot of the fused steps; the x updates in step mode (last template argument 0), not the
# state initialisation
SLOT_KERNEL = {
    "y_fwd_S": r"k_ypass<", "y_inv_R": r"k_ypass<", "y_fwd_F": r"k_ypass<", "y_inv_eta": r"k_ypass<",
    "y_fwd": r"k_ypass<", "y_inv": r"k_ypass<", "x_contract": r"k_xcontract(_r)?<|k_xcontract_split<\d+, \d+, 0>", "zmid_force": r"k_zmid_mart<",
    "x_update_g0": r"k_xupdate_mart<\d+, \d+, 0>|k_xupdate_split<\d+, 0>", "x_update": r"k_xupdate_prec<\d+, \d+, 0>",
    "zinv_eta_sq": r"k_zinv_sq<", "zinv_local": r"k_zinv_local_r<", "plane_mid": r"k_plane_mid<", "plane_sq": r"k_plane_sq<",
    "sub_mid": r"k_sub_mid<", "sub_sq": r"k_sub_sq<",
}
CALIB_MIB = 256


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calib_copy(torch, mib):
    """A device-to-device copy of `mib` MiB: reads and writes exactly that many bytes (the
    reference the HBM counters are scaled by)."""
    nb = mib << 20
    src = torch.empty(nb // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    return nb


def measure_traffic(a, slot):
    """HBM bytes per launch of the kernel behind timing slot `slot`: two short rocprofv3 counter
    runs of the same workload (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, averaged over the
    kernel's dispatches), each scaled by what that counter reads for a 256 MiB device-to-device
    copy in the same run. None when rocprofv3 is missing or a run fails."""
    import csv
    import glob
    import re
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    under_profiler = any(k.startswith("ROCPROF") for k in os.environ) or "rocprof" in os.environ.get("LD_PRELOAD", "")
    if not prof or slot not in SLOT_KERNEL or under_profiler:
        return None
    out = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="gsk_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
               os.path.abspath(__file__), "--workload", a.workload, "--n", str(a.n), "--nvar", str(a.nvar),
               "--noise", str(a.noise), "--steps", "2", "--warmup", "1", "--profile-steps", "1", "--no-cpu",
               "--no-plain", "--no-traffic", "--repeats", "1", "--calib-copy-mib", str(CALIB_MIB)]
        try:
            subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, timeout=150, check=True)
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            kern, calib = [], []
            for r in csv.DictReader(open(files[0])):
                v = float(r["Counter_Value"])
                if re.search(SLOT_KERNEL[slot], r["Kernel_Name"]):
                    kern.append(v)
                elif "copyBuffer" in r["Kernel_Name"]:
                    calib.append(v)
        except Exception as e:  # noqa: BLE001 - the counter run is optional
            log(f"traffic: {ctr} run failed ({type(e).__name__})")
            return None
        finally:
            shutil.rmtree(d, ignore_errors=True)
        if not kern or not calib:
            return None
        # the largest copy dispatch is the calibration copy: CALIB_MIB MiB read and written
        out[ctr] = sum(kern) / len(kern) * (CALIB_MIB * 1024.0 / max(calib)) * 1024.0
    return round(out["FETCH_SIZE"] + out["WRITE_SIZE"])


def make_sim(gsk, n, nvar, path, seed=0, noise=0.0):
    cfg = gsk.default_config(gsk.PHYS_MART, n=n, nvar=nvar, path=path, device=0, noise=noise, seed=seed)
    s = gsk.Sim(cfg)
    s.build_btensor(gsk.cubic_modulus(0.7, 0.3, 0.2), gsk.mart_eigenstrains(nvar))
    s.init_uniform("eta", -0.01, 0.01, seed)
    return s


def timed(sim, steps, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_ms = sim.time_steps(steps)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3, ev_ms


def timed_repeats(sim, steps, repeats, torch):
    """`repeats` timed passes of exactly `steps` steps each; returns the pass of median wall time
    (wall ms, event ms) and every pass's wall ms."""
    runs = [timed(sim, steps, torch) for _ in range(max(1, repeats))]
    order = sorted(range(len(runs)), key=lambda i: runs[i][0])
    med = runs[order[len(runs) // 2]]
    return med[0], med[1], [r[0] for r in runs]


def repeat_stats(walls, steps):
    ms = sorted(w / steps for w in walls)
    return {"count": len(ms), "ms_per_step": [round(x, 4) for x in ms],
            "spread": round((ms[-1] - ms[0]) / ms[len(ms) // 2], 4)}


def cpu_threads():
    """Threads the host offers this process: OMP_NUM_THREADS when set (a shared machine may set it
    to the process's CPU share), else the CPUs in the process's affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(n, nvar, budget_s=8.0, max_steps=10):
    """The FP64 oracle (C++/OpenMP, its own FFT) on the host cores, on a bounded sample of the
    same workload: steps are run one at a time until `budget_s` seconds have passed (at most
    `max_steps`) on every offered thread, then one step on a single thread."""
    import numpy as np

    import gsk
    import oracle as O

    threads = cpu_threads()
    C = gsk.cubic_modulus(0.7, 0.3, 0.2)
    eps = gsk.mart_eigenstrains(nvar)
    O.set_threads(threads)
    B = O.btensor(n, C, eps)
    eta = np.stack([O.uniform_field(0, v, n**3, -0.01, 0.01).reshape(n, n, n) for v in range(nvar)]).astype(np.float64)
    p = O.mart_params()
    steps, dt = 0, 0.0
    while steps < max_steps and dt < budget_s:
        t0 = time.perf_counter()
        eta = O.mart_run(eta, B, p, 1, step0=steps)
        dt += time.perf_counter() - t0
        steps += 1
    O.set_threads(1)
    t0 = time.perf_counter()
    O.mart_run(eta, B, p, 1, step0=steps)
    dt1 = time.perf_counter() - t0
    O.set_threads(threads)
    return {"value": n**3 * steps / dt / 1e6, "unit": "Mvoxel-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the {n}^3 {nvar}-variant martensite step in the FP64 oracle "
                      f"(C++/OpenMP, own FFT) on {threads} threads, {dt:.1f} s; then 1 step on 1 thread, {dt1:.1f} s",
            "single_thread_value": n**3 / dt1 / 1e6, "nproc": os.cpu_count(), "threads_offered": threads}


def ensemble_configs(gsk, master_seed=3, per_size=32, sizes=(64, 128, 192), nvar=3):
    """Members of the ensemble workload: per_size instances at each grid size. Member i
    draws A ~ U[0.15, 0.25], kappa ~ U[0.5, 1.5], dt ~ U[0.005, 0.01] from word 0 of the
    Philox4x32-10 blocks at counters (i, 0, 0, 2), (i, 1, 0, 2), (i, 2, 0, 2) and its 64-bit
    seed from words 0-1 of the block at (i, 3, 0, 2), all keyed by master_seed; u = (x >> 8) 2^-24."""
    cfgs, i = [], 0
    for n in sizes:
        for _ in range(per_size):
            A = 0.15 + 0.10 * gsk.philox_uniform(master_seed, i, 0)
            kap = 0.5 + 1.0 * gsk.philox_uniform(master_seed, i, 1)
            dt = 0.005 + 0.005 * gsk.philox_uniform(master_seed, i, 2)
            w = gsk.philox4x32([i, 3, 0, 2], master_seed)
            seed = int(w[0]) | (int(w[1]) << 32)
            cfgs.append(gsk.default_config(gsk.PHYS_MART, n=n, nvar=nvar, A=A, kappa=kap, dt=dt, seed=seed, device=0))
            i += 1
    return cfgs


def run_ensemble(a):
    """96 independent martensite instances (32 each at 64^3, 128^3, 192^3) on one GPU:
    autotuned schedule (hipGraph replay or streams) against the members run one after another."""
    import torch

    import gsk

    cfgs = ensemble_configs(gsk, per_size=a.ens_per_size)
    vox = sum(c.n**3 for c in cfgs)
    log(f"ensemble of {len(cfgs)} members, {vox} voxels per ensemble step")
    e = gsk.Ensemble(cfgs)
    e.build_btensor(gsk.cubic_modulus(0.7, 0.3, 0.2), gsk.mart_eigenstrains(3))
    for m in e.members:
        m.init_uniform("eta", -0.01, 0.01, m.cfg.seed)
    table = e.autotune(trial_steps=a.ens_trial)
    best = e.schedule()
    log(f"autotune picked mode={best[0]} branches={best[1]} steps/graph={best[2]}")
    e.step(a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.step(a.steps)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    rate = vox * a.steps / t / 1e6
    e.set_schedule(gsk.SCHED_SEQUENTIAL, 1, 1)
    t0 = time.perf_counter()
    e.step(a.ens_seq_steps)
    ts = time.perf_counter() - t0
    seq = vox * a.ens_seq_steps / ts / 1e6
    finite = all(bool(__import__("numpy").isfinite(m.get_field("eta")).all()) for m in e.members)
    e.close()
    # the same members one by one on the FP64 plain path (one kernel per operation, C2C transforms):
    # each member's own handle, a warm-up step, then ens_plain_steps steps timed with HIP events
    plain_ms = 0.0
    if a.ens_plain_steps > 0:
        C, eps = gsk.cubic_modulus(0.7, 0.3, 0.2), gsk.mart_eigenstrains(3)
        for c in cfgs:
            c.path = gsk.PATH_PLAIN64
            with gsk.Sim(c) as m:
                m.build_btensor(C, eps)
                m.init_uniform("eta", -0.01, 0.01, c.seed)
                m.step(1)
                plain_ms += m.time_steps(a.ens_plain_steps) / a.ens_plain_steps
        log(f"plain path one by one: {plain_ms:.2f} ms per ensemble step")
    names = {0: "sequential", 1: "streams", 2: "graph", 3: "batched", 4: "batched_graph"}
    ms = t / a.steps * 1e3
    result = {
        "metric": "aggregate Mvoxel-steps/s, 96-instance martensite ensemble (32 each at 64^3, 128^3, 192^3), FP32",
        "value": round(rate, 2), "unit": "Mvoxel-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (per-member Philox parameters and initial fields)",
        "config": {"workload": "ensemble: 32 x 64^3 + 32 x 128^3 + 32 x 192^3, 3 variants, semi-implicit AC + "
                               "Khachaturyan elasticity, A/kappa/dt drawn per member", "members": len(cfgs),
                   "voxels_per_step": vox, "parallelism": "single GPU"},
        "hbm_frac_canonical": round(264 * vox / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "schedule": {"mode": names[best[0]], "branches": best[1], "steps_per_graph": best[2]},
        "sequential": {"value": round(seq, 2), "steps": a.ens_seq_steps},
        "speedup_vs_sequential": round(rate / seq, 3),
        "plain_one_by_one": ({"value": round(vox / (plain_ms / 1e3) / 1e6, 2), "ms_per_step": round(plain_ms, 3),
                              "steps": a.ens_plain_steps, "path": "FP64 plain, members one after another"}
                             if plain_ms > 0 else None),
        "speedup_vs_plain_one_by_one": round(plain_ms / ms, 3) if plain_ms > 0 else None,
        "autotune": [{"mode": names[int(r[0])], "branches": int(r[1]), "steps_per_graph": int(r[2]),
                      "ms_per_step": round(r[3], 4)} for r in table],
        "finite": finite,
    }
    print(json.dumps(result), flush=True)


def run_prec(a):
    """Precipitate physics at 256^3 (4 order parameters + c, FP32 fused, S = A1), with the
    exact FP64 mass check; 242 B/voxel-step is the canonical fused count for V = 4."""
    import numpy as np
    import torch

    import gsk

    n, V = a.n, 4
    cfg = gsk.default_config(gsk.PHYS_PREC, n=n, nvar=V, M=1.0, L=1.0, kappa=1.5, kappa_c=1.5, dt=0.005,
                             noise=0.001, seed=5, S=54.0, device=0)
    s = gsk.Sim(cfg)
    s.build_btensor(gsk.cubic_modulus(3.5, 1.5, 1.0), 0.02 * np.eye(3)[None])
    s.init_uniform("concentration", 0.245, 0.255, 11)
    s.init_uniform("eta", 0.0, 0.01, 12)
    m0 = s.mass()
    s.step(a.warmup)
    wall_ms, ev_ms, walls = timed_repeats(s, a.steps, a.repeats, torch)
    ms = wall_ms / a.steps
    names, kms = s.kernel_times(a.profile_steps)
    kbytes = s.kernel_bytes()
    per = [(names[i], kms[i] / a.profile_steps, kbytes[i]) for i in range(len(names))]
    dom = max(per, key=lambda t: t[1])
    m1 = s.mass()
    finite = bool(np.isfinite(s.get_field("eta")).all())
    s.close()
    canon = (48 * (V + 1) + 2) * n**3
    result = {
        "metric": f"Mvoxel-steps/s, {n}^3 precipitate (CH + 4 AC order parameters, c-sourced elasticity), FP32 fused",
        "value": round(n**3 * a.steps / (wall_ms / 1e3) / 1e6, 2), "unit": "Mvoxel-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (Philox uniform c and eta)",
        "config": {"workload": f"precipitate {n}^3, 4 order parameters, M=L=1, kappa=1.5, dt=0.005, S=54",
                   "n": n, "nvar": V, "parallelism": "single GPU"},
        "hbm_frac_canonical": round(canon / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "event_ms_per_step": round(ev_ms / a.steps, 4),
        "repeats": repeat_stats(walls, a.steps),
        "kernels": {nm: {"ms": round(t, 4), "GBps": round(b / (t / 1e3) / 1e9, 1)} for nm, t, b in per},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(dom[2] / (dom[1] / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom[2] / (dom[1] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "mass_drift_rel": abs(m1 - m0) / m0,
        "finite": finite,
    }
    if a.calib_copy_mib:
        result["calib_copy_bytes"] = calib_copy(torch, a.calib_copy_mib)
    if not a.no_traffic:
        result["roofline"]["traffic"] = measure_traffic(a, dom[0])
    print(json.dumps(result), flush=True)


def run_box(a):
    """A box grid (nx x ny x nz, spacings 1, 0.5, 2): the 3-variant spectral martensite step on the box
    path, FP32 and FP64, inputs resident; one kernel per operation with C2C transforms per axis."""
    import torch

    import gsk

    dims = tuple(a.box)
    vox = dims[0] * dims[1] * dims[2]
    out = {}
    for path, name in ((gsk.PATH_FUSED32, "f32"), (gsk.PATH_PLAIN64, "f64")):
        cfg = gsk.default_config(gsk.PHYS_MART, n=dims[0], ny=dims[1], nz=dims[2], dx=1.0, dy=0.5, dz=2.0, nvar=3,
                                 path=path, device=0)
        with gsk.Sim(cfg) as s:
            s.build_btensor(gsk.cubic_modulus(0.7, 0.3, 0.2), gsk.mart_eigenstrains(3))
            s.init_uniform("eta", -0.01, 0.01, 0)
            s.step(2)
            steps = max(2, a.steps // (4 if name == "f32" else 10))
            wall, _, walls = timed_repeats(s, steps, a.repeats, torch)
        ms = wall / steps
        out[name] = {"value": round(vox / (ms / 1e3) / 1e6, 2), "ms_per_step": round(ms, 4), "steps": steps,
                     "repeats": repeat_stats(walls, steps)}
        log(f"box {dims} {name}: {ms:.3f} ms/step")
    result = {
        "metric": f"Mvoxel-steps/s, {dims[0]}x{dims[1]}x{dims[2]} box, 3-variant AC+elasticity, box path",
        "value": out["f32"]["value"], "unit": "Mvoxel-steps/s", "n_gpus": 1, "steps": out["f32"]["steps"],
        "warmup": 2, "ms_per_step": out["f32"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (Philox uniform initial fields, seed 0)",
        "config": {"workload": f"martensite box {dims[0]}x{dims[1]}x{dims[2]}, spacings 1/0.5/2, 3 variants, "
                               "box path (C2C per axis, one kernel per operation)", "parallelism": "single GPU"},
        "box_f32": out["f32"], "box_f64": out["f64"],
    }
    print(json.dumps(result), flush=True)


def run_ref(a):
    """The martensite plugin's own rule (explicit Euler with the elastic R from the fused transforms,
    src/dynamics_mart.cu:348-417) and the precipitate plugin's (src/dynamics_precipitate.cu:278-289)
    at n^3: what an unmodified reference script runs."""
    import torch

    import gsk

    n = a.n
    out = {}
    for phys, name in ((gsk.PHYS_MART, "mart_reference"), (gsk.PHYS_PREC, "prec_reference")):
        V = 3
        cfg = gsk.reference_config(phys, n=n, nvar=V, device=0, temperature=300.0, w_el=1e3, dt=0.002)
        with gsk.Sim(cfg) as s:
            s.build_btensor(gsk.cubic_modulus(3.5, 1.5, 1.0), gsk.mart_eigenstrains(V, base=0.0, axis=0.01))
            if phys == gsk.PHYS_PREC:
                s.init_uniform("concentration", 0.245, 0.255, 1)
            s.init_uniform("eta", 0.0, 0.1, 2)
            s.step(a.warmup)
            wall, _, walls = timed_repeats(s, a.steps, a.repeats, torch)
            ksteps = max(1, a.profile_steps)
            names, kms = s.kernel_times(ksteps)
            kms = kms / ksteps  # per step
        ms = wall / a.steps
        out[name] = {"value": round(n**3 / (ms / 1e3) / 1e6, 2), "ms_per_step": round(ms, 4),
                     "repeats": repeat_stats(walls, a.steps),
                     "kernels_us": {k: round(1e3 * float(t), 1) for k, t in zip(names, kms)}}
        log(f"{name} n={n}: {ms:.4f} ms/step")
    result = {
        "metric": f"Mvoxel-steps/s, {n}^3 3-variant martensite, the plugin's own rule (FP32)",
        "value": out["mart_reference"]["value"], "unit": "Mvoxel-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": out["mart_reference"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (Philox uniform fields)",
        "config": {"workload": f"reference rules at {n}^3, 3 variants", "parallelism": "single GPU"},
        "rules": out,
    }
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nvar", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--noise", type=float, default=0.0, help="noise amplitude (0.002 for the 512^3 configuration)")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="time budget of the multi-threaded CPU sample")
    ap.add_argument("--repeats", type=int, default=3, help="timed passes of --steps steps; the median is reported")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs for roofline.traffic")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this size once (counter calibration)")
    ap.add_argument("--workload", choices=["mart256", "ensemble", "prec256", "box", "reference"], default="mart256",
                    help="mart256: the headline 256^3 step; ensemble: 96 mixed-size instances; "
                         "prec256: the precipitate step at 256^3; box: a box grid on the box path; "
                         "reference: the plugins' own rules at --n")
    ap.add_argument("--box", type=int, nargs=3, default=[256, 256, 128], help="box workload: nx ny nz")
    ap.add_argument("--ens-per-size", type=int, default=32)
    ap.add_argument("--ens-trial", type=int, default=5)
    ap.add_argument("--ens-seq-steps", type=int, default=20)
    ap.add_argument("--ens-plain-steps", type=int, default=3,
                    help="ensemble: steps per member timed on the FP64 plain path, one member after another (0: skip)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: the library runs on one GPU "
              "(how the step would shard across GPUs is a design note only)", file=sys.stderr)
        sys.exit(2)

    if a.workload == "ensemble":
        run_ensemble(a)
        return
    if a.workload == "prec256":
        run_prec(a)
        return
    if a.workload == "box":
        run_box(a)
        return
    if a.workload == "reference":
        run_ref(a)
        return

    import torch

    import gsk

    n, V = a.n, a.nvar
    log(f"fused32 n={n} V={V}")
    s = make_sim(gsk, n, V, gsk.PATH_FUSED32, noise=a.noise)
    s.step(a.warmup)
    wall_ms, ev_ms, walls = timed_repeats(s, a.steps, a.repeats, torch)
    ms = wall_ms / a.steps
    names, kms = s.kernel_times(a.profile_steps)
    kbytes = s.kernel_bytes()
    per = [(names[i], kms[i] / a.profile_steps, kbytes[i]) for i in range(len(names))]
    dom = max(per, key=lambda t: t[1])
    import numpy as np

    eta = s.get_field("eta")
    finite = bool(np.isfinite(eta).all())
    s.close()
    mvox = n**3 * a.steps / (wall_ms / 1e3) / 1e6
    canon_bytes = (84 * V + V * (V + 1)) * n**3  # canonical fused schedule of nine kernels, bytes per step
    sched_bytes = (80 * V + V * (V + 1)) * n**3  # this schedule: the last z pass also transforms eta^2
    result = {
        "metric": f"Mvoxel-steps/s, {n}^3 {V}-variant AC+elasticity, FP32 fused spectral step",
        "value": round(mvox, 2),
        "unit": "Mvoxel-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox uniform initial fields, seed 0)",
        "config": {"workload": f"martensite {n}^3, {V} variants, semi-implicit AC + Khachaturyan elasticity, "
                               "FP32 fused spectral path", "n": n, "nvar": V, "dt": 0.01, "noise": a.noise,
                   "parallelism": "single GPU"},
        "hbm_frac_canonical": round(canon_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "hbm_frac_schedule": round(sched_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "event_ms_per_step": round(ev_ms / a.steps, 4),
        "repeats": repeat_stats(walls, a.steps),
        "kernels": {nm: {"ms": round(t, 4), "GBps": round(b / (t / 1e3) / 1e9, 1)} for nm, t, b in per},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(dom[2] / (dom[1] / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom[2] / (dom[1] / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None},
        "finite": finite,
    }
    if not a.no_plain:
        log(f"plain64 n={n} V={V}")
        p = make_sim(gsk, n, V, gsk.PATH_PLAIN64)
        p.step(2)
        pw, _ = timed(p, a.plain_steps, torch)
        p.close()
        pms = pw / a.plain_steps
        result["fp64_plain"] = {"value": round(n**3 / (pms / 1e3) / 1e6, 2), "ms_per_step": round(pms, 3)}
        result["speedup_vs_fp64_plain"] = round(pms / ms, 2)
    if not a.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(n, V, budget_s=a.cpu_seconds)
    if a.calib_copy_mib:
        result["calib_copy_bytes"] = calib_copy(torch, a.calib_copy_mib)
    if not a.no_traffic:
        log("traffic (rocprofv3 counters)")
        result["roofline"]["traffic"] = measure_traffic(a, dom[0])
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the hyperbolic-relaxation hot path on one MI355X.

Workload: the 128x128x16 SinhSymTP equal-mass boosted-puncture solve (m = 0.5 each at
z = +-5, P = (0, +-0.0953, 0), no spins; AMAX 1e6, SINHWAA 0.07, bScale 5, eta 20, CFL 0.7),
zero initial guess, FP64.  A "step" is one RK4 step (4 fused substep kernels, residual + L2
reduced every step as the reference driver does), graph-replayed from HBM-resident state.

Prints ONE JSON line: gp-substeps/s (4 * 262144 interior points * steps / s) for exactly
--steps steps after --warmup steps, the roofline of the fused substep kernel, the compiled
reference C timed on this host's cores on a bounded sample, and the time to log10 L2 <= -10.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-residual 1e-10 (NN-seeded vs plain FP64); grid-pt RK4 substeps/s, %HBM"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per gp-substep of the fused RK4 substep (FP64): stage k1 72 B, k2/k3 104 B,
# k4 88 B (reads of stage input, aux c/psi/AA2, y_n, running total; writes of stage output and
# running total), mean 92 B
BYTES_PER_GP_SUBSTEP = 92.0
# 9000 steps at 128x128x16 in 146.68 s on 6 cores (Tutorial-Start_to_Finish-NRPyElliptic.ipynb:1477)
PUBLISHED_GP_SUBSTEPS = 4 * 262144 * 9001 / 146.68


def cpu_baseline(steps: int, warmup: int, reps: int = 5, admbase: bool = True) -> dict:
    """The reference driver's loop body (residual + L2 + MoL step, main.c:138-187) of the
    compiled reference C at 128x128x16, or the C restatement when the reference build is
    absent, in `reps` repetitions of steps/reps steps. Runs in this (child) process; OMP
    threads come from the environment."""
    import oracle
    from nrpyelliptic_amd import config

    cfg = config.equal_mass_boosted()
    kind = "reference" if oracle.ref_lib() is not None else "port"
    if kind == "reference":
        R = oracle.RefProblem(cfg)

        def step():
            R.l2(R.residual(R.arr("y_n_gfs")))
            R.mol_step()
    else:
        O = oracle.OracleProblem(cfg)
        st = O.new_state()

        def step():
            O.log10_residual(st["y_n"])
            O.mol_step(st)
    for _ in range(warmup):
        step()
    per = max(1, steps // max(1, reps))
    times = []
    for _ in range(max(1, reps)):
        t0 = time.perf_counter()
        for _ in range(per):
            step()
        times.append(time.perf_counter() - t0)
    dt = sum(times)
    ms = sorted(1e3 * t / per for t in times)
    out = {"value": 4 * 262144 * per * len(times) / dt, "unit": "gp-substeps/s", "kind": kind, "seconds": dt,
           "ms_per_step": 1e3 * dt / (per * len(times)), "ms_per_step_min": ms[0],
           "ms_per_step_median": ms[len(ms) // 2], "repetitions": len(times), "steps_per_repetition": per}
    if admbase:
        out["admbase"] = cpu_admbase(cfg)
    return out


def cpu_admbase(cfg, n=1 << 20):
    """The consumer path (Cartesian points -> ADMBase) of the C restatement, one thread, on n
    uniform points of the evolution grid's box (the thorn's own C needs Cactus and does not
    build here)."""
    import numpy as np
    import oracle
    O = oracle.OracleProblem(cfg)
    X = np.random.default_rng(0).uniform(-20.0, 20.0, size=(n, 3))
    u = np.zeros(O.ntot)
    t0 = time.perf_counter()
    O.initialize_admbase(u, X[:, 0], X[:, 1], X[:, 2])
    dt = time.perf_counter() - t0
    return {"points_per_s": n / dt, "points": n, "seconds": dt, "cores": 1, "kind": "port"}


def _cpu_child(args, threads, timeout=600):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, os.path.abspath(__file__), *args], env=env, capture_output=True, text=True,
                       timeout=timeout)
    return json.loads(r.stdout.strip().splitlines()[-1])


def run_cpu_baseline_subprocess(steps=5000, warmup=5, reps=5, single_thread_steps=100):
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    try:
        out = _cpu_child(["--cpu-baseline-only", str(steps), str(warmup), str(reps)], cores)
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "gp-substeps/s", "cores": cores, "kind": None, "sample": f"failed: {e}"}
    try:  # one thread, a short sample (the reference's 1-thread rate)
        one = _cpu_child(["--cpu-baseline-only", str(single_thread_steps), "2", "1", "--no-admbase"], 1, timeout=300)
        out["single_thread"] = {"value": one["value"], "ms_per_step": one["ms_per_step"], "steps": single_thread_steps,
                                "cores": 1}
    except Exception as e:
        out["single_thread"] = {"value": None, "sample": f"failed: {e}"}
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    out.update(cores=cores, sample=f"{steps} RK4 steps in {reps} repetitions (+{warmup} warmup) of the 128x128x16 "
                                   f"solve from u=v=0, "
                                   f"residual+L2 every step as main.c; OpenMP {cores} threads on '{cpu}'")
    return out


def dropin_call_us(P, n=50):
    """Wall time per call of the reference-surface entry points on the problem's device
    arrays (each call ends with a device synchronize, as a C caller sees it)."""
    import ctypes as C
    from nrpyelliptic_amd import runtime
    L = runtime.lib()
    gd, p = P.gd, P.gd.params
    f = gd.gridfuncs
    ws = (C.c_double * 2)(p.uu_wavespeed_at_OB, p.vv_wavespeed_at_OB)
    finf = (C.c_double * 2)(0.0, 0.0)
    calls = {
        "rhs_eval": lambda: L.rhs_eval(C.byref(p), C.byref(gd.rfmstruct), f.auxevol_gfs, f.y_n_gfs, f.k_odd_gfs),
        "apply_bcs_outerradiation_and_inner": lambda: L.apply_bcs_outerradiation_and_inner(
            C.byref(p), C.byref(gd.bcstruct), gd.xx, ws, finf, f.y_n_gfs, f.k_odd_gfs),
        "residual_all_points": lambda: L.residual_all_points(C.byref(p), C.byref(gd.rfmstruct), f.auxevol_gfs,
                                                             f.y_n_gfs, f.diagnostic_output_gfs),
        "L2_norm_of_gf": lambda: L.L2_norm_of_gf(C.byref(p), 0, 100.0, gd.xx, f.diagnostic_output_gfs),
        "MoL_step_forward_in_time": lambda: L.MoL_step_forward_in_time(C.byref(gd), P.dt),
    }
    out = {}
    for name, fn in calls.items():
        fn()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        out[name] = round(1e6 * (time.perf_counter() - t0) / n, 2)
    return out


ADMBASE_BYTES_PER_POINT = 24 + 17 * 8  # read x, y, z; write the 17 ADMBase arrays


def admbase_throughput(P, n_side=128, half=20.0, repeats=5):
    """The Einstein Toolkit consumer path on the solved state: an n_side^3 Cartesian grid over
    [-half, half]^3 (x fastest, as a Cactus grid function), points and outputs resident on the
    device; wall time per synchronous nrpyell_Initialize_ADMBase call."""
    import ctypes as C
    import numpy as np
    from nrpyelliptic_amd import runtime
    g = np.linspace(-half, half, n_side)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    n = X.size
    dev = [runtime.DeviceArray(n).upload(a.ravel()) for a in (X, Y, Z)]
    outs = [runtime.DeviceArray(n) for _ in runtime.ADMBASE_NAMES]
    o = runtime.admbase_options()
    o.orbital_plane, o.initial_lapse, o.lapse_exponent_n, o.NRPy_epsilon, o.interpolation_order = 0, 0, -2.0, 1e-8, 4
    lsh = (C.c_int * 3)(n_side, n_side, n_side)
    L = runtime.lib()

    def call():
        runtime.check(L.nrpyell_Initialize_ADMBase(C.byref(P.gd), C.byref(o), lsh, *[d.as_real_p() for d in dev],
                                                   *[d.as_real_p() for d in outs]), "Initialize_ADMBase")
    call()
    t0 = time.perf_counter()
    for _ in range(repeats):
        call()
    sec = (time.perf_counter() - t0) / repeats
    for d in dev + outs:
        d.free()
    return {"points": n, "ms_per_call": 1e3 * sec, "points_per_s": n / sec,
            "bytes_per_point": ADMBASE_BYTES_PER_POINT,
            "achieved_GBs": ADMBASE_BYTES_PER_POINT * n / sec / 1e9, "interpolation_order": 4}


def pmc_probe():
    """A short run for rocprofv3 counter passes: 20 production steps of the 128x128x16 solve,
    then 5 plain device copies of 64 MiB (torch), whose FETCH_SIZE / WRITE_SIZE per dispatch
    calibrate the counters against a known byte count."""
    import numpy as np
    import torch
    from nrpyelliptic_amd import config, runtime
    P = runtime.Problem(config.equal_mass_boosted())
    S = P.solver(steps_per_graph=10)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    S.timed_steps(20)
    S1 = P.solver(steps_per_graph=10, substeps_per_launch=1)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    S1.timed_steps(20)
    a = torch.ones(8 * 1024 * 1024, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print("pmc probe: 20 steps of the production path (40 two-substep launches), 20 steps of one substep per "
          "launch (80 launches), 5 copies of 64 MiB", flush=True)


def nn_seeded_paths(P, n_held_out=256):
    """The learned initial guess against the plain method.

    Bench problem (equal-mass boosted punctures, 128x128x16): the reference's path (zero guess,
    configured eta, FP64) against the network's guess with the FP32 -> FP64 solver, at the
    configured eta and at the damping tuned to that start (autotune.EtaTuner, calibrated in FP32
    on a seed-901 configuration of the grid: neither the bench problem, the held-out set nor the
    training set). Per path: steps and seconds to log10 L2 <= -10 (seeding included), max |u - u*|
    against the reference's own u* (tests/golden/deep_equal_mass_boosted_128x128x16.npz: the
    reference's C relaxed to log10 L2 = -14), and the accuracy-matched time: the seconds of the
    first residual target in -10, -10.5, ... whose u is within 1e-10 of u*. A fourth path runs
    the damping schedule autotune.ScheduleTuner picks on the same seed-901 configuration for that
    accuracy bar (learned.solve_damping_schedule: a strong damping down to a switch residual,
    then a weaker one; the strong damping leaves long-wavelength error the masked residual hardly
    sees).

    Held-out (seed 900, n configurations at 128x128x16): learned.evaluate_batched, both starts
    on the batched FP64 solver, the network's at the damping EtaTuner gives it in FP64."""
    import numpy as np
    from nrpyelliptic_amd import config, learned, runtime
    from nrpyelliptic_amd.autotune import EtaTuner, ScheduleTuner
    from nrpyelliptic_amd.mlp import InitialGuessMLP
    w = os.path.join(ROOT, "nrpyelliptic_amd", "data", "mlp_initial_guess.npz")
    if not os.path.exists(w):
        return None
    net = InitialGuessMLP(seed=7)
    net.load(w)
    u_star = np.load(os.path.join(ROOT, "tests", "golden", "deep_equal_mass_boosted_128x128x16.npz"))["u_deep"]
    G0, G1, G2 = P.Ng

    t0 = time.perf_counter()
    calib = runtime.Problem(config.random_batch(1, seed=901, grids=False, Nxx=(128, 128, 16))[0])
    nn_start = lambda Q: learned.seed_with_fallback(net, Q)  # noqa: E731
    tuner32 = EtaTuner(cache_path=None, precision=32, fp32_switches=(-6.0, -7.0, -8.0))
    f_nn32 = tuner32.choose(calib, nn_start, "nn")
    sw_nn32 = tuner32.fp32_switch_for(calib, "nn")
    f_nn64 = EtaTuner(cache_path=None, precision=64).choose(calib, nn_start, "nn")
    f_zero = EtaTuner(cache_path=None, precision=64).choose(calib, lambda Q: Q.zero_guess(), "zero")
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    sched = ScheduleTuner(cache_path=None, precision=32).choose(calib, nn_start, "nn")
    sched_s = time.perf_counter() - t0
    calib.close()

    eta0 = P.gd.params.eta_damping

    def run(nn, factor, target):
        if nn == "schedule":
            f_hi, switch, f_lo = factor
            if target > switch - 0.5:
                return None  # the switch to the weaker damping comes first
            steps, fp32, sec, seed_s, conv, used = learned.solve_damping_schedule(net, P, f_hi, switch, target,
                                                                                    f_lo=f_lo)
            u = P.get("y_n_gfs")[: P.ntot].reshape(G2, G1, G0)[5:-5, 5:-5, 5:-5]
            return {"steps": steps, "fp32_steps": fp32, "seconds": sec + seed_s, "seed_seconds": seed_s,
                    "converged": conv, "nn_used": used, "max_abs_u_minus_ustar": float(np.max(np.abs(u - u_star)))}
        try:
            seed_s = 0.0
            if nn:
                t = time.perf_counter()
                used, _ = learned.seed_with_tuned_damping(net, P, factor, 1.0)
                seed_s = time.perf_counter() - t
            else:
                used = False
                P.gd.params.eta_damping = factor * eta0
                P.zero_guess()
            kw = {"fp32_switch_log10": sw_nn32} if nn and factor != 1.0 else {}
            S = P.solver(precision=32 if nn else 64, log10_target=target, max_steps=40000, **kw)
            r = S.solve()
            S.close()
        finally:
            P.gd.params.eta_damping = eta0
        u = P.get("y_n_gfs")[: P.ntot].reshape(G2, G1, G0)[5:-5, 5:-5, 5:-5]
        return {"steps": r.steps, "fp32_steps": r.fp32_steps, "seconds": r.seconds + seed_s, "seed_seconds": seed_s,
                "log10_l2": r.log10_l2, "converged": bool(r.converged), "nn_used": bool(used),
                "max_abs_u_minus_ustar": float(np.max(np.abs(u - u_star)))}

    # the damping schedule ScheduleTuner chose (learned.solve_damping_schedule)
    paths = {"plain_fp64": (False, 1.0), "nn_mixed_configured_eta": (True, 1.0),
             "nn_mixed_tuned_eta": (True, f_nn32), "nn_mixed_damping_schedule": ("schedule", sched["schedule"])}
    bench = {}
    for name, (nn, f) in paths.items():
        row = run(nn, f, -10.0) or {}
        for tgt in (-10.0, -10.5, -11.0, -11.5, -12.0, -12.5, -13.0, -13.5):
            acc = row if tgt == -10.0 else run(nn, f, tgt)
            if acc and acc["max_abs_u_minus_ustar"] <= 1e-10:
                row["accuracy_matched"] = {"target_log10": tgt, "steps": acc["steps"], "seconds": acc["seconds"],
                                           "max_abs_u_minus_ustar": acc["max_abs_u_minus_ustar"]}
                break
        bench[name] = row
    plain = bench["plain_fp64"]
    for name in ("nn_mixed_configured_eta", "nn_mixed_tuned_eta", "nn_mixed_damping_schedule"):
        r = bench[name]
        if "seconds" in r:
            r["speedup_time_to_residual"] = plain["seconds"] / r["seconds"]
        if "accuracy_matched" in r and "accuracy_matched" in plain:
            r["speedup_accuracy_matched"] = plain["accuracy_matched"]["seconds"] / r["accuracy_matched"]["seconds"]
    held = learned.evaluate_batched(
        net, config.random_batch(n_held_out, seed=900, grids=False, Nxx=(128, 128, 16)), f_nn64, f_zero,
        log=lambda s: None)
    return {"eta_factors": {"nn_fp32_fp64": f_nn32, "nn_fp64": f_nn64, "zero": f_zero},
            "fp32_switch_log10_tuned_eta": sw_nn32,
            "damping_schedule": {"eta_factor_before_switch": sched["schedule"][0],
                                 "switch_log10": sched["schedule"][1],
                                 "eta_factor_after_switch": sched["schedule"][2],
                                 "calibration": "ScheduleTuner on the seed-901 configuration: fewest steps until "
                                                "max|u - u*| <= 1e-10 (u* its zero-guess solve to -14)",
                                 "calibration_target_log10": sched["target"], "calibration_steps": sched["steps"],
                                 "calibration_seconds": sched_s},
            "calibration": "EtaTuner on a seed-901 configuration of the 128x128x16 grid: fewest steps to -10 among "
                           "the eta factors whose u is within half the reference path's error of a solve to -14, "
                           "in the precision the path runs; for the FP32 -> FP64 path also the deepest FP32 -> FP64 "
                           "switch (log10 L2 -6, -7, -8) with no more steps within the same margin",
            "calibration_seconds": tune_s, "bench_problem": bench, "held_out_batched": held}


def larger_grids(sizes=((256, 256, 32), (512, 256, 32)), steps=200, reps=3):
    """The production FP64 path (two-substep kernel, residual every step, graph replay) on grids
    larger than the bench problem (one seed-3 configuration each): microseconds per RK4 step and
    the rate by the same 92 B/pt-per-substep accounting as the roofline object."""
    from nrpyelliptic_amd import config, runtime
    out = {}
    for Nxx in sizes:
        P = runtime.Problem(config.random_batch(1, seed=3, grids=False, Nxx=Nxx)[0])
        try:
            P.zero_guess()
            S = P.solver(steps_per_graph=20)
            best = min(1e3 * S.timed_steps(steps)[0] / steps for _ in range(reps))
            spl = S.substeps_per_launch
            S.close()
        finally:
            P.close()
        gbs = 4 * BYTES_PER_GP_SUBSTEP * Nxx[0] * Nxx[1] * Nxx[2] / (best * 1e-6) / 1e9
        out["x".join(map(str, Nxx))] = {"us_per_step": round(best, 2), "substeps_per_launch": spl,
                                        "accounting_GBs": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 3)}
    return out


def mixed_precision_q4():
    """The FP32 -> FP64 solver on the unequal-mass workload (q = 4, D = 10, seed-7 spins,
    128x128x16) beside the FP64 solver: steps, FP32 steps and seconds to log10 L2 <= -10 from the
    zero guess, and each answer's max |du| and rms du against the reference's FP64 u at its own
    stopping step and against its u* (tests/golden/deep_q4_D10_seed7_128x128x16.npz, written
    from the reference's C)."""
    import numpy as np
    from nrpyelliptic_amd import config, runtime
    d = np.load(os.path.join(ROOT, "tests", "golden", "deep_q4_D10_seed7_128x128x16.npz"))
    u_star = d["u_deep"]
    u_ref = u_star + d["du_stop"].astype(np.float64)
    P = runtime.Problem(config.unequal_mass_q4())
    G0, G1, G2 = P.Ng
    out = {"workload": "q=4, D=10, seed-7 spins, 128x128x16", "reference_fp64_steps": int(d["steps_stop"])}
    try:
        for name, prec in (("fp64", 64), ("fp32_then_fp64", 32)):
            P.zero_guess()
            S = P.solver(precision=prec)
            r = S.solve()
            S.close()
            u = P.get("y_n_gfs")[: P.ntot].reshape(G2, G1, G0)[5:-5, 5:-5, 5:-5]
            du, ds = u - u_ref, u - u_star
            out[name] = {"steps": r.steps, "fp32_steps": r.fp32_steps, "seconds": r.seconds,
                         "log10_l2": r.log10_l2, "converged": bool(r.converged),
                         "max_abs_du_vs_reference_fp64": float(np.max(np.abs(du))),
                         "rms_du_vs_reference_fp64": float(np.sqrt(np.mean(du * du))),
                         "max_abs_u_minus_ustar": float(np.max(np.abs(ds)))}
        out["reference_max_abs_u_minus_ustar"] = float(np.max(np.abs(u_ref - u_star)))
        out["speedup_mixed"] = out["fp64"]["seconds"] / out["fp32_then_fp64"]["seconds"]
    finally:
        P.close()
    return out


def batched_solve(n=64, seed=42):
    """The batched configuration: n random punctures (seed 42: q log-uniform in [1,8], D in
    [6,20], Newtonian tangential momenta, random spins; grids mixing Nxx0 in {96,128,160},
    Nxx1 in {96,128}, Nxx2 in {8,16}) relaxed together to log10 L2 <= -10 with tile shapes from
    the autotuner; members that diverge at CFL 0.7 are relaxed again from the zero guess at the
    fallback CFL (runtime.relax_batch). Throughput counts every substep run."""
    from nrpyelliptic_amd import config, runtime
    from nrpyelliptic_amd.autotune import TileTuner
    probs = [runtime.Problem(c) for c in config.random_batch(n, seed=seed)]
    tuner = TileTuner(cache_path=None)
    t0 = time.perf_counter()
    tiles = tuner.choose_all(probs)
    tune_s = time.perf_counter() - t0
    res, sec, info = runtime.relax_batch(probs, tiles=tiles, log10_target=-10.0, max_steps=30000)
    conv = [(P, r) for P, r in zip(probs, res) if r.converged]
    # every substep run counts, the diverged first attempts of the fallback members included
    pts = sum(4.0 * P.cfg.Nxx[0] * P.cfg.Nxx[1] * P.cfg.Nxx[2] * (r.steps + i["first_steps"])
              for P, r, i in zip(probs, res, info))
    fell = [k for k in range(n) if info[k]["fallback"]]
    out = {"configurations": n, "converged": len(conv), "not_converged": n - len(conv), "seconds": sec,
           "gp_substeps_per_s": pts / sec, "mean_steps": sum(r.steps for _, r in conv) / max(len(conv), 1),
           "fallback_cfl": runtime.FALLBACK_CFL, "fallback_members": fell,
           "fallback_first_attempt_steps": [info[k]["first_steps"] for k in fell],
           "autotune_seconds": tune_s, "tiles": sorted({f"{t[0]}x{t[1]}" for t in tiles})}
    for P in probs:
        P.close()
    return out


def load_dropin_kernels():
    """Kernel durations of the reference-surface entry points from the committed rocprofv3
    summary (null when absent)."""
    try:
        with open(os.path.join(ROOT, "profiles", "dropin_kernels_128x128x16.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def load_traffic(spl=2):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary,
    scaled by the copy-kernel calibration recorded with it (null when absent)."""
    name = "substep2_pmc_128x128x16.json" if spl == 2 else "substep_pmc_128x128x16.json"
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-baseline-only", nargs="+", type=int, default=None,
                    help="STEPS WARMUP [REPS]: the CPU leg alone (child process)")
    ap.add_argument("--no-admbase", action="store_true", help="with --cpu-baseline-only: skip the consumer path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true")
    ap.add_argument("--timed-only", action="store_true",
                    help="only the warmup and the timed steps (the region the profiles cover)")
    ap.add_argument("--one-substep", action="store_true",
                    help="with --timed-only: time the one-substep-per-launch path instead of the production "
                         "two-substep path (the region the one-substep kernel profile covers)")
    ap.add_argument("--admbase-only", action="store_true",
                    help="only the solve and the Einstein Toolkit consumer path (ADMBase on 128^3 points)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the 64-configuration seed-42 batched solve (the region the batched kernel "
                         "profile covers)")
    ap.add_argument("--dropin-only", action="store_true",
                    help="only the reference-surface entry points, 200 calls each (the region the drop-in "
                         "kernel profile covers)")
    args = ap.parse_args()
    if args.pmc_probe:
        pmc_probe()
        return

    if args.batched_only:
        print(json.dumps({"batched_64_seed42": batched_solve()}), flush=True)
        return
    if args.cpu_baseline_only:
        a = args.cpu_baseline_only
        print(json.dumps(cpu_baseline(a[0], a[1], a[2] if len(a) > 2 else 5, admbase=not args.no_admbase)))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the solve is one GPU's full, unsharded "
              "problem); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from nrpyelliptic_amd import config, runtime

    cfg = config.equal_mass_boosted()
    P = runtime.Problem(cfg)
    npts = cfg.Nxx[0] * cfg.Nxx[1] * cfg.Nxx[2]
    S = P.solver(steps_per_graph=100, substeps_per_launch=1 if args.timed_only and args.one_substep else 0)

    if args.dropin_only:
        P.set("y_n_gfs", np.zeros(2 * P.ntot))
        S.timed_steps(1000)  # a relaxed, nonzero state
        print(json.dumps({"dropin_call_us": dropin_call_us(P, n=200)}), flush=True)
        return
    if args.admbase_only:
        # 3000 graph-replayed steps relax u to ~1e-10 (the solve stops at 3144); the kernel's
        # cost does not depend on the values
        P.set("y_n_gfs", np.zeros(2 * P.ntot))
        S.timed_steps(3000)
        print(json.dumps(admbase_throughput(P)), flush=True)
        return
    # warmup, then exactly --steps production steps from u = v = 0 (HIP events, solver stream)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    S.timed_steps(args.warmup)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    runtime.lib().nrpyell_device_synchronize()
    ms, _ = S.timed_steps(args.steps)
    runtime.lib().nrpyell_device_synchronize()
    sec = ms / 1e3
    value = 4 * npts * args.steps / sec
    # the dominant kernel does `spl` substeps per launch: 4 / spl launches per RK4 step
    spl = S.substeps_per_launch
    us_per_launch = 1e3 * ms / ((4 // spl) * args.steps)
    bytes_per_launch = spl * BYTES_PER_GP_SUBSTEP * npts
    achieved = bytes_per_launch / (us_per_launch * 1e-6) / 1e9

    if args.timed_only:
        print(json.dumps({"metric": METRIC, "value": value, "unit": "gp-substeps/s", "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": ms / args.steps, "us_per_launch": us_per_launch,
                          "achieved_GBs": achieved}), flush=True)
        return
    # time to residual 1e-10 from the zero guess (the reference's every-step criterion)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    r = S.solve()
    nn = nn_seeded_paths(P)
    # the Einstein Toolkit thorn's solver variant on its default grid (N0, N1, N2 = 128, 128, 16):
    # the same relaxation with the thorn's 6th-order Sommerfeld outer BC and wave speed
    S_et = P.solver(steps_per_graph=100, outer_bc="sommerfeld6")
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    r_et = S_et.solve()
    S_et.close()
    et_thorn = {"outer_bc": "sommerfeld6", "wave_speed": runtime.ET_SOMMERFELD_SPEED, "steps": r_et.steps,
                "seconds": r_et.seconds, "log10_l2": r_et.log10_l2, "converged": bool(r_et.converged)}
    # per-launch kernel durations (back-to-back launches, HIP events on the solver stream): the
    # two-substep kernels of the production path, and the one-substep kernels with the RK4 step
    # they make (the path of the Sommerfeld BC and of grids the two-substep kernel does not take)
    stage_us = {"stages12": round(S.time_substep(5, 200), 3), "stages34": round(S.time_substep(6, 200), 3)}
    S1 = P.solver(steps_per_graph=100, substeps_per_launch=1)
    stage_us.update({f"one_substep_stage{k}": round(S1.time_substep(k if k > 1 else 0, 200), 3) for k in (1, 2, 3, 4)})
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    stage_us["one_substep_per_launch_us_per_step"] = round(1e3 * S1.timed_steps(2000)[0] / 2000, 3)
    S1.close()
    dropin_us = dropin_call_us(P)
    admbase = admbase_throughput(P)
    batched = batched_solve()
    mixed_q4 = mixed_precision_q4()
    larger = larger_grids()

    cpu = None if args.no_cpu_baseline else run_cpu_baseline_subprocess()
    if cpu is not None and isinstance(cpu.get("admbase"), dict):
        admbase["cpu_oracle"] = cpu.pop("admbase")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "gp-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_GP_SUBSTEPS,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "equal_mass_boosted_128x128x16 SinhSymTP FP64 RK4 relaxation, zero initial guess, "
                               "residual+L2 every step", "Nxx": list(cfg.Nxx), "global_batch": 1,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(spl),
                     "kernel": ("substep2_kernel (two fused RK4 substeps per launch: RHS + radiation BC + stage "
                                "updates + ghost images, residual of y_n in the first)" if spl == 2 else
                                "substep_kernel (fused RHS + radiation BC + RK4 stage update + ghost images)"),
                     "substeps_per_launch": spl, "us_per_launch": us_per_launch,
                     "bytes_per_launch": bytes_per_launch, "stage_us": stage_us},
        "dropin_call_us": dropin_us,
        "dropin_kernel_us": load_dropin_kernels(),
        "et_consumer_admbase": admbase,
        "batched_64_seed42": batched,
        "mixed_precision_q4_seed7": mixed_q4,
        "larger_grids": larger,
        "cpu_baseline": cpu,
        "time_to_residual": {"target_log10": -10.0, "steps": r.steps, "seconds": r.seconds,
                             "log10_l2": r.log10_l2, "converged": bool(r.converged),
                             "reference_steps": 3144, "check_interval_steps": 1,
                             "learned_initial_guess": nn, "et_thorn_sommerfeld_bc": et_thorn},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Learned initial guess, end to end: generate training solutions with the batched FP64
solver, train the MLP on the GPU, and evaluate NN-seeded solves against zero-guess solves
with the fallback rule (zero guess whenever the NN guess has the larger initial residual).

    python -m nrpyelliptic_amd.learned --train-configs 4096 --eval-configs 16 --batched-eval 256 --out weights.npz
"""
from __future__ import annotations

import argparse
import json
import time
from typing import List

import numpy as np

from . import config, runtime
from .mlp import K0, Features, InitialGuessMLP
from .runtime import DeviceArray


def generate_dataset(configs: List[config.PunctureConfig], rows_per_config=8192, batch=128, target=-10.0, seed=0,
                     max_steps=30000, tuner=None, log=print):
    """Solve every configuration (batched, FP64, to `target`; members that diverge at the
    reference's CFL factor are solved again at runtime.FALLBACK_CFL, which leaves the converged
    u unchanged) and keep rows_per_config random (features, t) rows of each converged solution in
    a device dataset. A configuration that still does not converge is left out. Returns (X, T,
    rows, stats)."""
    N = len(configs) * rows_per_config
    X = DeviceArray(N * K0, np.uint16)
    T = DeviceArray(N, np.float32)
    row = 0
    stats = {"converged": 0, "skipped": [], "cfl_fallback": 0, "steps": 0, "points": 0.0, "seconds": 0.0}
    t0 = time.time()
    for b0 in range(0, len(configs), batch):
        chunk = configs[b0:b0 + batch]
        probs = [runtime.Problem(c) for c in chunk]
        tiles = tuner.choose_all(probs) if tuner is not None else None
        res, sec, info = runtime.relax_batch(probs, tiles=tiles, log10_target=target, max_steps=max_steps)
        stats["cfl_fallback"] += sum(i["fallback"] for i in info)
        for k, (P, r) in enumerate(zip(probs, res)):
            if r.converged:
                F = Features(P)
                t = F.targets(P)
                F.sample_into(t, rows_per_config, seed * 1000003 + b0 + k, X, T, row)
                row += rows_per_config
                stats["converged"] += 1
                stats["steps"] += r.steps
                stats["points"] += 4.0 * float(np.prod(P.cfg.Nxx)) * r.steps
            else:
                stats["skipped"].append(P.cfg.name)
            P.close()
        stats["seconds"] += sec
        conv = sum(r.converged for r in res)
        log(f"dataset: solved {b0 + len(chunk)}/{len(configs)} ({conv}/{len(chunk)} converged, "
            f"{sum(i['fallback'] for i in info)} at the fallback CFL, {sec:.2f}s batch, mean steps "
            f"{np.mean([r.steps for r in res]):.0f}, {time.time() - t0:.1f}s total)")
    return X, T, row, stats


def evaluate_batched(net: InitialGuessMLP, configs, f_nn: float, f_zero: float = 1.0, target=-10.0,
                     max_steps=30000, smooth_passes=2, log=print):
    """The held-out comparison on the batched FP64 solver: every configuration relaxed together
    from the zero guess at the configured eta, then from the network's guess at the damping tuned
    to that start (seed_with_tuned_damping, the zero guess and its damping whenever the guess has
    the larger initial residual); both batches apply the CFL fallback (runtime.relax_batch).
    Wall time of each batch includes its set-up of the starts (seeding, residual checks)."""
    probs = [runtime.Problem(c) for c in configs]
    eta0 = [P.gd.params.eta_damping for P in probs]
    try:
        t0 = time.perf_counter()
        for P in probs:
            P.zero_guess()
        rz, sz, iz = runtime.relax_batch(probs, log10_target=target, max_steps=max_steps)
        tz = time.perf_counter() - t0
        u_zero = [P.get("y_n_gfs")[: P.ntot] for P in probs]
        for P, e in zip(probs, eta0):  # the zero batch may have fallen back: start from the configuration again
            P.gd.params.eta_damping = e
            if P.cfl != P.cfg.CFL:
                P.set_wave_speed_cfl(P.cfg.CFL)
        t0 = time.perf_counter()
        used = [seed_with_tuned_damping(net, P, f_nn, f_zero, smooth_passes)[0] for P in probs]
        seed_s = time.perf_counter() - t0
        rn, sn, inn = runtime.relax_batch(probs, log10_target=target, max_steps=max_steps)
        tn = time.perf_counter() - t0
        du = [float(np.max(np.abs(P.get("y_n_gfs")[: P.ntot] - u))) for P, u in zip(probs, u_zero)]
    finally:
        for P, e in zip(probs, eta0):
            P.gd.params.eta_damping = e
            P.close()
    both = [k for k in range(len(configs)) if rz[k].converged and rn[k].converged]
    out = {"configurations": len(configs), "converged_zero": sum(bool(r.converged) for r in rz),
           "converged_nn": sum(bool(r.converged) for r in rn), "nn_guess_used": int(sum(used)),
           "fallback_to_zero_guess": len(configs) - int(sum(used)),
           "cfl_fallback_zero": sum(i["fallback"] for i in iz), "cfl_fallback_nn": sum(i["fallback"] for i in inn),
           "eta_factor_nn": f_nn, "eta_factor_zero": f_zero,
           "zero_guess_seconds": tz, "nn_seconds": tn, "nn_seed_seconds": seed_s,
           "solver_seconds_zero": sz, "solver_seconds_nn": sn, "speedup_time": tz / max(tn, 1e-12),
           "mean_steps_zero": float(np.mean([rz[k].steps for k in both])) if both else None,
           "mean_steps_nn": float(np.mean([rn[k].steps for k in both])) if both else None,
           "speedup_steps": (float(np.sum([rz[k].steps for k in both]) / np.sum([rn[k].steps for k in both]))
                             if both else None),
           "max_du_nn_vs_zero": max(du[k] for k in both) if both else None}
    log(json.dumps(out))
    return out


def seed_with_fallback(net: InitialGuessMLP, P: runtime.Problem, smooth_passes: int = 2):
    """Zero-guess residual vs NN-guess residual; keep the NN guess only if its residual is lower."""
    P.zero_guess()
    r0 = P.log10_residual()
    net.seed(P, smooth_passes)
    r1 = P.log10_residual()
    used = r1 < r0
    if not used:
        P.zero_guess()
    return used, r0, r1


def seed_with_tuned_damping(net: InitialGuessMLP, P: runtime.Problem, f_nn: float, f_zero: float = 1.0,
                            smooth_passes: int = 2):
    """The start of the fast path: eta = f_nn x the configured eta (a seed writes v = eta u, so
    eta is set first), then seed_with_fallback; when the zero guess is kept, eta = f_zero x the
    configured eta (factors from autotune.EtaTuner for the "nn" and "zero" starts). Returns
    (used, configured eta); the caller restores P.gd.params.eta_damping after its solve."""
    eta0 = P.gd.params.eta_damping
    P.gd.params.eta_damping = f_nn * eta0
    used, _, _ = seed_with_fallback(net, P, smooth_passes)
    if not used:  # u = v = 0: any eta is consistent with the state
        P.gd.params.eta_damping = f_zero * eta0
    return used, eta0


def solve_damping_schedule(net: InitialGuessMLP, P: runtime.Problem, f_hi: float, switch_log10: float = -9.0,
                           target: float = -12.0, precision: int = 32, smooth_passes: int = 2,
                           max_steps: int = 40000, f_lo: float = 1.0):
    """The network's start relaxed in two phases: at f_hi x the configured eta (a strong damping:
    fast on the grid-scale error the guess leaves) down to switch_log10, then at f_lo x the
    configured eta (a weaker one, which relaxes the long-wavelength error a strong damping leaves
    behind, the error the masked residual hardly sees) down to target, with u and
    u_t = v - eta u carried across the switch (Problem.set_eta_damping). autotune.ScheduleTuner
    picks (f_hi, switch_log10, f_lo) for accuracy. Returns (steps, fp32 steps, solver seconds,
    seed seconds, converged, network guess used); P's eta is the configured one afterwards."""
    eta0 = P.gd.params.eta_damping
    t0 = time.perf_counter()
    used, _ = seed_with_tuned_damping(net, P, f_hi, 1.0, smooth_passes)
    seed_s = time.perf_counter() - t0
    try:
        S = P.solver(precision=precision, log10_target=switch_log10, max_steps=max_steps)
        r1 = S.solve()
        S.close()
        P.set_eta_damping(f_lo * eta0)
        S = P.solver(log10_target=target, max_steps=max_steps)
        r2 = S.solve()
        S.close()
    finally:
        P.gd.params.eta_damping = eta0
    return (r1.steps + r2.steps, r1.fp32_steps, r1.seconds + r2.seconds, seed_s,
            bool(r1.converged and r2.converged), bool(used))


def _solve(S):
    try:
        return S.solve()
    except runtime.NrpyellError:  # NaN: reported as not converged
        return None


def evaluate(net: InitialGuessMLP, configs, precision=64, target=-10.0, max_steps=10000, smooth_passes=2,
             log=print, eta_factors=None):
    """Per configuration: steps and seconds to `target` from the zero guess (FP64, configured
    eta) and from the NN guess with the fallback (FP64 or mixed precision), single-configuration
    solver, and the max |du| between the two converged states. `eta_factors` = (f_nn, f_zero)
    from autotune.EtaTuner runs the NN-seeded solve with the damping tuned to its start
    (seed_with_tuned_damping); None keeps the configured eta."""
    rows = []
    for k, cfg in enumerate(configs):
        P = runtime.Problem(cfg)
        S = P.solver(log10_target=target, max_steps=max_steps)
        P.set("y_n_gfs", np.zeros(2 * P.ntot))
        rz = _solve(S)
        u_zero = P.get("y_n_gfs")[: P.ntot]
        eta0 = P.gd.params.eta_damping
        if eta_factors is None:
            used, r0, r1 = seed_with_fallback(net, P, smooth_passes)
        else:
            P.set("y_n_gfs", np.zeros(2 * P.ntot))
            r0 = P.log10_residual()
            used, _ = seed_with_tuned_damping(net, P, eta_factors[0], eta_factors[1], smooth_passes)
            r1 = P.log10_residual()
        S2 = P.solver(log10_target=target, precision=precision, max_steps=max_steps)
        rn = _solve(S2)
        P.gd.params.eta_damping = eta0
        ok = rz is not None and rn is not None and bool(rz.converged and rn.converged)
        du = float(np.max(np.abs(P.get("y_n_gfs")[: P.ntot] - u_zero))) if ok else None
        rows.append({"config": cfg.name, "zero_steps": rz.steps if rz else None, "zero_s": rz.seconds if rz else None,
                     "nn_steps": rn.steps if rn else None, "nn_fp32_steps": rn.fp32_steps if rn else None,
                     "nn_s": rn.seconds if rn else None, "nn_used": bool(used), "log10_res0_zero": r0,
                     "log10_res0_nn": r1, "max_du": du, "converged": ok})
        S.close()
        S2.close()
        P.close()
        if k < 8 or k % 16 == 0:
            log(json.dumps(rows[-1]))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-configs", type=int, default=1024)
    ap.add_argument("--rows-per-config", type=int, default=4096)
    ap.add_argument("--smooth", type=int, default=2)
    ap.add_argument("--eval-configs", type=int, default=64)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-segments", type=int, default=1,
                    help="split --steps into this many segments, halving the learning rate after each "
                         "(Adam moments carry over)")
    ap.add_argument("--target", type=float, default=-10.0)
    ap.add_argument("--max-steps", type=int, default=10000)
    ap.add_argument("--out", default="mlp_initial_guess.npz")
    ap.add_argument("--report", default=None)
    ap.add_argument("--batched-eval", type=int, default=0,
                    help="also compare zero and network starts on this many held-out configurations with the "
                         "batched solver (evaluate_batched)")
    args = ap.parse_args()
    log = lambda s: print(s, flush=True)  # noqa: E731
    train_cfgs = config.random_batch(args.train_configs, seed=1000)
    from .autotune import TileTuner
    tuner = TileTuner(cache_path=None)
    X, T, N, stats = generate_dataset(train_cfgs, rows_per_config=args.rows_per_config, target=args.target,
                                      max_steps=args.max_steps, tuner=tuner, log=log)
    log(f"dataset: {stats['converged']} converged solutions ({stats['cfl_fallback']} at the fallback CFL), {N} "
        f"rows; batched solver {stats['points'] / max(stats['seconds'], 1e-9):.3e} gp-substeps/s")
    net = InitialGuessMLP(seed=7)
    t0 = time.time()
    per = max(1, args.steps // max(1, args.lr_segments))
    parts = []
    for k in range(max(1, args.lr_segments)):
        parts.append(net.train(X, T, N, batch=args.batch, steps=per, lr=args.lr * 0.5 ** k, seed=11 + k))
        log(f"train: segment {k + 1}, lr {args.lr * 0.5 ** k:.2e}, loss {parts[-1][-50:].mean():.3e}")
    losses = np.concatenate(parts)
    train_s = time.time() - t0
    # training FLOPs: forward 2 M (64*256 + 4*256*256 + 256) per batch of M rows, backward twice that
    flops = 3.0 * 2.0 * args.batch * (64 * 256 + 4 * 256 * 256 + 256) * len(losses)
    log(f"train: {len(losses)} steps in {train_s:.1f}s ({flops / train_s / 1e12:.1f} TFLOP/s = "
        f"{100 * flops / train_s / 2.5e15:.2f} % of the 2.5 PFLOP/s dense BF16 peak), loss "
        f"{losses[:10].mean():.3e} -> {losses[-50:].mean():.3e}")
    net.save(args.out)
    eval_cfgs = config.random_batch(args.eval_configs, seed=900, grids=False, Nxx=(128, 128, 16))
    rows64 = evaluate(net, eval_cfgs, precision=64, target=args.target, smooth_passes=args.smooth, log=log)
    rows32 = evaluate(net, eval_cfgs, precision=32, target=args.target, smooth_passes=args.smooth, log=log)
    # the damping tuned to each start on a calibration configuration (seed 901)
    from .autotune import EtaTuner
    calib = runtime.Problem(config.random_batch(1, seed=901, grids=False, Nxx=(128, 128, 16))[0])
    # the network's start runs FP32 -> FP64 (rows32t), the zero guess FP64: each factor is
    # calibrated in the precision it will be used with
    factors = (EtaTuner(cache_path=None, target=args.target, precision=32).choose(
                   calib, lambda Q: seed_with_fallback(net, Q, args.smooth), "nn"),
               EtaTuner(cache_path=None, target=args.target, precision=64).choose(
                   calib, lambda Q: Q.set("y_n_gfs", np.zeros(2 * Q.ntot)), "zero"))
    calib.close()
    log(f"eta factors (nn, zero): {factors}")
    rows32t = evaluate(net, eval_cfgs, precision=32, target=args.target, smooth_passes=args.smooth, log=log,
                       eta_factors=factors)
    batched = None
    if args.batched_eval:
        batched = evaluate_batched(net, config.random_batch(args.batched_eval, seed=900, grids=False,
                                                            Nxx=(128, 128, 16)),
                                   factors[0], factors[1], target=args.target, smooth_passes=args.smooth, log=log)

    def summary(rows):
        n_all = len(rows)
        rows = [r for r in rows if r["converged"]]
        if not rows:
            return {"n": 0, "n_all": n_all}
        z = np.array([r["zero_s"] for r in rows])
        n = np.array([r["nn_s"] for r in rows])
        zs = np.array([r["zero_steps"] for r in rows])
        ns = np.array([r["nn_steps"] for r in rows])
        return {"mean_zero_steps": float(zs.mean()), "mean_nn_steps": float(ns.mean()),
                "speedup_time_mean": float(z.sum() / n.sum()), "speedup_steps_mean": float(zs.sum() / ns.sum()),
                "nn_used_frac": float(np.mean([r["nn_used"] for r in rows])),
                "max_du": float(max(r["max_du"] for r in rows)), "n": len(rows), "n_all": n_all}

    rep = {"train_configs": args.train_configs, "converged_train": stats["converged"],
           "cfl_fallback_train": stats["cfl_fallback"], "rows": N, "train_steps": int(len(losses)),
           "train_seconds": train_s, "train_tflops": flops / train_s / 1e12,
           "final_loss": float(losses[-50:].mean()), "batched_held_out": batched,
           "fp64": summary(rows64), "mixed": summary(rows32), "eta_factors": list(factors),
           "mixed_tuned_eta": summary(rows32t)}
    log(json.dumps(rep))
    if args.report:
        with open(args.report, "w") as f:
            json.dump({"summary": rep, "fp64": rows64, "mixed": rows32, "mixed_tuned_eta": rows32t,
                       "losses": losses[::10].tolist()}, f)


if __name__ == "__main__":
    main()

"""Benchmark: converged state points per second, batched mixed precision vs plain FP64.

Workload (one "step"): the 256-point pH sweep (pH 2..12) of 20-mer weak-acid
chains (pKa 4.5) plus hard-sphere solvent and 10 mM 1:1 salt in a hard-wall
slit H = 16 sigma, dz = sigma/50 (NP = 801), White Bear FMT + TPT1 + TPT1-MSA
+ Poisson, solved from the bulk guess to ||F(x)-x||_inf < 1e-8 with Anderson(5)
in one batched mixed-precision solve (FP32 until 1e-5, then FP64 polish).
Inputs are resident on the device before the timed region.

Reported beside it: plain one-at-a-time FP64 solves of the same points
(sample), the max relative Omega error of the mixed batch against a batched
FP64 solve, the roofline of the dominant kernel (HIP events on the library's
stream) and the CPU baseline (the FP64 NumPy restatement of the reference
algorithm on one host core, one state point).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged state points/sec at ≤1e-8 rel ΔΩ: batched mixed-prec vs plain FP64"
PEAK = {32: 157.3, 64: 78.6}  # MI355X published dense TFLOP/s (vector == matrix for FP32/FP64)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_probe(mib):
    """A plain device-to-device copy of a known size: rocprofv3's FETCH_SIZE / WRITE_SIZE for it
    give the counter-to-bytes factor applied to the kernels' readings."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: 5 copies of {mib} MiB ({mib * 1024} KiB read and written per copy)")


def traffic_from_profiles(kernel_name):
    """HBM bytes per launch of `kernel_name` measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and
    scaled by the copy-probe factor, as recorded in profiles/*.md lines of the form
    `traffic_bytes_per_launch: <kernel name> = <bytes>`; None when not measured."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    name, _, val = line[len("traffic_bytes_per_launch:"):].rpartition("=")
                    if name.strip() == kernel_name:
                        return float(val)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=256)
    ap.add_argument("--singles", type=int, default=6, help="one-at-a-time FP64 sample size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--kernel-only", choices=["chain32", "chain64", "fwd32", "fwd64", "back32", "back64"],
                    help="after one full-load evaluation, only re-launch this kernel class 21 times "
                         "(for rocprofv3 --kernel-trace / --pmc runs of that kernel)")
    ap.add_argument("--solve-only", action="store_true",
                    help="stop after the timed mixed solves (for rocprofv3 --kernel-trace of the solve itself)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; state points shard across GPUs with no "
                                   "communication, but only the single-GPU figure is measured"}))
        sys.exit(2)

    import numpy as np
    from softdft_amd import abi, configs
    from softdft_amd.problem import build_state_point

    t = time.time()
    sps = [build_state_point(d) for d in configs.config3(args.points)]
    log(f"setup: {len(sps)} state points in {time.time() - t:.1f} s")
    batch = abi.Batch(sps)
    x0 = batch.x0()
    num = dict(tole=1e-8, damping=0.1, mixing_max=0.02, switch_tole=100.0, anderm=5, max_iters=3000)

    if args.kernel_only:
        # one evaluation of every point in the class's precision with the banded operators, so that
        # every launch of the kernel in this run has all 256 points active (the full-load launch the
        # roofline figure times)
        prec = abi.PREC_FP32 if args.kernel_only.endswith("32") else abi.PREC_FP64
        batch.set_operator_band(1, 1)
        batch.objective(x0, prec)
        kc = {"chain32": abi.KP_CHAIN_FP32, "chain64": abi.KP_CHAIN_FP64, "fwd32": abi.KP_CONV_FWD_FP32,
              "fwd64": abi.KP_CONV_FWD_FP64, "back32": abi.KP_CONV_BACK_FP32, "back64": abi.KP_CONV_BACK_FP64}
        ms, fl = batch.kernel_profile(kc[args.kernel_only], reps=20)
        log(f"{args.kernel_only}: {ms * 1e3:.1f} us/launch, {fl:.4g} algorithmic flops/launch")
        batch.close()
        return
    for w in range(args.warmup):
        batch.solve(x0, precision=abi.PREC_MIXED, **num)
        log(f"warmup {w} done")
    times, iters = [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        x, it, err = batch.solve(x0, precision=abi.PREC_MIXED, **num)  # synchronous: returns after the device finished
        times.append(time.perf_counter() - t0)
        iters.append(it)
        log(f"step {k}: {times[-1] * 1e3:.1f} ms, iterations mean {it.mean():.1f} max {it.max()}")
    om_mixed, al_mixed, _ = batch.observables()
    total = sum(times)
    value = args.points * args.steps / total
    it32, it64 = batch.pipeline_iters()
    log(f"batch iterations with the FP32 / FP64 pipeline running: {it32} / {it64}")
    if args.solve_only:
        log(f"solve-only: {value:.1f} state points/s over {args.steps} timed solves")
        batch.close()
        return

    # per-launch time of the hot kernel classes, HIP events on the library stream, with the
    # operators the mixed solve used; weighted by launches per evaluation x iterations run
    # (one chain launch runs every propagator step of the chain species)
    prof = {}
    for name, kc, prec, per_eval, iters_run in [
            ("k_chain<float> (banded propagators, all steps)", abi.KP_CHAIN_FP32, 32, 1, it32),
            ("k_chain<double> (banded propagators, all steps)", abi.KP_CHAIN_FP64, 64, 1, it64),
            ("k_convgemm<float> (forward weighted densities)", abi.KP_CONV_FWD_FP32, 32, 1, it32),
            ("k_convgemm<double> (forward weighted densities)", abi.KP_CONV_FWD_FP64, 64, 1, it64),
            ("k_convgemm<float> (back projections)", abi.KP_CONV_BACK_FP32, 32, 1, it32),
            ("k_convgemm<double> (back projections)", abi.KP_CONV_BACK_FP64, 64, 1, it64)]:
        ms, fl = batch.kernel_profile(kc, reps=20)
        prof[name] = (ms, fl, prec, ms * per_eval * iters_run)
        log(f"kernel {name}: {ms * 1e3:.1f} us/launch, {fl / ms / 1e9:.2f} TFLOP/s, "
            f"{per_eval} launches/evaluation x {iters_run} iterations = {ms * per_eval * iters_run:.1f} ms per solve")

    # FP64 reference for the acceptance bound (untimed)
    xd, itd, errd = batch.solve(x0, precision=abi.PREC_FP64, **num)
    om_fp64, al_fp64, _ = batch.observables()
    rel = np.abs(om_mixed - om_fp64) / np.abs(om_fp64)
    log(f"max rel dOmega mixed vs FP64: {rel.max():.3e}")
    batch.close()

    # plain FP64, one state point at a time
    sel = np.linspace(0, args.points - 1, args.singles).astype(int)
    t0 = time.perf_counter()
    for i in sel:
        b1 = abi.Batch([sps[i]])
        b1.solve(b1.x0(), precision=abi.PREC_FP64, **num)
        b1.close()
    single_rate = len(sel) / (time.perf_counter() - t0)
    log(f"FP64 one-at-a-time: {single_rate:.2f} points/s")

    cpu = None
    if not args.no_cpu:
        from oracle.dft_oracle import OracleDFT, anderson_solve
        sp = sps[int(np.argmin(np.abs(np.linspace(2, 12, args.points) - 7.0)))]
        o = OracleDFT(sp)
        t0 = time.perf_counter()
        r = anderson_solve(o.objective, sp.x0(), sp.numerics)
        dt = time.perf_counter() - t0
        cpu = {"value": 1.0 / dt, "unit": "state points/s", "cores": 1, "kind": "port",
               "sample": f"1 of the {args.points} state points (pH 7.0), FP64 NumPy restatement of the reference "
                         f"algorithm, {r.iters} Anderson iterations in {dt:.1f} s on one host core"}
        log(f"cpu baseline: {dt:.1f} s for one point")

    # roofline of the dominant kernel: the largest estimated time per mixed solve
    name, (ms, fl, prec, _) = max(prof.items(), key=lambda kv: kv[1][3])
    achieved = fl / (ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "state points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"pH sweep 2..12, {args.points} state points, 20-mer weak acid + solvent + 10 mM salt, "
                               "H=16 sigma, dz=sigma/50 (NP=801), TPT1mfmt+TPT1msa+Poisson, Anderson(5) to 1e-8",
                   "points": args.points, "NP": 801, "precision": "FP32 -> FP64 polish at 1e-5",
                   "parallelism": "one GPU, batched state points"},
        "fp64_one_at_a_time": {"value": single_rate, "unit": "state points/s", "sample_points": int(len(sel))},
        "speedup_vs_fp64_one_at_a_time": value / single_rate,
        "max_rel_dOmega_vs_fp64": float(rel.max()),
        "mean_iterations": float(np.mean([i.mean() for i in iters])),
        "roofline": {"bound": "mfma", "kernel": name, "achieved": achieved, "peak": PEAK[prec],
                     "unit": "TFLOP/s", "frac": achieved / PEAK[prec], "traffic": traffic_from_profiles(name)},
        "cpu_baseline": cpu,
        "kernels_us": {k: v[0] * 1e3 for k, v in prof.items()},
        "kernels_ms_per_solve_est": {k: v[3] for k, v in prof.items()},
        "pipeline_iterations": {"fp32": it32, "fp64": it64},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

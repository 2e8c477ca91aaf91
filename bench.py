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


WARM_LAUNCHES = 100


def profile_warm(batch, kc, reps=20):
    """sdft_kernel_profile after WARM_LAUNCHES untimed launches of the same kernel class: the GPU's
    clock settles under sustained load (as in a solve, where the kernels run back to back), so the
    timed launches measure the kernel at that clock rather than the first launches after idle time."""
    batch.kernel_profile(kc, reps=WARM_LAUNCHES)
    return batch.kernel_profile(kc, reps=reps)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


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


# ---------------------------------------------------------------------------------------------
# per-config reports (`--config c1|c2|c4|c5`): one JSON line each, beside (not instead of) the
# benchmark line; the errors are against the FP64 paths and the oracle, the CPU figure is the NumPy
# oracle on one host core (a bounded sample where a full CPU solve would take minutes)
# ---------------------------------------------------------------------------------------------
def _finite(v):
    """JSON-safe copy: NaN / Inf (quantities a config does not have) become null."""
    import math
    if isinstance(v, float):
        return v if math.isfinite(v) else None
    if isinstance(v, dict):
        return {k: _finite(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_finite(x) for x in v]
    return v


def _solve_opts(num, **kw):
    d = dict(tole=num["tole"], damping=num["damping"], mixing_max=num["mixing_max"], switch_tole=num["switch_tole"],
             anderm=int(num["anderm"]), max_iters=5000)
    d.update(kw)
    return d


def _adsorption(sp, rho):
    """Excess adsorption per bead type: sum_K (rho_j(K) - rho_j^bulk) trapez_j(K) dz."""
    import numpy as np
    return np.sum((rho - sp.bulk.rho_beads[None, :]) * sp.trapez[:sp.NP], axis=0) * sp.dz


def _timed_solve(batch, x0, prec, opts):
    t0 = time.perf_counter()
    x, it, err = batch.solve(x0, precision=prec, raise_on_noconv=False, **opts)
    return x, it, err, time.perf_counter() - t0


def _oracle_eval_time(sp, x, reps=2):
    from oracle.dft_oracle import OracleDFT
    o = OracleDFT(sp)
    o.objective(x)
    t0 = time.perf_counter()
    for _ in range(reps):
        o.objective(x)
    return (time.perf_counter() - t0) / reps, o


def _oracle_solve(sp):
    """One full FP64 oracle solve (CPU baseline worker): iterations, Omega, bead densities and the
    chain species' state fractions of the converged point (compared with the GPU afterwards, off the
    clock)."""
    from oracle.dft_oracle import OracleDFT, anderson_solve
    o = OracleDFT(sp)
    r = anderson_solve(o.objective, sp.x0(), sp.numerics)
    return {"iters": r.iters, "omega": o.grand_potential(r.x), "rho": o.rho_beads.copy(),
            "alpha": o.state_fractions()}


def b_iter(sp, w, m=5):
    """Algorithmic HBM bytes of one SCF iteration of one state point: x in, F(x) out and the Anderson
    history (FP64 always), one propagator sweep stored and re-read, segment densities written and
    reduced, bead and bond densities written and read, Ext and trapez read; w = 4 (FP32 phase) or 8
    (FP64). The propagator count follows child_propagators.jl (S^2 (N - 1) per chain species)."""
    NP, NB, MB = sp.NP, sp.NB, sp.MB
    chains = sum(max(len(f) for f in s.state_family) ** 2 * (len(s.sequence) - 1) for s in sp.ms.species)
    segs = sum(max(len(f) for f in s.state_family) * len(s.sequence) for s in sp.ms.species)
    return (8 * sp.n * (3 * m + 5) + 2 * w * NP * chains + 2 * w * NP * segs + 2 * w * NP * (NB + 2 * MB)
            + w * NP * 2 * NB)


def _cpu_warm(k):
    """Starts a worker: imports the oracle and holds the worker briefly so that every worker spawns."""
    import oracle.dft_oracle  # noqa: F401
    time.sleep(0.5)
    return k


def cpu_points_parallel(points, cores):
    """Solve independent state points with the oracle, one spawned process per point (single-threaded
    NumPy each); returns (per-point results of _oracle_solve, wall seconds)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    old = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in old:
        os.environ[k] = "1"
    try:
        with ProcessPoolExecutor(max_workers=cores, mp_context=mp.get_context("spawn")) as ex:
            list(ex.map(_cpu_warm, range(cores)))  # every worker started (and numpy imported) before the clock
            t0 = time.perf_counter()
            res = list(ex.map(_oracle_solve, points))
            dt = time.perf_counter() - t0
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return res, dt


def report_single(name, ds, cpu_full, args):
    """C1 / C2: one state point. Plain FP64 GPU solve, mixed solve, CPU oracle (full solve or a bounded
    sample of objective evaluations scaled by the FP64 iteration count)."""
    import numpy as np
    from softdft_amd import abi
    from softdft_amd.problem import build_state_point
    from oracle.dft_oracle import anderson_solve
    sp = build_state_point(ds)
    b = abi.Batch([sp])
    if not args.no_tune:
        dec = b.autotune(reps=3)
        log(f"{name}: autotune " + "; ".join(f"kind {e['kind']} key {e['key']} prec {e['precision']} -> {e['choice']}"
                                             for e in dec))
    opts = _solve_opts(sp.numerics)
    x64, it64, e64, t64 = _timed_solve(b, b.x0(), abi.PREC_FP64, opts)
    om64, al64, _ = b.observables()
    rho64 = b.get_field(abi.FIELD_RHO_BEADS, 0)
    xm, itm, em, tm = _timed_solve(b, b.x0(), abi.PREC_MIXED, opts)
    omm, alm, _ = b.observables()
    rhom = b.get_field(abi.FIELD_RHO_BEADS, 0)
    b.close()
    log(f"{name}: FP64 {it64[0]} iterations {t64:.2f} s, mixed {itm[0]} iterations {tm:.2f} s")
    teval, o = _oracle_eval_time(sp, x64)
    fo = o.objective(x64)
    om_oracle = o.grand_potential(x64)
    res_oracle = float(np.max(np.abs(fo - x64)))
    rho_or = o.rho_beads
    # degree of ionization (charged-state fraction of the multistate species): GPU FP64, mixed, and the
    # oracle's at the GPU's FP64 solution
    al_or = o.state_fractions()
    ion = {}
    for u, fr in enumerate(al_or):
        if fr is None or len(fr) < 2:
            continue
        ion[u] = (float(al64[0, u, 1]), float(alm[0, u, 1]), float(fr[1]))
    if cpu_full:
        t0 = time.perf_counter()
        r = anderson_solve(o.objective, sp.x0(), sp.numerics)
        tcpu = time.perf_counter() - t0
        cpu = {"value": 1.0 / tcpu, "unit": "state points/s", "cores": 1, "kind": "port",
               "sample": f"full oracle solve, {r.iters} iterations in {tcpu:.1f} s"}
    else:
        tcpu = teval * int(it64[0])
        cpu = {"value": 1.0 / tcpu, "unit": "state points/s", "cores": 1, "kind": "port",
               "sample": f"3 oracle objective evaluations ({teval:.3f} s each) x the {int(it64[0])} iterations of the "
                         f"FP64 solve (estimate)"}
    scale = float(np.max(np.abs(rho64)))
    out = {
        "report": "config", "config": name, "NP": sp.NP,
        "fp64_gpu": {"value": 1.0 / t64, "unit": "state points/s", "iterations": int(it64[0]), "residual": float(e64[0])},
        "mixed_gpu": {"value": 1.0 / tm, "unit": "state points/s", "iterations": int(itm[0]), "residual": float(em[0])},
        "cpu_baseline": cpu,
        "speedup_mixed_vs_fp64_gpu": t64 / tm, "speedup_mixed_vs_cpu": tcpu / tm,
        "errors": {
            "omega_fp64_gpu": float(om64[0]), "omega_mixed_rel_vs_fp64": float(abs(omm[0] - om64[0]) / abs(om64[0])),
            "omega_fp64_rel_vs_oracle_at_same_x": float(abs(om64[0] - om_oracle) / abs(om_oracle)),
            "oracle_residual_at_gpu_fp64_solution": res_oracle,
            "density_max_rel_mixed_vs_fp64": float(np.max(np.abs(rhom - rho64)) / scale),
            "density_max_rel_fp64_vs_oracle": float(np.max(np.abs(rho64 - rho_or)) / scale),
            "adsorption_fp64": _adsorption(sp, rho64).tolist(),
            "adsorption_max_abs_mixed_vs_fp64": float(np.max(np.abs(_adsorption(sp, rhom) - _adsorption(sp, rho64)))),
            "ionization_fp64": [v[0] for v in ion.values()],
            "ionization_max_abs_mixed_vs_fp64": max((abs(v[1] - v[0]) for v in ion.values()), default=float("nan")),
            "ionization_max_abs_fp64_vs_oracle_at_same_x": max((abs(v[0] - v[2]) for v in ion.values()),
                                                              default=float("nan")),
        },
    }
    print(json.dumps(_finite(out)), flush=True)


def report_c4(args):
    """C4: 512 slit widths H = 4..40 (NP 201..2001) with neighbour warm starts, mixed precision, in waves."""
    import numpy as np
    from softdft_amd import abi, configs
    from softdft_amd.problem import build_state_points
    from softdft_amd.sweeps import width_sweep
    dss = configs.config4(args.points if args.points != 256 else 512)
    t0 = time.perf_counter()
    res = width_sweep(dss, waves=8, precision=abi.PREC_MIXED, workers=12, tune=not args.no_tune,
                      on_round=lambda r, e: log(f"c4 round {r}: max residual {e:.2e}"))
    tsw = time.perf_counter() - t0
    n = len(res)
    # plain FP64 one width at a time (sample) and the same widths in one batched cold FP64 solve
    sel = np.linspace(0, n - 1, 4).astype(int)
    sps = build_state_points([dss[i] for i in sel], device_bulk=True)
    t0 = time.perf_counter()
    for sp in sps:
        b1 = abi.Batch([sp])
        b1.solve(b1.x0(), precision=abi.PREC_FP64, raise_on_noconv=False, **_solve_opts(sp.numerics))
        b1.close()
    t1 = (time.perf_counter() - t0) / len(sps)
    b = abi.Batch(sps)
    x, it, err = b.solve(b.x0(), precision=abi.PREC_FP64, raise_on_noconv=False, **_solve_opts(sps[0].numerics))
    om, al, pc = b.observables()
    rho_gpu = [b.get_field(abi.FIELD_RHO_BEADS, k) for k in range(len(sps))]
    b.close()
    dom = [abs(res[i].omega - om[k]) / max(1.0, abs(om[k])) for k, i in enumerate(sel)]
    dpi = [abs(res[i].delta_p - (pc[k, 0] - sps[k].bulk.pressure)) for k, i in enumerate(sel)]
    # the FP64 GPU solves of the two narrowest sampled widths against full oracle solves (off the clock)
    from oracle.dft_oracle import OracleDFT, anderson_solve
    ora = []
    for k in range(2):
        sp = sps[k]
        o = OracleDFT(sp)
        r = anderson_solve(o.objective, sp.x0(), sp.numerics, max_iters=20000)
        om_o = o.grand_potential(r.x)
        rho_o = o.rho_beads.copy()
        ads_g, ads_o = _adsorption(sp, rho_gpu[k]), _adsorption(sp, rho_o)
        fr = o.state_fractions()
        da = max((float(np.max(np.abs(al[k, u, :len(f)] - f))) for u, f in enumerate(fr) if f is not None), default=0.0)
        ora.append({"H": float(sp.H), "NP": sp.NP, "oracle_iterations": int(r.iters),
                    "omega_rel": abs(om[k] - om_o) / max(1.0, abs(om_o)),
                    "rho_beads_rel": float(np.max(np.abs(rho_gpu[k] - rho_o)) / np.max(np.abs(rho_o))),
                    "adsorption_rel": float(np.max(np.abs(ads_g - ads_o)) / max(np.max(np.abs(ads_o)), 1e-300)),
                    "alpha_abs": da,
                    "pi_abs": abs((pc[k, 0] - sp.bulk.pressure) - (o.contact_pressure()[0] - sp.bulk.pressure))})
        log(f"c4 H={sp.H:.3f}: FP64 GPU vs oracle {ora[-1]}")
    teval, _ = _oracle_eval_time(sps[len(sps) // 2], x[b.offsets[len(sps) // 2]:b.offsets[len(sps) // 2 + 1]])
    mean_it = float(np.mean([r.iterations for r in res]))
    out = {"report": "config", "config": "C4", "widths": n, "NP_range": [min(s.NP for s in sps), max(s.NP for s in sps)],
           "mixed_gpu_sweep": {"value": n / tsw, "unit": "state points/s", "seconds": tsw, "mean_iterations": mean_it,
                               "unconverged": int(sum(r.residual >= 1e-8 for r in res))},
           "fp64_gpu_one_at_a_time": {"value": 1.0 / t1, "unit": "state points/s", "sample_points": len(sps)},
           "speedup_vs_fp64_one_at_a_time": (n / tsw) * t1,
           "cpu_baseline": {"value": 1.0 / (teval * mean_it), "unit": "state points/s", "cores": 1, "kind": "port",
                            "sample": f"3 oracle objective evaluations at NP={sps[len(sps) // 2].NP} "
                                      f"({teval:.3f} s each) x the sweep's mean {mean_it:.0f} iterations (estimate)"},
           "errors": {"omega_max_rel_sweep_vs_cold_fp64": float(max(dom)), "pi_max_abs_sweep_vs_cold_fp64": float(max(dpi)),
                      "fp64_gpu_vs_oracle": ora},
           "pi_curve": [[r.distance, r.delta_p] for r in res[::16]]}
    print(json.dumps(_finite(out)), flush=True)


def report_c5(args):
    """C5: 128 brush points at full size (NP 16384, 500-mers). Per-iteration throughput of the batch in
    FP32 and FP64 (banded, the mixed path's polish) and the plain FP64 (dense) evaluation of one point;
    convergence of the brush is reported from the solves themselves."""
    import numpy as np
    from softdft_amd import abi, configs
    from softdft_amd.problem import build_state_points
    dss = configs.config5()
    t0 = time.perf_counter()
    sps = build_state_points(dss, device_bulk=True)
    log(f"c5: {len(sps)} state points set up in {time.perf_counter() - t0:.1f} s")
    b = abi.Batch(sps)
    x0 = b.x0()
    res, conv_us = {}, {}
    b.set_operator_band(1, 1)
    # convolutions as dense operator-GEMMs (fixed rule) and as batched FFTs (mirror length 32768)
    for conv, shape in (("gemm", abi.CONV_MID), ("fft", abi.CONV_FFT)):
        b.set_conv_shape(shape)
        for prec, label in ((abi.PREC_FP32, "fp32"), (abi.PREC_FP64, "fp64_banded")):
            b.objective(x0, prec)
            t0 = time.perf_counter()
            for _ in range(3):
                b.objective(x0, prec)
            res[f"{label}_{conv}"] = (time.perf_counter() - t0) / 3
            log(f"c5 {label} ({conv} convolutions): {res[f'{label}_{conv}'] * 1e3:.1f} ms per batch evaluation")
            for kc, stage in (((abi.KP_CONV_FWD_FP32, abi.KP_CONV_FWD_FP64), "forward"),
                              ((abi.KP_CONV_BACK_FP32, abi.KP_CONV_BACK_FP64), "back")):
                ms, _ = b.kernel_profile(kc[0] if prec == abi.PREC_FP32 else kc[1], reps=5)
                conv_us[f"{label}_{conv}_{stage}"] = ms * 1e3
        log(f"c5 convolution stages (us per launch sequence): {conv_us}")
    res["fp32"], res["fp64_banded"] = res["fp32_fft"], res["fp64_banded_fft"]
    # the whole 128-point sweep solved as one mixed-precision batch from the bulk guess, capped at
    # args.c5_iters iterations: points that reach ||F(x)-x|| < 1e-8 are the converged subset
    b.set_conv_shape(-1)
    x, it_all, err_all, t_batch = _timed_solve(b, x0, abi.PREC_MIXED,
                                               _solve_opts(sps[0].numerics, max_iters=args.c5_iters))
    conv = err_all < 1e-8
    pH_all = np.repeat(np.linspace(3.0, 9.0, 16), 8)
    salt_all = np.tile(np.logspace(np.log10(0.001), np.log10(0.1), 8), 16)
    log(f"c5 batch solve: {int(conv.sum())}/{len(sps)} converged in {t_batch:.1f} s "
        f"(cap {args.c5_iters} iterations)")
    batch_solve = {"max_iters": args.c5_iters, "seconds": t_batch, "converged": int(conv.sum()),
                   "points": len(sps), "converged_points_per_s": float(conv.sum()) / t_batch,
                   "converged_pH_salt": [[float(pH_all[i]), float(salt_all[i]), int(it_all[i])]
                                         for i in range(len(sps)) if conv[i]],
                   "residual_of_the_rest": {"min": float(np.min(err_all[~conv])) if (~conv).any() else None,
                                            "median": float(np.median(err_all[~conv])) if (~conv).any() else None}}
    b.close()
    b1 = abi.Batch([sps[0]])
    b1.objective(b1.x0(), abi.PREC_FP64)
    t0 = time.perf_counter()
    b1.objective(b1.x0(), abi.PREC_FP64)
    t_plain = time.perf_counter() - t0
    b1.close()
    # one converged full-size solve: pH 3 at 1 mM (the first point), mixed precision, FFT convolutions
    bs = abi.Batch([sps[0]])
    bs.set_conv_shape(abi.CONV_FFT)
    x, it, err, ts = _timed_solve(bs, bs.x0(), abi.PREC_MIXED, _solve_opts(sps[0].numerics, max_iters=3000))
    om, al, _ = bs.observables()
    rho_m = bs.get_field(abi.FIELD_RHO_BEADS, 0)
    # the same point solved cold in plain FP64 (banded propagators, the operator the polish uses), and
    # the oracle's map at the mixed solution (banded direct-sum propagators: the FFT evaluation of a
    # 500-mer brush's propagators is round-off dominated); an oracle solve of this point takes hours,
    # so the error against the oracle is the fixed-point residual of its map
    x64, it64, err64, t64 = _timed_solve(bs, bs.x0(), abi.PREC_FP64, _solve_opts(sps[0].numerics, max_iters=4000,
                                                                                 band_fp32=2))
    om64, al64, _ = bs.observables()
    rho64 = bs.get_field(abi.FIELD_RHO_BEADS, 0)
    bs.close()
    from oracle.dft_oracle import OracleDFT
    o5 = OracleDFT(sps[0])
    o5.use_banded_direct_propagators()
    f_or = o5.objective(x)
    c5_err = {"fp64_cold_iterations": int(it64[0]), "fp64_residual": float(err64[0]), "fp64_seconds": t64,
              "omega_rel_mixed_vs_fp64": abs(om[0] - om64[0]) / max(1.0, abs(om64[0])),
              "rho_beads_rel_mixed_vs_fp64": float(np.max(np.abs(rho_m - rho64)) / np.max(np.abs(rho64))),
              "alpha_abs_mixed_vs_fp64": float(np.nanmax(np.abs(al[0] - al64[0]))),
              "oracle_map_residual_at_mixed_solution": float(np.max(np.abs(f_or - x)) / max(1.0, float(np.max(np.abs(x)))))}
    log(f"c5 errors: {c5_err}")
    per_point_fp32 = res["fp32"] / len(sps)
    out = {"report": "config", "config": "C5", "points": len(sps), "NP": sps[0].NP, "n_seg": 500,
           "batch_evaluation_ms": {k: v * 1e3 for k, v in res.items()},
           "convolution_stage_us": conv_us,
           "batch_iterations_per_s": {"fp32": 1.0 / res["fp32"], "fp64_banded": 1.0 / res["fp64_banded"]},
           "per_point_evaluation_ms": {"fp32_in_batch": per_point_fp32 * 1e3,
                                       "fp64_banded_in_batch": res["fp64_banded"] / len(sps) * 1e3,
                                       "plain_fp64_dense_one_point": t_plain * 1e3},
           "per_point_evaluation_speedup_fp32_batch_vs_plain_fp64": t_plain / per_point_fp32,
           "batch_solve": batch_solve,
           "single_point_mixed_solve": {"pH": 3.0, "salt_M": 0.001, "iterations": int(it[0]),
                                        "residual": float(err[0]), "seconds": ts, "omega": float(om[0]),
                                        "alpha": float(al[0, 0, 1]), "errors": c5_err},
           "note": "converged_points_per_s counts the converged subset of the capped batch solve over the "
                   "batch's wall time; the charged brushes at high pH and salt stall at O(0.1-10) with the "
                   "brush numerics, and the reference's algorithm (no propagator rescaling) overflows on them "
                   "from the bulk guess"}
    print(json.dumps(_finite(out)), flush=True)


def c5_fft_kernels(which):
    """One FFT convolution stage of the C5 batch (128 brushes, NP 16384, mirror length 32768): one
    evaluation of every point in the stage's precision, then 21 launch sequences of the stage
    (1 warm + 20 timed with HIP events on the library's stream)."""
    from softdft_amd import abi, configs
    from softdft_amd.problem import build_state_points
    sps = build_state_points(configs.config5(), device_bulk=True)
    b = abi.Batch(sps)
    b.set_operator_band(1, 1)
    b.set_conv_shape(abi.CONV_FFT)
    prec = abi.PREC_FP32 if which.endswith("32") else abi.PREC_FP64
    b.objective(b.x0(), prec)
    kc = {"c5fft_fwd32": abi.KP_CONV_FWD_FP32, "c5fft_fwd64": abi.KP_CONV_FWD_FP64,
          "c5fft_back32": abi.KP_CONV_BACK_FP32, "c5fft_back64": abi.KP_CONV_BACK_FP64}[which]
    ms, _ = b.kernel_profile(kc, reps=20)
    log(f"{which}: {ms * 1e3:.1f} us per launch sequence (3 kernels), {len(sps)} points, NP {sps[0].NP}")
    b.close()


def nd3_datasets():
    """The 3-D pore workload of report_nd: 16 pH values, 33 x 33 x 17 grid."""
    import numpy as np
    from softdft_amd import configs
    out = []
    for pH in np.linspace(3.0, 9.0, 16):
        d = configs.weak_polyelectrolyte(H=4.0, dz=0.125, n_seg=20, rho_poly=0.0954930, solvent=0.2864789, pH=float(pH))
        d["geometry"]["dimensions"] = [4.0, 4.0, 2.0]
        d["geometry"]["bin_width"] = [0.125, 0.125, 0.125]
        d["geometry"]["external_field"] = {"hard_walls": {"dims": [1, 2]}}
        out.append(d)
    return out


def nd3_kernels(which):
    """--kernel-only nd3_64 / nd3_32: 21 objective evaluations of the 3-D pore batch at the bulk guess
    (for rocprofv3 runs of the multi-dimensional kernels)."""
    from softdft_amd import abi
    from softdft_amd.geometry_nd import build_state_points_nd
    b = abi.Batch(build_state_points_nd(nd3_datasets(), device_bulk=True))
    prec = abi.PREC_FP64 if which.endswith("64") else abi.PREC_FP32
    x0 = b.x0()
    t0 = time.perf_counter()
    for _ in range(21):
        b.objective(x0, prec)
    log(f"{which}: 21 batch objective evaluations in {time.perf_counter() - t0:.3f} s")
    b.close()


def report_nd(args):
    """Three-dimensional pore: weak-acid 20-mers with solvent and 10 mM salt (the pH-sweep chemistry)
    in a square pore of 4 x 4 sigma (hard walls along x and y), 2 sigma along the open axis,
    h = sigma/8 (33 x 33 x 17 = 18513 points), 16 pH values 3..9 as one batch: mixed and plain FP64
    batch solves on the spectral multi-dimensional path, errors mixed vs FP64, the oracle's map at
    the FP64 solution of one point, and the CPU oracle's objective time (estimate: x iterations)."""
    import copy
    import numpy as np
    from softdft_amd import abi, configs
    from softdft_amd.geometry_nd import build_state_points_nd
    from oracle.dft_oracle_nd import OracleDFTND
    dss = nd3_datasets()
    t0 = time.perf_counter()
    sps = build_state_points_nd(dss, device_bulk=True)
    setup = time.perf_counter() - t0
    sp0 = sps[0]
    log(f"ND: {len(sps)} points on a {sp0.grid} grid ({sp0.NP} points), setup {setup:.2f} s")
    b = abi.Batch(sps)
    opts = _solve_opts(sp0.numerics, max_iters=6000)
    xm, itm, em, tm = _timed_solve(b, b.x0(), abi.PREC_MIXED, opts)
    omm, alm, _ = b.observables()
    rhom = [b.get_field(abi.FIELD_RHO_BEADS, k) for k in range(len(sps))]
    x64, it64, e64, t64 = _timed_solve(b, b.x0(), abi.PREC_FP64, opts)
    om64, al64, _ = b.observables()
    rho64 = [b.get_field(abi.FIELD_RHO_BEADS, k) for k in range(len(sps))]
    # time of one batched FP64 objective evaluation (all points)
    b.objective(x64, abi.PREC_FP64)
    t1 = time.perf_counter()
    for _ in range(3):
        b.objective(x64, abi.PREC_FP64)
    tobj = (time.perf_counter() - t1) / 3
    b.close()
    log(f"ND: mixed {tm:.2f} s (mean {itm.mean():.0f} it, max {itm.max()}), FP64 {t64:.2f} s (mean {it64.mean():.0f} it), "
        f"FP64 objective {tobj * 1e3:.1f} ms per batch evaluation")
    k = len(sps) // 2
    x_k = x64[int(sum(sp.n for sp in sps[:k])):int(sum(sp.n for sp in sps[:k + 1]))]
    o = OracleDFTND(sps[k])
    t2 = time.perf_counter()
    fo = o.objective(x_k)
    tor = time.perf_counter() - t2
    om_or = o.grand_potential(x_k)
    out = {
        "report": "config", "config": "ND3 pore", "grid": list(sp0.grid), "NP": sp0.NP, "points": len(sps),
        "mixed_gpu": {"value": len(sps) / tm, "unit": "state points/s", "mean_iterations": float(itm.mean()),
                      "max_iterations": int(itm.max()), "max_residual": float(em.max())},
        "fp64_gpu": {"value": len(sps) / t64, "unit": "state points/s", "mean_iterations": float(it64.mean()),
                     "max_residual": float(e64.max())},
        "fp64_objective_ms_per_batch": tobj * 1e3,
        "fp64_objective_ms_per_point": tobj * 1e3 / len(sps),
        "cpu_baseline": {"value": 1.0 / (tor * float(it64[k])), "unit": "state points/s", "cores": 1, "kind": "port",
                         "sample": f"one oracle objective evaluation ({tor:.2f} s, fftn on the mirrored grid) x the "
                                   f"{int(it64[k])} iterations of the point's FP64 solve (estimate)"},
        "errors": {
            "omega_max_rel_mixed_vs_fp64": float(np.max(np.abs(omm - om64) / np.abs(om64))),
            "density_max_rel_mixed_vs_fp64": float(max(np.max(np.abs(a - c)) / np.max(np.abs(c)) for a, c in zip(rhom, rho64))),
            "ionization_max_abs_mixed_vs_fp64": float(np.max(np.abs(alm[:, 0, 1] - al64[:, 0, 1]))),
            "oracle_residual_at_gpu_fp64_solution": float(np.max(np.abs(fo - x_k))),
            "omega_fp64_rel_vs_oracle_at_same_x": float(abs(om64[k] - om_or) / abs(om_or)),
            "density_max_rel_fp64_vs_oracle_at_same_x": float(np.max(np.abs(rho64[k] - o.rho_beads)) / np.max(np.abs(o.rho_beads))),
        },
    }
    print(json.dumps(_finite(out)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=256)
    ap.add_argument("--singles", type=int, default=6, help="one-at-a-time FP64 sample size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="fixed launch rules instead of sdft_autotune")
    ap.add_argument("--cpu-cores", type=int, default=16, help="host cores of the all-core CPU baseline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--kernel-only", choices=["chain32", "chain64", "fwd32", "fwd64", "back32", "back64", "acc32", "acc64",
                                              "pw64", "c5fft_fwd32", "c5fft_fwd64", "c5fft_back32", "c5fft_back64",
                                              "nd3_64", "nd3_32"],
                    help="after one full-load evaluation, only re-launch this kernel class 122 times: "
                         "1 + 100 untimed (clock under load), then 1 + 20 timed with HIP events "
                         "(for rocprofv3 --kernel-trace / --pmc runs of that kernel); c5fft_*: the FFT "
                         "convolution stages of the 128-point 500-mer brush batch (NP 16384)")
    ap.add_argument("--config", choices=["c1", "c2", "c4", "c5", "nd3"],
                    help="per-config report (one JSON line) instead of the benchmark line")
    ap.add_argument("--c5-iters", type=int, default=1500, help="iteration cap of the C5 batch solve")
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

    if args.kernel_only and args.kernel_only.startswith("nd3_"):
        nd3_kernels(args.kernel_only)
        return
    if args.kernel_only and args.kernel_only.startswith("c5fft_"):
        c5_fft_kernels(args.kernel_only)
        return
    if args.config:
        from softdft_amd import configs
        if args.config == "c1":
            report_single("C1", configs.config1(), cpu_full=True, args=args)
        elif args.config == "c2":
            report_single("C2", configs.config2(), cpu_full=False, args=args)
        elif args.config == "c4":
            report_c4(args)
        elif args.config == "nd3":
            report_nd(args)
        else:
            report_c5(args)
        return

    import numpy as np
    from softdft_amd import abi, configs
    from softdft_amd.problem import build_state_points

    t = time.time()
    # host setup with every bulk state solved on the device in one launch (sdft_bulk_solve)
    sps = build_state_points(configs.config3(args.points), device_bulk=True)
    setup_s = time.time() - t
    log(f"setup: {len(sps)} state points in {setup_s:.2f} s (bulk states on the device)")
    batch = abi.Batch(sps)
    x0 = batch.x0()
    # fp32_switch: the mixed solve hands a point to the FP64 pipeline once ||F(x)-x||_inf < 3e-5
    # (measured on this workload: 1e-6 / 3e-6 / 1e-5 / 3e-5 / 1e-4 gave 265 / 157 / 151 / 145 / 148 ms
    # per solve with the same Omega to 1.4e-10 of the banded FP64 solve; below ~1e-5 the FP32
    # iterations stall on single-precision round-off)
    num = dict(tole=1e-8, damping=0.1, mixing_max=0.02, switch_tole=100.0, anderm=5, max_iters=3000,
               fp32_switch=3e-5)
    tuning = []
    if not args.no_tune:
        t = time.time()
        tuning = batch.autotune(reps=3)
        log(f"autotune: {len(tuning)} decisions in {time.time() - t:.2f} s")
        for e in tuning:
            log(f"  kind {e['kind']} key {e['key']} precision {e['precision']}: choice {e['choice']}, "
                f"candidates (us) {[round(u, 1) for u in e['us'] if u >= 0]}")

    if args.kernel_only:
        # one evaluation of every point in the class's precision with the banded operators, so that
        # every launch of the kernel in this run has all 256 points active (the full-load launch the
        # roofline figure times)
        prec = abi.PREC_FP32 if args.kernel_only.endswith("32") else abi.PREC_FP64   # pw64: FP64 fields
        batch.set_operator_band(1, 1)
        batch.objective(x0, prec)
        kc = {"chain32": abi.KP_CHAIN_FP32, "chain64": abi.KP_CHAIN_FP64, "fwd32": abi.KP_CONV_FWD_FP32,
              "fwd64": abi.KP_CONV_FWD_FP64, "back32": abi.KP_CONV_BACK_FP32, "back64": abi.KP_CONV_BACK_FP64,
              "acc32": abi.KP_ACCUM_FP32, "acc64": abi.KP_ACCUM_FP64, "pw64": abi.KP_POINTWISE_FP64}
        ms, fl = profile_warm(batch, kc[args.kernel_only], reps=20)
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
    # path-level HBM view: algorithmic bytes of every point's FP32 and FP64 iterations of one solve
    sw = batch.switch_iters()
    path_bytes = float(sum(int(sw[p]) * b_iter(sps[p], 4) + int(it[p] - sw[p]) * b_iter(sps[p], 8)
                           for p in range(args.points)))
    step_s = total / args.steps
    log(f"path: {path_bytes / 1e9:.1f} GB algorithmic per solve, {path_bytes / step_s / 1e12:.3f} TB/s")
    # the mixed solution of the points the CPU baseline solves too (compared with the oracle below)
    cores = max(1, min(args.cpu_cores, os.cpu_count() or 1))
    cpu_idx = np.linspace(0, args.points - 1, cores).astype(int)
    rho_mixed = {int(i): batch.get_field(abi.FIELD_RHO_BEADS, int(i)) for i in cpu_idx}
    if args.solve_only:
        log(f"solve-only: {value:.1f} state points/s over {args.steps} timed solves")
        batch.close()
        return

    # the same solve launched kernel by kernel instead of as captured poll windows (hipGraph), and
    # with the FP32 / FP64 pipelines in turn instead of on two streams
    batch.set_launch_mode(graph=False)
    t0 = time.perf_counter()
    xe, ite, erre = batch.solve(x0, precision=abi.PREC_MIXED, **num)
    eager_ms = (time.perf_counter() - t0) * 1e3
    batch.set_launch_mode(graph=True, concurrent=False)
    t0 = time.perf_counter()
    xs_, its_, errs_ = batch.solve(x0, precision=abi.PREC_MIXED, **num)
    seq_ms = (time.perf_counter() - t0) * 1e3
    batch.set_launch_mode(concurrent=True)
    same = bool(np.array_equal(xe, x) and np.array_equal(ite, it) and np.array_equal(xs_, x))
    log(f"eager launches: {eager_ms:.1f} ms per solve, pipelines in turn: {seq_ms:.1f} ms "
        f"(default: {sum(times) / len(times) * 1e3:.1f} ms), bitwise equal: {same}")

    # per-launch time of the hot kernel classes, HIP events on the library stream, with the
    # operators the mixed solve used; weighted by launches per evaluation x iterations run
    # (one chain launch runs every propagator step of the chain species)
    prof = {}
    for name, kc, prec, per_eval, iters_run in [
            ("k_chain<float> (banded propagators, all steps)", abi.KP_CHAIN_FP32, 32, 1, it32),
            ("k_chain<double> (banded propagators, all steps)", abi.KP_CHAIN_FP64, 64, 1, it64),
            ("k_convgemm<float> (forward weighted densities)", abi.KP_CONV_FWD_FP32, 32, 1, it32),
            ("k_convgemm<double> (forward weighted densities)", abi.KP_CONV_FWD_FP64, 64, 1, it64),
            ("conv GEMMs<float> (back projections: shared partials and projections, one launch)", abi.KP_CONV_BACK_FP32, 32, 1, it32),
            ("conv GEMMs<double> (back projections: shared partials and projections, one launch)", abi.KP_CONV_BACK_FP64, 64, 1, it64)]:
        try:
            ms, fl = profile_warm(batch, kc, reps=20)
        except abi.SdftError as e:  # e.g. the autotuner chose the per-step propagators
            log(f"kernel {name}: not profiled ({e})")
            continue
        prof[name] = (ms, fl, prec, ms * per_eval * iters_run)
        log(f"kernel {name}: {ms * 1e3:.1f} us/launch, {fl / ms / 1e9:.2f} TFLOP/s, "
            f"{per_eval} launches/evaluation x {iters_run} iterations = {ms * per_eval * iters_run:.1f} ms per solve")

    # density accumulation (memory-bound): bytes of one launch = the stored propagators gC and gP
    # read once (2 N S^2 NP values per chain species) + the bead and bond densities written twice
    # (plain and trapezoid-weighted)
    sp0 = sps[0]
    prop_vals = sum(2 * len(s.sequence) * max(len(f) for f in s.state_family) ** 2
                    for s in sp0.ms.species if len(s.sequence) > 1)
    for name, kc, w, iters_run in [("k_accumulate<float> (densities)", abi.KP_ACCUM_FP32, 4, it32),
                                   ("k_accumulate<double> (densities)", abi.KP_ACCUM_FP64, 8, it64)]:
        ms, _ = profile_warm(batch, kc, reps=20)
        nbytes = args.points * sp0.NP * w * (prop_vals + 2 * (sp0.NB + 2 * sp0.MB))
        log(f"kernel {name}: {ms * 1e3:.1f} us/launch, {nbytes / ms / 1e9:.2f} TB/s of {nbytes / 1e6:.0f} MB, "
            f"x {iters_run} iterations = {ms * iters_run:.1f} ms per solve")

    # plain FP64 on the same batch (timed, one solve each): the dense bond operator (the reference's
    # full operator) and the banded one, so the mixed-precision gain is shown apart from batching;
    # the dense solve is also the reference of the acceptance bound
    t0 = time.perf_counter()
    xd, itd, errd = batch.solve(x0, precision=abi.PREC_FP64, **num)
    t_fp64 = time.perf_counter() - t0
    om_fp64, al_fp64, _ = batch.observables()
    t0 = time.perf_counter()
    xb, itb, errb = batch.solve(x0, precision=abi.PREC_FP64, band_fp32=2, **num)
    t_fp64b = time.perf_counter() - t0
    om_fp64b, _, _ = batch.observables()
    rel = np.abs(om_mixed - om_fp64) / np.abs(om_fp64)
    rel_b = float(np.max(np.abs(om_fp64b - om_fp64) / np.abs(om_fp64)))
    dal = float(np.nanmax(np.abs(al_mixed[:, 0, 1] - al_fp64[:, 0, 1])))  # chain species' charged fraction
    log(f"batched FP64: dense {t_fp64 * 1e3:.1f} ms ({args.points / t_fp64:.1f} points/s, mean {itd.mean():.1f} it), "
        f"banded {t_fp64b * 1e3:.1f} ms ({args.points / t_fp64b:.1f} points/s); banded vs dense dOmega {rel_b:.2e}")
    log(f"max rel dOmega mixed vs FP64: {rel.max():.3e}, max |d alpha|: {dal:.3e}")
    batch.close()

    # plain FP64, one state point at a time
    sel = np.linspace(0, args.points - 1, args.singles).astype(int)
    single_tuning = []
    if not args.no_tune:  # tuned once for the one-point layout, replayed on every one-point context
        b1 = abi.Batch([sps[sel[0]]])
        single_tuning = b1.autotune(reps=3)
        b1.close()
    t0 = time.perf_counter()
    for i in sel:
        b1 = abi.Batch([sps[i]])
        if single_tuning:
            b1.set_tuning(single_tuning)
        b1.solve(b1.x0(), precision=abi.PREC_FP64, **num)
        b1.close()
    single_rate = len(sel) / (time.perf_counter() - t0)
    log(f"FP64 one-at-a-time: {single_rate:.2f} points/s")

    cpu = cpu1 = vs_oracle = None
    if not args.no_cpu:
        from oracle.dft_oracle import OracleDFT, anderson_solve
        sp = sps[int(np.argmin(np.abs(np.linspace(2, 12, args.points) - 7.0)))]
        o = OracleDFT(sp)
        t0 = time.perf_counter()
        r = anderson_solve(o.objective, sp.x0(), sp.numerics)
        dt = time.perf_counter() - t0
        cpu1 = {"value": 1.0 / dt, "unit": "state points/s", "cores": 1, "kind": "port",
                "sample": f"1 of the {args.points} state points (pH 7.0), FP64 NumPy restatement of the reference "
                          f"algorithm, {r.iters} Anderson iterations in {dt:.1f} s on one host core"}
        log(f"cpu baseline: {dt:.1f} s for one point on one core")
        # all host cores: one process per state point (points are independent), a bounded sample
        idx = cpu_idx
        res, dt_all = cpu_points_parallel([sps[i] for i in idx], cores)
        n_it = sum(r["iters"] for r in res)
        cpu = {"value": len(idx) / dt_all, "unit": "state points/s", "cores": cores, "kind": "port",
               "sample": f"{len(idx)} of the {args.points} state points spread over pH 2..12, one process per point "
                         f"on {cores} host cores, FP64 NumPy restatement of the reference algorithm "
                         f"({n_it} Anderson iterations in total, {dt_all:.1f} s wall)"}
        log(f"cpu baseline: {len(idx)} points on {cores} cores in {dt_all:.1f} s")
        # the GPU mixed batch against the oracle's converged FP64 solutions of those points (both solved
        # to 1e-8): Omega, bead densities, excess adsorption, degree of ionization of the chain
        e_om, e_rho, e_ads, e_al = [], [], [], []
        for i, r in zip(idx, res):
            i = int(i)
            e_om.append(abs(om_mixed[i] - r["omega"]) / abs(r["omega"]))
            e_rho.append(float(np.max(np.abs(rho_mixed[i] - r["rho"])) / np.max(np.abs(r["rho"]))))
            ad_g, ad_o = _adsorption(sps[i], rho_mixed[i]), _adsorption(sps[i], r["rho"])
            e_ads.append(float(np.max(np.abs(ad_g - ad_o)) / np.max(np.abs(ad_o))))
            e_al.append(float(np.nanmax(np.abs(al_mixed[i, 0, :2] - r["alpha"][0][:2]))))
        vs_oracle = {"points": len(idx), "max_rel_dOmega": max(e_om), "max_rel_drho_beads": max(e_rho),
                     "max_rel_dadsorption": max(e_ads), "max_abs_dalpha": max(e_al),
                     "note": "GPU mixed-precision batch vs the FP64 oracle's own solves of the same points, "
                             "both to ||F(x)-x||_inf < 1e-8"}
        log(f"mixed vs oracle on {len(idx)} points: dOmega {max(e_om):.2e}, drho {max(e_rho):.2e}, "
            f"dadsorption {max(e_ads):.2e}, dalpha {max(e_al):.2e}")

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
                   "points": args.points, "NP": 801, "precision": "FP32 -> FP64 polish at 3e-5",
                   "parallelism": "one GPU, batched state points"},
        "fp64_one_at_a_time": {"value": single_rate, "unit": "state points/s", "sample_points": int(len(sel))},
        "speedup_vs_fp64_one_at_a_time": value / single_rate,
        "fp64_batched": {"dense": {"value": args.points / t_fp64, "unit": "state points/s",
                                   "mean_iterations": float(itd.mean())},
                         "banded": {"value": args.points / t_fp64b, "unit": "state points/s",
                                    "mean_iterations": float(itb.mean()), "max_rel_dOmega_vs_dense": rel_b}},
        "speedup_mixed_vs_fp64_batched": {"dense": value * t_fp64 / args.points,
                                          "banded": value * t_fp64b / args.points},
        "max_rel_dOmega_vs_fp64": float(rel.max()),
        "max_abs_dalpha_vs_fp64": dal,
        "mean_iterations": float(np.mean([i.mean() for i in iters])),
        "roofline": {"bound": "mfma", "kernel": name, "achieved": achieved, "peak": PEAK[prec],
                     "unit": "TFLOP/s", "frac": achieved / PEAK[prec], "traffic": traffic_from_profiles(name)},
        "roofline_hbm": {"bound": "hbm", "achieved": path_bytes / step_s / 1e9, "peak": 8000.0, "unit": "GB/s",
                         "frac": path_bytes / step_s / 8.0e12, "bytes_per_solve": path_bytes,
                         "note": "whole solve: algorithmic bytes of every point's FP32 and FP64 iterations "
                                 "(B_iter per iteration and point) over the solve's wall time"},
        "errors_vs_oracle": vs_oracle,
        "cpu_baseline": cpu,
        "cpu_baseline_one_core": cpu1,
        "setup_s": setup_s,
        "host": {"cpu_count": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0)),
                 "cpu_model": _cpu_model()},
        "kernels_us": {k: v[0] * 1e3 for k, v in prof.items()},
        "kernels_ms_per_solve_est": {k: v[3] for k, v in prof.items()},
        "pipeline_iterations": {"fp32": it32, "fp64": it64},
        "launch": {"mode": "hipGraph per poll window, FP32 and FP64 pipelines on two streams",
                   "eager_ms_per_step": eager_ms, "pipelines_in_turn_ms_per_step": seq_ms,
                   "bitwise_equal": same},
        "tuning": [{"kind": {1: "propagator", 2: "conv_fwd", 3: "conv_pre", 4: "conv_back"}[e["kind"]],
                    "key": e["key"], "precision": e["precision"], "choice": e["choice"],
                    "us": [round(u, 2) for u in e["us"] if u >= 0]} for e in tuning],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

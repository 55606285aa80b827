"""Benchmark: SCFT iterations/s on the 128^3 AB diblock (Ns = 100), FP64 and
mixed precision, with the HBM roofline of the contour step and the CPU
oracle timed on the host beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A "step" is one SCFT iteration (both propagator sweeps, Q, densities,
Anderson field update, F).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E peak

# 128^3 AB diblock, box 8.0^3 Rg, fA = 0.30, chiN = 25, Ns = 100 (ds = 0.01),
# seeded N(0, 0.1) fields plus the hexagonal-cylinder seed.
WORKLOAD = dict(shape=(128, 128, 128), box=8.0, fA=0.30, chiN=25.0, ds=0.01,
                sigma=0.1, seed=20251203, seed_kind="hex", amplitude=1.0)


def algorithmic_bytes_step(N: int, bp: int) -> int:
    """Bytes of one dual contour step: each of the two sweeps reads q_s and
    the w-factor and writes q_{s+1} once."""
    return 2 * 3 * bp * N


def algorithmic_bytes_iter(N: int, Ns: int, bp: int, m: int) -> int:
    """B_iter = N [6 Ns bp + (Ns+1) bp + 80 + (m > 0 ? (6m+4) 8 : 0)]."""
    return N * (6 * Ns * bp + (Ns + 1) * bp + 80 + ((6 * m + 4) * 8 if m > 0 else 0))


COUNTER_FILES = {"lines": "k_line_reg_128cubed_hbm_counters.md",
                 "planes": "k_plane_128cubed_hbm_counters.md"}


def counter_traffic(mode):
    """HBM bytes per dual contour step of sweep mode `mode` at 128^3 from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated on a copy
    of known size): ({"fp64": B, "fp32": B}, file) or (None, file)."""
    import re
    name = COUNTER_FILES[mode]
    try:
        m = re.search(r"traffic_per_dual_step_bytes: fp64=(\d+) fp32=(\d+)",
                      open(os.path.join(ROOT, "profiles", name)).read())
    except OSError:
        return None, name
    return (None if not m else {"fp64": int(m.group(1)), "fp32": int(m.group(2))}), name


ROOFLINE_KERNEL = {
    "planes": "k_plane + fused x line launch: one dual contour step (both sweeps; one launch per "
              "sweep step, plane launches of the z and y round trips of two steps alternating with "
              "fused x launches)",
    "lines": "k_line_reg / k_line_tile: one dual contour step (both sweeps; two line launches per "
             "sweep step: strided + fused boundary pair)"}


def make_engine(prec, w):
    from polyswift_plusplus_amd import DiblockSCFT, fields
    shape = w["shape"]
    cell = tuple(w["box"] / n for n in shape)
    e = DiblockSCFT(shape, cell, w["fA"], w["ds"], w["chiN"], prec=prec)
    wA, wB = fields.seeded_fields(shape, cell, sigma=w["sigma"], seed=w["seed"],
                                  seed_kind=w["seed_kind"], amplitude=w["amplitude"])
    e.setFields(wA, wB)
    return e


def time_iters(e, mix, warmup, steps):
    if warmup:
        e.iterate(warmup, mix, use_graph=True)
    t0 = time.perf_counter()
    h = e.iterate(steps, mix, use_graph=True)   # returns after a stream sync
    dt = time.perf_counter() - t0
    assert h["iters"] == steps
    return dt, h


def config5_relax_leg(tol_stress=1e-7, tol_scft=1e-10, seed_kind="hex", checkpoint=20):
    """Box relaxation at the differentiation workload's full size: 96^3, box
    (7.2, 7.6, 8.0) Rg to start, fA = 0.32, chiN = 22, Ns = 200 (ds = 0.005),
    FP64 sweeps with contour checkpoints every 20 steps, the trapezoid contour
    rule (under which the adjoint stress is the exact derivative of F), to
    |dF/dL| < 1e-7.  Seeded N(0, 0.1) fields plus an ordered seed: from the
    Gaussian fields alone the iteration settles on the disordered phase,
    whose stress vanishes at every box.  The fields are first converged with
    the mixed-precision protocol, then every outer step re-converges them in
    FP64 from the last ones."""
    from polyswift_plusplus_amd import DiblockSCFT, Mixing, fields
    from polyswift_plusplus_amd.relax import mixed_precision_solve, relax_box
    shape, box = (96, 96, 96), (7.2, 7.6, 8.0)
    cell = tuple(b / n for b, n in zip(box, shape))
    mix = Mixing("anderson", lambda0=0.2, depth=20, beta=1.0, anderson_start=0.05)
    e = DiblockSCFT(shape, cell, 0.32, 0.005, 22.0, prec="fp32", quadrature="trapezoid")
    e.setFields(*fields.seeded_fields(shape, cell, sigma=0.1, seed=5, seed_kind=seed_kind, amplitude=1.0))
    t0 = time.perf_counter()
    r0 = mixed_precision_solve(e, tol=tol_scft, switch_tol=3e-5, max_iters=20000, mixing=mix, keep_history=True)
    print(f"[bench] config5 relax: first solve {r0['fp32_iters']} FP32 + {r0['fp64_iters']} FP64 iterations, "
          f"err {r0['err']:.2e} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    if checkpoint:
        e.setCheckpointing(checkpoint)
    t1 = time.perf_counter()
    r = relax_box(e, tol_stress=tol_stress, tol_scft=tol_scft, max_outer=30, first_step=5.0, mixing=mix,
                  max_iters=20000, log=lambda m: print("[bench] config5 relax: " + m, file=sys.stderr, flush=True))
    dt = time.perf_counter() - t1
    e.close()
    hist = [dict(L=[float(x) for x in h["L"]], dFdL=[float(x) for x in h["dFdL"]], F=float(h["F"]),
                 iters=h["iters"], scft_err=h["scft_err"]) for h in r["history"]]
    return {"workload": "AB diblock 96^3, box (7.2, 7.6, 8.0) Rg to start, fA=0.32, chiN=22, Ns=200 "
                        f"(ds=0.005), FP64, checkpoint every {checkpoint} steps, trapezoid contour rule, "
                        f"N(0, 0.1) + {seed_kind} seed",
            "converged": r["converged"], "reason": r["reason"], "L": [float(x) for x in r["L"]],
            "max_abs_dFdL": float(np.max(np.abs(r["history"][-1]["dFdL"]))), "outer_steps": len(hist),
            "scft_iterations": int(sum(h["iters"] for h in hist)), "seconds": dt, "history": hist}


def cache_probe(reps=10, sizes_mb=(8, 16, 32, 64, 128, 256, 512, 1024, 2048)):
    """Back-to-back read-only sweeps and copies over buffers of each size
    (psx_probe_stream): GB/s of the bytes each sweep touches."""
    import ctypes
    from polyswift_plusplus_amd import _abi
    L = _abi.load()
    out = []
    for mb in sizes_mb:
        row = {"MiB": mb}
        for mode, name in ((0, "read"), (1, "copy")):
            ms = ctypes.c_double()
            _abi.check(L.psx_probe_stream(mb << 20, reps, mode, ctypes.byref(ms)))
            moved = (mb << 20) * (1 if mode == 0 else 2)
            row[name + "_us"] = ms.value * 1e3
            row[name + "_GBs"] = moved / (ms.value * 1e-3) / 1e9
        out.append(row)
        print(f"[bench] probe {mb} MiB: read {row['read_GBs']:.0f} GB/s, copy {row['copy_GBs']:.0f} GB/s",
              file=sys.stderr, flush=True)
    return out


def solution_state(e):
    """(phiA, phiB, Q, F) of one evaluation at the engine's fields, in its
    current precision."""
    e.solveQ()
    Q = e.calcBigQ()
    phiA, phiB = e.densities()
    return phiA, phiB, Q, e.getFreeE()


def state_error(a, b):
    """|dF|, max|dphi| and |dQ|/Q of state a against state b, and the
    largest difference of the Fourier moduli |phi^(k)|/N (translation
    invariant: independently converged runs can settle on slightly shifted
    copies of the same periodic pattern)."""
    spec = max(float(np.max(np.abs(np.abs(np.fft.rfftn(x)) - np.abs(np.fft.rfftn(y))))) / x.size
               for x, y in ((a[0], b[0]), (a[1], b[1])))
    return dict(dF=abs(a[3] - b[3]),
                max_dphi=float(max(np.max(np.abs(a[0] - b[0])), np.max(np.abs(a[1] - b[1])))),
                dQ_rel=abs(a[2] / b[2] - 1.0), max_dphi_fourier_moduli=spec)


def accuracy_leg(w, mix, tol, max_iters, switch_tol, max_fp32):
    """FP64 run to tol vs the mixed protocol (FP32 sweeps until the residual
    falls below switch_tol, then FP64 iterations to tol with a fresh, deeper
    Anderson history) from the same seed.  The FP64 engine also runs the
    mixed protocol's restart on its own (FP64 to switch_tol, then the fresh
    deeper history): the like-for-like time to solution.

    Each solution is evaluated once more in FP64 at its final fields; the
    mixed result states |dF|, max|dphi| and |dQ|/Q against the FP64 solution
    (whose evaluation the oracle reproduces to 1e-10 in the tests), and one
    FP32 evaluation at the FP64 solution gives the FP32 sweep's own error."""
    from polyswift_plusplus_amd.relax import tail_mixing
    tail = tail_mixing(mix)
    e64 = make_engine("fp64", w)
    t0 = time.perf_counter()
    h64 = e64.iterate(max_iters, mix, tol=tol)
    t64 = time.perf_counter() - t0
    s64 = solution_state(e64)
    e64.setPrecision("fp32")
    s32 = solution_state(e64)
    e64.close()
    er = make_engine("fp64", w)
    t0 = time.perf_counter()
    hr1 = er.iterate(max_iters, mix, tol=switch_tol)
    er.resetHistory()
    hr2 = er.iterate(max_iters, tail, tol=tol)
    tr = time.perf_counter() - t0
    sr = solution_state(er)
    er.close()
    em = make_engine("fp32", w)
    t0 = time.perf_counter()
    h32 = em.iterate(max_fp32, mix, tol=switch_tol)
    em.setPrecision("fp64")
    em.resetHistory()
    hfin = em.iterate(max_iters, tail, tol=tol)
    tm = time.perf_counter() - t0
    sm = solution_state(em)
    em.close()
    F64 = float(h64["F"][-1])
    Fm = float(hfin["F"][-1])
    err_m = state_error(sm, s64)
    return dict(F_fp64=F64, F_mixed=Fm, dF=err_m["dF"], max_dphi=err_m["max_dphi"], dQ_rel=err_m["dQ_rel"],
                max_dphi_fourier_moduli=err_m["max_dphi_fourier_moduli"],
                error_vs="FP64 solution evaluated at its fields (the oracle reproduces that evaluation to "
                         "1e-10 at 128^3 in tests/test_gpu_configs.py)",
                mixed_vs_fp64_restart=state_error(sm, sr),
                fp32_sweep_at_fp64_solution=state_error(s32, s64),
                fp64_iters=int(h64["iters"]), fp64_err=float(h64["err"][-1]),
                mixed_fp32_iters=int(h32["iters"]), mixed_fp64_iters=int(hfin["iters"]),
                mixed_err=float(hfin["err"][-1]), fp64_seconds=t64, mixed_seconds=tm,
                fp64_restart_iters=[int(hr1["iters"]), int(hr2["iters"])],
                fp64_restart_seconds=tr, mixed_vs_fp64_time=t64 / tm,
                mixed_vs_fp64_restart_time=tr / tm,
                anderson_depth=mix.depth, mixed_tail_depth=tail.depth, switch_tol=switch_tol,
                converged=bool(h64["err"][-1] < tol and hfin["err"][-1] < tol and hr2["err"][-1] < tol))


def batch_leg(n_inst=64, steps=5, warmup=2, seed=4, sequential=True):
    """Parameter sweep: n_inst instances, grid drawn from {32,48,64,96}^3,
    chiN ~ U[12, 30], fA ~ U[0.20, 0.50] rounded to 0.01 (the reference
    needs an integral fA/ds), box = grid x 0.08 Rg, Ns = 100, FP32 sweeps,
    Anderson m = 10, one hipGraph per batch iteration, autotuned line
    kernels -- against the same instances solved one after another with
    the plain FP64 engine."""
    from polyswift_plusplus_amd import BatchSCFT, DiblockSCFT, Mixing, fields
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_inst):
        n = int(rng.choice([32, 48, 64, 96]))
        specs.append(dict(shape=(n, n, n), cell=(0.08,) * 3, fA=round(float(rng.uniform(0.20, 0.50)), 2),
                          ds=0.01, chiN=float(rng.uniform(12.0, 30.0))))
    mix = Mixing("anderson", lambda0=0.2, depth=10, beta=1.0, anderson_start=0.05)
    b = BatchSCFT(specs, prec="fp32")
    for i, sp in enumerate(specs):
        b.instances[i].setFields(*fields.seeded_fields(sp["shape"], sp["cell"], sigma=0.1, seed=404,
                                                       instance=i))
    choice = b.autotune(5)
    order = []
    for sp in specs:
        if sp["shape"][0] not in order:
            order.append(sp["shape"][0])
    modes = {n: b.sweepMode(c) for c, n in enumerate(order)}
    b.iterate(warmup, mix)
    t0 = time.perf_counter()
    b.iterate(steps, mix)
    dt = time.perf_counter() - t0
    b.close()
    sizes = {n: sum(1 for sp in specs if sp["shape"][0] == n) for n in (32, 48, 64, 96)}
    if not sequential:
        return dict(instances=n_inst, sizes=sizes, steps=steps, aggregate_iter_per_s=n_inst * steps / dt,
                    autotune_choice=choice, sweep_modes=modes)
    seq = 0.0
    for i, sp in enumerate(specs):
        e = DiblockSCFT(sp["shape"], sp["cell"], sp["fA"], sp["ds"], sp["chiN"], prec="fp64")
        e.setFields(*fields.seeded_fields(sp["shape"], sp["cell"], sigma=0.1, seed=404, instance=i))
        e.iterate(warmup, mix)
        t0 = time.perf_counter()
        e.iterate(steps, mix)
        seq += time.perf_counter() - t0
        e.close()
    return dict(instances=n_inst, sizes=sizes, steps=steps,
                aggregate_iter_per_s=n_inst * steps / dt,
                sequential_fp64_iter_per_s=n_inst * steps / seq,
                speedup=seq / dt, autotune_choice=choice, sweep_modes=modes)


def config5_leg(iters=20, warmup=3):
    """The differentiation workload at full size: 96^3, box (7.2, 7.6, 8.0) Rg,
    fA = 0.32, chiN = 22, Ns = 200 (ds = 0.005), FP64, seeded lamellar
    fields.  Per contour mode (every slice stored; backward slices
    checkpointed every 20 steps): SCFT iterations/s (Anderson m = 10, one
    hipGraph each) and the time of one propagate + densities + adjoint box
    stress, against the slice memory the mode holds."""
    from polyswift_plusplus_amd import DiblockSCFT, Mixing, fields
    shape, box = (96, 96, 96), (7.2, 7.6, 8.0)
    cell = tuple(b / n for b, n in zip(box, shape))
    mix = Mixing("anderson", lambda0=0.1, depth=10, beta=0.1, anderson_start=1e300)
    out = {"workload": "AB diblock 96^3, box (7.2, 7.6, 8.0) Rg, fA=0.32, chiN=22, Ns=200 (ds=0.005), FP64"}
    for K in (0, 20):
        e = DiblockSCFT(shape, cell, 0.32, 0.005, 22.0, prec="fp64")
        e.setFields(*fields.seeded_fields(shape, cell, sigma=0.1, seed=5, seed_kind="lamellar", amplitude=1.0))
        if K:
            e.setCheckpointing(K)
        e.iterate(warmup, mix)
        t0 = time.perf_counter()
        e.iterate(iters, mix)
        dt = time.perf_counter() - t0
        e.solveQ()
        e.densities()
        e.boxStress()   # warm
        t0 = time.perf_counter()
        e.solveQ()
        e.densities()
        g = e.boxStress()
        ts = time.perf_counter() - t0
        Ns = e.Ns
        # stored: both propagators' Ns + 1 slices; checkpointed: the backward
        # checkpoints, K + 1 recomputed backward slices, two forward chunks of
        # K + 1 and three scratch slices
        slices = 2 * (Ns + 1) if not K else (Ns // K + 1 + (1 if Ns % K else 0)) + 3 * (K + 1) + 3
        out["stored" if not K else f"checkpoint_every_{K}"] = dict(
            iter_per_s=iters / dt, stress_seconds=ts, dFdL=[float(x) for x in g],
            slice_gib=slices * 8 * int(np.prod(shape)) / 2 ** 30, sweep_mode=e.sweepMode()[0])
        e.close()
    return out


def host_info():
    """Host CPU model, cores visible to this process, and the oracle's build."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "cpu_count": os.cpu_count(), "cores_available": avail,
            "compiler": "g++ -O2 -std=c++17 -fopenmp (oracle/Makefile)"}


def cpu_baseline_iterations(w, mix, threads, warmup=1, runs=5):
    """Full faithful oracle iterations (full-grid complex FFT pair and the
    separate element-wise passes of the reference, propagation, Q,
    quadrature, Anderson update and F) on the same grid and seed, each timed
    on the host's steady clock; the median of `runs` one-iteration runs after
    `warmup` untimed ones."""
    import oracle as O
    from polyswift_plusplus_amd import fields
    O.set_threads(threads)
    shape = w["shape"]
    cell = tuple(w["box"] / n for n in shape)
    wA, wB = fields.seeded_fields(shape, cell, sigma=w["sigma"], seed=w["seed"],
                                  seed_kind=w["seed_kind"], amplitude=w["amplitude"])
    P = O.Params(*shape, *cell, ds=w["ds"], fA=w["fA"], chiN=w["chiN"], mix=1, lambda0=mix.lambda0,
                 depth=mix.depth, beta=mix.beta, anderson_start=mix.anderson_start)
    r = O.scft_run_timed(P, wA, wB, n_iters=warmup + runs)
    O.set_threads(1)
    t = float(np.median(r["seconds"][warmup:]))
    return {"value": 1.0 / t, "unit": "SCFT iter/s", "cores": threads, "kind": "port",
            "sample": f"median of {runs} full faithful FP64 SCFT iterations of the bench workload "
                      f"({shape[0]}^3, Ns={int(round(1 / w['ds']))}, Anderson m={mix.depth}) after "
                      f"{warmup} untimed, {threads} OpenMP threads: {t:.2f} s per iteration "
                      f"(runs {[round(float(x), 2) for x in r['seconds'][warmup:]]})"}


def cpu_baseline_steps(w, seconds_target=12.0, threads=1):
    """Faithful oracle contour steps (full-grid complex FFT pair, separate
    element-wise passes) on `threads` host threads, timed on a bounded
    sample of the same grid and scaled to one iteration's 2 Ns steps (a
    full single-thread iteration takes minutes at 128^3; the O(N) work of an
    iteration is about 1% of it)."""
    import oracle as O
    O.set_threads(threads)
    shape = w["shape"]
    Ns = int(round(1.0 / w["ds"]))
    t1 = O.time_steps(shape, 1)
    steps = max(1, min(16, int(seconds_target / max(t1, 1e-3))))
    t = O.time_steps(shape, steps)
    per_step = t / steps
    it_s = 1.0 / (2 * Ns * per_step)
    O.set_threads(1)
    return {"value": it_s, "unit": "SCFT iter/s", "cores": threads, "kind": "port",
            "sample": f"{steps} faithful FP64 contour steps on {shape[0]}x{shape[1]}x{shape[2]} "
                      f"({per_step:.3f} s/step), scaled to 2*Ns={2 * Ns} steps per iteration"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--depth", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default line kernels (no per-pass autotuning)")
    ap.add_argument("--batch-profile", action="store_true",
                    help="only the batched half of the batch leg (for rocprofv3)")
    ap.add_argument("--config5", action="store_true",
                    help="only the differentiation workload (96^3, Ns = 200, FP64): iterations/s and "
                         "adjoint box stress time, with and without contour checkpointing")
    ap.add_argument("--config5-relax", action="store_true",
                    help="only the box relaxation of the differentiation workload (96^3, Ns = 200, FP64, "
                         "checkpointed) to |dF/dL| < 1e-7; prints progress per outer step")
    ap.add_argument("--cache-probe", action="store_true",
                    help="only the memory probe: read-only sweeps and copies over 8 MiB .. 2 GiB buffers "
                         "(--steps sweeps each), to tell cache-resident rates from HBM's")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3 counters: the dual contour step of "
                         "both precisions (5 reps each) and one FP64 copy of known size")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs one grid on one GPU; "
                          "there is no multi-GPU path (use --gpus 1)"}))
        sys.exit(2)

    from polyswift_plusplus_amd import Mixing

    w = dict(WORKLOAD)
    w["shape"] = (args.grid,) * 3
    N = int(np.prod(w["shape"]))
    if args.batch_profile:
        print(json.dumps(batch_leg(sequential=False)))
        return
    if args.config5:
        print(json.dumps(config5_leg()))
        return
    if args.config5_relax:
        print(json.dumps(config5_relax_leg()))
        return
    if args.cache_probe:
        print(json.dumps(cache_probe(reps=args.steps)))
        return
    if args.profile:
        # Counter calibration: one k_to_f64<double> over N cells moves exactly
        # 8N bytes in and 8N bytes out (psx_scft_get_slice_host on FP64).
        out = {}
        for prec in ("fp64", "fp32"):
            e = make_engine(prec, w)
            step_ms, pass_ms = e.timeStep(2)
            if prec == "fp64":
                e.slice(False, 1)
            out[prec] = dict(step_us=step_ms * 1e3, pass_us=[p * 1e3 for p in pass_ms],
                             mode=e.sweepMode()[0])
            e.close()
        out["calibration_copy_bytes_each_way"] = 8 * N
        print(json.dumps(out))
        return
    mix = Mixing("anderson", lambda0=0.1, depth=args.depth, beta=0.1, anderson_start=1e300)

    res = {}
    legs = {}
    for prec in ("fp64", "fp32"):
        e = make_engine(prec, w)
        choice = None if args.no_tune else e.autotune(5)
        mode = e.sweepMode()[0]
        print(f"[bench] {prec}: {mode}, line kernels {choice}, warmup {args.warmup}, timing {args.steps} "
              "iterations", file=sys.stderr, flush=True)
        dt, h = time_iters(e, mix, args.warmup, args.steps)
        step_ms, pass_ms = e.timeStep(10)
        Ns = e.Ns
        legs[prec] = dict(it_s=args.steps / dt, ms_per_iter=1e3 * dt / args.steps,
                          step_ms=step_ms, pass_ms=pass_ms, F_last=float(h["F"][-1]),
                          err_last=float(h["err"][-1]), line_kernels=choice, mode=mode)
        e.close()
        print(f"[bench] {prec}: {legs[prec]['it_s']:.2f} it/s, step {step_ms * 1e3:.1f} us "
              f"(passes {[round(p * 1e3, 1) for p in pass_ms]} us)", file=sys.stderr, flush=True)

    acc = None
    if not args.no_accuracy:
        print("[bench] accuracy leg (FP64 vs mixed protocol)", file=sys.stderr, flush=True)
        acc = accuracy_leg(w, Mixing("anderson", lambda0=0.2, depth=args.depth, beta=1.0,
                                     anderson_start=0.05),
                           tol=1e-8, max_iters=1500, switch_tol=2e-5, max_fp32=1500)
        print(f"[bench] accuracy: {acc}", file=sys.stderr, flush=True)

    batch = None
    if not args.no_batch:
        print("[bench] batch leg (64-instance parameter sweep)", file=sys.stderr, flush=True)
        batch = batch_leg()
        print(f"[bench] batch: {batch}", file=sys.stderr, flush=True)

    cpu = cpu_1 = host = None
    if not args.no_cpu:
        host = host_info()
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0")) or (host["cores_available"] or 1)
        nthr = max(1, min(nthr, 16))
        print(f"[bench] cpu baseline (oracle, full iterations, {nthr} threads)", file=sys.stderr, flush=True)
        cpu = cpu_baseline_iterations(w, Mixing("anderson", lambda0=0.1, depth=args.depth, beta=0.1,
                                                anderson_start=1e300), nthr)
        print(f"[bench] cpu baseline: {cpu['sample']}", file=sys.stderr, flush=True)
        print("[bench] cpu baseline (oracle contour steps, 1 thread)", file=sys.stderr, flush=True)
        cpu_1 = cpu_baseline_steps(w)

    m = legs["fp32"]
    d = legs["fp64"]
    traffic_m, src_m = counter_traffic(m["mode"])
    traffic_d, src_d = counter_traffic(d["mode"])
    bp_m, bp_d = 4, 8
    ach_m = algorithmic_bytes_step(N, bp_m) / (m["step_ms"] * 1e-3) / 1e9
    ach_d = algorithmic_bytes_step(N, bp_d) / (d["step_ms"] * 1e-3) / 1e9
    out = {
        "metric": "SCFT iter/s, 128^3 AB diblock Ns=100: FP64 vs mixed @|dF|<1e-8; % HBM peak",
        "value": m["it_s"],
        "unit": "SCFT iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_iter"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweep (FFT, propagator) + f64 Q/quadrature/update/F",
        "data": "synthetic (seeded N(0,0.1) + hexagonal-cylinder seed)",
        "config": {"workload": f"AB diblock {w['shape'][0]}^3, box {w['box']}^3 Rg, fA={w['fA']}, "
                               f"chiN={w['chiN']}, Ns=100 (ds={w['ds']}), Anderson m={args.depth}",
                   "grid": list(w["shape"]), "Ns": 100, "parallelism": "single GPU"},
        "fp64_iter_per_s": d["it_s"],
        "fp64_ms_per_step": d["ms_per_iter"],
        "mixed_vs_fp64": m["it_s"] / d["it_s"],
        "algorithmic_hbm_frac": {
            "fp64": algorithmic_bytes_iter(N, 100, 8, args.depth) * d["it_s"] / (HBM_PEAK_GBS * 1e9),
            "mixed": algorithmic_bytes_iter(N, 100, 4, args.depth) * m["it_s"] / (HBM_PEAK_GBS * 1e9)},
        "roofline": {"kernel": ROOFLINE_KERNEL[m["mode"]],
                     "bound": "hbm", "achieved": ach_m, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_m / HBM_PEAK_GBS,
                     "traffic": (traffic_m or {}).get("fp32") if args.grid == 128 else None,
                     "traffic_source": f"profiles/{src_m} (bytes per dual step)",
                     "sweep_mode": m["mode"],
                     "step_us": m["step_ms"] * 1e3, "pass_us": [p * 1e3 for p in m["pass_ms"]],
                     "line_kernels": m["line_kernels"],
                     "fp64": {"achieved": ach_d, "frac": ach_d / HBM_PEAK_GBS,
                              "traffic": (traffic_d or {}).get("fp64") if args.grid == 128 else None,
                              "traffic_source": f"profiles/{src_d} (bytes per dual step)",
                              "sweep_mode": d["mode"],
                              "step_us": d["step_ms"] * 1e3, "pass_us": [p * 1e3 for p in d["pass_ms"]],
                              "line_kernels": d["line_kernels"]}},
        "accuracy": acc,
        "batch": batch,
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu_1,
        "cpu_host": host,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

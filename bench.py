#!/usr/bin/env python3
"""Throughput of the evolver on the 256^3 Standard-Model workload.

One step = one leapfrog update of every site of the lattice (all fields).
The timed region runs exactly --steps steps (default 10^4) of the FP32-bulk
momentum-form kernel (FP32 arithmetic, phi/P/E stored as FP32 + 4-bit
residual) on data already resident in HBM, bracketed by stream
synchronisations; `value` is site-updates per second over that wall-clock
interval.  The Gauss-law violation of the metric's accuracy condition is
measured on a separate run of 10^4 steps (--gauss-steps) from the same initial
state, whatever --steps is.  Beside it:
the same kernel with plain FP32 storage, the FP64 momentum form (the plain
FP64 GPU evolver) and the FP64 two-slot reference-layout kernel on the same
state, the 64-member ensemble workload, the step kernel's roofline position
(HIP events on the library's own stream) and the reference CPU program timed
on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--n 256]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "site-updates/s + HBM GB/s, FP32-mixed vs FP64 at Gauss err<=1e-6, 256^3 SM"
BYTES_F32 = 280   # 35 reals read + 35 written per site-update, 4 B each (algorithmic)
STORED_F32C = 296  # what the compensated form moves: + one 64-bit word of 4-bit residuals of P and E each way
BYTES_F64 = 560
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gauss_rel(row):
    return max(row.gauss_int[m] / row.gauss_scale[m] for m in range(4) if row.gauss_scale[m] > 0)


def run_form(rt, desc, init, steps, warmup, launch=None, tune=False, occupancy=False):
    """Builds the initial state with init(ev), optionally autotunes the launch
    shape (untimed), runs `warmup` untimed steps, then times exactly `steps` steps.
    The timed steps continue the state the tuner advanced (a few hundred steps
    of the same lattice): starting again from t = 0 would put the timed steps
    right behind the initial-state kernels and a pause, and the first ~25
    steps after a pause run 3-10% slower while the GPU's clocks ramp up."""
    info = {}
    with rt.Evolver(desc) as ev:
        init(ev)
        e0 = ev.observe().energy[0] if desc.form >= 2 else None
        if launch:
            ev.set_launch(*launch)
            info["launch"] = list(launch)
        elif tune and desc.form >= 2:
            from gftevolver_amd.tune import tune as autotune
            best, timing = autotune(ev, fp32_sm=(desc.form in (2, 4) and desc.model == 0), occupancy=occupancy)
            info["launch"] = list(best)
            info["tune_ms"] = {",".join(map(str, k)): round(v, 4) for k, v in sorted(timing.items(), key=lambda kv: kv[1])[:6]}
        ev.evolve(warmup)
        ev.sync()
        t0 = time.perf_counter()
        ev.evolve(steps)
        ev.sync()
        wall = time.perf_counter() - t0
        avg_ms, _, nk = ev.last_step_ms()
        rows = ev.evolve(1, gauss_freq=1)
        e1 = ev.observe().energy[0] if desc.form >= 2 else None
    return {"wall_s": wall, "kernel_ms": avg_ms, "gauss_rel": gauss_rel(rows[0]), "after_steps": rows[0].time_iter + 1,
            "e0": e0, "e1": e1, **info}


def gauss_leg(rt, desc, init, launch, steps):
    """The headline's accuracy condition on its own run length: the same form
    and launch from the same initial state, `steps` steps (10^4 by default,
    whatever --steps times), then the relative Gauss-law violation and the
    energy drift.  Untimed; the step kernel's HIP-event average is reported."""
    with rt.Evolver(desc) as ev:
        init(ev)
        if launch:
            ev.set_launch(*launch)
        e0 = ev.observe().energy[0]
        t0 = time.perf_counter()
        ev.evolve(steps - 1)
        rows = ev.evolve(1, gauss_freq=1)
        wall = time.perf_counter() - t0
        kms = ev.last_step_ms()[0]
        e1 = ev.observe().energy[0]
    return {"steps": steps, "gauss_rel": gauss_rel(rows[0]), "energy_drift_rel": (e1 - e0) / abs(e0),
            "wall_s": wall, "kernel_ms_last_call": kms}


def measured_leg(rt, desc, init, launch, steps=1000, every=100, warmup=10):
    """The same step with the FP64 energy and Gauss-law reductions every `every`
    steps inside the timed region (the reference's analyser cadence)."""
    with rt.Evolver(desc) as ev:
        init(ev)
        if launch:
            ev.set_launch(*launch)
        ev.evolve(warmup)
        ev.sync()
        t0 = time.perf_counter()
        rows = ev.evolve(steps, energy_freq=every, gauss_freq=every)
        ev.sync()
        wall = time.perf_counter() - t0
    n = desc.nx * desc.ny * desc.nz
    return {"value": n * steps / wall, "unit": "site-updates/s", "ms_per_step": wall / steps * 1e3, "steps": steps,
            "energy_and_gauss_every": every, "measurements": len(rows)}


def reference_program_rate(model=0, n=32, nsteps=60, ns=4, nvd=5, pot2=None, gen=False):
    """Site-updates/s of the reference program (oracle/_ref/gft_reference, built
    from the reference's own sources) on a bounded sample, analysers off, one
    core, by its own 'Evolution finished in' timer; None if it is not built."""
    exe = os.path.join(ROOT, "oracle", "_ref", "gft_reference")
    if not os.path.exists(exe):
        return None
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcfg
    with tempfile.TemporaryDirectory() as td:
        refcfg.prepare(td)
        refcfg.write_model(td, model, 0.65, 0.35, lam=0.13, eta_sqr=1.0, pot2=pot2, gen=gen)
        refcfg.write_lattice(td, (n, n, n), (0.5, 0.5, 0.5), nsteps, 0.05, ns, energy=False, gauss=False,
                             progress=True, ic_range=(-0.1, 0.1), nvd=nvd)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=600, cwd=td)
    m = re.findall(r"Evolution finished in ([0-9.eE+-]+) seconds", r.stdout)
    if r.returncode != 0 or not m:
        return None
    sec = float(m[-1])
    return n ** 3 * nsteps / sec, sec


def cpu_baseline(nsteps=150):
    """The reference program on a bounded sample: SM 32^3, `nsteps` steps,
    analysers off, one core (the CPU restatement if the program is not built)."""
    n = 32
    ref = reference_program_rate(0, n, nsteps)
    if ref:
        rate, sec = ref
        return {"value": rate, "unit": "site-updates/s", "cores": 1, "kind": "reference",
                "sample": f"reference program built from its sources, SM {n}^3, {nsteps} steps, analysers off, "
                          f"single-threaded on one host core ({sec:.2f} s)"}
    # Fallback: the FP64 CPU restatement on one thread.
    import numpy as np
    from gftevolver_amd._abi import make_desc
    from gftevolver_amd.ic import gaussian_state
    from tests import oracle_lib as O
    d = make_desc(n=(n, n, n))
    s, v = gaussian_state((n, n, n))
    t0 = time.perf_counter()
    O.evolve(d, s, v, nsteps, nthreads=1)
    sec = time.perf_counter() - t0
    return {"value": n ** 3 * nsteps / sec, "unit": "site-updates/s", "cores": 1, "kind": "port",
            "sample": f"CPU restatement (FP64 two-slot), SM {n}^3, {nsteps} steps, one thread ({sec:.2f} s)"}


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_restatement_all_cores(n=128, nsteps=40):
    """The CPU restatement in the reference's own FP32 arithmetic (two-slot
    arrays, the reference's site order within each step) with its site loop
    spread over every host thread this process may use (OpenMP, static
    schedule): SM n^3, iid Gaussian state, `nsteps` steps, analysers off.
    Threads = the CPU affinity set, capped by OMP_NUM_THREADS when set."""
    import numpy as np
    from gftevolver_amd._abi import make_desc
    from gftevolver_amd.ic import gaussian_state
    from tests import oracle_lib as O
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    d = make_desc(n=(n, n, n))
    s, v = gaussian_state((n, n, n))
    s, v = s.astype(np.float32), v.astype(np.float32)
    O.evolve(d, s, v, 1, nthreads=threads)       # first touch of the pages, thread start-up
    t0 = time.perf_counter()
    O.evolve(d, s, v, nsteps, time_iter=1, nthreads=threads)
    sec = time.perf_counter() - t0
    return {"value": n ** 3 * nsteps / sec, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": host_cpu_model(),
            "sample": f"CPU restatement (FP32 two-slot, reference arithmetic), SM {n}^3, {nsteps} steps, "
                      f"OpenMP over {threads} threads of nproc={os.cpu_count()} ({sec:.2f} s)"}


def options_leg(rt, with_cpu):
    """The reference's other model options on the GPU, each beside the
    reference program on the host (one core, 32^3 sample): fixed faces in the
    fast path, the generator form (boundary conversion and the two-slot
    kernel), the 2HDM with SU(2) x U(1) links and with Higgs-family U(1) /
    SU(2) gauge fields in the reference-order two-slot form and in the fast
    path, the global O(4) model (two-slot form), and the 2HDM workload
    lattice (192^3) in the FP32 and FP64 momentum forms."""
    import numpy as np
    from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, GFT_FORM_TWO_SLOT_F32,
                                     GFT_MODEL_2HDM, GFT_MODEL_GLOBAL, TWO_HDM_POTENTIAL, make_desc)
    out = {}

    def timed(desc, steps, init=None, state=None):
        with rt.Evolver(desc) as ev:
            if init:
                init(ev)
            else:
                ev.import_two_slot(*state, 0)
            ev.evolve(2)
            ev.sync()
            t0 = time.perf_counter()
            ev.evolve(steps)
            ev.sync()
            wall = time.perf_counter() - t0
            kms = ev.last_step_ms()[0]
        N = desc.nx * desc.ny * desc.nz
        return {"value": N * steps / wall, "unit": "site-updates/s", "kernel_ms": kms, "steps": steps,
                "lattice": [desc.nx, desc.ny, desc.nz]}

    def grf(ev):
        ev.init_grf(seed=20261203, rms=0.1, slope=-1.0)

    n256 = (256,) * 3
    out["fixed_faces_xz"] = timed(make_desc(form=GFT_FORM_MOMENTUM_F32, n=n256, fixed_faces=1 | 2 | 16 | 32), 200,
                                  grf)
    # Generator form: the fast path converts at the boundary; the two-slot kernel steps it as the reference.
    n128 = (128,) * 3
    N = 128 ** 3
    rng = np.random.default_rng(1)
    s = np.zeros((2, N, 4), np.float32)
    s[:] = (rng.normal(size=(N, 4)) * 0.1).astype(np.float32)
    s[:, :, 2] += 1.0
    vg = np.zeros((2, N, 12), np.float32)
    dg = make_desc(form=GFT_FORM_MOMENTUM_F32, n=n128, gen_rep=1)
    with rt.Evolver(dg) as ev:
        t0 = time.perf_counter()
        ev.import_two_slot(s, vg, 0)
        t1 = time.perf_counter()
        ev.export_two_slot(np.float32)
        t2 = time.perf_counter()
    out["generator_form_boundary"] = {"lattice": list(n128), "import_s": t1 - t0, "export_s": t2 - t1}
    out["generator_form_two_slot_f32"] = timed(make_desc(form=GFT_FORM_TWO_SLOT_F32, n=n128, gen_rep=1), 50,
                                               state=(s, vg))
    vq = np.zeros((2, N, 15), np.float32)
    vq[..., 1::5] = 1.0
    out["quaternion_two_slot_f32"] = timed(make_desc(form=GFT_FORM_TWO_SLOT_F32, n=n128), 50, state=(s, vq))
    s8 = np.concatenate([s, s * 0.75], axis=-1)
    for g, nvd in ((4, 5), (5, 6), (7, 9)):
        v = np.zeros((2, N, 3, nvd), np.float32)
        v[..., 1] = 1.0
        if g == 7:
            v[..., 5] = 1.0
        d = make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_TWO_SLOT_F32, n=n128, pot2=TWO_HDM_POTENTIAL,
                      gauge_num=g, gpp=0.5)
        r = timed(d, 50, state=(s8, v.reshape(2, N, 3 * nvd)))
        if with_cpu:
            ref = reference_program_rate(1, 32, 10, 8, nvd, TWO_HDM_POTENTIAL)
            r["cpu_reference"] = ref[0] if ref else None
        out[f"2hdm_gauge{g}_two_slot_f32"] = r
        dm = make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_MOMENTUM_F32, n=n128, pot2=TWO_HDM_POTENTIAL,
                       gauge_num=g, gpp=0.5)
        out[f"2hdm_gauge{g}_momentum_f32"] = timed(dm, 200, state=(s8, v.reshape(2, N, 3 * nvd)))
    dgl = make_desc(model=GFT_MODEL_GLOBAL, form=GFT_FORM_TWO_SLOT_F32, n=n256, num_scalar=4, potential=1)
    Ng = 256 ** 3
    sg = (np.random.default_rng(2).normal(size=(2, Ng, 4)) * 0.1 + np.array([1.0, 0, 0, 0])).astype(np.float32)
    sg[0] = sg[1]
    r = timed(dgl, 100, state=(sg, np.zeros(0, np.float32)))
    if with_cpu:
        ref = reference_program_rate(2, 32, 60, 4, 1)
        r["cpu_reference"] = ref[0] if ref else None
    out["global_o4_two_slot_f32"] = r
    # The 2HDM workload lattice: 192^3, the two-Higgs-doublet potential, both
    # doublets GRFs of rms 0.1 around vevs 0.8 and 0.6; FP32-bulk against FP64
    # (accuracy over its 5000 steps: tests/test_gpu_workloads.py).
    n192 = (192,) * 3

    def grf2(ev):
        ev.init_grf(seed=20261204, rms=0.1, slope=-1.0, vevs=(0.8, 0.6))

    r32 = timed(make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_MOMENTUM_F32, n=n192, pot2=TWO_HDM_POTENTIAL), 200,
                grf2)
    r64 = timed(make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_MOMENTUM_F64, n=n192, pot2=TWO_HDM_POTENTIAL), 20,
                grf2)
    # 43 FP32 reals per site read and written once (two doublets' phi and P,
    # the links and their electric fields).
    r32["achieved_GBps"] = 344 * 192 ** 3 / (r32["kernel_ms"] * 1e-3) / 1e9
    out["2hdm_192_momentum_f32"] = r32
    out["2hdm_192_momentum_f64"] = r64
    out["2hdm_192_speedup_fp32_vs_fp64"] = r64["kernel_ms"] / r32["kernel_ms"]
    for k, v in out.items():
        log(f"option {k}: {v}")
    return out


def ensemble_leg(rt, steps, tiles=None):
    """The 64-member ensemble workload: aggregate FP32 throughput of the
    graph-replayed launch per step, and its ensemble-mean observables against
    the same members run one by one in FP64."""
    import numpy as np
    from gftevolver_amd import ensemble as E
    specs = E.members()
    sites = sum(int(np.prod(n)) for n, _, _ in specs)
    every = max(steps // 4, 1)
    o32, sec, kms, tiles = E.run_fp32(rt, specs, steps, every, log=log, tiles=tiles)
    value = sites * steps / sec
    log(f"ensemble: {len(specs)} members, {sites} sites, {steps} steps in {sec:.3f} s = {value:.3e} site-updates/s "
        f"({kms:.3f} ms per step by HIP events)")
    o64 = E.run_fp64_sequential(rt, specs, steps, every)
    m32, m64 = o32.mean(axis=1), o64.mean(axis=1)
    diff = np.abs(m32 - m64).max(axis=0)
    log(f"ensemble means at step {steps}: FP32 {m32[-1]}, FP64 {m64[-1]}, max |diff| over time {diff}")
    return {"value": value, "unit": "site-updates/s", "members": len(specs), "sites": sites, "steps": steps,
            "kernel_ms": kms, "achieved_GBps": BYTES_F32 * sites / (kms * 1e-3) / 1e9,
            "tiles_by_size": {str(k): list(v) for k, v in zip((32, 48, 64, 96), (tiles or {}).values())},
            "observables": list(E.OBSERVABLES),
            "mean_fp32_final": m32[-1].tolist(), "mean_fp64_final": m64[-1].tolist(),
            "max_abs_mean_diff_vs_fp64": diff.tolist(), "measured_every": every}


def read_traffic():
    """HBM bytes per step-kernel launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size)."""
    p = os.path.join(ROOT, "profiles", "mom_step_f32_256_pmc.txt")
    if not os.path.exists(p):
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(p).read())
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--gauss-steps", type=int, default=10000,
                    help="run length of the Gauss-law leg (the headline's accuracy condition); 0 skips it")
    ap.add_argument("--plain-steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--fp64-steps", type=int, default=100)
    ap.add_argument("--twoslot-steps", type=int, default=20)
    ap.add_argument("--ensemble-steps", type=int, default=2000,
                    help="steps of the 64-member ensemble workload (0 skips it)")
    ap.add_argument("--ensemble-tiles", type=str, default="",
                    help="bz x by per member size 32,48,64,96, e.g. 32x8,16x16,32x8,32x8 (skips the ensemble tuner)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-options", action="store_true",
                    help="skip the other model options (fixed faces, generator form, gaugeNum 5/7, O(N))")
    ap.add_argument("--only-f32", action="store_true")
    ap.add_argument("--launch", type=str, default="", help="bz,by,order,min_waves[,stage] (skips the autotuner)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--occupancy-variants", action="store_true",
                    help="let the tuner try the 3/4-wave kernels (last-bit different results)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device (a kernel of known traffic for PMC scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library runs every workload on a single MI355X "
              f"(--gpus {args.gpus} is not supported).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from gftevolver_amd import runtime as rt
    from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN, GFT_FORM_MOMENTUM_F64,
                                     GFT_FORM_TWO_SLOT_F64, make_desc)

    name, cus = rt.device_info()
    copy_gbps = None
    if args.calibrate_copy or not args.only_f32:
        # A plain copy of 1 GiB: its achieved bandwidth is the practical HBM
        # ceiling, and its PMC traffic calibrates FETCH_SIZE / WRITE_SIZE.
        cms = rt.bench_copy(1 << 30, 5)
        copy_gbps = 2 * (1 << 30) / (cms * 1e-3) / 1e9
        log(f"copy kernel: 1 GiB in {cms:.3f} ms = {copy_gbps:.0f} GB/s (read + write)")
    n = (args.n,) * 3
    N = args.n ** 3
    log(f"device {name}, {cus} CUs; GRF initial state {n} generated on the GPU")

    def init(ev):
        ev.init_grf(seed=20261203, rms=0.1, slope=-1.0)
    launch = tuple(int(x) for x in args.launch.split(",")) if args.launch else None

    d32 = make_desc(form=GFT_FORM_MOMENTUM_F32, n=n)
    r32 = run_form(rt, d32, init, args.steps, args.warmup, launch, tune=not args.no_tune,
                   occupancy=args.occupancy_variants)
    log(f"FP32 launch {r32.get('launch')}; tuner {r32.get('tune_ms')}")
    value = N * args.steps / r32["wall_s"]
    achieved = BYTES_F32 * N / (r32["kernel_ms"] * 1e-3) / 1e9
    log(f"FP32 momentum: {r32['wall_s'] / args.steps * 1e3:.3f} ms/step wall, kernel {r32['kernel_ms']:.3f} ms, "
        f"{value:.3e} site-updates/s, {achieved:.0f} GB/s algorithmic, Gauss rel {r32['gauss_rel']:.2e} "
        f"after {r32['after_steps']} steps")

    gl = None
    if args.gauss_steps:
        gl = gauss_leg(rt, d32, init, r32.get("launch"), args.gauss_steps)
        log(f"Gauss leg: {gl['steps']} steps from t = 0, relative Gauss violation {gl['gauss_rel']:.2e}, "
            f"energy drift {gl['energy_drift_rel']:.2e} ({gl['wall_s']:.1f} s)")

    stream_ms = None
    if args.calibrate_copy or not args.only_f32:
        # The HBM floor of this data set: the same arrays streamed once with
        # trivial arithmetic (gft_bench_streams), on a context of its own.
        with rt.Evolver(d32) as ev:
            if r32.get("launch"):
                ev.set_launch(*r32["launch"])
            stream_ms = ev.bench_streams(20)
        log(f"stream floor of the step's arrays: {stream_ms:.3f} ms = "
            f"{STORED_F32C * N / (stream_ms * 1e-3) / 1e9:.0f} GB/s; step kernel at "
            f"{stream_ms / r32['kernel_ms'] * 100:.1f}% of it")

    extra = {}
    if not args.only_f32:
        wm = measured_leg(rt, d32, init, r32.get("launch"))
        log(f"FP32 momentum with energy and Gauss every {wm['energy_and_gauss_every']} steps: "
            f"{wm['ms_per_step']:.3f} ms/step wall, {wm['value']:.3e} site-updates/s")
        extra["with_measurements"] = wm
        dpl = make_desc(form=GFT_FORM_MOMENTUM_F32_PLAIN, n=n)
        rpl = run_form(rt, dpl, init, args.plain_steps, args.warmup, launch, tune=not args.no_tune)
        vpl = N * args.plain_steps / rpl["wall_s"]
        log(f"FP32 plain storage: kernel {rpl['kernel_ms']:.3f} ms, {vpl:.3e} site-updates/s, "
            f"Gauss rel {rpl['gauss_rel']:.2e} after {rpl['after_steps']} steps")
        extra["fp32_plain_storage"] = {"value": vpl, "kernel_ms": rpl["kernel_ms"], "steps": args.plain_steps,
                                       "achieved_GBps": BYTES_F32 * N / (rpl["kernel_ms"] * 1e-3) / 1e9,
                                       "gauss_rel": rpl["gauss_rel"], "launch": rpl.get("launch")}
        d64 = make_desc(form=GFT_FORM_MOMENTUM_F64, n=n)
        r64 = run_form(rt, d64, init, args.fp64_steps, 2, tune=not args.no_tune)
        v64 = N * args.fp64_steps / r64["wall_s"]
        log(f"FP64 momentum: kernel {r64['kernel_ms']:.3f} ms, {v64:.3e} site-updates/s, "
            f"energy drift {(r64['e1'] - r64['e0']) / abs(r64['e0']):.2e}")
        dts = make_desc(form=GFT_FORM_TWO_SLOT_F64, n=n)
        rts = run_form(rt, dts, init, args.twoslot_steps, 2)
        vts = N * args.twoslot_steps / rts["wall_s"]
        log(f"FP64 two-slot: kernel {rts['kernel_ms']:.3f} ms, {vts:.3e} site-updates/s")
        extra.update({
            "fp64_momentum": {"value": v64, "kernel_ms": r64["kernel_ms"],
                              "achieved_GBps": BYTES_F64 * N / (r64["kernel_ms"] * 1e-3) / 1e9,
                              "gauss_rel": r64["gauss_rel"],
                              "energy_drift_rel": (r64["e1"] - r64["e0"]) / abs(r64["e0"])},
            "fp64_two_slot": {"value": vts, "kernel_ms": rts["kernel_ms"], "gauss_rel": rts["gauss_rel"]},
            "speedup_fp32_vs_fp64_momentum": value / v64,
            "speedup_fp32_vs_fp64_two_slot": value / vts,
        })

    if not args.only_f32 and args.ensemble_steps:
        tiles = None
        if args.ensemble_tiles:
            shapes = [tuple(int(v) for v in t.split("x")) for t in args.ensemble_tiles.split(",")]
            tiles = dict(zip((32, 48, 64, 96), shapes))
        extra["ensemble"] = ensemble_leg(rt, args.ensemble_steps, tiles)
    if not args.only_f32 and not args.no_options:
        extra["options"] = options_leg(rt, not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()
        cpu["all_cores"] = cpu_restatement_all_cores()
        print(f"cpu restatement over all cores: {cpu['all_cores']['value']:.3e} site-updates/s "
              f"({cpu['all_cores']['cores']} threads)", file=sys.stderr, flush=True)
    traffic = read_traffic()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall_s"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"SM {args.n}^3 periodic, FP32-bulk momentum-form leapfrog (FP32 arithmetic, phi/P/E "
                        f"stored FP32 + 4-bit residual); GRF Higgs IC "
                        f"(P(k)~1/k, k<=pi/(2dx), rms 0.1v), links identity, E=0; dx=0.5 dt=0.05 g=0.65 "
                        f"g'=0.35 lambda=0.13 v=1",
            "lattice": list(n),
            "launch": r32.get("launch", [64, 4, 2, 4, 2]),
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_site_update": BYTES_F32,
                     "stored_bytes_per_site_update": STORED_F32C,
                     "kernel_ms": r32["kernel_ms"], "copy_kernel_GBps": copy_gbps,
                     "stream_floor_ms": stream_ms,
                     "stream_floor_GBps": STORED_F32C * N / (stream_ms * 1e-3) / 1e9 if stream_ms else None},
        "cpu_baseline": cpu,
        # The accuracy condition of the metric, on the 10^4-step run of the Gauss
        # leg (the timed run's own end state when the leg is skipped).
        "gauss_rel_violation": gl["gauss_rel"] if gl else r32["gauss_rel"],
        "gauss_measured_after_steps": gl["steps"] if gl else r32["after_steps"],
        "energy_drift_rel": gl["energy_drift_rel"] if gl else (r32["e1"] - r32["e0"]) / abs(r32["e0"]),
        "timed_run_end": {"gauss_rel": r32["gauss_rel"], "after_steps": r32["after_steps"],
                          "energy_drift_rel": (r32["e1"] - r32["e0"]) / abs(r32["e0"])},
        **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

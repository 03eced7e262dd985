"""Benchmark of the vertical-mixing column step on one MI355X.

One step = one pass of the fused plain FP64 path (Wright EOS, PP81, KPP,
split, assemble, surface deposit, Thomas solve of T/S/passives, non-local
transport) over every wet column of the 1440 x 1080 x 75 z-level tile, with
the inputs resident in HBM.  The timed region is exactly ``--steps`` steps
replayed from one captured hipGraph, bracketed by stream synchronisation.

Prints ONE JSON line (the contract): value = wet column-steps per second.
``roofline`` uses the HIP-event duration of the step kernel (events recorded
on the library's stream inside the graph) and the algorithmic bytes
B_col = 8 (11 L + 3 (L + 1)) + 44 per wet column; ``cpu_baseline`` times the
CPU oracle (an op-for-op C restatement, OpenMP over columns, built -O3
-march=native for the host) on a bounded crop of the same workload, with all
threads and one, and on the 64 x 64 x 20 CPU case in full.  The learned
closure's T/S error is judged after 24 one-hour steps whatever --steps is.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

METRIC = "column-steps/s surrogate vs plain FP64 vmix+vdiff at \u22640.01 K T error; HBM GB/s"
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0  # dense BF16 MFMA
DT = 3600.0
ACCURACY_STEPS = 24  # the surrogate's T/S error is judged after 24 one-hour steps
TRAFFIC_FILE = ROOT / "profiles" / "traffic_colphys_step_tile.txt"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def read_traffic(workload: str, path: Path | None = None):
    """HBM bytes per step launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, each scaled by the copy-kernel calibration), if recorded."""
    path = path or TRAFFIC_FILE
    if not path.exists():
        return None
    vals = {}
    for line in path.read_text().splitlines():
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    if vals.get("workload") != workload:
        return None
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def host_cpu() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "nproc": os.cpu_count(), "cpus_allowed": aff}


def _time_oracle(O, ocean, max_steps: int, budget_s: float):
    """Oracle column steps on `ocean` until max_steps or budget_s (at least
    one step): (wet column-steps/s, steps, seconds)."""
    st = O.OracleState(ocean)
    steps = 0
    t0 = time.perf_counter()
    while True:
        st.step(DT, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= max_steps:
            break
    return ocean.n_wet * steps / el, steps, el


def cpu_baseline(ocean, budget_s: float, crop_cols: tuple[int, int]):
    """The CPU oracle (C, OpenMP over columns) built here with -O3
    -march=native -ffp-contract=off for this host, after checking that it
    keeps the bits of the default build; timed on the 64x64x20 CPU case in
    full (24 steps) and on a corner crop of the tile, with every allowed
    thread and with one."""
    from oracle import oracle as O
    from roundabout_amd import synth
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", cpu["cpus_allowed"]))
    native = O.build_native(tempfile.mkdtemp(prefix="rdb_oracle_native_"))
    small = synth.small(1)
    bits = []
    for path in (None, native):
        O.use_library(path)
        st = O.OracleState(small)
        st.step(DT, 2)
        bits.append((st.ocean.T.copy(), st.kt.copy()))
    same_bits = bool(np.array_equal(bits[0][0], bits[1][0]) and np.array_equal(bits[0][1], bits[1][1]))
    O.use_library(native)
    ci, cj = crop_cols
    sub = synth.crop(ocean, 0, min(ci, ocean.nx), 0, min(cj, ocean.ny))
    sub1 = synth.crop(ocean, 0, min(ci // 2, ocean.nx), 0, min(cj // 2, ocean.ny))
    O.set_threads(threads)
    used = O.num_threads()
    v_small, _, t_small = _time_oracle(O, small, 24, 1e9)
    v_tile, n_tile, t_tile = _time_oracle(O, sub, 24, budget_s)
    O.set_threads(1)
    v_small1, _, t_small1 = _time_oracle(O, small, 24, 1e9)
    v_tile1, n_tile1, t_tile1 = _time_oracle(O, sub1, 24, budget_s / 2)
    O.use_library(None)
    return {
        "value": v_tile,
        "unit": "column-steps/s",
        "cores": used,
        "kind": "port",
        "sample": f"oracle (C restatement of the reference path, gcc -O3 -march=native -ffp-contract=off, "
                  f"OpenMP over columns, {used} threads) on the {sub.nx}x{sub.ny}x{sub.nz} corner crop of the "
                  f"same tile ({sub.n_wet} wet columns), {n_tile} steps of dt={DT:.0f} s in {t_tile:.1f} s",
        "cpu_model": cpu["model"],
        "nproc": cpu["nproc"],
        "cpus_allowed": cpu["cpus_allowed"],
        "omp_num_threads": used,
        "one_thread_value": v_tile1,
        "one_thread_sample": f"{sub1.nx}x{sub1.ny}x{sub1.nz} corner crop ({sub1.n_wet} wet columns), "
                             f"{n_tile1} steps in {t_tile1:.1f} s",
        "cpu_case_64x64x20": {"column_steps_per_s": v_small, "seconds_24_steps": t_small,
                              "one_thread_column_steps_per_s": v_small1, "one_thread_seconds_24_steps": t_small1,
                              "wet_columns": small.n_wet},
        "native_build_bitwise_equal_to_default_build": same_bits,
    }


def momentum_leg(cp, ocean, reps: int = 10) -> dict:
    """The momentum vertical viscosity (rdb_colphys_vdiff_momentum, the
    closure's kv, ocean_vdiff_t defaults) on the same tile: wall time per
    call over `reps` calls, each including its natural-order copies of h and
    kv and the refresh of the centred velocities."""
    cp.vdiff_momentum(DT)  # first call: allocates the face buffers
    cp.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        cp.vdiff_momentum(DT)
    cp.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    nx, ny, nz = ocean.nx, ocean.ny, ocean.nz
    faces = (nx + 1) * ny + nx * (ny + 1)
    # per face and level: u read and written, h and kv of the two cells
    alg = faces * nz * 6 * 8
    return {"ms_per_call": ms, "face_columns": faces, "nz": nz,
            "face_column_steps_per_s": faces / (ms * 1e-3),
            "algorithmic_GBs": alg / (ms * 1e-3) / 1e9,
            "config": "kv of the plain closure, ocean_vdiff_t defaults (no folds)"}


def remap_leg(ocean, seed: int, reps: int = 5, cpu: bool = True) -> dict:
    """The conservative vertical (ALE) remap on the same tile: T and S
    remapped from the Lagrangian thicknesses after a step onto the z*
    target of the same column totals (synth.ale_pair), PPM with the default
    knobs and the heat/salt remap budget slots, as ocean_apply_ale_remap_centres
    runs it; then the face-velocity pass.  Kernel time from HIP events on the
    null stream the kernels run on, averaged over `reps` calls after one
    untimed call."""
    from roundabout_amd import synth
    from roundabout_amd.rdb import remap as R

    t0 = time.perf_counter()
    ho, hn = synth.ale_pair(ocean, seed)
    d_ho, d_hn = R.DeviceArray.from_host(ho), R.DeviceArray.from_host(hn)
    log(f"[bench] remap: ALE pair built and uploaded in {time.perf_counter() - t0:.1f} s")
    nx, ny, nz = ocean.nx, ocean.ny, ocean.nz
    cells = nx * ny * nz
    timer = R.Timer()

    def timed(fn):
        fn()
        R.synchronize()
        timer.start()
        for _ in range(reps):
            fn()
        return timer.stop_ms() / reps

    fields = {}
    for name, fld in (("T", ocean.T), ("S", ocean.S)):
        d_c = R.DeviceArray.from_host(fld)
        d_o, d_b = R.DeviceArray(d_c.shape), R.DeviceArray.zeros(d_c.shape)
        ms_out = timed(lambda: R.ocean_remap_tracer_field(d_ho, d_hn, d_c, R.REMAP_PPM, budget=d_b, c_out=d_o))
        ms_in = timed(lambda: R.ocean_remap_tracer_field(d_ho, d_hn, d_c, R.REMAP_PPM, budget=d_b))
        fields[name] = {"kernel_ms": ms_out, "in_place_ms": ms_in}
        for d in (d_c, d_o, d_b):
            d.free()
    # per cell: h_old, h_new and c read, c_new written, the budget slot read and written
    bytes_field = cells * 6 * 8
    ms_field = float(np.mean([f["kernel_ms"] for f in fields.values()]))
    achieved = bytes_field / (ms_field * 1e-3) / 1e9
    rng = np.random.default_rng(seed)
    d_u = R.DeviceArray.from_host(rng.normal(0.0, 0.1, (nx + 1, ny, nz)))
    d_v = R.DeviceArray.from_host(rng.normal(0.0, 0.1, (nx, ny + 1, nz)))
    ms_faces = timed(lambda: R.ocean_apply_ale_remap_faces(d_ho, d_hn, d_u, d_v, R.REMAP_PPM))
    face_cells = ((nx + 1) * ny + nx * (ny + 1)) * nz
    n_bad, worst_rel, worst_neg = R.ocean_remap_scan_preconditions(d_ho, d_hn)
    for d in (d_u, d_v, d_ho, d_hn):
        d.free()
    out = {
        "workload": f"{nx}x{ny}x{nz} tile, T and S, z* target of a {1.7:.1f} m free-surface anomaly over "
                    f"5 %-jittered Lagrangian thicknesses; PPM, default knobs, budget slots on",
        "kernel": "k_walk<CellCols, PPM, field> (one lane per column, register window merge walk)",
        "kernel_ms_per_field": ms_field,
        "fields": fields,
        "column_remaps_per_s": nx * ny / (ms_field * 1e-3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "algorithmic_bytes_per_launch": bytes_field,
                     "bytes_per_cell": 48,
                     "traffic": read_traffic(f"tile {nx}x{ny}x{nz}", TRAFFIC_FILE.parent / "traffic_remap_field_tile.txt")},
        "faces_ms": ms_faces,
        "face_cells": face_cells,
        "preconditions": {"n_bad": n_bad, "worst_rel": worst_rel, "worst_neg": worst_neg},
    }
    if cpu:
        from oracle import oracle as O
        sub = synth.crop(ocean, 0, min(360, nx), 0, min(270, ny))
        sho, shn = ho[:sub.nx, :sub.ny].copy(order="F"), hn[:sub.nx, :sub.ny].copy(order="F")
        O.set_threads(int(os.environ.get("OMP_NUM_THREADS", host_cpu()["cpus_allowed"])))
        O.remap_tracer_field(R.REMAP_PPM, sho, shn, sub.T)
        n, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < 3.0:
            O.remap_tracer_field(R.REMAP_PPM, sho, shn, sub.T, budget=np.zeros_like(sho, order="F"))
            n += 1
        dt = (time.perf_counter() - t1) / n
        out["cpu_baseline"] = {"value": sub.nx * sub.ny / dt, "unit": "column-remaps/s", "cores": O.num_threads(),
                               "kind": "port",
                               "sample": f"oracle remap_tracer_field (PPM, budget) on the {sub.nx}x{sub.ny}x{nz} "
                                         f"corner crop, {n} calls in {dt * n:.1f} s"}
    return out


CONTRIB_KNOBS = dict(epbl_enable=1, epbl_tke_diags=1, conv_enable=1, ddiff_enable=1, kv_ml_invz2=1.0e-2)


def contrib_leg(ocean, steps: int = 6, cpu: bool = True) -> dict:
    """The contributors off at the reference defaults, all on, on the same
    tile: EPBL (OM4 mstar, cube-root v*, MLD iteration, additive merge, TKE
    ledger kept) in place of KPP, KV_ML_INVZ2, convective adjustment and
    double diffusion.  The step runs the staged kernels; ms per step from
    the handle's HIP events around `steps` steps after one untimed step, and
    the coefficient build (PP81 + EPBL + merge/invz2/convection + split with
    double diffusion + assemble) timed alone the same way."""
    from roundabout_amd.rdb import ColumnPhysics, default_config

    cfg = default_config()
    for k, v in CONTRIB_KNOBS.items():
        setattr(cfg, k, v)
    nx, ny, nz = ocean.nx, ocean.ny, ocean.nz
    with ColumnPhysics.from_ocean(ocean, config=cfg) as cp:
        cp.step(DT, 1)
        cp.synchronize()
        cp.set_timing(True)
        cp.step(DT, steps)
        stats = cp.stats()
        ms_step = stats["sum_kernel_ms"] / max(stats["timed_launches"], 1)
        cp.set_timing(False)
        cp.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            cp.coefficients(DT)
        cp.synchronize()
        ms_coef = (time.perf_counter() - t0) * 1e3 / steps
        mld = cp.get_field("mld")
        wet = ocean.wet > 0.5
        n_wet = int(wet.sum())
        out = {
            "workload": f"{nx}x{ny}x{nz} tile, EPBL (OM4, cube-root v*, MLD iteration, add) + KV_ML_INVZ2 "
                        f"(1e-2 m^2/s, 20 m) + convective adjustment + double diffusion, dt={DT:.0f} s",
            "ms_per_step": ms_step,
            "column_steps_per_s": n_wet / (ms_step * 1e-3),
            "coefficients_ms": ms_coef,
            "mld_m": {"p5": float(np.percentile(mld[wet], 5)), "median": float(np.median(mld[wet])),
                      "p95": float(np.percentile(mld[wet], 95))},
            "kernels": "k_eos, k_epbl, k_coeff_tail (PP81 + merge + KV_ML_INVZ2 + convection + split with "
                       "double diffusion + assemble), k_surface_flux, k_vdiff (staged, one lane per column)",
        }
    if cpu:
        from oracle import oracle as O
        from roundabout_amd import synth
        sub = synth.crop(ocean, 0, min(240, nx), 0, min(180, ny))
        O.set_threads(int(os.environ.get("OMP_NUM_THREADS", host_cpu()["cpus_allowed"])))
        st = O.OracleState(sub, cfg)
        st.step(DT, 1)
        n, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < 5.0:
            st.step(DT, 1)
            n += 1
        dt = (time.perf_counter() - t1) / n
        out["cpu_baseline"] = {"value": sub.n_wet / dt, "unit": "column-steps/s", "cores": O.num_threads(),
                               "kind": "port",
                               "sample": f"oracle column step with the same contributors on the {sub.nx}x{sub.ny}x"
                                         f"{nz} corner crop, {n} steps in {dt * n:.1f} s"}
    return out


def live_rms(ocean, a, b) -> float:
    live = (ocean.h > 1.5e-4) & (ocean.wet > 0.5)[:, :, None]
    d = (a - b)[live]
    return float(np.sqrt(np.mean(d * d)))


def parse_schedule(text: str) -> list[tuple[int, float]]:
    """'180@1e-3,60@3e-4' -> [(180, 1e-3), (60, 3e-4)]: epochs at a learning rate, in order."""
    out = []
    for part in text.split(","):
        n, lr = part.split("@")
        out.append((int(n), float(lr)))
    if not out or any(n < 1 or lr <= 0 for n, lr in out):
        raise SystemExit(f"bad --lr-schedule {text!r}")
    return out


def train_schedule(cp, schedule: list[tuple[int, float]], guard: int) -> tuple[list, list]:
    """Train through `schedule` (epochs at a learning rate, in order), one
    mlp_train call per stage: a call keeps the network and its Adam moments,
    so the stages continue one training (each call draws the same held-out
    split and restarts the epoch shuffles from the configured seed)."""
    tl, vl, done = [], [], 0
    total = sum(n for n, _ in schedule)
    for n, lr in schedule:
        a, b = cp.mlp_train(epochs=n, guard_mode=guard, lr=lr)
        tl += a
        vl += b
        done += n
        log(f"[bench] surrogate: {done}/{total} epochs (lr {lr:g}), val loss {vl[-1]:.5f}")
    return tl, vl


def surrogate_leg(ocean, cfg, args, plain_value: float) -> dict:
    """The learned closure on the same tile: train on the GPU on columns the
    plain path generates from an independent state draw on the same grid,
    then run K guarded surrogate steps from the benchmark's initial state and
    compare them with K plain steps from that state."""
    from roundabout_amd import synth
    from roundabout_amd.rdb import CLOSURE_SURROGATE, GUARD_RANGE, GUARD_STRICT, GUARD_SURFACE, ColumnPhysics

    guard = {"strict": GUARD_STRICT, "surface": GUARD_SURFACE, "range": GUARD_RANGE}[args.guard]
    K = ACCURACY_STEPS  # the error horizon is fixed, whatever --steps times for the plain path
    schedule = parse_schedule(args.lr_schedule)
    epochs = sum(n for n, _ in schedule)
    net_path = Path(tempfile.gettempdir()) / f"rdb_bench_closure_{os.getpid()}.bin"
    t0 = time.perf_counter()
    with ColumnPhysics.from_ocean(synth.redraw(ocean, args.seed + 1000), config=cfg) as cp:
        for s in range(args.train_states):
            if s:
                cp.upload(synth.redraw(ocean, args.seed + 1000 + s))
            n = cp.mlp_collect(DT, args.train_steps, args.train_stride)
            log(f"[bench] surrogate: training state {s}: {n} columns at {time.perf_counter() - t0:.0f} s")
        cp.synchronize()
        t_col = time.perf_counter() - t0
        t1 = time.perf_counter()
        tl, vl = train_schedule(cp, schedule, guard)
        cp.synchronize()
        t_train = time.perf_counter() - t1
        cp.mlp_save(net_path)
    log(f"[bench] surrogate: {n} training columns in {t_col:.1f} s, {epochs} epochs in {t_train:.1f} s, "
        f"val loss {vl[0]:.4f} -> {vl[-1]:.4f}")
    with ColumnPhysics.from_ocean(ocean, config=cfg) as cp:
        cp.step(DT, K)
        T_ref, S_ref = cp.get_field("temperature"), cp.get_field("salinity")
    with ColumnPhysics.from_ocean(ocean, config=cfg) as cp:
        cp.mlp_load(net_path)
        cp.set_guard(guard)
        cp.set_closure(CLOSURE_SURROGATE)
        cp.step(DT, 1)  # warm-up: workspace and first-touch allocations
        cp.upload(ocean)
        h0, s0 = cp.heat_salt_totals()
        cp.set_timing(True)
        cp.synchronize()
        t2 = time.perf_counter()
        cp.step(DT, K)
        cp.synchronize()
        wall = time.perf_counter() - t2
        st = cp.stats()
        h1, s1 = cp.heat_salt_totals()
        T, S = cp.get_field("temperature"), cp.get_field("salinity")
    n_wet = int(st["wet_columns"])
    dep = K * DT / (cfg.rho0 * cfg.cp) * float((ocean.Q_heat * ocean.wet).sum())
    value = n_wet * K / wall
    t_rms, s_rms = live_rms(ocean, T, T_ref), live_rms(ocean, S, S_ref)
    mlp_ms = st["mlp_forward_ms"]
    tflops = st["mlp_flop_per_step"] / (mlp_ms * 1e-3) / 1e12 if mlp_ms > 0 else None
    # training work: per epoch, forward + backward (about twice the forward)
    # over the training columns and a forward over the held-out ones
    fwd_per_col = st["mlp_flop_per_step"] / max(st["wet_columns"], 1)
    n_val = round(0.05 * n)
    train_tflops = epochs * fwd_per_col * (3 * (n - n_val) + n_val) / t_train / 1e12
    return {
        "value": value,
        "unit": "column-steps/s",
        "steps": K,
        "error_horizon": f"{K} steps of dt={DT:.0f} s from the benchmark's initial state",
        "ms_per_step": wall * 1e3 / K,
        "speedup_vs_plain": value / plain_value,
        "T_rms_K": t_rms,
        "S_rms_psu": s_rms,
        "meets_T_error_bar": bool(t_rms <= 0.01),
        "meets_error_bar": bool(t_rms <= 0.01 and s_rms <= 0.002),
        "heat_rel_residual": abs((h1 - h0) - dep) / abs(h0),
        "salt_rel_residual": abs(s1 - s0) / abs(s0),
        "fallback_fraction_last_step": st["fallback_fraction"],
        "graph_captured": bool(st["graph_replay"]),
        "guard": args.guard,
        "training": {"columns": n, "tflops": train_tflops, "mfma_frac": train_tflops / PEAK_BF16_TFLOPS,
                     "source": f"{args.train_steps} plain steps of each of {args.train_states} "
                     f"independent state draws (seeds {args.seed + 1000}..{args.seed + 999 + args.train_states}) "
                     f"on the same grid, every {args.train_stride}. column of each step",
                     "epochs": epochs, "batch": 8192,
                     "lr_schedule": [{"epochs": n, "lr": lr} for n, lr in schedule], "seconds": t_train,
                     "val_loss_first": vl[0], "val_loss_last": vl[-1]},
        "mlp": {"forward_ms": mlp_ms, "flop_per_step": st["mlp_flop_per_step"], "achieved_tflops": tflops,
                "peak_tflops": PEAK_BF16_TFLOPS, "mfma_frac": (tflops / PEAK_BF16_TFLOPS) if tflops else None,
                "widths": "381-256-256-256-256-149"},
    }


def ensemble_sizes(n: int = 16):
    """Member grids from 256 x 256 x 50 to 1440 x 1080 x 75, evenly spaced."""
    out = []
    for i in range(n):
        f = i / max(n - 1, 1)
        out.append((int(round(256 + (1440 - 256) * f)), int(round(256 + (1080 - 256) * f)),
                    int(round(50 + 25 * f))))
    return out


def _member(args):
    from roundabout_amd import synth
    (nx, ny, nz), seed, redraw_seed = args
    o = synth.tile(seed, nx=nx, ny=ny, nz=nz)
    return synth.redraw(o, redraw_seed) if redraw_seed is not None else o


def ensemble_bench(args) -> dict:
    """The batched ensemble: 16 members of uneven grids, each seeded separately, stepped
    by one persistent graph-captured launch per step; then the guarded
    surrogate over the same ensemble, compared with the all-plain run."""
    from concurrent.futures import ProcessPoolExecutor

    from roundabout_amd.rdb import (CLOSURE_SURROGATE, GUARD_RANGE, GUARD_STRICT, GUARD_SURFACE, ColumnEnsemble,
                                    default_config)
    sizes = ensemble_sizes(16)
    K = args.steps
    workers = max(1, min(8, os.cpu_count() or 1))
    t0 = time.perf_counter()
    def generate(jobs, what):
        out = []
        with ProcessPoolExecutor(workers) as pool:
            for i, o in enumerate(pool.map(_member, jobs)):
                out.append(o)
                log(f"[bench] ensemble: {what} member {i} {o.nx}x{o.ny}x{o.nz} ({o.n_wet} wet) "
                    f"at {time.perf_counter() - t0:.0f} s")
        return out

    oceans = generate([(sz, args.seed + 100 * i, None) for i, sz in enumerate(sizes)], "generated")
    cfg = default_config()
    n_wet = sum(o.n_wet for o in oceans)
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        log(f"[bench] ensemble: {n_wet} wet columns uploaded at {time.perf_counter() - t0:.0f} s")
        ens.autotune(DT)
        ens.step(DT, K)  # builds and instantiates the graph
        ens.upload(oceans)
        ens.set_timing(True)
        ens.synchronize()
        t1 = time.perf_counter()
        ens.step(DT, K)
        ens.synchronize()
        wall_plain = time.perf_counter() - t1
        st_plain = ens.stats()
        if K != ACCURACY_STEPS:  # the surrogate is judged after a fixed horizon
            ens.upload(oceans)
            ens.step(DT, ACCURACY_STEPS)
        T_ref = [ens.member_get_field(m, "temperature") for m in range(len(oceans))]
        S_ref = [ens.member_get_field(m, "salinity") for m in range(len(oceans))]
    log(f"[bench] ensemble plain: {K} steps in {wall_plain * 1e3:.1f} ms "
        f"(block {st_plain['block_threads']}, lds {st_plain['lds_levels']})")
    guard = {"strict": GUARD_STRICT, "surface": GUARD_SURFACE, "range": GUARD_RANGE}[args.guard]
    net_path = Path(tempfile.gettempdir()) / f"rdb_bench_ensemble_closure_{os.getpid()}.bin"
    train_oceans = generate([(sz, args.seed + 100 * i, args.seed + 5000 + i) for i, sz in enumerate(sizes)],
                            "training state of")
    with ColumnEnsemble.from_oceans(train_oceans, config=cfg) as ens:
        del train_oceans
        n = ens.mlp_collect(DT, args.train_steps, stride=args.ens_train_stride)
        t2 = time.perf_counter()
        schedule = parse_schedule(args.lr_schedule)
        tl, vl = train_schedule(ens, schedule, guard)
        ens.synchronize()
        t_train = time.perf_counter() - t2
        ens.mlp_save(net_path)
    log(f"[bench] ensemble surrogate: {n} training columns, {sum(n for n, _ in schedule)} epochs in {t_train:.1f} s, "
        f"val loss {vl[0]:.4f} -> {vl[-1]:.4f}")
    with ColumnEnsemble.from_oceans(oceans, config=cfg) as ens:
        ens.mlp_load(net_path)
        ens.set_guard(guard)
        ens.set_closure(CLOSURE_SURROGATE)
        ens.step(DT, 1)
        ens.upload(oceans)
        h0, s0 = ens.heat_salt_totals()
        ens.set_timing(True)
        ens.synchronize()
        t3 = time.perf_counter()
        ens.step(DT, ACCURACY_STEPS)
        ens.synchronize()
        wall_s = time.perf_counter() - t3
        st_s = ens.stats()
        h1, s1 = ens.heat_salt_totals()
        t_sq = s_sq = 0.0
        cnt = 0
        per = []
        for m, o in enumerate(oceans):
            live = (o.h > 1.5e-4) & (o.wet > 0.5)[:, :, None]
            dT = (ens.member_get_field(m, "temperature") - T_ref[m])[live]
            dS = (ens.member_get_field(m, "salinity") - S_ref[m])[live]
            t_sq += float((dT * dT).sum())
            s_sq += float((dS * dS).sum())
            cnt += dT.size
            per.append(float(np.sqrt(np.mean(dT * dT))))
    dep = ACCURACY_STEPS * DT / (cfg.rho0 * cfg.cp) * sum(float((o.Q_heat * o.wet).sum()) for o in oceans)
    plain_value = n_wet * K / wall_plain
    sur_value = n_wet * ACCURACY_STEPS / wall_s
    t_rms, s_rms = float(np.sqrt(t_sq / cnt)), float(np.sqrt(s_sq / cnt))
    return {
        "metric": METRIC,
        "value": plain_value,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": 1,
        "ms_per_step": wall_plain * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, roundabout_amd/synth.py)",
        "config": {"workload": "batched ensemble: 16 z-level members, each seeded separately, from "
                               "256x256x50 to 1440x1080x75; one persistent graph-captured launch per step",
                   "members": [list(sz) for sz in sizes], "wet_columns": n_wet, "path": "plain",
                   "block_threads": st_plain["block_threads"], "lds_levels": st_plain["lds_levels"],
                   "graph_captured": bool(st_plain["graph_replay"])},
        "roofline": {"bound": "hbm", "achieved": st_plain["algorithmic_bytes_per_step"] / (wall_plain / K) / 1e9,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": st_plain["algorithmic_bytes_per_step"] / (wall_plain / K) / 1e9 / PEAK_HBM_GBS,
                     "traffic": None, "note": "wall-clock over the timed steps"},
        "surrogate": {"value": sur_value, "unit": "column-steps/s", "steps": ACCURACY_STEPS,
                      "ms_per_step": wall_s * 1e3 / ACCURACY_STEPS,
                      "speedup_vs_plain": sur_value / plain_value, "T_rms_K": t_rms, "S_rms_psu": s_rms,
                      "T_rms_K_per_member": per, "meets_T_error_bar": bool(t_rms <= 0.01),
                      "meets_error_bar": bool(t_rms <= 0.01 and s_rms <= 0.002),
                      "heat_rel_residual": abs((h1 - h0) - dep) / abs(h0), "salt_rel_residual": abs(s1 - s0) / abs(s0),
                      "fallback_fraction_last_step": st_s["fallback_fraction"], "guard": args.guard,
                      "graph_captured": bool(st_s["graph_replay"]),
                      "training": {"columns": n, "epochs": sum(e for e, _ in schedule), "val_loss_first": vl[0],
                                   "val_loss_last": vl[-1], "seconds": t_train},
                      "mlp": {"forward_ms": st_s["mlp_forward_ms"], "flop_per_step": st_s["mlp_flop_per_step"]}},
    }


def copy_probe(nbytes: int = 4 << 30, reps: int = 10) -> None:
    """Device-to-device copies of a known size (the runtime's copy kernel),
    the calibration run for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # warm-up
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    dt = (time.perf_counter() - t0) / reps
    hip.hipFree(a)
    hip.hipFree(b)
    # a copy reads and writes nbytes each
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps, "ms_per_copy": dt * 1e3,
                      "read_plus_write_GBs": 2 * nbytes / dt / 1e9}), flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", action="store_true", help="run only the copy-kernel calibration")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="tile", choices=["tile", "small", "thin"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "split"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-closure leg")
    ap.add_argument("--no-remap", action="store_true", help="skip the ALE remap leg")
    ap.add_argument("--remap-only", action="store_true", help="run only the ALE remap leg (for profiling)")
    ap.add_argument("--no-contrib", action="store_true", help="skip the EPBL/convection/double-diffusion leg")
    ap.add_argument("--contrib-only", action="store_true",
                    help="run only the EPBL/convection/double-diffusion leg (for profiling)")
    ap.add_argument("--ensemble", action="store_true",
                    help="the 16-member batched ensemble (about 10.5 M columns) instead of the single tile")
    ap.add_argument("--guard", default="surface", choices=["strict", "surface", "range"])
    ap.add_argument("--lr-schedule", default="180@1e-3,60@3e-4,60@1e-4,60@3e-5",
                    help="learned-closure training: comma-separated EPOCHS@LR stages, run in order")
    ap.add_argument("--train-steps", type=int, default=2, help="plain steps recorded per training state")
    ap.add_argument("--train-states", type=int, default=6, help="independent state draws the network trains on")
    ap.add_argument("--train-stride", type=int, default=1, help="record every n-th column of each training step")
    ap.add_argument("--ens-train-stride", type=int, default=1,
                    help="the ensemble's training: record every n-th column of each training step")
    args = ap.parse_args()

    if args.copy_probe:
        copy_probe()
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU: the column path shards by columns "
                          "(independent tiles, no halo), and a multi-GPU run is not implemented",
                          "n_gpus": args.gpus}))
        return 2

    if args.ensemble:
        print(json.dumps(ensemble_bench(args)), flush=True)
        return 0

    from roundabout_amd import synth
    from roundabout_amd.rdb import PREC_FP64, PREC_SPLIT, ColumnPhysics, default_config

    t0 = time.perf_counter()
    ocean = synth.WORKLOADS[args.workload](args.seed)
    if args.remap_only:
        print(json.dumps({"remap": remap_leg(ocean, args.seed, cpu=not args.no_cpu)}), flush=True)
        return 0
    if args.contrib_only:
        print(json.dumps({"contrib": contrib_leg(ocean, args.steps, cpu=not args.no_cpu)}), flush=True)
        return 0
    log(f"[bench] generated {args.workload} {ocean.nx}x{ocean.ny}x{ocean.nz}, wet {ocean.n_wet} "
        f"in {time.perf_counter() - t0:.1f} s")
    cfg = default_config()
    cp = ColumnPhysics.from_ocean(ocean, config=cfg)
    cp.set_precision(PREC_SPLIT if args.precision == "split" else PREC_FP64)
    log(f"[bench] uploaded in {time.perf_counter() - t0:.1f} s")
    if not args.no_autotune:
        cp.autotune(DT)
    cp.set_timing(True)
    if args.warmup > 0:
        cp.step(DT, args.warmup)
    cp.synchronize()
    # capture + instantiate the K-step graph outside the timed region (its
    # first replay is untimed too); the timed region replays the cached graph
    cp.step(DT, args.steps)
    cp.synchronize()
    cp.set_timing(False)  # reset the event accumulators
    cp.set_timing(True)
    cp.synchronize()
    t1 = time.perf_counter()
    cp.step(DT, args.steps)
    cp.synchronize()
    wall = time.perf_counter() - t1
    stats = cp.stats()
    log(f"[bench] {args.steps} steps in {wall * 1e3:.2f} ms; kernel avg "
        f"{stats['sum_kernel_ms'] / max(stats['timed_launches'], 1):.3f} ms; block {stats['block_threads']} "
        f"lds_levels {stats['lds_levels']}")
    heat, salt = cp.heat_salt_totals()
    n_wet = int(stats["wet_columns"])
    value = n_wet * args.steps / wall
    kern_ms = stats["sum_kernel_ms"] / max(stats["timed_launches"], 1)
    bytes_per_launch = stats["algorithmic_bytes_per_step"]  # live levels (bed fillers are static)
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    workload = f"{args.workload} {ocean.nx}x{ocean.ny}x{ocean.nz}"
    traffic = read_traffic(workload)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-assembly+f64-solve",
        "data": "synthetic (seeded generator, roundabout_amd/synth.py)",
        "config": {
            "workload": f"{workload} z-level tile, 30% land, GRF bathymetry 50-5500 m; plain PP81+KPP "
                        f"vmix_assemble + Thomas T/S/2 passives, dt={DT:.0f} s",
            "wet_columns": n_wet,
            "columns": ocean.nx * ocean.ny,
            "nz": ocean.nz,
            "path": "plain",
            "precision": args.precision,
            "kernel": "k_column_step (fused, one lane per column, hipGraph replay)",
            "block_threads": stats["block_threads"],
            "lds_levels": stats["lds_levels"],
            "graph_captured": bool(stats["graph_replay"]),
            "column_order": "deepest first" if stats["order_mode"] == 1 else "grid order",
            "live_levels_mean": stats["live_levels"] / max(stats["wet_columns"], 1),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "all_level_bytes_per_launch": stats["reference_bytes_per_step"],
            "kernel_ms": kern_ms,
        },
        "heat_salt_totals": [heat, salt],
    }
    if args.precision == "fp64":
        log("[bench] momentum vertical viscosity ...")
        out["momentum"] = momentum_leg(cp, ocean)
    cp.close()
    if not args.no_remap:
        log("[bench] ALE remap leg ...")
        out["remap"] = remap_leg(ocean, args.seed, cpu=not args.no_cpu)
    if not args.no_contrib and args.precision == "fp64":
        log("[bench] EPBL / convection / double-diffusion leg ...")
        out["contrib"] = contrib_leg(ocean, cpu=not args.no_cpu)
    if not args.no_surrogate and args.precision == "fp64":
        log("[bench] learned closure leg ...")
        out["surrogate"] = surrogate_leg(ocean, cfg, args, value)
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(ocean, args.cpu_seconds, (720, 540))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

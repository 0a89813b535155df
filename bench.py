#!/usr/bin/env python3
"""Benchmark of the coupled IBM-DEM time step on one MI355X.

Workload (the plain-FP64 GPU-vs-CPU configuration): triply periodic 256^3
grid, 2000 spheres of D = 12h with 452 markers each (904,000 markers),
seeded random sequential addition (solid fraction 10.8%), rho_p/rho_f = 2,
knCoeff 1e5 and e = 0.9 (reference units), Re_g = 40 gravity and zero-net-flux
bulk forcing. One "step" is one FirstOrderForwardEuler + CloudEvolutionMS:
5 multi-direct-forcing iterations, predictor, bulk forcing, divergence, MG-PCG
Poisson solve to rtol 1e-10, projection and 10 DEM sub-steps.

The FP64 form is the headline (value = steps/s); the mixed-precision form
(FP32 predictor and FP32 multigrid inside FP64 CG) is timed on the same
state and reported beside it. The timed steps run without per-launch
events; afterwards one step per profiled kernel (spread, level-0 smoother,
interpolation, predictor) is run with HIP events around each launch of that
kernel on the context stream. The roofline headline is the single kernel
(one launch shape) with the largest time per step; the level-0 smoother's
row sums its sweep variants and is listed, not ranked. Algorithmic bytes
per launch are computed by the library (s = 8 B, F = footprint cells x 3
components): spread 60 M + 4 s F (3 s F in the first iteration, which
reads no f_old); interpolation 72 M + s F; level-0 smoother 1.5 s cells;
predictor 7 s cells. `traffic` is the measured HBM bytes per launch of the
headline kernel (profiles/, calibrated FETCH_SIZE + WRITE_SIZE).

The other configurations run by default after the headline (fluidized bed
192x192x768 and cylinder in crossflow 512x256x64, each in FP64 and mixed
precision from the same state; the batched form over 64 uneven cases), and
the CPU baseline times the oracle on the workload itself: the GPU's FP64
state after the timed steps is uploaded into the oracle, which runs two
coupled steps on the host cores granted to this process (OpenMP) and one
step on one core.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--precision fp64|mixed|both]
                [--no-batched] [--no-configs]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ibm-dem_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=256, help="grid edge (256 is the workload)")
    ap.add_argument("--spheres", type=int, default=2000)
    ap.add_argument("--batch-profile", action="store_true",
                    help="also time the batched solve's launches by kind (ibmdem_batch_profile)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched form (64 uneven settling cases)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the open-boundary configurations (fluidized bed 192x192x768 with "
                         "10,000 spheres; cylinder in crossflow 512x256x64) in FP64 and mixed precision")
    ap.add_argument("--config-steps", type=int, default=10,
                    help="timed steps of the other configurations and the batch")
    ap.add_argument("--cylinder-steps", type=int, default=0,
                    help="also run the cylinder in crossflow this many steps in FP64 and mixed precision and "
                         "report its force coefficients over time (C_D, C_L, mean C_D, Strouhal number)")
    return ap.parse_args()


PROFILED = ("spread", "mg_smooth", "interp", "predictor")


def timed_run(setup, precision, steps, warmup, profile):
    """Warm-up, `steps` timed steps (no per-launch events: the step runs as
    captured HIP graphs), the state after them, then one profiled step per
    kernel in PROFILED with HIP events around each of its launches."""
    import ibmdem
    from ibmdem import abi, cases

    case = setup.case
    case.precision = precision
    sim = ibmdem.simulation(case)
    cases.load(sim, setup)
    iters = []
    for _ in range(warmup):
        iters.append(sim.step().iterations)
    sim.sync()
    t0 = time.perf_counter()
    info = None
    for _ in range(steps):
        info = sim.step()
        iters.append(info.iterations)
    sim.sync()
    t1 = time.perf_counter()
    state = (sim.download_fields(), sim.download_markers(), sim.download_bodies())
    prof = {}
    if profile:
        ids = {"spread": abi.K_SPREAD, "mg_smooth": abi.K_MG_SMOOTH, "interp": abi.K_INTERP,
               "predictor": abi.K_PREDICTOR}
        for name in PROFILED:
            sim.profile_kernel(ids[name])
            sim.step()
            prof[name] = sim.profile_read()     # (total ms, launches, bytes per launch)
        sim.profile_kernel(0)
    mem = sim.memory_bytes()
    sim.close()
    return (t1 - t0), iters, info, prof, state, mem


KERNEL_NAMES = {
    "spread": "k_spread_tiles (IBM spread + multi-direct-forcing update)",
    "mg_smooth": "k_rb_half, all level-0 variants (red-black Gauss-Seidel half-sweeps: plain, "
                 "first sweep, residual, r.z)",
    "interp": "k_interp (IBM interpolation)",
    "predictor": "k_predictor_lds (fused advection-diffusion predictor, LDS z-marching)",
}


GROUPS = ("mg_smooth",)   # rows that sum several kernels (variants)


def roofline_of(prof, traffic=None):
    """Per profiled kernel: average launch time and algorithmic GB/s; the
    single kernel with the largest time per step is the headline (grouped
    rows are listed only)."""
    rows = {}
    for name, (total_ms, launches, bpl) in prof.items():
        if not launches:
            continue
        avg_s = total_ms * 1e-3 / launches
        achieved = bpl / avg_s / 1e9 if bpl else None
        rows[name] = {
            "kernel": KERNEL_NAMES[name],
            "ms_per_step": total_ms,
            "avg_launch_us": avg_s * 1e6,
            "launches_per_step": launches,
            "algorithmic_bytes_per_launch": bpl,
            "achieved": achieved,
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
        }
        if name in GROUPS:
            rows[name]["group"] = True
        t = (traffic or {}).get(name)
        if t and bpl:
            rows[name]["traffic"] = t
            rows[name]["traffic_over_algorithmic"] = t / bpl
    single = [k for k in rows if k not in GROUPS]
    if not single:
        return None, rows
    top = max(single, key=lambda k: rows[k]["ms_per_step"])
    r = rows[top]
    tr = (traffic or {}).get(top)
    head = {"kernel": r["kernel"], "bound": "hbm", "achieved": r["achieved"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": r["frac"], "traffic": tr,
            "traffic_over_algorithmic": tr / r["algorithmic_bytes_per_launch"] if tr else None,
            "avg_launch_us": r["avg_launch_us"], "launches_per_step": r["launches_per_step"],
            "ms_per_step": r["ms_per_step"],
            "algorithmic_bytes_per_launch": r["algorithmic_bytes_per_launch"]}
    return head, rows


def measured_traffic():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy kernel), keyed like
    PROFILED; None where the summary has no figure."""
    path = os.path.join(ROOT, "profiles", "coupled_step_256cube_fp64_hbm_traffic.md")
    out = {}
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch":
                out[f[1]] = float(f[2])
    except OSError:
        pass
    return out


def batched_run(steps, warmup, profile=False):
    """The batched form on 64 independent settling cases of uneven size
    (grids 32..96 x 32..96 x 64..192, 1..40 spheres of D in [8h, 16h]):
    mixed precision with the autotuned multigrid cycle and solve-group split
    (the forcing loop, the solve and the post-solve part each run as
    launches over all cases of a group), timed over `steps` batch steps; the
    same cases in FP64 give the per-case error."""
    import numpy as np

    import ibmdem
    from ibmdem import cases
    from ibmdem.abi import FP64, MIXED

    out = {}
    states = {}
    for name, prec in (("mixed", MIXED), ("fp64", FP64)):
        setups = cases.config_batch(precision=prec)
        with ibmdem.batch([s.case for s in setups]) as b:
            for q, s in enumerate(setups):
                cases.load(b[q], s)
            t_tune = time.perf_counter()
            if prec == MIXED:
                b.autotune(2)
                groups = b.solve_groups()
            else:
                b.set_solve_groups(groups)   # the mixed run's tuned split
            t_tune = time.perf_counter() - t_tune
            b.run(warmup)
            for q in range(len(b)):
                b[q].sync()
            n0, ms0 = b.stats()
            t0 = time.perf_counter()
            infos = b.run(steps)
            for q in range(len(b)):
                b[q].sync()
            secs = time.perf_counter() - t0
            n1, ms1 = b.stats()
            states[name] = [(b[q].download_fields(), b[q].download_bodies().x) for q in range(len(b))]
            cells = sum(s.cells for s in setups)
            out[name] = {"seconds": secs, "batch_steps_per_s": steps / secs,
                         "gcell_per_s": cells * steps / secs / 1e9,
                         "cg_iterations_mean": float(np.mean([i.iterations for i in infos])),
                         "mg_cycles": sorted({b[q].mg_params() for q in range(len(b))}),
                         "solve_groups": b.solve_groups(),
                         "batched_group_solves": n1 - n0,
                         # GPU ms of the batched solves (summed over groups, which overlap), and
                         # host ms per phase of a batch step
                         "per_step_ms": {k: (ms1[k] - ms0[k]) / steps for k in ms1},
                         "autotune_s": t_tune}
            solve_profile = None
            if profile and prec == MIXED:
                # after the timed steps and the state download: the batched
                # solve launch by launch with an event after each (not from
                # its graph), one solve group, per kind over `steps` more
                # batch steps
                b.set_solve_groups(1)
                b.profile(True)
                b.run(1)
                p0 = b.profile_read()
                b.run(steps)
                p1 = b.profile_read()
                b.profile(False)
                solve_profile = {}
                for k in p1:
                    ms = (p1[k][0] - p0[k][0]) / steps
                    n = (p1[k][1] - p0[k][1]) / steps
                    by = (p1[k][2] - p0[k][2]) / steps
                    solve_profile[k] = {"ms_per_step": ms, "launches_per_step": n,
                                        "us_per_launch": 1e3 * ms / n if n else None,
                                        "algorithmic_GBps": by / (ms * 1e-3) / 1e9 if ms > 0 and by > 0 else None}
            if solve_profile:
                out[name]["solve_profile"] = solve_profile
            out["cases"] = len(setups)
            out["cells"] = cells
            out["markers"] = sum(s.case.n_markers for s in setups)

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

    errs = {"u": [], "v": [], "w": [], "p": [], "body_x": []}
    for (fm, bm), (ff, bf) in zip(states["mixed"], states["fp64"]):
        for k, a, b in zip("uvwp", fm, ff):
            errs[k].append(rel(a, b))
        errs["body_x"].append(rel(bm, bf))
    out["error_vs_fp64_max_over_cases"] = {k: max(v) for k, v in errs.items()}
    out["steps"] = steps
    out["warmup"] = warmup
    return out


def config_runs(steps, warmup):
    """The open-boundary configurations, each in FP64 and mixed precision
    from the same initial state: steps/s, Gcell/s, and the mixed form's
    error against FP64 after the run (velocity, pressure, divergence per
    cell relative to the velocity scale, body positions; for the cylinder,
    its drag and lift coefficients in both forms)."""
    import numpy as np

    import ibmdem
    from ibmdem import cases, diagnostics
    from ibmdem.abi import BC_OUTFLOW, FP64, MIXED

    out = {}
    for name, make in (("fluidized_bed", cases.config_fluidized_bed), ("cylinder", cases.config_cylinder)):
        t_gen = time.perf_counter()
        setup = make(precision=FP64)
        t_gen = time.perf_counter() - t_gen
        print(f"[bench] {name}: {setup.cells} cells, {setup.case.n_markers} markers, "
              f"generated in {t_gen:.1f} s", file=sys.stderr, flush=True)
        res = {}
        for prec_name, prec in (("fp64", FP64), ("mixed", MIXED)):
            setup.case.precision = prec
            with ibmdem.simulation(setup.case) as sim:
                cases.load(sim, setup)
                for _ in range(warmup):
                    sim.step()
                sim.sync()
                t0 = time.perf_counter()
                iters = []
                for _ in range(steps):
                    iters.append(sim.step().iterations)
                sim.sync()
                secs = time.perf_counter() - t0
                planes = {sd: sim.download_outflow(sd) for sd in (1, 3, 5)
                          if setup.case.btype[sd] == BC_OUTFLOW}
                res[prec_name] = dict(secs=secs, iters=iters, fields=sim.download_fields(),
                                      bodies=sim.download_bodies(), planes=planes)
            print(f"[bench] {name} {prec_name}: {steps} steps in {secs:.3f} s", file=sys.stderr, flush=True)
        f64, mx = res["fp64"], res["mixed"]
        vscale = max(float(np.abs(q).max()) for q in f64["fields"][:3])
        h = setup.case.dx

        def div_max(r):
            return float(np.abs(diagnostics.divergence(setup.case, r["fields"][:3], r["planes"])).max() * h / vscale)

        row = {
            "cells": setup.cells, "markers": int(setup.case.n_markers), "bodies": int(setup.case.n_bodies),
            "dt": setup.case.dt, "steps": steps, "warmup": warmup,
            "meta": {k: v for k, v in setup.meta.items() if isinstance(v, (int, float))},
        }
        for prec_name in ("fp64", "mixed"):
            r = res[prec_name]
            row[prec_name] = {"steps_per_s": steps / r["secs"], "gcell_per_s": setup.cells * steps / r["secs"] / 1e9,
                              "cg_iterations_mean": float(np.mean(r["iters"])), "divergence": div_max(r)}
        row["mixed_speedup"] = row["mixed"]["steps_per_s"] / row["fp64"]["steps_per_s"]
        row["mixed_error_vs_fp64"] = {
            "velocity": max(float(np.abs(a - b).max()) for a, b in zip(mx["fields"][:3], f64["fields"][:3])) / vscale,
            "pressure": float(np.abs(mx["fields"][3] - f64["fields"][3]).max() / max(np.abs(f64["fields"][3]).max(), 1e-300)),
            "body_x": float(np.abs(mx["bodies"].x - f64["bodies"].x).max() / np.abs(f64["bodies"].x).max()),
        }
        if name == "cylinder":
            b = setup.meta["cylinder_body"]
            m = setup.meta
            row["force_coefficients"] = {
                p: dict(zip(("C_D", "C_L"), diagnostics.force_coefficients(res[p]["bodies"], b, 1.0, m["U"], m["D"], m["span"])))
                for p in ("fp64", "mixed")}
        out[name] = row
    return out


def cylinder_shedding(steps, every=10):
    """The cylinder in crossflow (512x256x64, Re 200) over `steps` steps in
    FP64 and mixed precision from the same state. The cylinder's assembled
    force is recorded after every step in the device probe ring (no host
    wait inside the run); C_D = 2 F_x / (rho U^2 D L), C_L from F_y; mean
    C_D, rms C_L and the Strouhal number over the second half of the run;
    the histories every `every` steps."""
    import numpy as np

    import ibmdem
    from ibmdem import cases, diagnostics
    from ibmdem.abi import FP64, MIXED

    setup = cases.config_cylinder(precision=FP64)
    m = setup.meta
    body = m["cylinder_body"]
    q = 0.5 * setup.case.rhof * m["U"] ** 2 * m["D"] * m["span"]
    out = {"steps": steps, "dt": setup.case.dt, "sample_every": every}
    hist = {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        setup.case.precision = prec
        with ibmdem.simulation(setup.case) as sim:
            cases.load(sim, setup)
            sim.probe_bodies([body], steps)
            sim.sync()
            t0 = time.perf_counter()
            done = 0
            while done < steps:
                k = min(2000, steps - done)
                sim.run(k)
                done += k
                print(f"[bench] cylinder {name}: {done}/{steps} steps, {time.perf_counter() - t0:.0f} s",
                      file=sys.stderr, flush=True)
            sim.sync()
            secs = time.perf_counter() - t0
            first, rec = sim.probe_read()
        n = rec.shape[0]
        t = setup.case.dt * (first + np.arange(n) + 1)
        cd, cl = rec[:, 0, 9] / q, rec[:, 0, 10] / q
        stats = diagnostics.shedding_statistics(t, cd, cl, m["D"], m["U"])
        hist[name] = (cd, cl)
        out[name] = dict(stats, steps_per_s=steps / secs,
                         cd_history=[round(float(v), 6) for v in cd[every - 1::every]],
                         cl_history=[round(float(v), 6) for v in cl[every - 1::every]])
    (cd64, cl64), (cd32, cl32) = hist["fp64"], hist["mixed"]
    out["mixed_vs_fp64"] = {"max_abs_cd": float(np.abs(cd32 - cd64).max()),
                            "max_abs_cl": float(np.abs(cl32 - cl64).max()),
                            "mean_cd_rel": abs(out["mixed"]["mean_cd"] - out["fp64"]["mean_cd"]) / abs(out["fp64"]["mean_cd"]),
                            "strouhal_abs": (abs(out["mixed"]["strouhal"] - out["fp64"]["strouhal"])
                                             if out["mixed"]["strouhal"] and out["fp64"]["strouhal"] else None)}
    return out


def host_cores():
    """Cores this process may use (its CPU affinity set, capped by
    OMP_NUM_THREADS when that is set: a GPU box grants a share of its host)
    and the physical cores lscpu reports, for the record."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    phys = None
    try:
        import subprocess
        info = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(line.split(":", 1) for line in info.splitlines() if ":" in line)
        phys = int(kv["Core(s) per socket"].strip()) * int(kv["Socket(s)"].strip())
    except Exception:
        pass
    return n, phys


def cpu_baseline(setup, state, steps_multi=2, steps_one=1):
    """The oracle on the workload itself: the GPU's FP64 state after the
    timed steps (fields, markers, bodies) is uploaded into the oracle, which
    then runs `steps_multi` coupled steps with OpenMP on the granted cores and
    `steps_one` step on one core (results are the same bits either way)."""
    import ctypes

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from ibmdem.abi import FP64

    lib = oracle.library()
    lib.ibmdem_oracle_set_threads.restype = ctypes.c_int
    lib.ibmdem_oracle_set_threads.argtypes = [ctypes.c_int]
    cores, phys = host_cores()
    fields, markers, bodies = state
    setup.case.precision = FP64
    out = {}
    with oracle.simulation(setup.case) as sim:
        sim.upload_bodies(bodies)
        sim.upload_markers(markers)
        sim.upload_fields(*fields)
        for label, threads, steps in (("multi", cores, steps_multi), ("one", 1, steps_one)):
            used = lib.ibmdem_oracle_set_threads(threads)
            t0 = time.perf_counter()
            its = [sim.step().iterations for _ in range(steps)]
            secs = time.perf_counter() - t0
            out[label] = dict(threads=used, steps=steps, secs=secs, its=its)
            print(f"[bench] cpu oracle, {used} thread(s): {steps} steps in {secs:.1f} s, CG {its}",
                  file=sys.stderr, flush=True)
    m, o = out["multi"], out["one"]
    return {
        "value": m["steps"] / m["secs"],
        "unit": "steps/s",
        "gcell_per_s": setup.cells * m["steps"] / m["secs"] / 1e9,
        "cores": m["threads"],
        "physical_cores_lscpu": phys,
        "kind": "port",
        "sample": f"the 256^3 workload itself from the GPU's state after the timed steps: "
                  f"{m['steps']} steps on {m['threads']} cores in {m['secs']:.1f} s (CG iterations {m['its']})",
        "one_core": {"value": o["steps"] / o["secs"], "unit": "steps/s", "cores": 1,
                     "sample": f"{o['steps']} step in {o['secs']:.1f} s (CG iterations {o['its']})"},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design, not a run)"}))
        sys.exit(2)
    import numpy as np
    from ibmdem import cases
    from ibmdem.abi import FP64, MIXED

    setup = cases.config_suspension(n_spheres=args.spheres, n=args.n)
    cells = setup.cells
    out = {}
    results = {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        if args.precision not in (name, "both"):
            continue
        secs, iters, info, prof, state, mem = timed_run(setup, prec, args.steps, args.warmup, True)
        results[name] = dict(secs=secs, iters=iters, prof=prof, state=state, mem=mem)
        print(f"[bench] {name}: {args.steps} steps in {secs:.3f} s, "
              f"CG iterations {iters}", file=sys.stderr, flush=True)

    head = "fp64" if "fp64" in results else "mixed"
    r = results[head]
    sps = args.steps / r["secs"]
    roofline, kernels = roofline_of(r["prof"], measured_traffic() if args.n == 256 else None)
    line = {
        "metric": "steps/s",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["secs"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f64+f32",
        "data": "synthetic: seeded RSA sphere suspension (seed 20260002), fluid at rest at t=0",
        "config": {
            "workload": f"periodic suspension {args.n}^3, {args.spheres} spheres D=12h, "
                        f"{setup.case.n_markers} markers, FP64, MG-PCG rtol 1e-10",
            "grid": [args.n] * 3,
            "spheres": args.spheres,
            "markers": int(setup.case.n_markers),
            "dt": setup.case.dt,
        },
        "gcell_per_s": cells * sps / 1e9,
        "cg_iterations_mean": float(np.mean(r["iters"][args.warmup:])) if len(r["iters"]) > args.warmup else None,
        "device_memory_gb": r["mem"] / 1e9,
        "roofline": roofline,
        "kernels": kernels,
    }
    if "mixed" in results and head == "fp64":
        m = results["mixed"]
        msps = args.steps / m["secs"]
        uf, mf, bf = results["fp64"]["state"]
        um, mm, bm = m["state"]
        def rel(a, b):
            return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
        _, mk = roofline_of(m["prof"])
        line["mixed"] = {
            "steps_per_s": msps,
            "gcell_per_s": cells * msps / 1e9,
            "speedup_vs_fp64": msps / sps,
            "cg_iterations_mean": float(np.mean(m["iters"][args.warmup:])),
            "kernels": mk,
            "error_vs_fp64_after_run": {
                "u": rel(um[0], uf[0]), "v": rel(um[1], uf[1]), "w": rel(um[2], uf[2]),
                "p": rel(um[3], uf[3]), "marker_x": rel(mm.x, mf.x), "body_x": rel(bm.x, bf.x),
            },
        }
    csteps, cwarm = args.config_steps, min(args.warmup, 3)
    if not args.no_configs:
        line["configs"] = config_runs(csteps, cwarm)
    if not args.no_batched:
        line["batched"] = batched_run(csteps, cwarm, args.batch_profile)
    if args.cylinder_steps > 0:
        line["cylinder_shedding"] = cylinder_shedding(args.cylinder_steps)
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(setup, results["fp64"]["state"] if "fp64" in results else r["state"])
        except Exception as e:  # the baseline must not hide the GPU line
            line["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: zone-cycles/s of the FP64 MaHyCo time step on the 1792x768
three-material triple point, on one MI355X.

A step is one full Lagrange + Euler-ADI remap iteration over the whole mesh
with all data resident in HBM.  W warmup steps run first, then the state is
re-initialised and exactly K steps are timed (K=500 is the configuration's
run), bracketed by device synchronisations.  Per-kernel durations come from
HIP events recorded on the library's stream around every launch of the timed
region.  Untimed runs of the configuration's 500 steps, in FP64 and in the
FP32-bulk precision form, give the accuracy fields: their densities against
the C++ oracle's (a port of the reference arithmetic) after the same 500
steps.  That oracle run, on all host threads, is the CPU baseline; one
thread is timed on a 20-step sample.  The `sod` field is the 400x20 Sod
shock tube run to t = 0.2 on the GPU, with its L1 density error against the
exact Riemann solution; the oracle times the same run inside the CPU
baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "arcane-benchs_amd"))

import numpy as np  # noqa: E402

from mahyco_amd import abi, cases, load_gpu, roofline  # noqa: E402

NX, NY, NENV = 1792, 768, 3


def kernel_times(sim):
    kt = sim.L.optional["kernel_times"]
    import ctypes as C
    cap = 32
    names = (C.c_char_p * cap)()
    ms = (C.c_double * cap)()
    cnt = (C.c_int32 * cap)()
    n = kt(sim.ctx, -1, names, ms, cnt, cap)
    return {names[i].decode(): (ms[i], cnt[i]) for i in range(n)}


def set_timing(sim, on):
    sim.L.optional["kernel_times"](sim.ctx, 1 if on else 0, None, None, None, 0)


def oracle_lib():
    """The C++ oracle, loaded as the tests load it (the checker of the GPU runs)."""
    from importlib import util
    spec = util.spec_from_file_location("oracle_lib", os.path.join(ROOT, "tests", "helpers.py"))
    H = util.module_from_spec(spec)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    spec.loader.exec_module(H)
    return H.oracle()


def cpu_baseline(gpu_density, steps, steps_1thread, sod_gpu=None, fp32_density=None):
    """Oracle on the same triple point for `steps` steps on all host threads
    (the accuracy reference of the GPU runs after the same steps), then one
    thread for `steps_1thread` steps; and, if sod_gpu (the GPU's Sod run) is
    given, the whole Sod case on all threads and on one."""
    L = oracle_lib()
    threads = L.optional["set_threads"](0)
    o = cases.triple_point(L, NX, NY)
    t0 = time.perf_counter()
    done = 0
    while done < steps:  # progress line every 100 steps
        done += o.run(min(100, steps - done))
        print(f"oracle triple point: {done}/{steps} steps, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    rho = o.download(["density"])["density"]
    o.close()

    def l1(a):
        return float(np.sum(np.abs(a - rho)) / np.sum(np.abs(rho)))

    acc = {"after_steps": steps, "fp64": l1(gpu_density)}
    if fp32_density is not None:
        acc["fp32"] = l1(fp32_density)
        acc["fp32_linf"] = float(np.max(np.abs(fp32_density - rho)) / np.max(np.abs(rho)))
    out = {"value": NX * NY * steps / (t1 - t0), "unit": "zone-cycles/s", "cores": threads,
           "kind": "port",
           "sample": f"C++ oracle (OpenMP, {threads} threads) on the same 1792x768 triple point, "
                     f"steps 1..{steps}, {t1 - t0:.1f} s"}
    L.optional["set_threads"](1)
    o = cases.triple_point(L, NX, NY)
    t0 = time.perf_counter()
    o.run(steps_1thread)
    t1 = time.perf_counter()
    o.close()
    out["value_1thread"] = NX * NY * steps_1thread / (t1 - t0)
    out["sample_1thread"] = f"the same, 1 thread, steps 1..{steps_1thread}, {t1 - t0:.1f} s"
    if sod_gpu is not None:
        sod = {}
        for th in (threads, 1):
            L.optional["set_threads"](th)
            o = cases.sod(L, SOD_NX, SOD_NY, **SOD_OPT)
            t0 = time.perf_counter()
            n = o.run(1 << 30)
            t1 = time.perf_counter()
            rho_o = o.download(["density"])["density"]
            o.close()
            sod[f"zone_cycles_per_s_{th}_threads"] = SOD_NX * SOD_NY * n / (t1 - t0)
        sod["steps"] = n
        sod["l1_density_gpu_vs_oracle"] = float(np.sum(np.abs(sod_gpu - rho_o)) / np.sum(np.abs(rho_o)))
        out["sod"] = sod
    L.optional["set_threads"](threads)
    return out, acc


# Sod shock tube (SODService.cc:11-39, cas-test 3): 400x20 cells on [0,1]x[0,0.05],
# deltat-init 1e-5, deltat-min 1e-8, deltat-max 0.01, final time 0.2 (Data.1.arc),
# artificial viscosity coefficients 0.5 (linear) and 1.0 (quadratic; the
# reference hard-codes (gamma + 1) / 2 = 1.2, MahycoModule.cc:517-528)
SOD_NX, SOD_NY = 400, 20
SOD_OPT = dict(deltat_init=1e-5, deltat_min=1e-8, deltat_max=0.01, final_time=0.2, visc_quad_coeff=1.0)


def sod_l1(st):
    from mahyco_amd import riemann
    rho = np.asarray(st["density"])
    xc = (np.arange(SOD_NX) + 0.5) / SOD_NX
    rex, _, _ = riemann.sample(xc, st["global_time"], 0.5, (1, 0, 1), (0.125, 0, 0.1), 1.4)
    return float(np.mean(np.abs(rho.reshape(SOD_NY, SOD_NX) - rex[None, :])))


def sod_run(gpu, sync):
    """The Sod case to t = 0.2 on the GPU (FP64): zone-cycles/s of the whole
    run and its L1 density error against the exact Riemann solution; also
    the L1 error with the reference's own quadratic coefficient."""
    s = cases.sod(gpu, SOD_NX, SOD_NY, **SOD_OPT)
    s.run(3)  # first launches and graph capture out of the timed region
    s.close()
    s = cases.sod(gpu, SOD_NX, SOD_NY, **SOD_OPT)
    sync()
    t0 = time.perf_counter()
    n = s.run(1 << 30)
    sync()
    t1 = time.perf_counter()
    st = s.download()
    s.close()
    rho = np.asarray(st["density"])
    ref = dict(SOD_OPT, visc_quad_coeff=-1.0)
    s = cases.sod(gpu, SOD_NX, SOD_NY, **ref)
    s.run(1 << 30)
    l1_ref = sod_l1(s.download())
    s.close()
    return {"config": "Sod 400x20 on [0,1]x[0,0.05], gamma 1.4, order 2 minmod, viscosity 0.5 / 1.0, "
                      "to t = 0.2 (cas-test 3)",
            "steps": n, "time": st["global_time"], "value": SOD_NX * SOD_NY * n / (t1 - t0),
            "unit": "zone-cycles/s", "l1_density_vs_exact_riemann": sod_l1(st),
            "l1_density_vs_exact_riemann_quadratic_coeff_1.2": l1_ref}, rho


def timed_run(gpu, precision, args, sync):
    """Autotune, W warm-up steps, re-initialise, then K steps timed (graph
    replay); then, unless --no-events, the same K steps again launched one
    kernel at a time between HIP events for the per-kernel durations."""
    opt_kw = {"precision": precision}
    sim = cases.triple_point(gpu, NX, NY, **opt_kw)
    tuning = sim.autotune(10)
    print("autotune:", tuning, file=sys.stderr)
    if args.warmup:
        sim.L.optional["run_async"](sim.ctx, args.warmup)
        sim.L.optional["synchronize"](sim.ctx)
    sim.init_triple_point(cases.TP_RHO, cases.TP_P)
    set_timing(sim, False)
    sync()
    sim.L.optional["synchronize"](sim.ctx)
    t0 = time.perf_counter()
    r = sim.L.optional["run_async"](sim.ctx, args.steps)
    assert r == 0, sim.L.last_error(sim.ctx)
    sim.L.optional["synchronize"](sim.ctx)
    sync()
    t1 = time.perf_counter()
    st = sim.download(["density", "cell_mass"])
    assert st["global_iteration"] == args.steps, st["global_iteration"]
    assert np.all(np.isfinite(st["density"]))
    kt = {}
    if not args.no_events:
        sim.init_triple_point(cases.TP_RHO, cases.TP_P)
        set_timing(sim, True)
        r = sim.L.optional["run_async"](sim.ctx, args.steps)
        assert r == 0, sim.L.last_error(sim.ctx)
        sim.L.optional["synchronize"](sim.ctx)
        kt = kernel_times(sim)
        set_timing(sim, False)
        st2 = sim.download(["density"])
        assert np.array_equal(st2["density"], st["density"]), "graph and direct launches differ"
    return {"sim": sim, "elapsed": t1 - t0, "tuning": tuning, "kt": kt, "density": st["density"]}


def p3d_run(gpu, args, sync):
    """The 3D random-field case in FP64 and in the FP32-bulk precision form:
    throughput of each (graph-replayed steps) and per-field L1/Linf of the
    precision form against the FP64 path after the same steps."""
    n, steps = args.p3d_n, args.p3d_steps
    t0 = time.perf_counter()
    fields = cases.p3d_fields(n)
    t_fields = time.perf_counter() - t0
    out = {"config": f"{n}^3 unit cube, gamma 5/3, rho = exp(G) (corr. 8 cells, sigma 0.3), p = 1, "
                     f"solenoidal u (rms 0.1), reflecting walls, {steps} steps, seed {cases.P3D_SEED}",
           "fields_s": round(t_fields, 1)}
    states = {}
    for name, prec in [("fp64", abi.PREC_FP64), ("fp32", abi.PREC_FP32_BULK)]:
        s = cases.p3d(gpu, n, fields=fields, precision=prec)
        # warm-up of one graph group: captures the step graph before timing
        s.L.optional["run_async"](s.ctx, 12)
        s.L.optional["synchronize"](s.ctx)
        s.init_cells(np.zeros(n ** 3, dtype=np.int32), fields[0], np.ones(n ** 3), fields[1])
        sync()
        t0 = time.perf_counter()
        r = s.L.optional["run_async"](s.ctx, steps)
        assert r == 0, s.L.last_error(s.ctx)
        s.L.optional["synchronize"](s.ctx)
        sync()
        el = time.perf_counter() - t0
        st = s.download(["density", "internal_energy", "pressure", "velocity"])
        assert st["global_iteration"] == steps
        states[name] = st
        out[name] = {"value": n ** 3 * steps / el, "unit": "zone-cycles/s", "ms_per_step": 1e3 * el / steps}
        s.close()
    out["speedup_fp32_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    err = {}
    for f in ["density", "internal_energy", "pressure", "velocity"]:
        a = np.asarray(states["fp32"][f], dtype=np.float64)
        b = np.asarray(states["fp64"][f], dtype=np.float64)
        err[f] = {"l1": float(np.sum(np.abs(a - b)) / np.sum(np.abs(b))),
                  "linf": float(np.max(np.abs(a - b)) / np.max(np.abs(b)))}
    out["fp32_error_vs_fp64"] = err
    out["roofline_frac_fp64"] = round(out["fp64"]["value"] * roofline.bytes_per_zone_cycle(1, dim=3)
                                      / (roofline.PEAK_HBM_GBS * 1e9), 4)
    # short untimed runs from the same fields, compared with the oracle by p3d_oracle
    short = {}
    if args.p3d_oracle_steps and not args.no_cpu_baseline:
        for name, prec in [("fp64", abi.PREC_FP64), ("fp32", abi.PREC_FP32_BULK)]:
            s = cases.p3d(gpu, n, fields=fields, precision=prec)
            s.run(args.p3d_oracle_steps)
            short[name] = s.download(P3D_FIELDS)
            s.close()
    return out, fields, short


P3D_FIELDS = ["density", "internal_energy", "pressure", "velocity"]


def p3d_oracle(n, fields, steps, short):
    """The oracle (all host threads) on the 3D random-field case for `steps`
    steps from the same fields; per-field relative L1 / Linf of the GPU's
    FP64 and FP32-bulk runs of the same steps against it."""
    L = oracle_lib()
    threads = L.optional["set_threads"](0)
    o = cases.p3d(L, n, fields=fields)
    t0 = time.perf_counter()
    for k in range(steps):
        o.run(1)
        print(f"oracle {n}^3: {k + 1}/{steps} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    ref = o.download(P3D_FIELDS)
    o.close()
    out = {"after_steps": steps, "oracle_s": round(el, 1), "oracle_threads": threads,
           "oracle_zone_cycles_per_s": n ** 3 * steps / el}
    for name, st in short.items():
        assert st["global_iteration"] == steps
        err = {}
        for f in P3D_FIELDS:
            a = np.asarray(st[f], dtype=np.float64)
            b = np.asarray(ref[f], dtype=np.float64)
            err[f] = {"l1": float(np.sum(np.abs(a - b)) / np.sum(np.abs(b))),
                      "linf": float(np.max(np.abs(a - b)) / np.max(np.abs(b)))}
        out[name] = err
    return out


def batch_run(gpu, args, sync):
    """Batching form: the seeded batch of 2D Riemann problems, every problem to
    its own t_end, in FP32 bulk and in FP64.  Throughput = cells x iterations
    summed over the problems / wall time of the whole run (one mhg_batch_run
    call, host list updates included); the launch shapes are autotuned first
    (state restored).  The FP32 densities are compared with the FP64 batch's,
    over all cells of all problems."""
    specs = cases.riemann_batch_specs(args.batch_n)
    out = {"config": f"{args.batch_n} 2D x-Riemann problems, nx ~ U{{64..1024}}, ny ~ U{{8..64}}, "
                     f"rho ~ logU[0.1,10], p ~ logU[0.01,10], u ~ U[-1,1], x0 ~ U[0.3,0.7], "
                     f"gamma ~ U[1.2,1.67], reflecting walls, each to t_end = {cases.BATCH_TEND}, "
                     f"seed {cases.BATCH_SEED}",
           "cells": int(sum(sp["nx"] * sp["ny"] for sp in specs))}
    dens = {}
    for name, prec in [("fp32", abi.PREC_FP32_BULK), ("fp64", abi.PREC_FP64)]:
        t0 = time.perf_counter()
        b = cases.riemann_batch(gpu, specs, precision=prec)
        t_init = time.perf_counter() - t0
        tuning = b.autotune(2)
        print(f"batch {name} autotune:", tuning, file=sys.stderr)
        sync()
        t0 = time.perf_counter()
        nrun, zc = b.run()
        sync()
        el = time.perf_counter() - t0
        assert nrun == 0
        its = np.array([b.status(i)["iterations"] for i in range(b.n)])
        dens[name] = np.concatenate([b.problem(i).download(["density"])["density"] for i in range(b.n)])
        assert np.all(np.isfinite(dens[name]))
        v = zc / el
        bpz = (roofline.bytes_per_zone_cycle_fp32(1) if prec == abi.PREC_FP32_BULK
               else roofline.bytes_per_zone_cycle(1))
        out[name] = {"value": v, "unit": "zone-cycles/s", "wall_s": round(el, 4), "zone_cycles": int(zc),
                     "iterations_max": int(its.max()), "iterations_mean": round(float(its.mean()), 1),
                     "init_s": round(t_init, 2), "bytes_per_zone_cycle": bpz,
                     "roofline_frac": round(v * bpz / (roofline.PEAK_HBM_GBS * 1e9), 4),
                     "tuning": tuning.strip()}
        b.close()
    out["speedup_fp32_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    a, r = dens["fp32"], dens["fp64"]
    out["l1_density_err_fp32_vs_fp64"] = float(np.sum(np.abs(a - r)) / np.sum(np.abs(r)))
    out["linf_density_err_fp32_vs_fp64"] = float(np.max(np.abs(a - r)) / np.max(np.abs(r)))
    return out


def learned_run(gpu, args):
    """Learned form: the network trained by mahyco_amd.learned (weights in
    mahyco_amd/data/limiter_mlp.npz) on the triple point at 896x384 against
    the plain method at 1x, 1.5x and 2x that resolution; errors are relative
    L1 densities on the 896x384 grid against a plain run at 4x (block means),
    every run ending at t exactly (land_on_final_time)."""
    from mahyco_amd import learned
    path = os.path.join(ROOT, "arcane-benchs_amd", "mahyco_amd", "data", "limiter_mlp.npz")
    d = np.load(path)
    params = d["params"]
    train = json.loads(str(d["summary"]))
    ev = learned.evaluate_tp(gpu, params, args.learned_t, log=lambda s: print(s, file=sys.stderr))
    keep = ["problems", "fine", "samples", "epochs", "batch", "lr", "validation_loss", "validation_loss_central",
            "validation_loss_minmod", "validation_loss_first_order", "data_s", "train_s", "fine_zone_cycles",
            "fine_run_s"]
    return {"evaluation": ev, "training": {k: train[k] for k in keep if k in train},
            "network": "128-64-64-2 ReLU, BF16 weights, FP32 accumulation (MFMA)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--accuracy-steps", type=int, default=500,
                    help="steps of the untimed triple-point runs whose densities are compared "
                         "(FP64 and FP32 bulk against the oracle, FP32 against FP64)")
    ap.add_argument("--cpu-steps-1thread", type=int, default=20,
                    help="steps of the one-thread oracle sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="do not record per-kernel events")
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the FP32-bulk precision form run")
    ap.add_argument("--p3d-n", type=int, default=256, help="cells per side of the 3D random-field case")
    ap.add_argument("--p3d-steps", type=int, default=200)
    ap.add_argument("--no-p3d", action="store_true", help="skip the 3D random-field case")
    ap.add_argument("--p3d-oracle-steps", type=int, default=3,
                    help="steps of the 3D runs compared with the oracle (0: none; needs the CPU baseline)")
    ap.add_argument("--batch-n", type=int, default=1024, help="problems of the Riemann batch")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching form")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned form")
    ap.add_argument("--no-sod", action="store_true", help="skip the Sod case")
    ap.add_argument("--learned-t", type=float, default=0.5, help="final time of the learned-form evaluation")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device with torch: the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counts to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        sync()
        del a, b
    gpu = load_gpu()
    res = timed_run(gpu, abi.PREC_FP64, args, sync)
    sim, elapsed, tuning, kt = res["sim"], res["elapsed"], res["tuning"], res["kt"]
    # precision form: FP32 bulk fields, FP64 coordinates, geometry and dt
    p32 = None
    if not args.no_precision_form:
        r32 = timed_run(gpu, abi.PREC_FP32_BULK, args, sync)
        v32 = sim.nc * args.steps / r32["elapsed"]
        v64 = sim.nc * args.steps / elapsed
        p32 = {"value": v32, "unit": "zone-cycles/s", "ms_per_step": 1e3 * r32["elapsed"] / args.steps,
               "speedup_vs_fp64": v32 / v64,
               "reference": "densities after the same untimed steps of the FP64 path and of the oracle",
               "dtype": "f32 bulk; f64 coordinates, cell geometry, dt",
               "kernel_ms": {k: round(v[0], 5) for k, v in r32["kt"].items()}}
        r32["sim"].close()
    nc, nn = sim.nc, sim.nn
    value = nc * args.steps / elapsed
    # accuracy: the configuration's full run (500 steps), FP64 and FP32 bulk
    rho_acc, drift = {}, {}
    for name, prec in [("fp64", abi.PREC_FP64)] + ([] if args.no_precision_form else
                                                    [("fp32", abi.PREC_FP32_BULK)]):
        g2 = cases.triple_point(gpu, NX, NY, precision=prec)
        c0 = g2.conservation()
        g2.run(args.accuracy_steps)
        c1 = g2.conservation()
        st2 = g2.download(["density"])
        assert st2["global_iteration"] == args.accuracy_steps
        rho_acc[name] = st2["density"]
        drift[name] = (c1[0] - c0[0]) / c0[0]
        g2.close()
    if p32 is not None:
        a, r = rho_acc["fp32"], rho_acc["fp64"]
        p32["l1_density_err_vs_fp64"] = float(np.sum(np.abs(a - r)) / np.sum(np.abs(r)))
        p32["linf_density_err_vs_fp64"] = float(np.max(np.abs(a - r)) / np.max(np.abs(r)))
        p32["after_steps"] = args.accuracy_steps
        # total mass (compensated sum over the cells) after the same steps,
        # relative to the initial mass; the FP64 path's drift is the
        # reference scheme's (clamps and 1e-10 thresholds)
        p32["mass_drift"] = drift["fp32"]
        p32["mass_drift_fp64"] = drift["fp64"]
    p3, p3_fields, p3_short = (None, None, None) if args.no_p3d else p3d_run(gpu, args, sync)
    pb = None if args.no_batch else batch_run(gpu, args, sync)
    pl = None if args.no_learned else learned_run(gpu, args)
    if pb is not None:
        pb["fp32_vs_triple_point_fp64"] = pb["fp32"]["value"] / value

    rl = None
    if kt:
        tot = {k: v[0] * v[1] for k, v in kt.items()}
        dom = max(tot, key=tot.get)
        avg_ms, launches = kt[dom]
        b = roofline.kernel_bytes(dom, nc, nn, NENV)
        ach = b / (avg_ms * 1e-3) / 1e9
        rl = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1),
              "peak": roofline.PEAK_HBM_GBS, "unit": "GB/s",
              "frac": round(ach / roofline.PEAK_HBM_GBS, 4),
              "traffic": roofline.measured_traffic(dom, os.path.join(ROOT, "profiles")),
              "avg_ms": round(avg_ms, 5), "launches": launches,
              "bytes_per_launch": b}
    step_frac = value * roofline.bytes_per_zone_cycle(NENV) / (roofline.PEAK_HBM_GBS * 1e9)
    step_bytes = roofline.measured_step_bytes(os.path.join(ROOT, "profiles"))
    step_meas = None if step_bytes is None else {
        "bytes_per_step": step_bytes,
        "frac": round(step_bytes / (elapsed / args.steps) / (roofline.PEAK_HBM_GBS * 1e9), 4),
        "source": "FETCH_SIZE + WRITE_SIZE per step kernel, copy-calibrated, from the committed "
                  "triple-point traffic profile, over this run's ms_per_step"}

    ps, sod_rho = (None, None) if args.no_sod else sod_run(gpu, sync)
    cpu, acc = None, None
    if not args.no_cpu_baseline:
        cpu, acc = cpu_baseline(rho_acc["fp64"], args.accuracy_steps, args.cpu_steps_1thread, sod_rho,
                                rho_acc.get("fp32"))
        if p32 is not None:
            p32["l1_density_err_vs_oracle"] = acc["fp32"]
            p32["linf_density_err_vs_oracle"] = acc["fp32_linf"]
        if p3_short:
            p3["vs_oracle"] = p3d_oracle(args.p3d_n, p3_fields, args.p3d_oracle_steps, p3_short)

    out = {
        "metric": "zone-cycles/s on 1792x768 triple-point; L1 density err vs FP64 oracle",
        "value": value, "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"triple point 1792x768, 3 perfect-gas materials, plain FP64 "
                               f"Lagrange + Euler-ADI remap (order 2, minmod): {args.steps} steps timed, "
                               f"densities compared after {args.accuracy_steps} steps",
                   "cells": nc, "nodes": nn, "steps": args.steps, "parallelism": "none (1 GPU)",
                   "tuning": {part.split(":")[0].strip(): part.split("->")[-1].strip()
                              for part in tuning.split(";") if "->" in part} if tuning else None},
        "l1_density_err_vs_oracle": (None if acc is None else
                                     {"value": acc["fp64"], "after_steps": acc["after_steps"]}),
        "step_roofline_frac": round(step_frac, 4),
        "step_measured_hbm": step_meas,
        "kernel_ms": {k: round(v[0], 5) for k, v in kt.items()},
        "kernel_ms_source": "HIP events around each launch on the library stream, second pass "
                            "of the same K steps launched one kernel at a time",
        "roofline": rl,
        "cpu_baseline": cpu,
        "precision_form": p32,
        "precision_form_3d": p3,
        "batch_form": pb,
        "learned_form": pl,
        "sod": ps,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""RK3 steps/s of the compressible Hall-MHD step at 256^3 on one MI355X.

Workload (the 256^3 Hall-MHD case): periodic box [0,2pi)^3, gamma = 5/3,
rho0 = 1, P0 = 1, B0 = (1,0,0), seeded random-phase modes in shells 1-16
with amplitude |k|^(-5/3), rms |db| = 0.3 B0 and rms |du| = 0.3, d_i = 0.05,
nu = eta = 2e-4 (implicit), phase-shift dealiasing (dealias_option 3: the
nonlinear terms averaged over the grid and the grid shifted by half a cell,
open sqrt(2)/3 ellipsoid), CFL 0.2. A step is one full RK3 step (three
substages, CFL dt on the device).

CFL 0.2: the Hall limit of the CFL dt is cfl dx^2 rho / (d_i |B|), and the
whistler at wave number k turns at omega = d_i k^2 |B| / rho, so omega dt =
cfl (k dx)^2. RK3 is stable for |omega dt| <= sqrt(3). The open sqrt(2)/3
ellipsoid keeps k_x up to 120 at 256^3 (k dx = 2.95), which needs cfl <= 0.2;
at 0.4 the highest kept whistlers grow and the dt sequence turns erratic. The
2/3 rule keeps k_x <= 85 (0.4 would sit at the limit); both options run at 0.2.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--dealias 3|1]

Prints one JSON line on stdout: `value` is FP64 steps/s with the state
resident in HBM; the mixed-precision run (FP32 transforms and nonlinear
products, FP64 RK3 accumulation) is reported under "mixed" with its relative
L2 distance to the FP64 run after 100 steps (--compare-steps), taken with the
FP64 run's dt sequence replayed (equal time levels). The same two runs
with the reference's 2/3 rule (dealias_option 1) are reported under
"dealias_2_3", and the 96-box ensemble (32 each of 32^3, 64^3, 96^3, one
hipGraph launch per step) under "ensemble". "closure" is the learned subgrid
closure trained on box-filtered 128^3 runs and run on 32^3: shell-energy
deviations from the filtered fine run with and without it, and its cost
against the fine steps it replaces. `roofline` is the dominant kernel's algorithmic bytes over its
HIP-event time; `cpu_baseline` is the CPU oracle timed on this host for three
steps of the headline workload, on every CPU of the process's host share.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_params(n, precision, dealias=1, cfl=0.2):
    from laps_amd.params import default_params
    return default_params(nx=n, ny=n, nz=n, Lx=2 * math.pi, Ly=2 * math.pi, Lz=2 * math.pi,
                          adiabatic_index=5.0 / 3.0, cfl=cfl, dealias_option=dealias,
                          if_visc=1, viscosity=2e-4, if_resis=1, resistivity=2e-4,
                          if_hall=1, ion_inertial_length=0.05, precision=precision)


def initial_state(n, gamma):
    from laps_amd.ic import cons_from_prim, seeded_primitive
    prim = seeded_primitive((n, n, n), seed=1, kmin=1, kmax=16, du=0.3, db=0.3)
    return cons_from_prim(prim, gamma)


def shell_spectrum(uhat, n):
    """kinetic (rho u)^2/2 + magnetic B^2/2 energy per integer shell, Hermitian-weighted"""
    nxh = uhat.shape[0]
    mx = np.arange(nxh)[:, None, None]
    m = np.fft.fftfreq(n, 1.0 / n)
    kk = np.sqrt(mx ** 2 + m[None, :, None] ** 2 + m[None, None, :] ** 2)
    w = np.where(mx == 0, 1.0, 2.0) * np.ones_like(kk)
    e = sum(np.abs(uhat[..., v]) ** 2 for v in (1, 2, 3, 4, 5, 6)) * 0.5 * w
    shell = np.rint(kk).astype(np.int64).ravel()
    return np.bincount(shell, weights=e.ravel())


def run_gpu(n, precision, steps, warmup, uu, pr, dealias=1):
    import torch
    from laps_amd.runtime import Laps
    p = workload_params(n, precision, dealias)
    g = Laps(p)
    g.set_state(uu, pr)
    log(f"precision={'fp64' if precision == 0 else 'mixed'}: warmup {warmup} steps")
    g.step(warmup)
    g.sync()
    torch.cuda.synchronize()  # one process, one GPU: the barrier is this synchronisation
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the per-kernel breakdown: the same number of steps again with the library's HIP events
    # around every launch (kept out of the timed steps above, whose launches run back to back)
    g.timing_enable(True)
    g.step(steps)
    g.sync()
    timing = g.timing()
    g.timing_enable(False)
    sc = g.scalars()
    spec = g.spectral()
    g.close()
    log(f"  {steps} steps in {t1 - t0:.3f} s -> {steps / (t1 - t0):.2f} steps/s, t={sc['time']:.5f} dt={sc['dt']:.3e}")
    return dict(seconds=t1 - t0, timing=timing, scalars=sc, spectral=spec)


def dt_replay(n, steps, uu, pr, dealias):
    """Precision check at equal time levels. The CFL dt is computed from each
    run's own state, so an FP32 rounding difference at a step where the 2%
    hysteresis is close to switching gives the mixed run another dt sequence
    (the runs then end at different times). Here the FP64 run is repeated step
    by step with the CFL dt recorded, and the mixed run replays that dt
    sequence with the CFL reduction off (laps_set_fixed_dt)."""
    from laps_amd.params import LAPS_FP64, LAPS_MIXED
    from laps_amd.runtime import Laps
    g = Laps(workload_params(n, LAPS_FP64, dealias))
    g.set_state(uu, pr)
    dts = []
    for _ in range(steps):
        g.step(1)
        dts.append(g.scalars()["dt"])
    ref = g.spectral()
    g.close()
    m = Laps(workload_params(n, LAPS_MIXED, dealias))
    m.set_state(uu, pr)
    for dt in dts:
        m.set_fixed_dt(dt)
        m.step(1)
    spec = m.spectral()
    tm = m.scalars()["time"]
    m.close()
    log(f"  dt replay: {steps} steps, FP64 dt in [{min(dts):.4e}, {max(dts):.4e}], mixed t={tm:.5f}")
    return ref, spec, dts


def ensemble_params(steps_seed=100):
    """The batched ensemble: 96 independent periodic boxes (32 each of 32^3, 64^3 and 96^3),
    seeds 100..195, db/B0 = du drawn uniform in [0.1, 0.5], d_i alternating 0 / 0.05, mixed
    precision (FP32 transforms and nonlinear products, FP64 RK3 accumulation), 2/3 rule, CFL 0.4,
    nu = eta = 1e-3; k^-5/3 shells 1..n/8."""
    from laps_amd.params import default_params, LAPS_MIXED
    rng = np.random.default_rng(steps_seed)
    out = []
    for n in (32, 64, 96):
        for i in range(32):
            amp = float(rng.uniform(0.1, 0.5))
            di = 0.05 if i % 2 else 0.0
            p = default_params(nx=n, ny=n, nz=n, Lx=2 * math.pi, Ly=2 * math.pi, Lz=2 * math.pi,
                               adiabatic_index=5.0 / 3.0, cfl=0.4, dealias_option=1, if_visc=1, viscosity=1e-3,
                               if_resis=1, resistivity=1e-3, if_hall=1 if di else 0, ion_inertial_length=di,
                               precision=LAPS_MIXED)
            out.append((p, 100 + len(out), amp))
    return out


def run_ensemble(steps):
    """Box-steps/s of the ensemble advanced by hipGraph replays, and of the same boxes stepped one
    context at a time (laps_step, each box on its own stream); the two must agree bit for bit."""
    from laps_amd.ic import cons_from_prim, seeded_primitive
    from laps_amd.runtime import Ensemble, Laps
    prev_plan = os.environ.get("LAPS_FFT_PLAN")
    os.environ["LAPS_FFT_PLAN"] = "0"  # same FFT plan in both runs, so they can be compared bitwise
    spec = ensemble_params()
    states = []
    for p, seed, amp in spec:
        n = p.nx
        prim = seeded_primitive((n, n, n), seed, kmin=1, kmax=n // 8, du=amp, db=amp)
        states.append(cons_from_prim(prim, p.adiabatic_index))
    log(f"ensemble: {len(spec)} boxes, initial states built")
    res = {}
    for graph in (False, True):
        if graph:
            ens = Ensemble([p for p, _, _ in spec])
            boxes = ens.boxes
            ngroups = ens.groups()
        else:
            boxes = [Laps(p) for p, _, _ in spec]
        for b, (uu, pr) in zip(boxes, states):
            b.set_state(uu, pr)
        # warm-up: the first (per-box) step, the plan set-up and, for the ensemble, the capture
        if graph:
            ens.step(2)
            ens.sync()
        else:
            for b in boxes:
                b.step(2)
            for b in boxes:
                b.sync()
        t0 = time.perf_counter()
        if graph:
            ens.step(steps)
            ens.sync()
        else:
            for b in boxes:
                b.step(steps)
            for b in boxes:
                b.sync()
        dt = time.perf_counter() - t0
        res[graph] = dict(seconds=dt, scalars=[b.scalars() for b in boxes],
                          spectra=[boxes[i].spectral() for i in (0, 47, 95)])
        if graph:
            ens.close()
        else:
            for b in boxes:
                b.close()
        log(f"ensemble {'graph' if graph else 'per-box'}: {steps} steps of {len(spec)} boxes in {dt:.3f} s")
    if prev_plan is None:
        os.environ.pop("LAPS_FFT_PLAN", None)
    else:
        os.environ["LAPS_FFT_PLAN"] = prev_plan
    g, e = res[True], res[False]
    same = g["scalars"] == e["scalars"] and all(np.array_equal(a, b) for a, b in zip(g["spectra"], e["spectra"]))
    pts = sum(p.nx * p.ny * p.nz for p, _, _ in spec)
    nb = len(spec)
    return dict(workload="96 independent boxes: 32 x 32^3, 32 x 64^3, 32 x 96^3 (seeds 100-195, db/B0 = du "
                         "uniform in [0.1, 0.5], d_i 0 / 0.05), mixed precision, 2/3 rule, CFL 0.4, own CFL dt per box",
                steps=steps, launch_groups=ngroups, box_steps_per_s=round(nb * steps / g["seconds"], 1),
                box_steps_per_s_per_box_streams=round(nb * steps / e["seconds"], 1),
                graph_speedup=round(e["seconds"] / g["seconds"], 3),
                grid_point_steps_per_s=round(pts * steps / g["seconds"], 1),
                ms_per_ensemble_step=round(1000.0 * g["seconds"] / steps, 4),
                graph_equals_per_box_steps=bool(same),
                time_range=[min(x["time"] for x in g["scalars"]), max(x["time"] for x in g["scalars"])])


def run_output_staging(n, uu, pr, dealias, steps=40, every=10):
    """Steps/s of the FP64 workload with an output every `every` steps, the payload of
    outNNN.dat read (a) with laps_get_state (the host waits for the inverse transforms and a
    pageable download) and (b) staged on the device (laps_stage_output: the copy into pinned
    memory runs while the next step does; laps_staged_output then finds it done), against the
    rate without outputs. The payload is written to /dev/null in both cases."""
    from laps_amd.params import LAPS_FP64
    from laps_amd.runtime import Laps
    out = {}
    with Laps(workload_params(n, LAPS_FP64, dealias)) as g, open(os.devnull, "wb") as sink:
        g.set_state(uu, pr)
        g.step(2)
        g.sync()
        for mode in ("none", "get_state", "staged"):
            wait = 0.0
            pending = False
            t0 = time.perf_counter()
            for i in range(steps):
                g.step(1)
                if pending:
                    t1 = time.perf_counter()
                    w = g.staged_output(copy=False)[0]
                    wait += time.perf_counter() - t1
                    w.T.tofile(sink)
                    pending = False
                g.scalars()   # the host loop reads time and dt every step (mhd.f90:216-285)
                if (i + 1) % every:
                    continue
                if mode == "get_state":
                    t1 = time.perf_counter()
                    u, q = g.state()
                    u[..., 1:4] = q[..., 0:3]
                    u[..., 7] = q[..., 3]
                    wait += time.perf_counter() - t1
                    u.T.tofile(sink)
                elif mode == "staged":
                    g.stage_output(True)
                    pending = True
            if pending:
                g.staged_output(copy=False)[0].T.tofile(sink)
            g.sync()
            sec = time.perf_counter() - t0
            out[mode] = dict(steps_per_s=round(steps / sec, 3), host_wait_s_per_output=round(wait / (steps // every), 4)
                             if mode != "none" else 0.0)
            log(f"output staging {mode}: {steps} steps in {sec:.3f} s")
    return dict(workload=f"the FP64 {n}^3 workload, {steps} steps with an outNNN.dat payload "
                         f"({8 * n ** 3 * 8 / 1e9:.2f} GB) every {every} steps", **out)


def config2_params(precision):
    """512^3 mixed-precision Hall-MHD with the corotating expanding box: R0 = 1, Ur0 = 0.5 (Ur0/R0 = 0.5
    in box units), d_i = 0.02, nu = eta = 5e-5, 2/3 rule, CFL 0.3 (the Hall whistler at the 2/3 rule's
    k dx = 2.09 turns at cfl (k dx)^2 per step; 0.4 would sit at RK3's limit sqrt(3))."""
    from laps_amd.params import default_params
    L = 2 * math.pi
    return default_params(nx=512, ny=512, nz=512, Lx=L, Ly=L, Lz=L, adiabatic_index=5.0 / 3.0, cfl=0.3,
                          dealias_option=1, if_visc=1, viscosity=5e-5, if_resis=1, resistivity=5e-5, if_hall=1,
                          ion_inertial_length=0.02, if_aeb=1, if_corotating=1, radius0=1.0, Ur0=0.5,
                          precision=precision)


def config2_state(seed=2):
    """Outward Alfven-wave spectrum, amplitude k^-1 in shells 1-32, db/B0 = 0.5, du = -db/sqrt(rho),
    rho = p = 1, B0 = (1, 0, 0). The field has no mode beyond shell 32, so it is generated on 128^3 (same
    seeded modes, same rms) and brought to 512^3 by the library's own inverse transform of its spectrum."""
    from laps_amd.ic import band_limited_state
    from laps_amd.params import LAPS_FP64
    from laps_amd.runtime import Laps
    with Laps(config2_params(LAPS_FP64)) as g:
        return band_limited_state(512, 128, 5.0 / 3.0, g.fft_c2r, seed=seed, kmin=1, kmax=32, slope=-1.0,
                                  du=0.5, db=0.5, alfvenic=True)


def run_config2(steps):
    """FP64 and mixed runs of the 512^3 expanding-box workload: steps/s of each, and the mixed run's
    relative L2 distance to the FP64 run after `steps` steps at equal time levels."""
    from laps_amd.params import LAPS_FP64, LAPS_MIXED
    from laps_amd.runtime import Laps
    t0 = time.perf_counter()
    uu, pr = config2_state()
    log(f"config 2: 512^3 state built in {time.perf_counter() - t0:.1f} s")
    out = {}
    dts = []
    for prec in (LAPS_FP64, LAPS_MIXED):
        key = "fp64" if prec == LAPS_FP64 else "mixed"
        with Laps(config2_params(prec)) as g:
            g.set_state(uu, pr)
            g.step(1)   # first step: plan autotune; not timed
            g.sync()
            g.timing_enable(True)
            if prec == LAPS_FP64:
                dts.append(g.scalars()["dt"])
                t1 = time.perf_counter()
                for _ in range(steps - 1):
                    g.step(1)
                    dts.append(g.scalars()["dt"])
            else:
                t1 = time.perf_counter()
                for dt in dts[1:]:
                    g.set_fixed_dt(dt)
                    g.step(1)
            g.sync()
            sec = time.perf_counter() - t1
            timing = g.timing()
            g.timing_enable(False)
            sc = g.scalars()
            ns = steps - 1
            kern = {k: dict(ms_per_step=round(v[0] / ns, 4), gbs=round(v[1] / (v[0] * 1e-3) / 1e9, 1)
                            if v[0] > 0 else 0.0, launches=v[2]) for k, v in timing.items() if v[2]}
            out[key] = dict(steps_per_s=round(ns / sec, 3), time=sc["time"], radius=sc["radius"],
                            max_divb=g.max_divb(), spectral=g.spectral(), kernels=kern,
                            roofline=roofline(timing, "config2_" + key))
            log(f"config 2 {key}: {steps - 1} steps in {sec:.2f} s, t={sc['time']:.5f}, R={sc['radius']:.5f}")
    a, b = out["mixed"].pop("spectral"), out["fp64"].pop("spectral")
    rel = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
    return dict(workload="512^3 Hall-MHD, corotating expanding box (R0=1, Ur0=0.5), d_i=0.02, nu=eta=5e-5, "
                         "outward Alfven k^-1 shells 1-32, db/B0=0.5, 2/3 rule, CFL 0.3",
                steps=steps, fp64=out["fp64"], mixed=out["mixed"], mixed_rel_l2_vs_fp64=rel,
                compared_at="the mixed run replays the FP64 run's CFL dt sequence (timed with the per-step "
                            "host reads of dt and set_fixed_dt)",
                dt_range=[min(dts), max(dts)])


def roofline(timing, precision):
    name, (ms, by, nl) = max(timing.items(), key=lambda kv: kv[1][0])
    per_launch_s = ms * 1e-3 / max(nl, 1)
    achieved = (by / max(nl, 1)) / per_launch_s / 1e9 if per_launch_s > 0 else 0.0
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tfile):
        try:
            t = json.load(open(tfile)).get(precision, {}).get(name, {})
            traffic = t.get("bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    return dict(kernel=name, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                algorithmic_bytes_per_launch=by / max(nl, 1), avg_launch_us=round(per_launch_s * 1e6, 2),
                launches=nl)


def host_threads():
    """CPUs this process may run on (its affinity set), capped by OMP_NUM_THREADS when set: the host
    share given to a one-GPU job."""
    try:
        cpus = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cpus = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        cpus = min(cpus, int(omp))
    return max(cpus, 1)


def cpu_baseline(n, uu, pr, threads, dealias):
    from oracle import binding as ob
    ob.set_threads(threads)
    p = workload_params(n, 0, dealias)
    o = ob.Oracle(p)
    log(f"cpu baseline: oracle set_state at {n}^3 on {threads} threads")
    o.set_state(uu, pr)
    nstep = 3
    t0 = time.perf_counter()
    o.step(nstep)
    dt = time.perf_counter() - t0
    o.close()
    log(f"cpu baseline: {nstep} steps in {dt:.2f} s")
    return dict(value=round(nstep / dt, 5), unit="steps/s", cores=threads, kind="port",
                host_cpus=os.cpu_count(), omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                sample=f"{nstep} RK3 steps of the same {n}^3 Hall-MHD workload from the same initial state "
                       f"(FP64 C oracle, OpenMP, {threads} threads: every CPU of this process's host share)")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--dealias", type=int, choices=[1, 3], default=3,
                    help="3: phase-shift dealiasing (the workload's own); 1: the reference's 2/3 rule")
    ap.add_argument("--no-compare", action="store_true", help="skip the runs with the other dealiasing option")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=200,
                    help="steps of the 96-box ensemble reported under 'ensemble' (0: skip it)")
    ap.add_argument("--output-every", type=int, default=10,
                    help="steps between outputs in the output-staging comparison (0: skip it)")
    ap.add_argument("--no-closure", action="store_true",
                    help="skip the learned-closure experiment reported under 'closure'")
    ap.add_argument("--config2", type=int, default=0, metavar="STEPS",
                    help="also run the 512^3 expanding-box workload for STEPS steps (under 'config2')")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (default: every CPU this process may run on, capped by "
                         "OMP_NUM_THREADS when that is set)")
    ap.add_argument("--compare-steps", type=int, default=100,
                    help="steps after which the mixed run is compared with the FP64 run (dt sequence replayed)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run y = 2x over 1 GiB (1 GiB read + 1 GiB written) for HBM counter calibration")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU path "
              "(the box would shard as slabs with one all-to-all per 3D transform; not built).",
              file=sys.stderr)
        sys.exit(2)
    n = args.n
    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 2.0, out=y)
        torch.cuda.synchronize()
        log("calibration: 3 x elementwise y = 2x over 1 GiB (1 GiB read + 1 GiB write) done")
        del x, y
    from laps_amd.params import LAPS_FP64, LAPS_MIXED
    log(f"building {n}^3 seeded initial condition")
    uu, pr = initial_state(n, 5.0 / 3.0)
    dl = args.dealias

    def runs(dealias):
        res = {}
        if args.precision in ("fp64", "both"):
            res["fp64"] = run_gpu(n, LAPS_FP64, args.steps, args.warmup, uu, pr, dealias)
        if args.precision in ("mixed", "both"):
            res["mixed"] = run_gpu(n, LAPS_MIXED, args.steps, args.warmup, uu, pr, dealias)
        return res

    def compare(a, b, dealias):
        rel = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
        ea, eb = shell_spectrum(a, n), shell_spectrum(b, n)
        kmax = int(n * (math.sqrt(2.0) / 3.0 if dealias == 3 else 1.0 / 3.0))
        sel = np.arange(len(eb)) <= kmax
        sel &= eb > 1e-10 * eb[1:].max()      # shells holding energy above round-off level
        sel[0] = False
        return rel, float(np.max(np.abs(ea[sel] - eb[sel]) / eb[sel]))

    def mixed_report(res, sps, dealias):
        m, f = res["mixed"], res["fp64"]
        nstep = args.compare_steps
        ref, spec, dts = dt_replay(n, nstep, uu, pr, dealias)
        rel, dev = compare(spec, ref, dealias)
        rel_free, _ = compare(m["spectral"], f["spectral"], dealias)
        msps = args.steps / m["seconds"]
        return dict(value=round(msps, 3), unit="steps/s", speedup_vs_fp64=round(msps / sps, 3),
                    rel_l2_vs_fp64=rel, steps_compared=nstep, max_shell_energy_dev=dev,
                    compared_at="equal time levels: the mixed run replays the FP64 run's CFL dt sequence",
                    fp64_stepwise_identical=(bool(np.array_equal(ref, f["spectral"]))
                                             if nstep == 2 * args.steps + args.warmup else None),
                    free_running=dict(rel_l2_vs_fp64=rel_free, time_fp64=f["scalars"]["time"],
                                      time_mixed=m["scalars"]["time"]),
                    roofline=roofline(m["timing"], "mixed" if dealias == 1 else "mixed_ps"),
                    ms_per_step=round(1000.0 * m["seconds"] / args.steps, 4), kernels=kernels(m))

    def kernels(r):
        return {k: dict(ms_per_step=round(v[0] / args.steps, 4), gbs=round(v[1] / (v[0] * 1e-3) / 1e9, 1)
                        if v[0] > 0 else 0.0, launches=v[2]) for k, v in r["timing"].items() if v[2]}

    dname = {3: "phase-shift dealiasing (half-cell shifted grid, open sqrt(2)/3 ellipsoid)",
             1: "2/3-rule dealiasing (the reference's option 1)"}
    res = runs(dl)
    main_key = "fp64" if "fp64" in res else "mixed"
    r = res[main_key]
    sps = args.steps / r["seconds"]
    tkey = main_key if dl == 1 else main_key + "_ps"
    out = {
        "metric": "RK3 steps/s at 256^3 Hall-MHD, FP64 vs mixed; % of HBM roofline",
        "value": round(sps, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * r["seconds"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_key == "fp64" else "f32/f64",
        "data": "synthetic (seeded k^-5/3 shells 1-16 initial condition)",
        "config": {"workload": f"compressible Hall-MHD RK3 step, {n}^3 periodic, d_i=0.05, nu=eta=2e-4, "
                               f"{dname[dl]}, CFL 0.2",
                   "grid": [n, n, n], "precision": main_key, "dealias_option": dl},
        "roofline": roofline(r["timing"], tkey),
        "kernels": kernels(r),
    }
    if "fp64" in res and "mixed" in res:
        out["mixed"] = mixed_report(res, sps, dl)
    if not args.no_compare and "fp64" in res:
        other = 1 if dl == 3 else 3
        ro = runs(other)
        so = args.steps / ro["fp64"]["seconds"]
        rep = dict(dealias_option=other, value=round(so, 3), unit="steps/s",
                   ms_per_step=round(1000.0 * ro["fp64"]["seconds"] / args.steps, 4),
                   roofline=roofline(ro["fp64"]["timing"], "fp64" if other == 1 else "fp64_ps"),
                   kernels=kernels(ro["fp64"]))
        if "mixed" in ro:
            rep["mixed"] = mixed_report(ro, so, other)
        for v in ro.values():
            v.pop("spectral", None)
        out["dealias_2_3" if other == 1 else "dealias_phase_shift"] = rep
    if args.output_every > 0:
        out["output_staging"] = run_output_staging(n, uu, pr, dl, every=args.output_every)
    if args.ensemble_steps > 0:
        out["ensemble"] = run_ensemble(args.ensemble_steps)
        m23 = out.get("dealias_2_3", {}).get("mixed") if dl == 3 else out.get("mixed")
        if m23:
            # the same precision and dealiasing on one 256^3 box: grid-point steps per second
            single = m23["value"] * float(n) ** 3
            out["ensemble"]["single_box_grid_point_steps_per_s"] = round(single, 1)
            out["ensemble"]["vs_single_box"] = round(out["ensemble"]["grid_point_steps_per_s"] / single, 3)
    if not args.no_closure:
        from laps_amd.closure import experiment
        out["closure"] = experiment(log=log)
    if args.config2 > 0:
        out["config2"] = run_config2(args.config2)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_threads()
        try:
            out["cpu_baseline"] = cpu_baseline(n, uu, pr, threads, dl)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e}")
            out["cpu_baseline"] = None
    for v in res.values():
        v.pop("spectral", None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

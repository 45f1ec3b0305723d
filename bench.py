"""Benchmark of the real-time lattice update loop on one MI355X.

Workload: SU(3) 192^3, links exp(i eps^a lambda^a/2) with eps^a ~ N(0, 0.15^2)
and E^a ~ N(0, 0.25^2) Gauss-projected (seed 5005); leapfrog dt = 0.04; every
50 leapfrog steps the hot links are cooled on a copy by 200 gradient-flow
steps (dtau = 0.02) and Delta N_CS between consecutive cooled configurations
is measured on their geodesic (Simpson 1-4-1 of the improved E.B).

One bench step is one such measurement interval: 50 leapfrog steps, one
200-step cooling episode and the geodesic Delta N_CS.  The timed job is the
initial cooling (CoolingMethod::Start) and --steps intervals, so the default
20 steps are the whole 1000-step trajectory: 1000 leapfrog and 4200 flow link
updates per link.  The FP32 form (FP32 storage and arithmetic, FP64
reunitarisation and FP64 sums) is timed first, then the FP64 form on the same
initial condition.

value = link updates per second of the FP32 form over the whole timed job
(leapfrog + flow link updates / wall time).  One JSON line on stdout.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_link_update(nc, kind, fp32):
    """Algorithmic HBM bytes per link update: each link and E read once and
    written once for leapfrog/Langevin, links only for the fused flow."""
    su, se = (4, 3) if nc == 2 else (18, 8)
    w = 4 if fp32 else 8
    if kind in ("leapfrog", "langevin"):
        return 2 * su * w + 2 * se * w
    if kind == "flow":
        return 2 * su * w
    raise ValueError(kind)


def measured_traffic(nc, N, kind, fp32):
    """HBM bytes per launch of the dominant kernel, as measured by the FETCH_SIZE
    and WRITE_SIZE counters (calibrated against a copy kernel) in a separate
    profiler run whose summary is kept in profiles/; None if none matches."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None, None
    for f in sorted(os.listdir(d)):
        if not f.endswith("_traffic.json"):
            continue
        try:
            with open(os.path.join(d, f)) as fh:
                t = json.load(fh)
        except (OSError, ValueError):
            continue
        if (t.get("kernel"), t.get("dtype"), t.get("nc"), t.get("N")) == (kind, "f32" if fp32 else "f64", nc, N):
            return t["traffic_bytes"], "profiles/" + t.get("source", f)
    return None, None


def measured_flops(nc, N, kind, form):
    """FP32 and FP64 VALU flops per launch of the dominant kernel, from the
    SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F{32,64} counters of a separate profiler
    run whose summary is kept in profiles/; None if none matches."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None, None
    for f in sorted(os.listdir(d)):
        if not f.endswith("_valu.json"):
            continue
        try:
            with open(os.path.join(d, f)) as fh:
                t = json.load(fh)
        except (OSError, ValueError):
            continue
        if (t.get("kernel"), t.get("dtype"), t.get("nc"), t.get("N"), t.get("form")) == (kind, "f32", nc, N, form):
            return t["flop_f32"] + t["flop_f64"], "profiles/" + t.get("source", f)
    return None, None


def path_roofline(nc, V, r):
    """The update kernels of the timed job together (every leapfrog and flow
    launch): algorithmic bytes over their summed HIP-event time, and over the
    job's wall time (measurements included)."""
    tot_b, tot_s = 0.0, 0.0
    for kind, n in (("leapfrog", r["n_lf"]), ("flow", r["n_flow"])):
        ms, launches = r["stats"][kind]
        tot_b += 3 * V * bytes_per_link_update(nc, kind, True) * launches
        tot_s += ms / 1e3
    if tot_s <= 0:
        return None
    job_b = 3 * V * (bytes_per_link_update(nc, "leapfrog", True) * r["n_lf"]
                     + bytes_per_link_update(nc, "flow", True) * r["n_flow"])
    return {"achieved": tot_b / tot_s / 1e9, "frac": tot_b / tot_s / 1e9 / HBM_PEAK_GBS, "unit": "GB/s",
            "kernel_time_s": tot_s, "job_achieved": job_b / r["wall"] / 1e9,
            "job_frac": job_b / r["wall"] / 1e9 / HBM_PEAK_GBS,
            "basis": "algorithmic bytes of every leapfrog and flow launch over their summed kernel time "
                     "(job_*: over the wall time of the timed job, measurements included)"}


def make_initial_condition(T, nc, N, seed, su, se, maxiter):
    g = T.Solver(nc, N, T.FP64)
    g.random_links(seed, su)
    g.random_efield(seed, se)
    t0 = time.time()
    # GaussRestore in chunks of 1000 iterations so that the run reports its
    # progress; a chunk restarts from the current field, so the stop can come
    # at most one iteration earlier than in a single call
    it, fm = 0, float("inf")
    while it < maxiter:
        n, fm = g.gauss_law_restore(maxiter=min(1000, maxiter - it), tol=1e-14, check=False)
        it += n
        log("# GaussRestore: %d iterations, signed max %.3e (%.1f s)" % (it, fm, time.time() - t0))
        if fm <= 1e-14:
            break
    log("# initial condition: GaussRestore %d iterations, signed max %.3e (%.1f s)" % (it, fm, time.time() - t0))
    U, E = g.get_links(), g.get_efield()
    g.close()
    return U, E, it, fm


def run_form(T, nc, N, precision, U, E, steps, warmup, dt, cool_every, cool_steps, dtau, timing, form=0):
    """`steps` measurement intervals (cool_every leapfrog steps, a cool_steps
    cooling episode, the geodesic Delta N_CS and the energy) after the initial
    cooling, all timed; `warmup` untimed intervals run first on copies.
    `form`: the update kernel form (T.FORM_AUTO or T.FORM_LINK)."""
    g = T.Solver(nc, N, precision)
    g.set_update_form(form)
    g.set_links(U)
    g.set_efield(E)
    # warm-up on copies of the fields: the same calls as a measurement interval
    g.copy_links(T.U_AUX, T.U_HOT)
    g.copy_efield(T.E_AUX, T.E_HOT)
    for _ in range(warmup):
        g.dynamics_update(cool_every, dt, u=T.U_AUX, e=T.E_AUX)
        g.cool_n_save(cool_steps, dtau, u_hot=T.U_AUX, u_cold=T.U_COOL)
        g.ncs_dot(u=T.U_AUX, e=T.E_AUX)
        g.energy(u=T.U_AUX, e=T.E_AUX)
    g.sync()
    if timing:
        g.timing(True)
        g.timing_reset()
    energies, dncs_series = [], []
    n_lf = n_flow = 0
    ncs = 0.0
    g.sync()
    t0 = time.perf_counter()
    g.cooling_method_start(cool_steps, dtau)
    n_flow += cool_steps
    for k in range(1, steps + 1):
        # the leapfrog steps of one interval in one call, then the measurement
        g.dynamics_update(cool_every, dt)
        n_lf += cool_every
        d = g.cooling_method_update(cool_steps, dtau, measure=True)
        n_flow += cool_steps
        ncs += d
        dncs_series.append(ncs)
        energies.append(g.energy())
        if k % 4 == 0:
            log("#   interval %d (leapfrog step %d): T00 %.10e  Delta N_CS %.6e  (%.1f s)"
                % (k, n_lf, energies[-1], ncs, time.perf_counter() - t0))
    g.sync()
    wall = time.perf_counter() - t0
    stats, forms = {}, {}
    if timing:
        for name in ("leapfrog", "flow", "edotb", "geodesic", "copy", "bulk", "reduce"):
            stats[name] = g.kernel_time(name)
        # the update kernels of the plaquette form are timed under their own names
        for name in ("leapfrog", "flow"):
            ms, n = g.kernel_time(name + "_plaq")
            forms[name] = "plaquette" if n else "link"
            if n:
                stats[name] = (stats[name][0] + ms, stats[name][1] + n)
        g.timing(False)
    umax = g.unitarity_violation()[0]
    gauss = g.gauss_violation()
    g.close()
    return dict(wall=wall, n_lf=n_lf, n_flow=n_flow, energies=energies, ncs=dncs_series,
                stats=stats, forms=forms, unitarity=umax, gauss=gauss)


def batched_sizes(count, seed):
    """Lattice extents L^3, L drawn uniformly from {16, 24, 32, 40, 48} by a
    Philox generator with the given seed (numpy's Philox-4x64)."""
    rng = np.random.Generator(np.random.Philox(seed))
    return [int(L) for L in rng.choice([16, 24, 32, 40, 48], size=count)]


def run_batched(T, count, seed, precision, lang_steps, seg, cool_every, cool_steps, dt, dtau, damping, temp, tune):
    """The batching form: `count` independent SU(2) lattices of uneven sizes in
    one batch (one persistent launch per pass over all of them); Langevin
    steps in graph-replayed segments of `seg` steps, each ending with NCsDot of
    every lattice, and a cooled Delta N_CS of every lattice every `cool_every`
    steps (`cool_steps` flow steps).  Returns the timings and trajectories."""
    sizes = batched_sizes(count, seed)
    b = T.Batch(2, sizes, precision)
    t0 = time.perf_counter()
    iters = b.initial_condition(seed, 0.2, 0.3, maxiter=10000, tol=1e-14)
    t_ic = time.perf_counter() - t0
    choice, best_us, fchoice, fbest_us = None, None, None, None
    if tune:
        choice, best_us = b.tune(mode=1, dt=dt, reps=3)
        fchoice, fbest_us = b.tune(mode=3, dt=dtau, reps=3)
    # warm-up: one segment on the same path (the graph is captured here), then restart the fields
    links = [b.get_links(e) for e in range(count)]
    efield = [b.get_efield(e) for e in range(count)]
    b.evolve(1, 1, seg, dt, damping, temp, seed + 1000)
    b.cooling_start(2, dtau)
    for e in range(count):
        b.set_links(e, links[e])
        b.set_efield(e, efield[e])
    b.set_step(0)
    b.sync()
    b.timing(True)
    b.timing_reset()
    ncsdot, dncs = [], []
    t0 = time.perf_counter()
    b.cooling_start(cool_steps, dtau)
    n_flow = cool_steps
    done = 0
    while done < lang_steps:
        nseg = min(cool_every, lang_steps - done) // seg
        ncsdot.append(b.evolve(1, nseg, seg, dt, damping, temp, seed + 1000))
        done += nseg * seg
        dncs.append(b.cooling_update(cool_steps, dtau))
        n_flow += cool_steps
    b.sync()
    wall = time.perf_counter() - t0
    kern = {}
    for name in ("batch_langevin", "batch_langevin_plaq", "batch_flow", "batch_flow_plaq", "batch_edotb",
                 "batch_geodesic", "batch_reduce", "copy"):
        ms, n = b.kernel_time(name)
        if n:
            kern[name] = {"launches": n, "avg_us": 1e3 * ms / n}
    b.timing(False)
    un = b.unitarity()
    ga = b.gauss()
    b.close()
    links_total = sum(3 * L ** 3 for L in sizes)
    return dict(sizes=sizes, wall=wall, t_ic=t_ic, iters=iters, choice=choice, best_us=best_us, fchoice=fchoice,
                fbest_us=fbest_us,
                lu=links_total * (lang_steps + n_flow) / wall, links=links_total, n_lang=lang_steps, n_flow=n_flow,
                ncsdot=np.concatenate(ncsdot, axis=1), dncs=np.cumsum(np.array(dncs).T, axis=1), kernels=kern,
                unitarity=float(un[:, 0].max()), gauss=float(np.abs(ga[:, 0]).max()))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_sample(O, nc, N, U, E, dt, dtau, budget_s, threads):
    """leapfrog then flow steps of the oracle, each kind until about budget_s/2 s
    are spent; the rate weights them as the timed job does (1000 : 4200)"""
    lat = O.Lattice(nc, N, nthreads=threads)
    lat.U[:] = U
    lat.E[:] = E
    n_lf = n_fl = 0
    t0 = time.perf_counter()
    while True:
        lat.leapfrog(1, dt)
        n_lf += 1
        if time.perf_counter() - t0 > budget_s / 2:
            break
    t1 = time.perf_counter()
    while True:
        lat.flow(1, dtau)
        n_fl += 1
        if time.perf_counter() - t1 > budget_s / 2:
            break
    t2 = time.perf_counter()
    V = N[0] * N[1] * N[2]
    t_lf, t_fl = (t1 - t0) / n_lf, (t2 - t1) / n_fl
    return dict(n_lf=n_lf, n_fl=n_fl, wall=t2 - t0, lf=3 * V / t_lf, fl=3 * V / t_fl,
                mix=3 * V * 5200 / (1000 * t_lf + 4200 * t_fl))


def cpu_baseline(nc, N, U, E, dt, dtau, budget_s):
    """The oracle (FP64 CPU restatement) on the host cores, and on one core:
    leapfrog and flow steps on the same lattice for about budget_s seconds
    each run; the rate is that of the job's 1000 : 4200 leapfrog : flow mix."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    a = cpu_sample(O, nc, N, U, E, dt, dtau, budget_s, threads)
    one = cpu_sample(O, nc, N, U, E, dt, dtau, budget_s, 1)
    return dict(value=a["mix"], unit="link-updates/s", cores=threads, kind="port", cpu=cpu_model(),
                sample="%d leapfrog + %d flow steps of the oracle on SU(%d) %dx%dx%d (%.1f s, FP64, OpenMP, "
                       "%d threads), rate weighted 1000 : 4200 as the timed job"
                       % (a["n_lf"], a["n_fl"], nc, N[0], N[1], N[2], a["wall"], threads),
                leapfrog_lu_s=a["lf"], flow_lu_s=a["fl"],
                single_thread={"value": one["mix"], "leapfrog_lu_s": one["lf"], "flow_lu_s": one["fl"],
                               "sample": "%d leapfrog + %d flow steps on 1 thread (%.1f s)"
                                         % (one["n_lf"], one["n_fl"], one["wall"])})


SMALL_WORKLOADS = {
    "su2-16": dict(nc=2, L=16, seed=1001, su=0.2, se=0.3, forms=("fp64",), phases=[("leapfrog", 200, 0.05)],
            text="SU(2) 16^3 FP64, seeded links (sigma 0.2) and E (sigma 0.3) Gauss-projected; 200 leapfrog steps "
                 "dt=0.05 with the hot N_CS integral (E.B at every step) on the device"),
    "su2-64": dict(nc=2, L=64, seed=2002, su=0.2, se=0.3, forms=("fp32", "fp64"),
            phases=[("leapfrog", 1000, 0.05), ("langevin", 1000, 0.05)],
            text="SU(2) 64^3 FP64 vs FP32-mixed, seeded IC as the SU(2) 16^3 job; 1000 leapfrog steps dt=0.05 then 1000 Langevin "
                 "steps (damping 0.1, T 0.5, Philox noise), hot N_CS integral at every step, energy every 100 steps"),
    "su3-96": dict(nc=3, L=96, seed=3003, su=0.15, se=0.25, forms=("fp32", "fp64"),
            phases=[("leapfrog", 500, 0.04), ("flow", 500, 0.02)],
            text="SU(3) 96^3 FP64 vs FP32-mixed, links sigma 0.15, E sigma 0.25 Gauss-projected; 500 leapfrog steps "
                 "dt=0.04 (hot N_CS integral) then 500 gradient-flow steps dtau=0.02 on a copy"),
}


def run_small(T, name, prec, U, E):
    """One form of a small configuration: the phases in calls of 100 steps,
    the energy after each call; the wall time of the steps only."""
    c = SMALL_WORKLOADS[name]
    N = (c["L"],) * 3
    g = T.Solver(c["nc"], N, prec)
    g.set_links(U)
    g.set_efield(E)
    g.sync()
    ncs, energy, steps = [0.0], [g.energy()], 0
    wall = 0.0
    for kind, n, dt in c["phases"]:
        if kind == "flow":
            g.copy_links(T.U_COOL, T.U_HOT)
        done = 0
        while done < n:
            k = min(100, n - done)
            g.sync()
            t0 = time.perf_counter()
            if kind == "leapfrog":
                d = g.evolve_ncs(0, k, dt)
            elif kind == "langevin":
                d = g.evolve_ncs(1, k, dt, damping=0.1, T=0.5, seed=77, step0=done)
            else:
                g.gradient_flow(k, dt, u=T.U_COOL, e=T.E_COOL)
                d = 0.0
            g.sync()
            wall += time.perf_counter() - t0
            done += k
            steps += k
            if kind != "flow":
                ncs.append(ncs[-1] + d)
                energy.append(g.energy())
    out = dict(wall=wall, lu=3 * g.V * steps / wall, ncs=ncs, energy=energy, steps=steps,
               gauss=g.gauss_violation()[0], unitarity=g.unitarity_violation()[0])
    g.close()
    return out


def run_workload(T, name, cpu_seconds):
    """The smaller jobs of the path (SU(2) 16^3 and 64^3, SU(3) 96^3), timed:
    one JSON line with the same fields as the default workload's."""
    c = SMALL_WORKLOADS[name]
    N = (c["L"],) * 3
    log("# %s: %s" % (name, c["text"]))
    U, E, it, fm = make_initial_condition(T, c["nc"], N, c["seed"], c["su"], c["se"], 10000)
    res = {}
    for f in c["forms"]:
        run_small(T, name, T.FP32 if f == "fp32" else T.FP64, U, E)      # warm-up run
        res[f] = run_small(T, name, T.FP32 if f == "fp32" else T.FP64, U, E)
        log("# %s: %.3f s, %.3e link-updates/s" % (f, res[f]["wall"], res[f]["lu"]))
    main_form = "fp32" if "fp32" in res else "fp64"
    r = res[main_form]
    out = {
        "metric": "link-updates/s + HBM GB/s, FP32-mixed vs FP64 GPU at |dN_CS|<=1e-3",
        "value": r["lu"], "unit": "link-updates/s", "n_gpus": 1, "steps": r["steps"], "warmup": 1,
        "ms_per_step": 1e3 * r["wall"] / r["steps"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if main_form == "fp32" else "f64",
        "data": "synthetic (seeded Philox initial condition, Gauss-projected)",
        "config": {"workload": c["text"], "gauss_restore_iterations": it,
                   "gauss_restore_signed_max": fm},
    }
    for f, x in res.items():
        out[f] = {"wall_s": x["wall"], "link_updates_per_s": x["lu"], "gauss_signed_max": x["gauss"],
                  "unitarity_max": x["unitarity"], "hot_ncs_final": x["ncs"][-1]}
    if "fp32" in res and "fp64" in res:
        a, b = res["fp32"], res["fp64"]
        out["fp32_vs_fp64"] = {"speedup": a["lu"] / b["lu"],
                               "max_abs_hot_ncs_diff": max(abs(x - y) for x, y in zip(a["ncs"], b["ncs"])),
                               "max_rel_energy_diff": max(abs(x - y) / abs(y) for x, y in zip(a["energy"], b["energy"]))}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        smp = cpu_sample(O, c["nc"], N, U, E, c["phases"][0][2], 0.02, cpu_seconds, threads)
        rate = smp["lf"] if name != "su3-96" else 1000.0 / (500.0 / smp["lf"] + 500.0 / smp["fl"])
        out["cpu_baseline"] = {"value": rate, "unit": "link-updates/s", "cores": threads, "kind": "port",
                               "cpu": cpu_model(),
                               "sample": "%d leapfrog + %d flow steps of the oracle (FP64, OpenMP, %d threads, %.1f s)%s"
                                         % (smp["n_lf"], smp["n_fl"], threads, smp["wall"],
                                            ", rate weighted 500 : 500" if name == "su3-96" else ", leapfrog rate")}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="measurement intervals (50 leapfrog steps + one cooling each)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed measurement intervals, on copies of the fields")
    ap.add_argument("--N", type=int, default=192)
    ap.add_argument("--nc", type=int, default=3)
    ap.add_argument("--dt", type=float, default=0.04)
    ap.add_argument("--dtau", type=float, default=0.02)
    ap.add_argument("--cool-every", type=int, default=50)
    ap.add_argument("--cool-steps", type=int, default=200)
    ap.add_argument("--seed", type=int, default=5005)
    ap.add_argument("--restore-maxiter", type=int, default=10000)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 form")
    ap.add_argument("--no-link-form", action="store_true",
                    help="skip the FP32 run with the per-link update kernels (the comparison of the two forms)")
    ap.add_argument("--cpu-seconds", type=float, default=16.0, help="CPU baseline budget per run (0 = skip)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-ensembles workload")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--workload", default="su3-192", choices=("su3-192", "su2-16", "su2-64", "su3-96"),
                    help="su3-192: the SU(3) 192^3 leapfrog + cooling job (default; with the batched SU(2) "
                         "ensembles); the others: the smaller jobs of the path, timed alone")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import topologicalcymsolver_amd as T
    if args.workload != "su3-192":
        run_workload(T, args.workload, args.cpu_seconds)
        return
    nc, N = args.nc, (args.N, args.N, args.N)
    V = N[0] * N[1] * N[2]
    log("# SU(%d) %d^3, %d intervals of %d leapfrog steps and a %d-step cooling (+%d warm-up intervals)" %
        (nc, args.N, args.steps, args.cool_every, args.cool_steps, args.warmup))
    U, E, it, fm = make_initial_condition(T, nc, N, args.seed, 0.15, 0.25, args.restore_maxiter)

    r32 = run_form(T, nc, N, T.FP32, U, E, args.steps, args.warmup, args.dt, args.cool_every,
                   args.cool_steps, args.dtau, timing=True)
    lu32 = 3 * V * (r32["n_lf"] + r32["n_flow"]) / r32["wall"]
    log("# FP32 form: %.3f s, %.3e link-updates/s (update kernels: %s)" % (r32["wall"], lu32, r32["forms"]))
    r32l = None
    if not args.no_link_form and "plaquette" in r32["forms"].values():
        r32l = run_form(T, nc, N, T.FP32, U, E, args.steps, args.warmup, args.dt, args.cool_every,
                        args.cool_steps, args.dtau, timing=True, form=T.FORM_LINK)
        log("# FP32 form, per-link update kernels: %.3f s, %.3e link-updates/s"
            % (r32l["wall"], 3 * V * (r32l["n_lf"] + r32l["n_flow"]) / r32l["wall"]))
    r64 = None
    if not args.no_fp64:
        r64 = run_form(T, nc, N, T.FP64, U, E, args.steps, args.warmup, args.dt, args.cool_every,
                       args.cool_steps, args.dtau, timing=True)
        lu64 = 3 * V * (r64["n_lf"] + r64["n_flow"]) / r64["wall"]
        log("# FP64 form: %.3f s, %.3e link-updates/s" % (r64["wall"], lu64))

    # roofline of the dominant kernel of the FP32 form (HIP events on the context's stream)
    kinds = {"leapfrog": r32["stats"]["leapfrog"], "flow": r32["stats"]["flow"]}
    dom = max(kinds, key=lambda k: kinds[k][0])
    tot_ms, launches = kinds[dom]
    avg_s = tot_ms / max(launches, 1) / 1e3
    bpl = bytes_per_link_update(nc, dom, True)
    achieved = 3 * V * bpl / avg_s / 1e9
    traffic, traffic_src = measured_traffic(nc, args.N, dom, True)
    flops, flops_src = measured_flops(nc, args.N, dom, r32["forms"].get(dom))
    kernels = {}
    for k, (ms, n) in r32["stats"].items():
        if n:
            kernels[k] = {"launches": n, "avg_us": 1e3 * ms / n}
    for fk, rr in (("fp32", r32), ("fp32_link_form", r32l), ("fp64", r64)):
        if rr is None:
            continue
        for kk in ("leapfrog", "flow"):
            ms, n = rr["stats"][kk]
            if n:
                b = bytes_per_link_update(nc, kk, fk != "fp64")
                kernels["%s_%s" % (kk, fk)] = {"avg_us": 1e3 * ms / n, "launches": n, "form": rr["forms"].get(kk),
                                               "GB/s": 3 * V * b / (ms / n / 1e3) / 1e9,
                                               "LU/s": 3 * V / (ms / n / 1e3)}

    out = {
        "metric": "link-updates/s + HBM GB/s, FP32-mixed vs FP64 GPU at |dN_CS|<=1e-3",
        "value": lu32,
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r32["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox initial condition, Gauss-projected)",
        "config": {
            "workload": "SU(%d) %d^3 FP32-mixed leapfrog dt=%g, %d-step gradient-flow cooling dtau=%g every %d steps, "
                        "geodesic Delta N_CS" % (nc, args.N, args.dt, args.cool_steps, args.dtau, args.cool_every),
            "step": "one measurement interval: %d leapfrog steps + a %d-step cooling + geodesic Delta N_CS"
                    % (args.cool_every, args.cool_steps),
            "leapfrog_steps": r32["n_lf"], "flow_steps": r32["n_flow"],
            "gauss_restore_iterations": it, "gauss_restore_signed_max": fm,
        },
        "roofline": {
            "bound": "hbm", "kernel": "%s (%s form)" % (dom, r32["forms"].get(dom)), "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_algorithmic": 3 * V * bpl, "traffic_source": traffic_src,
            "bytes_per_link_update": bpl, "avg_launch_us": avg_s * 1e6,
            "frac_measured_traffic": (traffic / avg_s / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "bytes_basis": "algorithmic: every link read once and written once per step (2 s_U w = %d B); between the "
                           "steps of one call the kernel moves columns 0 and 1 of the links only, so the measured "
                           "traffic is lower" % bpl,
            "valu": None if flops is None else {
                "achieved": flops / avg_s / 1e12, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": flops / avg_s / 1e12 / VALU_PEAK_TFLOPS, "flop_per_launch": flops, "source": flops_src,
                "basis": "FP32 + FP64 VALU flops per launch (SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F{32,64} counters, "
                         "FMA = 2) over the live average launch time; peak = non-packed vector math, 256 CUs x 4 "
                         "SIMDs x 16 lanes x 2 x 2.4 GHz (packed FP32 is disabled in this build)"},
        },
        "kernels": kernels,
        "roofline_update_path": path_roofline(nc, V, r32),
        "fp32": {"wall_s": r32["wall"], "link_updates_per_s": lu32, "unitarity_max": r32["unitarity"],
                 "gauss_signed_max": r32["gauss"][0], "gauss_rms": r32["gauss"][1],
                 "delta_ncs": r32["ncs"][-1] if r32["ncs"] else None},
    }
    if r32l is not None:
        lul = 3 * V * (r32l["n_lf"] + r32l["n_flow"]) / r32l["wall"]
        out["fp32_link_form"] = {
            "wall_s": r32l["wall"], "link_updates_per_s": lul, "plaquette_form_speedup": lu32 / lul,
            "max_abs_delta_ncs_diff_vs_plaquette_form": max(abs(a - b) for a, b in zip(r32["ncs"], r32l["ncs"])),
            "max_rel_energy_diff_vs_plaquette_form": max(abs(a - b) / abs(b) for a, b in zip(r32["energies"],
                                                                                           r32l["energies"])),
        }
    if r64 is not None:
        e32, e64 = r32["energies"], r64["energies"]
        drift = max(abs(a - b) / abs(b) for a, b in zip(e32, e64)) if e64 else None
        dn = max(abs(a - b) for a, b in zip(r32["ncs"], r64["ncs"])) if r64["ncs"] else None
        out["fp64"] = {"wall_s": r64["wall"], "link_updates_per_s": lu64, "unitarity_max": r64["unitarity"],
                       "gauss_signed_max": r64["gauss"][0], "gauss_rms": r64["gauss"][1],
                       "delta_ncs": r64["ncs"][-1] if r64["ncs"] else None}
        out["fp32_vs_fp64"] = {"speedup": lu32 / lu64, "max_abs_delta_ncs_diff": dn,
                               "max_rel_energy_diff": drift}
    if not args.no_batched:
        log("# batched: %d SU(2) lattices, %d Langevin steps ..." % (args.batch_count, args.batch_steps))
        bkw = dict(lang_steps=args.batch_steps, seg=10, cool_every=100, cool_steps=200, dt=0.05, dtau=0.02,
                   damping=0.1, temp=0.5)
        b32 = run_batched(T, args.batch_count, 4004, T.FP32, tune=True, **bkw)
        log("# batched FP32: %.3f s, %.3e link-updates/s, tuner %s / flow %s" % (b32["wall"], b32["lu"], b32["choice"],
                                                                                   b32["fchoice"]))
        b64 = run_batched(T, args.batch_count, 4004, T.FP64, tune=True, **bkw)
        log("# batched FP64: %.3f s, %.3e link-updates/s" % (b64["wall"], b64["lu"]))
        kd = max(b32["kernels"], key=lambda k: b32["kernels"][k]["avg_us"] * b32["kernels"][k]["launches"])
        kind = {"batch_langevin": "langevin", "batch_flow": "flow", "batch_langevin_plaq": "langevin",
                "batch_flow_plaq": "flow"}.get(kd)
        roof = None
        if kind:
            bpl = bytes_per_link_update(2, kind, True)
            ach = b32["links"] * bpl / (b32["kernels"][kd]["avg_us"] * 1e-6) / 1e9
            roof = {"bound": "hbm", "kernel": kd, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": ach / HBM_PEAK_GBS, "bytes_per_link_update": bpl,
                    "avg_launch_us": b32["kernels"][kd]["avg_us"],
                    "note": "the batch's hot links and E fields (%.0f MB in FP32) fit in the 256 MB Infinity Cache, "
                            "so this kernel can run above the HBM roofline" % (b32["links"] * (16 + 12) / 1e6)}
        out["batched"] = {
            "workload": "%d independent SU(2) lattices L^3, L in {16,24,32,40,48} (counts %s), FP32-mixed; "
                        "%d Langevin steps dt=0.05 damping=0.1 T=0.5 in graph-replayed segments of 10 steps "
                        "ending with NCsDot, 200-step cooling dtau=0.02 every 100 steps (geodesic Delta N_CS)"
                        % (args.batch_count, {L: b32["sizes"].count(L) for L in (16, 24, 32, 40, 48)}, args.batch_steps),
            "link_updates_per_s": b32["lu"], "wall_s": b32["wall"], "links": b32["links"],
            "langevin_steps": b32["n_lang"], "flow_steps": b32["n_flow"],
            "tuner_choice": {"langevin": {"shape": b32["choice"][0], "blocks_per_cu": b32["choice"][1],
                                          "xcd_split": b32["choice"][2], "best_us": b32["best_us"]},
                             "flow": {"shape": b32["fchoice"][0], "blocks_per_cu": b32["fchoice"][1],
                                      "xcd_split": b32["fchoice"][2], "best_us": b32["fbest_us"]}},
            "kernels": b32["kernels"], "roofline": roof,
            "fp64_link_updates_per_s": b64["lu"], "fp32_vs_fp64": b32["lu"] / b64["lu"],
            "max_abs_delta_ncs_diff_vs_fp64": float(np.abs(b32["dncs"] - b64["dncs"]).max()),
            "max_abs_ncsdot_diff_vs_fp64": float(np.abs(b32["ncsdot"] - b64["ncsdot"]).max()),
            "delta_ncs_final": {"min": float(b32["dncs"][:, -1].min()), "max": float(b32["dncs"][:, -1].max())},
            "unitarity_max": b32["unitarity"], "gauss_abs_max": b32["gauss"],
            "gauss_restore_iterations": {"min": int(min(b32["iters"])), "max": int(max(b32["iters"]))},
        }
    if args.cpu_seconds > 0:
        log("# CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(nc, N, U, E, args.dt, args.dtau, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

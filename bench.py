"""Benchmark of the real-time lattice update loop on one MI355X.

Workload (one step = one leapfrog step of the hot lattice):
  SU(3) 192^3, links exp(i eps^a lambda^a/2) with eps^a ~ N(0, 0.15^2) and
  E^a ~ N(0, 0.25^2) Gauss-projected (seed 5005); leapfrog dt = 0.04; every
  50 steps the hot links are cooled on a copy by 200 gradient-flow steps
  (dtau = 0.02) and Delta N_CS between consecutive cooled configurations is
  measured on their geodesic (Simpson 1-4-1 of the improved E.B).
The FP32 form (FP32 storage and arithmetic, FP64 reunitarisation and FP64
sums) is timed first, then the FP64 form on the same initial condition.

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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


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


def run_form(T, nc, N, precision, U, E, steps, warmup, dt, cool_every, cool_steps, dtau, timing):
    g = T.Solver(nc, N, precision)
    g.set_links(U)
    g.set_efield(E)
    # warm-up on copies of the fields: leapfrog and a short cooling episode
    g.copy_links(T.U_AUX, T.U_HOT)
    g.copy_efield(T.E_AUX, T.E_HOT)
    if warmup > 0:
        g.dynamics_update(warmup, dt, u=T.U_AUX, e=T.E_AUX)
        g.cool_n_save(min(warmup, cool_steps), dtau, u_hot=T.U_AUX, u_cold=T.U_COOL)
        g.ncs_dot(u=T.U_AUX, e=T.E_AUX)
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
        g.dynamics_update(1, dt)
        n_lf += 1
        if k % cool_every == 0 or k == steps:
            d = g.cooling_method_update(cool_steps, dtau, measure=True)
            n_flow += cool_steps
            ncs += d
            dncs_series.append(ncs)
            energies.append(g.energy())
            if k % (4*cool_every) == 0:
                log("#   step %d: T00 %.10e  Delta N_CS %.6e  (%.1f s)" % (k, energies[-1], ncs, time.perf_counter() - t0))
    g.sync()
    wall = time.perf_counter() - t0
    stats = {}
    if timing:
        for name in ("leapfrog", "flow", "edotb", "geodesic", "copy", "bulk", "reduce"):
            stats[name] = g.kernel_time(name)
        g.timing(False)
    umax = g.unitarity_violation()[0]
    gauss = g.gauss_violation()
    g.close()
    return dict(wall=wall, n_lf=n_lf, n_flow=n_flow, energies=energies, ncs=dncs_series,
                stats=stats, unitarity=umax, gauss=gauss)


def cpu_baseline(nc, N, U, E, dt, dtau, budget_s):
    """The oracle (FP64 CPU restatement) on the host cores: leapfrog and flow
    steps on the same lattice, until about budget_s seconds are spent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
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
    lu = 3 * V * (n_lf + n_fl)
    return dict(value=lu / (t2 - t0), unit="link-updates/s", cores=threads, kind="port",
                sample="%d leapfrog + %d flow steps of the oracle on SU(%d) %dx%dx%d (%.1f s, FP64, OpenMP)"
                       % (n_lf, n_fl, nc, N[0], N[1], N[2], t2 - t0),
                leapfrog_lu_s=3 * V * n_lf / (t1 - t0), flow_lu_s=3 * V * n_fl / (t2 - t1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--N", type=int, default=192)
    ap.add_argument("--nc", type=int, default=3)
    ap.add_argument("--dt", type=float, default=0.04)
    ap.add_argument("--dtau", type=float, default=0.02)
    ap.add_argument("--cool-every", type=int, default=50)
    ap.add_argument("--cool-steps", type=int, default=200)
    ap.add_argument("--seed", type=int, default=5005)
    ap.add_argument("--restore-maxiter", type=int, default=10000)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 form")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline budget (0 = skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import topologicalcymsolver_amd as T
    nc, N = args.nc, (args.N, args.N, args.N)
    V = N[0] * N[1] * N[2]
    log("# SU(%d) %d^3, %d steps (+%d warm-up), cooling %d flow steps every %d" %
        (nc, args.N, args.steps, args.warmup, args.cool_steps, args.cool_every))
    U, E, it, fm = make_initial_condition(T, nc, N, args.seed, 0.15, 0.25, args.restore_maxiter)

    r32 = run_form(T, nc, N, T.FP32, U, E, args.steps, args.warmup, args.dt, args.cool_every,
                   args.cool_steps, args.dtau, timing=True)
    lu32 = 3 * V * (r32["n_lf"] + r32["n_flow"]) / r32["wall"]
    log("# FP32 form: %.3f s, %.3e link-updates/s" % (r32["wall"], lu32))
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
    kernels = {}
    for k, (ms, n) in r32["stats"].items():
        if n:
            kernels[k] = {"launches": n, "avg_us": 1e3 * ms / n}
    for fk, rr in (("fp32", r32), ("fp64", r64)):
        if rr is None:
            continue
        for kk in ("leapfrog", "flow"):
            ms, n = rr["stats"][kk]
            if n:
                b = bytes_per_link_update(nc, kk, fk == "fp32")
                kernels["%s_%s" % (kk, fk)] = {"avg_us": 1e3 * ms / n, "launches": n,
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
            "leapfrog_steps": r32["n_lf"], "flow_steps": r32["n_flow"],
            "gauss_restore_iterations": it, "gauss_restore_signed_max": fm,
        },
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_algorithmic": 3 * V * bpl, "traffic_source": traffic_src,
            "bytes_per_link_update": bpl, "avg_launch_us": avg_s * 1e6,
        },
        "kernels": kernels,
        "fp32": {"wall_s": r32["wall"], "link_updates_per_s": lu32, "unitarity_max": r32["unitarity"],
                 "gauss_signed_max": r32["gauss"][0], "gauss_rms": r32["gauss"][1],
                 "delta_ncs": r32["ncs"][-1] if r32["ncs"] else None},
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
    if args.cpu_seconds > 0:
        log("# CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(nc, N, U, E, args.dt, args.dtau, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark: wall time for the charge Q_8(t=10) to reach 1e-3 relative error
on one MI355X, learned flow warper vs plain warper.

Workload (the configuration the metric is quoted on): single-orbital Anderson
impurity, flat-band bath Gamma=1, T=0, V=0, eps_d=0, g0 tabulated analytically
on 10^4 (->10001) time points in [-10, 10]; observable compute_charge_Q_direct
at time 10, order 8.  Estimator: randomised QMC -- each step is one replica of
`--points` scrambled Sobol points (rng_seed_shift = replica index) pushed
through the warper chain and the Keldysh-sum kernel on the device.

Plain warper chain (the reference's fitted product-1D recipe):
    [plasma_uv(10), product_1d_simple_inverse(10, 1.0),
     make_warper_from_proj(order 8, 500 bins, 2^14 pilot points, sigma 0.1)]

Learned estimator: a 6-layer affine-coupling flow (hidden width 64, bf16 MFMA
products) appended to the plain chain (applied first to the Sobol point),
trained on the device for 2000 Adam steps of 2^14 points on 2^20 samples of
the plain warped integrand (scramble seed 1000003, disjoint from the
replicas), toward the density |f|^alpha for alpha in 1, 1/2, 1/4.  The plain
estimator and the candidates are scored on 128 pilot replicas of 2^16 points
(seeds 20000+) by spread^2 x seconds per point; the best flow is measured, and
it is the selected estimator only if its score is below 0.8 x the plain one's
(plain is the fallback).

For each estimator: time to 1e-3 = (points needed for 1e-3 relative error) /
(points per second).  The points needed come from a doubling search: 256
(--target-replicas) scrambled replicas at M points, M = 2^12, 2^13, ... until the relative spread
of one replica's estimate is <= 1e-3, then log-log interpolation between the
last two M.  Points per second come from the timed steps (`--points` per step,
replica seeds 0..steps-1, the same Sobol points for both estimators).  value =
the selected estimator's time (the plain one's with --no-flow).  One-off set-up
(g0 tabulation, pilot fit, flow training) is reported separately.

Side figure (estimator.plain_fp32_lu, skipped by --no-fp32): the plain
estimator with the determinants LU'd in FP32 and the Keldysh sum and the
accumulation in FP64, on the same replicas -- its points per second, its time
to 1e-3 at the FP64 search's points needed, and each replica's deviation from
the FP64 estimate of the same points.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector
TARGET_REL = 1e-3
PILOT_M = 1 << 16  # points per pilot replica when judging flow candidates
PILOT_REPLICAS = 128
SELECT_MARGIN = 0.8  # a flow is selected when its pilot score is below this fraction of the plain one
# committed rocprofv3 --pmc summaries holding "pmc-traffic <kernel> ..." lines
PMC_SUMMARIES = [os.path.join(ROOT, "profiles", f) for f in
                 ("integrand_ctree_kernel_n8_fp64_pmc.md", "integrand_wsplit_kernel_n8_fp64_pmc.md",
                  "integrand_split_kernel_n8_fp64_pmc.md",
                  "integrand_kernel_n8_fp64_pmc.md")]


class _stdout_to_stderr:
    """The host library prints the reference's model warnings on stdout (as
    keldy does); keep them off the one-JSON-line stdout of this benchmark."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)


def pmc_traffic(kernel, points_per_launch):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary: FETCH_SIZE and WRITE_SIZE (KiB per dispatch, averaged over the
    full-size launches), each scaled by the factor that makes the same
    counter read a plain device copy of known size correctly.  None when the
    summary does not cover this kernel at this launch size."""
    for path in PMC_SUMMARIES:
        try:
            text = open(path).read()
        except OSError:
            continue
        for line in text.splitlines():
            f = line.split()
            if len(f) < 2 or f[0] != "pmc-traffic" or f[1] != kernel:
                continue
            kv = dict(x.split("=", 1) for x in f[2:])
            if int(kv["points_per_launch"]) != points_per_launch:
                continue
            fetch = float(kv["fetch_kib"]) * 1024 * float(kv["fetch_factor"])
            write = float(kv["write_kib"]) * 1024 * float(kv["write_factor"])
            return {"bytes": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
                    "fetch_factor": float(kv["fetch_factor"]), "write_factor": float(kv["write_factor"]),
                    "source": os.path.relpath(path, ROOT)}
    return None


def charge_truth(order):
    """The committed high-precision Q_8(10) (tests/golden/q8_truth.json, made by
    tests/golden/make_q8_truth.py: 16 replicas of 2^28 points), or None."""
    path = os.path.join(ROOT, "tests", "golden", "q8_truth.json")
    if order != 8 or not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def model_params():
    return dict(beta=-1.0, bias_V=0.0, eps_d=0.0, Gamma=1.0, alpha=0.0, half_bandwidth=10.0,
                bath_type="flatband", time_max=10.0, nr_time_points_gf=10000, ft_method="analytic")


def plain_computer(order, t_max, pilot_points, precision):
    """The plain estimator: the reference's fitted product-1D recipe
    (projection_test.py:77-98): the pilot evaluates the warped integrand
    without the positive-time restriction on a g0 window of order * t_max."""
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw

    model = kio.G0Model(kio.G0ModelOmega(model_params()), False)
    wide = dict(model_params(), time_max=order * t_max, nr_time_points_gf=order * 10000)
    pilot = kio.ComputeChargeQDirect(kio.G0Model(kio.G0ModelOmega(wide), False), t_max, order, 0.0)
    pilot.warper.emplace_back(kw.WarperPlasmaUvT(t_max))
    pilot.warper.emplace_back(kw.make_product_1d_simple_inverse(t_max, 1.0))
    proj = kw.make_warper_from_proj(pilot, order, 500, pilot_points, 0.1, keep_u_hypercube=False)
    comp = kio.ComputeChargeQDirect(model, t_max, order, 0.0)
    comp.set_precision(precision)
    comp.warper.emplace_back(kw.WarperPlasmaUvT(t_max))
    comp.warper.emplace_back(kw.make_product_1d_simple_inverse(t_max, 1.0))
    comp.warper.emplace_back(proj)
    return model, comp, proj


def run_replicas(comp, first_seed, steps, points):
    """Run `steps` RQMC replicas (scramble seeds first_seed, first_seed+1, ...);
    returns per-replica estimates, in-domain counts, LU-kernel ms and the
    per-replica wall times (device work included)."""
    import numpy as np

    est, dom, lu_ms, walls = [], [], [], []
    for r in range(steps):
        comp.reset_rng("sobol", 0, True, True, first_seed + r)
        n0 = comp.reduce_nr_points_run()
        s0 = comp.reduce_result() * n0 if n0 else 0j
        d0 = comp.reduce_nr_points_in_domain()
        t0 = time.perf_counter()
        comp.run(points)
        walls.append(time.perf_counter() - t0)
        s1 = comp.reduce_result() * comp.reduce_nr_points_run()
        est.append((s1 - s0) / points)
        dom.append(comp.reduce_nr_points_in_domain() - d0)
        lu_ms.append(comp.last_timing()[1])
    return np.array(est), np.array(dom), np.array(lu_ms), np.array(walls)


def points_for_target(comp, replicas, target, m0=1 << 12, m_max=1 << 22):
    """Doubling search: RQMC spread over `replicas` scrambles at M points per
    replica until the relative spread of one replica falls below `target`.
    Returns (points needed by log-log interpolation, table of (M, rel spread),
    median wall time of one replica at the first M that reached the target)."""
    import numpy as np

    table = []
    m = m0
    while True:
        est, _, _, walls = run_replicas(comp, 10_000 + m, replicas, m)
        rel = est.std(ddof=1) / abs(est.mean())
        table.append((m, rel, float(np.median(walls))))
        if rel <= target or m >= m_max:
            break
        m *= 2
    if len(table) == 1:
        m1, r1, _ = table[0]
        return m1 * (r1 / target) ** 2, table, table[-1][2]
    (m1, r1, _), (m2, r2, _) = table[-2], table[-1]
    slope = np.log(r2 / r1) / np.log(m2 / m1)  # ~ -0.5 (MC) ... -1 (QMC)
    n = m2 * (target / r2) ** (1.0 / slope) if slope < 0 else m2 * (r2 / target) ** 2
    return float(n), table, table[-1][2]


def kernel_method_line(args, model, proj, t_max):
    """Throughput of the kernel method (compute_gf_kernel: per-vertex cofactor
    entries binned by vertex time and Keldysh index, 10 bins) at --order on the
    plain chain: points/s and the cofactor kernel's FP64 rate (its own flop
    count) over --steps replicas of --points points.  One JSON line."""
    import numpy as np

    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw

    comp = kio.ComputeGfKernel(model, t_max, args.order, 10)
    comp.warper.emplace_back(kw.WarperPlasmaUvT(t_max))
    comp.warper.emplace_back(kw.make_product_1d_simple_inverse(t_max, 1.0))
    comp.warper.emplace_back(proj)
    for r in range(args.warmup):
        comp.reset_rng("sobol", 0, True, True, 1000 + r)
        comp.run(args.points)
    walls, lu, dom = [], [], []
    for r in range(args.steps):
        comp.reset_rng("sobol", 0, True, True, r)
        d0 = comp.reduce_nr_points_in_domain()
        t0 = time.perf_counter()
        comp.run(args.points)
        walls.append(time.perf_counter() - t0)
        lu.append(comp.last_timing()[1])
        dom.append(comp.reduce_nr_points_in_domain() - d0)
    wall, lu_s = float(np.sum(walls)), float(np.sum(lu)) * 1e-3
    fpp = comp.flops_per_point()
    achieved = float(np.sum(dom)) * fpp / lu_s / 1e12
    launches = max(1, -(-args.points // (1 << 21)))
    b = comp.reduce_result()
    print(json.dumps({
        "metric": f"kernel-method points per second at n={args.order}", "value": args.steps * args.points / wall,
        "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "dtype": "f64",
        "config": {"workload": f"compute_gf_kernel order {args.order}, t={t_max}, 10 bins, plain fitted chain",
                   "points_per_replica": args.points},
        "roofline": {"bound": "valu-fp64", "kernel": f"cofactor_kernel<{args.order}>", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "flops_per_point": fpp, "lu_ms_per_launch": float(np.mean(lu)) / launches},
        "check": {"nr_values_dropped": int(b.get_nr_values_dropped()),
                  "sum_of_bins": complex(b.get_data().sum()).real},
    }))


def cpu_threads():
    """Threads for the CPU baseline and what they come from: the host cores this
    process may run on (sched_getaffinity), capped by OMP_NUM_THREADS when the
    environment sets it (the share of a shared GPU host's cores given to one
    GPU's jobs)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = affinity
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return threads, {"nproc": nproc, "affinity": affinity, "omp_num_threads": omp}


def cpu_baseline(model, proj, order, t_max, gpu_points_needed, seconds):
    """The CPU oracle (a C port of the reference path, FP64, partial-pivoting
    LUs per configuration) on the same warper chain and Sobol points, timed on
    the host cores three ways: points split over the threads in contiguous Sobol
    ranges (the analogue of the reference's MPI round robin,
    integrator.hpp:82-88); the reference's own OpenMP structure, every point in
    turn with its Wick rows and 2^n configurations spread over the threads
    (wick_direct.cpp:81,93,110); and one core.  The baseline is the faster of
    the two parallel structures.  Each is timed on a bounded sample sized to
    about its share of `seconds` from a calibration run."""
    import numpy as np

    from oracle import oracle as O

    threads, where = cpu_threads()
    g0 = O.G0Table(model.g0_lesser[0, 0], model.g0_greater[0, 0], model.time_max)
    f = O.charge_integrand(g0, t_max)
    maps = []
    for a in range(order):
        t, f1, F, cut = proj.get_1d_warper(a).knots()
        maps.append(O.Map1d("nearest", 1.0, t=np.array(t), f=np.array(f1), F=np.array(F), n_inv=cut + 1))
    train = O.Train([O.plasma(t_max), O.product(O.Map1d("inverse", t_max, w=1.0)), O.product(maps)])

    def timed(n, **kw):
        s = O.Sobol(order, 0, True, True, 0)
        t0 = time.perf_counter()
        val, _, ndom = O.run(f, train, s, n, **kw)
        return time.perf_counter() - t0, val, ndom

    out = {}
    for name, kw, share in (("points_split", dict(threads=threads), 0.55),
                            ("intrapoint_openmp", dict(threads=threads, intrapoint=True), 0.3),
                            ("single_core", dict(threads=1), 0.15)):
        calib = 1024
        dt, _, _ = timed(calib, **kw)
        n = int(max(calib, min(1 << 22, calib * seconds * share / max(dt, 1e-6))))
        dt, val, ndom = timed(n, **kw)
        out[name] = {"points_per_s": n / dt, "sample_points": n, "seconds": dt, "threads": kw["threads"],
                     "sum": val, "n_in_domain": ndom}
    best = max(("points_split", "intrapoint_openmp"), key=lambda k: out[k]["points_per_s"])
    pps = out[best]["points_per_s"]
    return {"pps": pps, "cores": threads, "structure": best, "where": where, "structures": out,
            "value": gpu_points_needed / pps, "sample_points": out[best]["sample_points"],
            "seconds": out[best]["seconds"]}


def workload_rates():
    """Keldysh-sum throughput of the other two device workloads, per order or
    time, counting only the points inside the integration domain (the rest cost
    the sampling stage only):
    - orders 1..10, floor(2^24 / n^2) points each, [plasma_uv(10), inverse(10, 1)],
      FP32 LU (FP64 from order 10), each order's launch shape autotuned; every
      order run on its own (in-domain points / Keldysh-sum kernel time), then
      all orders as one hipGraph launch;
    - the time-resolved current at order 10: 128 external times in (0, 20],
      2^20 points each, dot-lead semicircle g0 by contour integration,
      [plasma_uv(t), inverse(t, 1)]; summed over the times, then all times as
      one hipGraph launch;
    - the same with the time-conditioned flow appended (current_learned)."""
    import numpy as np

    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw

    def one(c, m):
        d0, f0 = c.reduce_nr_points_in_domain(), c.reduce_nr_fallback()
        c.run(m)
        lu_ms = c.last_timing()[1]
        return c.reduce_nr_points_in_domain() - d0, c.reduce_nr_fallback() - f0, lu_ms

    t = 10.0
    with _stdout_to_stderr():
        model = kio.G0Model(kio.G0ModelOmega(dict(model_params())), False)
    orders = list(range(1, 11))
    counts = [(1 << 24) // (n * n) for n in orders]
    comps, per_order = [], []
    for n, m in zip(orders, counts):
        c = kio.ComputeChargeQDirect(model, t, n, 0.0)
        c.set_precision("fp32")
        c.warper.emplace_back(kw.WarperPlasmaUvT(t))
        c.warper.emplace_back(kw.make_product_1d_simple_inverse(t, 1.0))
        c.reset_rng("sobol", 0, True, True, 11)
        lanes, wpc, _ = c.autotune(1 << 16)
        one(c, m)  # warm-up
        dom, fb, lu_ms = one(c, m)
        fpp = c.flops_per_point()
        per_order.append({"order": n, "points": m, "algorithm": c.get_algorithm(), "lanes": int(lanes),
                          "waves_per_cu": int(wpc), "in_domain": int(dom), "lu_ms": lu_ms,
                          "in_domain_points_per_s": dom / (lu_ms * 1e-3), "fallback_configurations": int(fb),
                          "tflops": dom * fpp / (lu_ms * 1e-3) / 1e12})
        c.reset_rng("sobol", 0, True, True, 11)
        comps.append(c)
    kio.run_batch(comps, counts)
    graph_ms = kio.run_batch(comps, counts)

    p = dict(beta=-1.0, bias_V=0.6, eps_d=-0.5, Gamma=1.0, time_max=20.0, nr_time_points_gf=2001, alpha=0.0,
             half_bandwidth=2.0, bath_type="semicircle", ft_method="contour")
    with _stdout_to_stderr():
        dot = kio.G0Model(kio.G0ModelOmega(p), True)
    times = np.linspace(20.0 / 128, 20.0, 128)
    cur, m = [], 1 << 20
    dom_t, fb_t, lu_t = 0, 0, 0.0
    for tt in times:
        c = kio.ComputeCurrentJDirect(dot, float(tt), 10, 0.0)
        c.warper.emplace_back(kw.WarperPlasmaUvT(float(tt)))
        c.warper.emplace_back(kw.make_product_1d_simple_inverse(float(tt), 1.0))
        c.reset_rng("sobol", 0, True, True, 5)
        dom, fb, lu_ms = one(c, m)
        dom_t, fb_t, lu_t = dom_t + dom, fb_t + fb, lu_t + lu_ms
        c.reset_rng("sobol", 0, True, True, 5)
        cur.append(c)
    kio.run_batch(cur, [m] * len(cur))
    cur_ms = kio.run_batch(cur, [m] * len(cur))
    learned = current_learned(cur, times, m)
    return {
        "orders_1_to_10": {"per_order": per_order, "one_graph_launch_ms": graph_ms},
        "current_order10_128_times": {"points": 128 * m, "in_domain": int(dom_t), "lu_ms_sum": lu_t,
                                      "in_domain_points_per_s": dom_t / (lu_t * 1e-3),
                                      "fallback_configurations": int(fb_t), "algorithm": cur[0].get_algorithm(),
                                      "one_graph_launch_ms": cur_ms},
        "current_order10_128_times_learned": learned,
    }


def current_learned(cur, times, m, reps=4, seed0=300):
    """The time-conditioned flow on the order-10 current (the 128 computers of
    workload_rates): a 2-layer flow with c = t / 20 trained on 2^18 samples
    per time (weights |f|^(1/4)), then `reps` replicas of m points per time
    with and without it on the same seeds, each replica set one hipGraph
    launch: per-time variance ratio plain/learned at equal points, replica-set
    wall times and in-domain fractions."""
    import numpy as np

    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw

    def replicas():
        est = np.zeros((reps, len(cur)), dtype=complex)
        dom = np.zeros(len(cur))
        secs = []
        for r in range(reps):
            prev = []
            for c in cur:
                c.reset_rng("sobol", 0, True, True, seed0 + r)
                n0 = c.reduce_nr_points_run()
                prev.append((c.reduce_result() * n0 if n0 else 0j, c.reduce_nr_points_in_domain()))
            t0 = time.perf_counter()
            kio.run_batch(cur, [m] * len(cur))
            secs.append(time.perf_counter() - t0)
            for i, c in enumerate(cur):
                est[r, i] = (c.reduce_result() * c.reduce_nr_points_run() - prev[i][0]) / m
                dom[i] += c.reduce_nr_points_in_domain() - prev[i][1]
        return est, dom / (reps * m), float(np.median(secs))

    conds = [float(t) / 20.0 for t in times]
    plain, dom_p, sec_p = replicas()
    t0 = time.perf_counter()
    flow, losses = kw.train_flow_warper_conditioned(cur, conds, n_layers=2, samples_per_computer=1 << 18,
                                                    steps=3000, batch=1 << 13, lr=1e-3, lr_final=1e-5,
                                                    weight_power=0.25)
    t_train = time.perf_counter() - t0
    for c, cv in zip(cur, conds):
        c.warper.emplace_back(flow.with_cond(cv))
    learned, dom_l, sec_l = replicas()
    vp = plain.real.var(axis=0, ddof=1) + plain.imag.var(axis=0, ddof=1)
    vl = learned.real.var(axis=0, ddof=1) + learned.imag.var(axis=0, ddof=1)
    ok = (vp > 0) & (vl > 0) & (dom_p * m >= 100)
    ratio = vp[ok] / vl[ok]
    gm = float(np.exp(np.mean(np.log(ratio)))) if ok.any() else float("nan")
    return {"points_per_time": m, "replicas": reps, "layers": 2, "train_s": t_train,
            "in_domain_fraction_plain": float(dom_p.mean()), "in_domain_fraction_learned": float(dom_l.mean()),
            "times_compared": int(ok.sum()), "variance_ratio_plain_over_learned_geomean": gm,
            "variance_ratio_quantiles_10_50_90": [float(x) for x in np.quantile(ratio, [0.1, 0.5, 0.9])] if ok.any() else None,
            "replica_set_s_plain": sec_p, "replica_set_s_learned": sec_l,
            "time_to_target_ratio_learned_over_plain": (sec_l / sec_p) / gm if ok.any() else None}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="timed RQMC replicas")
    ap.add_argument("--warmup", type=int, default=2, help="untimed replicas")
    ap.add_argument("--points", type=int, default=1 << 22, help="Sobol points per replica")
    ap.add_argument("--order", type=int, default=8)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--pilot-points", type=int, default=1 << 14)
    ap.add_argument("--target-replicas", type=int, default=256,
                    help="RQMC replicas per size in the doubling search for the points needed")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flow", action="store_true", help="plain estimator only")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the per-order rates of the orders 1..10 and time-resolved current workloads")
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the FP32-LU side measurement of the plain estimator")
    ap.add_argument("--flow-steps", type=int, default=2000)
    ap.add_argument("--flow-samples", type=int, default=1 << 20)
    ap.add_argument("--flow-batch", type=int, default=1 << 14)
    ap.add_argument("--flow-lr", type=float, default=1e-3)
    ap.add_argument("--flow-alphas", type=float, nargs="+", default=[1.0, 0.5, 0.25],
                    help="weight powers of the flow candidates (density ~ |f|^alpha)")
    ap.add_argument("--algorithm", default=None, choices=["config", "pair", "split", "wsplit", "ctree", "tree"],
                    help="Keldysh-sum algorithm (default: the library's)")
    ap.add_argument("--observable", default="charge", choices=["charge", "kernel"],
                    help="kernel: only time the kernel method (compute_gf_kernel, 10 bins) on the same chain")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM byte counters), then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_launch": n}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(points would shard by contiguous Sobol index ranges with one 32-byte all-reduce).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    t_max = 10.0
    t0 = time.perf_counter()
    with _stdout_to_stderr():
        model, comp, proj = plain_computer(args.order, t_max, args.pilot_points, args.precision)
    if args.algorithm:
        comp.set_algorithm(args.algorithm)
    setup_s = time.perf_counter() - t0
    if args.observable == "kernel":
        kernel_method_line(args, model, proj, t_max)
        return

    def measure(label):
        if args.warmup:
            run_replicas(comp, 1000, args.warmup, args.points)
        fb0 = comp.reduce_nr_fallback()
        est, dom, lu_ms, walls = run_replicas(comp, 0, args.steps, args.points)
        fb = comp.reduce_nr_fallback() - fb0
        timing = comp.last_timing()
        wall = float(walls.sum())
        n_total = args.steps * args.points
        pps = n_total / wall
        n_needed, table, wall_at_target = points_for_target(comp, args.target_replicas, TARGET_REL)
        # the time to the target measured directly: one run of the points needed
        # (median over replicas), launches and host round trips included; the
        # estimate needs no kernel timing, whose events cost ~20 us of a run
        comp.set_timing(0)
        walls_at = run_replicas(comp, 30_000, 32, int(np.ceil(n_needed)))[3]
        comp.set_timing(2)
        mean = est.mean()
        return {"warper": label, "est": est, "dom": dom, "lu_ms": lu_ms, "wall": wall, "pps": pps,
                "n_needed": n_needed, "t_needed": float(np.median(walls_at)), "t_throughput": n_needed / pps,
                "fallback": fb, "table": table,
                "wall_at_target": wall_at_target, "mean": mean,
                "charge": (-1j * (1j ** args.order) * mean).real,
                "rel_rep": est.std(ddof=1) / abs(mean) if len(est) > 1 else float("nan"),
                "in_domain_fraction": float(dom.sum()) / n_total,
                "timing": timing}

    def summary(m):
        return {"warper": m["warper"], "time_to_1e-3_s": m["t_needed"],
                "time_to_1e-3_from_throughput_s": m["t_throughput"], "charge_estimate": m["charge"],
                "rel_err_per_replica_at_points": m["rel_rep"], "points_needed_1e-3": m["n_needed"],
                "points_per_s": m["pps"], "ms_per_replica": m["wall"] / args.steps * 1e3,
                "rqmc_spread_table": [[a, r] for a, r, _ in m["table"]],
                "wall_s_one_replica_at_target": m["wall_at_target"],
                "in_domain_fraction": m["in_domain_fraction"],
                "fallback_configurations_in_timed_replicas": int(m["fallback"])}

    plain = measure("plain")
    # The mixed-precision variant (LU in FP32, the Keldysh sum and the
    # accumulation in FP64) on the same chain and the same scrambles as the
    # plain FP64 replicas: throughput, and each replica's deviation from its
    # FP64 estimate.  The points needed are the FP64 search's (the two
    # estimators differ by far less than their spread).
    fp32_side = None
    if args.precision == "fp64" and not args.no_fp32:
        comp.set_precision("fp32")
        run_replicas(comp, 1000, 1, args.points)
        est32, _, lu32, walls32 = run_replicas(comp, 0, args.steps, args.points)
        comp.set_timing(0)  # as for the FP64 time to target: no timing events
        walls32_at = run_replicas(comp, 30_000, 32, int(np.ceil(plain["n_needed"])))[3]
        comp.set_timing(2)
        comp.set_precision("fp64")
        pps32 = args.steps * args.points / float(walls32.sum())
        dev = np.abs(est32 - plain["est"]) / abs(plain["mean"])
        fp32_side = {"dtype": "f32-lu/f64-sum", "time_to_1e-3_s": float(np.median(walls32_at)),
                     "time_to_1e-3_from_throughput_s": plain["n_needed"] / pps32,
                     "points_per_s": pps32, "speedup_vs_fp64": pps32 / plain["pps"],
                     "lu_ms_per_launch": float(lu32.mean()) / max(1, -(-args.points // (1 << 21))),
                     "max_rel_deviation_from_fp64_per_replica": float(dev.max()),
                     "rel_spread_per_replica": plain["rel_rep"]}
    learned = None
    learned_selected = False
    selection = []
    if not args.no_flow:
        from keldy_amd import warpers as kw

        def set_chain(flow):
            tr = kw.WarperTrainT()
            tr.emplace_back(kw.WarperPlasmaUvT(t_max))
            tr.emplace_back(kw.make_product_1d_simple_inverse(t_max, 1.0))
            tr.emplace_back(proj)
            if flow is not None:
                tr.emplace_back(flow)
            comp.warper = tr

        def pilot():
            """(relative spread, device seconds per point) on the pilot replicas"""
            est, _, _, walls = run_replicas(comp, 20_000, PILOT_REPLICAS, PILOT_M)
            return float(est.std(ddof=1) / abs(est.mean())), float(walls.sum()) / (PILOT_REPLICAS * PILOT_M)

        # Candidates: flows trained toward |f|^alpha.  Each is scored on pilot
        # replicas (scramble seeds disjoint from the measured ones) by
        # spread^2 x seconds per point (the time to a target error, up to a
        # common factor); the best flow is measured, and it is the selected
        # estimator only if its score beats the plain one's by SELECT_MARGIN
        # (the pilot spreads carry ~10% noise); otherwise plain stays.
        set_chain(None)
        sp0, spp0 = pilot()
        plain_score = sp0 * sp0 * spp0
        selection.append({"candidate": "plain", "pilot_rel_spread": sp0, "pilot_s_per_point": spp0,
                          "score": plain_score})
        best, best_score, train_s, best_alpha = None, float("inf"), 0.0, None
        for alpha in args.flow_alphas:
            set_chain(None)
            t0 = time.perf_counter()
            flow, ls = kw.train_flow_warper(comp, n_layers=6, nr_samples=args.flow_samples, steps=args.flow_steps,
                                            batch=args.flow_batch, lr=args.flow_lr, lr_final=args.flow_lr * 1e-2,
                                            weight_power=alpha)
            ts = time.perf_counter() - t0
            train_s += ts
            set_chain(flow)
            sp, spp = pilot()
            score = sp * sp * spp
            selection.append({"candidate": f"flow alpha={alpha}", "pilot_rel_spread": sp, "pilot_s_per_point": spp,
                              "score": score, "train_s": ts, "loss_first": float(ls[0]),
                              "loss_last": float(np.mean(ls[-20:]))})
            if score < best_score:
                best, best_score, best_alpha = flow, score, alpha
        if best is not None:
            set_chain(best)
            learned = measure(f"learned flow (alpha={best_alpha})")
            learned_selected = best_score < SELECT_MARGIN * plain_score
            if not learned_selected:
                set_chain(None)
    head = learned if (learned and learned_selected) else plain

    fpp = comp.flops_per_point()
    flops = float(head["dom"].sum()) * fpp
    lu_s = head["lu_ms"].sum() * 1e-3
    achieved = flops / lu_s / 1e12
    peak = FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS
    launches_per_step = max(1, -(-args.points // (1 << 21)))
    algo = comp.get_algorithm()
    prefix = {"tree": "tree_", "split": "split_", "wsplit": "wsplit_", "ctree": "ctree_"}.get(algo, "")
    kname = (f"integrand_{prefix}kernel"
             f"<{args.order},{'double' if args.precision == 'fp64' else 'float'}{',pair' if algo == 'pair' else ''}>")
    traffic = pmc_traffic(kname, min(args.points, 1 << 21))

    estimator = {"plain": summary(plain), "setup_s": setup_s,
                 "end_to_end_s": {"plain": setup_s + plain["t_needed"]}}
    truth = charge_truth(args.order)
    if truth:
        qp = (-1j * (1j ** args.order) * plain["est"]).real
        err = float(np.hypot(qp.std(ddof=1) / np.sqrt(len(qp)), truth["charge_std_error"]))
        estimator["plain_vs_truth"] = {
            "truth": truth["charge_mean"], "truth_std_error": truth["charge_std_error"],
            "truth_points": truth["replicas"] * truth["points_per_replica"], "estimate": float(qp.mean()),
            "rel_deviation": float(qp.mean() / truth["charge_mean"] - 1.0),
            "deviation_over_sigma": float(abs(qp.mean() - truth["charge_mean"]) / err)}
    if fp32_side:
        estimator["plain_fp32_lu"] = fp32_side
    if selection:
        estimator["flow_selection"] = {"pilot_points": PILOT_M, "pilot_replicas": PILOT_REPLICAS, "candidates": selection,
                                       "selected": head["warper"], "select_margin": SELECT_MARGIN,
                                       "training_s_total": train_s}
    if learned:
        estimator["end_to_end_s"]["learned"] = setup_s + train_s + learned["t_needed"]
        pe, le = plain["est"], learned["est"]
        # the charge per replica: Q_n = -i i^n I_n (compute_obs_test.cpp:133)
        qp, ql = (-1j * (1j ** args.order) * pe).real, (-1j * (1j ** args.order) * le).real
        err = np.hypot(qp.std(ddof=1), ql.std(ddof=1)) / np.sqrt(len(qp))
        estimator.update({
            "learned": summary(learned),
            "speedup_learned_vs_plain": plain["t_needed"] / learned["t_needed"],
            "speedup_learned_vs_plain_from_throughput": plain["t_throughput"] / learned["t_throughput"],
            "variance_ratio_plain_over_learned_at_points": float(pe.var(ddof=1) / le.var(ddof=1)),
            "unbiasedness": {"charge_mean_plain": float(qp.mean()), "charge_mean_learned": float(ql.mean()),
                             "difference_over_sigma": float(abs(ql.mean() - qp.mean()) / err),
                             "replicas": len(qp), "points_per_replica": args.points},
            "flow_training": {"steps": args.flow_steps, "samples": args.flow_samples, "batch": args.flow_batch,
                              "lr": args.flow_lr, "layers": 6, "hidden": 64},
        })
    out = {
        "metric": "time to 1e-3 rel. error on charge at n=8: learned vs plain warper, 1 GPU",
        "value": head["t_needed"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] / args.steps * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-lu/f64-sum",
        "data": "synthetic (scrambled Sobol points; analytic flat-band g0)",
        "config": {"workload": "charge Q_n, n=8, t=10, flatband Gamma=1 T=0 V=0 eps_d=0, "
                               + ("learned 6-layer coupling flow on the plain fitted product-1D chain"
                                  if head is learned else "plain fitted product-1D warper"),
                   "order": args.order, "points_per_replica": args.points, "replicas": args.steps,
                   "pilot_points": args.pilot_points},
        "estimator": estimator,
        "roofline": {"bound": "valu-fp64" if args.precision == "fp64" else "valu-fp32",
                     "kernel": kname,
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_detail": traffic, "flops_per_point": fpp,
                     "lu_ms_per_launch": float(head["lu_ms"].mean()) / launches_per_step},
    }
    if not args.no_workloads:
        out["workloads"] = workload_rates()
    if not args.no_cpu:
        # the CPU runs the plain estimator: its time to 1e-3 = plain points needed / CPU points/s
        cb = cpu_baseline(model, proj, args.order, t_max, plain["n_needed"], args.cpu_seconds)
        out["cpu_baseline"] = {"value": cb["value"], "unit": "s", "cores": cb["cores"], "kind": "port",
                               "sample": f"{cb['sample_points']} scrambled Sobol points (seed 0) of the plain "
                                         f"chain in {cb['seconds']:.1f} s ({cb['pps']:.0f} points/s, "
                                         f"{cb['structure']})",
                               "points_per_s": cb["pps"], "structure": cb["structure"], "host": cb["where"],
                               "structures": {k: {kk: vv for kk, vv in v.items() if kk != "sum"}
                                              for k, v in cb["structures"].items()}}
        out["plain_gpu_vs_cpu"] = cb["value"] / plain["t_needed"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()

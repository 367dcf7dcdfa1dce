#!/usr/bin/env python3
"""HMC trajectories/s at L1=L2=32 (N=1024), Nt=128 on one MI355X: mixed precision vs the plain
FP64 path, with the force, dH and log-det errors of the mixed path against FP64.

Workload (one "step" = one HMC trajectory): triangular lattice 32x32, Nt=128, beta=40/eV,
J=(0.10,-0.03,0.02) eV, lambda=0.3 J (input line 0.3 0.3 0.3), mu0=0, omega0=0.006 eV, Nmd=10,
trajectory length pi/2 (Fourier acceleration default), 2 flavours, the partial product
of every slice stored for the force sweep (checkpoint k=1; one force launch covers 32 slices),
initial field = free-boson draw with seed 3.  Momentum normals and accept uniforms are synthetic
(numpy, seed 3) and resident in host memory; the chain state (field, M_hat^{-1}) lives in HBM.

Each precision runs eagerly (per-kernel HIP-event timings) and with the trajectory body replayed
as a captured hipGraph; the headline value is the graph-replayed mixed chain (--no-graph: the
eager one).  Prints one JSON line (rank 0, one GPU).
"""
import argparse
import collections
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "electron-phonon-hmc_amd"))

L, NT, NMD, K_CKPT, SPAN, SEED = 32, 128, 10, 1, 32, 3
BETA, J, LAM, MU0, OMEGA0, FLAVORS = 40.0, (0.10, -0.03, 0.02), (0.3, 0.3, 0.3), 0.0, 0.006, 2
PEAK_FP64 = 78.6e12      # MI355X FP64 (vector and matrix), FLOP/s
PEAK_FP32 = 157.3e12     # MI355X FP32 published dense peak
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "adjoint_fp32_32x32_nt128_traffic.json")
AUTOTUNE = True          # launch shapes measured per context before its chain starts (--no-autotune: rules)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def stencil_flops_adjoint_per_launch(ns, k):
    """Algorithmic FP operations of one adjoint launch (one segment of k slices, all columns),
    recomputation from the checkpoint excluded: per slice and column site
    3 stencils for A^b T, 4 for A^a W (13 flops each), 16 for the S_a sums, 8 for W_{tau-1},
    48 for the half-bond sums."""
    return (3 * 13 + 4 * 13 + 16 + 8 + 48) * k * ns * ns


def chain_flops(ns, nt):
    """60 flops per site, column and slice (two Taylor pairs: 4 stencils + 2 combines)."""
    return 60 * ns * ns * nt


def run_chain(E, precision, normals_list, uniforms, x0_normals, warmup, steps, graph=False):
    """Chain of `steps` timed trajectories after `warmup`.  Eager launches record the kernel
    groups with HIP events on the context stream; graph mode (the trajectory body replayed as one
    captured hipGraph) records no per-kernel events, only the wall time."""
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, NT)
    ctx = E.Context(L, L, NT, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=precision,
                    checkpoint=K_CKPT, graph=int(graph))
    if AUTOTUNE:
        log("autotune", precision, ctx.autotune())
    ctx.chain_init(x0_normals)
    for i in range(max(warmup, 1 if graph else 0)):
        ctx.trajectory(normals_list[i % len(normals_list)], uniforms[i % len(uniforms)], wait=False)
    ctx.sync()
    if not graph:
        ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(steps):
        j = (warmup + i) % len(normals_list)
        ctx.trajectory(normals_list[j], uniforms[j], wait=False)
    ctx.sync()
    el = time.perf_counter() - t0
    groups = {}
    if not graph:
        groups = {g: ctx.timing_get(g) for g in ("chain", "gj", "adjoint", "adjoint_kernel", "force_reduce", "fft",
                                                 "metropolis")}
        ctx.timing(False)
    acc = ctx.last_record()
    return ctx, el, groups, acc


def accuracy(E, x_host, normals, uniform, nobs=40):
    """Mixed vs FP64 on the same field and the same trajectory inputs:
    force_rel_err       |F32 - F64| / |F64| at the chain's field, same M_hat^{-1};
    dH_abs_err          Metropolis dH of the mixed path against the FP64 path evaluated at the
                        same leapfrog endpoint (the mixed chain evaluates H in FP64);
    dH_md_drift         dH of a trajectory integrated with FP32 forces minus the same trajectory
                        with FP64 forces (the force error carried along the molecular dynamics);
    obs_*               means and standard errors of log det / N and the boson action per link over
                        nobs trajectories of each chain, started from the same field with the same
                        random inputs."""
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, NT)
    out = {}
    ctxs = {}
    for name, prec in (("fp64", E.FP64), ("mixed", E.MIXED)):
        c = E.Context(L, L, NT, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=prec, checkpoint=K_CKPT)
        c.chain_set_x(x_host)
        ctxs[name] = c
    c64, c32 = ctxs["fp64"], ctxs["mixed"]
    minv = c64.chain_minv()
    ld = c64.chain_logdet()
    F64 = c64.force(x_host, minv, ld)
    F32 = c32.force(x_host, minv, ld)
    out["force_rel_err"] = float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64))
    out["logdet_rel_err"] = abs(c32.chain_logdet() - ld) / abs(ld)

    def H(c, x, p):
        lx = c.logdet_inv(c.M_hat(x))[0]
        return c.hamilton(x, lx, c.energy_momenta(p))

    p0 = c64.sample_fourier_momenta(normals)
    x1, p1 = c32.leap_frog(x_host, p0)             # mixed molecular dynamics
    dH_mixed = H(c32, x1, p1) - H(c32, x_host, p0)
    dH_fp64_same = H(c64, x1, p1) - H(c64, x_host, p0)
    x1d, p1d = c64.leap_frog(x_host, p0)           # FP64 molecular dynamics
    dH_fp64 = H(c64, x1d, p1d) - H(c64, x_host, p0)
    out["dH_mixed"] = dH_mixed
    out["dH_fp64"] = dH_fp64
    out["dH_abs_err"] = abs(dH_mixed - dH_fp64_same)
    out["dH_md_drift"] = abs(dH_mixed - dH_fp64)
    r64 = c64.trajectory(normals, uniform)
    r32 = c32.trajectory(normals, uniform)
    out["same_decision"] = bool(r64.accepted == r32.accepted)
    # short chains from the same state with the same random inputs
    rng = np.random.default_rng(SEED + 1)
    draws = [(rng.standard_normal(len(x_host)), float(rng.random())) for _ in range(nobs)]
    for name, c in ctxs.items():
        c.chain_set_x(x_host)
        lds, sbs, acc = [], [], 0
        for nrm, u in draws:
            rec = c.trajectory(nrm, u)
            acc += rec.accepted
            lds.append(rec.logdet / (L * L))
            sbs.append(c.hamilton(c.chain_x(), 0.0, 0.0) / len(x_host))
        for key, v in (("logdet_per_site", lds), ("boson_action_per_link", sbs)):
            v = np.asarray(v)
            out["obs_%s_%s" % (key, name)] = [float(v.mean()), float(v.std(ddof=1) / np.sqrt(len(v)))]
        out["obs_acceptance_%s" % name] = acc / nobs
    for c in ctxs.values():
        c.close()
    return out


def force_only(n):
    """n mixed-precision forces at the bench size (no torch, no graphs): hardware-counter passes."""
    import ephmc as E
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, NT)
    x = np.random.default_rng(SEED).standard_normal(NT * 3 * L * L) * 0.3
    ctx = E.Context(L, L, NT, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=E.MIXED, checkpoint=K_CKPT)
    ld, _, minv = ctx.logdet_inv(ctx.M_hat(x))
    for _ in range(n):
        F = ctx.force(x, minv, ld)
    ctx.sync()
    log("force-only: %d forces, |F|_1 = %.6g" % (n, float(np.abs(F).sum())))
    ctx.close()


def cpu_baseline(sample_threads):
    """The oracle (CPU restatement of the reference algorithm, C/OpenMP) timed on this host:
    one M_hat + LU log-det/inverse and one force at the full size, extrapolated to one
    trajectory as (Nmd+1) (M_hat + LU) + Nmd force (boson FFT work is below 1%)."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ["OMP_NUM_THREADS"] = str(sample_threads)
    from _oracle import OracleModel
    om = OracleModel(L, L, NT)
    x = np.random.default_rng(SEED).standard_normal(om.dim) * 0.5
    t0 = time.perf_counter()
    M = om.M_hat(x)
    ld, _, minv = om.logdet_inv(M)
    t_mlu = time.perf_counter() - t0
    t0 = time.perf_counter()
    om.force(x, minv, ld)
    t_force = time.perf_counter() - t0
    t_traj = (NMD + 1) * t_mlu + NMD * t_force
    return {"value": 1.0 / t_traj, "unit": "traj/s", "cores": sample_threads, "kind": "port",
            "sample": "one M_hat+LU (%.2f s) and one force (%.2f s) at 32x32, Nt=128 on %d OpenMP threads, "
                      "extrapolated to (Nmd+1)(M_hat+LU)+Nmd force = %.1f s per trajectory; CPU %s"
                      % (t_mlu, t_force, sample_threads, t_traj, platform.processor() or platform.machine())}


def copy_calibration():
    """A plain device-to-device copy of 512 MiB, five times: 512 MiB read + 512 MiB written per
    launch, the reference point for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy calibration: 5 x %d bytes" % (a.numel() * 8))


def correlator_leg(E, x_host, reps=5):
    """cc_correlator at 32x32, Nt=128 on the chain's field (FP64 M_hat^{-1} and log det, resident in
    HBM): ms per measurement (2 Nt slice applications + the 9-component contraction per slice, and
    the 9 Nt values copied back), and its agreement with a second call (the kernels are deterministic)."""
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, NT)
    c = E.Context(L, L, NT, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=E.FP64)
    c.chain_set_x(x_host)
    minv, ld = c.chain_minv(), c.chain_logdet()
    # the field and M_hat^{-1} stay in HBM, as in the host's trajectory loop
    dx, dm = c.array(c.dim, x_host), c.array(c.mat, minv)
    g0 = c.cc_correlator(dx, dm, ld)                # warm-up: allocates the transfer matrices
    c.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        g = c.cc_correlator(dx, dm, ld)
    el = (time.perf_counter() - t0) / reps
    c.close()
    return {"ms_per_measurement": 1e3 * el, "reps": reps, "repeatable": bool(np.array_equal(g, g0)),
            "g_00_tau0": float(np.asarray(g).ravel()[0])}


def tl_leg(E, nsamples):
    """Transient localisation at 32x32 (N=1024), 512 bins, beta=40, seed-7 disorder normals:
    disorder samples per second of the batched eigensolve + current matrix elements + binning."""
    nbins = 512
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, nbins, static_disorder=True)
    ctx = E.Context(L, L, 1, t, kappa, mu, omega, flavors=1, steps=1)
    width = -2 * mu / nbins + sum(8 * abs(k) / nbins for k in kappa)
    a, b = (1.0, 0.0), (0.5, 0.866)
    cm = [a[0], 0.5 * (a[0] + b[0]), 0.5 * (b[0] - a[0]), a[1], 0.5 * (a[1] + b[1]), 0.5 * (b[1] - a[1])]
    normals = np.random.default_rng(7).standard_normal((nsamples, 3 * L * L))
    ctx.tl_hist(normals, cm, width, nbins)       # warm-up: workspace of the full batch, code load
    ctx.sync()
    ctx.timing(True)
    t0 = time.perf_counter()
    g, z = ctx.tl_hist(normals, cm, width, nbins)
    el = time.perf_counter() - t0
    groups = {k: ctx.timing_get(k) for k in ("tl", "tl_tridiag", "tl_ql", "tl_back", "tl_pairs")}
    ctx.close()
    return {"samples_per_s": nsamples / el, "samples": nsamples, "seconds": el,
            "config": "32x32 (N=1024), 512 bins, beta=40, FP64",
            "kernel_ms_per_batch": {k: round(v[0], 3) for k, v in groups.items()},
            "greens_sum": float(g.sum())}


def c4_leg(E, ntraj, nchains=64, graph=True):
    """Independent chains of uneven size (L from {8,12,16,20,24}, Nt from {32,64,128}, drawn with
    MT19937-64 seed 0; momentum normals from numpy generators seeded 1000+i), mixed precision,
    each context on its own HIP stream, trajectories issued round-robin without waiting:
    aggregate trajectories/s."""
    import ctypes as C
    host = C.CDLL(os.path.join(ROOT, "electron-phonon-hmc_amd", "libephmc_host.so"))
    host.init_genrand64.argtypes = [C.c_ulong]
    host.genrand64_int64.restype = C.c_ulong
    host.init_genrand64(0)
    Ls, Nts = (8, 12, 16, 20, 24), (32, 64, 128)
    chains = []
    for i in range(nchains):
        l = Ls[host.genrand64_int64() % 5]
        nt = Nts[host.genrand64_int64() % 3]
        t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, nt)
        ctx = E.Context(l, l, nt, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=E.MIXED,
                        checkpoint=K_CKPT, graph=int(graph))
        rng = np.random.default_rng(1000 + i)
        dim = nt * 3 * l * l
        draws = [rng.standard_normal(dim) for _ in range(ntraj + 1)]
        us = rng.random(ntraj + 1)
        if AUTOTUNE:
            ctx.autotune()
        ctx.chain_init(rng.standard_normal(dim))
        chains.append((ctx, draws, us, l, nt))
    for ctx, draws, us, _, _ in chains:              # warm-up (graph capture)
        ctx.trajectory(draws[0], float(us[0]), wait=False)
    for ctx, *_ in chains:
        ctx.sync()
    t0 = time.perf_counter()
    for k in range(1, ntraj + 1):
        for ctx, draws, us, _, _ in chains:
            ctx.trajectory(draws[k], float(us[k]), wait=False)
    for ctx, *_ in chains:
        ctx.sync()
    el = time.perf_counter() - t0
    sizes = collections.Counter((l, nt) for *_, l, nt in chains)
    for ctx, *_ in chains:
        ctx.close()
    return {"traj_per_s": nchains * ntraj / el, "chains": nchains, "trajectories_per_chain": ntraj,
            "seconds": el, "launch": "graph" if graph else "eager",
            "sizes": {"%dx%d,Nt=%d" % (l, l, nt): n for (l, nt), n in sorted(sizes.items())}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-tl", action="store_true", help="skip the transient-localisation leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-chain leg")
    ap.add_argument("--batch-traj", type=int, default=3)
    ap.add_argument("--tl-samples", type=int, default=1024)
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the chains with the trajectory body replayed as a hipGraph (the headline value "
                         "then comes from the eager chain)")
    ap.add_argument("--tl-only", action="store_true",
                    help="run only the transient-localisation leg (--tl-samples samples) and print its JSON")
    ap.add_argument("--force-only", type=int, default=0, metavar="N",
                    help="run N mixed-precision forces at 32x32, Nt=128 (eager, no graphs) and exit: the "
                         "workload of the hardware-counter passes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 512 MiB device copy (known bytes) for HBM counter calibration")
    ap.add_argument("--copy-only", action="store_true", help="run only the calibration copy and exit")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the rule-based launch shapes instead of tuning each context")
    args = ap.parse_args()
    global AUTOTUNE
    AUTOTUNE = not args.no_autotune
    if args.copy_only:
        copy_calibration()
        return
    if args.force_only:
        force_only(args.force_only)
        return
    if args.tl_only:
        import ephmc as E
        print(json.dumps({"tl": tl_leg(E, args.tl_samples)}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; a single chain does not span GPUs here "
              "(independent chains and disorder samples shard trivially, see the design notes)", file=sys.stderr)
        sys.exit(2)
    import ephmc as E
    if args.copy_calibration:
        copy_calibration()
    rng = np.random.default_rng(SEED)
    dim = NT * 3 * L * L
    nbuf = min(args.steps + args.warmup, 8)
    normals_list = [rng.standard_normal(dim) for _ in range(nbuf)]
    uniforms = rng.random(nbuf)
    x0 = rng.standard_normal(dim)

    res = {}
    for name, prec in (("fp64", E.FP64), ("mixed", E.MIXED)):
        ctx, el, groups, rec = run_chain(E, prec, normals_list, uniforms, x0, args.warmup, args.steps)
        res[name] = {"el": el, "groups": groups, "ctx": ctx}
        log(name, "%.2f ms/traj" % (1e3 * el / args.steps), {g: "%.3f ms x %d" % v for g, v in groups.items()})
    if not args.no_graph:
        for name, prec in (("fp64_graph", E.FP64), ("mixed_graph", E.MIXED)):
            ctx, el, groups, rec = run_chain(E, prec, normals_list, uniforms, x0, args.warmup, args.steps, graph=True)
            res[name] = {"el": el, "groups": groups, "ctx": ctx}
            log(name, "%.2f ms/traj" % (1e3 * el / args.steps))
    x_state = res["fp64"]["ctx"].chain_x()
    for r in res.values():
        r["ctx"].close()
    acc = accuracy(E, x_state, normals_list[0], float(uniforms[0]))
    log("accuracy", acc)
    corr = correlator_leg(E, x_state)
    log("correlator", corr)

    el64, elmx = res["fp64"]["el"], res["mixed"]["el"]
    tps64, tpsmx = args.steps / el64, args.steps / elmx
    # dominant kernel of the measured (mixed) path: the adjoint force sweep, one launch per
    # span of 32 slices; HIP-event average over the timed region
    per_launch_ms, adj_calls = res["mixed"]["groups"]["adjoint_kernel"]
    flops = stencil_flops_adjoint_per_launch(L * L, SPAN)
    achieved = flops / (per_launch_ms * 1e-3)
    roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32, "traffic": None,
            "kernel": "adjoint_kernel<float,4,1024,1,4> (FP32 VALU stencils; one launch = 32 slices)"}
    # HBM bytes per launch of the same kernel from the two counter passes (FETCH_SIZE, WRITE_SIZE of
    # `bench.py --force-only 3`, each scaled by the counter's reading for a 512 MiB copy)
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            tr = json.load(fh)
        roof["traffic"] = tr["bytes_per_launch"]
        roof["traffic_source"] = os.path.relpath(TRAFFIC_FILE, ROOT)
    out = {
        "metric": "HMC traj/s at 32x32, Nt=128: mixed-prec vs FP64 GPU; force rel err vs FP64",
        "value": tpsmx, "unit": "traj/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elmx / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (numpy normals, seed 3)",
        "config": {"workload": "HMC 32x32 triangular, Nt=128, Nmd=10, traj=pi/2, 2 flavours, "
                               "every slice stored (k=1), mixed precision", "global_batch": 1, "seq_len": NT,
                   "parallelism": "none (1 GPU)"},
        "fp64_traj_per_s": tps64, "fp64_ms_per_step": 1e3 * el64 / args.steps,
        "mixed_over_fp64": tpsmx / tps64,
        "force_rel_err": acc["force_rel_err"], "dH_abs_err": acc["dH_abs_err"],
        "dH_md_drift": acc["dH_md_drift"],
        "logdet_rel_err_metropolis": acc["logdet_rel_err"],
        "observables": {k[4:]: v for k, v in acc.items() if k.startswith("obs_")},
        "kernel_ms": {p: {g: round(v[0], 4) for g, v in res[p]["groups"].items()} for p in res if res[p]["groups"]},
        "cc_correlator": corr,
        "roofline": roof,
    }
    if not args.no_graph:
        # headline: the trajectory body replayed as a captured hipGraph (the same kernels; the eager
        # chains above give the per-kernel HIP-event timings)
        out["eager_traj_per_s"] = {"fp64": tps64, "mixed": tpsmx}
        out["graph_traj_per_s"] = {"fp64": args.steps / res["fp64_graph"]["el"], "mixed": args.steps / res["mixed_graph"]["el"]}
        out["value"] = out["graph_traj_per_s"]["mixed"]
        out["ms_per_step"] = 1e3 * res["mixed_graph"]["el"] / args.steps
        out["fp64_traj_per_s"] = out["graph_traj_per_s"]["fp64"]
        out["fp64_ms_per_step"] = 1e3 * res["fp64_graph"]["el"] / args.steps
        out["mixed_over_fp64"] = out["value"] / out["fp64_traj_per_s"]
        out["launch"] = "hipGraph replay of the trajectory body"
    if not args.no_batch:
        out["chains64"] = c4_leg(E, args.batch_traj)
        log("chains64", out["chains64"])
    if not args.no_tl:
        out["tl"] = tl_leg(E, args.tl_samples)
        log("tl", out["tl"])
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

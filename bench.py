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


def run_chain(E, precision, normals_list, uniforms, x0_normals, warmup, steps, graph=False, checkpoint=K_CKPT):
    """Chain of `steps` timed trajectories after `warmup`.  Eager launches record the kernel
    groups with HIP events on the context stream; graph mode (the trajectory body replayed as one
    captured hipGraph) records no per-kernel events, only the wall time."""
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, NT)
    ctx = E.Context(L, L, NT, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=precision,
                    checkpoint=checkpoint, graph=int(graph))
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


def accuracy(E, x_host, ndraw=20, nobs=100, nsame=40):
    """Mixed precision against FP64 at 32x32, Nt=128 (every number compares two different
    computations; nothing is evaluated twice the same way):
    force_rel_err        |F_mixed - F_64| / |F_64| of the molecular-dynamics force at the chain's
                         field x, each built as the leapfrog builds it (mixed: FP32 chain -> M_hat ->
                         FP64 Gauss-Jordan -> FP32 adjoint sweep; FP64: all FP64);
    force_rel_err_sweep  the FP32 sweep alone: FP32 chain + sweep fed the FP64 M_hat^{-1};
    dH_drift_*           |dH_mixed - dH_64| of whole trajectories integrated from the same (x, p)
                         with FP32 and with FP64 forces (both dH evaluated with the FP64 Metropolis
                         H), over ndraw momentum draws: mean, max, and the mean |dH| for scale;
    obs_*                two independent chains (different random inputs, nobs trajectories each,
                         from the same start) per precision: means, binned standard errors (bins of
                         10) of log det / N, the boson action per link and the acceptance, and the
                         z-score of each difference;
    same_inputs_*        the two precisions with identical random inputs over nsame trajectories:
                         fraction of identical accept decisions and the largest log-det difference."""
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
    F64 = c64.md_force(x_host)
    F32 = c32.md_force(x_host)
    out["force_rel_err"] = float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64))
    out["force_rel_err_sweep"] = float(np.linalg.norm(c32.force(x_host, minv, ld) - F64) / np.linalg.norm(F64))
    out["_F64"], out["_F32"], out["_logdet"] = F64, F32, ld

    def H(x, p):
        return c64.hamilton(x, c64.logdet_inv(c64.M_hat(x))[0], c64.energy_momenta(p))

    rng = np.random.default_rng(SEED + 2)
    drift, dhs = [], []
    for _ in range(ndraw):
        p0 = c64.sample_fourier_momenta(rng.standard_normal(len(x_host)))
        H0 = H(x_host, p0)
        x1, p1 = c32.leap_frog(x_host, p0)
        x1d, p1d = c64.leap_frog(x_host, p0)
        d32, d64 = H(x1, p1) - H0, H(x1d, p1d) - H0
        drift.append(abs(d32 - d64))
        dhs.append(abs(d64))
    out["dH_drift_mean"] = float(np.mean(drift))
    out["dH_drift_max"] = float(np.max(drift))
    out["dH_abs_mean"] = float(np.mean(dhs))
    out["dH_drift_trajectories"] = ndraw

    def chain(c, seed, n):
        c.chain_set_x(x_host)
        r = np.random.default_rng(seed)
        lds, sbs, acc = [], [], []
        for _ in range(n):
            rec = c.trajectory(r.standard_normal(len(x_host)), float(r.random()))
            acc.append(rec.accepted)
            lds.append(rec.logdet / (L * L))
            sbs.append(c.hamilton(c.chain_x(), 0.0, 0.0) / len(x_host))
        return np.array(lds), np.array(sbs), np.array(acc, float)

    def binned(v, b=10):
        m = len(v) // b
        blocks = v[:m * b].reshape(m, b).mean(1)
        return float(v.mean()), float(blocks.std(ddof=1) / np.sqrt(m))

    series = {"fp64": chain(c64, SEED + 10, nobs), "mixed": chain(c32, SEED + 20, nobs)}
    for k, key in enumerate(("logdet_per_site", "boson_action_per_link", "acceptance")):
        a, b = binned(series["fp64"][k]), binned(series["mixed"][k])
        out["obs_" + key] = {"fp64": a, "mixed": b,
                             "z": abs(a[0] - b[0]) / max(math.hypot(a[1], b[1]), 1e-300)}
    out["obs_trajectories"] = nobs
    r64, r32 = [], []
    for c, recs in ((c64, r64), (c32, r32)):
        c.chain_set_x(x_host)
        r = np.random.default_rng(SEED + 30)
        for _ in range(nsame):
            rec = c.trajectory(r.standard_normal(len(x_host)), float(r.random()))
            recs.append((rec.accepted, rec.logdet))
    out["same_inputs_decisions_equal"] = sum(a[0] == b[0] for a, b in zip(r64, r32)) / nsame
    out["same_inputs_max_logdet_diff"] = float(max(abs(a[1] - b[1]) for a, b in zip(r64, r32)))
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


def graph_only(n):
    """n mixed-precision trajectories at the bench size with the trajectory body replayed as the
    captured hipGraph (the headline path; no torch): the workload of the counter passes on it."""
    import ephmc as E
    rng = np.random.default_rng(SEED)
    dim = NT * 3 * L * L
    normals_list = [rng.standard_normal(dim) for _ in range(min(n + 1, 8))]
    uniforms = rng.random(len(normals_list))
    ctx, el, _, _ = run_chain(E, E.MIXED, normals_list, uniforms, rng.standard_normal(dim), 1, n, graph=True)
    log("graph-only: %d trajectories, %.2f ms each" % (n, 1e3 * el / max(n, 1)))
    ctx.close()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _omp_threads(n):
    """Threads of the oracle's OpenMP regions (the runtime the oracle library was linked with)."""
    import ctypes as C
    try:
        C.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
    except OSError:
        os.environ["OMP_NUM_THREADS"] = str(n)


def cpu_baseline(threads, x_state, gpu):
    """The oracle (the C/OpenMP restatement of the reference algorithm) timed on this host's cores,
    and used as the checker of the GPU at the full size.

    Timed sample (32x32, Nt=128, `threads` OpenMP threads): one complete trajectory with Nmd=10
    through the oracle's trajectory driver (momentum draw and colouring, leapfrog with ten forces,
    eleven M_hat + LU, Metropolis), about 23 s on 16 threads; one M_hat + LU and one force at the
    chain's field x are timed as well (they are the checks below).  The oracle's own LU and FFT run
    (no LAPACK or FFTW).  Also at 1 thread: the 6x6, Nt=32 chain (Nmd=10), 200 trajectories in full.

    Checks at x (the GPU values computed in the accuracy leg): the GPU's FP64 log det and its FP64
    and mixed molecular-dynamics forces against the oracle's."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import HMC, OracleModel, oracle
    lib = oracle()
    _omp_threads(threads)
    om = OracleModel(L, L, NT)
    h = HMC()
    assert lib.orc_hmc_init(C.byref(h), C.byref(om.m), SEED, NMD, 0.0, None) == 0
    dH = C.c_double()
    t0 = time.perf_counter()
    lib.orc_trajectory(C.byref(h), None, None, C.byref(dH), None, None)
    t_traj = time.perf_counter() - t0
    lib.orc_hmc_free(C.byref(h))
    t0 = time.perf_counter()
    ld, _, minv = om.logdet_inv(om.M_hat(x_state))
    t_mlu = time.perf_counter() - t0
    t0 = time.perf_counter()
    F = om.force(x_state, minv, ld)
    t_force = time.perf_counter() - t0
    checks = {"logdet_rel_err_fp64_vs_oracle": abs(gpu["logdet"] - ld) / abs(ld),
              "force_rel_err_fp64_vs_oracle": float(np.linalg.norm(gpu["F64"] - F) / np.linalg.norm(F)),
              "force_rel_err_mixed_vs_oracle": float(np.linalg.norm(gpu["F32"] - F) / np.linalg.norm(F))}
    # 6x6, Nt=32 on one thread, in full
    _omp_threads(1)
    om1 = OracleModel(6, 6, 32)
    h = HMC()
    assert lib.orc_hmc_init(C.byref(h), C.byref(om1.m), 1, NMD, 0.0, None) == 0
    n1 = 200
    t0 = time.perf_counter()
    for _ in range(n1):
        lib.orc_trajectory(C.byref(h), None, None, C.byref(dH), None, None)
    t_c1 = (time.perf_counter() - t0) / n1
    lib.orc_hmc_free(C.byref(h))
    _omp_threads(threads)
    base = {"value": 1.0 / t_traj, "unit": "traj/s", "cores": threads, "kind": "port",
            "sample": "32x32, Nt=128 on %d OpenMP threads: one complete Nmd=10 trajectory (%.1f s; one M_hat+LU "
                      "%.2f s, one force %.2f s).  Oracle's own LU and FFT.  CPU %s (%s logical CPUs visible)"
                      % (threads, t_traj, t_mlu, t_force, cpu_model(), os.cpu_count())}
    c1 = {"traj_per_s": 1.0 / t_c1, "threads": 1, "trajectories": n1,
          "config": "6x6, Nt=32, Nmd=10, seed 1, FP64 (oracle, complete trajectories)"}
    return base, c1, checks


def c2_stab_leg(E, ndraw=10, ntherm=5):
    """16x16, Nt=64, Nmd=10 (the stabilised configuration): FP64, mixed, and mixed with the slice
    product stabilised by an FP64 QR every 8 slices (ephmc_config.stabilize = 8).  From a field
    after ntherm FP64 trajectories (seed 2): the MD force of each mixed path against FP64, and the
    dH of ndraw trajectories integrated from the same (x, p), each evaluated with the FP64 H:
    |dH_path - dH_64|, mean and max; the time of one MD force of each path."""
    l, nt = 16, 64
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, nt)
    mk = lambda prec, **kw: E.Context(l, l, nt, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=prec,
                                      checkpoint=K_CKPT, **kw)
    c64, c32, cst = mk(E.FP64), mk(E.MIXED), mk(E.MIXED, stabilize=8)
    rng = np.random.default_rng(2)
    dim = nt * 3 * l * l
    c64.chain_init(rng.standard_normal(dim))
    for _ in range(ntherm):
        c64.trajectory(rng.standard_normal(dim), float(rng.random()))
    x = c64.chain_x()
    out = {"config": "16x16, Nt=64, Nmd=10, seed 2, %d FP64 trajectories of thermalisation" % ntherm}
    F = {}
    for name, c in (("fp64", c64), ("mixed", c32), ("stabilised", cst)):
        c.md_force(x)
        c.sync()
        t0 = time.perf_counter()
        F[name] = c.md_force(x)
        out["md_force_ms_" + name] = 1e3 * (time.perf_counter() - t0)
    for name in ("mixed", "stabilised"):
        out["force_rel_err_" + name] = float(np.linalg.norm(F[name] - F["fp64"]) / np.linalg.norm(F["fp64"]))

    def H(xx, p):
        return c64.hamilton(xx, c64.logdet_inv(c64.M_hat(xx))[0], c64.energy_momenta(p))

    drift = {"mixed": [], "stabilised": []}
    for _ in range(ndraw):
        p0 = c64.sample_fourier_momenta(rng.standard_normal(dim))
        H0 = H(x, p0)
        x64, p64 = c64.leap_frog(x, p0)
        d64 = H(x64, p64) - H0
        for name, c in (("mixed", c32), ("stabilised", cst)):
            x1, p1 = c.leap_frog(x, p0)
            drift[name].append(abs(H(x1, p1) - H0 - d64))
    for name, v in drift.items():
        out["dH_drift_" + name] = {"mean": float(np.mean(v)), "max": float(np.max(v))}
    out["trajectories"] = ndraw
    for c in (c64, c32, cst):
        c.close()
    return out


def c2_chains(E, ntraj=500):
    """16x16, Nt=64, Nmd=10, seed 2: the FP64 chain (the drop-in) and the mixed-precision chain run
    over the same ntraj trajectories' random inputs (numpy generator seeded 2), trajectory body
    replayed as a hipGraph, each trajectory's record read back before the next: rates, acceptance,
    the fraction of identical accept/reject decisions,
    the largest log det difference, <exp(-dH)> and the mean log det per site of each chain (binned
    errors, bins of 25) with the z-score of the difference."""
    l, nt = 16, 64
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, nt)
    dim = nt * 3 * l * l
    out = {"config": "16x16, Nt=64, Nmd=10, seed 2, %d trajectories per chain, same random inputs" % ntraj}
    recs = {}
    for name, prec in (("fp64", E.FP64), ("mixed", E.MIXED)):
        ctx = E.Context(l, l, nt, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=prec, checkpoint=K_CKPT,
                        graph=1)
        rng = np.random.default_rng(2)
        ctx.chain_init(rng.standard_normal(dim))
        r = []
        t0 = time.perf_counter()
        for _ in range(ntraj):
            rec = ctx.trajectory(rng.standard_normal(dim), float(rng.random()))
            r.append((rec.accepted, rec.dH, rec.logdet))
        out[name + "_traj_per_s"] = ntraj / (time.perf_counter() - t0)
        recs[name] = np.array(r, dtype=float)
        ctx.close()

    def binned(v, b=25):
        m = len(v) // b
        bm = v[:m * b].reshape(m, b).mean(axis=1)
        return float(v.mean()), float(bm.std(ddof=1) / np.sqrt(m))

    for name, a in recs.items():
        out[name] = {"acceptance": float(a[:, 0].mean()), "exp_minus_dH": binned(np.exp(-a[:, 1])),
                     "logdet_per_site": binned(a[:, 2] / (l * l))}
    a, b = recs["fp64"], recs["mixed"]
    out["decisions_equal"] = float((a[:, 0] == b[:, 0]).mean())
    out["max_logdet_rel_diff"] = float(np.max(np.abs(a[:, 2] - b[:, 2]) / np.abs(a[:, 2])))
    (m0, e0), (m1, e1) = out["fp64"]["logdet_per_site"], out["mixed"]["logdet_per_site"]
    out["logdet_per_site_z"] = float(abs(m0 - m1) / np.hypot(e0, e1)) if e0 > 0 or e1 > 0 else 0.0
    return out


def c1_gpu(E, ntraj=200):
    """The 6x6, Nt=32 FP64 chain (Nmd=10) on the GPU, seed-1 normals from numpy, graph replay."""
    t, kappa, mu, omega = E.scale_params(BETA, J, LAM, MU0, OMEGA0, 32)
    ctx = E.Context(6, 6, 32, t, kappa, mu, omega, flavors=FLAVORS, steps=NMD, precision=E.FP64, graph=1)
    rng = np.random.default_rng(1)
    draws = [(rng.standard_normal(ctx.dim), float(rng.random())) for _ in range(ntraj + 1)]
    ctx.chain_init(rng.standard_normal(ctx.dim))
    ctx.trajectory(*draws[0], wait=False)
    ctx.sync()
    t0 = time.perf_counter()
    for nrm, u in draws[1:]:
        ctx.trajectory(nrm, u, wait=False)
    ctx.sync()
    el = time.perf_counter() - t0
    ctx.close()
    return {"traj_per_s": ntraj / el, "trajectories": ntraj, "config": "6x6, Nt=32, Nmd=10, FP64, graph replay"}


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


def c4_chains(E, ntraj, nchains, graph=True):
    """The 64 uneven chains: L from {8,12,16,20,24}, Nt from {32,64,128}, drawn with MT19937-64
    seed 0; momentum normals and uniforms from numpy generators seeded 1000+i; mixed precision,
    autotuned unless --no-autotune; initialised."""
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
    return chains


def c4_leg(E, ntraj, nchains=64, graph=True):
    """The 64 uneven chains (c4_chains), each context on its own HIP stream, trajectories issued
    round-robin without waiting: aggregate trajectories/s."""
    chains = c4_chains(E, ntraj, nchains, graph)
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


def c4_batch_leg(E, ntraj, nchains=64):
    """The same 64 chains and inputs as c4_leg, advanced as one batch (ephmc_batch_*: every kernel
    of a trajectory launched once over all chains of a group): aggregate trajectories/s."""
    chains = c4_chains(E, ntraj, nchains)
    ctxs = [c[0] for c in chains]
    b = E.Batch(ctxs)
    info = b.info()
    b.trajectory([c[1][0] for c in chains], [float(c[2][0]) for c in chains])     # warm-up
    t0 = time.perf_counter()
    for k in range(1, ntraj + 1):
        b.trajectory([c[1][k] for c in chains], [float(c[2][k]) for c in chains], wait=False)
        b.collect()
    el = time.perf_counter() - t0
    b.close()
    for ctx in ctxs:
        ctx.close()
    return {"traj_per_s": nchains * ntraj / el, "chains": nchains, "trajectories_per_chain": ntraj,
            "seconds": el, "groups": info["groups"], "launches_per_trajectory": info["launches"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-tl", action="store_true", help="skip the transient-localisation leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-chain leg")
    ap.add_argument("--no-c2", action="store_true", help="skip the 16x16 stabilised-product leg")
    ap.add_argument("--no-k16", action="store_true", help="skip the chains with partial products every 16 slices")
    ap.add_argument("--batch-traj", type=int, default=10)
    ap.add_argument("--c4-only", action="store_true",
                    help="run only the 64-chain legs (streams and one batch) and print their JSON")
    ap.add_argument("--c4-batch-only", action="store_true",
                    help="run only the 64-chain batch leg and print its JSON (the workload of its kernel trace, "
                         "with EPHMC_BATCH_EAGER=1 for plain launches)")
    ap.add_argument("--c2-only", action="store_true",
                    help="run only the 16x16, Nt=64 legs (stabilised product, 500-trajectory chains) and print their JSON")
    ap.add_argument("--tl-samples", type=int, default=4096)
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the chains with the trajectory body replayed as a hipGraph (the headline value "
                         "then comes from the eager chain)")
    ap.add_argument("--tl-only", action="store_true",
                    help="run only the transient-localisation leg (--tl-samples samples) and print its JSON")
    ap.add_argument("--force-only", type=int, default=0, metavar="N",
                    help="run N mixed-precision forces at 32x32, Nt=128 (eager, no graphs) and exit: the "
                         "workload of the hardware-counter passes")
    ap.add_argument("--graph-only", type=int, default=0, metavar="N",
                    help="run N mixed-precision trajectories at 32x32, Nt=128 replayed as the captured hipGraph "
                         "(the headline path) and exit: the workload of the counter passes on it")
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
    if args.graph_only:
        graph_only(args.graph_only)
        return
    if args.c4_batch_only:
        import ephmc as E
        print(json.dumps({"chains64_batch": c4_batch_leg(E, args.batch_traj)}), flush=True)
        return
    if args.c4_only:
        import ephmc as E
        print(json.dumps({"chains64": c4_leg(E, args.batch_traj),
                          "chains64_batch": c4_batch_leg(E, args.batch_traj)}), flush=True)
        return
    if args.c2_only:
        import ephmc as E
        print(json.dumps({"c2_stabilised": c2_stab_leg(E), "c2_chains": c2_chains(E)}), flush=True)
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
    # the stated checkpoint interval of the force check (k = 16: partial products every 16 slices,
    # the others recomputed inside the sweep), same chains
    k16 = {}
    if not args.no_k16:
        for name, prec in (("fp64", E.FP64), ("mixed", E.MIXED)):
            ctx, el, groups, rec = run_chain(E, prec, normals_list, uniforms, x0, args.warmup, args.steps,
                                             graph=not args.no_graph, checkpoint=16)
            k16[name] = args.steps / el
            ctx.close()
            log("k16", name, "%.2f ms/traj" % (1e3 * el / args.steps))
    x_state = res["fp64"]["ctx"].chain_x()
    for r in res.values():
        r["ctx"].close()
    acc = accuracy(E, x_state)
    gpu_vals = {"F64": acc.pop("_F64"), "F32": acc.pop("_F32"), "logdet": acc.pop("_logdet")}
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
    roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32, "traffic": None,
            "peak_unpacked": PEAK_FP64 / 1e12, "frac_unpacked": achieved / PEAK_FP64,
            "kernel": "adjoint_kernel<float,4,1024,1,4> (FP32 vector-ALU stencils, no matrix cores; one launch = "
                      "32 slices; the published 157.3 TFLOP/s FP32 peak assumes packed FP32, compiled out here, "
                      "so 78.6 TFLOP/s is its unpacked ceiling)",
            "flop_per_launch": flops, "avg_launch_ms": per_launch_ms, "launches": adj_calls}
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
        "force_rel_err": acc["force_rel_err"], "force_rel_err_sweep": acc["force_rel_err_sweep"],
        "dH_drift": {"mean": acc["dH_drift_mean"], "max": acc["dH_drift_max"], "mean_abs_dH": acc["dH_abs_mean"],
                     "trajectories": acc["dH_drift_trajectories"], "gate": 1e-8,
                     "met": acc["dH_drift_max"] <= 1e-8},
        "observables": {k[4:]: v for k, v in acc.items() if k.startswith("obs_")},
        "same_inputs": {"decisions_equal": acc["same_inputs_decisions_equal"],
                        "max_logdet_diff": acc["same_inputs_max_logdet_diff"]},

        "kernel_ms": {p: {g: round(v[0], 4) for g, v in res[p]["groups"].items()} for p in res if res[p]["groups"]},
        "cc_correlator": corr,
        "roofline": roof,
    }
    if k16:
        out["checkpoint16_traj_per_s"] = k16
        out["checkpoint16_mixed_over_fp64"] = k16["mixed"] / k16["fp64"]
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
        out["chains64_batch"] = c4_batch_leg(E, args.batch_traj)
        log("chains64_batch", out["chains64_batch"])
    if not args.no_c2:
        out["c2_stabilised"] = c2_stab_leg(E)
        out["c2_chains"] = c2_chains(E)
        log("c2 stabilised", out["c2_stabilised"])
    if not args.no_tl:
        out["tl"] = tl_leg(E, args.tl_samples)
        log("tl", out["tl"])
    if not args.no_cpu:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or min(16, os.cpu_count() or 1))
        out["c1_gpu"] = c1_gpu(E)
        base, c1, checks = cpu_baseline(threads, x_state, gpu_vals)
        out["cpu_baseline"] = base
        out["cpu_c1_1thread"] = c1
        out["full_size_checks_vs_oracle"] = checks
        log("cpu", base, c1, checks)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

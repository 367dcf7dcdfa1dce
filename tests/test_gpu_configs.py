"""The 16x16 (N=256), Nt=64 and batched-chain configurations on the GPU.

16x16, Nt=64 (seed 2, Nmd=10): the seeded FP64 chain against the oracle's chain of the same seed
(committed fixture tests/golden/c2_fp64_16x16_nt64_seed2.json, written by
tests/golden/make_chain_fixtures.py): identical accept/reject decisions, dH within 1e-8 and log det
within 1e-10 relative over 20 trajectories; then the mixed-precision errors against FP64: the force
of one molecular-dynamics step (FP32 chain -> M_hat -> FP64 inverse -> FP32 sweep) within 1e-5
relative (asserted), and the dH of the mixed trajectory against the FP64 trajectory from the same
(x, p) (reported; bounded loosely as a regression check).

Batched chains: contexts of uneven lattice and Nt (8..24, Nt 32..128), autotuned, with the
trajectory body replayed as a hipGraph, run interleaved on their own streams: every chain's records
are bit-identical to the same context run alone, and the 8x8 chains match the oracle."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from _gpu import gpu_context, rel, say
from _oracle import HMC, MT, OracleModel, oracle, ptr

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_fp64_16x16_nt64_seed2.json")


def mt_stream(seed):
    lib = oracle()
    m = MT()
    lib.orc_mt_seed(C.byref(m), seed)
    return lib, m


def draw(lib, m, n):
    v = np.zeros(n)
    lib.orc_random_vector(C.byref(m), ptr(v), n)
    return v


def test_c2_fp64_chain_matches_oracle_fixture():
    fx = json.load(open(FIXTURE))
    cfg = fx["config"]
    say("C2 FP64 chain vs oracle fixture", cfg)
    om = OracleModel(cfg["l1"], cfg["l2"], cfg["nt"])
    ctx = gpu_context(om, steps=cfg["steps"], checkpoint=0)
    lib, m = mt_stream(cfg["seed"])
    ctx.chain_init(draw(lib, m, om.dim))
    assert ctx.chain_logdet() == pytest.approx(fx["logdet0"], rel=1e-10)
    worst = 0.0
    for it, (acc, dH, ld) in enumerate(zip(fx["accepted"], fx["dH"], fx["logdet"])):
        nrm = draw(lib, m, om.dim)
        u = lib.orc_mt_real2(C.byref(m))
        rec = ctx.trajectory(nrm, u)
        assert rec.accepted == acc, it
        assert abs(rec.dH - dH) < 1e-8, (it, rec.dH, dH)
        worst = max(worst, abs(rec.logdet - ld) / abs(ld))
    say("C2: worst log det rel diff", worst)
    assert worst < 1e-10


def test_c2_mixed_precision_errors():
    """Force of one MD step, mixed vs FP64, within 1e-5; dH drift of whole trajectories reported."""
    om = OracleModel(16, 16, 64)
    c64 = gpu_context(om, steps=10, checkpoint=0)
    c32 = gpu_context(om, steps=10, checkpoint=0, precision=1)
    lib, m = mt_stream(2)
    x = c64.sample_free_bosons(draw(lib, m, om.dim))
    F64, F32 = c64.md_force(x), c32.md_force(x)
    err = rel(F32, F64)
    say("C2 MD force rel err (mixed vs FP64)", err)
    assert err < 1e-5
    drifts = []
    for k in range(6):
        p = c64.sample_fourier_momenta(draw(lib, m, om.dim))
        H0 = c64.hamilton(x, c64.logdet_inv(c64.M_hat(x))[0], c64.energy_momenta(p))
        dH = []
        for c in (c64, c32):
            x1, p1 = c.leap_frog(x, p)
            dH.append(c64.hamilton(x1, c64.logdet_inv(c64.M_hat(x1))[0], c64.energy_momenta(p1)) - H0)
        drifts.append(abs(dH[1] - dH[0]))
        x = x1
    say("C2 dH drift mixed vs FP64: mean %.3g max %.3g" % (np.mean(drifts), np.max(drifts)))
    assert np.max(drifts) < 1e-5


# ---- batched chains -----------------------------------------------------------------------------

CHAINS = [(8, 32, 0), (8, 64, 1), (12, 128, 1), (16, 32, 1), (20, 64, 1), (24, 32, 1), (12, 64, 0), (16, 128, 1)]
NTRAJ = 3


def test_c4_uneven_chains_interleaved_equal_solo():
    ctxs, streams = [], []
    for i, (l, nt, prec) in enumerate(CHAINS):
        om = OracleModel(l, l, nt)
        ctx = gpu_context(om, steps=10, checkpoint=0, precision=prec, graph=1)
        ctx.autotune()
        lib, m = mt_stream(1000 + i)
        init = draw(lib, m, om.dim)
        draws = []
        for _ in range(NTRAJ):
            nrm = draw(lib, m, om.dim)
            draws.append((nrm, lib.orc_mt_real2(C.byref(m))))
        ctxs.append(ctx)
        streams.append((om, init, draws))
    say("C4: %d contexts tuned" % len(ctxs))
    solo = []
    for ctx, (om, init, draws) in zip(ctxs, streams):
        ctx.chain_init(init)
        recs = []
        for nrm, u in draws:
            r = ctx.trajectory(nrm, u)
            recs.append((r.accepted, r.dH, r.logdet))
        solo.append((recs, ctx.chain_x()))
    # the same chains again, interleaved: every context on its own stream, nothing waits
    for ctx, (om, init, draws) in zip(ctxs, streams):
        ctx.chain_init(init)
    inter = [[] for _ in ctxs]
    for k in range(NTRAJ):
        for ctx, (om, init, draws) in zip(ctxs, streams):
            ctx.trajectory(draws[k][0], draws[k][1], wait=False)
        for j, ctx in enumerate(ctxs):
            r = ctx.last_record()
            inter[j].append((r.accepted, r.dH, r.logdet))
    for j, ctx in enumerate(ctxs):
        assert inter[j] == solo[j][0], CHAINS[j]
        np.testing.assert_array_equal(ctx.chain_x(), solo[j][1])
    say("C4: interleaved records equal the solo runs")
    # 8x8 chains against the oracle with the same random stream
    for j, (l, nt, prec) in enumerate(CHAINS):
        if l != 8:
            continue
        om = streams[j][0]
        lib = oracle()
        h = HMC()
        assert lib.orc_hmc_init(C.byref(h), C.byref(om.m), 1000 + j, 10, 0.0, None) == 0
        dH = C.c_double()
        for it in range(NTRAJ):
            acc = lib.orc_trajectory(C.byref(h), None, None, C.byref(dH), None, None)
            a, d, ld = solo[j][0][it]
            assert a == acc, (CHAINS[j], it)
            tol_dh, tol_ld = (1e-8, 1e-10) if prec == 0 else (1e-5, 1e-6)
            assert abs(d - dH.value) < tol_dh, (CHAINS[j], it, d, dH.value)
            assert abs(ld - h.logdet) <= tol_ld * abs(h.logdet), (CHAINS[j], it)
        lib.orc_hmc_free(C.byref(h))
    for ctx in ctxs:
        ctx.close()

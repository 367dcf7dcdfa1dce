"""Many chains advanced with one launch per phase (ephmc_batch_*).

Contexts of uneven lattice and Nt (L 8..24, Nt 32..128), FP64 and mixed, with and without
checkpoints, some autotuned and one with fewer MD steps, run the same random inputs alone
(ephmc_trajectory, with the force width the batch set) and as one batch: every chain's records, final
field and M_hat^{-1} are bit-identical over several trajectories (the batch runs the single-chain
kernels' bodies with each chain's own grid for every sum).  The solo path itself is pinned against the oracle by the
configuration tests.  The same holds at the 64-chain workload's shape (64 chains of L 8..24, Nt 32..128) and for
non-square lattices across the batch's range (N = 33 .. 1024).
Modes the batch does not run are refused with EPHMC_ERR_UNSUPPORTED, and a context held by a batch can
be neither tuned nor destroyed until the batch is."""
import ctypes as C

import numpy as np
import pytest

from _gpu import ephmc, gpu_context, say
from _oracle import MT, OracleModel, oracle, ptr

pytestmark = pytest.mark.gpu

# (L, Nt, precision, MD steps, checkpoint k, autotune)
CHAINS = [(8, 32, 1, 10, 0, False), (8, 64, 0, 10, 0, True), (12, 128, 1, 10, 16, False), (16, 32, 1, 10, 0, True),
          (20, 64, 1, 10, 0, False), (24, 32, 1, 10, 16, True), (12, 64, 0, 10, 8, False), (16, 128, 1, 6, 0, False)]
NTRAJ = 3


def draws_for(seed, dim, n):
    lib = oracle()
    m = MT()
    lib.orc_mt_seed(C.byref(m), seed)
    out = []
    for _ in range(n + 1):
        v = np.zeros(dim)
        lib.orc_random_vector(C.byref(m), ptr(v), dim)
        out.append((v, lib.orc_mt_real2(C.byref(m))))
    return out


def rec_tuple(r):
    return (r.accepted, r.boltzmann, r.energy, r.logdet, r.dH)


def test_batch_equals_solo():
    E = ephmc()
    ctxs, inputs = [], []
    for i, (l, nt, prec, steps, k, tune) in enumerate(CHAINS):
        om = OracleModel(l, l, nt)
        ctx = gpu_context(om, steps=steps, checkpoint=k, precision=prec, graph=1)
        if tune:
            ctx.autotune()
        ctxs.append(ctx)
        inputs.append(draws_for(2000 + i, om.dim, NTRAJ))
    say("batch: %d contexts" % len(ctxs))
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
    # the batch sets the contexts' force width; the contexts then run alone between batch trajectories
    b = E.Batch(ctxs)
    solo = []
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
        recs = [rec_tuple(ctx.trajectory(nrm, u)) for nrm, u in d[1:]]
        solo.append((recs, ctx.chain_x(), ctx.chain_minv()))
    say("batch: solo runs done")
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
    info = b.info()
    say("batch:", info)
    # groups: (FP64 or mixed) x workgroup class x force columns x MD steps
    assert info["groups"] >= 4 and info["launches"] > 0
    got = [[] for _ in ctxs]
    for t in range(1, NTRAJ + 1):
        recs = b.trajectory([d[t][0] for d in inputs], [d[t][1] for d in inputs], wait=(t != 2))
        if recs is None:
            recs = b.collect()
        for j, r in enumerate(recs):
            got[j].append(rec_tuple(r))
    b.close()
    for j, ctx in enumerate(ctxs):
        assert got[j] == solo[j][0], (CHAINS[j], got[j], solo[j][0])
        np.testing.assert_array_equal(ctx.chain_x(), solo[j][1])
        np.testing.assert_array_equal(ctx.chain_minv(), solo[j][2])
    acc = sum(r[0] for g in got for r in g)
    say("batch: records, fields and inverses equal the solo runs (%d of %d accepted)" % (acc, NTRAJ * len(ctxs)))
    # the contexts go on alone after the batch
    for ctx, d in zip(ctxs, inputs):
        ctx.trajectory(d[1][0], d[1][1])
    for ctx in ctxs:
        ctx.close()


def test_batch_refuses_unsupported():
    E = ephmc()
    om = OracleModel(8, 8, 32)
    ready = gpu_context(om, steps=4)
    ready.chain_init(draws_for(7, om.dim, 0)[0][0])
    fresh = gpu_context(om, steps=4)
    with pytest.raises(E.EphmcError, match="error -1"):
        E.Batch([ready, fresh])                 # no chain state yet
    nofa = gpu_context(om, steps=4, no_fourier_acc=1)
    nofa.chain_init(draws_for(8, om.dim, 0)[0][0])
    with pytest.raises(E.EphmcError, match="error -3"):
        E.Batch([ready, nofa])
    small = OracleModel(4, 4, 32)
    tiny = gpu_context(small, steps=4)
    tiny.chain_init(draws_for(9, small.dim, 0)[0][0])
    with pytest.raises(E.EphmcError, match="error -3"):
        E.Batch([tiny])
    with pytest.raises(E.EphmcError, match="error -1"):
        E.Batch([ready, ready])                 # one context twice
    for arr in ((C.c_void_p * 2)(None, ready.h.value), (C.c_void_p * 2)(ready.h.value, None)):
        out = C.c_void_p()
        assert E.LIB.ephmc_batch_create(arr, 2, C.byref(out)) == -1 and not out.value   # a null chain
    # a context held by a batch can be neither tuned nor destroyed
    other = gpu_context(om, steps=4)
    other.chain_init(draws_for(10, om.dim, 0)[0][0])
    b = E.Batch([ready, other])
    with pytest.raises(E.EphmcError, match="belongs to a batch"):
        ready.autotune()
    assert E.LIB.ephmc_destroy(other.h) == -1
    b.close()
    ready.autotune()
    for c in (ready, fresh, nofa, tiny, other):
        c.close()


def test_batch_64_uneven_chains_equal_solo():
    """The 64-chain workload's shape (L from {8, 12, 16, 20, 24}, Nt from {32, 64, 128}, mixed precision,
    Nmd = 10): two trajectories of every chain as one batch equal the chains run alone, bit for bit."""
    E = ephmc()
    rng = np.random.default_rng(0)
    Ls, Nts = (8, 12, 16, 20, 24), (32, 64, 128)
    ctxs, inputs = [], []
    for i in range(64):
        l, nt = Ls[rng.integers(5)], Nts[rng.integers(3)]
        om = OracleModel(l, l, nt)
        ctxs.append(gpu_context(om, steps=10, checkpoint=0, precision=1, graph=1))
        inputs.append(draws_for(3000 + i, om.dim, 2))
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
    b = E.Batch(ctxs)
    say("batch of 64:", b.info())
    solo = []
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
        recs = [rec_tuple(ctx.trajectory(nrm, u)) for nrm, u in d[1:]]
        solo.append((recs, ctx.chain_x()))
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
    got = [[] for _ in ctxs]
    for t in (1, 2):
        for j, r in enumerate(b.trajectory([d[t][0] for d in inputs], [d[t][1] for d in inputs])):
            got[j].append(rec_tuple(r))
    b.close()
    for j, ctx in enumerate(ctxs):
        assert got[j] == solo[j][0], j
        np.testing.assert_array_equal(ctx.chain_x(), solo[j][1])
    say("batch of 64: records and fields equal the solo runs")
    for ctx in ctxs:
        ctx.close()


def test_batch_ragged_sizes_equal_solo():
    """Non-square lattices across the batch's range, 33 <= N <= 1024 (N = 33, 63, 221, 675 and 1024:
    partial waves, partial Gauss-Jordan blocks, every workgroup class), FP64 and mixed: two
    trajectories as one batch equal the chains run alone, bit for bit."""
    E = ephmc()
    specs = [(3, 11, 32, 1), (7, 9, 64, 0), (13, 17, 32, 1), (25, 27, 32, 1), (32, 32, 32, 0), (3, 11, 64, 0)]
    ctxs, inputs = [], []
    for i, (l1, l2, nt, prec) in enumerate(specs):
        om = OracleModel(l1, l2, nt)
        ctxs.append(gpu_context(om, steps=6, checkpoint=0, precision=prec, graph=1))
        inputs.append(draws_for(4000 + i, om.dim, 2))
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
    b = E.Batch(ctxs)
    say("batch ragged:", b.info())
    solo = []
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
        recs = [rec_tuple(ctx.trajectory(nrm, u)) for nrm, u in d[1:]]
        solo.append((recs, ctx.chain_x(), ctx.chain_minv()))
    for ctx, d in zip(ctxs, inputs):
        ctx.chain_init(d[0][0])
    got = [[] for _ in ctxs]
    for t in (1, 2):
        for j, r in enumerate(b.trajectory([d[t][0] for d in inputs], [d[t][1] for d in inputs])):
            got[j].append(rec_tuple(r))
    b.close()
    for j, ctx in enumerate(ctxs):
        assert got[j] == solo[j][0], specs[j]
        np.testing.assert_array_equal(ctx.chain_x(), solo[j][1])
        np.testing.assert_array_equal(ctx.chain_minv(), solo[j][2])
    say("batch ragged: records, fields and inverses equal the solo runs")
    for ctx in ctxs:
        ctx.close()

"""Transient localisation on the GPU (ephmc_tl_hist): eigensolve + current matrix elements +
binning of hist_trans_loc (organic_trans_loc.c:25-83) against the oracle on small lattices and
against numpy (LAPACK eigh + dense current matrices + bincount) at N = 1024 and N = 1280.
Tolerances: bins 1e-9 relative (FP64; eigenvalue rounding can only move a pair across a bin edge
with probability ~1e-13 per pair), partition sums 1e-12."""
import ctypes as C

import numpy as np
import pytest

import _dense as D
from _gpu import gpu_context, rel, say
from _oracle import OracleModel, oracle, ptr

pytestmark = pytest.mark.gpu

CM = np.array([7.2, 3.6, -3.6, 0.0, 7.15, 7.15])   # contraction matrix of a = (1,0), b = (0.5,0.866) x 7.2


def tl_model(l1, l2, nbins):
    om = OracleModel(l1, l2, nbins, beta=40.0, static_disorder=True)
    width = oracle().orc_tl_width(om.ref, nbins)
    return om, width


def oracle_tl(om, xs, width, nbins):
    g = np.zeros(nbins)
    z = [oracle().orc_hist_trans_loc(om.ref, ptr(np.ascontiguousarray(x)), ptr(CM), width, ptr(g), nbins) for x in xs]
    return g, np.array(z)


def numpy_tl(om, x, width, nbins):
    n = om.ns
    H = D.hopping_matrix(om.nnt, n, x, om.t, om.kappa, -om.mu)
    E, V = np.linalg.eigh(H)
    nb = om.nnt[:6 * n].reshape(n, 6)
    G = [np.zeros((n, n)), np.zeros((n, n))]
    idx = np.arange(n)
    for k in range(3):
        f = om.t[k] - om.kappa[k] * x[3 * idx + k]
        Ik = np.zeros((n, n))
        np.add.at(Ik, (idx, nb[:, 2 * k]), f)
        np.add.at(Ik, (nb[:, 2 * k], idx), -f)
        for c in range(2):
            G[c] += CM[3 * c + k] * Ik
    J = [V.T @ Gc @ V for Gc in G]                    # J[c][j, i] = v_j . G_c v_i
    Ej, Ei = E[:, None], E[None, :]
    fr = (Ej - Ei) / width + 0.5
    ok = (Ej >= Ei) & (fr < nbins)
    val = (np.exp(-Ei) + np.exp(-Ej)) * (J[0] ** 2 + J[1] ** 2) / (width * n)
    g = np.bincount(fr[ok].astype(np.int64), weights=val[ok], minlength=nbins)
    return g, np.exp(-E).sum() / n


def draws(om, nsamples, seed):
    return np.random.default_rng(seed).standard_normal((nsamples, 3 * om.ns))


@pytest.mark.parametrize("l1,l2,ns_", [(1, 1, 3), (1, 2, 3), (2, 1, 3), (2, 2, 4), (3, 4, 5), (4, 4, 5), (6, 6, 5), (5, 7, 4),
                                       (8, 8, 3), (9, 7, 3), (10, 11, 2)])
def test_tl_small_lattices_vs_oracle(l1, l2, ns_):
    """N = 1 .. 110 against the oracle.  From N = 63 on (here 63, 64 and 110: a ragged last row
    block, a full one, four reflector blocks) the back-transformation runs in compact-WY blocks on
    the matrix cores; below, one reflector at a time."""
    nbins = 32
    say("tl", l1, l2, ns_)
    om, width = tl_model(l1, l2, nbins)
    xs = draws(om, ns_, l1 * 10 + l2)
    ctx = gpu_context(om)
    g, z = ctx.tl_hist(xs, CM, width, nbins)
    go, zo = oracle_tl(om, xs, width, nbins)
    assert np.abs(z - zo).max() <= 1e-12 * np.abs(zo).max()
    assert rel(g, go) < 1e-9
    np.testing.assert_allclose(g, go, rtol=1e-8, atol=1e-12 * np.abs(go).max())


def test_tl_many_batches_vs_oracle(monkeypatch):
    """More samples than one batch holds (a 10 MiB workspace budget: batches of 183 samples at 6x6,
    four batches; fixed-order accumulation)."""
    nbins = 24
    om, width = tl_model(6, 6, nbins)
    xs = draws(om, 600, 5)
    monkeypatch.setenv("EPHMC_TL_BATCH_MB", "10")
    say("tl 6x6, 600 samples")
    ctx = gpu_context(om)
    g, z = ctx.tl_hist(xs, CM, width, nbins)
    go, zo = oracle_tl(om, xs, width, nbins)
    assert np.abs(z - zo).max() <= 1e-12 * np.abs(zo).max()
    assert rel(g, go) < 1e-9


@pytest.mark.parametrize("l1,l2,nsamples", [(16, 16, 3), (19, 27, 2), (23, 31, 2), (32, 32, 2), (32, 40, 1)])
def test_tl_large_vs_numpy(l1, l2, nsamples):
    """N = 256, 513 and 713 (ragged: 9 and 12 row blocks of the lower-triangle panels, the last
    one of 1 and 9 rows), 1024 (the 32x32 size; all with the compact-WY back-transformation) and
    1280 (two rows per thread, reflectors one at a time) against LAPACK eigh."""
    nbins = 512
    om, width = tl_model(l1, l2, nbins)
    xs = draws(om, nsamples, 7)
    say("tl numpy", l1, l2, nsamples)
    ctx = gpu_context(om)
    g, z = ctx.tl_hist(xs, CM, width, nbins)
    gn, zn = np.zeros(nbins), []
    for x in xs:
        gi, zi = numpy_tl(om, x, width, nbins)
        gn += gi
        zn.append(zi)
    assert np.abs(z - np.array(zn)).max() <= 1e-11 * max(zn)
    assert rel(g, gn) < 1e-9


def test_tl_wide_tiles_vs_numpy():
    """4096 bins at N = 256: pair tiles whose bins span 512 or more take the binary-search binning,
    the others the row-ordered histogram; both against LAPACK eigh."""
    nbins = 4096
    om, width = tl_model(16, 16, nbins)
    xs = draws(om, 2, 13)
    ctx = gpu_context(om)
    g, z = ctx.tl_hist(xs, CM, width, nbins)
    gn = np.zeros(nbins)
    for x in xs:
        gn += numpy_tl(om, x, width, nbins)[0]
    assert rel(g, gn) < 1e-9


def test_tl_repeatable_and_additive():
    """Two identical calls give identical bins (no atomics); bins of a split sample set add up."""
    nbins = 64
    om, width = tl_model(8, 8, nbins)
    xs = draws(om, 20, 11)
    ctx = gpu_context(om)
    g1, z1 = ctx.tl_hist(xs, CM, width, nbins)
    g2, z2 = ctx.tl_hist(xs, CM, width, nbins)
    np.testing.assert_array_equal(g1, g2)
    np.testing.assert_array_equal(z1, z2)
    ga, za = ctx.tl_hist(xs[:7], CM, width, nbins)
    gb, zb = ctx.tl_hist(xs[7:], CM, width, nbins)
    np.testing.assert_array_equal(np.concatenate([za, zb]), z1)
    assert rel(ga + gb, g1) < 1e-14

"""organic_conv (greens.bin post-processing; the reference's simulation/organic_conv.c, rewritten
with its own quadrature) against the reference's tool compiled from its sources into
oracle/_ref/ref_conv, on synthetic histograms (the reference holds no greens.bin fixture) and on the
reference's own smearing-list layout (simulation/smear.txt format).  Both print %g (six significant
digits).  The frequency table is the same closed formula: 2e-5 relative covers the last printed
digit.  The imaginary-time table integrates the Lorentzian-smeared kernel with a different rule
(adaptive Gauss-Kronrod here, the reference's double-exponential sum there); the reference's rule
stops early at wide smearing (3e-5 relative at g = 40 /beta, where scipy's adaptive quad agrees
with this tool to every printed digit), so that table is compared at 6e-5 relative, and this tool's
table is pinned separately against scipy.integrate.quad at the printed precision."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONV = os.path.join(ROOT, "electron-phonon-hmc_amd", "host", "organic_conv")
REF_CONV = os.path.join(ROOT, "oracle", "_ref", "ref_conv")


def write_greens(path, g, h):
    with open(path, "wb") as f:
        f.write(np.uint32(len(g)).tobytes())
        f.write(np.float64(h).tobytes())
        f.write(np.asarray(g, np.float64).tobytes())


@pytest.fixture(scope="module")
def tools():
    if not os.path.exists(REF_CONV):
        pytest.skip("oracle/_ref/ref_conv not built (needs the reference tree)")
    assert os.path.exists(CONV), "organic_conv not built"
    return CONV, REF_CONV


@pytest.mark.parametrize("n,h,widths,seed", [
    (48, 0.35, "40 48\n0\n2.7e-3\n6e-3\n1\n", 1),        # the layout of simulation/smear.txt
    (512, 0.04, "40 512\n0\n1e-3\n2e-2\n", 2),
    (7, 1.3, "12.5 6\n0.5\n", 3),
])
def test_conv_matches_reference(tools, n, h, widths, seed):
    ours, ref = tools
    g = np.abs(np.random.default_rng(seed).standard_normal(n)) * 1e-2
    with tempfile.TemporaryDirectory() as d:
        gb, sw = os.path.join(d, "greens.bin"), os.path.join(d, "smear.txt")
        write_greens(gb, g, h)
        open(sw, "w").write(widths)
        outs = []
        for exe in (ours, ref):
            w, t = os.path.join(d, "w_%s.txt" % os.path.basename(exe)), os.path.join(d, "t_%s.txt" % os.path.basename(exe))
            r = subprocess.run([exe, gb, sw, w, t], capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            outs.append((open(w).read(), open(t).read()))
    for ours_txt, ref_txt, tol in zip(outs[0], outs[1], (2e-5, 6e-5)):
        a = [[float(v) for v in line.split()] for line in ours_txt.splitlines()]
        b = [[float(v) for v in line.split()] for line in ref_txt.splitlines()]
        assert len(a) == len(b) and all(len(r) == len(q) for r, q in zip(a, b))
        np.testing.assert_allclose(np.array(a), np.array(b), rtol=tol, atol=0)
    assert len(outs[0][0].splitlines()) == n


def test_conv_usage_message(tools):
    ours, ref = tools
    a = subprocess.run([ours], capture_output=True, text=True, timeout=30)
    b = subprocess.run([ref], capture_output=True, text=True, timeout=30)
    assert a.stdout == b.stdout and a.returncode == b.returncode == 0


def _smeared_kernel_scipy(t, w, g):
    """K_g(t, w) = int L_g(E - w) cosh((1/2 - t) E) / cosh(E/2) dE by scipy's adaptive quadrature
    over E = w + g tan(theta)."""
    from scipy.integrate import quad

    def k(E):
        a = abs(E)
        return (np.exp(-t * a) + np.exp(-(1 - t) * a)) / (1 + np.exp(-a))
    if g == 0:
        return k(w)
    return quad(lambda th: k(w + g * np.tan(th)), -np.pi / 2, np.pi / 2, epsabs=1e-14, epsrel=1e-13,
                limit=500)[0] / np.pi


@pytest.mark.parametrize("n,h,widths", [(24, 0.35, "40 16\n0\n2.7e-3\n6e-3\n1\n"), (9, 1.1, "10 12\n0.05\n0.9\n")])
def test_time_table_against_scipy_quad(tools, n, h, widths):
    ours, _ = tools
    g = np.abs(np.random.default_rng(n).standard_normal(n)) * 1e-2
    beta, nt = [float(v) for v in widths.split()[:2]]
    gs = [float(v) * beta for v in widths.split()[2:]]
    with tempfile.TemporaryDirectory() as d:
        gb, sw = os.path.join(d, "greens.bin"), os.path.join(d, "smear.txt")
        write_greens(gb, g, h)
        open(sw, "w").write(widths)
        w, t = os.path.join(d, "w.txt"), os.path.join(d, "t.txt")
        subprocess.run([ours, gb, sw, w, t], check=True, timeout=60)
        tab = np.loadtxt(t, ndmin=2)
    assert tab.shape == (int(nt) // 2 + 1, 1 + len(gs))
    for i, tt in enumerate(tab[:, 0]):
        for k, gg in enumerate(gs):
            ref = h * sum(g[j] * _smeared_kernel_scipy(tt, j * h, gg) for j in range(n))
            assert tab[i, 1 + k] == pytest.approx(ref, rel=6e-6), (tt, gg)
    # t = 0: K_g = 1 for every width (the Lorentzian has unit mass), so G(0) = h sum sigma
    np.testing.assert_allclose(tab[0, 1:], h * g.sum(), rtol=6e-6)

"""organic_conv (greens.bin post-processing, the reference's simulation/organic_conv.c) against
the reference's own tool compiled from its sources into oracle/_ref/ref_conv: identical text
output for the frequency and imaginary-time tables, on synthetic histograms (the reference holds
no greens.bin fixture) and on the reference's own smearing list (simulation/smear.txt format)."""
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
    assert outs[0][0] == outs[1][0]
    assert outs[0][1] == outs[1][1]
    assert len(outs[0][0].splitlines()) == n


def test_conv_usage_message(tools):
    ours, ref = tools
    a = subprocess.run([ours], capture_output=True, text=True, timeout=30)
    b = subprocess.run([ref], capture_output=True, text=True, timeout=30)
    assert a.stdout == b.stdout and a.returncode == b.returncode == 0

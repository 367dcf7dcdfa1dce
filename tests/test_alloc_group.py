"""Host-side check of the correlator's all-or-none buffer allocation (csrc/alloc_group.h): with an
allocator that fails at the k-th buffer, every buffer made before it is released and every pointer
of the group is null afterwards, so a later call allocates again instead of running on a partial
set.  Compiled with g++ (no GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "electron-phonon-hmc_amd", "csrc")

SRC = r"""
#include "alloc_group.h"
#include <cstdio>
#include <cstdlib>
#include <set>
int main(int argc, char **argv){
	const int fail_at = atoi(argv[1]);
	int calls = 0;
	std::set<void *> live;
	double *a = (double *)1, *b = (double *)1, *c = (double *)1;   // stale values
	void **ptrs[3] = {(void **)&a, (void **)&b, (void **)&c};
	const size_t bytes[3] = {16, 32, 64};
	const bool ok = ephmc::alloc_all_or_none(ptrs, bytes, 3,
		[&](void **p, size_t n){ if(calls++ == fail_at) return false; *p = malloc(n); live.insert(*p); return true; },
		[&](void *p){ live.erase(p); free(p); });
	printf("%d %d %d %d %zu\n", ok, a != nullptr, b != nullptr, c != nullptr, live.size());
	for(void *p : live) free(p);
	return 0;
}
"""


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    d = tmp_path_factory.mktemp("alloc")
    src = d / "t.cpp"
    src.write_text(SRC)
    exe = d / "t"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + HDR, str(src), "-o", str(exe)], check=True)
    return str(exe)


@pytest.mark.parametrize("fail_at,expect", [(-1, "1 1 1 1 3"), (0, "0 0 0 0 0"), (1, "0 0 0 0 0"), (2, "0 0 0 0 0")])
def test_all_or_none(prog, fail_at, expect):
    out = subprocess.run([prog, str(fail_at)], check=True, capture_output=True, text=True).stdout.strip()
    assert out == expect

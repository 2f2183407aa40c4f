"""The header swap, compile-only (CPU): each of the reference's six numa-*.C
programs (and numa-BFS-async-pipe.C), copied from the reference tree into a scratch directory beside
include/polymer.h (oracle/header_swap.py), compiles with hipcc --hipstdpar and no
change to its text; numa-PageRank.C, numa-BFS.C and numa-BP.C also compile with
their host-side edge loops pointed at the header's edge maps (the form that runs
on the GPU, tests/test_gpu_header_swap.py).  Runs where the reference sources
exist (the build container); skips elsewhere."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import header_swap as hs  # noqa: E402

needs_ref = pytest.mark.skipif(not hs.available(), reason="reference sources not present")


@needs_ref
def test_six_programs_compile_with_only_the_header_swapped():
    res = hs.build_all(reroute=False, link=False, jobs=6)
    bad = {a: log for a, rc, log in res if rc != 0}
    assert not bad, bad


@needs_ref
def test_rerouted_programs_compile():
    for app in ("numa-PageRank", "numa-BFS", "numa-BP"):
        a, rc, log = hs.compile_one(app, reroute=True, link=False)
        assert rc == 0, (a, log)

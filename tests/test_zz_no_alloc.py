"""pn_totals_noalloc (the totals on the library's file-scope scratch, no allocation) against
pn_totals through the C ABI: bit-identical on two calls in a row, the second on a smaller
disk, for both precisions."""
from __future__ import annotations

import numpy as np
import pytest

from penguin_2_0_amd import abi, configs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("prec", [abi.PN_FP64, abi.PN_FP32], ids=["fp64", "fp32"])
def test_totals_noalloc_equals_totals_on_two_calls(prec):
    # the larger disk runs the reduction's full 2048 blocks, the smaller one 27
    for w in (configs.c1(), configs.small(48, 144)):
        say("case: totals_noalloc", w.name, "fp64" if prec == abi.PN_FP64 else "fp32")
        d = abi.init_disk(w.grid, w.params, w.seed, w.noise)
        with abi.Context(w.grid, w.params, prec) as ctx:
            ctx.load(d)
            st = ctx.run_begin()
            ctx.run(3, st, graph=True)
            got = ctx.totals_noalloc()
            want = ctx.totals()
            again = ctx.totals_noalloc()
        say(f"  mass {got[0]!r} / {want[0]!r}, angmom {got[1]!r} / {want[1]!r}")
        assert got == want and again == want
        assert np.isfinite(got[0]) and got[0] > 0.0

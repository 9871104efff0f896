"""GPU: cmx_op_time, the measurement entry point behind the bench (every selector runs and returns a
positive time; an unknown selector is an argument error)."""
import math

import numpy as np
import pytest

from chroma_amd import lib as cmx

pytestmark = pytest.mark.gpu


def test_op_time_selectors():
    L = (8, 8, 8, 16)
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=41, eps=0.3), (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, 0.1, 1.0, sloppy_prec=cmx.HALF)
    for prec in (cmx.DOUBLE, cmx.SINGLE, cmx.HALF):
        for what in (0, 1, 2, 3, 5):
            ms = op.time(what, prec, 3)
            print(f"prec {prec} what {what}: {ms * 1e3:.1f} us", flush=True)
            assert math.isfinite(ms) and ms > 0
    ms = op.time(4, cmx.HALF, 2)
    assert math.isfinite(ms) and ms > 0
    with pytest.raises(cmx.CmxError):
        op.time(6, cmx.HALF, 1)

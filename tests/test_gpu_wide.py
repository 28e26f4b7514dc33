"""Very wide disparity ranges: lane groups R = 32 and 64 (D > 1024, up to the
4096 of GpuStereo.plan, GpuStereo.java:43-47) through the resident,
strip-streamed and compact plans, bit-exact to the oracle.  D = 3000 with 12
rows also covers last-sweep flushes of fewer than 18 steps while the flush
arguments are spilled (64 disparities per lane)."""
import pytest

from oracle import oracle as O
from pose_amd import sgm
from tests.cases import shifted
from tests.test_gpu_parity import assert_same, to_sgm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D,agg", [(1500, 0), (3000, 0), (3000, 4)])
def test_wide_ranges_all_plans(gpu_context, D, agg):
    """agg 4: 32-bit aggregates (the generic sweep); the compact plan refuses them."""
    w, h = D + 120, 12
    a, b = shifted(500 + D, w, h, 37)
    op = O.params(disparities=D, min_disparity=-20, aggregate_bytes=agg)
    print("wide D", D, "oracle", flush=True)
    want = O.compute(op, a, b, w, h)
    p = to_sgm(op)
    print("wide D", D, "resident", flush=True)
    assert_same(gpu_context.match_pair(p, a, b), want, "resident D=%d" % D)
    for mode, rows in ((sgm.PLAN_STREAMED, 5), (sgm.PLAN_ESGM, 5)):
        print("wide D", D, "agg", agg, "plan", mode, flush=True)
        gpu_context.plan_mode(mode, rows)
        try:
            if mode == sgm.PLAN_ESGM and agg == 4:
                with pytest.raises(sgm.SgmError):
                    gpu_context.match_pair(p, a, b)
                continue
            got = gpu_context.match_pair(p, a, b)
        finally:
            gpu_context.plan_mode(sgm.PLAN_AUTO)
        assert_same(got, want, "plan %d D=%d agg %d" % (mode, D, agg))



def test_disparity_limit_reaches_4096(gpu_context):
    """GpuStereo.limit's binary search (GpuStereo.java:86-94) tops out at 4096
    when memory allows."""
    assert gpu_context.disparity_limit(to_sgm(O.params(disparities=64)), 4200, 12) == 4096

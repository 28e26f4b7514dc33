"""CPU checks of the C-ABI library: it loads, exports every declared symbol, and its
struct layouts match the header.  No compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from pose_amd import sgm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sgm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sgm_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sgm.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(sgm.EXPORTS)


def test_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "sgm_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(sgm_params), sizeof(sgm_stats),
         sizeof(sgm_device_info), sizeof(sgm_plan), sizeof(sgm_pair), sizeof(sgm_request_info),
         offsetof(sgm_params, uniqueness), offsetof(sgm_stats, valid),
         sizeof(sgm_filter_params), offsetof(sgm_filter_params, median_max_change));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(sgm.SgmParams), ctypes.sizeof(sgm.SgmStats), ctypes.sizeof(sgm.SgmDeviceInfo),
            ctypes.sizeof(sgm.SgmPlan), ctypes.sizeof(sgm.SgmPair), ctypes.sizeof(sgm.SgmRequestInfo),
            sgm.SgmParams.uniqueness.offset, sgm.SgmStats.valid.offset,
            ctypes.sizeof(sgm.SgmFilterParams), sgm.SgmFilterParams.median_max_change.offset]
    assert got == want


def test_oracle_params_share_the_layout():
    from oracle import oracle as O
    assert [f for f, _ in O.OracleParams._fields_] == [f for f, _ in sgm.SgmParams._fields_]
    assert ctypes.sizeof(O.OracleParams) == ctypes.sizeof(sgm.SgmParams)
    assert ctypes.sizeof(O.OracleStats) == ctypes.sizeof(sgm.SgmStats)
    assert [f for f, _ in O.OracleFilter._fields_][0] == "median"
    assert ctypes.sizeof(O.OracleFilter) == ctypes.sizeof(sgm.SgmFilterParams)
    assert all(getattr(O.OracleFilter, f).offset == getattr(sgm.SgmFilterParams, f).offset
               for f, _ in O.OracleFilter._fields_)


def test_error_paths_without_device():
    if sgm.device_count() > 0:
        pytest.skip("a GPU is visible; the no-device path is not reachable")
    with pytest.raises(sgm.SgmError) as e:
        sgm.Context(0)
    assert e.value.code == -2
    assert "No GPU" in str(e.value)
    # the reference's GpuSgm.devices returns an empty list without a runtime
    assert sgm.device_count() == 0


def test_null_arguments_are_rejected():
    lib = sgm.lib()
    rc = lib.sgm_match_pair(None, None, None, None, None, None, 10, 10, None, None, None, None,
                            None, None)
    assert rc == -1
    assert b"null" in lib.sgm_last_error()


def test_options_mirror_reference_units():
    o = sgm.Options(census=9, p1=10, p2=120, rawPenalties=True, censusWidth=9, censusHeight=7)
    p = sgm.params_from_options(o)
    assert (p.window_w, p.window_h, p.p1, p.p2) == (9, 7, 10, 120)
    o = sgm.Options(census=5)
    p = sgm.params_from_options(o)
    assert (p.p1, p.p2, p.aggregate_bytes, p.census_weight) == (4, 24, 2, 500)
    with pytest.raises(ValueError):
        sgm.StereoCost.aggregateBytes(sgm.Options(p2=10000, sgmAggregate="16"))

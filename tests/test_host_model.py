"""CPU replay of the randomised GPU sweep through the host model of the
library's table indexing (tests/host_model.py): every build, attach, scan,
scatter, start-table read, pending list and voxel workspace of the sweep's
seeded cases checked against the allocation it addresses, with the exact-kNN
engine's handle (reserve, no-sync builds, synchronising retries) carried from
case to case as in one test process.
"""
import re
from pathlib import Path

import numpy as np
import pytest

import host_model as hm

CSRC = Path(__file__).resolve().parents[1] / "flashknn_exprepo_amd" / "csrc"


def test_model_constants_match_the_sources():
    """The sizes the model assumes are the ones the kernels are built with."""
    build = (CSRC / "fk_build.hip").read_text()
    internal = (CSRC / "fk_internal.h").read_text()
    api = (CSRC / "fk_api.hip").read_text()
    assert re.search(r"constexpr int kBlock = 256;", build)
    assert re.search(r"constexpr int kScanItems = 16;", build)
    assert re.search(r"constexpr int kRankBlock = 1024;", internal)
    assert re.search(r"constexpr int kRankSlots = 2 \* kRankBlock;", internal)
    assert re.search(r"return B <= 1024 \? 16 : 1;", internal)
    assert re.search(r"constexpr int64_t kHardCellCap = 1ll << 30;", api)
    assert re.search(r"int64_t n = need \+ need / 4 \+ 64;", api)
    assert re.search(r"const int64_t c = cells \+ cells / 4 \+ 64;", api)
    assert re.search(r"constexpr int kMaxLaneK = 64;", internal)


@pytest.mark.parametrize("lo,hi", [(0, 24), (24, 96)])
def test_sweep_cases_stay_in_bounds(oracle, lo, hi):
    eng = hm.Engine()
    for seed in range(lo, hi):
        case = hm.sweep_case(seed)
        sel = oracle.voxel_first(case["parts"][0], case["vox"])
        try:
            hm.run_case(eng, case, sel)
        except hm.Violation as e:
            raise AssertionError(f"seed {seed}: {e}") from None


def test_overflowing_no_sync_build_then_retry():
    """A no-sync build whose boxes outgrow the reserved cells raises
    FK_ERR_GRID_TOO_LARGE with kTotalCells clamped to the reserve (bases of
    later clouds clamped too), every query kernel exits, the synchronising
    rebuild grows the cell tables past the query tables, and the next no-sync
    build on the same handle stays within both."""
    rng = np.random.default_rng(7)
    h = hm.Handle()
    h.reserve(30000, 40000, 3, 30000)
    parts = [rng.integers(0, 20, (5000, 3)), rng.integers(0, 400, (5000, 3)),
             rng.integers(0, 20, (5000, 3))]
    cell = np.concatenate(parts).astype(np.int64)
    cloud = np.repeat(np.arange(3), 5000).astype(np.int64)
    assert h.build(cell, cloud, 3, hm.NO_SYNC, True) == hm.OK
    assert h.dev_status == hm.ERR_TOO_LARGE
    assert h.total == h.c("counts") - 1 == h.build_cap
    h.query(16, True, False)
    assert h.status() == hm.ERR_TOO_LARGE
    assert h.build(cell, cloud, 3, 0, True) == hm.OK       # synchronising retry
    assert h.total == 8000 + 400 ** 3 + 8000
    assert h.c("counts") > h.c("qcounts")
    h.query(16, True, False)
    small = cell % 20
    assert h.build(small, cloud, 3, hm.NO_SYNC, True) == hm.OK
    assert h.build_cap == h.c("qcounts") - 1
    h.attach(small, cloud, hm.NO_SYNC)
    h.query(16, True, True)
    assert h.status() == hm.OK


def test_model_catches_a_short_scan_launch():
    """A planted mistake (on the CPU, in the model): a scan launched for fewer
    entries than kTotalCells leaves the total unwritten."""
    h = hm.Handle()
    h.ensure_cells(100000)
    with pytest.raises(hm.Violation, match="scan total"):
        h._scan("counts", "cell_start", 50000, 40000, 0)


def test_attach_refuses_a_build_beyond_the_query_tables():
    """A no-sync build held only to the cell tables (a planted mistake in the
    model: the query-table hold of finish_build left out) after a
    synchronising build grew them: the no-sync attach refuses it with
    FK_ERR_CAPACITY instead of bucketing past the query counts."""
    h = hm.Handle()
    h.reserve(1000, 5000, 1, 1000)
    cell = np.stack([np.arange(1000) * 7, np.zeros(1000), np.zeros(1000)], 1).astype(np.int64)
    cloud = np.zeros(1000, np.int64)
    assert h.build(cell, cloud, 1, 0, True) == hm.OK       # grows counts to 7000+
    real = h.c("qcounts")
    assert h.c("counts") > real
    h.cap["qcounts"] = h.c("counts")                        # the planted mistake
    assert h.build(cell, cloud, 1, hm.NO_SYNC, True) == hm.OK
    h.cap["qcounts"] = real
    with pytest.raises(hm.LibError) as e:
        h.attach(cell, cloud, hm.NO_SYNC)
    assert e.value.status == hm.ERR_CAPACITY

"""The drop-in GPU bodies inside the reference's own query unit.

oracle/_ref/libdropin_tpch.so is the reference's query_q*.cpp with
bespokeolap_artifacts_amd/dropin/tpch_run_q.cpp in place of the Q1, Q3, Q6,
Q14, Q17 and Q18 files: the reference's parse_qN, then the drop-in run_qN
(device mirror of the Database + libbolap) and write_qN_csv.  For the same
Database and request lines, its CSVs must equal those of the unchanged
reference (oracle/_ref/libref_tpch.so), byte for byte.
"""
from __future__ import annotations

import pytest

from tests import tpch_data as T
from tests import tpch_ref as R

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (R.LIB.exists() and R.DROPIN.exists()),
                                 reason="oracle/_ref not built")]

LINES = ['1 90', '1 "3"', '1 "2000"', '3 "BUILDING" "1995-03-15"', '3 "MACHINERY" "1993-01-01"',
         '3 "NOSUCH" "1995-03-15"', '6 "1994-01-01" "0.06" "24"', '6 "1997-01-01" "0.01" "40"',
         '6 "1993-02-28" "0.10" "1.5"', '14 "1995-09-01"', '14 "1992-01-01"', '14 "2001-01-01"',
         '17 "Brand#23" "MED BOX"', '17 "Brand#11" "SM CASE"', '17 "Brand#99" "SM CASE"',
         '18 "200"', '18 "300"', '18 "<<NULL>>"', '4 "1995-01-01"']


@pytest.mark.parametrize("orders", [1, 20_001, 250_001])
def test_dropin_equals_the_reference(orders):
    print(f"dropin vs reference: orders={orders}", flush=True)
    r = T.raw(71 + orders, orders, parts=2000, custs=1501, supps=100)
    want = R.RefTpch(r).run(LINES)
    dropin = R.RefTpch(r, R.DROPIN)
    got = dropin.run(LINES)
    for line, w, g in zip(LINES, want, got):
        assert g == w, line
    dropin.close()
    if orders > 1:
        assert sum(w.count("\n") > 1 for w in want) >= len(LINES) - 4


def test_dropin_mirror_follows_the_database():
    """Two databases in a row through one process: the second call must not
    see the first one's device mirror."""
    lines = ['1 90', '6 "1994-01-01" "0.06" "24"', '3 "BUILDING" "1995-03-15"']
    for seed in (3, 4):
        r = T.raw(seed, 5000, parts=500, custs=301, supps=40)
        want = R.RefTpch(r).run(lines)
        d = R.RefTpch(r, R.DROPIN)
        assert d.run(lines) == want
        d.close()

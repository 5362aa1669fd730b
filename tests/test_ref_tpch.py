"""The oracle against the reference's own TPC-H pipelines.

oracle/_ref/libref_tpch.so is the reference's query_q*.cpp compiled unchanged
(with a shim that fills its Database; tests/tpch_ref.py).  Every served
pipeline's CSV, as the oracle plus Python restatements of the reference
writers produce it, must equal the CSV that the reference's run_qN and
write_qN_csv write for the same database and request line, byte for byte.
Skipped where the reference could not be built (no sources / no Arrow headers).
"""
from __future__ import annotations

import pytest

from tests import tpch_data as T
from tests import tpch_ref as R
from tests import test_tpch_driver_gpu as D

pytestmark = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")

Q16_S1 = ["49", "14", "23", "45", "19", "3", "36", "9"]
Q16_S2 = ["1", "2", "<<NULL>>", "x7", "3", "4", "5", "6"]
Q22_A = ["13", "31", "23", "29", "30", "18", "17"]
Q22_B = ["AB", "20", "7", "<<NULL>>", "20", "AB", "34"]


def quoted(v):
    return " ".join(f'"{x}"' for x in v)


def cases(db, r):
    """(request line, oracle-side CSV) for every pipeline from Q2 on."""
    return [
        ('2 "15" "BRASS" "EUROPE"', lambda: D.csv_q2(db, r, "15", "BRASS", "EUROPE")),
        ('2 "3" "TIN" "ASIA"', lambda: D.csv_q2(db, r, "3", "TIN", "ASIA")),
        ('2 "7" "" "AFRICA"', lambda: D.csv_q2(db, r, "7", "", "AFRICA")),
        ('2 "15" "BRASS" "ATLANTIS"', lambda: D.csv_q2(db, r, "15", "BRASS", "ATLANTIS")),
        ('11 "GERMANY" "0.0001"', lambda: D.csv_q11(db, r, "GERMANY", "0.0001")),
        ('11 "PERU" "0.02"', lambda: D.csv_q11(db, r, "PERU", "0.02")),
        ('11 "ATLANTIS" "0.0001"', lambda: D.csv_q11(db, r, "ATLANTIS", "0.0001")),
        ('15 "1996-01-01" "0"', lambda: D.csv_q15(db, r, "1996-01-01")),
        ('15 "1993-11-30" "7"', lambda: D.csv_q15(db, r, "1993-11-30")),
        ('15 "2010-01-01" "0"', lambda: D.csv_q15(db, r, "2010-01-01")),
        ('20 "forest" "1994-01-01" "CANADA"',
         lambda: D.csv_q20(db, r, "forest", "1994-01-01", "CANADA")),
        ('20 "" "1995-01-01" "GERMANY"', lambda: D.csv_q20(db, r, "", "1995-01-01", "GERMANY")),
        ('20 "g" "1996-02-29" "ATLANTIS"',
         lambda: D.csv_q20(db, r, "g", "1996-02-29", "ATLANTIS")),
        ('20 "<<NULL>>" "1994-01-01" "CANADA"',
         lambda: D.csv_q20(db, r, "<<NULL>>", "1994-01-01", "CANADA")),
        ('4 "1995-01-01"', lambda: D.csv_q4(db, "1995-01-01")),
        ("4 1993-07-31", lambda: D.csv_q4(db, "1993-07-31")),
        ('5 "ASIA" "1994-01-01"', lambda: D.csv_q5(db, "ASIA", "1994-01-01")),
        ("5 EUROPE 1996-02-29", lambda: D.csv_q5(db, "EUROPE", "1996-02-29")),
        ('5 "ATLANTIS" "1994-01-01"', lambda: D.csv_q5(db, "ATLANTIS", "1994-01-01")),
        ('7 "FRANCE" "GERMANY"', lambda: D.csv_q7(db, r, "FRANCE", "GERMANY")),
        ('7 "JAPAN" "JAPAN"', lambda: D.csv_q7(db, r, "JAPAN", "JAPAN")),
        ('7 "ATLANTIS" "FRANCE"', lambda: D.csv_q7(db, r, "ATLANTIS", "FRANCE")),
        ('8 "BRAZIL" "AMERICA" "ECONOMY ANODIZED STEEL"',
         lambda: D.csv_q8(db, r, "BRAZIL", "AMERICA", "ECONOMY ANODIZED STEEL")),
        ('8 "CHINA" "ASIA" "NO SUCH TYPE"', lambda: D.csv_q8(db, r, "CHINA", "ASIA", "NO SUCH TYPE")),
        ('9 "green"', lambda: D.csv_q9(db, r, "green")),
        ('9 "zzz"', lambda: D.csv_q9(db, r, "zzz")),
        ('10 "1993-10-01"', lambda: D.csv_q10(db, r, "1993-10-01")),
        ("10 2010-01-01", lambda: D.csv_q10(db, r, "2010-01-01")),
        ('12 "MAIL" "SHIP" "1994-01-01"', lambda: D.csv_q12(db, "MAIL", "SHIP", "1994-01-01")),
        ("12 AIR NOSUCH 1995-01-01", lambda: D.csv_q12(db, "AIR", "NOSUCH", "1995-01-01")),
        ('13 "special" "requests"', lambda: D.csv_q13(db, "special", "requests")),
        ('13 "x" ""', lambda: D.csv_q13(db, "x", "")),
        (f'16 "Brand#45" "MEDIUM POLISHED" {quoted(Q16_S1)}',
         lambda: D.csv_q16(db, "Brand#45", "MEDIUM POLISHED", Q16_S1)),
        (f'16 "Brand#11" "PROMO" {quoted(Q16_S2)}', lambda: D.csv_q16(db, "Brand#11", "PROMO", Q16_S2)),
        ('19 "1" "10" "20" "Brand#12" "Brand#23" "Brand#34"',
         lambda: D.csv_q19(db, ("1", "10", "20"), ("Brand#12", "Brand#23", "Brand#34"))),
        ('19 "5" "<<NULL>>" "25" "Brand#11" "Brand#21" "Brand#55"',
         lambda: D.csv_q19(db, ("5", "<<NULL>>", "25"), ("Brand#11", "Brand#21", "Brand#55"))),
        ('21 "SAUDI ARABIA"', lambda: D.csv_q21(db, r, "SAUDI ARABIA")),
        ('21 "ATLANTIS"', lambda: D.csv_q21(db, r, "ATLANTIS")),
    ] + ([
        (f"22 {quoted(Q22_A)}", lambda: D.csv_q22(db, Q22_A)),
        (f"22 {quoted(Q22_B)}", lambda: D.csv_q22(db, Q22_B)),
    ] if q22_defined(r) else [])


def q22_defined(r) -> bool:
    """The reference's Q22 reads orders_by_customer_offsets[custkey + 1] for
    every candidate (query_q22.cpp:351-355), one past the array for a customer
    key above the largest order custkey (builder_impl.cpp:631-633), so its
    answer is defined only when no customer key exceeds that bound."""
    return int(r["c_custkey"].max()) <= int(r["o_custkey"].max())


@pytest.mark.parametrize("orders", [3, 12_001])
def test_oracle_csv_equals_the_reference(orders):
    r = T.raw(53 + orders, orders, parts=700, custs=901, supps=60)
    db = T.build(r)
    ref = R.RefTpch(r)
    todo = cases(db, r)
    got = ref.run([line for line, _ in todo])
    for (line, want), out in zip(todo, got):
        assert out is not None, line
        assert out == want(), line
    if orders > 3:
        # the data reaches every pipeline's non-empty branch
        assert q22_defined(r)
        nonempty = {line.split()[0] for (line, _), out in zip(todo, got) if out.count("\n") > 1}
        assert nonempty == {line.split()[0] for line, _ in todo}


def test_shim_reports_reference_errors():
    """A request the reference's parser rejects comes back as its message."""
    r = T.raw(5, 3, parts=20, custs=10, supps=8)
    ref = R.RefTpch(r)
    with pytest.raises(RuntimeError, match="Q4"):
        ref.run(["4"])
    assert ref.run(["99 x"]) == [None]


def test_dropin_compiles_and_links_against_the_reference_headers(tmp_path):
    """bespokeolap_artifacts_amd/dropin/tpch_run_q.cpp -- run_qN / write_qN_csv
    of Q1, Q3, Q6, Q14, Q17, Q18 over libbolap -- compiles against the
    reference's own query_qN.hpp / builder_impl.hpp / args_parser.hpp and links
    against libbolap.so and the HIP runtime with no undefined symbol."""
    import shutil
    import subprocess

    from bespokeolap_artifacts_amd import build as B
    src = B.REFERENCE / "bespoke_tpch"
    if not (src / "query_q1.hpp").exists() or B._arrow_dir() is None:
        pytest.skip("reference headers or Arrow headers absent")
    if not B.LIB.exists():
        pytest.skip("libbolap.so not built")
    out = tmp_path / "libdropin_only.so"
    cmd = [shutil.which("g++") or "c++", "-std=c++20", "-O1", "-fPIC", "-shared", "-Wall",
           "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{B.ROOT / 'include'}", f"-I{src}", f"-I{B._arrow_dir() / 'include'}",
           str(B.DROPIN_SRC), "-o", str(out), f"-L{B.PKG}", "-lbolap", "-L/opt/rocm/lib",
           "-lamdhip64", "-Wl,--no-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    syms = subprocess.run(["nm", "-DC", "--defined-only", str(out)], capture_output=True,
                          text=True).stdout
    for q in (1, 3, 6, 14, 17, 18):
        assert f"run_q{q}(Database const&, Q{q}Args const&)" in syms
        assert f"write_q{q}_csv(" in syms

"""The C-ABI library loads and exports exactly what include/bolap.h declares.

CPU-only: no call here touches a GPU.
"""
import ctypes as C
import pathlib
import re

import pytest

from bespokeolap_artifacts_amd import native

ROOT = pathlib.Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "bolap.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(bolap_[a-z0-9_]+)\s*\(", text))


def test_library_is_built():
    assert native.LIB_PATH.exists(), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(str(native.LIB_PATH))
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_mirror_the_header():
    assert set(native.PROTOTYPES) == header_functions()
    lib = native.load()
    assert lib.bolap_abi_version() == 1
    assert isinstance(lib.bolap_last_error(), bytes)


def test_struct_sizes_match_c_layout():
    # Sizes fixed by the C declarations (natural alignment, x86-64 / amdgcn).
    assert C.sizeof(native.Pred) == 80
    assert C.sizeof(native.Q1Group) == 64
    assert C.sizeof(native.GroupRow) == 16
    assert C.sizeof(native.JoinSumArgs) == 8 + 8 + 8 + 80 * 4 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8


STRUCTS = {  # ctypes mirror -> C type name in bolap.h
    "DeviceInfo": "bolap_device_info", "TraceEntry": "bolap_trace_entry", "Pred": "bolap_pred",
    "Q1Args": "bolap_q1_args", "Q1Group": "bolap_q1_group",
    "JoinSumArgs": "bolap_join_sum_args", "JoinSumResult": "bolap_join_sum_result",
    "ChainArgs": "bolap_chain_args", "GroupRow": "bolap_group_row",
    "ChainStats": "bolap_chain_stats", "GenLineitemCols": "bolap_gen_lineitem_cols",
    "StarDim": "bolap_star_dim", "StarArgs": "bolap_star_args", "StarStats": "bolap_star_stats",
    "CebDb": "bolap_ceb_db", "CebQuery": "bolap_ceb_query", "CebResult": "bolap_ceb_result",
    "CebCol": "bolap_ceb_col", "CebQ10aArgs": "bolap_ceb_q10a_args",
    "EstColumn": "bolap_est_column",
    "Zone": "bolap_zone", "Q6Args": "bolap_q6_args", "Q6Result": "bolap_q6_result",
    "Q6Params": "bolap_q6_params", "Q18Args": "bolap_q18_args", "Q18Row": "bolap_q18_row",
    "Q18Stats": "bolap_q18_stats", "Q1Row": "bolap_q1_row", "Q14Args": "bolap_q14_args",
    "Q14Result": "bolap_q14_result", "Q17Args": "bolap_q17_args",
    "Q17Result": "bolap_q17_result", "OrderSide": "bolap_order_side", "Q4Args": "bolap_q4_args",
    "Q5Args": "bolap_q5_args", "Q10Args": "bolap_q10_args", "Q10Row": "bolap_q10_row",
    "Q12Args": "bolap_q12_args", "Q12Result": "bolap_q12_result", "Q19Group": "bolap_q19_group",
    "Q19Args": "bolap_q19_args", "Q21Args": "bolap_q21_args", "Q22Args": "bolap_q22_args",
    "Q22Result": "bolap_q22_result", "Q7Args": "bolap_q7_args", "Q8Args": "bolap_q8_args",
    "Q9Args": "bolap_q9_args", "Q13Args": "bolap_q13_args", "Q16Args": "bolap_q16_args",
    "Q2Args": "bolap_q2_args", "Q11Args": "bolap_q11_args", "Q15Args": "bolap_q15_args",
    "Q20Args": "bolap_q20_args", "Q20Stats": "bolap_q20_stats",
    "MinmaxArgs": "bolap_minmax_args", "MinmaxRow": "bolap_minmax_row",
    "MinmaxStats": "bolap_minmax_stats", "CebPred": "bolap_ceb_pred", "CebTerm": "bolap_ceb_term",
    "CebCountArgs": "bolap_ceb_count_args", "CebCountResult": "bolap_ceb_count_result",
}


def test_every_struct_matches_the_compiled_layout(tmp_path):
    """Compiles the header with gcc and compares sizeof/offsetof of every
    field with the ctypes mirror."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bolap.h"', 'int main(void){']
    expect = []
    for py, cname in STRUCTS.items():
        cls = getattr(native, py)
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(cls))
        for fname, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append(getattr(cls, fname).offset)
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == expect


def test_loader_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setattr(native, "LIB_PATH", tmp_path / "libbolap.so")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        native.load()

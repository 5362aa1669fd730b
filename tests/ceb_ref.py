"""The reference's own CEB pipelines as a checker (oracle/_ref/libref_ceb.so).

`RefCeb(tables)` fills the reference's Database (bespoke_ceb/builder_impl.hpp:
14-123) from a `tests.ceb_data.tables()` draw through the C shim
oracle/ref_ceb.cpp, which builds it as the reference's build() does (int32
columns with is_null bytes and non-null min / max, strings interned into the
StringPool, the builder's table orders).  `run(lines)` answers request lines
with the reference's parse_qN + run_qN and returns the CSV texts the
reference's query() writes for them.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import tempfile

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "_ref" / "libref_ceb.so"


def available() -> bool:
    return LIB.exists()


def _load():
    lib = C.CDLL(str(LIB))
    vp, cp, i64 = C.c_void_p, C.c_char_p, C.c_int64
    lib.ref_ceb_db_new.restype = vp
    lib.ref_ceb_db_free.argtypes = [vp]
    lib.ref_ceb_set_i32.argtypes = [vp, cp, cp, vp, vp, i64]
    lib.ref_ceb_set_str.argtypes = [vp, cp, cp, cp, vp, vp, i64]
    lib.ref_ceb_finish.argtypes = [vp]
    lib.ref_ceb_run.argtypes = [vp, cp, cp, cp, C.c_uint64]
    lib.ref_ceb_count_timed.argtypes = [vp, cp, C.POINTER(C.c_int64), C.POINTER(C.c_double), cp,
                                        C.c_uint64]
    return lib


class RefCeb:
    def __init__(self, tables: dict):
        self.lib = _load()
        self.db = self.lib.ref_ceb_db_new()
        for tname, cols in tables.items():
            for cname, v in cols.items():
                if isinstance(v, tuple):
                    vals = np.ascontiguousarray(v[0], np.int32)
                    nul = None if v[1] is None else np.ascontiguousarray(v[1], np.uint8)
                    self.lib.ref_ceb_set_i32(self.db, tname.encode(), cname.encode(),
                                             vals.ctypes.data,
                                             None if nul is None else nul.ctypes.data, len(vals))
                elif isinstance(v, dict):   # {"chars": uint8, "offsets": int64, "null"}
                    off = np.ascontiguousarray(v["offsets"], np.int64)
                    nul = v.get("null")
                    nul = None if nul is None else np.ascontiguousarray(nul, np.uint8)
                    chars = np.ascontiguousarray(v["chars"], np.uint8)
                    self.lib.ref_ceb_set_str(self.db, tname.encode(), cname.encode(),
                                             C.c_char_p(chars.ctypes.data), off.ctypes.data,
                                             None if nul is None else nul.ctypes.data,
                                             len(off) - 1)
                else:
                    enc = [b"" if s is None else s.encode() for s in v]
                    off = np.zeros(len(enc) + 1, np.int64)
                    np.cumsum([len(e) for e in enc], out=off[1:])
                    nul = np.array([s is None for s in v], np.uint8)
                    self.lib.ref_ceb_set_str(self.db, tname.encode(), cname.encode(),
                                             b"".join(enc), off.ctypes.data, nul.ctypes.data,
                                             len(enc))
        self.lib.ref_ceb_finish(self.db)

    def close(self):
        if self.db:
            self.lib.ref_ceb_db_free(self.db)
            self.db = None

    __del__ = close

    def count_timed(self, line: str) -> tuple[int, float]:
        """(COUNT, ms of the reference's run_qN alone) of one COUNT request."""
        n, ms = C.c_int64(), C.c_double()
        err = C.create_string_buffer(4096)
        rc = self.lib.ref_ceb_count_timed(self.db, line.encode(), C.byref(n), C.byref(ms), err,
                                          len(err))
        if rc != 0:
            raise RuntimeError(f"reference could not time {line!r}: {err.value.decode()}")
        return n.value, ms.value

    def run(self, lines: list[str]) -> list[str | None]:
        """CSV text per request line; None for an id the reference does not serve."""
        out = []
        err = C.create_string_buffer(4096)
        with tempfile.TemporaryDirectory() as d:
            for i, line in enumerate(lines):
                path = pathlib.Path(d) / f"result{i + 1}.csv"
                rc = self.lib.ref_ceb_run(self.db, line.encode(), str(path).encode(), err,
                                          len(err))
                if rc == 1:
                    raise RuntimeError(f"reference threw on {line!r}: {err.value.decode()}")
                out.append(path.read_text() if rc == 0 else None)
        return out

"""Build recipe of libbolap (HIP, gfx950) and of the CPU oracle.

libbolap.so is compiled in-tree with hipcc for gfx950 only; the oracle is
plain C compiled with gcc.  Both outputs sit next to their sources so they
travel with the repository snapshot.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libbolap.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 instruction selection is disabled for every device compile.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: pathlib.Path) -> pathlib.Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_lib(jobs: int | None = None) -> pathlib.Path:
    """Compiles every csrc/*.hip and links libbolap.so (always recompiles)."""
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


DRIVER_SRC = PKG / "driver" / "tpch_query.cpp"
DRIVER = PKG / "bolap_tpch"
CEB_DRIVER_SRC = PKG / "driver" / "ceb_query.cpp"
CEB_DRIVER = PKG / "bolap_ceb"


def _arrow_dir() -> pathlib.Path | None:
    """pyarrow's bundled Arrow C++ (headers and libarrow/libparquet), if present."""
    try:
        import pyarrow
    except ImportError:
        return None
    d = pathlib.Path(pyarrow.__file__).resolve().parent
    if (d / "include" / "parquet" / "arrow" / "reader.h").exists() and \
            list(d.glob("libparquet.so.*")):
        return d
    return None


def build_driver() -> pathlib.Path | None:
    """The TPC-H and CEB load/build/query drivers (host C++ over libbolap and
    Arrow); skipped when Arrow C++ is not installed."""
    arrow = _arrow_dir()
    if arrow is None:
        print("bolap_tpch / bolap_ceb skipped: no Arrow C++ (pyarrow) in this environment")
        return None
    _build_driver(arrow, DRIVER_SRC, DRIVER)
    _build_driver(arrow, CEB_DRIVER_SRC, CEB_DRIVER)
    return DRIVER


def _build_driver(arrow: pathlib.Path, src: pathlib.Path, out: pathlib.Path) -> pathlib.Path:
    libs = sorted(arrow.glob("libarrow.so.*")) + sorted(arrow.glob("libparquet.so.*"))
    arrow_so = next(x.name for x in libs if x.name.startswith("libarrow.so.") and
                    x.name.count(".") == 2)
    parquet_so = next(x.name for x in libs if x.name.startswith("libparquet.so.") and
                      x.name.count(".") == 2)
    tmp = out.with_suffix(".tmp")
    cmd = [HIPCC, "-x", "c++", "-std=c++20", "-O2", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", f"-I{ROOT / 'include'}", f"-I{arrow / 'include'}",
           str(src), "-o", str(tmp), f"-L{PKG}", "-lbolap", "-Wl,-rpath,$ORIGIN",
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", f"-L{arrow}",
           f"-l:{arrow_so}", f"-l:{parquet_so}", f"-Wl,-rpath,{arrow}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{out.name} build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


def oracle_sources() -> list[pathlib.Path]:
    return sorted(ORACLE_DIR.glob("*.c"))


def build_oracle(out: pathlib.Path | None = None, march: str = "x86-64-v2") -> pathlib.Path:
    """Compiles the C oracle (test infrastructure) into oracle/liboracle.so, or
    into `out` (the CPU baseline builds its own copy with -march=native on
    the host that times it)."""
    out = out or ORACLE_LIB
    cc = shutil.which("gcc") or "cc"
    tmp = out.with_suffix(".so.tmp")
    cmd = [cc, "-O3", f"-march={march}", "-std=c11", "-fPIC", "-shared",
           *map(str, oracle_sources()), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


REFERENCE = pathlib.Path(os.environ.get("BOLAP_REFERENCE", "/root/reference"))
REF_DIR = ORACLE_DIR / "_ref"
REF_TPCH = REF_DIR / "libref_tpch.so"
DROPIN_TPCH = REF_DIR / "libdropin_tpch.so"
DROPIN_SRC = PKG / "dropin" / "tpch_run_q.cpp"


REF_CEB = REF_DIR / "libref_ceb.so"


def _ref_compile(out: pathlib.Path, extra: list[str], suite: str = "tpch") -> pathlib.Path:
    src = REFERENCE / f"bespoke_{suite}"
    tmp = out.with_suffix(".so.tmp")
    cmd = [shutil.which("g++") or "c++", "-std=c++20", "-O2", "-fPIC", "-shared", "-pthread",
           "-w", "-Wl,-Bsymbolic", f"-I{src}", f"-I{_arrow_dir() / 'include'}",
           str(ORACLE_DIR / f"ref_{suite}.cpp"), "-o", str(tmp), *extra]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{out.name} build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


def build_ref() -> list[pathlib.Path]:
    """oracle/_ref/: the reference's own TPC-H run_qN / write_qN_csv
    (bespoke_tpch/query_q*.cpp, compiled unchanged from where they lie) behind
    the C shim oracle/ref_tpch.cpp -- a checker for tests only --, its CEB
    run_qN (bespoke_ceb/query_q*.cpp) behind oracle/ref_ceb.cpp, and the same
    unit with the drop-in GPU bodies (dropin/tpch_run_q.cpp) in place of six
    pipelines, linked against libbolap.  Skipped where the reference sources or
    pyarrow's Arrow headers (which loader_impl.hpp includes; no Arrow library
    is linked) are absent."""
    if not (REFERENCE / "bespoke_tpch" / "query_q1.cpp").exists() or _arrow_dir() is None:
        print("oracle/_ref skipped: reference sources or Arrow headers absent")
        return []
    REF_DIR.mkdir(exist_ok=True)
    out = [_ref_compile(REF_TPCH, [])]
    if (REFERENCE / "bespoke_ceb" / "query_q1a.cpp").exists():
        out.append(_ref_compile(REF_CEB, [], suite="ceb"))
    if LIB.exists():
        out.append(_ref_compile(DROPIN_TPCH, [
            "-DREF_TPCH_DROPIN", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
            f"-I{ROOT / 'include'}", f"-I{DROPIN_SRC.parent}", f"-L{PKG}", "-lbolap",
            "-Wl,-rpath,$ORIGIN/../../bespokeolap_artifacts_amd", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
            "-Wl,--no-undefined"]))
    return out


def build_all() -> None:
    build_oracle()
    build_lib()
    build_driver()
    build_ref()


if __name__ == "__main__":
    build_all()
    print(LIB)
    print(ORACLE_LIB)

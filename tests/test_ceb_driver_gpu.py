"""bolap_ceb (CEB load / build / query over libbolap) against the reference.

Seeded IMDB-shaped Parquet tables with nulls go through the driver: Arrow
load, string pool and nullable int32 columns, device mirror, the ceb_count
kernels.  Every 1a / 2a answer must equal, byte for byte, the file the
reference's own run_q1a / run_q2a (oracle/_ref/libref_ceb.so, compiled from
its sources) writes for the same tables; ids without a GPU pipeline get the
reference's stub file.
"""
from __future__ import annotations

import subprocess

import pytest

from bespokeolap_artifacts_amd import build as B
from tests import ceb_data as CD
from tests import ceb_ref as CR
from tests.test_ref_ceb import QUERIES, line

pytestmark = pytest.mark.gpu


def requests():
    out = []
    for q in QUERIES:
        out += [line("1a", q), line("2a", q)]
    odd = dict(QUERIES[0], ID1=["3", "<<NULL>>", "abc", "8"], INFO1=["Drama", "NULL", "Nope"],
               KIND=["movie", "tv series", "video game"], GENDER=["f", "null"], YEAR2=1900)
    out += [line("1a", odd), line("2a", odd)]
    empty = dict(QUERIES[0], ROLE=["nobody"])
    out += [line("1a", empty), line("2a", empty)]
    late = dict(QUERIES[1], YEAR1=1950, YEAR2=2000)
    out += [line("2a", late), "1a ('3') ('8') ('Drama') ('USA') () ('actor') ('m') \"2019\" \"1950\""]
    return out


@pytest.mark.parametrize("movies", [40, 60_000])
def test_ceb_driver_equals_the_reference(tmp_path, movies):
    if not B.CEB_DRIVER.exists():
        pytest.skip("bolap_ceb not built (no Arrow C++)")
    if not CR.available():
        pytest.skip("oracle/_ref not built")
    import pyarrow.parquet as pq
    print(f"ceb driver vs reference: movies={movies}", flush=True)
    t = CD.tables(11 + movies, movies)
    data = tmp_path / "data"
    data.mkdir()
    for name, tab in CD.parquet(t).items():
        pq.write_table(tab, data / f"{name}.parquet")
    lines = requests() + ["10a ('3') ('Drama') ('%a%') ('movie') ('actor')"]
    run = tmp_path / "run"
    run.mkdir()
    res = subprocess.run([str(B.CEB_DRIVER), str(data) + "/"], input="\n".join(lines) + "\n\n",
                         capture_output=True, text=True, cwd=run, timeout=300)
    assert res.returncode == 0, res.stderr
    assert [ln.split(" | ")[0] for ln in res.stdout.splitlines()] == \
        [str(i + 1) for i in range(len(lines))]
    want = CR.RefCeb(t).run(lines[:-1])
    for i, w in enumerate(want):
        assert (run / f"result{i + 1}.csv").read_text() == w, lines[i]
    assert (run / f"result{len(lines)}.csv").read_text() == "result\n0\n"
    assert "query 10a is not served" in res.stderr
    if movies > 40:
        assert sum(w != "count_star()\n0\n" for w in want) >= 6


def test_ceb_driver_rejects_an_unknown_id(tmp_path):
    if not B.CEB_DRIVER.exists():
        pytest.skip("bolap_ceb not built (no Arrow C++)")
    res = subprocess.run([str(B.CEB_DRIVER), str(tmp_path) + "/"], input="99z x\n\n",
                         capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert res.returncode == 1 and "Unsupported query id: 99z" in res.stderr

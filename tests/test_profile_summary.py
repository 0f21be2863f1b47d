"""The rocprofv3 summary helpers behind the profiles/ notes, on small
synthetic result databases with the tables and columns they read."""
import sqlite3

import pytest

from vship_amd import profile_summary as P


def _kernels_db(path, rows):
    c = sqlite3.connect(path)
    c.execute("create table kernels (start integer, end integer, name text)")
    c.executemany("insert into kernels values (?, ?, ?)", rows)
    c.commit()
    c.close()


def test_busy_union_and_concurrency(tmp_path):
    db = str(tmp_path / "k.db")
    # window 1: [0, 100) and [50, 150) overlap, [200, 300) after a gap of 50
    # -> span 300, covered 250, summed durations 300
    # window 2: [1000, 1100) alone -> covered = span; a copy kernel is filtered out
    rows = [(0, 100, "void vship::(anonymous namespace)::blur_strip_kernel<float>(...)"),
            (50, 150, "frontend_single_kernel<float, 1>"),
            (120, 130, "__amd_rocclr_copyBuffer"),
            (200, 300, "blur_strip_kernel<float>"),
            (1000, 1100, "frontend_single_kernel<float, 1>"),
            (1050, 1100, "blur_strip_kernel<float>"),
            (1200, 1300, "blur_strip_kernel<float>")]
    _kernels_db(db, rows)
    out = P.busy(db, 3, ["frontend_single", "blur_strip"])
    assert len(out) == 2
    k, span, frac, conc = out[0]
    assert k == 0 and span == pytest.approx(300 / 1e6)
    assert frac == pytest.approx(250 / 300) and conc == pytest.approx(300 / 300)
    k, span, frac, conc = out[1]
    assert k == 3 and span == pytest.approx(300 / 1e6)
    assert frac == pytest.approx(200 / 300) and conc == pytest.approx(250 / 300)
    # no filter: the copy kernel counts, windows of all 7 dispatches
    (k, span, frac, conc), = P.busy(db, 7, [])
    assert span == pytest.approx(1300 / 1e6) and frac == pytest.approx(450 / 1300)


def test_pmc_mean_per_dispatch(tmp_path):
    db = str(tmp_path / "p.db")
    c = sqlite3.connect(db)
    c.execute("create table counters_collection (kernel_name text, dispatch_id integer, counter_name text, "
              "value real)")
    # two dispatches of one kernel, counter summed over its instances (e.g. per XCD), then averaged
    c.executemany("insert into counters_collection values (?, ?, ?, ?)",
                  [("void vship::(anonymous namespace)::k<float>(int)", 1, "FETCH_SIZE", 10.0),
                   ("void vship::(anonymous namespace)::k<float>(int)", 1, "FETCH_SIZE", 5.0),
                   ("void vship::(anonymous namespace)::k<float>(int)", 2, "FETCH_SIZE", 25.0),
                   ("void vship::(anonymous namespace)::k<float>(int)", 2, "WRITE_SIZE", 99.0)])
    c.commit()
    c.close()
    assert P.pmc("FETCH_SIZE", db) == {"k<float>": pytest.approx(20.0)}

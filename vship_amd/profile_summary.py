"""Summaries of rocprofv3 result databases for the profiles/ notes.

python -m vship_amd.profile_summary trace RESULTS_DB
    per-kernel calls, total and average duration (the --stats table)
python -m vship_amd.profile_summary pmc COUNTER KERNEL_DB COPY_DB COPY_MIB
    mean KiB per dispatch of COUNTER for every kernel, and the same counter
    scaled by what it reads for a plain copy kernel of COPY_MIB MiB
"""
from __future__ import annotations

import sqlite3
import sys


def _short(name: str) -> str:
    return name.replace("void ", "").replace("vship::(anonymous namespace)::", "").split("(")[0]


def trace(db: str) -> list[tuple[str, int, float, float, float]]:
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    return [(_short(n), int(k), t, a, p) for n, k, t, a, p in rows]  # durations in µs


def pmc(counter: str, db: str) -> dict[str, float]:
    """Mean over dispatches of the counter summed over its instances (KiB for FETCH/WRITE_SIZE)."""
    c = sqlite3.connect(db)
    q = ("select kernel_name, dispatch_id, sum(value) from counters_collection where counter_name = ? "
         "group by dispatch_id")
    acc: dict[str, list[float]] = {}
    for name, _d, v in c.execute(q, (counter,)):
        acc.setdefault(_short(name), []).append(v)
    return {k: sum(v) / len(v) for k, v in acc.items()}


def main(argv=None) -> int:
    a = sys.argv[1:] if argv is None else argv
    if a and a[0] == "trace":
        print("| kernel | calls | total µs | average µs | % |")
        print("|---|---|---|---|---|")
        for n, k, t, av, p in trace(a[1]):
            print(f"| `{n}` | {k} | {t:.1f} | {av:.3f} | {p:.1f} |")
        return 0
    if a and a[0] == "pmc":
        counter, db, copy_db, mib = a[1], a[2], a[3], float(a[4])
        copy = pmc(counter, copy_db).get("copy_kernel")
        expect_kib = mib * 1024.0  # the copy reads and writes mib MiB once
        scale = expect_kib / copy if copy else float("nan")
        print(f"{counter}: copy kernel of {mib:g} MiB reads {copy:.1f} KiB per dispatch -> scale {scale:.4f}")
        for k, v in sorted(pmc(counter, db).items()):
            print(f"| `{k}` | {v:.1f} | {v * scale * 1024:.0f} |")
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main())

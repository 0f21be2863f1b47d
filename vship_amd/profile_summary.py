"""Summaries of rocprofv3 result databases for the profiles/ notes.

python -m vship_amd.profile_summary trace RESULTS_DB
    per-kernel calls, total and average duration (the --stats table)
python -m vship_amd.profile_summary pmc COUNTER KERNEL_DB COPY_DB COPY_MIB
    mean KiB per dispatch of COUNTER for every kernel, and the same counter
    scaled by what it reads for a plain copy kernel of COPY_MIB MiB
python -m vship_amd.profile_summary valu KERNEL_DB SIMDS XCDS
    VALU busy fraction per kernel from a run that collected SQ_ACTIVE_INST_VALU,
    SQ_INSTS_VALU and GRBM_GUI_ACTIVE: SQ_ACTIVE_INST_VALU counts quad-cycles
    (a wave64 VALU instruction holds a 16-lane SIMD for 4 cycles) summed over
    the chip, GRBM_GUI_ACTIVE counts the dispatch's cycles once per XCD, so
    busy = 4 * SQ_ACTIVE_INST_VALU / (SIMDS * GRBM_GUI_ACTIVE / XCDS)
python -m vship_amd.profile_summary counters OUT_JSON WORKLOAD PAIRS_PER_LAUNCH DB [DB...]
    every SQ_/GRBM_ counter of the runs DB... (one counter group per run,
    same workload) per kernel, mean per dispatch, with the derived issue
    fractions (SIMDS = 1024, XCDS = 8): VALU busy as above, MFMA busy =
    SQ_VALU_MFMA_BUSY_CYCLES / (SIMDS * GRBM_GUI_ACTIVE / XCDS), and VALU
    lane-operations per pair = SQ_INSTS_VALU * 64 / PAIRS_PER_LAUNCH; merged
    into OUT_JSON under WORKLOAD, and printed as a Markdown table
python -m vship_amd.profile_summary busy RESULTS_DB WINDOW [NAME...]
    GPU occupancy of a kernel trace, per window of WINDOW consecutive
    dispatches of the kernels whose names contain one of NAME (all kernels
    when none is given): the window's span, the fraction of it during which
    at least one of them runs, and the mean number running at once
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


SIMDS, XCDS = 1024, 8


def counters(dbs: list[str]) -> dict[str, dict[str, float]]:
    """{kernel: {counter: mean per dispatch}} over the runs (one group each)."""
    out: dict[str, dict[str, float]] = {}
    for db in dbs:
        c = sqlite3.connect(db)
        q = ("select kernel_name, counter_name, dispatch_id, sum(value) from counters_collection "
             "group by dispatch_id, counter_name")
        acc: dict[tuple[str, str], list[float]] = {}
        for name, counter, _d, v in c.execute(q):
            acc.setdefault((_short(name), counter), []).append(v)
        for (k, counter), v in acc.items():
            out.setdefault(k, {})[counter] = sum(v) / len(v)
    return out


def derived(cnt: dict[str, float], pairs_per_launch: float) -> dict[str, float]:
    d = dict(cnt)
    grbm = cnt.get("GRBM_GUI_ACTIVE")
    if grbm:
        cyc = grbm / XCDS
        d["cycles"] = cyc
        if "SQ_ACTIVE_INST_VALU" in cnt:
            d["valu_busy"] = 4.0 * cnt["SQ_ACTIVE_INST_VALU"] / (SIMDS * cyc)
        if "SQ_VALU_MFMA_BUSY_CYCLES" in cnt:
            d["mfma_busy"] = cnt["SQ_VALU_MFMA_BUSY_CYCLES"] / (SIMDS * cyc)
    if "SQ_INSTS_VALU" in cnt:
        d["valu_insts_per_pair"] = cnt["SQ_INSTS_VALU"] / pairs_per_launch
        d["valu_lane_ops_per_pair"] = cnt["SQ_INSTS_VALU"] * 64.0 / pairs_per_launch
    if "SQ_INSTS_MFMA" in cnt:
        d["mfma_insts_per_pair"] = cnt["SQ_INSTS_MFMA"] / pairs_per_launch
    return d


def busy(db: str, window: int, names: list[str]) -> list[tuple[int, float, float, float]]:
    """(first dispatch, span ms, busy fraction, mean kernels in flight) per window."""
    c = sqlite3.connect(db)
    rows = [(s, e) for s, e, n in c.execute("select start, end, name from kernels order by start")
            if not names or any(x in n for x in names)]
    out = []
    for k in range(0, len(rows) - window + 1, window):
        iv = sorted(rows[k:k + window])
        t0, t1 = iv[0][0], max(e for _, e in iv)
        covered, cs, ce = 0, iv[0][0], iv[0][1]
        for s, e in iv[1:]:
            if s > ce:
                covered += ce - cs
                cs, ce = s, e
            else:
                ce = max(ce, e)
        covered += ce - cs
        out.append((k, (t1 - t0) / 1e6, covered / (t1 - t0), sum(e - s for s, e in iv) / (t1 - t0)))
    return out


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
    if a and a[0] == "valu":
        db, simds, xcds = a[1], float(a[2]), float(a[3])
        act, ins, grbm = pmc("SQ_ACTIVE_INST_VALU", db), pmc("SQ_INSTS_VALU", db), pmc("GRBM_GUI_ACTIVE", db)
        print("| kernel | VALU instructions | SQ_ACTIVE_INST_VALU | cycles | VALU busy |")
        print("|---|---|---|---|---|")
        for k in sorted(act):
            if k not in grbm or k not in ins or grbm[k] <= 0:
                continue
            cyc = grbm[k] / xcds
            print(f"| `{k}` | {ins[k]:.0f} | {act[k]:.0f} | {cyc:.0f} | {4.0 * act[k] / (simds * cyc):.3f} |")
        return 0
    if a and a[0] == "counters":
        import json
        import os
        out_json, workload, ppl, dbs = a[1], a[2], float(a[3]), a[4:]
        table = {k: derived(v, ppl) for k, v in sorted(counters(dbs).items()) if "copyBuffer" not in k}
        doc = json.load(open(out_json)) if os.path.exists(out_json) else {}
        doc[workload] = {"pairs_per_launch": ppl, "kernels": table}
        with open(out_json, "w") as f:
            json.dump(doc, f, indent=1, sort_keys=True)
        cols = ["SQ_INSTS_VALU", "SQ_INSTS_MFMA", "valu_lane_ops_per_pair", "valu_busy", "mfma_busy", "cycles"]
        print("| kernel | " + " | ".join(cols) + " |")
        print("|---|" + "---|" * len(cols))
        for k, d in table.items():
            cells = [(f"{d[c]:.3f}" if c.endswith("busy") else f"{d[c]:,.0f}") if c in d else "" for c in cols]
            print(f"| `{k}` | " + " | ".join(cells) + " |")
        return 0
    if a and a[0] == "busy":
        print("| first dispatch | span ms | GPU busy | kernels in flight (mean) |")
        print("|---|---|---|---|")
        for k, span, frac, conc in busy(a[1], int(a[2]), a[3:]):
            print(f"| {k} | {span:.3f} | {frac:.3f} | {conc:.2f} |")
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main())

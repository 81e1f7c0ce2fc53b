"""Summaries of rocprofv3 output for the library's kernels.

kernel_stats_markdown(): the `--kernel-trace --stats` kernel table as text.
traffic_per_launch(): HBM bytes per launch of one GEMM class from two
`--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs, each scaled by what the same
counter reads for a device copy of known size in the same program.

Usage (after the rocprofv3 runs of `bench.py --profile-step`):
  python -m dcvc_amd.profiling stats <kernel_stats.csv> <title> <command...>
  python -m dcvc_amd.profiling traffic <fetch.csv> <write.csv> <class> <copy_bytes>
"""
from __future__ import annotations

import csv
import json
import re
import sys
from collections import defaultdict

EPI_NAMES = {0: "bias", 1: "wsilu", 2: "shortcut", 3: "shortcut2", 4: "shortcut_quant", 5: "quant", 6: "quad"}
# Tile<BM, BN, WM, WN, STAGES, EPI, AMODE, DMODE[, PREC]>
_TILE = re.compile(r"Tile<(\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+)(?:, (\d+))?>")


# FfnCfg<C, CM, EPI>: the fused DepthConvBlock FFN
_FFN = re.compile(r"FfnCfg<(\d+), (\d+), (\d+)>")


def gemm_class(kernel_name: str) -> str | None:
    """Kernel class of a GEMM instantiation, named as in the launch trace."""
    f = _FFN.search(kernel_name)
    if f:
        return f"ffn_{EPI_NAMES[int(f.group(3))]}"
    m = _TILE.search(kernel_name)
    if not m:
        return None
    epi, amode, dmode = int(m.group(6)), int(m.group(7)), int(m.group(8))
    kind = "conv_s2" if amode in (1, 2) else ("subpel" if dmode == 1 else "conv1x1")
    return f"{kind}_{EPI_NAMES[epi]}"


def short_name(kernel_name: str) -> str:
    m = _TILE.search(kernel_name)
    if m:
        return f"gemm_kernel<{','.join(x for x in m.groups() if x is not None)}> [{gemm_class(kernel_name)}]"
    f = _FFN.search(kernel_name)
    if f:
        return f"ffn_kernel<{','.join(f.groups())}> [{gemm_class(kernel_name)}]"
    return kernel_name.split("(")[0][:90]


def kernel_stats_markdown(stats_csv: str, title: str, command: str, top: int = 25) -> str:
    rows = list(csv.DictReader(open(stats_csv)))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", "",
             f"Total kernel time {total / 1e6:.3f} ms over {sum(int(r['Calls']) for r in rows)} dispatches.", "",
             "| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        lines.append(f"| {short_name(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(lines) + "\n"


def _per_dispatch(pmc_csv: str, counter: str):
    vals = defaultdict(float)
    names = {}
    for r in csv.DictReader(open(pmc_csv)):
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        vals[d] += float(r["Counter_Value"])
        names[d] = r["Kernel_Name"]
    return vals, names


def traffic_per_launch(fetch_csv: str, write_csv: str, kernel_class: str, copy_bytes: int) -> dict:
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) averaged over the dispatches of
    `kernel_class`, scaled so the calibration copy (the program's last
    dispatch, a device copy of copy_bytes) reads copy_bytes and writes
    copy_bytes."""
    out = {"kernel_class": kernel_class}
    for counter, path, key in (("FETCH_SIZE", fetch_csv, "fetch"), ("WRITE_SIZE", write_csv, "write")):
        vals, names = _per_dispatch(path, counter)
        last = max(vals)
        ref = vals[last] * 1024.0
        factor = copy_bytes / ref
        mine = [v for d, v in vals.items() if gemm_class(names[d]) == kernel_class]
        out[f"{key}_kib_raw_avg"] = sum(mine) / len(mine) if mine else None
        out[f"{key}_scale"] = factor
        out[f"{key}_bytes_per_launch"] = (sum(mine) / len(mine)) * 1024.0 * factor if mine else None
        out["dispatches"] = len(mine)
        out["calibration_kernel"] = names[last]
    if out["fetch_bytes_per_launch"] is not None and out["write_bytes_per_launch"] is not None:
        out["bytes_per_launch"] = int(out["fetch_bytes_per_launch"] + out["write_bytes_per_launch"])
    return out


def main(argv):
    if argv[0] == "stats":
        print(kernel_stats_markdown(argv[1], argv[2], " ".join(argv[3:])), end="")
    elif argv[0] == "traffic":
        print(json.dumps(traffic_per_launch(argv[1], argv[2], argv[3], int(argv[4])), indent=1))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])

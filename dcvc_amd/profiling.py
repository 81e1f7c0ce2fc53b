"""Summaries of rocprofv3 output for the library's kernels.

`bench.py --profile-step` brackets its timed steps with two launches of a
marker kernel (torch's spin kernel); every summary here keeps only the
dispatches between the last two markers, so warm-up steps, intra frames and
weight conversions stay out.

window_stats_markdown(): per-kernel table of the window from a
`--kernel-trace` CSV.
traffic_per_launch(): HBM bytes per launch of one kernel from two
`--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs, each scaled by what the same
counter reads for a device copy of known size in the same program (its last
dispatch).
counters_per_launch(): any other counters (e.g. SQ_VALU_MFMA_BUSY_CYCLES,
SQ_BUSY_CYCLES, GRBM_GUI_ACTIVE) averaged over the window's dispatches of
one kernel.

Usage (after the rocprofv3 runs of `bench.py --profile-step`):
  python -m dcvc_amd.profiling stats <kernel_trace.csv> <title> <command...>
  python -m dcvc_amd.profiling traffic <fetch.csv> <write.csv> <kernel> <copy_bytes>
  python -m dcvc_amd.profiling counters <pmc.csv> <kernel> <counter,...>
  python -m dcvc_amd.profiling launch <epi> <M> <K> <N> <times>
  python -m dcvc_amd.profiling launch_dw <H> <W> <C> <times>
  python -m dcvc_amd.profiling launch_hts_batch <n> <H> <W> <chunks>
A <kernel> is a short name as printed by `stats` (e.g.
gemm_kernel<128,256,2,2,3,6,0,0,0>) or a kernel class (e.g. conv1x1_quad).
`launch` runs one 1x1-conv GEMM shape (N weight rows; epi bias, wsilu,
shortcut or quad) `times` times on N(0,1) data, the program to put under a
counter pass for one kernel alone.
"""
from __future__ import annotations

import csv
import json
import re
import sys
from collections import defaultdict

EPI_NAMES = {0: "bias", 1: "wsilu", 2: "shortcut", 3: "shortcut2", 4: "shortcut_quant", 5: "quant", 6: "quad"}
# Tile<BM, BN, WM, WN, STAGES, EPI, AMODE, DMODE[, PREC[, KB]]>
_TILE = re.compile(r"Tile<(\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+)(?:, (\d+))?(?:, (\d+))?>")


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
        # the K tile depth is shown only when it is not the default 32
        g = [x for x in m.groups()[:9] if x is not None]
        kb = m.group(10)
        tag = f",k{kb}" if kb is not None and kb != "32" else ""
        return f"gemm_kernel<{','.join(g)}{tag}> [{gemm_class(kernel_name)}]"
    f = _FFN.search(kernel_name)
    if f:
        return f"ffn_kernel<{','.join(f.groups())}> [{gemm_class(kernel_name)}]"
    n = kernel_name.split("(")[0]
    if n.startswith("void "):
        n = n[5:]
    return n.replace("dcvc::", "")[:90]


MARKER = "spin_kernel"


def _window(ids_names):
    """Dispatch ids strictly between the last two marker dispatches."""
    marks = sorted(d for d, n in ids_names if MARKER in n)
    if len(marks) < 2:
        raise ValueError("fewer than two marker dispatches: not a bench.py --profile-step run")
    return marks[-2], marks[-1]


def _matches(name: str, kernel: str) -> bool:
    return short_name(name).split(" ")[0] == kernel or gemm_class(name) == kernel


def window_rows(trace_csv: str):
    rows = list(csv.DictReader(open(trace_csv)))
    lo, hi = _window([(int(r["Dispatch_Id"]), r["Kernel_Name"]) for r in rows])
    return [r for r in rows if lo < int(r["Dispatch_Id"]) < hi]


def window_stats(trace_csv: str) -> dict:
    """short name -> {calls, total_ns, avg_ns, class} over the window."""
    out = {}
    for r in window_rows(trace_csv):
        n = short_name(r["Kernel_Name"]).split(" ")[0]
        e = out.setdefault(n, {"calls": 0, "total_ns": 0, "class": gemm_class(r["Kernel_Name"])})
        e["calls"] += 1
        e["total_ns"] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    for e in out.values():
        e["avg_ns"] = e["total_ns"] / e["calls"]
    return dict(sorted(out.items(), key=lambda kv: -kv[1]["total_ns"]))


def window_stats_markdown(trace_csv: str, title: str, command: str, top: int = 25) -> str:
    st = window_stats(trace_csv)
    rows = window_rows(trace_csv)
    total = sum(e["total_ns"] for e in st.values())
    span = max(int(r["End_Timestamp"]) for r in rows) - min(int(r["Start_Timestamp"]) for r in rows)
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", "",
             f"Window between the two markers: {len(rows)} dispatches, {span / 1e6:.3f} ms wall, "
             f"{total / 1e6:.3f} ms summed kernel time.", "",
             "| kernel | class | calls | total ms | avg us | % |", "|---|---|---:|---:|---:|---:|"]
    for n, e in list(st.items())[:top]:
        lines.append(f"| {n} | {e['class'] or ''} | {e['calls']} | {e['total_ns'] / 1e6:.3f} | "
                     f"{e['avg_ns'] / 1e3:.2f} | {100.0 * e['total_ns'] / total:.2f} |")
    return "\n".join(lines) + "\n"


def _per_dispatch(pmc_csv: str, counter: str):
    vals = defaultdict(float)
    names = {}
    for r in csv.DictReader(open(pmc_csv)):
        d = int(r["Dispatch_Id"])
        names[d] = r["Kernel_Name"]
        if r["Counter_Name"] == counter:
            vals[d] += float(r["Counter_Value"])
    return vals, names


def traffic_per_launch(fetch_csv: str, write_csv: str, kernel: str, copy_bytes: int) -> dict:
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) averaged over the window's
    dispatches of `kernel`, scaled so the calibration copy (the program's
    last dispatch, a device copy of copy_bytes) reads copy_bytes and writes
    copy_bytes."""
    out = {"kernel": kernel}
    for counter, path, key in (("FETCH_SIZE", fetch_csv, "fetch"), ("WRITE_SIZE", write_csv, "write")):
        vals, names = _per_dispatch(path, counter)
        lo, hi = _window(names.items())
        last = max(vals)
        ref = vals[last] * 1024.0
        factor = copy_bytes / ref
        mine = [v for d, v in vals.items() if lo < d < hi and _matches(names[d], kernel)]
        out[f"{key}_kib_raw_avg"] = sum(mine) / len(mine) if mine else None
        out[f"{key}_scale"] = factor
        out[f"{key}_bytes_per_launch"] = (sum(mine) / len(mine)) * 1024.0 * factor if mine else None
        out["dispatches"] = len(mine)
        out["calibration_kernel"] = names[last]
    if out["fetch_bytes_per_launch"] is not None and out["write_bytes_per_launch"] is not None:
        out["bytes_per_launch"] = int(out["fetch_bytes_per_launch"] + out["write_bytes_per_launch"])
    return out


def counters_per_launch(pmc_csv: str, kernel: str, counters: list[str]) -> dict:
    """Each counter averaged over the window's dispatches of `kernel`."""
    out = {"kernel": kernel}
    for c in counters:
        vals, names = _per_dispatch(pmc_csv, c)
        lo, hi = _window(names.items())
        mine = [v for d, v in vals.items() if lo < d < hi and _matches(names[d], kernel)]
        out[c] = sum(mine) / len(mine) if mine else None
        out["dispatches"] = len(mine)
    return out


def launch(epi: str, M: int, K: int, N: int, times: int):
    import numpy as np
    import torch
    from . import kernels
    dev = torch.device("cuda:0")
    r = np.random.Generator(np.random.PCG64(0))
    a = torch.from_numpy(r.standard_normal((M, K)).astype(np.float16)).to(dev)
    w = torch.from_numpy((r.standard_normal((N, K)) / np.sqrt(K)).astype(np.float16)).to(dev)
    b = torch.zeros(N, dtype=torch.float32, device=dev)
    n_out = N // 4 if epi == "quad" else N
    s1 = torch.from_numpy(r.standard_normal((M, n_out)).astype(np.float16)).to(dev) if epi == "shortcut" else None
    out = torch.empty((M, n_out), dtype=torch.float16, device=dev)
    for _ in range(times):
        kernels.conv1x1(a, w, b, epi=epi, s1=s1, out=out)
    torch.cuda.synchronize()


def launch_dw(H: int, W: int, C: int, times: int):
    """The depthwise 3x3 conv on one H x W x C picture, `times` times."""
    import numpy as np
    import torch
    from . import kernels
    dev = torch.device("cuda:0")
    r = np.random.Generator(np.random.PCG64(0))
    x = torch.from_numpy(r.standard_normal((H * W, C)).astype(np.float16)).to(dev)
    wt = torch.from_numpy((r.standard_normal((9, C)) * 0.3).astype(np.float16)).to(dev)
    y = torch.empty_like(x)
    for _ in range(times):
        kernels.d3x3(x, 1, H, W, wt, out=y)
    torch.cuda.synchronize()


def launch_hts_batch(n: int, H: int, W: int, steps: int):
    """n HT-S streams of H x W coded as one batch (DMCHTSBatchProxy pair),
    `steps` chunks between two marker launches after two warm-up chunks: a
    --kernel-trace of it shows one launch per layer for the n streams."""
    import torch
    from . import batching, codec
    sdi, sdp = codec.make_state_dicts(seed=0)
    gs, _ = batching.build_groups(sdi, sdp, [(H, W)] * n, n_chunks=2 + steps, seed=0)
    g = gs.groups[0]
    for c in range(2):
        g.code_chunk(gs.chunks[0][c], gs.qps[0])
    torch.cuda.synchronize()
    torch.cuda._sleep(1000)  # window marker
    for c in range(steps):
        g.code_chunk(gs.chunks[0][2 + c], gs.qps[0])
    torch.cuda._sleep(1000)
    torch.cuda.synchronize()


def main(argv):
    if argv[0] == "stats":
        print(window_stats_markdown(argv[1], argv[2], " ".join(argv[3:])), end="")
    elif argv[0] == "traffic":
        print(json.dumps(traffic_per_launch(argv[1], argv[2], argv[3], int(argv[4])), indent=1))
    elif argv[0] == "launch":
        launch(argv[1], int(argv[2]), int(argv[3]), int(argv[4]), int(argv[5]))
    elif argv[0] == "launch_dw":
        launch_dw(int(argv[1]), int(argv[2]), int(argv[3]), int(argv[4]))
    elif argv[0] == "launch_hts_batch":
        launch_hts_batch(int(argv[1]), int(argv[2]), int(argv[3]), int(argv[4]))
    elif argv[0] == "counters":
        print(json.dumps(counters_per_launch(argv[1], argv[2], argv[3].split(",")), indent=1))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])

"""Roofline reference points and rocprofv3 summaries for the horizon march.

``probe_peaks(session)`` measures, on the running device, the streaming copy
rate and the 2x2-gather rate of a float32 grid resident in L2, in the MALL and
in HBM (the march's memory access pattern; ``lsh_probe_gather``).

Command line:

* ``python -m lunarscout_amd.profiling copy --mib M`` runs one device copy of
  M MiB (the known-size stream that scales rocprofv3's FETCH_SIZE and
  WRITE_SIZE counters into bytes);
* ``python -m lunarscout_amd.profiling workload --tiles K --times T`` runs the
  full-site march on K tiles (segments + one march launch each, degrees) and,
  with T > 0, one lighting launch over them with T Sun + T Earth vectors
  (2-hour steps) plus their PSR reduction -- a short program for counter runs;
* ``python -m lunarscout_amd.profiling counters ...`` turns the ``--pmc`` passes
  of one kernel into VALU busy, FP64 rate, L2 hit rate and scaled HBM bytes;
* ``python -m lunarscout_amd.profiling summarize ...`` turns rocprofv3 CSV
  output (kernel trace stats and the FETCH_SIZE / WRITE_SIZE passes) into the
  short text and JSON summaries kept with the sources.
"""

from __future__ import annotations

import argparse
import csv
import json
import sys
from pathlib import Path

import numpy as np

MARCH_KERNEL = "march_kernel"
COPY_KERNEL = "copy_kernel"


def probe_peaks(session, repeats: int = 3) -> dict:
    """Copy and gather rates (GB/s, algorithmic bytes) on ``session``'s device."""
    from ._lib import DeviceBuffer, Event, check, load_library

    lib = load_library()
    out: dict = {}
    nbytes = 1 << 30
    src = DeviceBuffer.empty(session, (nbytes // 4,), np.float32)
    dst = DeviceBuffer.empty(session, (nbytes // 4,), np.float32)
    src.zero()
    best = float("inf")
    for _ in range(repeats + 1):
        e0 = Event(session).record()
        check(lib.lsh_probe_copy(session.handle, dst.ptr, src.ptr, nbytes), "lsh_probe_copy")
        e1 = Event(session).record()
        best = min(best, e0.elapsed_ms(e1))
    out["copy_gbs"] = round(2 * nbytes / (best / 1000.0) / 1e9, 1)
    dst.free()
    threads, taps = 256 * 4096, 256
    result = DeviceBuffer.empty(session, (threads,), np.float32)
    for name, side in (("l2", 512), ("mall", 4096), ("hbm", 16384)):
        window = src if side * side * 4 <= nbytes else None
        if window is None:
            continue
        best = float("inf")
        for _ in range(repeats + 1):
            e0 = Event(session).record()
            check(lib.lsh_probe_gather(session.handle, window.ptr, side, side, taps, threads,
                                       result.ptr), "lsh_probe_gather")
            e1 = Event(session).record()
            best = min(best, e0.elapsed_ms(e1))
        out[f"gather_{name}_gbs"] = round(16.0 * threads * taps / (best / 1000.0) / 1e9, 1)
        out[f"gather_{name}_window_mib"] = side * side * 4 / 2**20
    result.free()
    src.free()
    return out


def _copy(mib: int) -> int:
    from ._lib import DeviceBuffer, check, load_library
    from .hip_backend import HipSession

    session = HipSession(0)
    nbytes = mib << 20
    src = DeviceBuffer.empty(session, (nbytes // 4,), np.float32)
    dst = DeviceBuffer.empty(session, (nbytes // 4,), np.float32)
    src.zero()
    check(load_library().lsh_probe_copy(session.handle, dst.ptr, src.ptr, nbytes), "lsh_probe_copy")
    session.synchronize()
    print(json.dumps({"copy_bytes": nbytes}), flush=True)
    return 0


def _workload(args) -> int:
    from . import synthetic
    from ._lib import HORIZON_DEGREES, DeviceBuffer
    from .hip_backend import HipSession
    from .lighting import HipLightingSession, reduce_sun_vectors_for_psr

    dems = synthetic.site_dems(synthetic.FULL_SITE)
    primary = dems[0]
    session = HipSession(0)
    grid = [(x, y) for y in range(0, primary.height, 128) for x in range(0, primary.width, 128)]
    tiles = [grid[(i * 97) % len(grid)] for i in range(args.tiles)]
    layout = "azimuth-major" if args.azimuth_major else "pixel-major"
    shape = (len(tiles), 1440, 128 * 128) if args.azimuth_major else (len(tiles), 128 * 128, 1440)
    horizons = DeviceBuffer.empty(session, shape, np.float32)
    segments = None
    for slot, (c, r) in enumerate(tiles):
        segments = session.segments_device(dems, tile_column=c, tile_row=r, out=segments)
        session.march_device(dems, segments, tile_column=c, tile_row=r, flags=HORIZON_DEGREES,
                             output=horizons.slot(slot), azimuth_major=args.azimuth_major)
        print(f"march tile {slot + 1}/{len(tiles)} ({c}, {r}) enqueued", flush=True)
    session.synchronize()
    if args.times > 0:
        lighting = HipLightingSession(session)
        sun = synthetic.sun_vectors(primary, args.times, 2.0, seed=5)
        earth = synthetic.earth_vectors(args.times, 2.0, seed=6)
        psr, _ = reduce_sun_vectors_for_psr(primary, sun)
        bufs = [DeviceBuffer.from_host(session, v) for v in (sun, earth, psr)]
        out = lighting.lighting_tiles(primary, tiles, horizons, sun=bufs[0], earth=bufs[1],
                                      psr_vectors=bufs[2], sun_bytes=True, sun_mean=True,
                                      earth_visible=True, psr=True,
                                      azimuth_major=args.azimuth_major)
        session.synchronize()
        out.free()
        print(f"lighting: {len(tiles)} tiles x ({len(sun)} + {len(earth)} + {len(psr)}) vectors, "
              f"{layout}", flush=True)
    print(json.dumps({"tiles": tiles, "times": args.times, "layout": layout}), flush=True)
    horizons.free()
    session.close()
    return 0


def _rows(path: Path):
    with open(path, newline="") as f:
        yield from csv.DictReader(f)


def _find(directory: Path, suffix: str) -> Path:
    hits = sorted(Path(directory).rglob(f"*{suffix}"))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    return hits[0]


def counter_per_dispatch(directory: Path, kernel: str, counter: str) -> tuple[float, int]:
    """Mean counter value per dispatch of kernels whose name contains ``kernel``."""
    per: dict[str, float] = {}
    for row in _rows(_find(directory, "counter_collection.csv")):
        if kernel in row["Kernel_Name"] and row["Counter_Name"] == counter:
            key = row.get("Dispatch_Id") or row.get("Correlation_Id")
            per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    if not per:
        raise ValueError(f"{counter} of {kernel} not found under {directory}")
    return float(np.mean(list(per.values()))), len(per)


def march_dispatch_ms(directory: Path) -> list[float]:
    """Durations (ms) of the march dispatches of a --kernel-trace run, in dispatch order."""
    rows = [r for r in _rows(_find(directory, "kernel_trace.csv"))
            if f"{MARCH_KERNEL}<" in r["Kernel_Name"] or f"{MARCH_KERNEL}(" in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in rows]


def kernel_stats(directory: Path) -> list[dict]:
    return list(_rows(_find(directory, "kernel_stats.csv")))


def _kernel_counters(directory: Path, kernel: str) -> tuple[dict, int, float]:
    """Mean per-dispatch counter values of kernels whose name contains ``kernel``.

    Returns ({counter: mean}, dispatch count, mean dispatch duration in ms).
    """
    per: dict[str, dict[str, float]] = {}
    spans: dict[str, float] = {}
    for row in _rows(_find(directory, "counter_collection.csv")):
        if kernel not in row["Kernel_Name"]:
            continue
        key = row.get("Dispatch_Id") or row.get("Correlation_Id")
        per.setdefault(key, {})
        per[key][row["Counter_Name"]] = per[key].get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
        spans[key] = (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e6
    if not per:
        raise ValueError(f"no dispatch of {kernel} under {directory}")
    names = set().union(*per.values())
    means = {n: float(np.mean([d[n] for d in per.values() if n in d])) for n in names}
    return means, len(per), float(np.mean(list(spans.values())))


def counters(args) -> int:
    """Counter summary of one kernel: VALU busy, FP64 rate, L2 hits, HBM bytes."""
    merged: dict[str, float] = {}
    dispatches = []
    durations = []
    for directory in args.pmc:
        means, n, ms = _kernel_counters(Path(directory), args.kernel)
        merged.update(means)
        dispatches.append(n)
        durations.append(ms)
    duration_ms = float(np.mean(durations))
    cus, xccs = args.cu_count, args.xcc_count
    out: dict = {"kernel": args.kernel, "workload": args.workload,
                 "dispatches_per_pass": dispatches, "duration_ms": round(duration_ms, 3),
                 "pmc_commands": args.pmc_command}
    gui = merged.get("GRBM_GUI_ACTIVE")
    if gui and "SQ_ACTIVE_INST_VALU" in merged:
        # VALUBusy: VALU quad-cycles summed over the SIMDs of every CU, over one
        # XCC's busy cycles (GRBM_GUI_ACTIVE is summed over the XCCs)
        out["valu_busy"] = round(merged["SQ_ACTIVE_INST_VALU"] / cus / (gui / xccs), 4)
    if "SQ_THREAD_CYCLES_VALU" in merged and "SQ_ACTIVE_INST_VALU" in merged:
        out["valu_lane_utilization"] = round(
            merged["SQ_THREAD_CYCLES_VALU"] / (merged["SQ_ACTIVE_INST_VALU"] * 64.0), 4)
    if "SQ_INSTS_VALU" in merged:
        out["valu_instructions_per_dispatch"] = merged["SQ_INSTS_VALU"]
    fp64 = [merged.get(k) for k in ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64",
                                    "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_TRANS_F64")]
    if all(v is not None for v in fp64):
        flops = (2.0 * fp64[0] + fp64[1] + fp64[2] + fp64[3]) * 64.0
        out["fp64_flop_per_dispatch"] = flops
        out["fp64_tflops"] = round(flops / (duration_ms / 1000.0) / 1e12, 3)
        out["fp64_frac_of_peak"] = round(out["fp64_tflops"] / 78.6, 4)
        if "SQ_INSTS_VALU" in merged:
            out["fp64_share_of_valu_instructions"] = round(sum(fp64) / merged["SQ_INSTS_VALU"], 4)
    fp32 = [merged.get(k) for k in ("SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_MUL_F32",
                                    "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_TRANS_F32")]
    if all(v is not None for v in fp32):
        out["fp32_flop_per_dispatch"] = (2.0 * fp32[0] + fp32[1] + fp32[2] + fp32[3]) * 64.0
    if "TCC_HIT_sum" in merged and "TCC_MISS_sum" in merged:
        total = merged["TCC_HIT_sum"] + merged["TCC_MISS_sum"]
        out["l2_hit_rate"] = round(merged["TCC_HIT_sum"] / total, 4) if total else None
        out["l2_requests_per_dispatch"] = total
    if args.fetch and args.write:
        copy_bytes = args.copy_mib * 2**20
        copy_fetch_kib, _ = counter_per_dispatch(Path(args.copy_fetch), COPY_KERNEL, "FETCH_SIZE")
        copy_write_kib, _ = counter_per_dispatch(Path(args.copy_write), COPY_KERNEL, "WRITE_SIZE")
        fetch_scale = copy_bytes / (copy_fetch_kib * 1024.0)
        write_scale = copy_bytes / (copy_write_kib * 1024.0)
        fetch_kib, _ = counter_per_dispatch(Path(args.fetch), args.kernel, "FETCH_SIZE")
        write_kib, _ = counter_per_dispatch(Path(args.write), args.kernel, "WRITE_SIZE")
        out.update({"fetch_scale": round(fetch_scale, 4), "write_scale": round(write_scale, 4),
                    "hbm_fetch_bytes_per_launch": fetch_kib * 1024.0 * fetch_scale,
                    "hbm_write_bytes_per_launch": write_kib * 1024.0 * write_scale})
        out["hbm_bytes_per_launch"] = out["hbm_fetch_bytes_per_launch"] + out["hbm_write_bytes_per_launch"]
        out["hbm_gbs"] = round(out["hbm_bytes_per_launch"] / (duration_ms / 1000.0) / 1e9, 1)
    lines = [f"# {args.title}", "", f"Workload: {args.workload}", "", "Commands:", ""]
    lines += [f"    {c}" for c in args.pmc_command]
    lines += ["", f"Mean per dispatch of `{args.kernel}` (duration {duration_ms:.3f} ms):", ""]
    lines += ["| figure | value |", "|---|---|"]
    lines += [f"| {k} | {v} |" for k, v in out.items()
              if k not in ("kernel", "workload", "pmc_commands", "dispatches_per_pass")]
    lines += ["", "Raw counter means per dispatch:", "", "| counter | value |", "|---|---|"]
    lines += [f"| {k} | {merged[k]:.6g} |" for k in sorted(merged)]
    if args.note:
        lines += ["", args.note]
    Path(args.out_text).write_text("\n".join(lines) + "\n")
    if args.out_json:
        Path(args.out_json).write_text(json.dumps(out, indent=1) + "\n")
    print("\n".join(lines))
    return 0


def summarize(args) -> int:
    stats = kernel_stats(Path(args.trace))
    lines = [f"# {args.title}", "", "Command:", "", f"    {args.trace_command}", "",
             "| kernel | calls | average ms | min ms | max ms | share % |",
             "|---|---|---|---|---|---|"]
    march_avg_ms = None
    for row in stats:
        name = row["Name"]
        short = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
        avg = float(row["AverageNs"]) / 1e6
        if MARCH_KERNEL in name and march_avg_ms is None:
            march_avg_ms = avg
        lines.append(f"| `{short}` | {row['Calls']} | {avg:.3f} | {float(row['MinNs']) / 1e6:.3f} | "
                     f"{float(row['MaxNs']) / 1e6:.3f} | {float(row['Percentage']):.2f} |")
    result = {"march_kernel_average_ms": march_avg_ms, "trace_command": args.trace_command}
    if args.march_window:
        first, count = (int(v) for v in args.march_window.split(":"))
        durations = march_dispatch_ms(Path(args.trace))[first:first + count]
        window_ms = float(np.mean(durations))
        result.update({"march_kernel_average_ms_all": march_avg_ms,
                       "march_kernel_average_ms": window_ms,
                       "march_window": {"first": first, "count": len(durations)}})
        lines += ["", f"`march_kernel` dispatches {first + 1}-{first + len(durations)} in dispatch order "
                  f"(the timed tiles of the first workload; {first} warm-up launches precede them): "
                  f"average {window_ms:.3f} ms, min {min(durations):.3f}, max {max(durations):.3f}."]
    if args.fetch:
        copy_bytes = args.copy_mib * 2**20
        copy_fetch_kib, _ = counter_per_dispatch(Path(args.copy_fetch), COPY_KERNEL, "FETCH_SIZE")
        copy_write_kib, _ = counter_per_dispatch(Path(args.copy_write), COPY_KERNEL, "WRITE_SIZE")
        fetch_scale = copy_bytes / (copy_fetch_kib * 1024.0)
        write_scale = copy_bytes / (copy_write_kib * 1024.0)
        fetch_kib, n_fetch = counter_per_dispatch(Path(args.fetch), MARCH_KERNEL, "FETCH_SIZE")
        write_kib, n_write = counter_per_dispatch(Path(args.write), MARCH_KERNEL, "WRITE_SIZE")
        fetch_bytes = fetch_kib * 1024.0 * fetch_scale
        write_bytes = write_kib * 1024.0 * write_scale
        result.update({
            "copy_bytes": copy_bytes,
            "copy_fetch_kib": copy_fetch_kib, "copy_write_kib": copy_write_kib,
            "fetch_scale": fetch_scale, "write_scale": write_scale,
            "march_dispatches": {"fetch": n_fetch, "write": n_write},
            "march_fetch_kib_per_dispatch": fetch_kib,
            "march_write_kib_per_dispatch": write_kib,
            "hbm_fetch_bytes_per_launch": fetch_bytes,
            "hbm_write_bytes_per_launch": write_bytes,
            "hbm_bytes_per_launch": fetch_bytes + write_bytes,
            "pmc_commands": args.pmc_command,
        })
        lines += ["", "HBM traffic of `march_kernel` (rocprofv3 --pmc, one counter per run):", ""]
        lines += [f"    {c}" for c in args.pmc_command]
        lines += ["",
                  f"- copy probe of {copy_bytes} B: FETCH_SIZE {copy_fetch_kib:.0f} KiB, "
                  f"WRITE_SIZE {copy_write_kib:.0f} KiB -> scale {fetch_scale:.4f} / {write_scale:.4f}",
                  f"- march, mean of {n_fetch} dispatches: FETCH_SIZE {fetch_kib:.0f} KiB -> "
                  f"{fetch_bytes / 1e6:.1f} MB; WRITE_SIZE {write_kib:.0f} KiB -> {write_bytes / 1e6:.1f} MB",
                  f"- HBM bytes per launch: {(fetch_bytes + write_bytes) / 1e6:.1f} MB"]
    if args.note:
        lines += ["", args.note]
    Path(args.out_text).write_text("\n".join(lines) + "\n")
    if args.out_json:
        Path(args.out_json).write_text(json.dumps(result, indent=1) + "\n")
    print("\n".join(lines))
    return 0


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = p.add_subparsers(dest="command", required=True)
    c = sub.add_parser("copy", help="one device copy of --mib MiB")
    c.add_argument("--mib", type=int, default=1024)
    w = sub.add_parser("workload", help="full-site march tiles and one lighting launch")
    w.add_argument("--tiles", type=int, default=4)
    w.add_argument("--times", type=int, default=4380)
    w.add_argument("--azimuth-major", action="store_true",
                   help="march into and light from [azimuth][pixel] horizons")
    k = sub.add_parser("counters", help="summarize --pmc passes of one kernel")
    k.add_argument("--kernel", required=True, help="substring of the kernel name")
    k.add_argument("--pmc", action="append", required=True, help="directory of one --pmc pass")
    k.add_argument("--pmc-command", action="append", default=[])
    k.add_argument("--fetch", help="directory of the kernel's FETCH_SIZE pass")
    k.add_argument("--write", help="directory of the kernel's WRITE_SIZE pass")
    k.add_argument("--copy-fetch", help="directory of the copy probe's FETCH_SIZE pass")
    k.add_argument("--copy-write", help="directory of the copy probe's WRITE_SIZE pass")
    k.add_argument("--copy-mib", type=int, default=1024)
    k.add_argument("--cu-count", type=int, default=256)
    k.add_argument("--xcc-count", type=int, default=8)
    k.add_argument("--workload", default="")
    k.add_argument("--title", default="counters")
    k.add_argument("--note", default="")
    k.add_argument("--out-text", required=True)
    k.add_argument("--out-json")
    s = sub.add_parser("summarize", help="summarize rocprofv3 CSV output")
    s.add_argument("--trace", required=True, help="directory of the --kernel-trace --stats run")
    s.add_argument("--trace-command", required=True)
    s.add_argument("--fetch", help="directory of the --pmc FETCH_SIZE run")
    s.add_argument("--write", help="directory of the --pmc WRITE_SIZE run")
    s.add_argument("--copy-fetch", help="directory of the copy probe's FETCH_SIZE run")
    s.add_argument("--copy-write", help="directory of the copy probe's WRITE_SIZE run")
    s.add_argument("--copy-mib", type=int, default=1024)
    s.add_argument("--pmc-command", action="append", default=[])
    s.add_argument("--march-window", help="FIRST:COUNT march dispatches (dispatch order) to average")
    s.add_argument("--title", default="march_kernel, full site")
    s.add_argument("--note", default="")
    s.add_argument("--out-text", required=True)
    s.add_argument("--out-json")
    args = p.parse_args(argv)
    if args.command == "copy":
        return _copy(args.mib)
    if args.command == "workload":
        return _workload(args)
    if args.command == "counters":
        return counters(args)
    return summarize(args)


if __name__ == "__main__":
    sys.exit(main())

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
    return summarize(args)


if __name__ == "__main__":
    sys.exit(main())

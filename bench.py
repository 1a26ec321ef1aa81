#!/usr/bin/env python3
"""Full-site horizon + lighting benchmark on one MI355X.

Workload (synthetic, seeded): a 2048x2048 float32 primary DEM at 5 m/px (fBm
Hurst 0.8, 300 m RMS, 40 paraboloid craters, seed 1) and a 4096x4096 float32
regional DEM at 60 m/px (seed 2), both centred on an 89.5 S site in south-polar
stereographic; 256 horizon tiles of 128x128 px x 1440 azimuths over both DEM
passes, then from those horizons the lightmap (finite-disk sun fraction bytes
per time), the Earth-visibility fraction and the PSR mask for 125 Sun and 125
Earth Moon-ME vectors at 6-hour steps over 31 days.

A step is one batch of tiles (``--tiles-per-step``, default 32): FP64 segment
generation, the fused two-pass march into HBM (degrees), and the streamed
lighting pass over the batch. Tiles are taken in the order i * 97 mod 256, so
any run of consecutive steps samples rows and columns of the whole site evenly;
the default 8 timed steps march every one of the 256 tiles once. Each marched
tile keeps its own slot in HBM (24.2 GB for the whole site, azimuth-major), so
the year of lighting (4380 Sun + 4380 Earth vectors and the PSR test) runs over
all 256 tiles in one launch. DEMs and vectors are resident in HBM before the
timed region.

Besides the headline the line carries: the march's roofline (FP64 VALU: the
method's FLOPs per hierarchy test and per level-0 sample, counted by the
kernel's counting build on timed tiles, over the HIP-event time), the lighting
roofline of the 365-day pass, the batched persistent path on the 48-site
workload in one launch against the per-tile launches, the end-to-end
``generate_horizons`` path with compressed tile writes, the adjoint, and the
CPU baselines. The CPU leg also checks one timed full-site tile and one batched
landing-site tile against the float64 oracle (64 px x 1440 az each) and writes
the errors into the horizon and batched sections.

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "horizon tiles/s + lightmap px·times/s, 1 MI355X; batched vs plain at ≤0.01° err"
FP64_VALU_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
AZIMUTHS = 1440
PIXELS = 128 * 128
# FP64 operations the method performs (the float64 oracle's hierarchy_maximum,
# oracle/lsh_oracle.cpp): per hierarchy test, the position (two quartics, planar
# distance, chord, footprint: 41) and the test (two tangents, cell exit, cull
# slope: 54); per level-0 sample, the bilinear height, the slope and the
# adaptive step (37). A square root or a division counts as one operation.
MARCH_FLOP_PER_TEST = 95
MARCH_FLOP_PER_SAMPLE = 37
# lighting per pixel-time: Sun = frame transform, azimuth, elevation (27 FP64)
# + the 16-slice float32 disk walk (100); Earth = frame + az/el + horizon
# interpolation and margin (33). The PSR vectors are not counted (the mask
# stops at the first lighting vector, so its work is data dependent).
LIGHT_FLOP_SUN = 127
LIGHT_FLOP_EARTH = 33


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=8)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--tiles-per-step", type=int, default=32)
    p.add_argument("--cpu-sample-pixels", type=int, default=256,
                   help="pixels of one tile the CPU baseline marches (0 disables it)")
    p.add_argument("--counter-tiles", type=int, default=4,
                   help="timed tiles re-run with traversal counters for the roofline (0 skips)")
    p.add_argument("--no-probes", action="store_true",
                   help="skip the copy / gather peak probes")
    p.add_argument("--batch-tiles", type=int, default=48,
                   help="tiles of the 48-site workload timed plain and batched (0 skips)")
    p.add_argument("--pipeline-tiles", type=int, default=12,
                   help="tiles (a block of the site's top-left corner, 4 across) of the "
                        "end-to-end generate_horizons and product legs (0 skips)")
    p.add_argument("--no-products", dest="products", action="store_false",
                   help="skip the lightmap / PSR products over the generated tiles")
    p.add_argument("--pixel-major", action="store_true",
                   help="keep resident horizons [pixel][azimuth] (the tile file layout) "
                        "instead of [azimuth][pixel]")
    p.add_argument("--no-year-lighting", dest="year_lighting", action="store_false",
                   help="skip the 365-day lighting pass over resident horizons")
    p.add_argument("--no-adjoint", dest="adjoint", action="store_false",
                   help="skip the DEM-sensitivity (adjoint) leg")
    return p.parse_args()


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(tiles shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        return 2
    from lunarscout_amd import synthetic
    from lunarscout_amd._lib import DeviceBuffer, Event, HORIZON_COUNTERS, HORIZON_DEGREES
    from lunarscout_amd.hip_backend import HipSession
    from lunarscout_amd.lighting import HipLightingSession, reduce_sun_vectors_for_psr

    t0 = time.time()
    dems = synthetic.site_dems(synthetic.FULL_SITE)
    primary = dems[0]
    sun = synthetic.sun_vectors(primary, 125, 6.0, seed=3)
    earth = synthetic.earth_vectors(125, 6.0, seed=4)
    psr_vectors, _ = reduce_sun_vectors_for_psr(primary, sun)
    log(f"inputs generated in {time.time() - t0:.1f}s; PSR vectors {len(psr_vectors)}")

    session = HipSession(0)
    log("device:", session.device_name())
    for d in dems:
        session.dem_id(d)
    session.synchronize()
    lighting = HipLightingSession(session)
    sun_d = DeviceBuffer.from_host(session, sun)
    earth_d = DeviceBuffer.from_host(session, earth)
    psr_d = DeviceBuffer.from_host(session, psr_vectors)

    grid = [(x, y) for y in range(0, primary.height, 128) for x in range(0, primary.width, 128)]
    tiles = [grid[(i * 97) % len(grid)] for i in range(len(grid))]
    per_step = args.tiles_per_step
    major = not args.pixel_major
    # every marched tile keeps its own slot in HBM (all 256 with the defaults:
    # 24.2 GB); the i-th tile marched goes to slot i mod slots
    slots = min(len(tiles) // per_step, args.warmup + args.steps) * per_step
    if slots < per_step:
        raise SystemExit("--tiles-per-step must not exceed the site's 256 tiles")
    horizons = DeviceBuffer.empty(session, (slots, AZIMUTHS, PIXELS) if major
                                  else (slots, PIXELS, AZIMUTHS), np.float32)
    resident = {}  # slot -> the tile its horizons belong to
    segments = None
    outputs = None
    stats = {"march_ms": 0.0, "march_launches": 0, "segments_ms": 0.0, "lighting_ms": 0.0,
             "lighting_launches": 0, "tiles": 0}

    def step(index: int, record: bool):
        nonlocal segments, outputs
        batch = [tiles[(index * per_step + i) % len(tiles)] for i in range(per_step)]
        first = (index * per_step) % slots
        events = []
        for i, (c, r) in enumerate(batch):
            e0 = Event(session).record()
            segments = session.segments_device(dems, tile_column=c, tile_row=r, out=segments)
            e1 = Event(session).record()
            session.march_device(dems, segments, tile_column=c, tile_row=r,
                                 flags=HORIZON_DEGREES, output=horizons.slot(first + i),
                                 azimuth_major=major)
            e2 = Event(session).record()
            events.append((e0, e1, e2))
            resident[first + i] = (c, r)
        l0 = Event(session).record()
        outputs = lighting.lighting_tiles(primary, batch, horizons.slots(first, per_step),
                                          sun=sun_d, earth=earth_d,
                                          psr_vectors=psr_d, sun_bytes=True, sun_mean=True,
                                          earth_visible=True, psr=True, outputs=outputs,
                                          azimuth_major=major)
        l1 = Event(session).record()
        if record:
            for e0, e1, e2 in events:
                stats["segments_ms"] += e0.elapsed_ms(e1)
                stats["march_ms"] += e1.elapsed_ms(e2)
                stats["march_launches"] += 1
            stats["lighting_ms"] += l0.elapsed_ms(l1)
            stats["lighting_launches"] += 1
            stats["tiles"] += len(batch)

    for w in range(args.warmup):
        step(w, record=False)
    session.synchronize()
    log(f"warmup done ({args.warmup} steps)")
    start = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k, record=True)
        log(f"step {k + 1}/{args.steps} done")
    session.synchronize()
    elapsed = time.perf_counter() - start

    tiles_per_s = stats["tiles"] / elapsed
    march_ms = stats["march_ms"] / max(stats["march_launches"], 1)
    lighting_ms = stats["lighting_ms"] / max(stats["lighting_launches"], 1)
    light_px_times = per_step * PIXELS * (len(sun) + len(earth) + len(psr_vectors))
    resident = [resident[i] for i in range(slots)]  # slot order

    # traversal counters on timed tiles -> the method's FLOPs and gathered bytes per tile
    counted = None
    if args.counter_tiles > 0:
        counters = DeviceBuffer.empty(session, (4,), np.int64)
        scratch = DeviceBuffer.empty(session, (PIXELS, AZIMUTHS), np.float32)
        samples = tests = 0
        timed = tiles[args.warmup * per_step:(args.warmup + args.steps) * per_step] or tiles
        pick = [timed[i] for i in np.linspace(0, len(timed) - 1, args.counter_tiles).astype(int)]
        for c, r in pick:
            counters.zero()
            segments = session.segments_device(dems, tile_column=c, tile_row=r, out=segments)
            session.march_device(dems, segments, tile_column=c, tile_row=r, flags=HORIZON_COUNTERS,
                                 output=scratch, counters=counters)
            v = counters.to_host()
            samples += int(v[0])
            tests += int(v[2])  # every hierarchy test (culls, nodata skips, descents, samples)
        scratch.free()
        counters.free()
        counted = {"tiles": len(pick), "samples": samples / len(pick), "tests": tests / len(pick)}
    roofline = _march_roofline(counted, march_ms, len(dems))
    # the tile file layout [pixel][azimuth] of one resident tile, for the codec and CPU legs
    cpu_light_tile = horizons.slot(0).to_host()
    if major:
        cpu_light_tile = np.ascontiguousarray(cpu_light_tile.T)
    codec = _codec_leg(cpu_light_tile)
    long_lighting = reducers = None
    if args.year_lighting:
        long_lighting = _year_lighting(session, lighting, primary, horizons, resident, major)
        reducers = _durations_leg(session, lighting, primary, horizons, resident, major)
    horizons.free()
    if segments is not None:
        segments.free()
    if outputs is not None:
        outputs.free()
    if not args.no_probes:
        from lunarscout_amd.profiling import probe_peaks

        peaks = probe_peaks(session)
        roofline["measured_peaks"] = peaks
        if roofline.get("gathered_bytes_per_launch"):
            roofline["gather_frac_of_l2_peak"] = round(
                roofline["gathered_bytes_per_launch"] / (march_ms / 1000.0) / 1e9
                / peaks["gather_l2_gbs"], 4)

    result = {
        "metric": METRIC,
        "value": round(tiles_per_s, 4),
        "unit": "tiles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1000.0, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (ray geometry, heights, slopes; DEMs stored f32; lighting f32 disk slices)",
        "data": "synthetic (seeded fBm + craters DEMs, synthetic Moon-ME vectors)",
        "config": {
            "workload": "full site, plain HIP path: 2048^2 @5 m primary + 4096^2 @60 m regional, "
                        "256 tiles x 1440 az, lighting + Earth + PSR from 125 vectors",
            "tiles_per_step": per_step,
            "resident_tiles": slots,
            "seeds": {"primary": 1, "regional": 2, "sun": 3, "earth": 4},
        },
        "horizon": {
            "march_ms_per_tile": round(march_ms, 3),
            "segments_ms_per_tile": round(stats["segments_ms"] / max(stats["march_launches"], 1), 3),
            "counted": counted,
        },
        "tile_codec": codec,
        "lighting": {
            "value": round(light_px_times / (lighting_ms / 1000.0), 1),
            "unit": "px·times/s (sun + earth + PSR vectors)",
            "kernel_ms_per_batch": round(lighting_ms, 3),
            "times": {"sun": len(sun), "earth": len(earth), "psr": len(psr_vectors)},
            "year": long_lighting,
            "year_reducers": reducers,
        },
        "roofline": roofline,
    }
    checks = {"horizon": (dems, resident[0], cpu_light_tile)}
    result["horizon"]["max_err_deg_vs_oracle"] = None  # filled by the CPU leg's oracle check
    if args.batch_tiles > 0:
        checks["batched"], result["batched"] = _batched_vs_plain(session, dems[1],
                                                                 args.batch_tiles)
    if args.pipeline_tiles > 0:
        corner = [(c * 128, r * 128) for r in range(0, (args.pipeline_tiles + 3) // 4)
                  for c in range(4)][:args.pipeline_tiles]
        with tempfile.TemporaryDirectory(prefix="lsh-bench-") as root:
            result["generate_horizons"] = _pipeline_leg(session, dems, corner, Path(root))
            if args.products:
                result["products"] = _products_leg(primary, corner, Path(root))
    if args.adjoint:
        result["adjoint"] = _adjoint_leg(session)
    if args.cpu_sample_pixels > 0:
        result["cpu_baseline"] = _cpu_baseline(dems, args.cpu_sample_pixels, cpu_light_tile,
                                               resident[0])
        for name, error in _oracle_checks(checks).items():
            result[name]["max_err_deg_vs_oracle"] = error
    print(json.dumps(result), flush=True)
    return 0


def _profile_summary(name: str):
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return None


def _march_roofline(counted, march_ms: float, passes: int) -> dict:
    """FP64 VALU roofline of the march from the counted tests and samples."""
    pmc = _profile_summary("march_full_site_pmc.json") or {}
    out = {"kernel": "march_kernel", "bound": "valu", "unit": "TFLOP/s",
           "peak": FP64_VALU_PEAK_TFLOPS, "achieved": None, "frac": None,
           "traffic": pmc.get("hbm_bytes_per_launch"),
           "flop_model": {"per_hierarchy_test": MARCH_FLOP_PER_TEST,
                          "per_level0_sample": MARCH_FLOP_PER_SAMPLE},
           "valu_busy_profile": pmc.get("valu_busy"),
           "fp64_tflops_profile": pmc.get("fp64_tflops")}
    if counted:
        flops = MARCH_FLOP_PER_TEST * counted["tests"] + MARCH_FLOP_PER_SAMPLE * counted["samples"]
        achieved = flops / (march_ms / 1000.0) / 1e12
        rays = PIXELS * AZIMUTHS
        # bytes the kernel gathers: one stored 4-byte bound per hierarchy test,
        # the four 4-byte level-0 taps per sample, the segment corners and the output
        gathered = 4.0 * counted["tests"] + 16.0 * counted["samples"] + 256.0 * rays * passes \
            + 4.0 * rays
        # the method's 2x2 gather: 16 B per bound test and per sample (the kernel
        # reads a test's bound as one stored 4 B value instead)
        algorithmic = 16.0 * (counted["tests"] + counted["samples"]) + 256.0 * rays * passes \
            + 4.0 * rays
        seconds = march_ms / 1000.0
        out.update({"achieved": round(achieved, 3), "frac": round(achieved / FP64_VALU_PEAK_TFLOPS, 4),
                    "flop_per_launch": flops, "gathered_bytes_per_launch": gathered,
                    "gathered_gbs": round(gathered / seconds / 1e9, 1),
                    "algorithmic_bytes_per_launch": algorithmic,
                    "algorithmic_gbs": round(algorithmic / seconds / 1e9, 1)})
        if out["traffic"]:
            out["traffic_frac_of_hbm_peak"] = round(out["traffic"] / seconds / 1e9 / HBM_PEAK_GBS, 5)
    return out


def _codec_leg(degrees):
    """.cbin encode / decode of one marched full-site tile on the host threads."""
    from lunarscout_amd import tiles

    payload = tiles.encode_cbin(degrees)
    t0 = time.perf_counter()
    for _ in range(3):
        payload = tiles.encode_cbin(degrees)
    encode_s = (time.perf_counter() - t0) / 3
    out = np.empty_like(degrees)
    t0 = time.perf_counter()
    for _ in range(3):
        tiles.decode_cbin(payload, out=out)
    decode_s = (time.perf_counter() - t0) / 3
    return {"encode_ms_per_tile": round(encode_s * 1000, 2), "decode_ms_per_tile": round(decode_s * 1000, 2),
            "compressed_bytes": len(payload), "raw_bytes": int(degrees.nbytes),
            "threads": len(os.sched_getaffinity(0))}


def _year_lighting(session, lighting, primary, horizons, batch, major):
    """Long-duration lighting from stored horizons: 4380 Sun and 4380 Earth
    vectors (2-hour steps over 365 days, seeds 5 and 6) over the resident tiles,
    per-time sun-fraction bytes + Earth visibility + PSR in one streamed launch."""
    from lunarscout_amd import synthetic
    from lunarscout_amd._lib import DeviceBuffer, Event
    from lunarscout_amd.lighting import reduce_sun_vectors_for_psr

    sun = synthetic.sun_vectors(primary, 4380, 2.0, seed=5)
    earth = synthetic.earth_vectors(4380, 2.0, seed=6)
    psr_vectors, _ = reduce_sun_vectors_for_psr(primary, sun)
    bufs = [DeviceBuffer.from_host(session, v) for v in (sun, earth, psr_vectors)]
    kwargs = dict(sun=bufs[0], earth=bufs[1], psr_vectors=bufs[2], sun_bytes=True, sun_mean=True,
                  earth_visible=True, psr=True, azimuth_major=major)
    outputs = lighting.lighting_tiles(primary, batch, horizons, **kwargs)  # warm-up
    e0 = Event(session).record()
    outputs = lighting.lighting_tiles(primary, batch, horizons, outputs=outputs, **kwargs)
    e1 = Event(session).record()
    ms = e0.elapsed_ms(e1)
    outputs.free()
    for b in bufs:
        b.free()
    pixels = len(batch) * PIXELS
    px_times = pixels * (len(sun) + len(earth) + len(psr_vectors))
    flops = pixels * (LIGHT_FLOP_SUN * len(sun) + LIGHT_FLOP_EARTH * len(earth))
    achieved = flops / (ms / 1000.0) / 1e12
    return {"workload": f"{len(batch)} resident full-site tiles x (4380 Sun + 4380 Earth + "
                        f"{len(psr_vectors)} PSR vectors), 2-hour steps over 365 days",
            "layout": "[azimuth][pixel]" if major else "[pixel][azimuth]",
            "value": round(px_times / (ms / 1000.0), 1), "unit": "px·times/s",
            "kernel_ms": round(ms, 3),
            "roofline": {"kernel": "lighting_kernel", "bound": "valu", "unit": "TFLOP/s",
                         "achieved": round(achieved, 3), "peak": FP64_VALU_PEAK_TFLOPS,
                         "frac": round(achieved / FP64_VALU_PEAK_TFLOPS, 4),
                         "flop_model": {"per_sun_px_time": LIGHT_FLOP_SUN,
                                        "per_earth_px_time": LIGHT_FLOP_EARTH,
                                        "psr": "not counted"},
                         "valu_busy_profile": (_profile_summary("lighting_year_pmc.json") or {})
                         .get("valu_busy")}}


def _durations_leg(session, lighting, primary, horizons, batch, major):
    """Downstream reducers fused into the lighting walk over the resident tiles:
    the longest sunlight (fraction >= 0.5) AND Earth-visible (margin >= 0 deg)
    duration per calendar-month start interval, and the monthly safe haven (Sun
    fraction < 0.2 during an Earth outage, margin < 2 deg), from 4380 Sun and
    4380 Earth vectors (2-hour steps over 365 days, seeds 5 and 6)."""
    from datetime import datetime, timedelta, timezone

    from lunarscout_amd import durations, synthetic
    from lunarscout_amd._lib import DeviceBuffer, Event

    sun = synthetic.sun_vectors(primary, 4380, 2.0, seed=5)
    earth = synthetic.earth_vectors(4380, 2.0, seed=6)
    bufs = [DeviceBuffer.from_host(session, v) for v in (sun, earth)]
    start = datetime(2027, 1, 1, tzinfo=timezone.utc)
    times = [start + timedelta(hours=2 * i) for i in range(4380)]
    stop = times[-1] + timedelta(hours=2)
    plans = {
        "longest_sunlight_earth_duration": durations.longest_duration_plan(
            times, evaluation_start_utc=start, evaluation_stop_utc=stop,
            start_intervals=durations.monthly_candidate_intervals(start, stop),
            sun_threshold=0.5, earth_threshold_deg=0.0),
        "safe_haven": durations.safe_haven_plan(times, sunlight_threshold=0.2,
                                                earth_threshold_deg=2.0, time_step_hours=2.0),
    }
    result = {"workload": f"{len(batch)} resident full-site tiles x 4380 times (Sun + Earth), "
                          "2-hour steps over 365 days; 12 monthly bands each"}
    for name, plan in plans.items():
        out = lighting.duration_tiles(primary, batch, horizons, plan, bufs[0], bufs[1],
                                      azimuth_major=major)  # warm-up
        e0 = Event(session).record()
        out = lighting.duration_tiles(primary, batch, horizons, plan, bufs[0], bufs[1], out=out,
                                      azimuth_major=major)
        e1 = Event(session).record()
        ms = e0.elapsed_ms(e1)
        values = out.to_host()
        out.free()
        px_times = len(batch) * PIXELS * plan.time_count
        result[name] = {"value": round(px_times / (ms / 1000.0), 1), "unit": "px·times/s",
                        "ms": round(ms, 3), "bands": plan.bands,
                        "finite_fraction": round(float(np.isfinite(values).mean()), 4)}
    for b in bufs:
        b.free()
    return result


def _adjoint_leg(session):
    """DEM sensitivity of the mean smoothed sun fraction: 512^2 @10 m (seed 7) +
    1024^2 @60 m regional (seed 8), 16 tiles x 720 azimuths, 2190 Sun vectors
    (4-hourly over 365 days, seed 10); forward = march with winners + loss,
    backward = horizon gradient + re-march adjoint. Warm-up pass first."""
    from lunarscout_amd import synthetic
    from lunarscout_amd.adjoint import SunLossProblem

    dems = synthetic.site_dems(synthetic.ADJOINT_SITE)
    tiles = [(x, y) for y in range(0, 512, 128) for x in range(0, 512, 128)]
    sun = synthetic.sun_vectors(dems[0], 2190, 4.0, seed=10)
    problem = SunLossProblem(dems, tiles, sun, azimuth_count=720, smoothing_deg=0.08,
                             session=session)
    problem.forward()
    problem.backward()
    t0 = time.perf_counter()
    loss = problem.forward()
    forward_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    grads = problem.backward()
    backward_s = time.perf_counter() - t0
    sec = dict(problem.seconds)
    problem.close()
    return {
        "workload": "512^2 @10 m + 1024^2 @60 m, 16 tiles x 720 az, 2190 Sun vectors (4-hourly, "
                    "365 d), slice clamps rounded over 0.08 deg",
        "loss": loss,
        "forward_s": round(forward_s, 4), "backward_s": round(backward_s, 4),
        "backward_over_forward": round(backward_s / forward_s, 4),
        "kernel_s": {k: round(v, 4) for k, v in sec.items()},
        "grad_nonzero_nodes": [int(np.count_nonzero(g)) for g in grads],
    }


def _batched_vs_plain(session, regional, tile_count):
    """48 landing sites sharing the regional DEM: per-tile launches against one
    persistent launch (hipGraph, autotuned shape) over the same tiles."""
    from lunarscout_amd import synthetic
    from lunarscout_amd._lib import HORIZON_DEGREES, DeviceBuffer, Event
    from lunarscout_amd.batch import HorizonBatch, Site, autotune

    t0 = time.time()
    sites = synthetic.landing_sites()
    site_dems = [synthetic.landing_site_dem(s) for s in sites]
    work = []  # every tile of every patch
    for i, s in enumerate(sites):
        for k in range(s.tile_count):
            work.append((i, (k % s.tiles_x) * 128, (k // s.tiles_x) * 128))
    pick = np.linspace(0, len(work) - 1, min(tile_count, len(work))).astype(int)
    sample = [work[i] for i in pick]
    n = len(sample)
    batch_sites = [Site((d, regional)) for d in site_dems]
    for d in site_dems:
        session.dem_id(d)
    log(f"48-site workload: {len(work)} tiles, timing {n} ({time.time() - t0:.1f}s setup)")
    batch = HorizonBatch(session, batch_sites, max_tiles=n)
    out = DeviceBuffer.empty(session, (n, PIXELS, AZIMUTHS), np.float32)
    tune = autotune(batch, sample[: min(8, n)], out, key="landing-sites-48", use_cache=False)
    log(f"autotune: shape {tune['config']} workgroups {tune['workgroups']}")
    # plain: one segments + one march launch per tile, each into its own slot
    plain = DeviceBuffer.empty(session, (n, PIXELS, AZIMUTHS), np.float32)
    segments = None
    kernel_ms = 0.0
    session.synchronize()
    w0 = Event(session).record()
    marks = []
    for slot, (site, c, r) in enumerate(sample):
        dl = [site_dems[site], regional]
        e0 = Event(session).record()
        segments = session.segments_device(dl, tile_column=c, tile_row=r, out=segments)
        session.march_device(dl, segments, tile_column=c, tile_row=r, flags=HORIZON_DEGREES,
                             output=plain.slot(slot))
        e1 = Event(session).record()
        marks.append((e0, e1))
    w1 = Event(session).record()
    session.synchronize()
    plain_ms = w0.elapsed_ms(w1)
    kernel_ms = sum(a.elapsed_ms(b) for a, b in marks)
    log(f"plain: {plain_ms / n:.1f} ms/tile, GPU busy {kernel_ms / plain_ms:.4f}")
    # batched: all sampled tiles in one persistent launch (graph replay after the capture)
    batch.run(sample, out, config=tune["config"], workgroups=tune["workgroups"], degrees=True,
              use_graph=True)
    session.synchronize()
    e0 = Event(session).record()
    batch.run(sample, out, config=tune["config"], workgroups=tune["workgroups"], degrees=True,
              use_graph=True)
    e1 = Event(session).record()
    session.synchronize()
    batched_ms = e0.elapsed_ms(e1)
    log(f"batched: {batched_ms / n:.1f} ms/tile")
    identical = 0
    for slot in range(n):
        a, b = plain.slot(slot).to_host(), out.slot(slot).to_host()
        identical += int(np.all((a == b) | (np.isneginf(a) & np.isneginf(b))))
    # the first timed tile's batched degrees, for the oracle check of the CPU leg
    site, c, r = sample[0]
    check = ([site_dems[site], regional], (c, r), out.slot(0).to_host())
    batch.close()
    out.free()
    plain.free()
    if segments is not None:
        segments.free()
    busy = kernel_ms / plain_ms
    return check, {
        "workload": "48 landing-site patches (tile counts log-uniform 1..64, 5-20 m/px, seeds "
                    "100-147, centres within +-60 km) sharing the 4096^2 @60 m regional DEM; "
                    f"{len(work)} tiles, {n} timed (evenly spaced over the sites)",
        "plain_tiles_per_s": round(n / (plain_ms / 1000.0), 4),
        "batched_tiles_per_s": round(n / (batched_ms / 1000.0), 4),
        "speedup": round(plain_ms / batched_ms, 4),
        "batched_launches": 1,
        "plain_gpu_busy_fraction": round(busy, 5),
        "plain_launch_gap_ms_per_tile": round((plain_ms - kernel_ms) / n, 4),
        "autotuned": {"shape": tune["config"], "workgroups": tune["workgroups"],
                      "trials": [(t["shape"], t["workgroups"], round(t["ms"], 1))
                                 for t in tune["trials"]]},
        "bit_identical_tiles": f"{identical}/{n}",
        "max_err_deg_vs_oracle": None,  # filled by the CPU leg's oracle check
    }


def _pipeline_leg(session, dems, tiles, root):
    """generate_horizons' path end to end on a block of the full site: the
    bounded pipeline (GPU segments + march, degrees copied back, .cbin encode
    and the staged, fsynced tile publish in the writer thread) into the empty
    directory ``root``; tiles/s from the wall clock."""
    from lunarscout_amd.pipeline import PatchDescriptor, run_bounded_pipeline
    from lunarscout_amd.tiles import HorizonTileStore

    patches = [PatchDescriptor(i, c, r, c // 128, r // 128, 128, 128) for i, (c, r) in enumerate(tiles)]

    def factory(_worker):
        def process(patch, _prepared):
            return session.horizon_tile(dems, tile_column=patch.tile_x, tile_row=patch.tile_y,
                                        degrees=True)
        return process

    started = time.perf_counter()
    result = run_bounded_pipeline(patches, store=HorizonTileStore(root),
                                  prepare_patch=lambda _p: None, processor_factory=factory,
                                  compress=True, skip_existing=False,
                                  prepared_queue_capacity=1, writer_queue_capacity=1)
    wall = time.perf_counter() - started
    size = sum(p.stat().st_size for p in Path(root).rglob("*.cbin"))
    return {"workload": f"{len(tiles)} full-site tiles (the top-left block) through "
                        "generate_horizons' bounded pipeline, compressed (.cbin) staged writes",
            "tiles_per_s": round(len(tiles) / wall, 4), "wall_s": round(wall, 3),
            "compute_s": round(result.compute_seconds, 3), "write_s": round(result.write_seconds, 3),
            "bytes_written": int(size)}


def _products_leg(primary, tiles, root):
    """generate_psr and generate_lightmap end to end (hip backend) over the
    tiles the generate_horizons leg wrote: the DEM block those tiles cover, tile
    reads + .cbin decodes on 4 threads, the GPU calculation, compressed GeoTIFF
    blocks and batched durable checkpoints, overlapped by the bounded pipeline.
    PSR from a year of 2-hourly Sun vectors (4380, reduced); the lightmap from
    two years of 6-hourly vectors (2921 bands), the reference's two published
    product workloads; patches/s from the wall clock."""
    from lunarscout_amd import synthetic
    from lunarscout_amd.geometry import DemGrid
    from lunarscout_amd.products import generate_lightmap, generate_psr

    height = max(r for _, r in tiles) + 128
    width = max(c for c, _ in tiles) + 128
    dem = DemGrid(np.ascontiguousarray(primary.elevation_m[:height, :width]),
                  primary.geo_transform, primary.projection)
    out = {"workload": f"{len(tiles)} patches of the full site's {width}x{height} px top-left "
                       "block, from the generate_horizons leg's .cbin tiles"}
    for name, fn, count, hours in (("psr", generate_psr, 4380, 2.0),
                                   ("lightmap", generate_lightmap, 2921, 6.0)):
        vectors = synthetic.sun_vectors(primary, count, hours, seed=5)
        path = root / f"{name}.tif"
        started = time.perf_counter()
        fn(dem, root, path, sun_vectors_m=vectors, backend="hip", overwrite=True)
        wall = time.perf_counter() - started
        out[name] = {"vectors": count, "bands": 1 if name == "psr" else count,
                     "patches_per_s": round(len(tiles) / wall, 4), "wall_s": round(wall, 3),
                     "bytes_written": path.stat().st_size}
        path.unlink()
        log(f"{name} product: {len(tiles)} patches in {wall:.2f} s")
    return out


def _oracle_checks(checks, pixels: int = 64) -> dict:
    """Max |GPU - float64 oracle| in degrees on ``pixels`` pixels x 1440 az of one
    timed tile per leg: {leg: {"value", "within_1e-3_deg", "sample"}}."""
    import oracle

    out = {}
    px = np.linspace(0, PIXELS - 1, pixels).astype(np.int32)
    for name, (dems, (c, r), degrees) in checks.items():
        segs = oracle.segments(dems, tile_column=c, tile_row=r, tile_width=128,
                               azimuth_count=AZIMUTHS)
        ref, _ = oracle.horizon_tile([oracle.OraclePyramid(d) for d in dems], segs, tile_column=c,
                                     tile_row=r, pixels=px, float64=True)
        ref_deg = np.degrees(np.arctan(ref))
        got = degrees[px].astype(np.float64)
        finite = np.isfinite(ref_deg)
        delta = np.abs(got - ref_deg)[finite]
        out[name] = {"value": float(delta.max()) if delta.size else 0.0,
                     "within_1e-3_deg": float(np.mean(delta <= 1e-3)) if delta.size else 1.0,
                     "horizonless_bins_differing": int(np.count_nonzero(finite != np.isfinite(got))),
                     "sample": f"{pixels} px x {AZIMUTHS} az of tile ({c}, {r})"}
        log(f"oracle check {name}: max |delta| {out[name]['value']:.3g} deg")
    return out


def _cpu_baseline(dems, sample_pixels, horizon_tile, tile):
    """The C++/OpenMP oracle (float32 reference arithmetic) on bounded samples."""
    import oracle

    cores = oracle.cpu_threads()
    pyramids = [oracle.OraclePyramid(d) for d in dems]
    march_tile = (1024, 1024)
    t0 = time.perf_counter()
    segs = oracle.segments(dems, tile_column=march_tile[0], tile_row=march_tile[1], tile_width=128,
                           azimuth_count=AZIMUTHS, threads=cores)
    t_seg = time.perf_counter() - t0
    pixels = np.linspace(0, PIXELS - 1, sample_pixels).astype(np.int32)
    t0 = time.perf_counter()
    oracle.horizon_tile(pyramids, segs, tile_column=march_tile[0], tile_row=march_tile[1],
                        pixels=pixels, mode=oracle.FLOAT32 | oracle.REFERENCE_SLOPE, threads=cores)
    t_march = time.perf_counter() - t0
    per_tile = t_seg + t_march * PIXELS / sample_pixels
    from lunarscout_amd import synthetic

    sun = synthetic.sun_vectors(dems[0], 4380, 2.0, seed=5)
    rows = np.maximum(horizon_tile, np.float32(-50.0)).reshape(128, 128, AZIMUTHS)
    t0 = time.perf_counter()
    oracle.lighting_patch(dems[0], rows, sun, tile_y=tile[1], tile_x=tile[0],
                          float32_slices=True, threads=cores)
    t_light = time.perf_counter() - t0
    return {
        "value": round(1.0 / per_tile, 6),
        "unit": "tiles/s",
        "cores": cores,
        "kind": "port",
        "sample": f"tile {march_tile}: full FP64 segment generation ({t_seg:.2f}s) + float32 march "
                  f"of {sample_pixels} of 16384 px x 1440 az x 2 passes ({t_march:.2f}s), scaled "
                  f"to one tile; lighting: marched full-site tile {tile} x 4380 Sun times "
                  f"({t_light:.2f}s)",
        "lighting_px_times_per_s": round(PIXELS * len(sun) / t_light, 1),
    }


if __name__ == "__main__":
    sys.exit(main())

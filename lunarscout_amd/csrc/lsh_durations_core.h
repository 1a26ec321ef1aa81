// lsh_durations_core.h -- per-pixel reductions of the lighting time series,
// shared by the HIP kernel and the multi-threaded host path.
//
// Restates, for one pixel and without materialising the [time][pixel] signal
// stream (paths under src/lunarscout/_numba_horizon/):
//   * reduce_longest_candidate_duration_stream (mission_duration.py:178-256):
//     for every candidate-start interval k, the longest run of consecutive
//     samples satisfying the condition that starts at a true sample inside k
//     (a run already true when k opens starts at k's first sample) and goes on
//     while true, past k's end; the run still true at the last sample is
//     credited through the evaluation stop. The condition is
//     signal >= threshold, optionally AND Earth margin >= threshold
//     (mission_duration_pipeline.py:289-318), compared in float32 as the
//     reference compares its float32 tiles;
//   * reduce_safe_haven_patch_stream (safe_haven.py:79-243): for every month
//     band, the longest run of low-Sun samples (fraction < threshold) that
//     contains, inside that month, a sample with the Earth below its threshold
//     (margin < threshold); NaN where the month has no Earth outage sample or
//     no sample with the Earth above.
// Both walk the time series once per pixel and touch per-band state only when
// a run ends (longest durations) or the month changes (safe-haven flags).
// Durations are sums of the sample hours in time order from the candidate's
// first sample, as the reference accumulates them, so results are identical.
#pragma once

#include "lsh_lighting_core.h"

namespace lsh {

struct DurationArgs {
  LightArgs light;     // frame, horizons, sun (and Earth) vectors; sun_count = T
  int32_t reducer;     // LSH_REDUCE_LONGEST_DURATION or LSH_REDUCE_SAFE_HAVEN
  int32_t sun_signal;  // LSH_SIGNAL_SUN_FRACTION or LSH_SIGNAL_SUN_MARGIN
  int32_t use_earth;
  int32_t T;
  float sun_threshold, earth_threshold;
  // longest durations
  const double* sample_hours;  // [T]
  int32_t K;
  const int32_t* first;  // [K]
  const int32_t* end;    // [K]
  int32_t output_days;
  // safe haven
  const int32_t* month_of;  // [T]
  int32_t months_sorted;    // month_of is non-decreasing (calendar bands of sorted times)
  int32_t B;
  float step_hours;
  // workspace (per band, per tile, per pixel) and output [band][tile][16384]
  double* longest;   // [K][tiles][16384]
  int32_t* best;     // [B][tiles][16384]
  uint8_t* flags;    // [B][tiles][16384]: 1 Earth outage seen, 2 Earth above seen, 4 run touches
  float* out;
};

constexpr uint8_t kHadOutage = 1, kWasAbove = 2, kTouching = 4;

template <bool kAzimuthMajor>
struct PixelSignals {
  const LightArgs& a;
  Frame f;
  HorizonRow<kAzimuthMajor> hz;
  // per-sector decisions of the sun signal (fill_sectors): for the fraction the
  // disk bounds (0 below, 1 above); for the centre margin the bounds of its
  // sign against the threshold (-inf below, +inf above: the reducers only
  // compare the signal with the threshold). Off for short series.
  SectorTable sectors;
  bool decide;
  // per-sector decisions of the Earth margin's sign against its threshold (8
  // sectors of 45 deg: the workgroup's LDS holds both tables without costing
  // the kernel a resident workgroup)
  SectorTable earth_sectors;
  bool decide_earth;
  // sun fraction (BuilderSunFraction) or centre margin el - lerp(horizon), float32
  LSH_LHD float sun(int t, int signal) const {
    double e[3], horizontal;
    body_vector(f, a.sun + 3 * t, e, horizontal);
    if (decide) {
      const int j = sector_of(e[0], e[1]) * sectors.stride;
      const bool fraction = signal == LSH_SIGNAL_SUN_FRACTION;
      if (e[2] < horizontal * static_cast<double>(sectors.dark[j])) return fraction ? 0.0f : -INFINITY;
      if (e[2] > horizontal * static_cast<double>(sectors.lit[j])) return fraction ? 1.0f : INFINITY;
    }
    double az, el;
    azimuth_elevation_of(e, horizontal, az, el);
    if (signal == LSH_SIGNAL_SUN_FRACTION)
      return sun_fraction(hz, a.N, static_cast<float>(az), static_cast<float>(el), a);
    return static_cast<float>(el - horizon_at(hz, a.N, az));
  }
  LSH_LHD float earth(int t) const {
    double e[3], horizontal;
    body_vector(f, a.earth + 3 * t, e, horizontal);
    if (decide_earth) {
      const int j = (sector_of(e[0], e[1]) >> 1) * earth_sectors.stride;
      if (e[2] < horizontal * static_cast<double>(earth_sectors.dark[j])) return -INFINITY;
      if (e[2] > horizontal * static_cast<double>(earth_sectors.lit[j])) return INFINITY;
    }
    double az, el;
    azimuth_elevation_of(e, horizontal, az, el);
    return static_cast<float>(el - horizon_at(hz, a.N, az));
  }
};

// One run of true samples [s0, b) of the longest-duration condition ends.
LSH_LHD void close_candidate_run(const DurationArgs& a, int64_t slot, int64_t stride, int s0,
                                 int b) {
  for (int k = 0; k < a.K; ++k) {
    const int s = s0 > a.first[k] ? s0 : a.first[k];
    if (!(s < a.end[k] && s < b)) continue;  // no true sample inside interval k
    double d = 0.0;
    for (int t = s; t < b; ++t) d += a.sample_hours[t];
    double& l = a.longest[k * stride + slot];
    if (d > l) l = d;
  }
}

// A low-Sun run [s0, b) of length L ends: months it touched during an outage
// take L (safe_haven.py:196-215), and their touch marks are cleared.
LSH_LHD void credit_month(const DurationArgs& a, int64_t slot, int64_t stride, int m, int L) {
  uint8_t& fl = a.flags[m * stride + slot];
  if (fl & kTouching) {
    int32_t& bm = a.best[m * stride + slot];
    if (L > bm) bm = L;
    fl = static_cast<uint8_t>(fl & ~kTouching);
  }
}

LSH_LHD void close_low_sun_run(const DurationArgs& a, int64_t slot, int64_t stride, int s0, int b,
                               int L) {
  if (a.months_sorted) {  // the run's months are the range month_of[s0] .. month_of[b - 1]
    const int m0 = a.month_of[s0] > 0 ? a.month_of[s0] : 0;
    for (int m = m0; m <= a.month_of[b - 1]; ++m) credit_month(a, slot, stride, m, L);
    return;
  }
  int last = -2;
  for (int t = s0; t < b; ++t) {
    const int m = a.month_of[t];
    if (m < 0 || m == last) continue;
    last = m;
    credit_month(a, slot, stride, m, L);
  }
}

template <bool kAzimuthMajor>
LSH_LHD void duration_pixel(const DurationArgs& a, int tile, int pixel, lsh_tile origin,
                            SectorTable sectors, SectorTable earth_sectors) {
  const LightArgs& la = a.light;
  const int line = pixel / 128, sample = pixel % 128;
  const int row = origin.tile_row + line, column = origin.tile_column + sample;
  const int64_t stride = static_cast<int64_t>(la.tile_count) * kPatchPixels;
  const int64_t slot = static_cast<int64_t>(tile) * kPatchPixels + pixel;
  const int bands = a.reducer == LSH_REDUCE_SAFE_HAVEN ? a.B : a.K;
  if (!(row < la.dem_height && column < la.dem_width)) {  // edge padding
    for (int k = 0; k < bands; ++k) a.out[k * stride + slot] = NAN;
    return;
  }
  PixelSignals<kAzimuthMajor> sig{la, pixel_frame(la, row, column), {}, sectors, false,
                                  earth_sectors, false};
  sig.hz.N = la.N;
  sig.hz.base = kAzimuthMajor ? la.horizons + static_cast<int64_t>(tile) * kPatchPixels * la.N + pixel
                              : la.horizons + slot * la.N;
  // the safe-haven reducer reads the fraction whatever sun_signal says
  const int signal =
      a.reducer == LSH_REDUCE_SAFE_HAVEN ? LSH_SIGNAL_SUN_FRACTION : a.sun_signal;
  sig.decide = a.T >= kDecideMinTimes;
  if (sig.decide && signal == LSH_SIGNAL_SUN_FRACTION) {
    float c = 0.0f;  // the tallest slice half-height
    for (int i = 0; i < 16; ++i) c = la.half_circle[i] > c ? la.half_circle[i] : c;
    fill_sectors(sig.hz, la.N, -static_cast<double>(c) - 2e-3, static_cast<double>(c) + 2e-3,
                 true, sectors);
  } else if (sig.decide) {
    fill_sectors(sig.hz, la.N, static_cast<double>(a.sun_threshold) - 1e-4,
                 static_cast<double>(a.sun_threshold) + 1e-4, false, sectors);
  }
  // the safe haven reads the Earth margin at every time (the longest durations
  // only where the Sun holds, too seldom to repay the table)
  sig.decide_earth = sig.decide && a.reducer == LSH_REDUCE_SAFE_HAVEN;
  if (sig.decide_earth)
    fill_sectors<kSectors / 2>(sig.hz, la.N, static_cast<double>(a.earth_threshold) - 1e-4,
                               static_cast<double>(a.earth_threshold) + 1e-4, false,
                               earth_sectors);

  if (a.reducer == LSH_REDUCE_LONGEST_DURATION) {
    for (int k = 0; k < a.K; ++k) a.longest[k * stride + slot] = 0.0;
    int start = -1;
    for (int t = 0; t < a.T; ++t) {
      bool c = sig.sun(t, a.sun_signal) >= a.sun_threshold;
      if (c && a.use_earth) c = sig.earth(t) >= a.earth_threshold;
      if (c) {
        if (start < 0) start = t;
      } else if (start >= 0) {
        close_candidate_run(a, slot, stride, start, t);
        start = -1;
      }
    }
    if (start >= 0) close_candidate_run(a, slot, stride, start, a.T);  // right-censored
    for (int k = 0; k < a.K; ++k) {
      double v = a.longest[k * stride + slot];
      if (a.output_days) v /= 24.0;
      a.out[k * stride + slot] = static_cast<float>(v);
    }
    return;
  }

  // safe haven
  for (int m = 0; m < a.B; ++m) {
    a.best[m * stride + slot] = 0;
    a.flags[m * stride + slot] = 0;
  }
  int run = 0, run_start = -1;
  int cached = -1;  // the month whose flags are held in `cf`
  uint8_t cf = 0;
  for (int t = 0; t < a.T; ++t) {
    const bool low = sig.sun(t, LSH_SIGNAL_SUN_FRACTION) < a.sun_threshold;
    const bool below = sig.earth(t) < a.earth_threshold;
    const int m = a.month_of[t];
    if (m != cached) {
      if (cached >= 0) a.flags[cached * stride + slot] = cf;
      cached = m;
      cf = m >= 0 ? a.flags[m * stride + slot] : 0;
    }
    if (m >= 0) {
      cf = static_cast<uint8_t>(cf | (below ? kHadOutage : kWasAbove));
      if (below && low) cf = static_cast<uint8_t>(cf | kTouching);
    }
    if (low) {
      if (run == 0) run_start = t;
      ++run;
    } else if (run > 0) {
      if (cached >= 0) a.flags[cached * stride + slot] = cf;
      close_low_sun_run(a, slot, stride, run_start, t, run);
      if (cached >= 0) cf = a.flags[cached * stride + slot];
      run = 0;
    }
  }
  if (cached >= 0) a.flags[cached * stride + slot] = cf;
  if (run > 0) close_low_sun_run(a, slot, stride, run_start, a.T, run);  // right-censored
  for (int m = 0; m < a.B; ++m) {
    const uint8_t fl = a.flags[m * stride + slot];
    const bool posed = (fl & kHadOutage) && (fl & kWasAbove);
    a.out[m * stride + slot] =
        posed ? static_cast<float>(a.best[m * stride + slot]) * a.step_hours : NAN;
  }
}

}  // namespace lsh

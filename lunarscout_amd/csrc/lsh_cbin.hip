// lsh_cbin.hip -- the .cbin horizon tile codec, host side, multi-threaded.
//
// Format (file_format.py:39-73 encoder, 135-182 decoder; the C#
// HorizonCompressor): per pixel a little-endian u16 block length, then the
// block: the first sample as a big-endian int16 of round-half-away(clamp(v,
// +-50) * 32767/50), then each following sample as the wrapped int16 delta to
// the previous quantized value, clamped to [-16384, 16383]: one byte for
// 0 <= d < 63 (d) or -64 <= d < 0 (d & 0x7F), else two bytes
// ((d >> 8) | 0x80, d & 0xFF) whose 15-bit payload is sign-extended on decode.
// Decoded value = float32(q) * (50 / 32767).
//
// Pixels are independent once block boundaries are known, so encode splits the
// pixels into contiguous chunks encoded into per-chunk buffers and joined in
// order, and decode walks the u16 lengths once (a 16384-step scan) and then
// decodes chunks in parallel. Pure host code: no device is touched.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "lsh_internal.h"

namespace lsh {
namespace {

const float kShortScale = 32767.0f / 50.0f;      // _SHORT_SCALE (file_format.py:23)
const float kElevationScale = 50.0f / 32767.0f;  // _ELEVATION_SCALE (file_format.py:24)

inline int quantize(float raw) {
  double value = static_cast<double>(raw);
  value = value < -50.0 ? -50.0 : (value > 50.0 ? 50.0 : value);
  const double scaled = static_cast<double>(static_cast<float>(value) * kShortScale);
  double q = scaled >= 0.0 ? std::floor(scaled + 0.5) : std::ceil(scaled - 0.5);
  q = q < -32767.0 ? -32767.0 : (q > 32767.0 ? 32767.0 : q);
  return static_cast<int>(q);
}

// Encodes one pixel's n samples; returns the block length (<= 2n).
inline int encode_block(const float* h, int n, uint8_t* out) {
  int previous = quantize(h[0]);
  const uint16_t first = static_cast<uint16_t>(previous);
  out[0] = static_cast<uint8_t>(first >> 8);
  out[1] = static_cast<uint8_t>(first & 0xFF);
  int k = 2;
  for (int i = 1; i < n; ++i) {
    int delta = quantize(h[i]) - previous;
    delta = delta < -16384 ? -16384 : (delta > 16383 ? 16383 : delta);
    previous = static_cast<int16_t>(static_cast<uint16_t>(previous + delta));
    if (delta >= 0 && delta < 63) {
      out[k++] = static_cast<uint8_t>(delta);
    } else if (delta < 0 && delta >= -64) {
      out[k++] = static_cast<uint8_t>(delta & 0x7F);
    } else {
      const uint16_t d = static_cast<uint16_t>(delta);
      out[k++] = static_cast<uint8_t>((d >> 8) | 0x80);
      out[k++] = static_cast<uint8_t>(d & 0xFF);
    }
  }
  return k;
}

// Decodes one block of `length` bytes into n samples; false if malformed.
inline bool decode_block(const uint8_t* b, int length, int n, float* out) {
  int acc = static_cast<int16_t>(static_cast<uint16_t>((b[0] << 8) | b[1]));
  out[0] = static_cast<float>(acc) * kElevationScale;
  int k = 2, i = 1;
  while (k < length && i < n) {
    const int first = b[k++];
    int delta;
    if (first & 0x80) {
      if (k >= length) return false;
      delta = ((first & 0x7F) << 8) | b[k++];
      if (delta & 0x4000) delta -= 0x8000;
    } else {
      delta = first & 0x40 ? (first & 0x7F) - 0x80 : first;
    }
    acc = static_cast<int16_t>(static_cast<uint16_t>(acc + delta));
    out[i++] = static_cast<float>(acc) * kElevationScale;
  }
  return i == n && k == length;
}

int thread_count(int32_t requested, int64_t work_items) {
  int t = requested > 0 ? requested : static_cast<int>(std::thread::hardware_concurrency());
  t = std::max(1, std::min(t, 64));
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(t, work_items)));
}

template <class F>
void parallel_chunks(int threads, int64_t items, F&& body) {
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    const int64_t begin = items * t / threads, end = items * (t + 1) / threads;
    if (t == threads - 1) {
      body(t, begin, end);
    } else {
      pool.emplace_back([&body, t, begin, end] { body(t, begin, end); });
    }
  }
  for (auto& th : pool) th.join();
}

}  // namespace
}  // namespace lsh

using namespace lsh;

extern "C" {

int lsh_cbin_encode(const float* degrees, int32_t pixels, int32_t azimuths, uint8_t* payload,
                    int64_t capacity, int64_t* payload_bytes, int32_t threads) {
  LSH_REQUIRE(degrees && payload && payload_bytes, "invalid arguments");
  LSH_REQUIRE(pixels > 0 && azimuths > 0 && azimuths <= 32767, "invalid tile shape");
  const int64_t block_max = 2 * static_cast<int64_t>(azimuths);
  const int nt = thread_count(threads, pixels);
  std::vector<std::vector<uint8_t>> parts(nt);
  std::vector<int> failed(nt, 0);
  parallel_chunks(nt, pixels, [&](int t, int64_t begin, int64_t end) {
    std::vector<uint8_t>& part = parts[t];
    part.resize(static_cast<size_t>((end - begin) * (block_max + 2)));
    int64_t used = 0;
    for (int64_t p = begin; p < end; ++p) {
      const float* h = degrees + p * azimuths;
      for (int i = 0; i < azimuths; ++i)
        if (h[i] != h[i]) failed[t] = 1;  // NaN has no quantized value
      uint8_t* block = part.data() + used + 2;
      const int length = encode_block(h, azimuths, block);
      part[used] = static_cast<uint8_t>(length & 0xFF);
      part[used + 1] = static_cast<uint8_t>(length >> 8);
      used += 2 + length;
    }
    part.resize(static_cast<size_t>(used));
  });
  int64_t total = 0;
  for (int t = 0; t < nt; ++t) {
    LSH_REQUIRE(!failed[t], "degrees must not contain NaN values");
    total += static_cast<int64_t>(parts[t].size());
  }
  LSH_REQUIRE(total <= capacity, "payload capacity is too small");
  int64_t offset = 0;
  for (int t = 0; t < nt; ++t) {
    std::memcpy(payload + offset, parts[t].data(), parts[t].size());
    offset += static_cast<int64_t>(parts[t].size());
  }
  *payload_bytes = total;
  return LSH_OK;
}

int lsh_cbin_decode(const uint8_t* payload, int64_t bytes, int32_t pixels, int32_t azimuths,
                    float* degrees, int32_t threads) {
  LSH_REQUIRE(payload && degrees && bytes >= 0, "invalid arguments");
  LSH_REQUIRE(pixels > 0 && azimuths > 0 && azimuths <= 32767, "invalid tile shape");
  std::vector<int64_t> offsets(static_cast<size_t>(pixels));
  int64_t off = 0;
  for (int32_t p = 0; p < pixels; ++p) {
    LSH_REQUIRE(off + 2 <= bytes, "compressed horizon ended while reading a block length");
    const int length = payload[off] | (payload[off + 1] << 8);
    LSH_REQUIRE(length >= 2 && length <= 2 * azimuths, "compressed horizon block length is invalid");
    offsets[p] = off + 2;
    off += 2 + length;
    LSH_REQUIRE(off <= bytes, "compressed horizon ended inside a block");
  }
  LSH_REQUIRE(off == bytes, "compressed horizon contains trailing bytes");
  const int nt = thread_count(threads, pixels);
  std::vector<int> failed(nt, 0);
  parallel_chunks(nt, pixels, [&](int t, int64_t begin, int64_t end) {
    for (int64_t p = begin; p < end; ++p) {
      const uint8_t* block = payload + offsets[p];
      const int length = block[-2] | (block[-1] << 8);
      if (!decode_block(block, length, azimuths, degrees + p * azimuths)) failed[t] = 1;
    }
  });
  for (int t = 0; t < nt; ++t)
    LSH_REQUIRE(!failed[t], "compressed horizon block did not decode to the azimuth count");
  return LSH_OK;
}

int lsh_cbin_complete(const uint8_t* payload, int64_t bytes, int32_t pixels, int32_t azimuths,
                      int32_t* complete) {
  LSH_REQUIRE(payload && complete && bytes >= 0 && pixels > 0 && azimuths > 0,
                "invalid arguments");
  // HorizonTileStore.is_complete (file_format.py:335-356): every block length in
  // [1, 2N], every block present, nothing after the last block.
  int64_t off = 0;
  *complete = 0;
  for (int32_t p = 0; p < pixels; ++p) {
    if (off + 2 > bytes) return LSH_OK;
    const int length = payload[off] | (payload[off + 1] << 8);
    if (length < 1 || length > 2 * azimuths) return LSH_OK;
    off += 2 + length;
    if (off > bytes) return LSH_OK;
  }
  *complete = off == bytes ? 1 : 0;
  return LSH_OK;
}

}  // extern "C"

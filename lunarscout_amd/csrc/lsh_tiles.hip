// lsh_tiles.hip -- horizon tile files on the host: completeness, reading and a
// staged, fsynced publish, next to the .cbin codec they use.
//
// The file contract (file_format.py:18-30, 283-415; the C# HorizonTileStore):
//   * a raw tile (.bin) is pixels x azimuths little-endian float32 values,
//     optionally followed by 28 bytes of legacy metadata (seven floats) that
//     readers ignore; it is complete when its size is one of those two;
//   * a compressed tile (.cbin) is complete when every u16 block length is in
//     [1, 2 * azimuths], every block is present and nothing follows the last;
//   * a writer never exposes a partial file: the bytes go to a uniquely named
//     file in the destination's directory, are flushed to storage and then
//     renamed over the destination (a crash leaves at most a stray staging
//     file, never a truncated tile under the final name).
// The kind of a file is its extension (.bin / .cbin, any letter case).
// Pure host code: POSIX file I/O, no device is touched.

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lsh_internal.h"

namespace {

enum class Kind { kRaw, kCompressed, kOther };

Kind kind_of(const std::string& path) {
  const size_t dot = path.find_last_of('.');
  const size_t slash = path.find_last_of('/');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return Kind::kOther;
  std::string ext = path.substr(dot);
  for (char& c : ext) c = static_cast<char>(c >= 'A' && c <= 'Z' ? c - 'A' + 'a' : c);
  return ext == ".bin" ? Kind::kRaw : (ext == ".cbin" ? Kind::kCompressed : Kind::kOther);
}

// An open file descriptor closed on scope exit.
struct Fd {
  int fd = -1;
  explicit Fd(int f) : fd(f) {}
  ~Fd() {
    if (fd >= 0) ::close(fd);
  }
  Fd(const Fd&) = delete;
  Fd& operator=(const Fd&) = delete;
};

// Reads exactly `bytes` bytes from the start of the file into dst.
bool read_exact(int fd, void* dst, int64_t bytes) {
  uint8_t* p = static_cast<uint8_t*>(dst);
  int64_t done = 0;
  while (done < bytes) {
    const ssize_t got = ::pread(fd, p + done, static_cast<size_t>(bytes - done), done);
    if (got < 0 && errno == EINTR) continue;
    if (got <= 0) return false;
    done += got;
  }
  return true;
}

bool write_all(int fd, const void* src, int64_t bytes) {
  const uint8_t* p = static_cast<const uint8_t*>(src);
  int64_t done = 0;
  while (done < bytes) {
    const ssize_t put = ::write(fd, p + done, static_cast<size_t>(bytes - done));
    if (put < 0 && errno == EINTR) continue;
    if (put <= 0) return false;
    done += put;
  }
  return true;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int io_error(const std::string& what) {
  lsh::set_error(what + ": " + std::strerror(errno));
  return LSH_EIO;
}

}  // namespace

extern "C" {

int lsh_tile_file_status(const char* path, int32_t pixels, int32_t azimuths, int32_t* status) {
  LSH_REQUIRE(path && status && pixels > 0 && azimuths > 0, "invalid arguments");
  *status = LSH_TILE_ABSENT;
  const std::string name(path);
  struct stat st;
  if (::stat(path, &st) != 0 || !S_ISREG(st.st_mode)) return LSH_OK;
  *status = LSH_TILE_INCOMPLETE;
  const int64_t raw = 4ll * pixels * azimuths;
  const Kind kind = kind_of(name);
  if (kind == Kind::kRaw) {
    if (st.st_size == raw || st.st_size == raw + 28) *status = LSH_TILE_COMPLETE;
    return LSH_OK;
  }
  if (kind != Kind::kCompressed || st.st_size <= 0) return LSH_OK;
  Fd f(::open(path, O_RDONLY | O_CLOEXEC));
  if (f.fd < 0) return LSH_OK;
  std::vector<uint8_t> payload(static_cast<size_t>(st.st_size));
  if (!read_exact(f.fd, payload.data(), st.st_size)) return LSH_OK;
  int32_t complete = 0;
  if (int r = lsh_cbin_complete(payload.data(), st.st_size, pixels, azimuths, &complete)) return r;
  if (complete) *status = LSH_TILE_COMPLETE;
  return LSH_OK;
}

int lsh_tile_file_read(const char* path, int32_t pixels, int32_t azimuths, float* degrees,
                       int32_t threads, double* seconds) {
  LSH_REQUIRE(path && degrees && pixels > 0 && azimuths > 0, "invalid arguments");
  const std::string name(path);
  const Kind kind = kind_of(name);
  LSH_REQUIRE(kind != Kind::kOther, "horizon tile must have a .bin or .cbin extension");
  const auto t0 = std::chrono::steady_clock::now();
  Fd f(::open(path, O_RDONLY | O_CLOEXEC));
  struct stat st;
  if (f.fd < 0 || ::fstat(f.fd, &st) != 0) {
    lsh::set_error("unable to read horizon tile: " + name);
    return LSH_EINVAL;
  }
  const int64_t raw = 4ll * pixels * azimuths;
  double read_s = 0.0, decode_s = 0.0;
  if (kind == Kind::kRaw) {
    LSH_REQUIRE(st.st_size == raw || st.st_size == raw + 28,
                "uncompressed horizon file size is invalid");
    // little-endian float32 on this host: the file bytes are the values
    if (!read_exact(f.fd, degrees, raw)) {
      lsh::set_error("unable to read horizon tile: " + name);
      return LSH_EINVAL;
    }
    read_s = seconds_since(t0);
  } else {
    std::vector<uint8_t> payload(static_cast<size_t>(st.st_size));
    if (!read_exact(f.fd, payload.data(), st.st_size)) {
      lsh::set_error("unable to read horizon tile: " + name);
      return LSH_EINVAL;
    }
    read_s = seconds_since(t0);
    const auto t1 = std::chrono::steady_clock::now();
    if (int r = lsh_cbin_decode(payload.data(), st.st_size, pixels, azimuths, degrees, threads))
      return r;
    decode_s = seconds_since(t1);
  }
  if (seconds) {
    seconds[0] = read_s;
    seconds[1] = decode_s;
  }
  return LSH_OK;
}

int lsh_tile_file_write(const char* path, const float* degrees, int32_t pixels, int32_t azimuths,
                        int32_t compress, int32_t threads) {
  LSH_REQUIRE(path && degrees && pixels > 0 && azimuths > 0 && azimuths <= 32767,
              "invalid arguments");
  const std::string name(path);
  const int64_t count = static_cast<int64_t>(pixels) * azimuths;
  for (int64_t i = 0; i < count; ++i)
    LSH_REQUIRE(!std::isnan(degrees[i]), "degrees must not contain NaN values");
  std::vector<uint8_t> packed;
  const void* bytes = degrees;
  int64_t size = 4 * count;
  if (compress) {
    packed.resize(static_cast<size_t>(pixels) * (2 * static_cast<size_t>(azimuths) + 2));
    int64_t used = 0;
    if (int r = lsh_cbin_encode(degrees, pixels, azimuths, packed.data(),
                                static_cast<int64_t>(packed.size()), &used, threads))
      return r;
    bytes = packed.data();
    size = used;
  }
  // staging name: <final>.<6 random characters>.staging in the same directory
  std::string staging = name + ".XXXXXX.staging";
  std::vector<char> buf(staging.begin(), staging.end());
  buf.push_back('\0');
  const int fd = ::mkstemps(buf.data(), 8);
  if (fd < 0) return io_error("unable to create a staging file beside " + name);
  staging.assign(buf.data());
  bool published = false;
  {
    Fd f(fd);
    if (write_all(f.fd, bytes, size) && ::fsync(f.fd) == 0) {
      f.fd = -1;
      if (::close(fd) == 0 && ::rename(staging.c_str(), path) == 0) published = true;
    }
  }
  if (!published) {
    const int saved = errno;
    ::unlink(staging.c_str());
    errno = saved;
    return io_error("unable to write horizon tile " + name);
  }
  return LSH_OK;
}

}  // extern "C"

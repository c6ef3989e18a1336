// .bin vector files: uint8 iscomplex, uint32 n, n float64 real parts,
// [n float64 imaginary parts] (matlab/writebin_vector.m:35-44,
// readbin_vector.m:30-40).  Host memory only.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/avionic.h"

void avionic_set_error(const char *msg);

namespace {

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

int fail(const std::string &msg) {
  avionic_set_error(msg.c_str());
  return AVIONIC_E_INVALID;
}

}  // namespace

extern "C" int avionic_bin_read(const char *path, double *re, double *im, size_t capacity,
                                size_t *n, int *iscomplex) {
  if (!path || !n) return fail("avionic_bin_read: null argument");
  File fh;
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return fail(std::string("cannot open ") + path);
  uint8_t cplx = 0;
  uint32_t len = 0;
  if (std::fread(&cplx, 1, 1, fh.f) != 1 || std::fread(&len, 4, 1, fh.f) != 1)
    return fail(std::string("truncated header in ") + path);
  *n = len;
  if (iscomplex) *iscomplex = (cplx & 1) ? 1 : 0;
  // the payload must be exactly n (or 2n) doubles
  const long start = std::ftell(fh.f);
  std::fseek(fh.f, 0, SEEK_END);
  const long end = std::ftell(fh.f);
  const long want = (long)len * 8 * ((cplx & 1) ? 2 : 1);
  if (end - start != want)
    return fail(std::string("size mismatch in ") + path + ": header says " +
                std::to_string(len) + ((cplx & 1) ? " complex" : " real") + " values, file holds " +
                std::to_string(end - start) + " payload bytes");
  if (!re) return AVIONIC_OK;
  std::fseek(fh.f, start, SEEK_SET);
  const size_t m = std::min<size_t>(capacity, len);
  if (m == 0) return AVIONIC_OK;
  if (std::fread(re, 8, m, fh.f) != m) return fail(std::string("read error in ") + path);
  if (im) {
    if (cplx & 1) {
      std::fseek(fh.f, start + (long)len * 8, SEEK_SET);
      if (std::fread(im, 8, m, fh.f) != m) return fail(std::string("read error in ") + path);
    } else {
      std::memset(im, 0, m * sizeof(double));
    }
  }
  return AVIONIC_OK;
}

extern "C" int avionic_bin_write(const char *path, const double *re, const double *im, size_t n) {
  if (!path || (!re && n)) return fail("avionic_bin_write: null argument");
  if (n > 0xffffffffu) return fail("avionic_bin_write: more than 2^32-1 values");
  File fh;
  fh.f = std::fopen(path, "wb");
  if (!fh.f) return fail(std::string("cannot create ") + path);
  const uint8_t cplx = im ? 1 : 0;
  const uint32_t len = (uint32_t)n;
  // (no payload call for n = 0: re / im may then be null, which fwrite's
  // pointer argument must never be)
  bool ok = std::fwrite(&cplx, 1, 1, fh.f) == 1 && std::fwrite(&len, 4, 1, fh.f) == 1 &&
            (n == 0 || (std::fwrite(re, 8, n, fh.f) == n && (!im || std::fwrite(im, 8, n, fh.f) == n)));
  if (!ok) return fail(std::string("write error in ") + path);
  return AVIONIC_OK;
}

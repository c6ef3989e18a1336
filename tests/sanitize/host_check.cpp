// AddressSanitizer / UndefinedBehaviorSanitizer run of the library's host-only
// C-ABI code (.bin reader / writer, the normalisation median), built from the
// library's own sources with g++ by tests/test_sanitizers.py.  Exits 0 when
// every case behaves; a sanitizer report aborts with a non-zero status.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/avionic.h"

// The library keeps its last error in capi.cpp; the host-only build stores it here.
static std::string g_err;
void avionic_set_error(const char *msg) { g_err = msg ? msg : ""; }

#define CHECK(cond, what)                                  \
  do {                                                     \
    if (!(cond)) {                                         \
      std::fprintf(stderr, "host_check: %s failed\n", what); \
      return 1;                                            \
    }                                                      \
  } while (0)

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  // .bin round trips: complex and real vectors, and a capacity below n
  for (size_t n : {1u, 7u, 1000u}) {
    std::vector<double> re(n), im(n), r2(n), i2(n);
    for (size_t i = 0; i < n; ++i) {
      re[i] = 0.5 * i - 3.0;
      im[i] = -1.0 / (i + 1.0);
    }
    const std::string path = dir + "/v" + std::to_string(n) + ".bin";
    CHECK(avionic_bin_write(path.c_str(), re.data(), im.data(), n) == 0, "complex write");
    size_t got = 0;
    int cplx = -1;
    CHECK(avionic_bin_read(path.c_str(), nullptr, nullptr, 0, &got, &cplx) == 0 && got == n &&
              cplx == 1,
          "complex header");
    CHECK(avionic_bin_read(path.c_str(), r2.data(), i2.data(), n, &got, &cplx) == 0,
          "complex read");
    for (size_t i = 0; i < n; ++i) CHECK(r2[i] == re[i] && i2[i] == im[i], "complex values");
    std::vector<double> part(n / 2 + 1), parti(n / 2 + 1);
    CHECK(avionic_bin_read(path.c_str(), part.data(), parti.data(), part.size(), &got, &cplx) == 0 &&
              part.back() == re[part.size() - 1] && parti.back() == im[part.size() - 1],
          "partial read");
    CHECK(avionic_bin_write(path.c_str(), re.data(), nullptr, n) == 0, "real write");
    CHECK(avionic_bin_read(path.c_str(), r2.data(), i2.data(), n, &got, &cplx) == 0 && cplx == 0 &&
              got == n && i2[n - 1] == 0.0,
          "real read");
  }
  {
    const std::string path = dir + "/empty.bin";
    CHECK(avionic_bin_write(path.c_str(), nullptr, nullptr, 0) == 0, "empty write");
    size_t got = 1;
    int cplx = -1;
    double dummy = 0.0;
    CHECK(avionic_bin_read(path.c_str(), &dummy, nullptr, 1, &got, &cplx) == 0 && got == 0,
          "empty read");
  }
  // truncated and missing files are refused
  {
    const std::string path = dir + "/trunc.bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    const unsigned char c = 1;
    const unsigned n = 100;
    std::fwrite(&c, 1, 1, f);
    std::fwrite(&n, 4, 1, f);
    const double v = 1.0;
    std::fwrite(&v, 8, 1, f);
    std::fclose(f);
    std::vector<double> re(100), im(100);
    size_t got = 0;
    int cplx = 0;
    CHECK(avionic_bin_read(path.c_str(), re.data(), im.data(), 100, &got, &cplx) != 0, "truncated");
    CHECK(avionic_bin_read((dir + "/none.bin").c_str(), re.data(), im.data(), 100, &got, &cplx) != 0,
          "missing");
  }
  // normalisation median: ordinary, all-equal, single value, and refusals
  {
    std::vector<float> m(257);
    for (size_t i = 0; i < m.size(); ++i) m[i] = (float)((i * 37) % 101);
    float med = 0.f;
    CHECK(avionic_normalization_factor(m.data(), m.size(), &med) == 0 && std::isfinite(med),
          "median");
    std::vector<float> e(5, 2.5f);
    CHECK(avionic_normalization_factor(e.data(), e.size(), &med) == 0 && med == 2.5f, "equal");
    CHECK(avionic_normalization_factor(e.data(), 1, &med) == 0 && med == 2.5f, "single");
    CHECK(avionic_normalization_factor(e.data(), 0, &med) != 0, "empty refused");
    CHECK(avionic_normalization_factor(nullptr, 3, &med) != 0, "null refused");
    e[2] = NAN;
    CHECK(avionic_normalization_factor(e.data(), e.size(), &med) != 0 || std::isfinite(med),
          "non-finite");
  }
  std::printf("host_check: ok\n");
  return 0;
}

// Raw-data normalisation factor (RawDataPreparation::FindNormalizationFactor,
// raw_data_preparation.cpp:661-686, with utils::Median, utils.cpp:1442-1462).
// Host memory only.
//
// The reference walks the magnitudes once: it starts at the FIRST element
// equal to the maximum and keeps that element and every later element
// >= 0.9 * max.  Elements above the threshold that come before the first
// maximum are not kept; this function keeps that behaviour, since it decides
// which values the median is taken over.  The threshold is the double product
// 0.9 * max rounded to float, as the reference's comparator stores it
// (isLargerThan holds a `const RType &` with RType = float,
// raw_data_preparation.cpp:649-658), and the comparison is in float; the
// median of an even count is
// the mean of the two middle values (their float sum halved).  A median that
// is not positive (all magnitudes zero) is an error here; the reference would
// go on with an infinite factor.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/avionic.h"

void avionic_set_error(const char *msg);

extern "C" int avionic_normalization_factor(const float *magnitude, size_t n, float *median) {
  if (!median || (!magnitude && n)) {
    avionic_set_error("avionic_normalization_factor: null argument");
    return AVIONIC_E_INVALID;
  }
  if (n == 0) {
    avionic_set_error("avionic_normalization_factor: no values");
    return AVIONIC_E_INVALID;
  }
  size_t imax = 0;
  for (size_t i = 1; i < n; ++i)
    if (magnitude[i] > magnitude[imax]) imax = i;
  const float mx = magnitude[imax];
  if (!std::isfinite(mx)) {
    avionic_set_error("avionic_normalization_factor: non-finite maximum");
    return AVIONIC_E_INVALID;
  }
  const float thr = (float)((double)mx * 0.9);
  std::vector<float> keep;
  keep.push_back(mx);
  for (size_t i = imax + 1; i < n; ++i)
    if (magnitude[i] >= thr) keep.push_back(magnitude[i]);
  const size_t mid = keep.size() / 2;
  std::nth_element(keep.begin(), keep.begin() + mid, keep.end());
  const float hi = keep[mid];
  if (keep.size() % 2 == 0) {
    const float lo = *std::max_element(keep.begin(), keep.begin() + mid);
    const float sum = hi + lo;  // float sum, then 0.5 * it in double
    *median = (float)(0.5 * (double)sum);
  } else {
    *median = hi;
  }
  if (!(*median > 0.f)) {
    avionic_set_error(
        "avionic_normalization_factor: no valid normalisation factor (the median magnitude is "
        "not positive)");
    return AVIONIC_E_INVALID;
  }
  return AVIONIC_OK;
}

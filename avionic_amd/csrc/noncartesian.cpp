// Placeholder until the Kaiser-Bessel gridding operator lands.
#include "solver.hpp"

namespace avionic {
Operator *make_noncartesian(int, int, int, int, int, int, const float *, const float *, int,
                            double) {
  throw Error(AVIONIC_E_UNSUPPORTED, "non-Cartesian operator not built yet");
}
}  // namespace avionic

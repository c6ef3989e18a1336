// K (BackwardOperation) of the Cartesian operator without any allocation, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  The operator's
// column pass writes its compact k-space columns to a scratch buffer; here
// that buffer is a file-scope device array, so a call launches kernels only.
//
// The array is per GPU: every device the library runs on holds its own copy
// for the life of the process.  Host code must not pass the array's name to a
// kernel (on the host it names a stand-in, not device memory); its device
// address is looked up once per device with hipGetSymbolAddress and that
// pointer is what the kernels receive.  Every call on one device shares the
// array, so calls that overlap on one GPU (two streams, or two host threads)
// race on it; calls on one stream are ordered and safe.
#include <array>
#include <mutex>

#include "../../include/avionic.h"
#include "solver.hpp"

namespace avionic {

// Compact k-space elements (complex float) the scratch holds: 64 MiB.
constexpr size_t kNoAllocScratchElems = (size_t(64) << 20) / sizeof(float2);

static __device__ __align__(256) float2 scratch[kNoAllocScratchElems];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<float2 *, kMaxDevices> g_ptr{};

// Device address of `scratch` on the current device (looked up on the first
// call from that device).
float2 *scratch_ptr() {
  int dev = 0;
  AV_HIP(hipGetDevice(&dev));
  AV_CHECK(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    AV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    AV_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_ptr[dev] = static_cast<float2 *>(p);
  }
  return g_ptr[dev];
}

// Zero fill of the dense output (a kernel, not hipMemsetAsync, so that a
// captured graph replays it correctly).
__global__ void k_zero_fill(float2 *out, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = make_float2(0.f, 0.f);
}

}  // namespace

size_t noalloc_scratch_elems() { return kNoAllocScratchElems; }

void cartesian_backward_noalloc(const CartesianOperator &op, const float2 *img, float2 *kdense,
                                const float2 *b1, hipStream_t s) {
  const CartGeom &g = op.plan().g;
  const long K = op.compact_elems();
  AV_CHECK((size_t)K <= kNoAllocScratchElems,
           "the operator's compact k-space (" + std::to_string(K) +
               " complex values) exceeds the no-allocation scratch of " +
               std::to_string(kNoAllocScratchElems) + " (64 MiB)");
  float2 *kcol = scratch_ptr();
  const long n = op.dense_elems();
  const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_zero_fill, dim3(std::max(1u, blocks)), dim3(256), 0, s, kdense, n);
  AV_LAUNCH_CHECK();
  launch_col_forward(g, img, nullptr, b1, kcol, s);
  launch_row_pass(kRowApplyK, g, kcol, nullptr, nullptr, nullptr, kdense, nullptr, nullptr,
                  nullptr, s);
}

}  // namespace avionic

// shaw_rom_project_noalloc: the Galerkin projection of shaw_rom_project on
// device pointers, for callers that cannot allocate device memory (code under
// HIP graph capture, serving loops that forbid hipMalloc). It runs the same
// kernels (projectTN, proj.hip) with the same chunking, so its operators equal
// shaw_rom_project's bit for bit. The split-K partial tiles live in the
// file-scope array `scratch` below: every GPU the library runs on holds its
// own copy (16 MiB) until the process ends, and every call on one GPU shares
// it, so calls that overlap on one GPU race. Each partial entry a call reads
// was written earlier in the same call, so nothing needs clearing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "rom_internal.hpp"
#include "shaw_common.hpp"

namespace {
// partial tiles of the projection: kv * ks * chunks doubles at most
constexpr uint64_t kScratchDoubles = (16ull << 20) / sizeof(double);
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace shaw {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
double* g_scratchPtr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device with that device current. In host code the array's name stands for
// a host-side shadow, never for device memory, so only this pointer is passed
// to kernels.
double* scratchOnCurrentDevice() {
  int dev = 0;
  SHAW_HIP(hipGetDevice(&dev));
  SHAW_REQUIRE(dev >= 0 && dev < kMaxDevices, SHAW_ERR_HIP, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    SHAW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SHAW_REQUIRE(p != nullptr, SHAW_ERR_HIP, "hipGetSymbolAddress returned no address for the projection scratch");
    g_scratchPtr[dev] = static_cast<double*>(p);
  }
  return g_scratchPtr[dev];
}
}  // namespace
}  // namespace shaw

extern "C" int shaw_rom_project_noalloc(uint64_t nr, uint64_t nth, const double* coefVp, const double* coefSp,
                                        const double* phiV, uint64_t kv, const double* phiS, uint64_t ks, double* arV,
                                        double* arS, void* stream) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(coefVp && coefSp && phiV && phiS && arV && arS, SHAW_ERR_ARG, "null argument");
    SHAW_REQUIRE(nr >= 2 && nth >= 2 && nr < (1ull << 31) && nth < (1ull << 31), SHAW_ERR_ARG, "bad mesh size");
    SHAW_REQUIRE(kv > 0 && ks > 0 && kv < (1ull << 31) && ks < (1ull << 31), SHAW_ERR_ARG, "bad number of modes");
    const uint64_t nVp = nr * nth, nSp = (nr - 1) * nth + nr * (nth - 1);
    const uint64_t need = std::max(projPartDoubles(nVp, kv, ks), projPartDoubles(nSp, ks, kv));
    SHAW_REQUIRE(need <= kScratchDoubles, SHAW_ERR_ARG,
                 "shaw_rom_project_noalloc: " + std::to_string(kv) + " x " + std::to_string(ks) +
                     " modes need more than its 16 MiB scratch (kv * ks must be at most 2097152)");
    double* part = scratchOnCurrentDevice();
    hipStream_t s = static_cast<hipStream_t>(stream);
    projectTN(0, (uint32_t)nr, (uint32_t)nth, nVp, nSp, coefVp, phiV, (uint32_t)kv, nullptr, phiS, (uint32_t)ks, part,
              arV, s);
    projectTN(1, (uint32_t)nr, (uint32_t)nth, nSp, nVp, coefSp, phiS, (uint32_t)ks, nullptr, phiV, (uint32_t)kv, part,
              arS, s);
  });
}

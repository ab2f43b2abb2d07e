// Error plumbing shared by the C ABI entry points: a thread-local message,
// HIP status checks that throw, and a wrapper that turns exceptions into
// status codes so nothing propagates across the extern "C" boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

#include "shaw_amd.h"

namespace shaw {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void setLastError(const std::string& msg);
void clearLastError();

#define SHAW_HIP(call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::shaw::Error(e_ == hipErrorOutOfMemory ? SHAW_ERR_MEMORY : SHAW_ERR_HIP,       \
                          std::string(#call) + ": " + hipGetErrorString(e_));               \
  } while (0)

#define SHAW_REQUIRE(cond, code, msg)             \
  do {                                            \
    if (!(cond)) throw ::shaw::Error(code, msg);  \
  } while (0)

template <class F>
int guarded(F&& f) {
  try {
    clearLastError();
    f();
    return SHAW_OK;
  } catch (const Error& e) {
    setLastError(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    setLastError("host allocation failed");
    return SHAW_ERR_MEMORY;
  } catch (const std::exception& e) {
    setLastError(e.what());
    return SHAW_ERR_ARG;
  } catch (...) {
    setLastError("unknown error");
    return SHAW_ERR_ARG;
  }
}

// XCD-aware block order: hardware hands block b to XCD b % 8; return a
// logical index such that each XCD receives one contiguous range of logical
// blocks (so neighbouring tiles share that XCD's L2).
__device__ __forceinline__ unsigned xcdLogicalBlock(unsigned b, unsigned G) {
  constexpr unsigned kXcd = 8;
  const unsigned xcd = b % kXcd, idx = b / kXcd;
  const unsigned q = G / kXcd, rem = G % kXcd;
  return xcd < rem ? xcd * (q + 1) + idx : rem * (q + 1) + (xcd - rem) * q + idx;
}

}  // namespace shaw

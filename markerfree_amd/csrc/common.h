// Shared definitions of the MI355X alignment library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "markerfree/mf_api.h"

namespace mf {

// Error carried to the C-ABI boundary, where it becomes an mf_status.
struct Error : public std::runtime_error {
  mf_status code;
  Error(mf_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file,
                            int line);

#define MF_HIP(call)                                              \
  do {                                                            \
    hipError_t _e = (call);                                       \
    if (_e != hipSuccess) ::mf::throw_hip(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define MF_REQUIRE(cond, msg)                                     \
  do {                                                            \
    if (!(cond)) throw ::mf::Error(MF_EINVAL, (msg));             \
  } while (0)

// D2R as the reference defines it (Util.h:33).
constexpr double kD2R = 3.14159265358979323846 / 180.0;
// Backprojection's angle factor (CalcReproj.cu:12).
constexpr float kBPD2R = 0.01745f;
constexpr float kInvalid = -1e30f;

__host__ __device__ inline int pad_x(int nx) { return (nx / 2 + 1) * 2; }
__host__ __device__ inline int cdiv(long long a, long long b) {
  return (int)((a + b - 1) / b);
}

// Float cos/sin of a float argument evaluated in double on the host; every
// index-deciding trigonometric value reaches the kernels through these.
inline float hcos(float a) { return (float)std::cos((double)a); }
inline float hsin(float a) { return (float)std::sin((double)a); }

}  // namespace mf

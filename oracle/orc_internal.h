// Internal helpers of the CPU oracle (test infrastructure only).
#pragma once
#include <cmath>
#include <complex>
#include <vector>
#include <cstdint>
#include <cstring>

namespace orc {

// D2R as the reference defines it (Util.h:33).
static const double kD2R = 4.0 * std::atan(1.0) / 180.0;
// Backprojection's angle factor (CalcReproj.cu:12, quirk Q12).
static const float kBPD2R = 0.01745f;

inline float invalid() { return -1e30f; }

// Padded row pitch of an in-place R2C image (CalcCC.cu:215).
inline int pad_x(int nx) { return (nx / 2 + 1) * 2; }

// Float cos/sin of a float argument, evaluated in double and rounded.  The
// HIP path takes the same host tables.
inline float fcos(float a) { return (float)std::cos((double)a); }
inline float fsin(float a) { return (float)std::sin((double)a); }

// GetFrameIdxFromTilt (Util.cpp:8-20).
int zero_tilt_index(const float* tilts, int nz, float tilt = 0.0f);
// FindRefIndex (Util.cpp:22-35).
int find_ref_index(const float* tilts, int nz, int z);

// Stretch matrix of CalcCC::Stretch (CalcCC.cu:318-345): inverse of
// R diag(s,1) R^T computed in double, cast to float.
void stretch_matrix(float stretch, float axis_deg, float m[3]);

// FFT (double precision internally) -------------------------------------
typedef std::complex<double> cd;
void fft_c2c(cd* data, int n, int sign);  // sign -1 forward, +1 inverse
void fft_r2c_line(const float* in, int n, float* out_cmp);  // out n/2+1 cplx
void fft_c2r_line(const float* in_cmp, int n, float* out);

}  // namespace orc

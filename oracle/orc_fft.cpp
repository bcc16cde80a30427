// Double-precision DFT for the CPU oracle (test infrastructure).
//
// The reference calls cuFFT (third-party, closed).  The oracle restates the
// transform from its definition X_k = sum_n x_n exp(-2 pi i n k / N): mixed
// radix decimation in time for factors up to 61 and Bluestein's chirp-z for
// larger primes.  Pinned in tests against numpy.fft and the direct sum.
// cuFFT's R2C/C2R conventions (unnormalised, n/2+1 outputs, in-place padded
// rows) follow method/Cufft1&2.cu:17-76 and :109-165.
#include <map>
#include <mutex>

#include "oracle.h"
#include "orc_internal.h"

namespace orc {

namespace {

int smallest_factor(int n) {
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0) return p;
  return n;
}

void fft_rec(cd* a, int n, int sign, cd* scratch);

void bluestein(cd* a, int n, int sign) {
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  std::vector<cd> w(n), A(m, 0.0), B(m, 0.0), s(m);
  for (int k = 0; k < n; k++) {
    long long q = ((long long)k * k) % (2LL * n);
    w[k] = std::polar(1.0, sign * M_PI * (double)q / n);
  }
  for (int k = 0; k < n; k++) A[k] = a[k] * w[k];
  for (int k = 0; k < n; k++) {
    B[k] = std::conj(w[k]);
    if (k) B[m - k] = std::conj(w[k]);
  }
  fft_rec(A.data(), m, -1, s.data());
  fft_rec(B.data(), m, -1, s.data());
  for (int k = 0; k < m; k++) A[k] *= B[k];
  fft_rec(A.data(), m, +1, s.data());
  for (int k = 0; k < n; k++) a[k] = A[k] * w[k] / (double)m;
}

void fft_rec(cd* a, int n, int sign, cd* scratch) {
  if (n == 1) return;
  int p = smallest_factor(n);
  if (p > 61) {
    bluestein(a, n, sign);
    return;
  }
  const int m = n / p;
  // Split into p decimated subsequences (scratch), transform each in place.
  for (int r = 0; r < p; r++)
    for (int j = 0; j < m; j++) scratch[r * m + j] = a[j * p + r];
  std::vector<cd> sub(scratch, scratch + n);
  std::vector<cd> tmp(n);
  for (int r = 0; r < p; r++) fft_rec(sub.data() + r * m, m, sign, tmp.data());
  const double ang = sign * 2.0 * M_PI / n;
  for (int k = 0; k < n; k++) {
    cd acc = 0.0;
    const int km = k % m;
    for (int r = 0; r < p; r++)
      acc += sub[r * m + km] *
             std::polar(1.0, ang * (double)(((long long)r * k) % n));
    a[k] = acc;
  }
}

}  // namespace

void fft_c2c(cd* data, int n, int sign) {
  std::vector<cd> scratch(n);
  fft_rec(data, n, sign, scratch.data());
}

void fft_r2c_line(const float* in, int n, float* out_cmp) {
  std::vector<cd> a(n);
  for (int i = 0; i < n; i++) a[i] = in[i];
  fft_c2c(a.data(), n, -1);
  for (int k = 0; k <= n / 2; k++) {
    out_cmp[2 * k] = (float)a[k].real();
    out_cmp[2 * k + 1] = (float)a[k].imag();
  }
}

void fft_c2r_line(const float* in_cmp, int n, float* out) {
  // Hermitian extension of the n/2+1 stored values; imaginary parts of the
  // DC and Nyquist terms are ignored, as a real-output transform must.
  std::vector<cd> a(n);
  for (int k = 0; k <= n / 2; k++) a[k] = cd(in_cmp[2 * k], in_cmp[2 * k + 1]);
  a[0] = cd(a[0].real(), 0.0);
  if (n % 2 == 0) a[n / 2] = cd(a[n / 2].real(), 0.0);
  for (int k = n / 2 + 1; k < n; k++) a[k] = std::conj(a[n - k]);
  fft_c2c(a.data(), n, +1);
  for (int i = 0; i < n; i++) out[i] = (float)a[i].real();
}

}  // namespace orc

using namespace orc;

extern "C" void orc_fft_r2c_2d(float* buf, int nx, int ny) {
  const int pitch = pad_x(nx), nc = nx / 2 + 1;
  std::vector<cd> spec((size_t)nc * ny);
#pragma omp parallel for
  for (int y = 0; y < ny; y++) {
    std::vector<cd> a(nx);
    for (int x = 0; x < nx; x++) a[x] = buf[(size_t)y * pitch + x];
    fft_c2c(a.data(), nx, -1);
    for (int k = 0; k < nc; k++) spec[(size_t)y * nc + k] = a[k];
  }
#pragma omp parallel for
  for (int k = 0; k < nc; k++) {
    std::vector<cd> a(ny);
    for (int y = 0; y < ny; y++) a[y] = spec[(size_t)y * nc + k];
    fft_c2c(a.data(), ny, -1);
    for (int y = 0; y < ny; y++) {
      buf[(size_t)y * pitch + 2 * k] = (float)a[y].real();
      buf[(size_t)y * pitch + 2 * k + 1] = (float)a[y].imag();
    }
  }
}

extern "C" void orc_fft_c2r_2d(float* buf, int nx, int ny) {
  const int pitch = pad_x(nx), nc = nx / 2 + 1;
  std::vector<cd> spec((size_t)nc * ny);
#pragma omp parallel for
  for (int k = 0; k < nc; k++) {
    std::vector<cd> a(ny);
    for (int y = 0; y < ny; y++)
      a[y] = cd(buf[(size_t)y * pitch + 2 * k], buf[(size_t)y * pitch + 2 * k + 1]);
    fft_c2c(a.data(), ny, +1);
    for (int y = 0; y < ny; y++) spec[(size_t)y * nc + k] = a[y];
  }
#pragma omp parallel for
  for (int y = 0; y < ny; y++) {
    std::vector<float> c(2 * nc);
    for (int k = 0; k < nc; k++) {
      c[2 * k] = (float)spec[(size_t)y * nc + k].real();
      c[2 * k + 1] = (float)spec[(size_t)y * nc + k].imag();
    }
    // Keep the column-transformed values in double: redo the row in double.
    std::vector<cd> a(nx);
    for (int k = 0; k < nc; k++) a[k] = spec[(size_t)y * nc + k];
    a[0] = cd(a[0].real(), 0.0);
    a[nx / 2] = cd(a[nx / 2].real(), 0.0);
    for (int k = nc; k < nx; k++) a[k] = std::conj(a[nx - k]);
    fft_c2c(a.data(), nx, +1);
    for (int x = 0; x < nx; x++) buf[(size_t)y * pitch + x] = (float)a[x].real();
  }
}

extern "C" void orc_fft_r2c_1d(float* buf, int n, int nlines) {
  const int pitch = pad_x(n);
#pragma omp parallel for
  for (int l = 0; l < nlines; l++) {
    std::vector<float> in(buf + (size_t)l * pitch, buf + (size_t)l * pitch + n);
    fft_r2c_line(in.data(), n, buf + (size_t)l * pitch);
  }
}

extern "C" void orc_fft_c2r_1d(float* buf, int n, int nlines) {
  const int pitch = pad_x(n);
#pragma omp parallel for
  for (int l = 0; l < nlines; l++) {
    std::vector<float> in(buf + (size_t)l * pitch,
                          buf + (size_t)l * pitch + 2 * (n / 2 + 1));
    fft_c2r_line(in.data(), n, buf + (size_t)l * pitch);
  }
}

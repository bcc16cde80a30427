// Seeded synthetic tilt series with known alignment (host C++, OpenMP).
//
// Phantom: isotropic 3-D Gaussian blobs.  The line integral of a blob of
// amplitude A and width sigma is a 2-D Gaussian of amplitude A*sigma*sqrt(2pi),
// so projections are exact (no voxel projector).  The aligned projection at
// tilt t places a blob centred at (x0, y0, z0) at detector position
// u0 = x0 cos t + z0 sin t, v0 = y0 -- the geometry the reference's
// backprojector assumes (method/CalcReproj.cu:37-50).  The raw image is the
// exact inverse of the reference's correction mGCorrect
// (method/CorrStack.cu:142-179): raw(q) = P(R(axis)^-1 (q - shift - c)).
//
// Random numbers: PCG32 and Box-Muller written here, so every platform draws
// the same series.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {
typedef struct {
  int nx, ny, nz;           // image width, height, number of tilts
  int vol_x, vol_y, vol_z;  // phantom extent (voxels)
  int nblobs;
  float sigma_min, sigma_max, amp_min, amp_max;
  float tilt_min, tilt_step;  // tilts = tilt_min + i*tilt_step (ascending)
  float axis_min, axis_max;   // axis drawn U(axis_min, axis_max) (degrees)
  float shift_max;            // shifts U(-shift_max, shift_max) px per axis
  float mag_jitter;           // per-tilt isotropic magnification 1+U(-j,j)
  int noise_kind;             // 0 none, 1 Gaussian at 'snr', 2 Poisson
  float snr;                  // var(signal)/var(noise) for Gaussian noise
  float dose;                 // electrons per pixel per tilt (Poisson)
  int int16_counts;           // 1: quantise to int16 counts (MRC mode 1)
  uint64_t seed_phantom, seed_geom, seed_noise;
} mfs_config;

int mfs_generate(const mfs_config* cfg, float* stack, float* tilts,
                 float* shift_x, float* shift_y, float* axis_out,
                 float* scale_out);
// The raw-image positions of the first 'count' blob centres in every
// section (pixel x at position x, as the images are drawn), NaN where the
// centre falls outside the image: raw_xy [count][nz][2].  The same seeds
// give the same blobs and geometry as mfs_generate.  Returns the number of
// blobs written.
int mfs_fiducials(const mfs_config* cfg, int count, float* raw_xy);
}

namespace {

struct Pcg32 {
  uint64_t state, inc;
  explicit Pcg32(uint64_t seed, uint64_t seq = 54u) {
    state = 0u;
    inc = (seq << 1u) | 1u;
    next();
    state += seed;
    next();
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
  }
  double uniform() { return (next() + 0.5) / 4294967296.0; }  // (0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  // Poisson by inversion for small means, normal approximation above 60.
  int poisson(double mu) {
    if (mu <= 0) return 0;
    if (mu > 60.0) {
      double v = std::floor(mu + std::sqrt(mu) * normal() + 0.5);
      return v < 0 ? 0 : (int)v;
    }
    double L = std::exp(-mu), p = 1.0;
    int k = 0;
    do {
      k++;
      p *= uniform();
    } while (p > L);
    return k - 1;
  }
};

struct Blob {
  double x, y, z, sigma, amp;
};

// The phantom and the geometry the seeds draw (shared by mfs_generate and
// mfs_fiducials).
struct Scene {
  std::vector<Blob> blobs;
  double axis;
  std::vector<double> sx, sy, mag;
  Scene(const mfs_config* c) {
    Pcg32 rp(c->seed_phantom, 1), rg(c->seed_geom, 2);
    blobs.resize(c->nblobs);
    for (auto& b : blobs) {
      b.x = rp.uniform(-0.5, 0.5) * c->vol_x;
      b.y = rp.uniform(-0.5, 0.5) * c->vol_y;
      b.z = rp.uniform(-0.5, 0.5) * c->vol_z;
      b.sigma = rp.uniform(c->sigma_min, c->sigma_max);
      b.amp = rp.uniform(c->amp_min, c->amp_max);
    }
    axis = c->axis_min == c->axis_max ? c->axis_min
                                      : rg.uniform(c->axis_min, c->axis_max);
    const int nz = c->nz;
    sx.resize(nz);
    sy.resize(nz);
    mag.resize(nz);
    for (int i = 0; i < nz; i++) {
      sx[i] = rg.uniform(-c->shift_max, c->shift_max);
      sy[i] = rg.uniform(-c->shift_max, c->shift_max);
      mag[i] = c->mag_jitter > 0
                   ? 1.0 + rg.uniform(-c->mag_jitter, c->mag_jitter)
                   : 1.0;
    }
  }
  // Raw position of blob b in section i: q = R(axis) p + shift + c, with p
  // the aligned detector position (x0 cos t + z0 sin t, y0), magnified.
  void raw_pos(const mfs_config* c, const Blob& b, int i, double* qx,
               double* qy) const {
    const double t = (c->tilt_min + i * c->tilt_step) * M_PI / 180.0;
    const double ca = std::cos(axis * M_PI / 180.0),
                 sa = std::sin(axis * M_PI / 180.0);
    const double u0 = (b.x * std::cos(t) + b.z * std::sin(t)) * mag[i];
    const double v0 = b.y * mag[i];
    *qx = ca * u0 - sa * v0 + sx[i] + c->nx * 0.5;
    *qy = sa * u0 + ca * v0 + sy[i] + c->ny * 0.5;
  }
};

}  // namespace

int mfs_fiducials(const mfs_config* c, int count, float* raw_xy) {
  if (!c || !raw_xy || count < 0) return -1;
  const Scene sc(c);
  const int n = std::min(count, (int)sc.blobs.size());
  for (int k = 0; k < n; k++)
    for (int i = 0; i < c->nz; i++) {
      double qx, qy;
      sc.raw_pos(c, sc.blobs[k], i, &qx, &qy);
      const bool in = qx >= 0 && qx <= c->nx - 1 && qy >= 0 && qy <= c->ny - 1;
      raw_xy[2 * ((size_t)k * c->nz + i)] = in ? (float)qx : NAN;
      raw_xy[2 * ((size_t)k * c->nz + i) + 1] = in ? (float)qy : NAN;
    }
  return n;
}

int mfs_generate(const mfs_config* c, float* stack, float* tilts,
                 float* shift_x, float* shift_y, float* axis_out,
                 float* scale_out) {
  if (!c || !stack || c->nx <= 0 || c->ny <= 0 || c->nz <= 0) return 1;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const Scene sc(c);
  const std::vector<Blob>& blobs = sc.blobs;
  const double axis = sc.axis;
  const std::vector<double>&sx = sc.sx, &sy = sc.sy, &mag = sc.mag;
  if (axis_out) *axis_out = (float)axis;
  const double ca = std::cos(axis * M_PI / 180.0), sa = std::sin(axis * M_PI / 180.0);
  const double cxr = nx * 0.5, cyr = ny * 0.5;
  const size_t plane = (size_t)nx * ny;

#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < nz; i++) {
    const double t = (c->tilt_min + i * c->tilt_step) * M_PI / 180.0;
    const double ct = std::cos(t), st = std::sin(t);
    float* img = stack + (size_t)i * plane;
    std::fill(img, img + plane, 0.0f);
    std::vector<double> acc(plane, 0.0);
    for (const Blob& b : blobs) {
      // Aligned detector position (centred coordinates), magnified.
      const double u0 = (b.x * ct + b.z * st) * mag[i];
      const double v0 = b.y * mag[i];
      const double sg = b.sigma * mag[i];
      const double a = b.amp * b.sigma * std::sqrt(2.0 * M_PI);
      // Raw position of the blob centre: q = R(axis) p + shift + c.
      const double qx = ca * u0 - sa * v0 + sx[i] + cxr;
      const double qy = sa * u0 + ca * v0 + sy[i] + cyr;
      const double r = 4.0 * sg + 1.0;
      int x0 = (int)std::floor(qx - r), x1 = (int)std::ceil(qx + r);
      int y0 = (int)std::floor(qy - r), y1 = (int)std::ceil(qy + r);
      if (x1 < 0 || y1 < 0 || x0 >= nx || y0 >= ny) continue;
      x0 = std::max(x0, 0);
      y0 = std::max(y0, 0);
      x1 = std::min(x1, nx - 1);
      y1 = std::min(y1, ny - 1);
      const double inv2s2 = 1.0 / (2.0 * sg * sg);
      for (int y = y0; y <= y1; y++) {
        const double dy = y - qy;
        for (int x = x0; x <= x1; x++) {
          const double dx = x - qx;
          acc[(size_t)y * nx + x] += a * std::exp(-(dx * dx + dy * dy) * inv2s2);
        }
      }
    }
    for (size_t k = 0; k < plane; k++) img[k] = (float)acc[k];
  }

  // Noise: Gaussian at a per-tilt SNR, or Poisson at 'dose' electrons/px.
  // Each section draws from its own stream, so the sections run in parallel
  // with the same result.
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < nz; i++) {
    float* img = stack + (size_t)i * plane;
    Pcg32 rn(c->seed_noise + 7919ULL * (uint64_t)i, 3);
    if (c->noise_kind == 1 && c->snr > 0) {
      double m = 0, m2 = 0;
      for (size_t k = 0; k < plane; k++) {
        m += img[k];
        m2 += (double)img[k] * img[k];
      }
      m /= plane;
      double var = m2 / plane - m * m;
      double sd = std::sqrt(std::max(var, 1e-30) / c->snr);
      for (size_t k = 0; k < plane; k++) img[k] += (float)(sd * rn.normal());
    } else if (c->noise_kind == 2) {
      double mx = 0;
      for (size_t k = 0; k < plane; k++) mx = std::max(mx, (double)img[k]);
      // Signal modulates a uniform dose: counts ~ Poisson(dose*(1+img/max)).
      for (size_t k = 0; k < plane; k++)
        img[k] = (float)rn.poisson(c->dose * (1.0 + img[k] / (mx + 1e-30)));
    }
  }
  float scale = 1.0f;
  if (c->int16_counts) {
    float mn = stack[0], mx = stack[0];
    const long long total = (long long)plane * nz;
#pragma omp parallel for reduction(min : mn) reduction(max : mx)
    for (long long k = 0; k < total; k++) {
      mn = std::min(mn, stack[k]);
      mx = std::max(mx, stack[k]);
    }
    // Map [mn, mx] to [-30000, 30000] counts, rounded to integers.
    scale = 60000.0f / std::max(mx - mn, 1e-20f);
#pragma omp parallel for
    for (long long k = 0; k < total; k++)
      stack[k] = std::nearbyint((stack[k] - mn) * scale - 30000.0f);
  }
  if (scale_out) *scale_out = scale;
  for (int i = 0; i < nz; i++) {
    if (tilts) tilts[i] = c->tilt_min + i * c->tilt_step;
    if (shift_x) shift_x[i] = (float)sx[i];
    if (shift_y) shift_y[i] = (float)sy[i];
  }
  return 0;
}

// CPU restatement of the stack correction (CorrStack.cu) -- test
// infrastructure.  Correction (mGCorrect + mGRandom), b x b sum binning
// (mGBinImage), R-weighting (GRWeight between a row R2C and C2R) and unpadding.
#include "oracle.h"
#include "orc_internal.h"
#include "orc_stages.h"

namespace orc {

namespace {

// mGRandom (CorrStack.cu:116-140): deterministic neighbour pick for pixels
// that map outside the raw image; window 51, LCG next = (19 next + 57) mod
// 2601 (quirk Q10: the offset uses next mod 2601, not mod 51).
float random_fill_corr(int x, int y, int nx, int ny, const float* raw) {
  if (x < 0) x = -x;
  if (y < 0) y = -y;
  if (x >= nx) x = nx - 1 - (x % nx);
  if (y >= ny) y = ny - 1 - (y % ny);
  const int win = 51, size = win * win;
  uint32_t next = (uint32_t)(y * nx + x);
  for (int i = 0; i < 20; i++) {
    next = (next * 19u + 57u) % (uint32_t)size;
    int ix = (int)(next % size) - win / 2 + x;
    if (ix < 0 || ix >= nx) continue;
    int iy = (int)(next / win) - win / 2 + y;
    if (iy < 0 || iy >= ny) continue;
    return raw[(size_t)iy * nx + ix];
  }
  return raw[(size_t)y * nx + x];
}

}  // namespace

void corr_sizes(int nx, int ny, int bin, CorrSizes* s) {
  s->nx = nx;
  s->ny = ny;
  s->pad = pad_x(nx);
  s->cx = (s->pad / 2 - 1) * 2;  // corrected width (GCorrPatchShift::SetSizes)
  s->bx = nx / bin / 2 * 2;      // BinImage::GetBinSize on the aligned size
  s->by = ny / bin / 2 * 2;
  s->padb = pad_x(s->bx);
  s->binx = s->cx / s->bx;       // BinImage::SetupSizes
  s->biny = ny / s->by;
}

void correct_one(const float* raw, const CorrSizes& s, float shift_x,
                 float shift_y, float axis_deg, bool shift_only,
                 bool random_fill, bool rweight, float* out) {
  const int nx = s.nx, ny = s.ny;
  // mCorrectProj (CorrStack.cu:370-397): angle in radians as float.
  float theta = shift_only ? 0.0f : (float)(axis_deg * kD2R);
  const float c = fcos(theta), sn = fsin(theta);
  std::vector<float> corr((size_t)s.pad * ny, 0.0f);
  const float hx = (float)s.cx * 0.5f, hy = (float)ny * 0.5f;
  const float ox = shift_x + (float)nx * 0.5f, oy = shift_y + (float)ny * 0.5f;
  // Rows in parallel when called for one image; inside correct_stack's
  // per-image loop the nested region runs on one thread.
#pragma omp parallel for schedule(static)
  for (int y = 0; y < ny; y++) {
    for (int i = 0; i < s.cx; i++) {
      // mGCorrect (CorrStack.cu:142-179).
      float px = (float)i + 0.5f - hx;
      float py = (float)y + 0.5f - hy;
      float qx = px * c - py * sn;
      float qy = px * sn + py * c;
      qx += ox;
      qy += oy;
      int x = (int)(qx + 0.5f), yy = (int)(qy + 0.5f);
      float v;
      if (x >= 0 && x < nx && yy >= 0 && yy < ny)
        v = raw[(size_t)yy * nx + x];
      else if (random_fill)
        v = random_fill_corr(x, yy, nx, ny, raw);
      else
        v = invalid();
      corr[(size_t)y * s.pad + i] = v;
    }
  }
  // mGBinImage (CorrStack.cu:40-63): b x b sum, -1e30 if any invalid.
  std::vector<float> bin((size_t)s.padb * s.by, 0.0f);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < s.by; y++) {
    for (int x = 0; x < s.bx; x++) {
      float sum = 0.0f;
      bool bad = false;
      for (int iy = 0; iy < s.biny && !bad; iy++) {
        const float* row = corr.data() + (size_t)(y * s.biny + iy) * s.pad;
        for (int ix = 0; ix < s.binx; ix++) {
          float v = row[x * s.binx + ix];
          if (v < -1e10f) {
            bad = true;
            break;
          }
          sum += v;
        }
      }
      bin[(size_t)y * s.padb + x] = bad ? invalid() : sum;
    }
  }
  if (rweight) {
    // RWeight::DoIt (CorrStack.cu:431-462): per row R2C (length bx), scale
    // by 1/N and the ramp 2f(0.55 + 0.45 cos 2 pi f), f = k/N; C2R.
    const int n = s.bx, nc = n / 2 + 1;
    const float fn = 2.0f * ((float)nc - 1.0f);
    std::vector<float> filt(nc);
    for (int k = 0; k < nc; k++) {
      float r = (float)k / fn;
      filt[k] = 2.0f * r * (0.55f + 0.45f * fcos(6.2831852f * r));
    }
#pragma omp parallel
    {
      std::vector<float> cmp(2 * nc);
#pragma omp for schedule(static)
      for (int y = 0; y < s.by; y++) {
        float* row = bin.data() + (size_t)y * s.padb;
        fft_r2c_line(row, n, cmp.data());
        for (int k = 0; k < nc; k++) {
          cmp[2 * k] = cmp[2 * k] / fn * filt[k];
          cmp[2 * k + 1] = cmp[2 * k + 1] / fn * filt[k];
        }
        fft_c2r_line(cmp.data(), n, row);
      }
    }
  }
  // CorrUtil::Unpad (CorrStack.cu:158-168).
  for (int y = 0; y < s.by; y++)
    std::memcpy(out + (size_t)y * s.bx, bin.data() + (size_t)y * s.padb,
                sizeof(float) * s.bx);
}

void correct_stack(const float* stack, int nx, int ny, int nz, int bin,
                   const float* sx, const float* sy, const float* rot,
                   bool shift_only, bool random_fill, bool rweight,
                   std::vector<float>& out, CorrSizes* sizes) {
  CorrSizes s;
  corr_sizes(nx, ny, bin, &s);
  out.assign((size_t)s.bx * s.by * nz, 0.0f);
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < nz; z++)
    correct_one(stack + (size_t)z * nx * ny, s, sx[z], sy[z], rot[z],
                shift_only, random_fill, rweight,
                out.data() + (size_t)z * s.bx * s.by);
  if (sizes) *sizes = s;
}

}  // namespace orc

using namespace orc;

extern "C" void orc_binned_size(int nx, int ny, int bin, int* out_size) {
  CorrSizes s;
  corr_sizes(nx, ny, bin, &s);
  out_size[0] = s.bx;
  out_size[1] = s.by;
}

extern "C" void orc_correct_proj(const float* raw, int nx, int ny,
                                 float shift_x, float shift_y, float axis_deg,
                                 int shift_only, int random_fill, int bin,
                                 int rweight, float* out, int* out_size) {
  CorrSizes s;
  corr_sizes(nx, ny, bin, &s);
  correct_one(raw, s, shift_x, shift_y, axis_deg, shift_only != 0,
              random_fill != 0, rweight != 0, out);
  if (out_size) {
    out_size[0] = s.bx;
    out_size[1] = s.by;
  }
}

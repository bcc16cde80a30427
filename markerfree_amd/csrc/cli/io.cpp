// File formats of the drop-in command line (libmf_io.so): MRC stacks, tilt
// angle lists and the alignment parameter files.  Host code, plain stdio.
#include "markerfree/mf_io.h"

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <sys/types.h>
#include <fstream>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

// The 1024-byte MRC header (mrc/mrcheader.h:23-96), packed, little-endian.
#pragma pack(push, 1)
struct MrcHeader {
  int32_t nx, ny, nz, mode;
  int32_t nxstart, nystart, nzstart;
  int32_t mx, my, mz;
  float xlen, ylen, zlen;
  float alpha, beta, gamma;
  int32_t mapc, mapr, maps;
  float amin, amax, amean;
  int16_t ispg, nsymbt;
  int32_t next;
  int16_t creatid;
  char blank[30];
  int16_t nint, nreal, sub, zfac;
  float min2, max2, min3, max3, min4, max4;
  int16_t idtype, lens, nd1, nd2, vd1, vd2;
  float tiltangles[6];
  float xorg, yorg, zorg;
  char cmap[4], stamp[4];
  float rms;
  int32_t nlabl;
  char labels[10][80];
};
#pragma pack(pop)
static_assert(sizeof(MrcHeader) == 1024, "MRC header must be 1024 bytes");

size_t mode_bytes(int mode) {
  switch (mode) {
    case 0: return 1;
    case 1: return 2;
    case 2: return 4;
    case 6: return 2;
    default: return 0;
  }
}

int read_header(FILE* f, const char* path, MrcHeader* h) {
  if (std::fread(h, sizeof(MrcHeader), 1, f) != 1)
    return fail(std::string("short MRC header: ") + path);
  if (h->nx <= 0 || h->ny <= 0 || h->nz <= 0)
    return fail(std::string("bad MRC dimensions: ") + path);
  if (h->next < 0) return fail(std::string("bad MRC extended header: ") + path);
  if (!mode_bytes(h->mode)) return fail("File type unknown!");
  return 0;
}

struct File {
  FILE* f;
  explicit File(FILE* f_) : f(f_) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

extern "C" {

const char* mfio_last_error(void) { return g_err.c_str(); }

int mfio_mrc_read_info(const char* path, mfio_mrc_info* info) {
  if (!path || !info) return fail("null argument");
  File fh(std::fopen(path, "rb"));
  if (!fh.f) return fail(std::string("cannot open ") + path);
  MrcHeader h;
  if (read_header(fh.f, path, &h)) return -1;
  info->nx = h.nx;
  info->ny = h.ny;
  info->nz = h.nz;
  info->mode = h.mode;
  info->next = h.next;
  return 0;
}

int mfio_mrc_read_slices(const char* path, int z0, int count, float* out) {
  if (!path || !out) return fail("null argument");
  File fh(std::fopen(path, "rb"));
  if (!fh.f) return fail(std::string("cannot open ") + path);
  MrcHeader h;
  if (read_header(fh.f, path, &h)) return -1;
  if (z0 < 0 || count < 0 || z0 + count > h.nz)
    return fail(std::string("MRC image range out of bounds: ") + path);
  const size_t plane = (size_t)h.nx * h.ny, psize = mode_bytes(h.mode);
  const long long at = (long long)sizeof(MrcHeader) + h.next +
                       (long long)z0 * (long long)(plane * psize);
  if (fseeko(fh.f, (off_t)at, SEEK_SET) != 0)
    return fail(std::string("cannot seek to MRC data: ") + path);
  std::vector<unsigned char> buf(plane * psize);
  for (int z = 0; z < count; z++) {
    if (std::fread(buf.data(), psize, plane, fh.f) != plane)
      return fail(std::string("short MRC data: ") + path);
    float* o = out + (size_t)z * plane;
    switch (h.mode) {
      case 0:
        for (size_t i = 0; i < plane; i++) o[i] = (float)buf[i];
        break;
      case 1: {
        const int16_t* p = (const int16_t*)buf.data();
        for (size_t i = 0; i < plane; i++) o[i] = (float)p[i];
        break;
      }
      case 2:
        std::memcpy(o, buf.data(), plane * 4);
        break;
      case 6: {
        const uint16_t* p = (const uint16_t*)buf.data();
        for (size_t i = 0; i < plane; i++) o[i] = (float)p[i];
        break;
      }
    }
  }
  return 0;
}

int mfio_mrc_read(const char* path, float* out) {
  mfio_mrc_info info;
  if (mfio_mrc_read_info(path, &info)) return -1;
  return mfio_mrc_read_slices(path, 0, info.nz, out);
}

int mfio_mrc_write(const char* path, const float* data, int nx, int ny,
                   int nz, float stats[3]) {
  if (!path || !data || nx <= 0 || ny <= 0 || nz <= 0)
    return fail("bad arguments");
  MrcHeader h;
  std::memset(&h, 0, sizeof(h));
  // InitializeHeader (mrcstack.cpp:421-454) + SetSize.
  h.mode = 2;
  h.mx = nx;
  h.my = ny;
  h.mz = nz;
  h.xlen = h.ylen = h.zlen = 1;
  h.alpha = h.beta = h.gamma = 90;
  h.mapc = 1;
  h.mapr = 2;
  h.maps = 3;
  h.creatid = 1000;
  h.nx = nx;
  h.ny = ny;
  h.nz = nz;
  // UpdateHeader (mrcstack.cpp:456-500): statistics of the middle slice.
  const size_t plane = (size_t)nx * ny;
  const float* mid = data + (size_t)(int)(nz * .5) * plane;
  float amin = FLT_MAX, amax = FLT_MIN;
  long double amean = 0;
  int count = 0;
  for (size_t i = plane; i--;) {
    const float v = mid[i];
    if (v < (float)-1e10) continue;
    if (v > amax) amax = v;
    if (v < amin) amin = v;
    amean += v;
    count++;
  }
  amean /= count;
  h.amin = amin;
  h.amax = amax;
  h.amean = (float)amean;
  h.nlabl = 0;
  if (stats) {
    stats[0] = h.amin;
    stats[1] = h.amax;
    stats[2] = h.amean;
  }
  File fh(std::fopen(path, "wb"));
  if (!fh.f) return fail(std::string("cannot create ") + path);
  if (std::fwrite(&h, sizeof(h), 1, fh.f) != 1 ||
      std::fwrite(data, sizeof(float), plane * nz, fh.f) != plane * nz)
    return fail(std::string("write failed: ") + path);
  if (std::fclose(fh.f) != 0) {
    fh.f = nullptr;
    return fail(std::string("write failed: ") + path);
  }
  fh.f = nullptr;
  return 0;
}

int mfio_read_angles(const char* path, float* out, int cap) {
  if (!path) return fail("null argument");
  std::ifstream in(path);
  if (!in.good()) return fail(std::string("cannot open ") + path);
  int n = 0;
  while (in.good()) {
    float v;
    in >> v;
    if (in.fail()) break;
    if (out && n < cap) out[n] = v;
    n++;
  }
  return n;
}

int mfio_param_path(const char* outfile, int mode, char* buf, size_t size) {
  if (!outfile || !buf) return fail("null argument");
  std::string name(outfile);
  const char* suffix = mode == 0 ? ".txt" : ".xf";
  const size_t pos = name.find(".mrc");
  if (pos == std::string::npos)
    name += suffix;
  else
    name = name.substr(0, pos) + suffix;  // the reference's strcpy at ".mrc"
  if (name.size() + 1 > size) return fail("path buffer too small");
  std::memcpy(buf, name.c_str(), name.size() + 1);
  return 0;
}

int mfio_save_txt(const char* path, int nx, int ny, int nz,
                  const float* rotate, const double* shift_x,
                  const double* shift_y, const float* tilts,
                  float angle_offset, int out_bin, float elapsed_s) {
  if (!path || !rotate || !shift_x || !shift_y || !tilts)
    return fail("null argument");
  File fh(std::fopen(path, "wt"));
  if (!fh.f) return fail(std::string("Unable to open ") + path + ".");
  FILE* f = fh.f;
  std::fprintf(f, "# Markerfree Alignment\n");
  std::fprintf(f, "# RawSize = %d %d %d\n", nx, ny, nz);
  std::fprintf(f, "# SEC       ROT          TX        TY         TILT\n");
  for (int i = 0; i < nz; i++) {
    // GetParam(1, ...): shifts times the bin factor 1.
    const float sx = (float)shift_x[i], sy = (float)shift_y[i];
    std::fprintf(f, "%5d   %9.4f  %9.3f  %9.3f   %8.2f\n", i, rotate[i], sx,
                 sy, tilts[i]);
  }
  std::fprintf(f, "angleoffset: %f\n", angle_offset);
  std::fprintf(f, "Output Bin: %d\n", out_bin);
  std::fprintf(f, "Elapsed time: %f seconds\n", elapsed_s);
  return 0;
}

int mfio_save_xf(const char* path, int nx, int nz, const float* rotate,
                 const double* shift_x, const double* shift_y, int inv) {
  if (!path || !rotate || !shift_x || !shift_y) return fail("null argument");
  File fh(std::fopen(path, "wt"));
  if (!fh.f) return fail(std::string("Unable to open ") + path + ".");
  const float width = (float)nx, height = (float)nx;
  const float cx = width / 2.0f, cy = height / 2.0f;
  for (int k = 0; k < nz; k++) {
    const int i = inv ? nz - 1 - k : k;
    const float s0 = (float)shift_x[i], s1 = (float)shift_y[i];
    const float th = (float)(rotate[i] * M_PI / 180.0f);
    const float c = std::cos(th), s = std::sin(th);
    const float t1 = -s0 - cx + c * cx - s * cy;
    const float t2 = -s1 - cy + s * cx + c * cy;
    const float p4 = t1 * c + t2 * s - cx + c * cx + s * cy;
    const float p5 = -t1 * s + t2 * c - cy - s * cx + c * cy;
    std::fprintf(fh.f, "%9.3f  %9.3f  %9.3f  %9.3f  %9.3f  %9.3f \n", c, s, -s,
                 c, p4, p5);
  }
  return 0;
}

}  // extern "C"

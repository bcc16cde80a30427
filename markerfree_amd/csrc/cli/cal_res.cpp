// cal_res: fiducial residuals of a Markerfree alignment.
//
// Counterpart of the reference's tool (tool/src/main/main.cpp:35-237,
// options tool/src/main/opts.h): it reads the alignment .txt (-x), the stack
// (-i, for its size) and the tilt angles (-a), and reports how far the
// alignment leaves each tracked marker from one 3-D position.  The reference
// detects and tracks gold markers itself (OpenCV templates, -d DIAMETER);
// here the tracks come from a text file (-f): one "marker section x y" line
// per observation, raw-image pixel coordinates, sections 0-based in the
// .txt's (ascending tilt) order.
#include <getopt.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "markerfree/mf_io.h"

namespace {

void usage() {
  std::printf(
      "[-i INPUT][-a INITANGLE][-x TXTINPUT][-f FIDUCIALS]([-d DIAMETER])"
      "([-m MINVIEWS])([-c 0|1])([-v VERBOSE])\n"
      "  [--input(-i) INPUT]             stack (its header gives the size)\n"
      "  [--initangle(-a) INIT ANGLE]    tilt angles (.rawtlt)\n"
      "  [--txtinput(-x) TXT INPUT]      Markerfree alignment (.txt)\n"
      "  [--fiducials(-f) FIDUCIALS]     marker tracks: marker section x y\n"
      "  ([--diameter(-d) DIAMETER])     marker size (accepted; detection is"
      " not done here)\n"
      "  ([--minviews(-m) N])            views a marker needs (default 70%% of"
      " the sections, as the reference's tracks)\n"
      "  ([--centre(-c) 0|1])            fit the common rotation-centre offset"
      " (default 1)\n"
      "  ([--verbose(-v) 0|1])           print every section\n");
}

}  // namespace

int main(int argc, char** argv) {
  static struct option longopts[] = {
      {"help", no_argument, nullptr, 'h'},
      {"input", required_argument, nullptr, 'i'},
      {"initangle", required_argument, nullptr, 'a'},
      {"txtinput", required_argument, nullptr, 'x'},
      {"fiducials", required_argument, nullptr, 'f'},
      {"diameter", required_argument, nullptr, 'd'},
      {"minviews", required_argument, nullptr, 'm'},
      {"centre", required_argument, nullptr, 'c'},
      {"verbose", required_argument, nullptr, 'v'},
      {nullptr, 0, nullptr, 0}};
  const char *input = nullptr, *angles = nullptr, *txt = nullptr,
             *fids = nullptr;
  int min_views = -1, centre = 1, verbose = 0;
  int ch;
  while ((ch = getopt_long(argc, argv, "hi:a:x:f:d:m:c:v:", longopts,
                           nullptr)) != -1) {
    switch (ch) {
      case 'i': input = optarg; break;
      case 'a': angles = optarg; break;
      case 'x': txt = optarg; break;
      case 'f': fids = optarg; break;
      case 'd': break;
      case 'm': min_views = std::atoi(optarg); break;
      case 'c': centre = std::atoi(optarg); break;
      case 'v': verbose = std::atoi(optarg); break;
      case 'h': usage(); return 0;
      default: usage(); return 1;
    }
  }
  if (!input || !txt || !fids) {
    usage();
    return 1;
  }
  mfio_mrc_info info;
  if (mfio_mrc_read_info(input, &info) != 0) {
    std::fprintf(stderr, "%s\n", mfio_last_error());
    return 1;
  }
  const int nz = info.nz;
  std::vector<float> rot(nz), tilts(nz);
  std::vector<double> sx(nz), sy(nz);
  float offset = 0.0f;
  int out_bin = 1;
  const int rows = mfio_read_txt(txt, nz, rot.data(), sx.data(), sy.data(),
                                 tilts.data(), &offset, &out_bin);
  if (rows != nz) {
    std::fprintf(stderr, "%s: %d parameter rows for %d sections\n", txt, rows,
                 nz);
    return 1;
  }
  if (angles) {
    std::vector<float> a(2 * nz + 1);
    const int n = mfio_read_angles(angles, a.data(), (int)a.size());
    if (n < nz) {
      std::fprintf(stderr, "%s: %d angles for %d sections\n", angles, n, nz);
      return 1;
    }
    // The .txt rows are in ascending order; a descending file is reversed.
    const bool desc = a[0] > a[1];
    for (int i = 0; i < nz; i++) tilts[i] = desc ? a[nz - 1 - i] : a[i];
  }
  // The .txt holds raw-pixel shifts (GetParam's bin factor 1) and the
  // angles without the offset.
  for (int i = 0; i < nz; i++) tilts[i] += offset;
  std::map<long, std::vector<float>> tracks;  // marker -> [nz][2]
  FILE* f = std::fopen(fids, "r");
  if (!f) {
    std::fprintf(stderr, "cannot read %s\n", fids);
    return 1;
  }
  char line[256];
  while (std::fgets(line, sizeof line, f)) {
    long m;
    int sec;
    float x, y;
    if (line[0] == '#' || std::sscanf(line, "%ld %d %f %f", &m, &sec, &x, &y) != 4)
      continue;
    if (sec < 0 || sec >= nz) continue;
    auto it = tracks.find(m);
    if (it == tracks.end())
      it = tracks.emplace(m, std::vector<float>(2 * nz, NAN)).first;
    it->second[2 * sec] = x;
    it->second[2 * sec + 1] = y;
  }
  std::fclose(f);
  std::vector<float> xy;
  for (auto& t : tracks) xy.insert(xy.end(), t.second.begin(), t.second.end());
  if (min_views < 0) min_views = (int)std::ceil(0.7 * nz);
  std::vector<double> per(nz), summary(6);
  if (mfio_fiducial_residuals(info.nx, info.ny, nz, rot.data(), sx.data(),
                              sy.data(), tilts.data(), (int)tracks.size(),
                              xy.data(), min_views, centre, per.data(),
                              nullptr, summary.data()) != 0) {
    std::fprintf(stderr, "bad fiducial input\n");
    return 1;
  }
  if (verbose) {
    std::printf("# SEC    TILT    RMS(px)\n");
    for (int i = 0; i < nz; i++)
      std::printf("%5d  %8.2f  %8.3f\n", i, tilts[i], per[i]);
  }
  std::printf("Markers used: %d of %zu\n", (int)summary[4], tracks.size());
  std::printf("Observations: %d\n", (int)summary[3]);
  std::printf("Mean residual: %.4f px\n", summary[0]);
  std::printf("RMS residual: %.4f px\n", summary[1]);
  std::printf("Max residual: %.4f px\n", summary[2]);
  std::printf("Centre offset: %.4f px\n", summary[5]);
  return 0;
}

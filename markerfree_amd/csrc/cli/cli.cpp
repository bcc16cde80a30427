// markerfree: the drop-in command line (main/main.cpp, main/opts.h,
// main/Process.cpp) over the C-ABI of libmarkerfree_hip.so and the file
// formats of libmf_io.so.
//
//   markerfree -i in.{st,mrc} -o out.mrc -a tilts.rawtlt
//              -g offset,pitch,zshift,thickness,AlignZ,OutBin,GPU
//              -p NPROJ -s {0,1}
//
// Flags, defaults, validation, messages and exit codes follow opts.h and
// main.cpp; the stage trace on stdout is printed by the library.  The
// reference's intermediate buf.mrc is not written: the preprocessed stack
// stays in device memory.
#include <getopt.h>
#include <unistd.h>
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <algorithm>
#include <string>
#include <vector>

#include "markerfree/mf_api.h"
#include "markerfree/mf_io.h"

namespace {

// main/opts.h:16-41.
struct Options {
  std::string input, output, initial, angle;
  float pitch_angle = 0, zshift = 0, offset = 0;
  int thickness = 0, align_z = 600, out_bin = 1, gpu = 0;
  int save_mode = 0, nproj = 10;
  std::string method;
  int iteration = 0;
  float gamma = 0;
  int reproj_mode = 8;  // not a reference flag: MF_REPROJ_MODE=0 selects
                        // the materialised slab, MF_REPROJ_MODE=1 the fused
                        // form bit-identical to the reference's evaluation
};

// UsageDual (opts.h:43-65).
void usage() {
  std::cout << "[-i INPUT FILENAME]\n"
            << "    MRC file for reconstruction\n"
            << "[-o OUTPUT FILENAME]\n"
            << "    MRC filename for result\n"
            << "[-a TILT ANGLE FILENAME]\n"
            << "    Tilt Angles\n"
            << "[-n INITIAL RECONSTRUCTION]\n"
            << "    MRC file as initial model (optional)\n"
            << "[-g O,P,Z,T,A,B,G]\n"
            << "    Geometry: offset,pitch_angle,zshift,thickness,AlignZ,OutBin,GPU\n"
            << "[-m METHODS (I,R)]\n"
            << "    BPT\n"
            << "    SART: SART iteration_number,relax_parameter\n"
            << "    RP\n"
            << "[-p NPROJ]\n"
            << "    Number of neighbor projections\n"
            << "[-s SAVE_MODE]\n"
            << "    SaveParameter (0 or 1)\n"
            << "[-h]\n"
            << "    Help\n";
}

// PrintOpts (opts.h:83-100).  iter/step are uninitialised in the reference;
// here they print 0.
void print_opts(const Options& o) {
  std::cout << "pitch_angle = " << o.pitch_angle << "\n"
            << "zshift = " << o.zshift << "\n"
            << "thickness = " << o.thickness << "\n"
            << "offset = " << o.offset << "\n"
            << "AlignZ = " << o.align_z << "\n"
            << "OutBin = " << o.out_bin << "\n"
            << "GPU = " << o.gpu << "\n"
            << "Savemode = " << o.save_mode << "\n"
            << "nProj = " << o.nproj << "\n"
            << "input = " << o.input << "\n"
            << "output = " << o.output << "\n"
            << "initial = " << o.initial << "\n"
            << "method = " << o.method << "\n"
            << "iter = " << o.iteration << "\n"
            << "step = " << o.gamma << "\n";
}

std::string first_word(const char* s) {
  std::string w;
  std::istringstream(s) >> w;
  return w;
}

// GetOpts (opts.h:102-243): 1 = run, 0 = help printed, -1 = error.
int get_opts(int argc, char** argv, Options* o) {
  static struct option longopts[] = {
      {"help", no_argument, nullptr, 'h'},
      {"input", required_argument, nullptr, 'i'},
      {"output", required_argument, nullptr, 'o'},
      {"initial", required_argument, nullptr, 'n'},
      {"angle", required_argument, nullptr, 'a'},
      {"geometry", required_argument, nullptr, 'g'},
      {"mode", required_argument, nullptr, 'm'},
      {"savemode", required_argument, nullptr, 's'},
      {"nproj", required_argument, nullptr, 'p'},
      {nullptr, 0, nullptr, 0}};
  int ch;
  while ((ch = getopt_long(argc, argv, "hi:o:n:a:g:m:s:p:", longopts,
                           nullptr)) != -1) {
    switch (ch) {
      case '?':
        printf("Invalid option '%s'.", argv[optind - 1]);
        fflush(stdout);
        return -1;
      case ':':
        printf("Missing option argument for '%s'.", argv[optind - 1]);
        fflush(stdout);
        return -1;
      case 'h':
        usage();
        return 0;
      case 'i':
        o->input = first_word(optarg);
        break;
      case 'o':
        o->output = first_word(optarg);
        break;
      case 'n':
        o->initial = first_word(optarg);
        break;
      case 'a':
        o->angle = first_word(optarg);
        break;
      case 'g': {
        // Seven comma-separated fields; fewer is an error.
        std::istringstream iss(optarg);
        std::string f[7];
        for (int k = 0; k < 7; k++) {
          const bool ok = k < 6 ? (bool)std::getline(iss, f[k], ',')
                                : (bool)std::getline(iss, f[k]);
          if (!ok) return -1;
        }
        o->offset = (float)atof(f[0].c_str());
        o->pitch_angle = (float)atof(f[1].c_str());
        o->zshift = (float)atof(f[2].c_str());
        o->thickness = atoi(f[3].c_str());
        o->align_z = atoi(f[4].c_str());
        o->out_bin = atoi(f[5].c_str());
        o->gpu = atoi(f[6].c_str());
        break;
      }
      case 's':
        o->save_mode = atoi(optarg);
        if (o->save_mode != 0 && o->save_mode != 1) {
          printf("Savemode must be 0 or 1.");
          fflush(stdout);
          return -1;
        }
        break;
      case 'p':
        o->nproj = atoi(optarg);
        if (o->nproj <= 0) {
          printf("nProj must be > 0.");
          fflush(stdout);
          return -1;
        }
        break;
      case 'm': {
        if (strcmp(optarg, "BPT") != 0 && strcmp(optarg, "RP") != 0) {
          std::istringstream iss(optarg);
          std::string tmp;
          if (!std::getline(iss, o->method, ',')) {
            printf("Invalid method argument.");
            fflush(stdout);
            return -1;
          }
          if (o->method == "SIRT" || o->method == "SART") {
            if (!std::getline(iss, tmp, ',')) {
              printf("Missing iteration number.");
              fflush(stdout);
              return -1;
            }
            o->iteration = atoi(tmp.c_str());
            if (!std::getline(iss, tmp)) {
              printf("Missing relax parameter.");
              fflush(stdout);
              return -1;
            }
            o->gamma = (float)atof(tmp.c_str());
          }
        } else {
          o->method = optarg;
        }
        break;
      }
      default:
        return -1;
    }
  }
  return 1;
}

double now_s() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec / 1e6;
}

// Host float buffer, page-locked when the runtime can provide it (faster,
// asynchronous copies), ordinary memory otherwise.
class HostBuf {
 public:
  explicit HostBuf(size_t n) {
    if (mf_host_alloc(n * sizeof(float), (void**)&p_) != MF_OK) {
      p_ = nullptr;
      v_.resize(n);
    }
  }
  ~HostBuf() {
    if (p_) mf_host_free(p_);
  }
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  float* data() { return p_ ? p_ : v_.data(); }

 private:
  float* p_ = nullptr;
  std::vector<float> v_;
};

// Process::DoIt (Process.cpp:34-92).  Returns the process exit code: 0 even
// after an input or angle-file failure, as the reference does.
int run(const Options& o) {
  const double t_start = now_s();
  // ReadStack (Process.cpp:129-174).
  mfio_mrc_info info;
  if (mfio_mrc_read_info(o.input.c_str(), &info) != 0) {
    if (std::string(mfio_last_error()) == "File type unknown!")
      printf("File type unknown!\n");
    else
      printf("[Error] Cannot open input file: %s\n", o.input.c_str());
    printf("[Abort] Input stack loading failed. Please check the file path.\n");
    printf("[Abort] DoIt() stopped due to input file loading failure.\n");
    return 0;
  }
  const int nx = info.nx, ny = info.ny, nz = info.nz;
  // ReadAngles (Process.cpp:58-68, 186-206).
  std::vector<float> angles(nz);
  const int na = mfio_read_angles(o.angle.c_str(), angles.data(), nz);
  if (na < 0) {
    printf("[Error] Cannot open angle file: %s\n", o.angle.c_str());
    printf("[Abort] DoIt() stopped due to angle file loading failure.\n");
    return 0;
  }
  if (na < nz) {
    // The reference reads past the end of its angle list here.
    printf("[Error] Angle file %s holds %d angles for %d images\n",
           o.angle.c_str(), na, nz);
    printf("[Abort] DoIt() stopped due to angle file loading failure.\n");
    return 0;
  }
  const bool inv = nz > 1 && angles[0] > angles[1];

  mf_ctx* ctx = nullptr;
  if (mf_ctx_create(o.gpu, 0, &ctx) != MF_OK) {
    printf("[Error] Cannot use GPU %d\n", o.gpu);
    return 1;
  }
  // Launch shapes from the autotuner's cache: MF_TUNING names a file ("" to
  // keep the built-in heuristics); by default the cache installed beside the
  // library, <bin>/../tuned/gfx950.json, when it exists.
  {
    std::string path;
    const char* env = getenv("MF_TUNING");
    if (env) {
      path = env;
    } else {
      char exe[4096];
      const ssize_t n = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
      if (n > 0) {
        exe[n] = 0;
        std::string dir(exe);
        dir = dir.substr(0, dir.rfind('/'));
        path = dir + "/../tuned/gfx950.json";
        if (access(path.c_str(), R_OK) != 0) path.clear();
      }
    }
    if (!path.empty() && mf_tune_load(ctx, path.c_str(), nullptr) != MF_OK)
      printf("[Warning] tuning file %s not applied: %s\n", path.c_str(),
             mf_last_error(ctx));
  }
  mf_series* s = nullptr;
  auto die = [&](const char* what) {
    printf("[Error] %s: %s\n", what,
           s ? mf_series_last_error(s) : mf_last_error(ctx));
    if (s) mf_series_destroy(s);
    mf_ctx_destroy(ctx);
    return 1;
  };
  if (mf_series_create(ctx, nx, ny, nz, nullptr, angles.data(), &s) != MF_OK)
    return die("series");
  // Streaming ingest: the stack is read in chunks of images into two
  // page-locked buffers in turn, and each chunk's upload runs while the next
  // chunk is read from the file.  The buffers are small (16 MB each): the
  // first host write to newly page-locked memory costs about 0.3 ms per MB.
  {
    const size_t plane = (size_t)nx * ny;
    const int chunk = (int)std::max<size_t>(
        1, std::min<size_t>(nz, ((size_t)16 << 20) / (plane * sizeof(float))));
    HostBuf buf[2] = {HostBuf((size_t)chunk * plane),
                      HostBuf((size_t)chunk * plane)};
    for (int z0 = 0, k = 0; z0 < nz; z0 += chunk, k++) {
      const int cnt = std::min(chunk, nz - z0);
      float* b = buf[k & 1].data();
      if (mfio_mrc_read_slices(o.input.c_str(), z0, cnt, b) != 0) {
        printf("[Error] Cannot open input file: %s (%s)\n", o.input.c_str(),
               mfio_last_error());
        printf("[Abort] Input stack loading failed. Please check the file "
               "path.\n");
        printf("[Abort] DoIt() stopped due to input file loading failure.\n");
        mf_series_upload(s, 0, 0, nullptr);
        mf_series_destroy(s);
        mf_ctx_destroy(ctx);
        return 0;
      }
      if (mf_series_upload(s, z0, cnt, b) != MF_OK) return die("upload");
    }
    if (mf_series_upload(s, 0, 0, nullptr) != MF_OK) return die("upload");
  }
  mf_options mo;
  mf_options_default(&mo);
  mo.offset = o.offset;
  mo.pitch_angle = o.pitch_angle;
  mo.zshift = o.zshift;
  mo.thickness = o.thickness;
  mo.align_z = o.align_z;
  mo.out_bin = o.out_bin;
  mo.nproj = o.nproj;
  mo.verbose = 1;
  mo.reproj_mode = o.reproj_mode;
  // mPreprocess + (AlignZ != 0) SetParam, mCoarseAlign, mProjAlign.  With
  // AlignZ 0 the reference dereferences unset parameters and crashes; here
  // the stack is corrected with zero parameters and no parameter file is
  // written.
  if (mf_run(s, &mo) != MF_OK) return die("alignment");

  // CorrectStack (Process.cpp:329-372).
  int bx = 0, by = 0;
  if (mf_correct(s, o.out_bin, nullptr, &bx, &by) != MF_OK)
    return die("correction");
  std::vector<float> out((size_t)bx * by * nz);
  if (mf_correct(s, o.out_bin, out.data(), &bx, &by) != MF_OK)
    return die("correction");
  printf("Update head!\n");
  float stats[3];
  if (mfio_mrc_write(o.output.c_str(), out.data(), bx, by, nz, stats) != 0) {
    printf("[Error] Cannot write output file: %s (%s)\n", o.output.c_str(),
           mfio_last_error());
    mf_series_destroy(s);
    mf_ctx_destroy(ctx);
    return 1;
  }
  printf("amin is %f, amax is %f, amean is %Lf\n", stats[0], stats[1],
         (long double)stats[2]);
  printf("Updating finished!\n");

  const double elapsed = now_s() - t_start;
  if (o.align_z != 0) {
    // CollectParam (Process.cpp:399-418) + SaveParam::GetParam.
    std::vector<double> sx(nz), sy(nz);
    std::vector<float> rot(nz), tilts(angles);
    float offset = 0;
    if (mf_get_params(s, sx.data(), sy.data(), rot.data(), &offset) != MF_OK)
      return die("parameters");
    if (inv) std::vector<float>(tilts.rbegin(), tilts.rend()).swap(tilts);
    char path[4096];
    if (mfio_param_path(o.output.c_str(), o.save_mode, path, sizeof path) ==
        0) {
      const int rc =
          o.save_mode == 0
              ? mfio_save_txt(path, nx, ny, nz, rot.data(), sx.data(),
                              sy.data(), tilts.data(), offset, o.out_bin,
                              (float)elapsed)
              : mfio_save_xf(path, nx, nz, rot.data(), sx.data(), sy.data(),
                             inv ? 1 : 0);
      if (rc != 0) {
        printf("Unable to open %s.\n", path);
        printf("Alignment data will not be saved\n\n");
      }
    }
  }
  std::cout << "Elapsed time: " << elapsed << " seconds" << std::endl;
  mf_series_destroy(s);
  mf_ctx_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  const char* rm = getenv("MF_REPROJ_MODE");
  if (rm && rm[0] == '0') o.reproj_mode = 0;
  if (rm && rm[0] == '1') o.reproj_mode = 1;
  if (get_opts(argc, argv, &o) <= 0) {
    printf("***WRONG INPUT.\n");
    fflush(stdout);
    return -1;
  }
  print_opts(o);
  std::cout.flush();
  return run(o);
}

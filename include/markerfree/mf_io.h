/* File formats of the drop-in command line: MRC stacks, tilt-angle lists and
 * the alignment parameter files.  Host code only (libmf_io.so); the
 * `markerfree` program and the tests call it.
 *
 * Reference interfaces replaced:
 *   MRC header          mrc/mrcheader.h:23-96 (1024 bytes, packed)
 *   MrcStackM           mrc/mrcstack.cpp:5-525 (read modes 0/1/2, write
 *                       float32, InitializeHeader :421-454, UpdateHeader
 *                       :456-500)
 *   Process::ReadAngles main/Process.cpp:186-206
 *   SaveParam           method/SaveParam.cpp:15-262 (.txt and .xf)
 */
#ifndef MARKERFREE_MF_IO_H
#define MARKERFREE_MF_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 0 on success; otherwise a message is available from mfio_last_error(). */
typedef struct {
  int nx, ny, nz;
  int mode;   /* 0 uint8, 1 int16, 2 float32, 6 uint16 */
  int next;   /* extended-header bytes between the header and the data */
} mfio_mrc_info;

const char* mfio_last_error(void);

/* MrcStackM::ReadHeader (mrcstack.cpp:43-49). */
int mfio_mrc_read_info(const char* path, mfio_mrc_info* info);
/* MrcStackM::ReadSlice for every slice (mrcstack.cpp:51-119): data after
 * 1024 + next bytes, converted to float (mode 0 unsigned, 1 signed int16,
 * 2 float32; mode 6 unsigned int16 in addition).  out: nx*ny*nz floats. */
int mfio_mrc_read(const char* path, float* out);
/* Images [z0, z0+count) only, as float [count][ny][nx], for streaming
 * ingest (reads no other image data). */
int mfio_mrc_read_slices(const char* path, int z0, int count, float* out);

/* The aligned-stack writer of Process::CorrectStack (Process.cpp:329-372):
 * InitializeHeader + SetSize (mrcstack.cpp:421-454), float32 data, then
 * UpdateHeader (mrcstack.cpp:456-500): amin/amax/amean from the middle
 * slice (index (int)(nz * 0.5)) skipping values < -1e10, amax starting at
 * FLT_MIN, amean accumulated in long double, nlabl = 0.  stats (may be
 * NULL) receives amin, amax, amean as written. */
int mfio_mrc_write(const char* path, const float* data, int nx, int ny,
                   int nz, float stats[3]);

/* Process::ReadAngles (Process.cpp:186-206): whitespace-separated floats
 * until the first that fails to parse.  Returns the count (<= cap values
 * stored), or -1 if the file cannot be opened. */
int mfio_read_angles(const char* path, float* out, int cap);

/* SaveParam::GetParam's file name (SaveParam.cpp:132-136, 150-154): the first ".mrc"
 * in 'outfile' replaced by ".txt" (mode 0) or ".xf" (mode 1), else the
 * suffix appended.  Returns 0, or -1 if 'size' is too small. */
int mfio_param_path(const char* outfile, int mode, char* buf, size_t size);

/* SaveHeader + SaveAllParam + SaveTime (SaveParam.cpp:185-262): the .txt
 * parameter file.  All arrays hold nz entries in ascending tilt order;
 * tilts are the angles as read from the file (ascending, no offset). */
int mfio_save_txt(const char* path, int nx, int ny, int nz,
                  const float* rotate, const double* shift_x,
                  const double* shift_y, const float* tilts,
                  float angle_offset, int out_bin, float elapsed_s);

/* SaveXf (SaveParam.cpp:15-120): one line per image, in file order when the
 * file's tilts descend ('inv' nonzero), with cx = cy = nx/2 (the reference
 * uses the width for both). */
int mfio_save_xf(const char* path, int nx, int nz, const float* rotate,
                 const double* shift_x, const double* shift_y, int inv);

/* Reads a .txt parameter file written by SaveAllParam (SaveParam.cpp:236-256;
 * this library's mfio_save_txt): the "SEC ROT TX TY TILT" rows in order, up
 * to cap of them, and the angleoffset and Output Bin lines (either pointer
 * may be NULL).  Returns the number of rows, or -1 if the file cannot be
 * opened. */
int mfio_read_txt(const char* path, int cap, float* rotate, double* shift_x,
                  double* shift_y, float* tilts, float* angle_offset,
                  int* out_bin);

/* Fiducial residuals of an alignment: the evaluation the reference's cal_res
 * tool makes (tool/src/main/main.cpp:35-237, triangulation
 * tool/src/nbundle/bundle_core.cpp:149-216) from marker tracks given as
 * coordinates (its OpenCV marker detection is not part of this library).
 * raw_xy [nmarkers][nz][2]: each marker's raw-image position per section
 * (pixel k at position k; NaN where it is not seen).  Each observation goes
 * to the aligned frame by the inverse of the stack correction,
 * p = R(-rotate) (q - (nx/2, ny/2) - shift); each marker seen in at least
 * min_views sections gets the 3-D position (X, Y, Z) minimising
 * sum |p - (X cos t + Z sin t + d, Y)|^2 over its views, t = tilts (with the
 * angle offset), d one x offset shared by all (fitted when fit_centre is
 * nonzero, else 0: the alignment leaves the rotation centre's x free).
 * Outputs: per_section[nz] RMS residual (px, NaN if unseen; may be NULL),
 * marker_xyz[nmarkers][3] (NaN for markers not used; may be NULL) and
 * summary[6] = mean, RMS and largest residual (px), observations used,
 * markers used, d.  Returns 0, or -1 on bad arguments. */
int mfio_fiducial_residuals(int nx, int ny, int nz, const float* rotate,
                            const double* shift_x, const double* shift_y,
                            const float* tilts, int nmarkers,
                            const float* raw_xy, int min_views,
                            int fit_centre, double* per_section,
                            double* marker_xyz, double* summary);

#ifdef __cplusplus
}
#endif
#endif /* MARKERFREE_MF_IO_H */

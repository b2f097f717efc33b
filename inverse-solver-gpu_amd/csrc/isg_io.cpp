// isg_io.cpp -- the reference's on-disk formats, for side-by-side runs.
//
//   phantom text   numunk lines "%f", raster order, real contrast before the
//                  reader's scale (setup_born.cpp:105-112 reads o = 0.1 * val)
//   cg_<k>.bin     numunk std::complex<double>, raster order (the unkmap
//                  gather of setup_born.cpp:209-218); cg_ref.bin the same for
//                  the true contrast (setup_born.cpp:115-121)
//   hist.txt       one line "%e %e %e %e %e\n" per DBIM iteration: misfit,
//                  object error, object cost, gradient norm, object norm
//                  (setup_born.cpp:239, opened by mom.cpp:220)
#include <cstdio>
#include <vector>

#include "../../include/isg.h"

extern "C" {

int isg_phantom_read_txt(const char* path, int numunk, double scale, double* raster_out) {
  if (!path || numunk < 1 || !raster_out) return ISG_E_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return ISG_E_ARG;
  int rc = ISG_OK;
  for (int n = 0; n < numunk; n++) {
    float v = 0;
    if (std::fscanf(f, "%f\n", &v) != 1) {
      rc = ISG_E_ARG;
      break;
    }
    raster_out[2 * n] = (double)v * scale;
    raster_out[2 * n + 1] = 0.0;
  }
  std::fclose(f);
  return rc;
}

int isg_phantom_write_txt(const char* path, int numunk, double scale, const double* raster_in) {
  if (!path || numunk < 1 || !raster_in || scale == 0.0) return ISG_E_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return ISG_E_ARG;
  bool ok = true;
  for (int n = 0; n < numunk && ok; n++) ok = std::fprintf(f, "%f\n", raster_in[2 * n] / scale) > 0;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? ISG_OK : ISG_E_ARG;
}

int isg_raster_bin_write(const char* path, int numunk, const double* raster) {
  if (!path || numunk < 1 || !raster) return ISG_E_ARG;
  FILE* f = std::fopen(path, "wb");
  if (!f) return ISG_E_ARG;
  bool ok = std::fwrite(raster, 2 * sizeof(double), numunk, f) == (size_t)numunk;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? ISG_OK : ISG_E_ARG;
}

int isg_raster_bin_read(const char* path, int numunk, double* raster) {
  if (!path || numunk < 1 || !raster) return ISG_E_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) return ISG_E_ARG;
  bool ok = std::fread(raster, 2 * sizeof(double), numunk, f) == (size_t)numunk;
  std::fclose(f);
  return ok ? ISG_OK : ISG_E_ARG;
}

int isg_hist_append(const char* path, const isg_dbim_stats* st, int create) {
  if (!path || !st) return ISG_E_ARG;
  FILE* f = std::fopen(path, create ? "w" : "a");
  if (!f) return ISG_E_ARG;
  bool ok = std::fprintf(f, "%e %e %e %e %e\n", st->misfit, st->object_error, st->object_cost, st->gradient_norm2,
                         st->object_norm2) > 0;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? ISG_OK : ISG_E_ARG;
}

}  // extern "C"

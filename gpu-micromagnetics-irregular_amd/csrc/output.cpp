// output.cpp -- on-disk formats of the reference's runs, so its plotters
// (src/mdyn2D.m, src/slit/plot_slit.py) read our output unchanged.
//   frames      2d_i4/2d_fft.cu:393-419 ("%f %d %d" header, one "%f %f %f"
//               line per cell in row-major order, blank line), windowed by w
//               for the polycrystal (2d_i3/2d_fft.cu:750-768); early/late
//               cadence ShouldWriteFrame
//   final state 2d_i2/2d_fft.cu:732-758 (the same frame with "%e")
//   aux files   mask.txt / grain_id.txt (2d_i3/2d_fft.cu:390-408)
#include <cstdio>
#include <cstdlib>

#include "llg/llg_output.h"

struct llg_frame_writer {
  FILE* fp;
};

extern "C" {

llg_frame_writer* LLG_FrameWriterOpen(const char* path) {
  FILE* fp = std::fopen(path, "w");
  if (!fp) return nullptr;
  std::setvbuf(fp, nullptr, _IOFBF, 1 << 20);
  llg_frame_writer* w = new llg_frame_writer{fp};
  return w;
}

int LLG_FrameWriterWrite(llg_frame_writer* w, double t, int nx, int ny, int ng,
                         const double* y, const double* weight, int sci) {
  if (!w || !w->fp || !y) return -1;
  const int ncell = ng * ny;
  std::fprintf(w->fp, "%f %d %d\n", t, nx, ny);
  const char* fmt = sci ? "%e %e %e\n" : "%f %f %f\n";
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < ng; i++) {
      const int c = j * ng + i;
      const double s = weight ? weight[c] : 1.0;
      std::fprintf(w->fp, fmt, s * y[c], s * y[ncell + c], s * y[2 * ncell + c]);
    }
  std::fprintf(w->fp, "\n");
  return 0;
}

int LLG_FrameWriterClose(llg_frame_writer* w) {
  if (!w) return -1;
  const int rc = w->fp ? std::fclose(w->fp) : 0;
  delete w;
  return rc;
}

int LLG_ShouldWriteFrame(long iout, double t, double early_until, int early_every,
                         int late_every) {
  if (t <= early_until) return early_every > 0 && (iout % early_every) == 0;
  return late_every > 0 && (iout % late_every) == 0;
}

int LLG_WriteMaskFile(const char* path, int ng, int ny, const double* w) {
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  std::fprintf(f, "%d %d\n", ng, ny);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < ng; i++) std::fprintf(f, "%f\n", w[j * ng + i]);
  return std::fclose(f);
}

int LLG_WriteGrainIdFile(const char* path, int ng, int ny, const int* grain_id) {
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  std::fprintf(f, "%d %d\n", ng, ny);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < ng; i++) std::fprintf(f, "%d\n", grain_id[j * ng + i]);
  return std::fclose(f);
}

}  // extern "C"

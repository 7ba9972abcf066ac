/*
 * llg_output.h -- the reference's text output formats.
 *
 * Replaces WriteFrame / ShouldWriteFrame (src/irregular/2d_fft/2d_i4/2d_fft.cu
 * :393-419; windowed form 2d_i3/2d_fft.cu:750-768), the final-state dump
 * (2d_i2/2d_fft.cu:732-758, sci = 1) and write_aux_files
 * (2d_i3/2d_fft.cu:390-408).  y is the host SoA state [mx|my|mz] of
 * ng*ny cells; nx is written in the header as the reference does
 * (NX_VAL = 3*ng in the reference's convention).
 */
#ifndef LLG_OUTPUT_H
#define LLG_OUTPUT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct llg_frame_writer llg_frame_writer;

llg_frame_writer* LLG_FrameWriterOpen(const char* path);
int LLG_FrameWriterWrite(llg_frame_writer* w, double t, int nx, int ny, int ng,
                         const double* y, const double* weight, int sci);
int LLG_FrameWriterClose(llg_frame_writer* w);
int LLG_ShouldWriteFrame(long iout, double t, double early_until, int early_every,
                         int late_every);
int LLG_WriteMaskFile(const char* path, int ng, int ny, const double* w);
int LLG_WriteGrainIdFile(const char* path, int ng, int ny, const int* grain_id);

#ifdef __cplusplus
}
#endif
#endif /* LLG_OUTPUT_H */

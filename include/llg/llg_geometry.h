/*
 * llg_geometry.h -- host builders for the geometry (active mask, compact
 * lists, polycrystal window/axes) and the initial magnetisation.
 *
 * Replaces the per-variant host code in main() (paths under
 * src/irregular/2d_fft/):
 *   LLG_GeomSquareHole   2d_i4/2d_fft.cu:497-523
 *   LLG_GeomRectangle    2d_i5/2d_fft.cu:516-535
 *   LLG_GeomEllipse      2d_i6/2d_i6.cu:462-533 (BuildEllipticalActiveLists)
 *   LLG_GeomPolycrystal  2d_i3/2d_fft.cu:253-387 (generate_grains,
 *                        nearest_grain, build_polycrystal)
 *   LLG_GeomRing         configs/sim_config.h:150-162 (GEOMETRY_RING knobs;
 *                        no variant implements it)
 *   LLG_ICTwoDomain      2d_i4/2d_fft.cu:567-610
 *   LLG_ICUniform        2d_i6/2d_i6.cu:671-711
 *   LLG_ICStripes        2d_i1/2d_fft.cu:625-657
 *   LLG_ICGrainBumps     2d_i3/2d_fft.cu:916-984
 *   LLG_ICTiltedZ / LLG_ICSplitX / LLG_ICRandomSphere: seeded initial states
 *                        of the plain 128x128, square-hole 1536x512 and
 *                        polycrystal workloads (no reference counterpart).
 * Arrays are cell-major (cell = j*ng + i); y and axis arrays are SoA
 * [x|y|z] planes of ng*ny.  Geometry builders return the active count.
 */
#ifndef LLG_GEOMETRY_H
#define LLG_GEOMETRY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct llg_grain {
  double sx, sy;       /* seed position (cells) */
  double ax, ay, az;   /* anisotropy axis */
  double nx, ny, nz;   /* DMI direction */
  int dead;
  int pad;
} llg_grain;

int LLG_GeomSquareHole(int ng, int ny, double cx_frac, double cy_frac, double halfside_frac,
                       unsigned char* active);
int LLG_GeomRectangle(int ng, int ny, double w_frac, double h_frac, unsigned char* active);
int LLG_GeomRing(int ng, int ny, double w_frac, double h_frac, double inner_w_frac,
                 double inner_h_frac, unsigned char* active);
int LLG_GeomEllipse(int ng, int ny, double rx_frac, double ry_frac, unsigned char* active);
/* grains[num_grains], grain_id[ng*ny]; active/w/msk/nsk may be NULL */
int LLG_GeomPolycrystal(int ng, int ny, int num_grains, double dead_frac, unsigned seed,
                        double z_bias, double eps_cells, llg_grain* grains, int* grain_id,
                        unsigned char* active, double* w, double* msk, double* nsk);

int LLG_ICUniform(int ng, int ny, const unsigned char* active, double mx, double my, double mz,
                  double* y);
int LLG_ICTwoDomain(int ng, int ny, const unsigned char* active, int j_split, double tail_my,
                    double* y);
int LLG_ICStripes(int ng, int ny, double left_frac, double right_frac, double eps,
                  unsigned seed, double* y);
int LLG_ICGrainBumps(int ng, int ny, const llg_grain* grains, const int* grain_id,
                     const double* w, double core_mz, unsigned seed, double* y);
int LLG_ICTiltedZ(int ng, int ny, const unsigned char* active, double max_angle, uint64_t seed,
                  double* y);
int LLG_ICSplitX(int ng, int ny, const unsigned char* active, double max_angle, uint64_t seed,
                 double* y);
int LLG_ICRandomSphere(int ng, int ny, const unsigned char* active, uint64_t seed, double* y);

#ifdef __cplusplus
}
#endif
#endif /* LLG_GEOMETRY_H */

/*
 * oph_oracle.h -- CPU oracle for the point-cloud hologram path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark load it, and only to check results / as the timed CPU baseline.
 *
 * Two independent restatements live here:
 *
 *  1. orc_gen_restate / orc_rs_diffraction / orc_fresnel_diffraction:
 *     the reference's OpenMP CPU path kept loop for loop (per-point parallel,
 *     omp-atomic FP64 adds, glibc libm), following
 *       ophgen/src/ophPointCloud.cpp:253-329  (driver, channel loop, scaling)
 *       ophgen/src/ophGen.cpp:361-445         (RS_Diffraction)
 *       ophgen/src/ophGen.cpp:447-514         (Fresnel_Diffraction)
 *     Every floating-point expression is evaluated in the same order as the
 *     reference so that the mask (box + RS range test) decisions are the
 *     same bits.  Build with -ffp-contract=off.
 *
 *  2. orc_field_exact: a pixel-parallel oracle that shares only the mask
 *     predicate with (1) and computes every term in x87 extended precision
 *     (64-bit mantissa) with the phase reduced in cycles and Neumaier
 *     compensated summation.  It is the "FP64 oracle" against which field
 *     relative-L2 errors are stated.
 *
 * plus the encoders (ophgen/src/ophGen.cpp:816-849, 1821-1864) and the 8-bit
 * normalisation (ophgen/src/ophGen.cpp:654-687).
 *
 * Parity pins (see tests/test_oracle_*.py): closed-form single-point fields
 * written as complex exponentials (numpy) and at 50 digits (mpmath golden
 * vectors), brute-force geometric mask predicates on tiny grids, linearity,
 * superposition and integer-pixel translation invariants, and known-answer
 * encodes.  Nothing here is "parity unpinned".
 */
#ifndef OPH_ORACLE_H
#define OPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 56-byte packed record as the reference's Vertex
 * (Openholo/src/struct.h:86-106): pos[3], color[3], phase, all double. */
typedef struct orc_vertex {
    double pos[3];
    double color[3];
    double phase;
} orc_vertex;

/* Hologram-plane geometry and point transform
 * (Openholo/src/Openholo.h:96-120 OphConfig, ophgen/src/ophGen.h:546-567). */
typedef struct orc_geom {
    int32_t pn_x, pn_y;     /* pixel_number                        */
    int32_t off_x, off_y;   /* offset (default 0)                  */
    double pp_x, pp_y;      /* pixel_pitch                         */
    double scale[3];        /* OphPointCloudConfig::scale          */
    double distance;        /* OphPointCloudConfig::distance       */
} orc_geom;

enum { ORC_DIFF_RS = 0, ORC_DIFF_FRESNEL = 1 };
enum { ORC_ENCODE_PHASE = 0, ORC_ENCODE_AMPLITUDE = 1, ORC_ENCODE_REAL = 2,
       ORC_ENCODE_IMAGINARY = 3 };

/* ---- (1) reference restatement ------------------------------------------ */

/* One point's RS contribution accumulated into field (interleaved re,im,
 * pn_x*pn_y complex).  pos is already scaled; distance is added inside,
 * as in ophGen.cpp:379. */
void orc_rs_diffraction(const orc_geom* g, const double pos[3], double* field,
                        double lambda, double distance, double amplitude);
void orc_fresnel_diffraction(const orc_geom* g, const double pos[3], double* field,
                             double lambda, double distance, double amplitude);

/* genCghPointCloudCPU: fields[ch] are zeroed by the caller (resetBuffer) and
 * accumulated here.  n_threads <= 0 keeps the OpenMP default. */
void orc_gen_restate(const orc_geom* g, const orc_vertex* v, int64_t n,
                     const double* lambdas, int n_ch, int diff_flag,
                     double** fields, int n_threads);

/* ---- (2) independent extended-precision oracle --------------------------- */

/* Field at an explicit list of pixels (pix[i] = row*pn_x + col).  When
 * use_phase != 0 the amplitude is color[ch]*exp(j*phase) (an extension; the
 * reference ignores Vertex::phase).  out is interleaved re,im. */
void orc_field_exact(const orc_geom* g, const orc_vertex* v, int64_t n,
                     double lambda, int channel, int diff_flag, int use_phase,
                     const int64_t* pix, int64_t n_pix, double* out, int n_threads);

/* Per-point mask membership as the reference decides it: 1 when point p
 * contributes to pixel (col,row).  Exposed for the brute-force mask tests. */
int orc_mask(const orc_geom* g, const double pos_scaled[3], double lambda,
             int diff_flag, int col, int row);

/* Integer box [x0,x1) x [y0,y1) of ophGen.cpp:384-407 / 467-490 (clamped). */
void orc_box(const orc_geom* g, const double pos_scaled[3], double lambda,
             int diff_flag, int64_t box[4]);

/* ---- encoders and 8-bit normalisation ------------------------------------ */

void orc_encode(const double* field, int64_t n_px, double* enc, int flag);
void orc_normalize(double* const* enc, int n_ch, int64_t n_px, uint8_t* const* out);

#ifdef __cplusplus
}
#endif
#endif

/*
 * oph_oracle.c -- CPU oracle for the point-cloud hologram path.
 * TEST INFRASTRUCTURE ONLY (see oph_oracle.h).  Never linked by the product.
 *
 * The restatement (section 1) follows the reference's BSD-2-licensed CPU path
 * (Copyright (C) 2017-2024, Korea Electronics Technology Institute), written
 * anew here in C; the expression order is kept so that mask decisions are
 * bit-identical.  Build flags: -O3 -fopenmp -ffp-contract=off (no -ffast-math).
 */
#define _GNU_SOURCE
#include "oph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Box bounds, identical arithmetic to ophGen.cpp:373-407 (RS) and
 * ophGen.cpp:463-490 (Fresnel).  Outputs the four doubles the reference
 * loops over: X in [xb[1], xb[0]), Y in [yb[1], yb[0]). */
static void ref_bounds(const orc_geom* g, const double pos[3], double lambda,
                       double distance, int diff_flag, double xb[2], double yb[2])
{
    const int pnX = g->pn_x, pnY = g->pn_y;
    const double ppX = g->pp_x, ppY = g->pp_y;
    const double ssX = pnX * ppX, ssY = pnY * ppY;
    const double x = -ssX / 2, y = -ssY / 2;
    const double z = pos[2] + distance;
    double hx, hy;
    if (diff_flag == ORC_DIFF_RS) {
        const double tx = lambda / (2 * ppX), ty = lambda / (2 * ppY);
        const double sqrtX = sqrt(1 - (tx * tx)), sqrtY = sqrt(1 - (ty * ty));
        hx = fabs(tx / sqrtX * z);
        hy = fabs(ty / sqrtY * z);
    } else {
        const double operand = lambda * z;
        hx = fabs(operand / (2 * ppX));
        hy = fabs(operand / (2 * ppY));
    }
    const double xbp = pos[0] + hx, xbm = pos[0] - hx;
    const double ybp = pos[1] + hy, ybm = pos[1] - hy;
    xb[0] = floor((xbp - x) / ppX) + 1;
    xb[1] = floor((xbm - x) / ppX) + 1;
    yb[0] = pnY - floor((ybm - y) / ppY);
    yb[1] = pnY - floor((ybp - y) / ppY);
    if (xb[0] > pnX) xb[0] = pnX;
    if (xb[1] < 0) xb[1] = 0;
    if (yb[0] > pnY) yb[0] = pnY;
    if (yb[1] < 0) yb[1] = 0;
}

void orc_box(const orc_geom* g, const double pos_scaled[3], double lambda,
             int diff_flag, int64_t box[4])
{
    double xb[2], yb[2];
    ref_bounds(g, pos_scaled, lambda, g->distance, diff_flag, xb, yb);
    /* integer set {c : xb[1] <= c < xb[0]} etc., as half-open int ranges */
    box[0] = (int64_t)ceil(xb[1]);
    box[1] = (int64_t)ceil(xb[0]);
    box[2] = (int64_t)ceil(yb[1]);
    box[3] = (int64_t)ceil(yb[0]);
    if (box[1] < box[0]) box[1] = box[0];
    if (box[3] < box[2]) box[3] = box[2];
}

/* RS per-pixel range test of ophGen.cpp:415-429 for pixel coordinates
 * (xxx, yyy) computed as the reference does. */
static int rs_range_ok(double px, double py, double zz, double tx_sx, double ty_sy,
                       double xxx, double yyy)
{
    const double rx0 = px + fabs(tx_sx * sqrt((yyy - py) * (yyy - py) + zz));
    const double rx1 = px - fabs(tx_sx * sqrt((yyy - py) * (yyy - py) + zz));
    const double ry0 = py + fabs(ty_sy * sqrt((xxx - px) * (xxx - px) + zz));
    const double ry1 = py - fabs(ty_sy * sqrt((xxx - px) * (xxx - px) + zz));
    return ((xxx < rx0) && (xxx > rx1)) && ((yyy < ry0) && (yyy > ry1));
}

int orc_mask(const orc_geom* g, const double pos[3], double lambda,
             int diff_flag, int col, int row)
{
    double xb[2], yb[2];
    ref_bounds(g, pos, lambda, g->distance, diff_flag, xb, yb);
    if (!((col >= xb[1]) && (col < xb[0]) && (row >= yb[1]) && (row < yb[0])))
        return 0;
    if (diff_flag != ORC_DIFF_RS) return 1;
    const int pnX = g->pn_x, pnY = g->pn_y;
    const double ppX = g->pp_x, ppY = g->pp_y;
    const double ssX = pnX * ppX, ssY = pnY * ppY;
    const double x = -ssX / 2, y = -ssY / 2;
    const double tx = lambda / (2 * ppX), ty = lambda / (2 * ppY);
    const double sqrtX = sqrt(1 - (tx * tx)), sqrtY = sqrt(1 - (ty * ty));
    const double z = pos[2] + g->distance;
    const double zz = z * z;
    const double yyy = y + ((pnY - row + g->off_y) * ppY);
    const double xxx = x + ((col - 1 + g->off_x) * ppX);
    return rs_range_ok(pos[0], pos[1], zz, tx / sqrtX, ty / sqrtY, xxx, yyy);
}

/* ------------------------------------------------------------------------ */
/* (1) restatement of ophGen::RS_Diffraction, ophGen.cpp:361-445 */
void orc_rs_diffraction(const orc_geom* g, const double pos[3], double* field,
                        double lambda, double distance, double amplitude)
{
    const int pnX = g->pn_x, pnY = g->pn_y;
    const double ppX = g->pp_x, ppY = g->pp_y;
    const double ssX = pnX * ppX, ssY = pnY * ppY;
    const int offsetX = g->off_x, offsetY = g->off_y;
    const double tx = lambda / (2 * ppX), ty = lambda / (2 * ppY);
    const double sqrtX = sqrt(1 - (tx * tx)), sqrtY = sqrt(1 - (ty * ty));
    const double x = -ssX / 2, y = -ssY / 2;
    const double k = (2 * M_PI) / lambda;
    const double z = pos[2] + distance;
    const double zz = z * z;
    const double ampZ = amplitude * z;

    double xb[2], yb[2];
    ref_bounds(g, pos, lambda, distance, ORC_DIFF_RS, xb, yb);
    if (!(xb[1] < xb[0]) || !(yb[1] < yb[0])) return;

    for (int yytr = (int)yb[1]; yytr < yb[0]; ++yytr) {
        const long long offset = (long long)yytr * pnX;
        const double yyy = y + ((pnY - yytr + offsetY) * ppY);
        const double rx0 = pos[0] + fabs(tx / sqrtX * sqrt((yyy - pos[1]) * (yyy - pos[1]) + zz));
        const double rx1 = pos[0] - fabs(tx / sqrtX * sqrt((yyy - pos[1]) * (yyy - pos[1]) + zz));
        for (int xxtr = (int)xb[1]; xxtr < xb[0]; ++xxtr) {
            const double xxx = x + ((xxtr - 1 + offsetX) * ppX);
            const double r = sqrt((xxx - pos[0]) * (xxx - pos[0]) +
                                  (yyy - pos[1]) * (yyy - pos[1]) + zz);
            const double ry0 = pos[1] + fabs(ty / sqrtY * sqrt((xxx - pos[0]) * (xxx - pos[0]) + zz));
            const double ry1 = pos[1] - fabs(ty / sqrtY * sqrt((xxx - pos[0]) * (xxx - pos[0]) + zz));
            if (((xxx < rx0) && (xxx > rx1)) && ((yyy < ry0) && (yyy > ry1))) {
                const double kr = k * r;
                const double operand = lambda * r * r;
                const double re = (ampZ * sin(kr)) / operand;
                const double im = (-ampZ * cos(kr)) / operand;
                double* d = field + 2 * (offset + xxtr);
#pragma omp atomic
                d[0] += re;
#pragma omp atomic
                d[1] += im;
            }
        }
    }
}

/* (1) restatement of ophGen::Fresnel_Diffraction, ophGen.cpp:447-514 */
void orc_fresnel_diffraction(const orc_geom* g, const double pos[3], double* field,
                             double lambda, double distance, double amplitude)
{
    const int pnX = g->pn_x, pnY = g->pn_y;
    const double ppX = g->pp_x, ppY = g->pp_y;
    const double ssX = pnX * ppX, ssY = pnY * ppY;
    const int offsetX = g->off_x, offsetY = g->off_y;
    const double k = (2 * M_PI) / lambda;
    const double x = -ssX / 2, y = -ssY / 2;
    const double z = pos[2] + distance;
    const double zz = z * z;
    const double operand = lambda * z;

    double xb[2], yb[2];
    ref_bounds(g, pos, lambda, distance, ORC_DIFF_FRESNEL, xb, yb);
    if (!(xb[1] < xb[0]) || !(yb[1] < yb[0])) return;

    for (int yytr = (int)yb[1]; yytr < yb[0]; ++yytr) {
        const double yyy = (y + (pnY - yytr + offsetY) * ppY) - pos[1];
        const long long offset = (long long)yytr * pnX;
        for (int xxtr = (int)xb[1]; xxtr < xb[0]; ++xxtr) {
            const double xxx = (x + (xxtr - 1 + offsetX) * ppX) - pos[0];
            const double p = k * (xxx * xxx + yyy * yyy + 2 * zz) / (2 * z);
            const double re = amplitude * sin(p) / operand;
            const double im = amplitude * (-cos(p)) / operand;
            double* d = field + 2 * (offset + xxtr);
#pragma omp atomic
            d[0] += re;
#pragma omp atomic
            d[1] += im;
        }
    }
}

/* (1) restatement of ophPointCloud::genCghPointCloudCPU,
 * ophPointCloud.cpp:253-329 (viewing-window transform off, as by default). */
void orc_gen_restate(const orc_geom* g, const orc_vertex* v, int64_t n,
                     const double* lambdas, int n_ch, int diff_flag,
                     double** fields, int n_threads)
{
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    for (int ch = 0; ch < n_ch; ++ch) {
        const double lambda = lambdas[ch];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double pc[3];
            pc[0] = v[i].pos[0] * g->scale[0];
            pc[1] = v[i].pos[1] * g->scale[1];
            pc[2] = v[i].pos[2] * g->scale[2];
            const double amplitude = v[i].color[ch];
            if (diff_flag == ORC_DIFF_RS)
                orc_rs_diffraction(g, pc, fields[ch], lambda, g->distance, amplitude);
            else
                orc_fresnel_diffraction(g, pc, fields[ch], lambda, g->distance, amplitude);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* (2) independent oracle.  Pixel-parallel; mask from the reference predicate
 * (the set of contributing pairs is part of the algorithm's definition), but
 * every term is computed in long double with the phase reduced in cycles and
 * summed with Neumaier compensation. */
typedef long double ld;

static inline void neumaier(ld* s, ld* c, ld x)
{
    const ld t = *s + x;
    if (fabsl(*s) >= fabsl(x)) *c += (*s - t) + x;
    else *c += (x - t) + *s;
    *s = t;
}

void orc_field_exact(const orc_geom* g, const orc_vertex* v, int64_t n,
                     double lambda, int channel, int diff_flag, int use_phase,
                     const int64_t* pix, int64_t n_pix, double* out, int n_threads)
{
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    const int pnX = g->pn_x, pnY = g->pn_y;
    const ld L = (ld)lambda;
    const ld two_pi = 6.283185307179586476925286766559005768L;
    const ld ssx_half = (ld)pnX * (ld)g->pp_x / 2, ssy_half = (ld)pnY * (ld)g->pp_y / 2;

#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t q = 0; q < n_pix; ++q) {
        const int col = (int)(pix[q] % pnX);
        const int row = (int)(pix[q] / pnX);
        /* exact pixel coordinates of the reference grid (ophGen.cpp:413,422) */
        const ld X = -ssx_half + (ld)(col - 1 + g->off_x) * (ld)g->pp_x;
        const ld Y = -ssy_half + (ld)(pnY - row + g->off_y) * (ld)g->pp_y;
        ld sr = 0, cr = 0, si = 0, ci = 0;
        for (int64_t i = 0; i < n; ++i) {
            double pc[3];
            pc[0] = v[i].pos[0] * g->scale[0];
            pc[1] = v[i].pos[1] * g->scale[1];
            pc[2] = v[i].pos[2] * g->scale[2];
            if (!orc_mask(g, pc, lambda, diff_flag, col, row)) continue;
            const ld z = (ld)pc[2] + (ld)g->distance;
            const ld dx = X - (ld)pc[0], dy = Y - (ld)pc[1];
            const ld rho2 = dx * dx + dy * dy;
            const ld a = (ld)v[i].color[channel];
            ld cyc, amp;
            if (diff_flag == ORC_DIFF_RS) {
                const ld r = sqrtl(rho2 + z * z);
                /* r/lambda = z/lambda + (r - z)/lambda, r - z = rho^2/(r + z) */
                cyc = z / L + (rho2 / (r + z)) / L;
                amp = a * z / (L * r * r);
            } else {
                cyc = rho2 / (2 * z * L) + z / L;
                amp = a / (L * z);
            }
            if (use_phase) cyc += (ld)v[i].phase / two_pi;
            cyc -= floorl(cyc);
            const ld ph = two_pi * cyc;
            /* (1/j) * amp * exp(j ph) = amp * (sin ph - j cos ph) */
            neumaier(&sr, &cr, amp * sinl(ph));
            neumaier(&si, &ci, -amp * cosl(ph));
        }
        out[2 * q] = (double)(sr + cr);
        out[2 * q + 1] = (double)(si + ci);
    }
}

/* ------------------------------------------------------------------------ */
/* Encoders, ophGen.cpp:1821-1864 (Phase: (atan2(im,re)+pi)/(2pi)). */
void orc_encode(const double* field, int64_t n_px, double* enc, int flag)
{
    const double pi2 = M_PI * 2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_px; ++i) {
        const double re = field[2 * i], im = field[2 * i + 1];
        switch (flag) {
        case ORC_ENCODE_PHASE: enc[i] = (atan2(im, re) + M_PI) / pi2; break;
        case ORC_ENCODE_AMPLITUDE: enc[i] = sqrt(re * re + im * im); break;
        case ORC_ENCODE_REAL: enc[i] = re; break;
        default: enc[i] = im; break;
        }
    }
}

/* ophGen::normalize, ophGen.cpp:654-687 with minOfArr/maxOfArr
 * (Openholo/src/function.h:79-99): global min/max over all channels, then
 * (uchar)(((v - min) / gap) * 255 + 0.5). */
void orc_normalize(double* const* enc, int n_ch, int64_t n_px, uint8_t* const* out)
{
    double minVal = 1.7976931348623158e+308, maxVal = 2.2250738585072014e-308;
    for (int ch = 0; ch < n_ch; ++ch) {
        double mn = 1.7976931348623158e+308, mx = 2.2250738585072014e-308;
        for (int64_t i = 0; i < n_px; ++i) {
            if (enc[ch][i] < mn) mn = enc[ch][i];
        }
        for (int64_t i = 0; i < n_px; ++i) {
            if (enc[ch][i] > mx) mx = enc[ch][i];
        }
        if (minVal > mn) minVal = mn;
        if (maxVal < mx) maxVal = mx;
    }
    const double gap = maxVal - minVal;
    for (int ch = 0; ch < n_ch; ++ch) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n_px; ++i)
            out[ch][i] = (uint8_t)(((enc[ch][i] - minVal) / gap) * 255 + 0.5);
    }
}

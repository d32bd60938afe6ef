// mc_cel.h - the n-level solve of a collisionally excited model ion
// (equilibrium, emission_mod.f90:1913-2114, with luSlv/lured/reslv
// 2119-2190), shared by the balance, line-luminosity and re-emission
// kernels.  One 64-lane wavefront works on one ion; rows and pivot updates
// are spread across the lanes, the back substitution runs on lane 0.
#pragma once
#include <hip/hip_runtime.h>

#include "mc_balance.h"

#pragma clang fp contract(off)

namespace mcb {

constexpr int kCelW = 64;

// natural cubic spline evaluation with precomputed second derivatives
// (splint, interpolation_mod.f90:268-316)
__device__ __forceinline__ double splint(const double *xa, const double *ya, const double *y2, int n, double x)
{
    if (x <= xa[0]) return ya[0];
    if (x >= xa[n - 1]) return ya[n - 1];
    int klo = 1, khi = n;
    while (khi - klo > 1) {
        int k = (khi + klo) / 2;
        if (xa[k - 1] > x) khi = k; else klo = k;
    }
    double h = xa[khi - 1] - xa[klo - 1];
    double a = (xa[khi - 1] - x) / h;
    double b = (x - xa[klo - 1]) / h;
    return a * ya[klo - 1] + b * ya[khi - 1] +
           ((a * a * a - a) * y2[klo - 1] + (b * b * b - b) * y2[khi - 1]) * (h * h) / 6.;
}

// Level populations of one model ion (equilibrium, emission_mod.f90:
// 1913-2114, with luSlv): S.y[1..nl] unnormalised, valid on every lane on
// return.  up = ionDen(ion+1)/ionDen(ion), not NaN.  S: any LDS block with
// the x, q, y and alev arrays of Lds.
template <class LdsT>
__device__ __forceinline__ void cel_levels(const Atomic &A, const CelIon &ci, LdsT &S, int lane, double Te,
                                           double Ne, double up)
{
    const int nl = ci.nl, nt = ci.nt;
    const double *E = A.cel_e + ci.off_e;
    const double *Aa = A.cel_a + ci.off_a;          // a(i,j) at [(i-1) + nl*(j-1)]
    const double *lT = A.cel_t + ci.off_t;
    const double *gg = A.cel_g + ci.off_g;
    const double sqrTe = sqrt(Te);
    const double log10Te = log10(Te);
    // level recombination coefficients (emission_mod.f90:1991-2002)
    for (int j = lane; j <= kMaxLev; j += kCelW) S.alev[j] = 0.0;
    __syncthreads();
    if (ci.br > 0.0) {
        for (int j = 2 + lane; j <= nl; j += kCelW) {
            double t4 = Te * 1.e-4;
            double zz = ci.z;
            double x1 = t4 / (zz * zz);
            S.alev[j] = 1.e-13 * ci.br * zz * ci.a_r[0] * pow(x1, ci.a_r[1]) /
                            (1. + ci.a_r[2] * pow(t4 / (zz * zz), ci.a_r[3])) +
                        1.0e-12 * ci.br *
                            (ci.a_d[0] / t4 + ci.a_d[1] + ci.a_d[2] * Te * 1.e-4 + ci.a_d[3] * (t4 * t4)) *
                            pow(t4, -1.5) * exp(-(ci.a_d[4] * 1.e4 / Te));
        }
        if (lane == 0 && nl >= 2) S.alev[1] = 1.0;
    }
    // collision rates q(lo,up) excitation, q(up,lo) de-excitation (2004-2045)
    const int npair = nl * (nl - 1) / 2;
    for (int p = lane; p < npair; p += kCelW) {
        int lo = 1, rem = p;
        while (rem >= nl - lo) { rem -= nl - lo; ++lo; }
        const int upL = lo + 1 + rem;
        const double *qq = A.cel_q + ci.off_q + (size_t)p * nt;
        const double *y2 = A.cel_y2 + ci.off_q + (size_t)p * nt;
        double cs;
        if (nt == 1) cs = qq[0];
        else if (nt == 2) cs = qq[0] + (qq[1] - qq[0]) / (lT[1] - lT[0]) * (log10Te - lT[0]);
        else cs = splint(lT, qq, y2, nt, log10Te);
        const double delTeK = (E[lo - 1] - E[upL - 1]) * 1.4388463;
        const double expFac = exp(delTeK / Te);
        S.q[lo][upL] = 8.63e-6 * cs * expFac / (gg[lo - 1] * sqrTe);
        S.q[upL][lo] = 8.63e-6 * cs / (gg[upL - 1] * sqrTe);
    }
    __syncthreads();
    // rate matrix rows 2..nl in the reference's j order, row 1 = ones
    for (int i = 2 + lane; i <= nl; i += kCelW) {
        for (int j = 1; j <= nl; ++j) S.x[i][j] = 0.0;
        for (int j = 1; j <= nl; ++j) {
            if (j != i) {
                S.x[i][j] = S.x[i][j] + Ne * S.q[j][i];
                S.x[i][i] = S.x[i][i] - Ne * S.q[i][j];
                if (j > i) S.x[i][j] = S.x[i][j] + Aa[(j - 1) + nl * (i - 1)];
                else S.x[i][i] = S.x[i][i] - Aa[(i - 1) + nl * (j - 1)];
            }
        }
        S.y[i] = -(Ne * up * S.alev[i]);
    }
    if (lane == 0) {
        for (int j = 1; j <= nl; ++j) S.x[1][j] = 1.0;
        S.y[1] = 1.0;
    }
    __syncthreads();
    // lured: Gaussian elimination without pivoting.  Each pivot step updates
    // the (nl-i)^2 entries below and right of the pivot, one entry per lane
    // at a time; every entry gets the reference's factor and operation, and
    // column i (the multipliers' numerators reslv reads) is left as it is.
    for (int i = 1; i <= nl - 1; ++i) {
        const int m = nl - i;
        for (int e = lane; e < m * m; e += kCelW) {
            const int k = i + 1 + e / m, j = i + 1 + e % m;
            const double factor = S.x[k][i] / S.x[i][i];
            S.x[k][j] = S.x[k][j] - S.x[i][j] * factor;
        }
        __syncthreads();
    }
    // reslv: forward substitution across lanes, back substitution on lane 0
    for (int i = 1; i <= nl - 1; ++i) {
        const double yi = S.y[i], xii = S.x[i][i];
        for (int j = i + 1 + lane; j <= nl; j += kCelW) S.y[j] = S.y[j] - yi * S.x[j][i] / xii;
        __syncthreads();
    }
    if (lane == 0) {
        S.y[nl] = S.y[nl] / S.x[nl][nl];
        for (int i = 1; i <= nl - 1; ++i) {
            const int k = nl - i;
            for (int j = k + 1; j <= nl; ++j) S.y[k] = S.y[k] - S.y[j] * S.x[k][j];
            S.y[k] = S.y[k] / S.x[k][k];
        }
    }
    __syncthreads();
}

}  // namespace mcb

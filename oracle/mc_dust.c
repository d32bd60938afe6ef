/*
 * mc_dust.c - CPU restatement of MOCASSIN's grain temperatures (getDustT,
 * update_mod.f90:1836-1945, with locate, interpolation_mod.f90:48-81), for
 * one dust chemistry without the photoelectric and resonance-line terms.
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU dust kernel.  Compiled
 * twice: in the reference's FP32 (mco_dust_temps_f32) and in FP64
 * (mco_dust_temps_f64).
 *
 * Per reached cell: radField = Jste/pi; for every species nS and size ai
 * the absorption integral sum_i sigma_abs(i) radField(i) is located in the
 * emission integrals E(T), T = 1..nTemps K, and interpolated linearly;
 * Tdust(nS,0) = sum_ai Tdust(nS,ai) weight(ai), Tdust(0,0) = sum_nS
 * Tdust(nS,0) abun(nS).  Unreached cells keep their temperatures.  With
 * the photoelectric effect (gdc, peh non-null, [cell][nS][ai]) the
 * absorption integral gains gasDustColl_d - photoelHeat_d of the cell's last
 * temperature iteration (update_mod.f90:1886-1889).
 */
#include <stdint.h>
#include <string.h>

#define PI_REF 3.141592654f     /* constants_mod.f90:32, a REAL */

/* locate(xa, x) for xa(1:n): n if x > xa(n), 0 if x < xa(1), else
 * max(minloc(xa - x, mask = xa - x > 0) - 1, 1), minloc giving 0 when no
 * element is above x (so x == xa(n) yields 1) */
#define LOCATE(R)                                                              \
    static int locate_##R(const R *xa, int n, R x)                             \
    {                                                                          \
        if (x > xa[n - 1]) return n;                                           \
        if (x < xa[0]) return 0;                                               \
        int best = 0;                                                          \
        R bd = 0;                                                              \
        for (int i = 0; i < n; ++i) {                                          \
            R d = xa[i] - x;                                                   \
            if (d > 0 && (best == 0 || d < bd)) { best = i + 1; bd = d; }      \
        }                                                                      \
        return best - 1 > 1 ? best - 1 : 1;                                    \
    }
LOCATE(float)
LOCATE(double)

/* J: scaled Jste, (0:nCells, 1:nbins) column-major (cell fastest).
 * Tdust: (0:nsMax, 0:nSizes, 0:nCells) column-major, in/out.
 * absx: [nS][ai][nbins]; em: (nSpecies, nSizes, nTemps) column-major. */
#define DUST(R, NAME)                                                                          \
    void NAME(const R *J, int nc, int nb, int nS, int nsMax, int nSizes, int nTemps,          \
              const float *absx, const float *weight, const float *abun, const float *em,     \
              int nSpeciesEm, const R *gdc, const R *peh, R *Tdust)                            \
    {                                                                                          \
        R xa[4096];                                                                            \
        if (nTemps > 4096) return;                                                             \
        const size_t tc = (size_t)(nsMax + 1) * (nSizes + 1);                                  \
        for (int c = 1; c <= nc; ++c) {                                                        \
            int hit = 0;                                                                       \
            for (int i = 0; i < nb; ++i)                                                       \
                if (J[c + (size_t)(nc + 1) * i] > 0) { hit = 1; break; }                       \
            if (!hit) continue;                                                                \
            R *T = Tdust + (size_t)c * tc;                                                     \
            for (size_t k = 0; k < tc; ++k) T[k] = 0;                                          \
            for (int s = 1; s <= nS; ++s) {                                                    \
                for (int ai = 1; ai <= nSizes; ++ai) {                                         \
                    R integ = 0;                                                               \
                    const float *sx = absx + ((size_t)(s - 1) * nSizes + (ai - 1)) * nb;       \
                    for (int i = 0; i < nb; ++i) {                                             \
                        R rf = J[c + (size_t)(nc + 1) * i] / (R)PI_REF;                        \
                        integ = integ + (R)sx[i] * rf;                                         \
                    }                                                                          \
                    if (gdc) {                                                                 \
                        const size_t q = (size_t)c * nS * nSizes + (size_t)(s - 1) * nSizes + ai - 1; \
                        integ = integ + gdc[q] - peh[q];                                       \
                    }                                                                          \
                    for (int t = 0; t < nTemps; ++t)                                           \
                        xa[t] = (R)em[(s - 1) + (size_t)nSpeciesEm * ((ai - 1) + (size_t)nSizes * t)]; \
                    int iT = locate_##R(xa, nTemps, integ);                                    \
                    R td;                                                                      \
                    if (iT <= 0) td = (R)1;                                                    \
                    else if (iT >= nTemps) td = (R)nTemps;                                     \
                    else td = (R)iT + (integ - xa[iT - 1]) * ((R)(iT + 1) - (R)iT) / (xa[iT] - xa[iT - 1]); \
                    T[s + (size_t)(nsMax + 1) * ai] = td;                                      \
                    T[s] = T[s] + td * (R)weight[ai - 1];                                      \
                }                                                                              \
                T[0] = T[0] + T[s] * (R)abun[s - 1];                                           \
            }                                                                                  \
        }                                                                                      \
    }
DUST(float, mco_dust_temps_f32)
DUST(double, mco_dust_temps_f64)

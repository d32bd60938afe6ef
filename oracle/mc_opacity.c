/*
 * mc_opacity.c - CPU restatement, in the reference's FP32, of MOCASSIN's
 * opacity tables (iterateMC's table step, iteration_mod.f90:117-228):
 * ionizationDriver per cell (ionization_mod.f90:26-128) with BoltGaunt,
 * getGauntFF (Hummer 1988), eDenSum and addOpacity (130-484), then the dust
 * add (iteration_mod.f90:164-228).
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU opacity kernel.
 *
 * Two ways to evaluate the free-free Gaunt factor:
 *   gaunt_cache = 1: as the reference does, recomputed only when Te moved by
 *     more than 10 % since the last recomputation (BoltGaunt's SAVEd TeOld2,
 *     ionization_mod.f90:164-170), so cells are processed in iterateMC's
 *     order and the result depends on it;
 *   gaunt_cache = 0: at every cell's own Te (the GPU's choice).
 * Either way only bin 1 carries free-free opacity: addOpacity assigns
 * FFOpacity(i) at i = 1 only (ionization_mod.f90:369-393).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define TE1RYD 1.578866e5f     /* constants_mod.f90:23 */

/* Hummer's Chebyshev coefficients d(8,11) (ionization_mod.f90:213-242) */
static const double HUMMER_D[88] = {
    8.986940175e+00, -4.009515855e+00, 8.808871266e-01, 2.640245111e-02, -4.580645915e-02,
    -3.568055702e-03, 2.827798067e-03, 3.365860195e-04, -8.006936989e-01, 9.466021705e-01,
    9.043402532e-02, -9.608451450e-02, -1.885629865e-02, 1.050313890e-02, 2.800889961e-03,
    -1.078209202e-03, -3.781305103e-01, 1.102726332e-01, -1.543619180e-02, 8.310561114e-03,
    2.179620525e-02, 4.259726289e-03, -4.181588794e-03, -1.770208330e-03, 1.877213132e-02,
    -1.004885705e-01, -5.483366378e-02, -4.520154409e-03, 8.366530426e-03, 3.700273930e-03,
    6.889320423e-04, 9.460313195e-05, 7.300158392e-02, 3.576785497e-03, -4.545307025e-03,
    -1.017965604e-02, -9.530211924e-03, -3.450186162e-03, 1.040482914e-03, 1.407073544e-03,
    -1.744671550e-03, 2.864013856e-02, 1.903394837e-02, 7.091074494e-03, -9.668371391e-04,
    -2.999107465e-03, -1.820642230e-03, -3.874082085e-04, -1.707268366e-02, -4.694254776e-03,
    1.311691517e-03, 5.316703136e-03, 5.178193095e-03, 2.451228935e-03, -2.277321615e-05,
    -8.182359057e-04, 2.567331664e-04, -9.155339970e-03, -6.997479192e-03, -3.571518641e-03,
    -2.096101038e-04, 1.553822487e-03, 1.509584686e-03, 6.212627837e-04, 4.098322531e-03,
    1.635218463e-03, -5.918883504e-04, -2.333091048e-03, -2.484138313e-03, -1.359996060e-03,
    -5.371426147e-05, 5.553549563e-04, 3.837562402e-05, 2.938325230e-03, 2.393747064e-03,
    1.328839809e-03, 9.135013312e-05, -7.137252303e-04, -7.656848158e-04, -3.504683798e-04,
    -8.491991820e-04, -3.615327726e-04, 3.148015257e-04, 8.909207650e-04, 9.869737522e-04,
    6.134671184e-04, 1.068883394e-04, -2.046080100e-04};

/* getGauntFF at one frequency: z = log10 of the nuclear charge, xlf =
 * log10(h nu / Ryd); g_prev is returned where no branch assigns g */
float mco_gaunt_ff(float z, float log10Te, float xlf, float g_prev)
{
#define D(j, i) HUMMER_D[((i) - 1) * 8 + (j) - 1]
    float xlrkt;
    if (log10Te < 2.5f) xlrkt = 5.1983649f - 2.5f;
    else if (log10Te > 8.f) xlrkt = 5.1983649f - 8.f;
    else xlrkt = 5.1983649f - log10Te;
    const float txg = 0.66666667f * (2.0f * z + xlrkt);
    const float gamma2 = powf(10.f, txg * 1.5f);
    const float con = 0.72727273f * xlrkt + 0.90909091f;
    float b[12], c[9];
    for (int j = 1; j <= 8; ++j) {
        int ir = 9;
        b[11] = (float)D(j, 11);
        b[10] = (float)((double)(txg * b[11]) + D(j, 10));
        for (int i = 1; i <= 9; ++i) {
            b[ir] = (float)((double)(txg * b[ir + 1] - b[ir + 2]) + D(j, ir));
            ir = ir - 1;
        }
        c[j] = 0.25f * (b[1] - b[3]);
    }
    const float txu = 0.72727273f * xlf + con;
    const float u = powf(10.f, (txu - .90909091f) / .72727273f);
    float g = g_prev;
    if (fabsf(txu) <= 2.0f) {
        int ir = 6;
        b[8] = c[8];
        b[7] = txu * b[8] + c[7];
        for (int j = 1; j <= 6; ++j) {
            b[ir] = txu * b[ir + 1] - b[ir + 2] + c[ir];
            ir = ir - 1;
        }
        g = b[1] - b[3];
    } else if (log10f(u) < -4.0f && gamma2 < 0.3f) {
        g = 0.551329f * logf(2.24593f / u);
    } else if (log10f(u) < -4.0f && gamma2 >= 0.3f) {
        g = 0.551329f * logf(0.944931f / u / sqrtf(gamma2));
    } else if (txu > 2.0f) {
        if (log10f(gamma2) < -3.f) {
            g = sqrtf(0.9549297f / u);
        } else if (log10f(gamma2) < 0.0f) {
            const float slope = (sqrtf(12.f * gamma2 / u) - sqrtf(0.9549297f / u)) / 3.0f;
            g = sqrtf(12.f * gamma2 / u) + slope * log10f(gamma2);
        } else {
            g = sqrtf(12.f * gamma2 / u);
            g = 1.05f < g ? 1.05f : g;
        }
    }
    return g;
#undef D
}

/* Gas opacity of every active cell, in FP32, cells in order 1..nc (the order
 * ionizationDriver meets them in iterateMC when the cell numbering follows
 * its loop).  Layouts: ionDen (0:nc, 1:nElemUsed, 1:nstages) and opacity
 * (0:nc, 1:nbins), column-major; elem_on/xref/abun by atomic number - 1;
 * seg_meta [nseg][6] (elem, ion, shell, nuLowP, nuHighP, xSecP) in
 * addOpacity's order, with sigma of segment s at seg_sigma[seg_off[s] + i -
 * nuLowP] for bins nuLowP..max(nuLowP, min(nuHighP, nbins)).
 * Te_cache: the Te of BoltGaunt's last recomputation before cell 1 (<= 0:
 * none); updated. */
void mco_opacity_f32(int nc, int nb, int nstages, int nElemUsed, const float *nu, const float *Hden,
                     const float *Te, const float *Ne, const float *ionDen, const int32_t *elem_on,
                     const int32_t *xref, const float *abun, int nseg, const int32_t *seg_meta,
                     const float *seg_sigma, const int64_t *seg_off, float brems1, int gaunt_cache,
                     float *Te_cache, float *gaunt_cache_val, float *opacity)
{
    const size_t nc1 = (size_t)nc + 1;
    float TeOld2 = *Te_cache, g1 = *gaunt_cache_val;
    const float xlf1 = log10f(nu[0]);
    for (int cell = 1; cell <= nc; ++cell) {
        float dens[31][10];
        memset(dens, 0, sizeof dens);
        const float T = Te[cell], N = Ne[cell], H = Hden[cell];
        for (int n = 1; n <= 30; ++n) {
            const int top = n < nstages ? n : nstages;
            for (int i = 1; i <= top; ++i) {
                if (!elem_on[n - 1]) break;
                dens[n][i] = ionDen[cell + nc1 * ((size_t)(xref[n - 1] - 1) + (size_t)nElemUsed * (i - 1))] *
                             abun[n - 1] * H;
            }
        }
        const float log10Te = log10f(T);
        const float sqrTe = sqrtf(T);
        const float expo = (TE1RYD / T) + nu[0];
        const float cb1 = expo > 84.f ? 0.f : expf(-expo);
        if (!gaunt_cache || fabsf(1.f - TeOld2 / T) > 0.10f) {
            g1 = mco_gaunt_ff(0.f, log10Te, xlf1, g1);
            TeOld2 = T;
        }
        float eSum = 0.f;
        for (int n = 3; n <= 30; ++n) {
            const int top = n < nstages - 1 ? n : nstages - 1;
            for (int i = 1; i <= top; ++i) eSum = eSum + (float)(i * i) * dens[n][i + 1];
        }
        const float fac1 = (N / 1.e10f) * ((dens[1][2] + dens[2][2] + 4.f * dens[2][3] + eSum) / 1.e10f) / sqrTe;
        const float fac2 = fac1 * TE1RYD / T;
        float ff1;
        if (cb1 < 0.995f) ff1 = (brems1 * (1.f - cb1) * g1) * fac1;
        else ff1 = (brems1 * nu[0] * g1) * fac2;
        for (int i = 1; i <= nb; ++i) opacity[cell + nc1 * (i - 1)] = 0.f + (i == 1 ? ff1 : 0.f);
        for (int s = 0; s < nseg; ++s) {
            const int32_t *m = seg_meta + 6 * s;
            const int e = m[0], ion = m[1], lo = m[3];
            if (e > 2 && !(dens[e][ion] > 0.f)) continue;
            int iup = m[4] < nb ? m[4] : nb;
            iup = iup > lo ? iup : lo;
            const float den = dens[e][ion];
            const float *sg = seg_sigma + seg_off[s];
            for (int i = lo; i <= iup; ++i) {
                float *o = opacity + cell + nc1 * (i - 1);
                *o = *o + sg[i - lo] * den;
            }
        }
    }
    *Te_cache = TeOld2;
    *gaunt_cache_val = g1;
}

/* The dust add (iteration_mod.f90:184-222), one chemistry, FP32: grains of
 * (species, size) below their sublimation temperature add grainAbun *
 * grainWeight * Ndust * sigma to scaOpac/absOpac, and opacity gains
 * scaOpac + absOpac.  absx/scax [nS][nSizes][nb]; Tdust (0:nsMax, 0:nSizes,
 * 0:nc); sca, opacity (0:nc, 1:nb) (sca zeroed here). */
void mco_dust_opacity_f32(int nc, int nb, int nS, int nsMax, int nSizes, const float *absx, const float *scax,
                          const float *weight, const float *gabun, const float *tsub, const float *Ndust,
                          const float *Tdust, float *sca, float *opacity)
{
    const size_t nc1 = (size_t)nc + 1;
    float ab[4096];
    if (nb > 4096) return;
    for (int cell = 1; cell <= nc; ++cell) {
        for (int f = 0; f < nb; ++f) { sca[cell + nc1 * f] = 0.f; ab[f] = 0.f; }
        for (int s = 1; s <= nS; ++s)
            for (int ai = 1; ai <= nSizes; ++ai) {
                const float td = Tdust[s + (size_t)(nsMax + 1) * (ai + (size_t)(nSizes + 1) * cell)];
                if (!(td < tsub[s - 1])) continue;
                const float w = gabun[s - 1] * weight[ai - 1] * Ndust[cell];
                const float *xs = scax + ((size_t)(s - 1) * nSizes + ai - 1) * nb;
                const float *xa = absx + ((size_t)(s - 1) * nSizes + ai - 1) * nb;
                for (int f = 0; f < nb; ++f) {
                    sca[cell + nc1 * f] = sca[cell + nc1 * f] + w * xs[f];
                    ab[f] = ab[f] + w * xa[f];
                }
            }
        for (int f = 0; f < nb; ++f)
            opacity[cell + nc1 * f] = opacity[cell + nc1 * f] + (sca[cell + nc1 * f] + ab[f]);
    }
}

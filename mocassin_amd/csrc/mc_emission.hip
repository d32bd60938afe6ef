// mc_emission.hip - the re-emission tables on MI355X: emissionDriver's gas
// part and setDiffusePDF (emission_mod.f90:32-1311) for every active cell,
// written into the transport's recPDF rows ([cell][nu]) and totalLines.
//
// One 64-lane wavefront per cell.  Lanes over frequency bins evaluate the
// free-bound, free-free and two-photon coefficients and the cell's partial
// sums (sumDiffuseHI/HeI/HeII and the dust term) into LDS; lanes over lines
// evaluate the recombination and forbidden-line terms.  The sums the
// reference forms in REAL, in a fixed order (normHI, normRec, normFor, the
// cumulative recPDF), are then taken by lane 0 in that order, so the tables
// follow the reference's rounding.  The level populations of the forbidden
// lines come from the balance's n-level solve (FP64).
#include "mc_cel.h"
#include "mc_emission.h"

#pragma clang fp contract(off)

namespace mce {

constexpr int W = 64;
constexpr float kHcRydK = 157893.94f;   // continuum_mod.f90:10
constexpr float kHcRyd = 2.1799153e-11f; // continuum_mod.f90:11
constexpr float kCRyd = 3.2898423e15f;   // continuum_mod.f90:12
constexpr float kHPlanck = 6.6262e-27f;  // constants_mod.f90:14
constexpr float kPi = 3.141592654f;      // constants_mod.f90:32
constexpr int kLev = mcb::kMaxLev;

struct CelLds {
    double x[kLev + 1][kLev + 1];
    double q[kLev + 1][kLev + 1];
    double y[kLev + 1];
    double alev[kLev + 1];
};

// One dynamic LDS block per wavefront.  Region A serves the forbidden-line
// solves first (CelLds and the terms of one model ion) and then the per-bin
// partial sums s1, s2, s3 (and the grains' sD); the log-gamma rows, the
// recombination-line terms and a reduction slot follow it.  Sharing region
// A keeps the block at about 12 KB on an 800-bin mesh, so that more
// wavefronts fit on a CU.
__host__ __device__ inline size_t region_a_bytes(int nb, bool dust)
{
    const size_t cel = sizeof(CelLds) + sizeof(float) * kLev * kLev;
    const size_t sums = (size_t)(dust ? 4 : 3) * nb * sizeof(float);
    return ((cel > sums ? cel : sums) + 15) / 16 * 16;
}

// hydro2phot (emission_mod.f90:553-663), REAL arithmetic
__device__ double hydro2phot(float nu, int Z, float Te, float Ne)
{
    float nu0, y, alpha, Ay, gNu, factor;
    if (Z == 1) {
        nu0 = 0.7496f;
        y = nu / nu0;
        if (nu >= nu0) return 0.0;
        alpha = 0.8368f * powf(Te * 1.e-4f, -0.723f);
        Ay = 202.0f * (y * (1.f - y) * (1.f - powf(4.f * y * (1.f - y), 0.8f)) +
                       0.88f * powf(y * (1.f - y), 1.53f) * powf(4.f * y * (1.f - y), 0.8f));
        gNu = kHPlanck * y * Ay / 8.2249f;
        gNu = gNu * 1e27f;
        float Q;
        if (Te <= 10000.f) Q = 5.31e-4f;
        else if (Te >= 20000.f) Q = 4.71e-4f;
        else {
            factor = log10f(4.71e-4f / 5.31e-4f) / log10f(2.f);
            Q = powf(10.f, log10f(5.31e-4f) + factor * log10f(Te / 10000.f));
        }
        return (double)(alpha * gNu / (1.f + (Ne * Q) / 8.23f));
    }
    nu0 = 3.00f;
    y = nu / nu0;
    if (nu >= nu0) return 0.0;
    if (Te <= 5000.f) alpha = 6.161f;
    else if (Te >= 30000.f) alpha = 2.035f;
    else if (Te > 5000.f && Te <= 10000.f) {
        factor = log10f(4.091f / 6.161f) / log10f(2.f);
        alpha = powf(10.f, log10f(6.161f) + factor * log10f(Te / 5000.f));
    } else if (Te > 10000.f && Te <= 15000.f) {
        factor = log10f(3.189f / 4.091f) / log10f(1.5f);
        alpha = powf(10.f, log10f(4.091f) + factor * log10f(Te / 10000.f));
    } else {
        factor = log10f(2.035f / 3.189f) / log10f(2.f);
        alpha = powf(10.f, log10f(3.189f) + factor * log10f(Te / 15000.f));
    }
    const float z6 = 64.f;
    Ay = z6 * 0.9994667f * 202.0f *
         (y * (1.f - y) * (1.f - powf(4.f * y * (1.f - y), 0.8f)) +
          0.88f * powf(y * (1.f - y), 1.53f) * powf(4.f * y * (1.f - y), 0.8f));
    const float A2qZ = 8.226f * z6;
    gNu = kHPlanck * y * Ay / A2qZ;
    gNu = gNu * 1e27f;
    return (double)(alpha * gNu);
}

// ffCoeff (emission_mod.f90:726-742)
__device__ __forceinline__ double ffcoeff(float nu, float Z, float g, float Te, float sqrTe)
{
    const double expF = exp((double)(-nu * kHcRydK / Te));
    return (double)(4.f * kPi * 54.43f * Z * Z * g) * expF / (double)sqrTe;
}

// getFlux's blackbody (continuum_mod.f90:359-400): the Planck function over h
__device__ __forceinline__ float blackbody(float energy, float T)
{
    const float constant = 0.5250229f / kHPlanck;
    if (kHcRydK * energy / T > 86.f)
        return (float)((double)(constant * energy * energy * energy) * exp((double)(-kHcRydK * energy / T)));
    const float den = expf(kHcRydK * energy / T) - 1.f;
    if (den <= 0.f) return 3.32154e-6f * energy * energy * T / kHPlanck;
    return constant * energy * energy * energy / den;
}

// locate (interpolation_mod.f90:48-81) of Te in the increasing tkGamma
__device__ int locate_tk(const float *xa, int n, float x)
{
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    int p = 0;
    float bd = 0.f;
    for (int i = 0; i < n; ++i) {
        const float d = xa[i] - x;
        if (d > 0.f && (p == 0 || d < bd)) { p = i + 1; bd = d; }
    }
    return p - 1 > 1 ? p - 1 : 1;
}

__global__ void __launch_bounds__(64) emission_kernel(const mcb::Atomic *__restrict__ Ap, EmArgs a)
{
    extern __shared__ __align__(16) unsigned char dyn[];
    const mcb::Atomic &A = *Ap;
    const int cell = blockIdx.x + 1;
    const int lane = threadIdx.x;
    const int nb = a.nb;
    const float *nu = a.nu;
    const bool dust = a.dabs != nullptr;
    // region A: the forbidden-line solves, then the per-bin sums
    CelLds &S = *reinterpret_cast<CelLds *>(dyn);
    float *fterm = reinterpret_cast<float *>(dyn + sizeof(CelLds));
    float *s1 = reinterpret_cast<float *>(dyn), *s2 = s1 + nb, *s3 = s1 + 2 * nb, *sD = s1 + 3 * nb;
    unsigned char *rest = dyn + region_a_bytes(nb, dust);
    float (*loc)[kMaxLim] = reinterpret_cast<float (*)[kMaxLim]>(rest);
    float *rterm = reinterpret_cast<float *>(rest + sizeof(float) * 3 * kMaxLim);
    float *red = rterm + kMaxRec;

    // the state as the host grid stores it (REAL)
    const float Te = (float)a.Te[cell], NeIn = (float)a.Ne[cell];
    const double *ionc = a.ion + (size_t)cell * a.nE * a.ns1;
#define XION(e, st) ((float)ionc[(e) * a.ns1 + (st)])
    float NeUsed = NeIn;
    const float sqrTe = sqrtf(Te);
    const float log10Te = log10f(Te);
    if (NeUsed < 1.e-5f) NeUsed = 1.e-5f;
    const float xHp = XION(a.eH, 2), xHe1 = XION(a.eHe, 2), xHe2 = XION(a.eHe, 3);
    if (!(Te > 0.f) || !isfinite(Te) || !isfinite(NeIn)) {
        // not a state the balance leaves (its floor is 1 K): refused
        if (lane == 0) a.err[cell] = 1;
        return;
    }

    float normFor = 0.f;
    // forbidden lines (forLines), each model ion's terms above 1e-35 summed
    // lower level outer, upper inner, as setDiffusePDF does
    const double kL = (double)1.9865e9f;
    for (int k = 0; k < A.ncel; ++k) {
        const mcb::CelIon &ci = A.cel[k];
        const float x = XION(ci.eidx, ci.ion);
        const float up = (ci.ion < a.ns1 - 1) ? XION(ci.eidx, ci.ion + 1) / x : 0.f;
        if (up != up) continue;                  // equilibrium's NaN test: no emission
        mcb::cel_levels(A, ci, S, lane, (double)Te, (double)NeUsed, (double)up);
        const int nl = ci.nl;
        const double *E = A.cel_e + ci.off_e;
        const double *Aa = A.cel_a + ci.off_a;
        double sumN = 0.0;
        for (int i = 1; i <= nl; ++i) sumN = sumN + S.y[i];
        const double ab = (double)(float)A.el[ci.eidx].abun;
        for (int p = lane; p < kLev * kLev; p += W) {
            const int lo = p / kLev + 1, hi = p % kLev + 1;
            float v = 0.f;
            if (lo < hi && hi <= nl) {
                const double av = Aa[(hi - 1) + nl * (lo - 1)];
                if (av != 0.0) {
                    const double Eji = E[hi - 1] - E[lo - 1];
                    double f = av * Eji * (S.y[hi] / sumN);
                    f = f * ab * (double)x;
                    f = f * kL;
                    if (f > (double)1.e-35f) v = (float)f;
                }
            }
            fterm[p] = v;
        }
        __syncthreads();
        if (lane == 0)
            for (int p = 0; p < kLev * kLev; ++p) normFor = normFor + fterm[p];
        __syncthreads();
    }
    __syncthreads();                             // region A now holds the per-bin sums

    // log gamma rows at Te (fb_ff, emission_mod.f90:306-336)
    for (int s = 0; s < 3; ++s)
        for (int i = lane; i < a.nlim[s]; i += W) {
            const float *g = a.lg[s];
            const int ntk = a.ntk;
            float v;
            if (Te <= a.tk[0]) v = g[ntk * i];
            else if (Te >= a.tk[ntk - 1]) v = g[(ntk - 1) + ntk * i];
            else {
                const int t = locate_tk(a.tk, ntk, Te);
                const double factor = (double)log10f(Te / a.tk[t - 1]);
                const double aFit = (double)((g[t + ntk * i] - g[(t - 1) + ntk * i]) / log10f(a.tk[t] / a.tk[t - 1]));
                v = g[(t - 1) + ntk * i] + (float)(aFit * factor);
            }
            loc[s][i] = v;
        }
    __syncthreads();

    // per bin: emission coefficients and the partial sums of setDiffusePDF
    const double constant = (double)4.9874105e-6f;
    const float statW[3] = { 2.f, 0.5f, 2.f };
    const float alpha21 = 6.23f * powf(Te / 10000.f, -0.827f);
    for (int i = lane + 1; i <= nb; i += W) {
        const float nui = nu[i - 1];
        double g[3];
        for (int s = 0; s < 3; ++s) {
            g[s] = 0.0;
            const int k = a.pair[s][i - 1];
            if (k >= 0) {
                const int ilow = a.edge[s][2 * k], iup = a.edge[s][2 * k + 1];
                const double aFit = (double)((loc[s][2 * k + 1] - loc[s][2 * k]) / (nu[iup - 1] - nu[ilow - 1]));
                const double v = (double)loc[s][2 * k + 1] + aFit * (double)(nui - nu[iup - 1]);
                g[s] = pow(10.0, v);
            }
        }
        if (i >= a.ip[0]) {
            // the Milne relation above the H, He I and He II edges, plus free-free
            const double factor = (double)((nui * nui * nui) / (Te * sqrTe));
            double expF = exp((double)((-nui + nu[a.ip[0] - 1]) * kHcRydK / Te));
            g[0] = (double)(4.f * kPi) * (double)a.sig[0][i - a.i0[0]] * (double)statW[0] * (double)kHcRyd * constant *
                   factor * expF * (double)1.e20f * (double)1.e20f;
            g[0] = g[0] + ffcoeff(nui, 1.f, a.gff[i - 1], Te, sqrTe);
            if (i >= a.i0[1]) {
                expF = exp((double)((-nui + nu[a.ip[1] - 1]) * kHcRydK / Te));
                g[1] = (double)(4.f * kPi) * (double)a.sig[1][i - a.i0[1]] * (double)statW[1] * (double)kHcRyd *
                       constant * factor * expF * (double)1.e20f * (double)1.e20f;
                g[1] = g[1] + ffcoeff(nui, 1.f, a.gff[i - 1], Te, sqrTe);
                if (i >= a.i0[2]) {
                    expF = exp((double)((-nui + nu[a.ip[2] - 1]) * kHcRydK / Te));
                    g[2] = (double)(4.f * kPi) * (double)a.sig[2][i - a.i0[2]] * (double)statW[2] * (double)kHcRyd *
                           constant * factor * expF * (double)1.e20f * (double)1.e20f;
                    g[2] = g[2] + ffcoeff(nui, 2.f, a.gffHeII[i - 1], Te, sqrTe);
                }
            }
        }
        // heavy-ion free-bound emission (fb_ff, emission_mod.f90:473-528)
        double gh = 0.0;
        for (int c = 0; c < a.nfb; ++c) {
            const int *m = a.fb_i + 4 * c;
            const int ipc = m[2];
            if (i >= ipc && i < m[3]) {
                const double expF = exp((double)((-nui + nu[ipc - 1]) * kHcRydK / Te));
                gh = gh + expF * (double)a.fb_sig[a.fb_off[c] + (i - ipc)] * (double)a.fb_f[2 * c] *
                              (double)XION(m[0], m[1]) * (double)a.fb_f[2 * c + 1];
            }
        }
        {
            const float factor = nui * nui * nui / (Te * sqrTe);
            gh = gh * (double)(4.f * kPi) * constant * (double)kHcRyd * (double)factor * (double)1.e20f *
                 (double)1.e20f;
        }
        // two-photon continua (twoPhoton, HeI2photSub)
        const double tHI = hydro2phot(nui, 1, Te, NeUsed);
        const double tHeII = hydro2phot(nui, 2, Te, NeUsed);
        double tHeI = 0.0;
        const int j = a.jHeI[i - 1];
        if (j > 0) {
            const float y = nui / 1.514f;
            const float fit1 = (a.Aydat[j] - a.Aydat[j - 1]) / (a.ydat[j] - a.ydat[j - 1]);
            const float fit2 = a.Aydat[j - 1] - a.ydat[j - 1] * fit1;
            float Ay = fit1 * y + fit2;
            if (Ay < 9.2163086E-03f) Ay = 0.f;
            tHeI = (double)(alpha21 * Ay * 0.66262f * y / 51.3f);
        }
        double eHI = (g[0] + tHI) * (double)a.abH * (double)xHp * (double)NeUsed + gh * (double)NeUsed;
        double eHeI = (g[1] + tHeI) * (double)a.abHe * (double)xHe1 * (double)NeUsed;
        double eHeII = (g[2] + tHeII) * (double)a.abHe * (double)xHe2 * (double)NeUsed;
        if (eHI < (double)1.e-15f) eHI = 0.0;
        if (eHeI < (double)1.e-15f) eHeI = 0.0;
        if (eHeII < (double)1.e-15f) eHeII = 0.0;
        const float cw = kCRyd * a.widFlx[i - 1];
        s1[i - 1] = (float)((double)cw * eHI / (double)1.e15f);
        s2[i - 1] = (float)((double)cw * eHeI / (double)1.e15f);
        s3[i - 1] = (float)((double)cw * eHeII / (double)1.e15f);
        if (dust) {
            // grains below their sublimation temperature, sizes outer and
            // species inner as setDiffusePDF loops
            float sd = 0.f;
            const size_t tb = (size_t)cell * (a.dnsMax + 1) * (a.dSizes + 1);
            for (int ai = 1; ai <= a.dSizes; ++ai)
                for (int sp = 1; sp <= a.dS; ++sp) {
                    const float td = (float)a.Tdust[tb + sp + (size_t)(a.dnsMax + 1) * ai];
                    if (!(td < a.dtsub[(sp - 1) * a.dSizes + ai - 1])) continue;
                    const float bb = blackbody(nui, td);
                    sd = sd + a.dabs[((size_t)(sp - 1) * a.dSizes + ai - 1) * nb + i - 1] * bb * a.widFlx[i - 1] *
                                  a.dweight[ai - 1] * a.dgabun[sp - 1];
                }
            sD[i - 1] = sd;
        }
    }

    // recombination-line terms in normRec's order (RecLinesEmission and
    // setDiffusePDF, emission_mod.f90:744-832, 1099-1117): H I 3..15 -> 2..8,
    // He II 3..30 -> 2..16, He I singlets
    {
        float Hbeta = powf(10.f, -0.870f * log10Te + 3.57f);
        Hbeta = Hbeta * NeUsed * xHp * a.abH;
        const float Lalpha = powf(10.f, -0.897f * log10Te + 5.05f);
        float h4686 = powf(10.f, -.997f * log10f(Te) + 5.16f);
        h4686 = h4686 * NeUsed * a.abHe * xHe2;
        const float T4 = Te / 10000.f;
        int denint;
        if (NeIn <= 100.f) denint = 0;
        else if (NeIn <= 1.e4f) denint = 1;
        else if (NeIn <= 1.e6f) denint = 2;
        else denint = 3;
        const float p100[4] = { 1.f, 100.f, 10000.f, 1.e6f };
        for (int t = lane; t < 70 + 315 + 34; t += W) {
            float v;
            if (t < 70) {
                int u = 3, r = t;
                while (r >= (u - 1 < 8 ? u - 1 : 8) - 1) { r -= (u - 1 < 8 ? u - 1 : 8) - 1; ++u; }
                const int l = 2 + r;
                double h = a.HIrec[(u - 3) + 28 * (l - 2)] * (double)Hbeta;
                if (u == 15 && l == 8) h = h + (double)(a.abH * xHp * NeUsed * Lalpha);
                v = (float)h;
            } else if (t < 70 + 315) {
                int u = 3, r = t - 70;
                while (r >= (u - 1 < 16 ? u - 1 : 16) - 1) { r -= (u - 1 < 16 ? u - 1 : 16) - 1; ++u; }
                const int l = 2 + r;
                v = (float)(a.HeIIrec[(u - 3) + 28 * (l - 2)] * (double)h4686);
            } else {
                const int i = t - 70 - 315 + 1;
#define HEIC(d, k) a.HeIc[((i) - 1) + 34 * (((d) - 1) + 3 * ((k) - 1))]
                double h;
                if (denint > 0 && denint < 3) {
                    const float x1 = HEIC(denint, 1) * powf(T4, HEIC(denint, 2)) * expf(HEIC(denint, 3) / T4);
                    const float x2 =
                        HEIC(denint + 1, 1) * powf(T4, HEIC(denint + 1, 2)) * expf(HEIC(denint + 1, 3) / T4);
                    h = (double)(x1 + ((x2 - x1) * (NeUsed - p100[denint]) / (p100[denint + 1] - p100[denint])));
                } else {
                    const int d = denint == 0 ? 1 : 3;
                    h = (double)(HEIC(d, 1) * powf(T4, HEIC(d, 2)) * expf(HEIC(d, 3) / T4));
                }
#undef HEIC
                h = h * (double)NeUsed * (double)a.abHe * (double)xHe1;
                v = (float)h;
            }
            rterm[t] = v;
        }
    }
    __syncthreads();
    float normRec = 0.f;
    if (lane == 0)
        for (int t = 0; t < 70 + 315 + 34; ++t) normRec = normRec + rterm[t];

    // lane 0: the partial sums in bin order, the He lines able to ionise H,
    // totalLines, and the cumulative recPDF (in s1)
    if (lane == 0) {
        float nHI = 0.f, nHeI = 0.f, nHeII = 0.f, nD = 0.f;
        for (int i = 0; i < nb; ++i) {
            nHI = nHI + s1[i];
            nHeI = nHeI + s2[i];
            nHeII = nHeII + s3[i];
        }
        const float T4 = Te * 1.e-4f;
        const float a2tS = 27.2f * powf(T4, -0.678f);
        const float corr = a2tS * NeUsed * a.abHe * xHe1 * 1.45673f * kHcRyd * 1e11f;
        s2[a.j2TsP - 1] = s2[a.j2TsP - 1] + corr;
        nHeI = nHeI + corr;
        float ly[kMaxLy];
        for (int k = 0; k < a.nly; ++k) ly[k] = a.lyData[k];
        ly[3] = powf(10.f, -0.792f * log10f(Te) + 6.01f);
        ly[3] = ly[3] * NeUsed * xHe2 * a.abHe;
        for (int k = 0; k < a.nly - 1; ++k) ly[k] = ly[k] * ly[3];
        for (int k = 0; k < a.nly; ++k) {
            s3[a.lyP[k] - 1] = s3[a.lyP[k] - 1] + ly[k];
            nHeII = nHeII + ly[k];
        }
        float normalize = nHI + nHeI + nHeII;
        if (dust) {
            // setDiffusePDF scales the raw grain sums by h 4 pi 1e25 cRyd
            // Ndust / Hden; the raw total is taken in bin order here (the
            // reference adds sizes, species, then bins)
            const float cst = kHPlanck * 4.f * kPi * 1.e25f * kCRyd;
            const float nd = a.ndust[cell], hd = (float)a.Hden[cell];
            for (int i = 0; i < nb; ++i) {
                nD = nD + sD[i];
                sD[i] = sD[i] * cst * nd / hd;
            }
            nD = nD * cst * nd / hd;
            normalize = normalize + nD;
            float r = 0.f;
            for (int i = 0; i < nb; ++i) {
                r = i == 0 ? s1[i] + s2[i] + s3[i] + sD[i] : r + s1[i] + s2[i] + s3[i] + sD[i];
                s1[i] = r;
            }
        } else {
            float r = 0.f;
            for (int i = 0; i < nb; ++i) {
                const float t = s1[i] + s2[i] + s3[i];
                r = i == 0 ? t : r + t;
                s1[i] = r;
            }
        }
        const float tl = normRec + normFor;
        normalize = normalize + tl;
        // the reference's fraction; at a few tens of K its fits give inf*0
        // and it is NaN, which newPhotonPacket's test random <= totalLines
        // never passes (photon_mod.f90:925-929): the same as 0
        float f = tl / normalize;
        a.err[cell] = isnan(f) ? 2 : 0;
        if (isnan(f)) f = 0.f;
        a.totalLines[cell] = f;
        red[0] = s1[nb - 1];
    }
    __syncthreads();
    // recPDF = cumulative / last (setDiffusePDF).  In barely lit cells the
    // iteration leaves at a few tens of K, the REAL cumulative sum
    // overflows (the Milne factor exp((nu_edge - nu) hcRyd/kT) of the bins
    // between a cross-section's start and its edge bin), so entries from the
    // overflow on are inf/inf = NaN; no finite total divides 0/0 = NaN
    // everywhere.  getNu2's scan (photon_mod.f90:745-752) stops at the
    // first entry random >= probDen fails, which a NaN entry fails as an
    // entry of 1 does: from the first NaN on the row is written 1, and the
    // transport draws the bins the reference draws.
    const float last = red[0];
    int m = nb;
    for (int i0 = 0; i0 < nb; i0 += W) {
        const int i = i0 + lane;
        const uint64_t bad = __ballot(i < nb && isnan(s1[i] / last));
        if (bad) { m = i0 + __ffsll((unsigned long long)bad) - 1; break; }
    }
    if (lane == 0 && m < nb) a.err[cell] = 2;
    float *row = a.rec + (size_t)cell * nb;
    for (int i = lane; i < nb; i += W) {
        float v = s1[i] / last;
        if (i >= m) v = 1.f;
        if (i == nb - 1) v = 1.f;
        if (v > 0.999998f) v = 1.f;
        row[i] = v;
    }
#undef XION
}

}  // namespace mce

size_t mce_lds_bytes(int nb, bool dust)
{
    return mce::region_a_bytes(nb, dust) + sizeof(float) * (3 * mce::kMaxLim + mce::kMaxRec + 4);
}

hipError_t mce_launch(const mcb::Atomic *dA, const mce::EmArgs &a, hipStream_t s)
{
    if (a.nc <= 0) return hipSuccess;
    hipLaunchKernelGGL(mce::emission_kernel, dim3(a.nc), dim3(64), mce_lds_bytes(a.nb, a.dabs != nullptr), s, dA, a);
    return hipGetLastError();
}

// mc_balance.hip - per-cell ionisation and thermal balance (MOCASSIN's
// updateCell, update_mod.f90:15-525) on MI355X, FP64.
//
// One 64-lane wavefront per cell (one wavefront per workgroup), cell state
// in LDS.  Lanes share the work of each stage of the reference's schedule:
//   * photoionisation rates and heating integrals: lanes over frequency
//     bins, wave reductions per channel (update_mod.f90:184-269, 1146-1234);
//   * recombination fits, charge exchange, ion ratios: lanes over the
//     (elem, ion) channels (calcAlpha, radRecFit, dielectronic, ionBalance);
//   * CEL cooling: the model ions one after another, each n-level system
//     built and eliminated cooperatively (rows and pivot updates across
//     lanes; equilibrium/luSlv of emission_mod.f90:1913-2190);
//   * the bracketed false-position T iteration and the fixed-point ion
//     iteration run uniformly on the wavefront (iterateT, ionBalance).
// The collision-strength spline second derivatives depend only on the
// atomic data, so they are computed once at setup (the reference recomputes
// them on every call; the values are the same).
#include "mc_balance.h"
#include "mc_cel.h"
#include "mc_device.h"
#include "mc_mlp.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>

#pragma clang fp contract(off)

namespace mcb {

constexpr int W = 64;

struct Lds {
    double Te, Ne, Hden;
    double ion[kMaxElem][kMaxStage];
    double nPhot[kMaxCh], heatNu[kMaxCh], alpha[kMaxCh], ratio[kMaxCh], chexv[kMaxCh];
    double diO[kMaxStage];
    double red[W];
    double x[kMaxLev + 1][kMaxLev + 1];
    double q[kMaxLev + 1][kMaxLev + 1];
    double y[kMaxLev + 1];
    double alev[kMaxLev + 1];
};

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ double rad_rec(const Channel &c, double Te)
{
    if (c.rr_kind == 0) {
        double tt = sqrt(Te / c.rr[2]);
        return c.rr[0] / (tt * pow(tt + 1.0, 1.0 - c.rr[1]) * pow(1.0 + sqrt(Te / c.rr[3]), 1.0 + c.rr[1]));
    }
    double tt = Te * 1.e-4;
    if (c.rr_kind == 2) return c.rr[0] / pow(tt, c.rr[1] + c.rr[2] * log10(tt));
    return c.rr[0] / pow(tt, c.rr[1]);
}

// dielectronic (update_mod.f90:1782-1834) for one (elem, ion): the last
// satisfying Nussbaumer-Storey row, else (or above 6e4 K) Aldrovandi-Pequignot
__device__ double di_rec(const Atomic &A, int ndr, const int *dr, int ap, double Te)
{
    double v = 0.0;
    double t = Te / 10000.;
    for (int r = 0; r < ndr; ++r) {
        const int i = dr[r];
        const int g = A.dr_g[i];
        if (g == 0 || (g == 1 && Te < 20000.) || (g == 2 && Te >= 20000.)) {
            const double *c = A.dr_real[i];
            v = 1.e-12 * (c[0] / t + c[1] + c[2] * t + c[3] * (t * t)) * pow(t, -1.5) * exp(-c[4] / t);
        }
    }
    if (Te > 60000. || v == 0.0) {
        if (ap >= 0) {
            const double *c = A.ap_real[ap];
            v = c[0] * pow(Te, -1.5) * exp(-c[2] / Te) * (1.0 + c[1] * exp(-c[3] / Te));
        } else {
            v = 0.0;
        }
    }
    return v;
}

// CEL cooling of one model ion [erg/s per H], levels 1..10 (forLines +
// equilibrium + the cooling sum of thermBalance)
__device__ __forceinline__ double cel_cool(const Atomic &A, const CelIon &ci, Lds &S, int lane, double Te)
{
    const int nl = ci.nl;
    const double *E = A.cel_e + ci.off_e;
    const double *Aa = A.cel_a + ci.off_a;
    const double xi = S.ion[ci.eidx][ci.ion];
    const double up = (ci.ion < A.nstages) ? S.ion[ci.eidx][ci.ion + 1] / xi : 0.0;
    if (up != up) return 0.0;
    cel_levels(A, ci, S, lane, Te, S.Ne, up);
    double cool = 0.0;
    if (lane == 0) {
        double sumN = 0.0;
        for (int i = 1; i <= nl; ++i) sumN = sumN + S.y[i];
        const double s1 = A.el[ci.eidx].abun, s2 = xi;
        const double kE = 1.9865e-16;
        for (int i = 1; i <= 10 && i <= nl - 1; ++i)
            for (int j = i + 1; j <= 10 && j <= nl; ++j) {
                const double a = Aa[(j - 1) + nl * (i - 1)];
                if (a != 0.0) {
                    const double Eji = E[j - 1] - E[i - 1];
                    cool = cool + a * Eji * (S.y[j] / sumN) * s1 * s2 * kE;
                }
            }
    }
    __syncthreads();
    return cool;   // valid on lane 0
}

// total cooling at temperature T for the cell's current ionisation and Ne
// (thermBalance, update_mod.f90:933-1113: H and He free-free and
// recombination fits, H collisional excitation, CEL lines); valid on lane 0
__device__ __forceinline__ double cooling(const Atomic &A, Lds &S, int lane, double T)
{
    double cool = 0.0;
    for (int k = 0; k < A.ncel; ++k) {
        double cc = cel_cool(A, A.cel[k], S, lane, T);
        cool = cool + cc;
    }
    if (lane != 0) return 0.0;
    const double Ne = S.Ne, kB = 1.380658e-16, hcRyd = 2.1799153e-11;
    const double lg = log10(T), Te4 = T / 1.e4;
    double Np = S.ion[0][2] * A.el[0].abun;
    double bFF = 1.0108464E-11 + 9.7930778E-13 * lg - 6.6433144E-13 * lg * lg +
                 2.4793747E-13 * lg * lg * lg - 2.3938215E-14 * lg * lg * lg * lg;
    double coolFF = Np * Ne * bFF * kB * T / sqrt(T);
    double bRec = 9.4255985E-11 - 4.04794384E-12 * lg - 1.0055237E-11 * lg * lg +
                  1.99266862E-12 * lg * lg * lg - 1.06681387E-13 * lg * lg * lg * lg;
    double coolRec = Np * Ne * bRec * kB * T / sqrt(T);
    Np = S.ion[1][3] * A.el[1].abun;
    const double ls = log10(T / 4.);
    bFF = 2. * (1.0108464E-11 + 9.7930778E-13 * ls - 6.6433144E-13 * ls * ls +
                2.4793747E-13 * ls * ls * ls - 2.3938215E-14 * ls * ls * ls * ls);
    coolFF = coolFF + Np * Ne * bFF * kB * T / sqrt(T / 4.);
    bRec = 2. * (9.4255985E-11 - 4.04794384E-12 * ls - 1.0055237E-11 * ls * ls +
                 1.99266862E-12 * ls * ls * ls - 1.06681387E-13 * ls * ls * ls * ls);
    coolRec = coolRec + Np * Ne * bRec * kB * T / sqrt(T / 4.);
    Np = S.ion[1][2] * A.el[1].abun;
    bFF = 1.070073e-11 - 2.5730207e-13 * lg + 2.109134e-13 * lg * lg;
    coolFF = coolFF + Np * Ne * bFF * kB * T / sqrt(T);
    bRec = 9.4255985E-11 - 4.04794384E-12 * lg - 1.0055237E-11 * lg * lg +
           1.99266862E-12 * lg * lg * lg - 1.06681387E-13 * lg * lg * lg * lg;
    coolRec = coolRec + Np * Ne * bRec * kB * T / sqrt(T);
    double coolColl = 0.0;
    if (T > 5000.)
        coolColl = (2.47e-8 * exp(-118338. / T) * pow(Te4, -0.228) +
                    1.32e-8 * exp(-140252. / T) * pow(Te4, -0.460)) *
                   hcRyd * S.ion[0][1] * Ne;
    coolColl = coolColl + cool;
    return coolFF + coolRec + coolColl;
}

// ---- grains: charge, photoelectric heating, gas-grain collisions ---------
// (update_mod.f90:527-932).  Per temperature iteration, after the ion fixed
// point: every grain row's potential (setGrainPotential's secant iteration,
// at most 100 steps), then the photoelectric heating and cooling
// (setPhotoelHeatCool) and the collision terms (setDustGasCollHeatCool).
// The cell's photon flux J/(h c Ryd nu) per bin is staged in LDS once per
// cell; the bin sums run across lanes, the collision sums with one lane per
// grain row in the reference's order of elements and stages.
constexpr double kPiR = (double)3.141592654f;             // constants_mod.f90:32 (REAL)
constexpr double kFourPiR = (double)(4.f * 3.141592654f);
constexpr double kHcRyd = 2.1799153e-11, kRyd2erg = 2.1799e-11, kKB = 1.380658e-16;
constexpr double kAmu = 1.66053e-24, kMe = 9.109382e-28;

// dynamic LDS of the grain variant, after the cell's photon flux pf[nbins]
struct GrainLds {
    double w[kMaxElem][kMaxStage];   // n(ion) v_mean of each (element, stage >= 2)
    double pot[kMaxGrainRows], gdc[kMaxGrainRows], peh[kMaxGrainRows];
    double gdcg, pehg;
};

__device__ __forceinline__ double minr(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double maxr(double a, double b) { return a > b ? a : b; }

// locate(nuArray, x) on the increasing frequency mesh (interpolation_mod.f90:
// 48-81): nb when x > nu(nb), 0 when x < nu(1), else the last i with
// nu(i) <= x, at least 1 (x == nu(nb) gives 1)
__device__ int locate_nu(const double *nu, int n, double x)
{
    if (x > nu[n - 1]) return n;
    if (x < nu[0]) return 0;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (nu[mid] > x) hi = mid; else lo = mid + 1;
    }
    const int p = lo < n ? lo + 1 : 0;
    return p - 1 > 1 ? p - 1 : 1;
}

// getGrainEmission (update_mod.f90:601-651) of row k at potential Vg, on every lane
__device__ double grain_emission(const Atomic &A, const Grains &G, const double *pf, int k, double Vg, int lane)
{
    const int nb = A.nbins, s = k / G.nSizes;
    const double vn = G.vn[s];
    const int ip = locate_nu(A.nu, nb, maxr(vn + Vg, vn)) + 1;
    const double *qa = G.qa + (size_t)k * nb, *yn = G.yn + (size_t)s * nb;
    double sum = 0.0;
    for (int f = ip + lane; f <= nb; f += W) {
        const double Yhat = yn[f - 1] * minr(1., maxr(0., 1. - Vg / (A.nu[f - 1] - vn)));
        const double photFlux = pf[f - 1] * kFourPiR;
        sum = sum + Yhat * photFlux * qa[f - 1];
    }
    return wave_sum(sum);
}

// getGrainRecombination (update_mod.f90:656-733), on every lane
__device__ double grain_recombination(const Atomic &A, const GrainLds &L, double Te, double eNe, double Vg, int lane)
{
    const double kT = 6.336e-6 * Te;
    double eta = -(Vg / kT);
    eta = eta <= 0. ? 1. - eta : exp(-eta);
    const double rec = eNe * eta;
    double ions = 0.0;
    for (int i = lane; i < A.nE * kMaxStage; i += W) {
        const int e = i / kMaxStage, st = i % kMaxStage;
        if (st < 2 || st > A.el[e].nst) continue;
        double et = (double)(st - 1) * Vg / kT;
        et = et <= 0. ? 1. - et : exp(-et);
        ions = ions + L.w[e][st] * et;
    }
    return rec - wave_sum(ions);
}

// the grain block of iterateT after ionBalance (update_mod.f90:367-388) and
// the two heat/cool terms it adds to thermBalance (1252-1255)
__device__ void grain_terms(const Atomic &A, const Grains &G, const double *pf, GrainLds &L, Lds &S, int lane,
                            int cell)
{
    const int nb = A.nbins, K = G.K;
    const double Te = S.Te, Ne = S.Ne, Hden = S.Hden;
    const double kT = 6.336e-6 * Te, e8 = 1.1045e-15 * Te / kPiR;
    // n(ion) v_mean of the ions; electrons ((Ne vmean) S) with S = 1
    for (int i = lane; i < A.nE * kMaxStage; i += W) {
        const int e = i / kMaxStage, st = i % kMaxStage;
        if (st < 2 || st > A.el[e].nst) continue;
        const double cpDen = S.ion[e][st] * A.el[e].abun * Hden;
        const double vm = sqrt(e8 / (G.aweight[A.el[e].z - 1] * kAmu));
        L.w[e][st] = cpDen * vm * 1.;
    }
    __syncthreads();
    const double eNe = Ne * sqrt(e8 / kMe) * 1.;
    // setGrainPotential: secant on emission - recombination, from 0 V
    for (int k = 0; k < K; ++k) {
        const double vn = G.vn[k / G.nSizes];
        double slope = 0., potOld = 0.;
        double emiOld = grain_emission(A, G, pf, k, 0., lane);
        double recOld = grain_recombination(A, L, Te, eNe, 0., lane);
        double pot = potOld + 0.05;
        for (int n = 1;; ++n) {
            const double threshold = maxr(vn + pot, vn);
            const double emi = grain_emission(A, G, pf, k, pot, lane);
            const double rec = grain_recombination(A, L, Te, eNe, pot, lane);
            double delta = emi - rec;
            double delta1 = fabs(delta / (0.5 * maxr(1.e-35, emi + rec)));
            if (delta1 < 0.005) break;
            if (pot != potOld) {
                const double fac = (emi - emiOld) - (rec - recOld);
                if (fac != 0.) slope = fac / (pot - potOld);
            }
            potOld = pot;
            recOld = rec;
            emiOld = emi;
            delta1 = -(delta / slope);
            delta = minr(fabs(delta1), 0.05 * threshold);
            delta = signbit(delta1) ? -fabs(delta) : fabs(delta);
            pot = pot + delta;
            if (n >= 100) break;
        }
        if (lane == 0) L.pot[k] = pot;
    }
    __syncthreads();
    // setPhotoelHeatCool: lanes over bins, one row after another
    double pehg = 0.0;
    for (int k = 0; k < K; ++k) {
        const int s = k / G.nSizes, ai = k % G.nSizes;
        const double pot = L.pot[k], vn = G.vn[s];
        int thP = locate_nu(A.nu, nb, maxr(vn + pot, vn));
        if (thP <= 0) thP = 1;
        const double *ax = G.absx + (size_t)k * nb, *yn = G.yn + (size_t)s * nb;
        double hd = 0.0, hg = 0.0;
        for (int f = thP + lane; f <= nb; f += W) {
            const double nu = A.nu[f - 1], Yn = yn[f - 1];
            const double Yhat = Yn * minr(1., maxr(0., 1. - pot / (nu - vn)));
            const double d = nu - vn;
            const double EY = Yn * 0.5 * minr(d, maxr(0., (d * d - pot * pot) / d));
            const double q = ax[f - 1] * pf[f - 1];
            hd = hd + q * EY * kHcRyd / kPiR;
            hg = hg + q * (EY - Yhat * pot) * G.gabun[s] * G.weight[ai];
        }
        hd = wave_sum(hd);
        hg = wave_sum(hg);
        if (lane == 0) L.peh[k] = hd;
        pehg = pehg + hg;
    }
    pehg = pehg * G.ndust[cell] * kHcRyd / Hden;
    // setDustGasCollHeatCool: one lane per grain row, elements and stages in
    // the reference's order; IPerg is set for ionised stages and kept for
    // the neutral stage that follows (it starts at 0)
    double gd = 0.0, gg = 0.0;
    if (lane < K) {
        const int k = lane, s = k / G.nSizes, ai = k % G.nSizes;
        const double pot = L.pot[k], rad = G.radius[ai], ms = G.msurf[s];
        const double Td = G.Tdust[(size_t)cell * (G.nsMax + 1) * (G.nSizes + 1) + (s + 1) + (G.nsMax + 1) * (ai + 1)];
        const double nd = G.ndust[cell];
        double IPerg = 0.0;
        for (int e = 0; e < A.nE; ++e) {
            const int z = A.el[e].z;
            const double mcp = G.aweight[z - 1] * kAmu;
            const double vmean = sqrt(e8 / mcp);
            for (int st = 1; st <= A.el[e].nst; ++st) {
                const double Z = (double)(st - 1);
                if (st > 1) IPerg = A.nu[A.ch[A.el[e].ch0 + st - 2].ip - 1] * kRyd2erg;
                const double S_ = st == 1 ? 2. * mcp * ms / ((mcp + ms) * (mcp + ms)) : 1.;
                const double psi = Z * pot / kT;
                double eta, xi;
                if (psi <= 0.) { eta = 1. - psi; xi = 1. - psi / 2.; }
                else { eta = exp(-psi); xi = (1. + psi / 2.) * eta; }
                const double x = S.ion[e][st], ab = A.el[e].abun;
                gd = gd + x * ab * Hden * kPiR * rad * rad * 1.e-8 * S_ * vmean *
                          (2. * kT * kRyd2erg * xi - eta * (Z * pot * kRyd2erg - IPerg + 2. * kKB * Td));
                gg = gg + x * G.weight[ai] * G.gabun[s] * nd * ab * kPiR * rad * rad * 1.e-8 * S_ * vmean *
                          (2. * kT * kRyd2erg * xi - eta * 2. * kKB * Td);
            }
        }
        // electrons
        const double vme = sqrt(e8 / kMe), Z = -1.;
        const double psi = Z * pot / kT;
        double eta, xi;
        if (psi <= 0.) { eta = 1. - psi; xi = 1. - psi / 2.; }
        else { eta = exp(-psi); xi = (1. + psi / 2.) * eta; }
        gd = gd + Ne * kPiR * rad * rad * 1.e-8 * 1. * vme * (2. * kT * kRyd2erg * xi - eta * (Z * pot * kRyd2erg));
        gg = gg + Ne * G.weight[ai] * G.gabun[s] * nd * kPiR * rad * rad * 1.e-8 * 1. * vme *
                  (2. * kT * kRyd2erg * xi) / Hden;
        L.gdc[k] = gd / kFourPiR;
    }
    gg = wave_sum(gg) / kFourPiR;
    if (lane == 0) { L.gdcg = gg; L.pehg = pehg; }
    __syncthreads();
}

template <bool PE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) balance_kernel(const Atomic *__restrict__ Ap, CellArgs a)
{
    __shared__ Lds S;
    extern __shared__ double dyn[];           // grain variant: photon flux per bin, then GrainLds
    const Atomic &A = *Ap;
    const int lane = threadIdx.x;
    const int cell = a.order ? a.order[blockIdx.x] : (int)blockIdx.x + 1;
    if (cell > a.ncells) return;
    const int ns = A.nstages, nE = A.nE, nch = A.nch;
    double *ionOut = a.ion + (size_t)cell * nE * (ns + 1);

    // hit test (update_mod.f90:104-154) and the channel integrals, both
    // from integrals_kernel
    if (!a.hit[cell]) {
        if (lane == 0) { a.flags[cell] = 16; a.nIter[cell] = 0; }
        return;
    }
    if (lane == 0) {
        S.Te = a.Te[cell];
        S.Ne = a.Ne[cell];
        S.Hden = a.Hden[cell];
    }
    for (int i = lane; i < nE * (ns + 1); i += W) S.ion[i / (ns + 1)][i % (ns + 1)] = ionOut[i];
    for (int k = lane; k < nch; k += W) {
        S.nPhot[k] = 1.e-20 + a.Rint[(size_t)cell * nch + k];
        S.heatNu[k] = a.Hint[(size_t)cell * nch + k];
    }
    double *pf = dyn;
    GrainLds &GL = *reinterpret_cast<GrainLds *>(dyn + A.nbins);
    if (PE) {
        // the cell's photon flux J/(h c Ryd nu) (getGrainEmission, setPhotoelHeatCool)
        const Grains &G = *a.grains;
        const unsigned pt = mcd::j3_point_of(mcd::j3_shape(G.nx, G.ny, G.nz), G.cell_xyz[cell]);
        for (int f = lane; f < A.nbins; f += W)
            pf[f] = G.J3[mcd::j3_index(A.nbins, f, pt)] * a.jscale / (kHcRyd * A.nu[f]);
    }
    __syncthreads();
    // lgConverged compares with the caller's H0 fraction (update_mod.f90:510-515)
    const double XOldHI = a.ion0 ? a.ion0[(size_t)cell * nE * (ns + 1) + 1] : S.ion[0][1];

    int nIterT = 0, flags = 0, lgIBConv = 1, nTot = 0;
    for (int attempt = 0;; ++attempt) {
        nIterT = 0;
        flags = 0;
        lgIBConv = 1;
        double rHigh = 0., rLow = 0., Thigh = 0., Tlow = 0.;        // reference schedule
        double Tp = 0., fp = 0., Tlo = 0., Thi = 0.;                // secant iteration
        bool havep = false, haveLo = false, haveHi = false;
        for (;;) {
            nIterT++;
            const double Te = S.Te;
            // calcAlpha: radiative + dielectronic (oxygen rows fill empty entries)
            for (int i = 1 + lane; i < ns; i += W)
                S.diO[i] = di_rec(A, A.odr_n[i], A.odr[i], A.oap[i], Te);
            __syncthreads();
            for (int k = lane; k < nch; k += W) {
                const Channel &c = A.ch[k];
                double al = rad_rec(c, Te);
                if (c.elem >= 3) {
                    double d = di_rec(A, c.ndr, c.dr, c.ap, Te);
                    if (d == 0.0) d = S.diO[c.ion];
                    al = al + d;
                }
                S.alpha[k] = al;
                // charge-exchange rate coefficient (ionBalance, 1448-1455)
                const double t4 = Te / 10000.;
                double cx = c.chex[0] * pow(t4, c.chex[1]) * (1. + c.chex[2] * exp(c.chex[3] * t4));
                if (cx < 0.) cx = 0.;
                if (Te < 6000. || Te > 5.e4) cx = 0.;
                S.chexv[k] = cx;
            }
            __syncthreads();
            // ionBalance fixed point (1320-1622)
            int nIterX = 1;
            for (;;) {
                const double HIOld = S.ion[0][1], HeIOld = S.ion[1][1], HeIIOld = S.ion[1][2];
                const double xH0 = S.ion[0][1];
                for (int k = lane; k < nch; k += W) {
                    const Channel &c = A.ch[k];
                    double collIon = 0.0;
                    if (k == 0) {
                        double expFact = 157893.94 / Te;
                        if (expFact > 75.) expFact = 75.;
                        collIon = 2.75E-16 * Te * sqrt(Te) * (157893.94 / Te + 2.) * exp(-expFact);
                    }
                    double revRate = 0.0;
                    if (c.dEk > 1.0)
                        revRate = S.chexv[k] * (1. - xH0) * S.Hden * 2. * exp(-c.dEk / Te) / c.g0g1;
                    S.ratio[k] = (S.nPhot[k] + collIon + revRate) /
                                 (S.Ne * S.alpha[k] + S.chexv[k] * S.Hden * xH0);
                }
                __syncthreads();
                if (lane < nE) {
                    const Element &e = A.el[lane];
                    double prod[kMaxStage];
                    double den = 1.0, p = 1.0;
                    for (int i = 0; i < e.nch; ++i) { p = 1.0; for (int q = 0; q <= i; ++q) p = p * S.ratio[e.ch0 + q]; prod[i] = p; }
                    for (int i = 0; i < e.nch; ++i) den = den + prod[i];
                    for (int st = 1; st <= e.nst; ++st) {
                        double v = (st == 1) ? 1.0 / den : prod[st - 2] / den;
                        if (v > 0.99999) v = 0.99999;
                        if (v < 1.e-20) v = 1.e-20;
                        S.ion[lane][st] = v;
                    }
                }
                __syncthreads();
                if (lane == 0) {
                    double Ne = 0.0;
                    for (int ei = 0; ei < nE; ++ei)
                        for (int st = 2; st <= A.el[ei].nst; ++st)
                            if (S.ion[ei][st] >= 1.e-10) Ne = Ne + (double)(st - 1) * A.el[ei].abun * S.ion[ei][st];
                    Ne = Ne * S.Hden;
                    if (Ne == 0.0) Ne = 1.0;
                    S.Ne = Ne;
                }
                __syncthreads();
                const double dHI = (S.ion[0][1] - HIOld) / HIOld;
                const double dHeI = (S.ion[1][1] - HeIOld) / HeIOld;
                const double dHeII = (S.ion[1][2] - HeIIOld) / HeIIOld;
                const bool big = fabs(dHI) > a.ion_tol || fabs(dHeI) > a.ion_tol || fabs(dHeII) > a.ion_tol;
                if (big && nIterX < 25) { nIterX++; continue; }
                lgIBConv = !(big && nIterX == 25);
                break;
            }
            if (PE) grain_terms(A, *a.grains, pf, GL, S, lane, cell);
            // thermBalance (933-1258)
            double cool = cooling(A, S, lane, S.Te), heat = 0.0;
            if (lane == 0) {
                for (int k = 0; k < nch; ++k) {
                    const Channel &c = A.ch[k];
                    heat = heat + S.heatNu[k] * S.ion[c.eidx][c.ion] * A.el[c.eidx].abun;
                }
                if (PE) {
                    // the grain terms of thermBalance (1252-1255)
                    cool = cool + GL.gdcg;
                    heat = heat + GL.pehg;
                }
                S.red[0] = heat;
                S.red[1] = cool;
            }
            __syncthreads();
            heat = S.red[0];
            cool = S.red[1];
            __syncthreads();
            // iterateT false-position update (398-458)
            const double res = heat - cool;
            double Tn = S.Te;
            bool again = false;
            if (a.solver == 0) {
                if (fabs(res) >= 0.01 * heat) {
                    if (nIterT < 25) {
                        if (res < 0.) {
                            rLow = res; Tlow = Tn;
                            Tn = (Thigh != 0.) ? Tlow - (Thigh - Tlow) * rLow / (rHigh - rLow) : Tn / 1.2;
                        } else {
                            rHigh = res; Thigh = Tn;
                            Tn = (Tlow != 0.) ? Tlow - (Thigh - Tlow) * rLow / (rHigh - rLow) : Tn * 1.2;
                        }
                        if (nIterT >= 6 && fabs(Thigh - Tlow) <= 0.002 * (Thigh + Tlow)) {
                            if (res < 0.) { Tn = Thigh - 100.; Thigh = 0.; }
                            else { Tn = Tlow + 100.; Tlow = 0.; }
                        }
                        if (Tn <= 0.) Tn = 1.;
                        again = true;
                    } else {
                        if (Thigh == 0.) Tn = Tlow;
                        else if (Tlow == 0.) Tn = Thigh;
                        else Tn = (Thigh + Tlow) * 0.5;
                        flags |= 2 | 4;
                    }
                }
            } else if (fabs(res) > a.tol * heat) {
                // secant steps on f(T) = heat - cool, kept inside the bracket
                // once both signs have been seen (f > 0: heating wins, T rises)
                const double f = res, T = Tn;
                if (f > 0.) { Tlo = T; haveLo = true; } else { Thi = T; haveHi = true; }
                const double bl = fmin(Tlo, Thi), bh = fmax(Tlo, Thi);
                // a warm-started attempt gets warm_iter iterations before
                // the restart from the caller's state
                const int cap = (attempt == 0 && a.Te0) ? min(a.max_iter, a.warm_iter) : a.max_iter;
                if (nIterT >= cap) {
                    if (haveLo && haveHi) Tn = 0.5 * (bl + bh);
                    flags |= 2 | 4;
                } else if (haveLo && haveHi && bh - bl <= 1.e-7 * T) {
                    Tn = 0.5 * (bl + bh);
                } else {
                    // first trial: a step of the relative residual (cooling
                    // grows about linearly with T near balance), divided by
                    // the expected slope d(f/heat)/d ln T, at most step0
                    const double d0 = fmin(a.step0, fmax(fabs(f) / heat / a.slope0, 1.e-4));
                    double Tc = (havep && f != fp) ? T - f * (T - Tp) / (f - fp)
                                                   : (f > 0. ? T * (1. + d0) : T / (1. + d0));
                    if (haveLo && haveHi) {
                        if (!(Tc > bl && Tc < bh)) Tc = 0.5 * (bl + bh);
                    } else {
                        if (f > 0. && !(Tc > T)) Tc = T * 1.2;
                        if (f < 0. && !(Tc < T)) Tc = T / 1.2;
                        Tc = fmin(fmax(Tc, T / 1.5), T * 1.5);
                    }
                    Tp = T;
                    fp = f;
                    havep = true;
                    Tn = Tc;
                    again = true;
                }
            }
            __syncthreads();
            if (lane == 0) S.Te = Tn;
            __syncthreads();
            if (!again) break;
        }
        nTot += nIterT;
        const bool failed = (flags & 4) != 0 || !lgIBConv;
        if (failed && attempt == 0 && a.Te0) {
            // restart from the caller's state (the reference's starting point)
            const double *src = a.ion0 + (size_t)cell * nE * (ns + 1);
            for (int i = lane; i < nE * (ns + 1); i += W) S.ion[i / (ns + 1)][i % (ns + 1)] = src[i];
            if (lane == 0) { S.Te = a.Te0[cell]; S.Ne = a.Ne0[cell]; }
            __syncthreads();
            continue;
        }
        if (attempt > 0) flags |= 32;
        break;
    }
    if (!lgIBConv) flags |= 2 | 8;
    const double dXHI = (S.ion[0][1] - XOldHI) / XOldHI;
    if (fabs(dXHI) <= (double)A.XHILimit) flags |= 1;
    for (int i = lane; i < nE * (ns + 1); i += W) ionOut[i] = S.ion[i / (ns + 1)][i % (ns + 1)];
    if (PE && lane < a.grains->K) {
        // terms of the last temperature iteration, for getDustT
        const Grains &G = *a.grains;
        G.gdc[(size_t)cell * G.K + lane] = GL.gdc[lane];
        G.peh[(size_t)cell * G.K + lane] = GL.peh[lane];
    }
    if (lane == 0) {
        a.Te[cell] = S.Te;
        a.Ne[cell] = S.Ne;
        a.flags[cell] = flags;
        a.nIter[cell] = nTot;
    }
}

// Per lattice point (one lane each, so that every read of J3[nu][xyz] is
// coalesced across the wavefront): the photoionisation rate and heating
// integrals of every channel (update_mod.f90:184-269, 1146-1234), the hit
// flag (104-154) and, when X is given, the warm-start network's features of
// the cell (row cell-1 of X [nCells][72]): log10 of J summed over 64
// contiguous groups of bins, log10 nH, log10 abundances of its seven
// elements (-30 when absent).  Channels are accumulated eight at a time.
__global__ void __launch_bounds__(64) integrals_kernel(const Atomic *__restrict__ Ap, CellArgs a,
                                                       const double *__restrict__ J3,
                                                       const int *__restrict__ active, mcd::J3Shape js,
                                                       double *__restrict__ Rint, double *__restrict__ Hint,
                                                       int *__restrict__ hit, float *__restrict__ X)
{
    constexpr int G = 8, C = 64, U = 16;
    __shared__ double cr[G][C], ch[G][C];              // coefficients of one chunk of bins
    const Atomic &A = *Ap;
    const int lane = threadIdx.x;
    const int xyz = mcd::j3_lattice(js, blockIdx.x * 64u + (unsigned)lane);
    const int cell = xyz >= 0 ? active[xyz] : 0;
    if (!__any(cell > 0)) return;
    const int nb = A.nbins, nch = A.nch;
    // this wavefront's tile of J3: [nu][64 points], contiguous
    const double *Jp = J3 + mcd::j3_index(nb, 0, blockIdx.x * 64u) + lane;
    float *brow = (X && cell > 0) ? X + (size_t)(cell - 1) * mcm::kIn : nullptr;
    int any = 0;
    for (int g0 = 0; g0 < nch || g0 == 0; g0 += G) {
        const int ng = nch - g0 < G ? nch - g0 : G;
        double r[G], h[G];
        for (int k = 0; k < G; ++k) { r[k] = 0.0; h[k] = 0.0; }
        // the first pass visits every bin (hit flag, band sums)
        int jlo = nb + 1, jhi = 0;
        for (int k = 0; k < ng; ++k) {
            const Channel &c = A.ch[g0 + k];
            jlo = c.ip < jlo ? c.ip : jlo;
            const int hi = c.hiR > c.hiH ? c.hiR : c.hiH;
            jhi = hi > jhi ? hi : jhi;
        }
        if (g0 == 0) { jlo = 1; jhi = nb; }
        double band = 0.0;
        int b = 0, bend = nb / mcm::kBands;
        for (int c0 = jlo; c0 <= jhi; c0 += C) {
            const int cn = jhi - c0 + 1 < C ? jhi - c0 + 1 : C;
            __syncthreads();
            for (int k = 0; k < ng; ++k) {
                const size_t o = (size_t)(g0 + k) * nb + c0 - 1 + lane;
                cr[k][lane] = lane < cn ? A.cR[o] : 0.0;
                ch[k][lane] = lane < cn ? A.cH[o] : 0.0;
            }
            __syncthreads();
            for (int u0 = 0; u0 < cn; u0 += U) {
                double Jv[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    Jv[u] = u0 + u < cn ? Jp[(size_t)(c0 + u0 + u - 1) * mcd::kJTile] * a.jscale : 0.0;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (u0 + u >= cn) break;
                    const double J = Jv[u];
                    for (int k = 0; k < ng; ++k) {
                        r[k] = r[k] + J * cr[k][u0 + u];
                        h[k] = h[k] + J * ch[k][u0 + u];
                    }
                    if (g0 == 0) {
                        const int j = c0 + u0 + u;
                        any |= J > 0.0;
                        band += J;
                        if (j == bend) {
                            if (brow) brow[b] = (float)log10(fmax(band, 1.e-30));
                            band = 0.0;
                            ++b;
                            bend = (b + 1) * nb / mcm::kBands;
                        }
                    }
                }
            }
        }
        if (cell > 0)
            for (int k = 0; k < ng; ++k) {
                Rint[(size_t)cell * nch + g0 + k] = r[k];
                Hint[(size_t)cell * nch + g0 + k] = h[k];
            }
        if (nch == 0) break;
    }
    if (cell <= 0) return;
    hit[cell] = any;
    if (brow) {
        float *row = brow;
        row[mcm::kBands] = (float)log10(a.Hden[cell]);
        for (int k = 0; k < 7; ++k) {
            float v = -30.f;
            for (int e = 0; e < A.nE; ++e)
                if (A.el[e].z == mcm::kSlotsZ[k] && A.el[e].abun > 0.0) v = (float)log10(A.el[e].abun);
            row[mcm::kBands + 1 + k] = v;
        }
    }
}

// Warm start of reached cells from the network's (unscaled) outputs Y
// [nCells][ldy]: Te = 10^y0, the seven elements' fractions 10^y normalised
// per element and clamped as ionBalance clamps them, Ne from them.
__global__ void warm_kernel(const Atomic *__restrict__ Ap, CellArgs a, const float *__restrict__ Y, int ldy,
                            const int *__restrict__ hit)
{
    const Atomic &A = *Ap;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (cell > a.ncells || !hit[cell]) return;
    const int ns = A.nstages, nE = A.nE;
    const float *y = Y + (size_t)(cell - 1) * ldy;
    double *ion = a.ion + (size_t)cell * nE * (ns + 1);
    a.Te[cell] = fmin(fmax(pow(10.0, (double)y[0]), 500.0), 2.e5);
    int base = 1;
    for (int k = 0; k < 7; ++k) {
        const int z = mcm::kSlotsZ[k], nslot = z + 1 < mcm::kStageMax ? z + 1 : mcm::kStageMax;
        for (int e = 0; e < nE; ++e) {
            if (A.el[e].z != z) continue;
            const int nst = A.el[e].nst;
            double v[kMaxStage], sum = 0.0;
            for (int st = 1; st <= nst; ++st) {
                v[st] = st <= nslot ? pow(10.0, fmin((double)y[base + st - 1], 0.0)) : 0.0;
                sum += v[st];
            }
            for (int st = 1; st <= nst; ++st) {
                double x = sum > 0.0 ? v[st] / sum : (st == 1 ? 1.0 : 0.0);
                x = fmin(fmax(x, 1.e-20), 0.99999);
                ion[e * (ns + 1) + st] = x;
            }
        }
        base += nslot;
    }
    double Ne = 0.0;
    for (int e = 0; e < nE; ++e)
        for (int st = 2; st <= A.el[e].nst; ++st) {
            const double x = ion[e * (ns + 1) + st];
            if (x >= 1.e-10) Ne = Ne + (double)(st - 1) * A.el[e].abun * x;
        }
    a.Ne[cell] = Ne * a.Hden[cell] > 0.0 ? Ne * a.Hden[cell] : 1.0;
}

// ---- analytic line luminosities (outputGas, output_mod.f90:302-520) -----
// One wavefront per block of consecutive cells.  Per model ion (outer loop,
// so that a lane's accumulators stay in registers): every cell's level
// populations (cel_levels), then lanes over the 17 x 17 (upper, lower)
// entries add forLines' emissivity times Hden * getVolume.  Each block writes
// its sums to its own row of `part`; lines_reduce_kernel adds the rows in
// block order, so the result does not depend on scheduling.
struct LineLds {
    double x[kMaxLev + 1][kMaxLev + 1];
    double q[kMaxLev + 1][kMaxLev + 1];
    double y[kMaxLev + 1];
    double alev[kMaxLev + 1];
};

constexpr int kLineR = (kLineEnt + W - 1) / W;       // entries per lane

__device__ __forceinline__ float line_volume(const LineArgs &a, int cell)
{
    // getVolume (grid_mod.f90:2876-2964) of the cell's lattice point: the
    // per-axis widths over 1e15 multiplied in FP32 as (dx * dy) * dz
    const int l = a.cell_xyz[cell];
    const int ix = l % a.nx, iy = (l / a.nx) % a.ny, iz = l / (a.nx * a.ny);
    return (a.dfac[ix] * a.dfac[a.nx + iy]) * a.dfac[a.nx + a.ny + iz];
}

__global__ void __launch_bounds__(64) lines_kernel(const Atomic *__restrict__ Ap, LineArgs a)
{
    __shared__ LineLds S;
    const Atomic &A = *Ap;
    const int lane = threadIdx.x;
    const int nE = A.nE, ns = A.nstages;
    const int c0 = 1 + blockIdx.x * a.cells_per_blk;
    const int c1 = min(a.ncells, c0 + a.cells_per_blk - 1);
    double *out = a.part + (size_t)blockIdx.x * a.ldp;
    // forLines' 1.9865e9 is a REAL literal
    const double kL = (double)1.9865e9f;
    for (int k = 0; k < A.ncel; ++k) {
        const CelIon &ci = A.cel[k];
        const int nl = ci.nl;
        const double *E = A.cel_e + ci.off_e;
        const double *Aa = A.cel_a + ci.off_a;
        const double abun = (double)(float)A.el[ci.eidx].abun;
        double acc[kLineR];
        for (int r = 0; r < kLineR; ++r) acc[r] = 0.0;
        for (int cell = c0; cell <= c1; ++cell) {
            if (a.black && (a.black[cell] & a.black_mask)) continue;
            // the state as the host grid stores it (REAL)
            const float te = (float)a.Te[cell], ne = (float)a.Ne[cell], hd = (float)a.Hden[cell];
            const double *ion = a.ion + ((size_t)cell * nE + ci.eidx) * (ns + 1);
            const float x = (float)ion[ci.ion];
            const float up = (ci.ion < ns) ? (float)ion[ci.ion + 1] / x : 0.f;
            if (up != up) continue;               // equilibrium's NaN test: no emission
            const float dv = line_volume(a, cell);
            cel_levels(A, ci, S, lane, (double)te, (double)ne, (double)up);
            double sumN = 0.0;
            for (int i = 1; i <= nl; ++i) sumN = sumN + S.y[i];
            for (int r = 0; r < kLineR; ++r) {
                const int e = lane + W * r;
                const int hi = e / kMaxLev + 1, lo = e % kMaxLev + 1;
                if (e < kLineEnt && lo < hi && hi <= nl) {
                    const double av = Aa[(hi - 1) + nl * (lo - 1)];
                    if (av != 0.0) {
                        const double Eji = E[hi - 1] - E[lo - 1];
                        double f = av * Eji * (S.y[hi] / sumN);
                        f = f * abun * (double)x;
                        f = f * kL;
                        acc[r] = acc[r] + f * (double)hd * (double)dv;
                    }
                }
            }
            __syncthreads();                      // S is rewritten by the next cell
        }
        for (int r = 0; r < kLineR; ++r) {
            const int e = lane + W * r;
            if (e < kLineEnt) out[(size_t)k * kLineEnt + e] = acc[r];
        }
    }
    // RecLinesEmission's Hbeta (output_mod.f90:2065-2079), HIRecLines(4,2) =
    // (fh/hb) * Hbeta = Hbeta, times Hden * dV
    double hb = 0.0;
    const double abunH = (double)(float)A.el[0].abun;
    for (int cell = c0 + lane; cell <= c1; cell += W) {
        if (a.black && (a.black[cell] & a.black_mask)) continue;
        const float te = (float)a.Te[cell], ne = (float)a.Ne[cell], hd = (float)a.Hden[cell];
        const float xp = (float)a.ion[((size_t)cell * nE) * (ns + 1) + 2];
        float h = 2530.f / powf(te, 0.833f);
        h = h * ne * xp * (float)abunH;
        hb = hb + (double)h * (double)hd * (double)line_volume(a, cell);
    }
    // lanes in order: a fixed summation order
    if (lane == 0) S.y[0] = 0.0;
    __syncthreads();
    for (int l = 0; l < W; ++l) {
        if (lane == l) S.y[0] = S.y[0] + hb;
        __syncthreads();
    }
    if (lane == 0) out[(size_t)A.ncel * kLineEnt] = S.y[0];
}

// out[t] = sum over blocks b (in order) of part[b][t]; four wavefronts per
// 64 outputs, each over every fourth block, combined in wavefront order
__global__ void __launch_bounds__(256) lines_reduce_kernel(const double *__restrict__ part, int ldp, int nblk,
                                                           int n, double *__restrict__ out)
{
    __shared__ double red[4][W];
    const int lane = threadIdx.x % W, w = threadIdx.x / W;
    const int t = blockIdx.x * W + lane;
    double sum = 0.0;
    if (t < n)
        for (int b = w; b < nblk; b += 4) sum = sum + part[(size_t)b * ldp + t];
    red[w][lane] = sum;
    __syncthreads();
    if (w == 0 && t < n) out[t] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

}  // namespace mcb

hipError_t mcb_launch_balance(const mcb::Atomic *dA, const mcb::CellArgs &a, hipStream_t s)
{
    if (a.ncells <= 0) return hipSuccess;
    if (a.grains) {
        // the caller passes the frequency-bin count through grains_nb
        const size_t lds = (size_t)a.grains_nb * sizeof(double) + sizeof(mcb::GrainLds);
        hipLaunchKernelGGL(mcb::balance_kernel<true>, dim3(a.ncells), dim3(64), lds, s, dA, a);
    } else {
        hipLaunchKernelGGL(mcb::balance_kernel<false>, dim3(a.ncells), dim3(64), 0, s, dA, a);
    }
    return hipGetLastError();
}

hipError_t mcb_launch_integrals(const mcb::Atomic *dA, const mcb::CellArgs &a, const double *J3, const int *active,
                               int nx, int ny, int nz, double *Rint, double *Hint, int *hit, float *X, hipStream_t s)
{
    const mcd::J3Shape js = mcd::j3_shape(nx, ny, nz);
    hipLaunchKernelGGL(mcb::integrals_kernel, dim3((unsigned)(mcd::j3_points(js) / mcd::kJTile)), dim3(64), 0, s,
                       dA, a, J3, active, js, Rint, Hint, hit, X);
    return hipGetLastError();
}

hipError_t mcb_launch_warm(const mcb::Atomic *dA, const mcb::CellArgs &a, const float *Y, int ldy, const int *hit,
                           hipStream_t s)
{
    if (a.ncells <= 0) return hipSuccess;
    hipLaunchKernelGGL(mcb::warm_kernel, dim3((a.ncells + 255) / 256), dim3(256), 0, s, dA, a, Y, ldy, hit);
    return hipGetLastError();
}

hipError_t mcb_launch_lines(const mcb::Atomic *dA, const mcb::LineArgs &a, int nblk, int ncel, double *out,
                            hipStream_t s)
{
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(mcb::lines_kernel, dim3(nblk), dim3(64), 0, s, dA, a);
    const int n = ncel * mcb::kLineEnt + 1;
    hipLaunchKernelGGL(mcb::lines_reduce_kernel, dim3((n + 63) / 64), dim3(256), 0, s, a.part, a.ldp, nblk, n, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ host tables
namespace mcb {
namespace {

// Charge-exchange fits (elem, ion, a, b, c, d) with a in units of 1e-9
// cm^3 s^-1, in the order ionBalance assigns them (update_mod.f90:1389-1446):
// the second Ar II/III pair replaces the first, as there.
struct ChexRow { int elem, ion; double a, b, c, d; };
const ChexRow kChex[] = {
    {2, 1, 7.47e-6, 2.06, 9.93, -3.89},    {2, 2, 1.e-5, 0., 0., 0.},
    {3, 2, 1.26, 0.96, 3.02, -0.65},       {3, 3, 1.e-5, 0., 0., 0.},
    {4, 2, 1.e-5, 0., 0., 0.},             {4, 3, 1.e-5, 0., 0., 0.},
    {4, 4, 5.17, 0.82, -.69, -1.12},       {5, 2, 2.e-2, 0., 0., 0.},
    {5, 3, 1.e-5, 0., 0., 0.},             {5, 4, 5.27e-1, 0.76, -0.63, -1.17},
    {6, 1, 1.76e-9, 8.33, 4278.78, -6.41}, {6, 2, 1.67e-4, 2.79, 304.72, -4.07},
    {6, 3, 3.25, 0.21, 0.19, -3.29},       {6, 4, 332.46, -0.11, -0.995, -1.58e-3},
    {7, 1, 1.01e-3, -0.29, -0.92, -8.38},  {7, 2, 3.05e-1, 0.60, 2.65, -0.93},
    {7, 3, 4.54, 0.57, -0.65, -0.89},      {7, 4, 3.28, 0.52, -0.52, -0.19},
    {8, 1, 1.04, 3.15e-2, -0.61, -9.73},   {8, 2, 1.04, 0.27, 2.02, -5.92},
    {8, 3, 3.98, 0.26, 0.56, -2.62},       {8, 4, 2.52e-1, 0.63, 2.08, -4.16},
    {9, 2, 1.e-5, 0., 0., 0.},             {9, 3, 9.86, 0.29, -0.21, -1.15},
    {9, 4, 7.15e-1, 1.21, -0.70, -0.85},   {10, 2, 1.e-5, 0., 0., 0.},
    {10, 3, 14.73, 4.52e-2, -0.84, -0.31}, {10, 4, 6.47, 0.54, 3.59, -5.22},
    {11, 2, 1.e-5, 0., 0., 0.},            {11, 3, 1.33, 1.15, 1.20, -0.32},
    {11, 4, 1.01e-1, 1.34, 10.05, -6.41},  {12, 2, 8.58e-5, 2.49e-3, 2.93e-2, -4.33},
    {12, 3, 6.49, 0.53, 2.82, -7.63},      {12, 4, 6.36, 0.55, 3.86, -5.19},
    {13, 2, 1.e-5, 0., 0., 0.},            {13, 3, 7.11e-5, 4.12, 1.72e4, -22.24},
    {13, 4, 7.52e-1, 0.77, 6.24, -5.67},   {14, 2, 1.23, 0.24, 3.17, 4.18e-3},
    {14, 3, 4.900e-1, -8.74e-2, -0.36, -0.79}, {14, 4, 7.58, 0.37, 1.06, -4.09},
    {16, 1, 3.82e-7, 11.10, 2.57e4, -8.22}, {16, 2, 1.e-5, 0., 0., 0.},
    {16, 3, 2.29, 4.02e-2, 1.59, -6.06},   {16, 4, 6.44, 0.13, 2.69, -5.69},
    {18, 2, 1.e-5, 0., 0., 0.},            {18, 3, 4.57, 0.27, -0.18, -1.57},
    {18, 4, 6.37, 2.12, 10.21, -6.22},     {18, 3, 3.17e-2, 2.12, 12.06, -0.40},
    {18, 4, 2.68, 0.69, -0.68, -4.47},     {26, 2, 1.26, 7.72e-2, -0.41, -7.31},
    {26, 3, 3.42, 0.51, -2.06, -8.99},
};

int err(char *e, size_t n, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(e, n, fmt, ap);
    va_end(ap);
    return code;
}

// natural cubic spline second derivatives (spline, interpolation_mod.f90:
// 182-266, with zero end derivatives of second order)
void spline_nat(const double *xa, const double *ya, int n, double *y2)
{
    double u[kMaxTemp];
    y2[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        const double sig = (xa[i] - xa[i - 1]) / (xa[i + 1] - xa[i - 1]);
        const double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        double v = ya[i + 1];
        v = v - ya[i];
        v = v / (xa[i + 1] - xa[i]);
        v = v - (ya[i] - ya[i - 1]) / (xa[i] - xa[i - 1]);
        u[i] = (6.0 * v / (xa[i + 1] - xa[i - 1]) - sig * u[i - 1]) / p;
    }
    y2[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

// dielectronic rows (direc_coeffs, aldropequi_coeffs) that fill (z, ion)
// (dielectronic, update_mod.f90:1782-1834); returns their number, or -1
// with a message in e
int dr_rows(const mc_atomic *in, int ns, int z, int ion, int *rows, int *ap, char *e, size_t n)
{
    int nr = 0;
    for (int i = 0; i < 25; ++i) {
        const int el = in->dr_int[3 * i], nn = in->dr_int[3 * i + 1];
        const int io = el + 1 - nn;
        if (el != z || io != ion || io > ns || io < 1) continue;
        if (nr == kMaxDr) {
            err(e, n, MC_ETABLE, "more than %d dielectronic rows for Z=%d ion %d", kMaxDr, z, ion);
            return -1;
        }
        rows[nr++] = i;
    }
    *ap = -1;
    for (int i = 0; i < 167; ++i) {
        const int el = in->ap_int[2 * i], nn = in->ap_int[2 * i + 1];
        const int io = el + 1 - nn;
        if (el == z && io == ion && io >= 1 && io <= ns) *ap = i;
    }
    return nr;
}

}  // namespace

int prepare(const mc_atomic *in, int nb, const float *nuArray, Prepared &P, char *e, size_t n)
{
    if (!in || !in->elemOn || !in->xref || !in->abun || !in->ch_meta || !in->ch_sigma || !in->rrec ||
        !in->rnew || !in->fe || !in->dr_int || !in->dr_real || !in->ap_int || !in->ap_real ||
        (in->ncel > 0 && (!in->cel_meta || !in->cel_dbl || !in->cel_flt)))
        return err(e, n, MC_EINVAL, "atomic descriptor has null arrays");
    const int ns = in->nstages;
    if (ns < 2 || ns > kMaxStage - 1) return err(e, n, MC_EUNSUP, "nstages %d outside 2..%d", ns, kMaxStage - 1);
    if (nb > kMaxNb || nb < mcm::kBands)
        return err(e, n, MC_EUNSUP, "nbins %d outside %d..%d", nb, mcm::kBands, kMaxNb);
    if (!in->elemOn[0] || !in->elemOn[1]) return err(e, n, MC_EUNSUP, "H and He must be switched on");
    Atomic &A = P.A;
    A = Atomic{};
    A.nbins = nb;
    A.nstages = ns;
    A.XHILimit = in->XHILimit;
    P.nu.assign(nuArray, nuArray + nb);
    // elements in atomic-number order, their channels ion = 1..min(z, ns-1)
    int eidx_of_z[31];
    for (int z = 0; z <= 30; ++z) eidx_of_z[z] = -1;
    int nE = 0, nch = 0;
    P.eidx_xref.clear();
    for (int z = 1; z <= 30; ++z) {
        if (!in->elemOn[z - 1]) continue;
        if (nE == kMaxElem) return err(e, n, MC_EUNSUP, "more than %d elements", kMaxElem);
        const int xr = in->xref[z - 1];
        if (xr < 1 || xr > in->nElemUsed) return err(e, n, MC_ETABLE, "elementXref(%d) = %d", z, xr);
        Element &E = A.el[nE];
        E.z = z;
        E.nst = std::min(z + 1, ns);
        E.ch0 = nch;
        E.nch = std::min(z, ns - 1);
        E.abun = (double)in->abun[z - 1];
        if (nch + E.nch > kMaxCh) return err(e, n, MC_EUNSUP, "more than %d channels", kMaxCh);
        for (int ion = 1; ion <= E.nch; ++ion) {
            Channel &c = A.ch[nch];
            c.elem = z; c.ion = ion; c.eidx = nE;
            c.ip = 1; c.hiR = 0; c.hiH = 0;
            int g0 = 0, g1 = 0, row = -1;
            for (int k = 0; k < in->nch; ++k)
                if (in->ch_meta[8 * k] == z && in->ch_meta[8 * k + 1] == ion) row = k;
            std::vector<double> sig(nb, 0.0);
            if (row >= 0) {
                const int32_t *m = in->ch_meta + 8 * row;
                g0 = m[2]; g1 = m[3];
                c.ip = m[4]; c.hiR = m[5]; c.hiH = m[6];
                if (c.ip < 1 || c.ip > nb || c.hiR > nb || c.hiH > nb)
                    return err(e, n, MC_ETABLE, "channel Z=%d ion %d: bins %d..%d/%d outside 1..%d", z, ion,
                               c.ip, c.hiR, c.hiH, nb);
                for (int j = 0; j < nb; ++j) sig[j] = (double)in->ch_sigma[(size_t)row * nb + j];
                // the heating integral stops at the first cross section below 1e-35
                for (int j = c.ip; j <= c.hiH; ++j)
                    if (sig[j - 1] < 1.e-35) { c.hiH = j - 1; break; }
            }
            // coefficients of the rate and heating integrals per bin:
            // rate J sigma / (h c Ryd nu) over ip..hiRate where sigma >= 1e-35,
            // heating J sigma (nu - nu_ip) / nu over ip..hiHeat
            {
                const double hcRyd = 2.1799153e-11;
                const double nuIP = (double)nuArray[c.ip - 1];
                for (int j = 1; j <= nb; ++j) {
                    const double nu = (double)nuArray[j - 1], ph = sig[j - 1];
                    P.cR.push_back(j >= c.ip && j <= c.hiR && ph >= 1.e-35 ? ph / (hcRyd * nu) : 0.0);
                    P.cH.push_back(j >= c.ip && j <= c.hiH ? ph * (nu - nuIP) / nu : 0.0);
                }
            }
            P.sigma.insert(P.sigma.end(), sig.begin(), sig.end());
            c.g0g1 = (double)g0 / (double)g1;
            // radiative recombination fit (radRecFit, update_mod.f90:1674-1780)
            const int nn = z - ion + 1;
            if (nn <= 3 || nn == 11 || (z > 5 && z < 9) || z == 10 || (z == 26 && nn > 11)) {
                c.rr_kind = 0;
                for (int i = 0; i < 4; ++i) c.rr[i] = (double)in->rnew[(z - 1) * 120 + (nn - 1) * 4 + i];
            } else if (z == 26 && nn <= 13) {
                c.rr_kind = 2;
                for (int i = 0; i < 3; ++i) c.rr[i] = (double)in->fe[(nn - 4) * 3 + i];
            } else {
                c.rr_kind = 1;
                for (int i = 0; i < 2; ++i) c.rr[i] = (double)in->rrec[(z - 1) * 60 + (nn - 1) * 2 + i];
            }
            const int nr = dr_rows(in, ns, z, ion, c.dr, &c.ap, e, n);
            if (nr < 0) return MC_ETABLE;
            c.ndr = nr;
            // charge exchange
            for (const ChexRow &r : kChex)
                if (r.elem == z && r.ion == ion) {
                    c.chex[0] = r.a * 1.e-9; c.chex[1] = r.b; c.chex[2] = r.c; c.chex[3] = r.d;
                }
            c.dEk = (z == 7 && ion == 1) ? 10863. : (z == 8 && ion == 1) ? 2205. : 0.;
            ++nch;
        }
        eidx_of_z[z] = nE;
        P.eidx_xref.push_back(xr);
        ++nE;
    }
    A.nE = nE;
    A.nch = nch;
    if (A.ch[0].elem != 1 || A.ch[0].ion != 1) return err(e, n, MC_EUNSUP, "first channel is not H I");
    for (int ion = 1; ion < ns; ++ion) {
        const int nr = dr_rows(in, ns, 8, ion, A.odr[ion], &A.oap[ion], e, n);
        if (nr < 0) return MC_ETABLE;
        A.odr_n[ion] = nr;
    }
    for (int i = 0; i < 25; ++i) {
        for (int k = 0; k < 5; ++k) A.dr_real[i][k] = (double)in->dr_real[5 * i + k];
        A.dr_g[i] = in->dr_int[3 * i + 2];
    }
    for (int i = 0; i < 167; ++i)
        for (int k = 0; k < 4; ++k) A.ap_real[i][k] = (double)in->ap_real[4 * i + k];
    // CEL model ions
    if (in->ncel < 0 || in->ncel > kMaxCel) return err(e, n, MC_EUNSUP, "more than %d CEL ions", kMaxCel);
    A.ncel = in->ncel;
    size_t od = 0, of = 0;
    for (int k = 0; k < in->ncel; ++k) {
        const int32_t *m = in->cel_meta + 6 * k;
        const int z = m[0], ion = m[1], nl = m[4], nt = m[5];
        if (z < 1 || z > 30 || eidx_of_z[z] < 0 || ion < 1 || ion > ns)
            return err(e, n, MC_ETABLE, "CEL ion %d: Z=%d ion %d not switched on", k, z, ion);
        if (nl < 2 || nl > kMaxLev || nt < 1 || nt > kMaxTemp)
            return err(e, n, MC_EUNSUP, "CEL ion %d: %d levels, %d temperatures", k, nl, nt);
        const double *ev = in->cel_dbl + od;
        const double *av = ev + nl;
        const double *qv = av + (size_t)nl * nl;
        const float *lt = in->cel_flt + of;
        const float *gv = lt + nt;
        CelIon &ci = A.cel[k];
        ci.eidx = eidx_of_z[z]; ci.ion = ion; ci.nl = nl; ci.nt = nt;
        ci.br = (double)gv[nl];
        ci.z = (double)gv[nl + 1];
        for (int i = 0; i < 4; ++i) ci.a_r[i] = (double)gv[nl + 2 + i];
        for (int i = 0; i < 5; ++i) ci.a_d[i] = (double)gv[nl + 6 + i];
        ci.off_e = (int)P.cel_e.size();
        ci.off_a = (int)P.cel_a.size();
        ci.off_q = (int)P.cel_q.size();
        ci.off_t = (int)P.cel_t.size();
        ci.off_g = (int)P.cel_g.size();
        P.cel_e.insert(P.cel_e.end(), ev, ev + nl);
        P.cel_a.insert(P.cel_a.end(), av, av + (size_t)nl * nl);
        std::vector<double> xa(nt);
        for (int t = 0; t < nt; ++t) xa[t] = (double)lt[t];
        for (int t = 1; t < nt; ++t)
            if (!(xa[t] > xa[t - 1])) return err(e, n, MC_ETABLE, "CEL ion %d: log T not increasing", k);
        P.cel_t.insert(P.cel_t.end(), xa.begin(), xa.end());
        for (int i = 0; i < nl; ++i) P.cel_g.push_back((double)gv[i]);
        // collision strengths of each pair (lo < up) and their spline curvature
        double qq[kMaxTemp], y2[kMaxTemp];
        for (int lo = 1; lo <= nl - 1; ++lo)
            for (int up = lo + 1; up <= nl; ++up) {
                for (int t = 0; t < nt; ++t) qq[t] = qv[t + (size_t)nt * ((lo - 1) + (size_t)nl * (up - 1))];
                if (nt >= 3) spline_nat(xa.data(), qq, nt, y2);
                else for (int t = 0; t < nt; ++t) y2[t] = 0.0;
                P.cel_q.insert(P.cel_q.end(), qq, qq + nt);
                P.cel_y2.insert(P.cel_y2.end(), y2, y2 + nt);
            }
        od += (size_t)nl + (size_t)nl * nl + (size_t)nt * nl * nl;
        of += (size_t)nt + nl + 2 + 4 + 5;
    }
    return MC_OK;
}

}  // namespace mcb

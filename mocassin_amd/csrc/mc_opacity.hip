// mc_opacity.hip - opacity tables on the device (ionizationDriver and
// addOpacity, ionization_mod.f90:26-484; the dust add of iterateMC,
// iteration_mod.f90:164-228), written straight into the transport's dense
// [nu][xyz] tables.
//
// One lane per lattice point, one wavefront per 64 consecutive points, so
// every bin's row of opac3 is written in one coalesced 256-byte store.  A
// lane forms its cell's ion densities in its LDS column once, then walks the
// bins; each bin's cross-section entries are the same for every lane
// (scalar loads), in addOpacity's order, and are added in FP32 in that
// order, so the tables equal the reference's bit for bit.  Bin 1 also
// carries the free-free term: addOpacity assigns FFOpacity at i = 1 only
// (ionization_mod.f90:369-393), with the Gaunt factor of the cell's own Te
// (Hummer 1988, getGauntFF, 195-318).
#include "mc_opacity.h"

#pragma clang fp contract(off)

namespace mco {

constexpr float kTe1Ryd = 1.578866e5f;   // constants_mod.f90:23

// Hummer's Chebyshev coefficients d(8,11), column-major (ionization_mod.f90:213-242)
__constant__ double kHummer[88] = {
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

// getGauntFF at one frequency for nuclear charge 10^z (ionization_mod.f90:195-318)
__device__ float gaunt_ff(float z, float log10Te, float xlf)
{
    float xlrkt;
    if (log10Te < 2.5f) xlrkt = 5.1983649f - 2.5f;
    else if (log10Te > 8.f) xlrkt = 5.1983649f - 8.f;
    else xlrkt = 5.1983649f - log10Te;
    const float txg = 0.66666667f * (2.0f * z + xlrkt);
    const float gamma2 = powf(10.f, txg * 1.5f);
    const float con = 0.72727273f * xlrkt + 0.90909091f;
    float c[9];
    for (int j = 1; j <= 8; ++j) {
        // Clenshaw recurrence b(ir) over ir = 11..1, three terms live
        float b2 = (float)kHummer[10 * 8 + j - 1];                              // b(11)
        float b1 = (float)((double)(txg * b2) + kHummer[9 * 8 + j - 1]);        // b(10)
        float bb1 = 0.f, bb3 = 0.f;
        for (int ir = 9; ir >= 1; --ir) {
            const float b0 = (float)((double)(txg * b1 - b2) + kHummer[(ir - 1) * 8 + j - 1]);
            if (ir == 3) bb3 = b0;
            if (ir == 1) bb1 = b0;
            b2 = b1;
            b1 = b0;
        }
        c[j] = 0.25f * (bb1 - bb3);
    }
    const float txu = 0.72727273f * xlf + con;
    const float u = powf(10.f, (txu - .90909091f) / .72727273f);
    float g = 0.f;
    if (fabsf(txu) <= 2.0f) {
        float b2 = c[8];
        float b1 = txu * b2 + c[7];
        float bb1 = 0.f, bb3 = 0.f;
        for (int ir = 6; ir >= 1; --ir) {
            const float b0 = txu * b1 - b2 + c[ir];
            if (ir == 3) bb3 = b0;
            if (ir == 1) bb1 = b0;
            b2 = b1;
            b1 = b0;
        }
        g = bb1 - bb3;
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
}

__global__ void __launch_bounds__(64) opacity_kernel(OpacArgs a)
{
    extern __shared__ float lds[];               // [nslot][64] densities, then [ndrow][64] weights
    const int lane = threadIdx.x;
    // kSplit wavefronts share a tile of 64 lattice points, each writing one
    // contiguous range of bins (more wavefronts in flight to hide latency)
    const int tile = blockIdx.x / kSplit, part = blockIdx.x % kSplit;
    const int xyz = mcd::j3_lattice(a.js, (unsigned)(tile * 64 + lane));   // -1: a padding point
    const bool in = xyz >= 0;
    const int cell = in ? a.active[xyz] : 0;
    float *dens = lds, *wr = lds + a.nslot * 64;
    float ff1 = 0.f;
    if (cell > 0) {
        const float *s = a.st + (size_t)cell * a.ldst;
        const float T = s[0], N = s[1], H = s[2];
        float eSum = 0.f, d22 = 0.f;
        for (int q = 0; q < a.nslot; ++q) {
            const float d = s[3 + q] * a.slot_abun[q] * H;
            dens[q * 64 + lane] = d;
            if (a.slot_efac[q] != 0.f) eSum = eSum + a.slot_efac[q] * d;
            if (q == a.slot_he2) d22 = d;
        }
        if (part == 0) {                          // bin 1's free-free term
            const float log10Te = log10f(T);
            const float sqrTe = sqrtf(T);
            const float expo = (kTe1Ryd / T) + a.nu1;
            const float cb1 = expo > 84.f ? 0.f : expf(-expo);
            const float g1 = gaunt_ff(0.f, log10Te, a.xlf1);
            // (density(1,2) + density(2,2) + 4 density(2,3) + eDenFFSum): the
            // H+ and He++ densities are never set in ionizationDriver's loop
            float t = 0.f + d22;
            t = t + 4.f * 0.f;
            t = t + eSum;
            const float fac1 = (N / 1.e10f) * (t / 1.e10f) / sqrTe;
            const float fac2 = fac1 * kTe1Ryd / T;
            if (cb1 < 0.995f) ff1 = (a.brems1 * (1.f - cb1) * g1) * fac1;
            else ff1 = (a.brems1 * a.nu1 * g1) * fac2;
        }
        if (a.ndrow) {
            const float *Td = a.Tdust + (size_t)cell * (a.nsMax + 1) * (a.nSizes + 1);
            for (int r = 0; r < a.ndrow; ++r) {
                const int sp = r / a.nSizes + 1, ai = r % a.nSizes + 1;
                const float td = Td[sp + (a.nsMax + 1) * ai];
                wr[r * 64 + lane] = (td < a.tsub[r]) ? a.dw[r] * a.ndust[cell] : -1.f;
            }
        }
    }
    const int per = (a.nb + kSplit - 1) / kSplit;
    const int i0 = part * per, i1 = min(a.nb, i0 + per);
    const Ent *ent = a.ent;
    for (int i = i0; i < i1; ++i) {
        float acc = 0.f, sca = 0.f;
        if (cell > 0) {
            acc = acc + (i == 0 ? ff1 : 0.f);
            const int e1 = a.bin_start[i + 1];
            int e = a.bin_start[i];
            // four entries' (slot, kind, sigma) fetched ahead of the
            // dependent FP32 sum, which stays in the reference's order
            for (; e + 4 <= e1; e += 4) {
                Ent en[4];
                float d[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) en[k] = ent[e + k];
#pragma unroll
                for (int k = 0; k < 4; ++k) d[k] = dens[(en[k].qk >> 1) * 64 + lane];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (!((en[k].qk & 1) && !(d[k] > 0.f))) acc = acc + en[k].sigma * d[k];
            }
            for (; e < e1; ++e) {
                const Ent en = ent[e];
                const float d = dens[(en.qk >> 1) * 64 + lane];
                if ((en.qk & 1) && !(d > 0.f)) continue;
                acc = acc + en.sigma * d;
            }
            if (a.ndrow) {
                float ab = 0.f;
                for (int r = 0; r < a.ndrow; ++r) {
                    const float w = wr[r * 64 + lane];
                    if (w < 0.f) continue;
                    sca = sca + w * a.dsca[(size_t)r * a.nb + i];
                    ab = ab + w * a.dabs[(size_t)r * a.nb + i];
                }
                acc = acc + (sca + ab);
            }
        }
        const size_t o = ((size_t)tile * a.nb + i) * 64 + lane;   // padding points get 0
        a.opac3[o] = acc;
        if (a.sca3) a.sca3[o] = sca;
    }
}

__global__ void state_kernel(const double *Te, const double *Ne, const double *Hden, const double *ion, int nE,
                             int nst1, const int *slot_e, const int *slot_stage, int nslot, int nc, float *st,
                             int ldst)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell > nc) return;
    float *s = st + (size_t)cell * ldst;
    s[0] = (float)Te[cell];
    s[1] = (float)Ne[cell];
    s[2] = (float)Hden[cell];
    const double *x = ion + (size_t)cell * nE * nst1;
    for (int q = 0; q < nslot; ++q) s[3 + q] = (float)x[slot_e[q] * nst1 + slot_stage[q]];
}

__global__ void f32_kernel(const double *in, size_t n, float *out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

}  // namespace mco

hipError_t mco_launch_opacity(const mco::OpacArgs &a, size_t lds_bytes, hipStream_t s)
{
    const unsigned tiles = (unsigned)(mcd::j3_points(a.js) / 64);
    hipLaunchKernelGGL(mco::opacity_kernel, dim3(tiles * mco::kSplit), dim3(64), lds_bytes, s, a);
    return hipGetLastError();
}

hipError_t mco_launch_state(const double *Te, const double *Ne, const double *Hden, const double *ion, int nE,
                            int nst1, const int *slot_e, const int *slot_stage, int nslot, int nc, float *st,
                            int ldst, hipStream_t s)
{
    hipLaunchKernelGGL(mco::state_kernel, dim3((nc + 1 + 255) / 256), dim3(256), 0, s, Te, Ne, Hden, ion, nE, nst1,
                       slot_e, slot_stage, nslot, nc, st, ldst);
    return hipGetLastError();
}

hipError_t mco_launch_f32(const double *in, size_t n, float *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mco::f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

namespace mco {
__global__ void gather_kernel(const float *t3, mcd::J3Shape js, const int *cell_xyz, int nc, int nb, float *out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nc1 = (size_t)nc + 1;
    if (i >= nc1 * nb) return;
    const size_t cell = i % nc1, nu = i / nc1;
    out[i] = cell == 0 ? 0.f : t3[mcd::j3_index(nb, (int)nu, mcd::j3_point_of(js, cell_xyz[cell]))];
}
}  // namespace mco

hipError_t mco_launch_gather(const float *t3, const mcd::J3Shape &js, const int *cell_xyz, int nc, int nb,
                             float *out, hipStream_t s)
{
    const size_t n = ((size_t)nc + 1) * nb;
    hipLaunchKernelGGL(mco::gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t3, js, cell_xyz, nc,
                       nb, out);
    return hipGetLastError();
}

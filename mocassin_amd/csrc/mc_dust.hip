// mc_dust.hip - grain temperatures (getDustT, source/update_mod.f90:1836-1945).
//
// One lane per lattice point, one wavefront per tile of 64 points of the
// tiled J3 estimator, so the wavefront reads each bin's 64 values in one
// 512-byte row.  The absorption cross-sections of every (species, size) are
// staged through LDS 64 bins at a time and shared by the 64 lanes; each lane
// keeps its G absorption integrals in registers and reads its cell's J once.
// The temperature is then located in the row of emission integrals
// (1..nTemps K) by binary search where the row never decreases (the usual
// case: the emission integral grows with T), else by the reference's
// minloc scan, and interpolated linearly.
//
// FP64 throughout, contraction off: every sum and product is in the order
// of the reference's loops, so the result equals the FP64 restatement of
// getDustT bit for bit.
#include "mc_device.h"
#include "mc_dust.h"

namespace mcg {

// constants_mod.f90:32 declares Pi as a REAL: 3.141592654 rounded to FP32
constexpr double kPiRef = (double)3.141592654f;

// locate(xa(1:n), x), interpolation_mod.f90:48-81: n when x > xa(n), 0 when
// x < xa(1), else max(p - 1, 1) where p is the 1-based position of the
// smallest xa(i) - x > 0 (p = 0 when there is none, so x == xa(n) gives 1)
__device__ int locate_row(const double *__restrict__ xa, int n, int mono, double x)
{
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    int p = 0;
    if (mono) {
        // first i with xa[i] > x: in a row that never decreases, the
        // smallest positive difference and its first occurrence
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (xa[mid] > x) hi = mid; else lo = mid + 1;
        }
        p = lo < n ? lo + 1 : 0;
    } else {
        double bd = 0.0;
        for (int i = 0; i < n; ++i) {
            const double d = xa[i] - x;
            if (d > 0.0 && (p == 0 || d < bd)) { p = i + 1; bd = d; }
        }
    }
    return p - 1 > 1 ? p - 1 : 1;
}

template <int G>
__global__ void __launch_bounds__(64) dust_kernel(DustArgs a, const double *__restrict__ J3,
                                                  const int *__restrict__ active)
{
    __shared__ double sx[G][64];              // cross-sections of 64 bins, G rows
    const int lane = threadIdx.x;
    const mcd::J3Shape js = mcd::j3_shape(a.nx, a.ny, a.nz);
    const int xyz = mcd::j3_lattice(js, blockIdx.x * 64u + (unsigned)lane);
    const int cell = xyz >= 0 ? active[xyz] : 0;
    if (!__any(cell > 0)) return;             // the whole tile is inactive
    const int nb = a.nb, K = a.nS * a.nSizes;
    const double *Jp = J3 + mcd::j3_index(nb, 0, blockIdx.x * 64u) + lane;
    double acc[G];
#pragma unroll
    for (int k = 0; k < G; ++k) acc[k] = 0.0;
    int hit = 0;
    for (int c0 = 0; c0 < nb; c0 += 64) {
        const int cn = nb - c0 < 64 ? nb - c0 : 64;
        __syncthreads();
        for (int k = 0; k < G; ++k) sx[k][lane] = (k < K && lane < cn) ? a.absx[(size_t)k * nb + c0 + lane] : 0.0;
        __syncthreads();
        for (int u = 0; u < cn; ++u) {
            const double J = Jp[(size_t)(c0 + u) * mcd::kJTile] * a.jscale;
            hit |= J > 0.0;
            const double rf = J / kPiRef;                 // radField = Jste / Pi
#pragma unroll
            for (int k = 0; k < G; ++k) acc[k] = acc[k] + sx[k][u] * rf;
        }
    }
    // park the integrals in this lane's column of the LDS buffer, so the
    // per-row work below indexes them at run time without private memory
    __syncthreads();
#pragma unroll
    for (int k = 0; k < G; ++k) sx[k][lane] = acc[k];
    if (cell <= 0) return;
    if (!hit) {
        if (a.flags) a.flags[cell] = 16;
        return;
    }
    const int ld = a.nsMax + 1;
    double *T = a.Tdust + (size_t)cell * ld * (a.nSizes + 1);
    const double old = T[0];
    for (int i = 0; i < ld * (a.nSizes + 1); ++i) T[i] = 0.0;
    double t00 = 0.0;
    for (int k = 0; k < K; ++k) {
        const int s = k / a.nSizes + 1, ai = k % a.nSizes + 1;
        double x = sx[k][lane];
        if (a.gdc) x = x + a.gdc[(size_t)cell * K + k] - a.peh[(size_t)cell * K + k];
        const double *xa = a.em + (size_t)k * a.nTemps;
        const int n = a.nTemps;
        const int iT = locate_row(xa, n, a.mono[k], x);
        double td;
        if (iT <= 0) td = 1.0;
        else if (iT >= n) td = (double)n;
        else td = (double)iT + (x - xa[iT - 1]) * ((double)(iT + 1) - (double)iT) / (xa[iT] - xa[iT - 1]);
        T[s + ld * ai] = td;
        T[s] = T[s] + td * a.weight[ai - 1];
        if (ai == a.nSizes) t00 = t00 + T[s] * a.abun[s - 1];
    }
    T[0] = t00;
    if (a.flags) {
        const double d = (t00 - old) / old;
        a.flags[cell] = fabs(d) <= a.XHILimit ? 1 : 0;
    }
}

}  // namespace mcg

hipError_t mcg_launch_dust(const mcg::DustArgs &a, const double *J3, const int *active, hipStream_t s)
{
    const int K = a.nS * a.nSizes;
    const dim3 grid((unsigned)(mcd::j3_points(mcd::j3_shape(a.nx, a.ny, a.nz)) / mcd::kJTile)), block(64);
    if (K <= 8) mcg::dust_kernel<8><<<grid, block, 0, s>>>(a, J3, active);
    else if (K <= 16) mcg::dust_kernel<16><<<grid, block, 0, s>>>(a, J3, active);
    else mcg::dust_kernel<32><<<grid, block, 0, s>>>(a, J3, active);
    return hipGetLastError();
}

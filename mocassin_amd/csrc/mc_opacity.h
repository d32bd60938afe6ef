// mc_opacity.h - the opacity tables of every lattice point on the device, in
// the reference's FP32 (ionizationDriver/addOpacity, ionization_mod.f90:
// 26-484, and the dust add of iterateMC, iteration_mod.f90:164-228).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mc_device.h"

namespace mco {

constexpr int kMaxSlots = 160;      // (element, ion) densities per cell
constexpr int kMaxDustRows = 32;    // grain (species, size) rows
constexpr int kSplit = 4;           // wavefronts per tile of 64 lattice points

// one cross-section entry of a bin: qk = slot * 2 + (1 for a heavy ion,
// added only when its density is > 0)
struct Ent {
    int qk;
    float sigma;
};

struct OpacArgs {
    int ncell3, nb, nc;
    mcd::J3Shape js;                 // the tables' layout (J3 tiles of 64 lattice points)
    const int *active;               // lattice point -> cell (0 inactive)
    // per-cell state, FP32: st[cell * ldst + 0..2] = Te, Ne, Hden; then the
    // ionic fractions of slot q at st[cell * ldst + 3 + q]
    const float *st;
    int ldst, nslot;
    const float *slot_abun;          // [nslot] elemental abundance of the slot's element
    const float *slot_efac;          // [nslot]: (stage-1)^2 for heavy ions above neutral, else 0
                                     // (eDenSum), in eDenSum's order
    int slot_he2;                    // slot of He+ (density(2,2) of fac1), -1 if absent
    // bins: entries [bin_start[i], bin_start[i+1]) of bin i (0-based), in
    // addOpacity's order
    const int *bin_start;
    const Ent *ent;
    float brems1, nu1, xlf1;         // free-free factor, nu(1), log10 nu(1)
    // grains (nullable): rows k = (species-1)*nSizes + size-1
    int ndrow, nSizes, nsMax;
    const float *dsca, *dabs;        // [ndrow][nb]
    const float *dw;                 // [ndrow] grainAbun * grainWeight
    const float *tsub;               // [ndrow] sublimation temperature of the row's species
    const float *ndust;              // (0:nc)
    const float *Tdust;              // (0:nsMax, 0:nSizes, 0:nc), FP32
    float *opac3, *sca3;             // outputs in the J3 layout (sca3 only with grains)
};

}  // namespace mco

hipError_t mco_launch_opacity(const mco::OpacArgs &a, size_t lds_bytes, hipStream_t s);
// st rows from the device balance state (FP64 Te, Ne, Hden, ion[cell][e][stage]),
// rounded to FP32 as the host stores them
hipError_t mco_launch_state(const double *Te, const double *Ne, const double *Hden, const double *ion, int nE,
                            int nst1, const int *slot_e, const int *slot_stage, int nslot, int nc, float *st,
                            int ldst, hipStream_t s);
// FP64 Tdust -> FP32
hipError_t mco_launch_f32(const double *in, size_t n, float *out, hipStream_t s);
// (0:nc, 1:nb) column-major copy of a J3-layout table (row 0 zero)
hipError_t mco_launch_gather(const float *t3, const mcd::J3Shape &js, const int *cell_xyz, int nc, int nb,
                             float *out, hipStream_t s);

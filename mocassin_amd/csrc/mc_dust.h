// mc_dust.h - grain temperatures of every reached cell on the device
// (getDustT, source/update_mod.f90:1836-1945; locate, interpolation_mod.f90:48-81).
#pragma once
#include <hip/hip_runtime.h>

namespace mcg {

// Absorption integrals per (species, size) in one pass over a cell's bins.
constexpr int kMaxRows = 32;          // nSpecies * nSizes
constexpr int kMaxTemps = 8192;       // length of an emission-integral row

struct DustArgs {
    int nx, ny, nz, nb, nc;           // lattice, frequency bins, cells
    int nS, nsMax, nSizes, nTemps;    // Tdust is (0:nsMax, 0:nSizes, 0:nc), column-major
    double jscale;                    // J = J3 * jscale (the host's normalisation of Jste)
    const double *absx;               // [nS*nSizes][nb], row (s-1)*nSizes + ai-1
    const double *em;                 // [nS*nSizes][nTemps]: emission integral at 1..nTemps K
    const int *mono;                  // [nS*nSizes]: 1 where the row never decreases
    const double *weight, *abun;      // grainWeight(1:nSizes), grainAbun(1:nS)
    const double *gdc, *peh;          // nullable [nc+1][nS*nSizes]: the absorption integral
                                      // becomes (integral + gdc) - peh (gas-grain collisions,
                                      // photoelectric loss; update_mod.f90:1886-1889)
    double *Tdust;                    // in/out; reached cells rewritten
    int *flags;                       // nullable (0:nc): bit0 converged, bit4 not reached
    double XHILimit;                  // convergence of Tdust(0,0) (grains without gas)
};

}  // namespace mcg

// Launch over the J3 tiles; the caller has checked the sizes against kMaxRows,
// kMaxTemps and the J3 layout.
hipError_t mcg_launch_dust(const mcg::DustArgs &a, const double *J3, const int *active, hipStream_t s);

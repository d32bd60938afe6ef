// mc_balance.h - device tables of the per-cell balance (updateCell), the
// host-side preparation that builds them from the C-ABI atomic descriptor,
// and the launcher of the balance kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/mocassin_amd.h"

namespace mcb {

constexpr int kMaxCh = 64;      // photoionisation channels (elem, ion)
constexpr int kMaxElem = 16;    // elements switched on
constexpr int kMaxStage = 10;   // ionisation stages + 1 (stage index 1..nstages)
constexpr int kMaxCel = 64;     // CEL model ions
constexpr int kMaxLev = 17;     // levels per model ion (nForLevels)
constexpr int kMaxTemp = 64;    // temperatures per collision-strength table
constexpr int kMaxDr = 4;       // dielectronic fit rows per channel
constexpr int kMaxNb = 65536;   // frequency bins

// one photoionisation channel (elem, ion) in the reference's loop order
struct Channel {
    int elem, ion, eidx;        // atomic number, stage, element index (0..nE-1)
    int ip, hiR, hiH;           // 1-based bins: rate sum ip..hiR, heat sum ip..hiH
    int rr_kind;                // 0 Verner-Ferland (4), 1 power law (2), 2 Fe fit (3)
    double rr[4];
    int ndr, dr[kMaxDr];        // dielectronic rows (direc_coeffs) filling (elem, ion)
    int ap;                     // last aldropequi row filling (elem, ion), -1 if none
    double chex[4], dEk;        // charge-exchange fit and reverse-reaction deltaE/k
    double g0g1;                // real(g0)/real(g1)
};

struct Element {
    int z, nst;                 // atomic number, stages written (min(z+1, nstages))
    int ch0, nch;               // its channels (stages 1..min(z, nstages-1))
    double abun;
};

struct CelIon {
    int eidx, ion, nl, nt;
    double br, z, a_r[4], a_d[5];
    int off_e, off_a, off_q, off_t, off_g;   // offsets: e, A(nl,nl), qom/y2(nt,npair), logT, g
};

struct Atomic {
    int nbins, nstages, nE, nch, ncel;
    float XHILimit;
    Channel ch[kMaxCh];
    Element el[kMaxElem];
    int odr_n[kMaxStage], odr[kMaxStage][kMaxDr], oap[kMaxStage];   // oxygen rows (fallback)
    double dr_real[25][5];
    int dr_g[25];
    double ap_real[167][4];
    CelIon cel[kMaxCel];
    // device arrays
    const double *sigma;        // [nch][nbins]
    const double *nu;           // [nbins]
    const double *cR, *cH;      // [nch][nbins] rate and heating coefficients per bin
    const double *cel_e, *cel_a, *cel_q, *cel_y2, *cel_t, *cel_g;
};

// Grains of one dust chemistry for the grain charge, photoelectric heating
// and gas-grain collision terms (update_mod.f90:527-932); rows k = (species
// - 1) * nSizes + size - 1.
constexpr int kMaxGrainRows = 32;
struct Grains {
    int nS, nSizes, nsMax, K;
    double vn[kMaxGrainRows], msurf[kMaxGrainRows];       // grainVn, MsurfAtom per species
    double gabun[kMaxGrainRows];                           // grainAbun per species
    double radius[kMaxGrainRows], weight[kMaxGrainRows];   // grainRadius [um], grainWeight per size
    double aweight[30];                                    // aWeight by atomic number - 1
    const double *absx;         // [K][nbins] absorption cross-sections
    const double *qa;           // [K][nbins] absx / (1e-8 radius^2)
    const double *yn;           // [nS][nbins] min(0.5 (1 - vn/nu), 0.2)
    const double *ndust;        // Ndust(0:nCells)
    const double *Tdust;        // Tdust(0:nsMax, 0:nSizes, 0:nCells) before the update
    double *gdc, *peh;          // [nCells+1][K]: gasDustColl_d, photoelHeat_d of the last T-iteration
    const double *J3;           // the estimator, tiled (mcd::j3_index)
    const int *cell_xyz;        // lattice point of each cell, (x-1) + nx*((y-1) + ny*(z-1))
    int nx, ny, nz;             // the lattice (J3's tiling)
};

struct CellArgs {
    int ncells;                 // active cells (1..nCells)
    double jscale;              // J = J3 * jscale
    const double *Rint, *Hint;  // [nCells+1][nch] channel integrals (integrals_kernel)
    const int *hit;             // [nCells+1] reached by a packet
    const double *Hden;         // [nCells+1]
    double *Te, *Ne;            // [nCells+1] in/out
    double *ion;                // [nCells+1][nE][nstages+1], stage 0 unused
    int *flags, *nIter;         // [nCells+1]
    // temperature solver: 0 the reference's schedule (iterateT), 1 secant
    // iteration to |heat - cool| <= tol * heat in at most max_iter steps
    int solver, max_iter;
    double tol, ion_tol;        // ion_tol: the ion fixed point's relative change (reference 0.01)
    double step0;               // secant: relative size of the first trial step
    double slope0;              // secant: expected |d((heat - cool)/heat)/d ln T| for that step
    int warm_iter;              // secant: iterations of a warm-started attempt before its restart
    // state to restart from when the solve from Te/Ne/ion fails (nullable)
    const double *Te0, *Ne0, *ion0;
    // grains with the photoelectric effect (nullable: gas alone); device
    // pointer, with the frequency-bin count for the launch's LDS size
    const Grains *grains;
    int grains_nb;
    // nullable: block b balances cell order[b] (a permutation of 1..ncells),
    // else cell b + 1; results do not depend on the order
    const int *order;
};

// Line luminosities of a state (outputGas's analytic forbidden lines and
// Hbeta, output_mod.f90:302-520).  Cells are split into blocks of
// cells_per_blk consecutive cells, one wavefront each; block b writes its
// sums to part[b * ldp ...]: per model ion k the 17 x 17 entries (upper - 1)
// * 17 + lower - 1 at k * 289, then Hbeta at ncel * 289 (ldp >= that + 1).
constexpr int kLineEnt = kMaxLev * kMaxLev;   // entries per model ion
struct LineArgs {
    int ncells, cells_per_blk;
    const double *Te, *Ne, *Hden;   // [nCells+1], rounded to FP32 as the host grid stores them
    const double *ion;              // [nCells+1][nE][nstages+1]
    const int *black;               // nullable: cell c is skipped when black[c] & black_mask
    int black_mask;
    const int *cell_xyz;            // lattice point of each cell, (x-1) + nx*((y-1) + ny*(z-1))
    int nx, ny;
    const float *dfac;              // per-axis getVolume widths / 1e15, axes x, y, z concatenated
    double *part;
    int ldp;
};

// Host-built tables: `A` with its device pointers unset, and the flat arrays
// they will point to.  eidx_xref[e] is the elementXref (1-based) of element
// index e.
struct Prepared {
    Atomic A{};
    std::vector<double> sigma, nu, cR, cH, cel_e, cel_a, cel_q, cel_y2, cel_t, cel_g;
    std::vector<int> eidx_xref;
};

// Builds the tables from the descriptor; returns MC_OK or an MC_E* code with
// a message in err.
int prepare(const mc_atomic *in, int nbins, const float *nuArray, Prepared &P, char *err,
            size_t errlen);

}  // namespace mcb

hipError_t mcb_launch_balance(const mcb::Atomic *dA, const mcb::CellArgs &a, hipStream_t s);
// channel integrals Rint/Hint [nCells+1][nch], hit flags [nCells+1] and
// (X non-null) the warm-start network's inputs X [nCells][72] of every
// active cell, from the tiled J3 (lattice nx x ny x nz) times a.jscale
hipError_t mcb_launch_integrals(const mcb::Atomic *dA, const mcb::CellArgs &a, const double *J3,
                                const int *active, int nx, int ny, int nz, double *Rint, double *Hint,
                                int *hit, float *X, hipStream_t s);
// Te, Ne, ion of reached cells from the network's outputs Y [nCells][ldy]
hipError_t mcb_launch_warm(const mcb::Atomic *dA, const mcb::CellArgs &a, const float *Y, int ldy,
                           const int *hit, hipStream_t s);
// line luminosities of every unblackened cell: out[ncel * 289 + 1] (model
// ion, upper level, lower level; then Hbeta), summed over cells in a fixed
// order; a.part holds nblk rows of a.ldp doubles
hipError_t mcb_launch_lines(const mcb::Atomic *dA, const mcb::LineArgs &a, int nblk, int ncel, double *out,
                            hipStream_t s);

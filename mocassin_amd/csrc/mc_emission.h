// mc_emission.h - the re-emission tables of every active cell on the device:
// emissionDriver's gas part (emission_mod.f90:32-173 with fb_ff, twoPhoton,
// free_free, RecLinesEmission, forLines) and setDiffusePDF (905-1311), in
// the reference's mixed REAL/DOUBLE precision, written into the transport's
// recPDF rows and totalLines.
#pragma once
#include <hip/hip_runtime.h>

#include "mc_balance.h"

namespace mce {

constexpr int kMaxLim = 64;      // entries of a log-gamma table row
constexpr int kMaxLy = 8;        // He II Lyman lines
constexpr int kMaxRec = 512;     // recombination lines summed into normRec
constexpr int kMaxDustRows = 32; // grain (species, size) rows

struct EmArgs {
    int nb, nc;
    // state: Te, Ne, Hden [nCells+1], ion [nCells+1][nE][nstages+1] (FP64,
    // rounded to FP32 as the host grid holds them)
    const double *Te, *Ne, *Hden, *ion;
    int nE, ns1, eH, eHe;                    // element indices of H and He
    float abH, abHe;
    const float *nu, *widFlx;
    // log gamma tables of H I, He I, He II: (ntk, nlim) column-major, their
    // edge bins, and per bin the last edge pair covering it (-1 none)
    int ntk;
    const float *tk;
    int nlim[3];
    const float *lg[3];
    const int *edge[3];
    const int *pair[3];
    int ip[3], i0[3];                        // ionisation bins, Milne starts
    const float *sig[3];                     // ground-state cross-sections from i0
    const float *gff, *gffHeII;
    // heavy-ion free-bound channels: [nfb][4] element index, stage of the
    // recombining ion, IPnuP, highNuP; [nfb] sigma offset; [nfb][2] g0/g1,
    // elemAbun
    int nfb;
    const int *fb_i, *fb_off;
    const float *fb_f, *fb_sig;
    const int *jHeI;                         // [nb] He I two-photon segment, 0 beyond the table
    const float *ydat, *Aydat;
    int j2TsP, nly;
    int lyP[kMaxLy];
    float lyData[kMaxLy];
    const double *HIrec, *HeIIrec;           // (3:30,2:8), (3:30,2:16)
    const float *HeIc;                       // (34,3,4)
    // grains (nullable): rows k = (species-1)*nSizes + size-1
    int dS, dSizes, dnsMax;
    const float *dabs;                       // [rows][nb]
    const float *dweight, *dgabun;           // grainWeight, grainAbun
    const float *dtsub;                      // [rows] TdustSublime of the row's species
    const float *ndust;                      // (0:nc)
    const double *Tdust;                     // (0:nsMax, 0:nSizes, 0:nc)
    // outputs
    float *rec;                              // [nCells+1][nb]
    float *totalLines;                       // [nCells+1]
    int *err;                                // cells whose recPDF has no positive total
};

}  // namespace mce

// One wavefront per active cell; lds = dynamic LDS bytes (mce_lds_bytes).
size_t mce_lds_bytes(int nb, bool dust);
hipError_t mce_launch(const mcb::Atomic *dA, const mce::EmArgs &a, hipStream_t s);

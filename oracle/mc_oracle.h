/*
 * mc_oracle.h - CPU restatement of MOCASSIN's photon-packet transport
 * (TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked into the product).
 *
 * Arrays use the reference's own Fortran layouts so that the restatement
 * reads like the reference:
 *   active(1:nx,1:ny,1:nz)           active[(i-1) + nx*((j-1) + ny*(k-1))]
 *   table(0:nCells,1:nbins)          tab[c + (nCells+1)*(nu-1)]
 *   escapedPackets(0:nCells,0:nbins,0:nAngleBins)
 *                                    esc[c + (nCells+1)*(nu + (nbins+1)*a)]
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MCO_MATH_LIBM = 0, MCO_MATH_OWN = 1 };
enum { MCO_J_FP32 = 0, MCO_J_FP64 = 1 };

typedef struct {
    int nx, ny, nz;
    const float *xAxis, *yAxis, *zAxis;
    const int32_t *active;
    int nCells;
    int nbins;
    const float *opacity;      /* (0:nCells,1:nbins) */
    const float *scaOpac;      /* (0:nCells,1:nbins) or NULL when no dust */
    const float *recPDF;       /* (0:nCells,1:nbins) */
    const float *totalLines;   /* (0:nCells) */
    const float *specCDF;      /* inSpectrumProbDen(iStar,1:nbins) */
    const float *nuArray;      /* (1:nbins) */
    const float *gSca;         /* (1:nbins) or NULL */
    float R_out;
    float geoCorrX, geoCorrY, geoCorrZ;
    float ionEdge1;
    float deltaE;
    float starPos[3];
    int starIdx[3];            /* 1-based cell indices of the source */
    int lgSymmetricXYZ, lgDust, lgGas, lgIsotropic, lgMultiStars;
    int math_mode;             /* MCO_MATH_* */
    int j_mode;                /* MCO_J_* */
    uint64_t seed;             /* Philox key */
    uint32_t stream;           /* Philox counter word 3 (source index) */
    /* viewing-angle bins of the escapes as initCartesianGrid sets them
     * (grid_mod.f90:416-467); nAngleBins = 0 leaves esc with plane 0 only */
    int nAngleBins, totAngleBinsTheta, totAngleBinsPhi;
    float dTheta, dPhi;
    const int32_t *viewPointPtheta;      /* (0:totAngleBinsTheta) */
    const int32_t *viewPointPphi;        /* (0:totAngleBinsPhi) */
    const float *viewPointTheta;         /* (0:nAngleBins) */
    const float *viewPointPhi;           /* (0:nAngleBins) */
} mco_problem;

typedef struct {
    double n_seg, n_abs, n_sca, n_esc, n_lost, n_line, n_trapped, n_flights;
    double qphot;
    int error;                 /* first fatal condition met ("stop" in the reference) */
} mco_counters;

/* Philox4x32-10 block, exposed for the known-answer test. */
void mco_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* k-th uniform of packet p: [0,1) with 24-bit resolution. */
float mco_uniform(uint64_t seed, uint32_t stream, uint64_t packet, uint32_t k);

/* getNu2 restatement (photon_mod.f90:720-764) on an explicit uniform
 * stream; returns nuP (1-based) and the number of uniforms consumed. */
int mco_getnu2(const float *cdf, int nbins, const float *u, int nu_avail, int *used);

/* Elementary functions in the selected mode. */
float mco_logf(int mode, float x);
void mco_sincosf(int mode, float a, float *s, float *c);
float mco_acosf(int mode, float x);
float mco_atanf(int mode, float x);

/* Geometry-only walk (pathSegment with zero opacity) from pos/dir in the
 * given start cell.  Writes up to max_len segments (cell x,y,z 1-based and
 * dS).  Returns the number of segments; *status: 1 escaped, 2 lost on the
 * upper wall snap, 3 safe limit, <0 error. */
int mco_walk(const mco_problem *P, const float pos[3], const float dir[3],
             const int cell[3], int32_t *cells_xyz, float *ds, int max_len,
             int *status);

/* Full transport of packets [first, first+n) of one source.
 * J and esc are FP32 (j_mode == MCO_J_FP32, summed in packet order exactly
 * as the reference does) or FP64; both (0:nCells,*) Fortran layouts. */
int mco_transport(const mco_problem *P, uint64_t first, uint64_t n,
                  void *J, void *esc, mco_counters *cnt);

/* ------------------------------------------------------------- balance */
#define MCO_MAX_CEL 64

/* Atomic inputs of updateCell as the reference holds them after setup
 * (arrays dumped by the reference harness; see tests/golden). */
typedef struct {
    int nbins, nstages, nElemUsed;
    const float *nuArray;
    int32_t elem_on_in[30], xref_in[30];   /* lgElementOn, elementXref (1..30) */
    float abun_in[30];                     /* elemAbun(1,1:30) */
    int nch;                               /* photoionisation channels */
    const int32_t *ch_meta;                /* [nch][8] elem ion g0 g1 IPnuP hiRate hiHeat outShell */
    const float *ch_sigma;                 /* [nch][nbins] cross sections on nuArray */
    const float *rrec;                     /* rrec(1:2, z, n)  as [z-1][n-1][2] */
    const float *rnew;                     /* rnew(1:4, z, n)  as [z-1][n-1][4] */
    const float *fe;                       /* fe(1:3, 4:13)    as [n-4][3] */
    const int32_t *dr_int; const float *dr_real;   /* direc_coeffs(25) */
    const int32_t *ap_int; const float *ap_real;   /* aldropequi_coeffs(167) */
    int ncel;                              /* CEL model ions in forLines order */
    const int32_t *cel_meta;               /* [ncel][6] elem ion ii jj nlev ntemp */
    const double *cel_dbl;                 /* per ion: e(nl), a(nl,nl), qom(nt,nl,nl) */
    const float *cel_flt;                  /* per ion: logtemp(nt), g(nl), br, z, a_r(4), a_d(5) */
    float XHILimit;
    /* filled by mco_atomic_init */
    int elemOn[31], xref[31];
    float abun[31];
    int32_t chan_g[31 * 10 * 2];
    int32_t chan_ip[31 * 10];              /* IPnuP of channel (elem, ion) */
    size_t cel_dbl_off[MCO_MAX_CEL], cel_flt_off[MCO_MAX_CEL];
} mco_atomic;

/* Grains of one dust chemistry for the grain charge, photoelectric heating
 * and gas-grain collision terms of updateCell (update_mod.f90:527-932). */
#define MCO_MAX_GROWS 32
typedef struct {
    int nS, nSizes;
    const float *absx;      /* [nS][nSizes][nbins]: xSecArray(dustAbsXsecP(nS,ai)+i-1) */
    const float *weight;    /* grainWeight(1:nSizes) */
    const float *gabun;     /* grainAbun(1,1:nS) */
    const float *vn;        /* grainVn(1:nS) [Ryd] */
    const float *msurf;     /* MsurfAtom(1:nS) */
    const float *radius;    /* grainRadius(1:nSizes) [um] */
    const float *aweight;   /* aWeight(1:30) */
} mco_grains;

int mco_atomic_init(mco_atomic *A);
int mco_balance_cell_f32(const mco_atomic *A, const float *J, float Hden, float *Te, float *Ne,
                         float *ionDen, int *nT);
int mco_balance_cell_f64(const mco_atomic *A, const double *J, double Hden, double *Te, double *Ne,
                         double *ionDen, int *nT);
/* The same with the grain terms: ndust = Ndust(cell), Tdust = Tdust(0:nsMax,
 * 0:nSizes, cell) before the update; out: gdc/peh[(nS-1)*nSizes + ai-1] =
 * gasDustColl_d, photoelHeat_d and pot = grainPot of the last T-iteration. */
int mco_balance_cell_pe_f32(const mco_atomic *A, const mco_grains *G, int nsMax, const float *J,
                            float Hden, float *Te, float *Ne, float *ionDen, int *nT, float ndust,
                            const float *Tdust, float *gdc, float *peh, float *pot);
int mco_balance_cell_pe_f64(const mco_atomic *A, const mco_grains *G, int nsMax, const double *J,
                            double Hden, double *Te, double *Ne, double *ionDen, int *nT, double ndust,
                            const double *Tdust, double *gdc, double *peh, double *pot);

/* The module data emissionDriver reads for the gas re-emission of a cell
 * (set up by the reference's ph_mod, hydro_mod and readData): the log gamma
 * tables of H I, He I and He II (ntk, nlim) column-major with their edge
 * bins, the free-free Gaunt factors, the ground-state cross-sections from
 * the Milne start bins (sHI from ipHI, sHeI from i0HeI, sHeII from i0HeII),
 * the heavy-ion free-bound channels of fb_ff ([nfb][4] elem, ion, IPnuP,
 * highNuP; g0/g1; cross-sections over IPnuP..highNuP-1, channels
 * concatenated), the He I two-photon table, the recombination-line data and
 * the bins of the He I 2s triplet and the He II Lyman lines. */
typedef struct {
    int nb, ntk, nHI, nHeI, nHeII, ipHI, ipHeI, ipHeII, i0HeI, i0HeII, nfb, nly, j2TsP;
    const float *nu, *widFlx, *tk, *gHI, *gHeI, *gHeII;
    const int32_t *eHI, *eHeI, *eHeII;
    const float *gff, *gffHeII, *sHI, *sHeI, *sHeII;
    const int32_t *fbm;
    const float *fbg, *fbs;
    const float *ydat, *Aydat, *lyData;
    const int32_t *lyP;
    const float *HeIc;                   /* HeIrecLineCoeff(34,3,4) */
    const double *HIrec, *HeIIrec;       /* HIRecLineData(3:30,2:8), HeIIRecLineData(3:30,2:16) */
} mco_emission;

/* The grains of one cell for setDiffusePDF's dust emission (one chemistry):
 * absx [nS][nSizes][nbins], grainWeight, grainAbun, TdustSublime per
 * species, the cell's Tdust(0:nsMax, 0:nSizes) block, Ndust and Hden. */
typedef struct {
    int nS, nSizes, nsMax;
    const float *absx, *weight, *gabun, *tsub, *Tdust;
    float Ndust, Hden;
} mco_emission_dust;

/* emissionDriver's gas part for one cell: its recPDF row rec[nb] and
 * totalLines (D: the cell's grains, or NULL); ion[(stage-1)*nElemUsed + xref-1].  _f32: the reference's
 * precision throughout; _f64: the level populations in FP64 (the GPU's). */
int mco_emission_cell_f32(const mco_atomic *A, const mco_emission *E, const mco_emission_dust *D, float Te,
                          float Ne, const float *ion, float *rec, float *totalLines);
int mco_emission_cell_f64(const mco_atomic *A, const mco_emission *E, const mco_emission_dust *D, float Te,
                          float Ne, const float *ion, float *rec, float *totalLines);

/* outputGas's analytic forbidden-line luminosities forbVol of every model ion
 * and HbetaVol (output_mod.f90:302-520) of n cells, accumulated into forb
 * [ncel][17][17] (model ion, upper level, lower level) and *hbeta. */
void mco_line_lum_f32(const mco_atomic *A, int n, const float *Hden, const float *Te, const float *Ne,
                      const float *ionDen, const float *dV, float *forb, float *hbeta);
void mco_line_lum_f64(const mco_atomic *A, int n, const double *Hden, const double *Te, const double *Ne,
                      const double *ionDen, const double *dV, double *forb, double *hbeta);

#ifdef __cplusplus
}
#endif
#endif

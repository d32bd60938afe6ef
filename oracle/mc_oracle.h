/*
 * mc_oracle.h - CPU restatement of MOCASSIN's photon-packet transport
 * (TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked into the product).
 *
 * Arrays use the reference's own Fortran layouts so that the restatement
 * reads like the reference:
 *   active(1:nx,1:ny,1:nz)           active[(i-1) + nx*((j-1) + ny*(k-1))]
 *   table(0:nCells,1:nbins)          tab[c + (nCells+1)*(nu-1)]
 *   escapedPackets(0:nCells,0:nbins) esc[c + (nCells+1)*nu]     (angle bin 0)
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

#ifdef __cplusplus
}
#endif
#endif

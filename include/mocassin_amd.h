/*
 * mocassin_amd.h - C ABI of the MI355X photon-packet transport and per-cell
 * balance library (libmocassin_amd.so).
 *
 * Drop-in boundary for MOCASSIN 2.02's transport/balance step.  The Fortran
 * host keeps every array it owns in its own layout; the library owns device
 * buffers, copies explicitly, and retains no host pointer across calls.
 * Layout conventions (all Fortran column-major, 1-based on the host side):
 *   active(1:nx,1:ny,1:nz)                      int32
 *   opacity/scaOpac/recPDF(0:nCells,1:nbins)    float,  cell index fastest
 *   totalLines(0:nCells)                        float
 *   specCDF(1:nbins, 1:nStars)                  float  (inSpectrumProbDen rows)
 *   J(0:nCells,1:nbins)                         double (Jste, unscaled)
 *   esc(0:nCells,0:nbins[,0:nAngleBins])        double (escapedPackets; planes
 *                                                after mc_set_view_angles)
 * Row 0 of every (0:nCells,...) array is the inactive-cell sink, as in the
 * reference.
 *
 * Every function returns 0 on success or a positive MC_E* code; the message
 * of the last failure on a context is available from mc_ctx_last_error().
 * The library never calls exit(): the host maps a non-zero status to its own
 * fail-stop ("stop" in the reference).
 */
#ifndef MOCASSIN_AMD_H
#define MOCASSIN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MC_ABI_VERSION 3

/* flags for mc_desc.flags */
#define MC_SYMMETRIC   0x01  /* lgSymmetricXYZ  (set_input_mod.f90:333) */
#define MC_DUST        0x02  /* lgDust */
#define MC_GAS         0x04  /* lgGas */
#define MC_ISOTROPIC   0x08  /* lgIsotropic: isotropic dust scattering */
#define MC_MULTISTARS  0x10  /* lgMultiStars: no octant folding of stellar directions */

/* error codes */
#define MC_OK          0
#define MC_EINVAL      1   /* bad argument or descriptor */
#define MC_ENODEV      2   /* no usable GPU / HIP error */
#define MC_ENOMEM      3   /* device allocation failed */
#define MC_ESTATE      4   /* call out of order (e.g. run before tables) */
#define MC_ETABLE      5   /* table rejected (NaN, non-monotone CDF, ...) */
#define MC_EPACKET     6   /* a packet met a fail-stop condition of the reference */
#define MC_EUNSUP      7   /* configuration outside the supported path */

typedef struct mc_ctx mc_ctx;

/* Grid and run description.  Replaces the module globals that
 * energyPacketDriver reads (photon_mod.f90:26; common_mod.f90:241-302). */
typedef struct {
    int32_t abi_version;        /* MC_ABI_VERSION */
    int32_t nx, ny, nz;         /* grid(1)%nx, ny, nz */
    const float *xAxis;         /* grid(1)%xAxis(1:nx) [cm] */
    const float *yAxis;
    const float *zAxis;
    const int32_t *active;      /* grid(1)%active, Fortran order */
    int32_t nCells;             /* grid(1)%nCells */
    int32_t nbins;              /* nbins */
    int32_t flags;              /* MC_* flags */
    float R_out;                /* R_out [cm] */
    float geoCorr[3];           /* grid(1)%geoCorrX/Y/Z */
    float ionEdge1;             /* ionEdge(1) [Ryd] */
    const float *nuArray;       /* nuArray(1:nbins) [Ryd] */
    int32_t nStars;             /* nStars */
    const float *starPos;       /* starPosition(1:nStars) as (3,nStars) [cm] */
    const int32_t *starIdx;     /* starIndeces(1:nStars,1:3) as (3,nStars), 1-based */
    const float *deltaE;        /* deltaE(1:nStars) [1e36 erg/s], finite and > 0 */
    uint64_t seed;              /* key of the counter-based random streams */
    int32_t device;             /* HIP device ordinal */
} mc_desc;

/* Counters of one transport call (summed over its packets). */
typedef struct {
    double n_seg;       /* cell-crossing or partial segments walked */
    double n_abs;       /* absorptions (absInt) */
    double n_sca;       /* dust scatterings (scaInt) */
    double n_esc;       /* escapes recorded in esc */
    double n_lost;      /* packets dropped by the reference's wall-snap return or safeLimit */
    double n_line;      /* diffuse packets turned into (escaping) line packets */
    double n_trapped;   /* packets that hit recursionLimit */
    double n_flights;   /* pathSegment calls */
    double qphot;       /* Qphot: sum deltaE/(h nu) over stellar packets with nu > 1 Ryd */
    double kernel_ms;   /* device time of the transport kernel */
    /* the scheduling of the launch (no counterpart in the reference):
     * wavefront walk passes and the lanes that walked in them, wavefront
     * emission passes and the lanes that emitted in them; lanes per pass =
     * lanes / passes */
    double n_walk_passes, n_walk_lanes, n_emit_passes, n_emit_lanes;
} mc_counters;

/* Launch shape of the transport kernel (chosen by the autotuner,
 * mocassin_amd/autotune.py; no counterpart in the reference):
 *   blocks_per_cu  resident workgroups per compute unit requested (1..32),
 *                  counted in 256-thread units (wavefronts per SIMD)
 *   threads        workgroup size, a multiple of 64 (64..256)
 *   emit_min       an emission pass runs when at least this many lanes of a
 *                  wavefront wait for one (1..64), or half the live lanes do
 *                  (the pool kernels: when at least this many lanes do not
 *                  walk, or none does)
 *   walk_steps     walk passes between two emission checks (1..64)
 *   kernel         MC_KERNEL_AUTO: two packets per lane when blocks_per_cu
 *                  is at most 3, else the pool kernel where the grid allows
 *                  it, else one packet per lane;
 *                  MC_KERNEL_ONE: one packet per lane;
 *                  MC_KERNEL_SPARE: two packets per lane (a lane whose
 *                  flight ends walks its spare while the next emission pass
 *                  waits; 3 wavefronts per SIMD at most);
 *                  MC_KERNEL_POOL: one packet per lane and a pool of 32
 *                  packets per wavefront in LDS (a lane whose flight ends
 *                  takes a walking packet from the pool; emission passes run
 *                  on nearly full wavefronts); needs a grid of at most
 *                  2047 x 2047 x 1023 cells and 65535 bins (MC_EUNSUP).
 * Results do not depend on these values (each packet has its own stream).
 * A new context starts with 4, 256, 16, 16, MC_KERNEL_AUTO. */
#define MC_KERNEL_AUTO 0
#define MC_KERNEL_ONE 1
#define MC_KERNEL_SPARE 2
#define MC_KERNEL_POOL 3
typedef struct {
    int32_t blocks_per_cu;
    int32_t threads;
    int32_t emit_min;
    int32_t walk_steps;
    int32_t kernel;
} mc_tuning;

/* Atomic inputs of the per-cell balance, as the host holds them after its
 * setup (readData, makeElements, initXSecArray; the tables are read by
 * updateCell and its callees, update_mod.f90:15-2027, emission_mod.f90:
 * 1913-2190).  All arrays are plain C order as stated:
 *   elemOn/xref/abun[30]     lgElementOn, elementXref, elemAbun(1,:) by Z-1
 *   ch_meta[nch][8]          elem, ion, g0, g1, IPnuP, hiRate, hiHeat, outShell
 *                            of each photoionisation channel: statistical
 *                            weights of the ground states, first bin above the
 *                            edge and the last bins of the rate and heating
 *                            integrals (1-based), outer shell
 *   ch_sigma[nch][nbins]     its cross section on nuArray [cm^2]
 *   rrec[30][30][2]          rrec(1:2, z, n) at [z-1][n-1][i-1]
 *   rnew[30][30][4]          rnew(1:4, z, n) at [z-1][n-1][i-1]
 *   fe[10][3]                Fe fits of radRecFit for n = 4..13
 *   dr_int[25][3]/dr_real[25][5]    direc_coeffs: elem, n, Te-range flag / a..f
 *   ap_int[167][2]/ap_real[167][4]  aldropequi_coeffs: elem, n / A, B, T0, T1
 *   cel_meta[ncel][6]        elem, ion, ii, jj, nlev, ntemp of each CEL
 *                            model ion in forLines order
 *   cel_dbl                  per model ion: e(nl) [cm^-1], a(nl,nl) [s^-1]
 *                            column-major, qom(nt,nl,nl) column-major
 *   cel_flt                  per model ion: logtemp(nt), g(nl), br, z,
 *                            a_r(4), a_d(5) (level recombination fit) */
typedef struct {
    int32_t nstages;            /* nstages (<= 9) */
    int32_t nElemUsed;          /* nElementsUsed */
    const int32_t *elemOn;
    const int32_t *xref;
    const float *abun;
    int32_t nch;
    const int32_t *ch_meta;
    const float *ch_sigma;
    const float *rrec, *rnew, *fe;
    const int32_t *dr_int;
    const float *dr_real;
    const int32_t *ap_int;
    const float *ap_real;
    int32_t ncel;
    const int32_t *cel_meta;
    const double *cel_dbl;
    const float *cel_flt;
    float XHILimit;             /* XHILimit (convergence of the H0 fraction) */
} mc_atomic;

/* Create a context on desc->device and upload the grid.
 * Replaces: grid setup consumed by energyPacketDriver (photon_mod.f90:26). */
int mc_ctx_create(const mc_desc *desc, mc_ctx **out);

/* Upload the per-sweep tables (frozen during a transport sweep).
 * scaOpac/gSca may be NULL without dust; recPDF/totalLines NULL without gas.
 * opacity (and scaOpac) NULL keep the device's current tables: those
 * mc_opacity_build wrote, or an earlier upload's.
 * Replaces: the reads of grid%opacity, grid%scaOpac, grid%recPDF,
 * grid%totalLines, inSpectrumProbDen, gSca inside photon_mod.f90:768-1812;
 * the tables are produced by ionizationDriver (ionization_mod.f90:349-484),
 * emissionDriver (emission_mod.f90:905-1311) and setProbDen
 * (continuum_mod.f90:418-474). */
int mc_upload_tables(mc_ctx *ctx, const float *opacity, const float *scaOpac,
                     const float *recPDF, const float *totalLines,
                     const float *specCDF, const float *gSca);

/* Zero the device J and escape accumulators.
 * Replaces: iteration_mod.f90:458-472. */
int mc_zero_estimators(mc_ctx *ctx);

/* Transport npackets stellar packets of source istar (1-based), packet
 * indices packet_offset .. packet_offset+npackets-1 (each index selects its
 * own random stream, so disjoint ranges are independent and results do not
 * depend on launch shape).  Accumulates into the device estimators; if J_out
 * / esc_out are non-NULL the running totals are copied back to them
 * (layouts above).  c may be NULL.
 * Replaces: energyPacketDriver(iStar, n, grid) (photon_mod.f90:26-284) with
 * energyPacketRun/newPhotonPacket/pathSegment (289-2872). */
int mc_transport_run(mc_ctx *ctx, int32_t istar, int64_t npackets, uint64_t packet_offset,
                     double *J_out, double *esc_out, mc_counters *c);

/* Count the path segments walked in each cell during transport (1) or stop
 * counting (0); counts are zeroed with the estimators.  Off by default: it
 * costs one extra atomic per segment.  No counterpart in the reference; it
 * gives the per-cell packet crossings that decide which cells are sampled
 * well enough to judge convergence. */
int mc_count_crossings(mc_ctx *ctx, int32_t enable);
/* crossings(0:nCells): the counts per cell (row 0: all inactive points). */
int mc_fetch_crossings(mc_ctx *ctx, uint32_t *crossings);

/* Launch shape of a batched transport (chosen by the autotuner; no
 * counterpart in the reference):
 *   waves_per_cu  resident one-wavefront workgroups requested per compute
 *                 unit (1..32)
 *   emit_min, walk_steps  as in mc_tuning
 *   chunk         packets handed to a wavefront at a time (64..2^24)
 * Results do not depend on these values. */
typedef struct {
    int32_t waves_per_cu;
    int32_t emit_min;
    int32_t walk_steps;
    int64_t chunk;
} mc_batch_tuning;

/* Transport the packets of n independent models (contexts on one device)
 * in one persistent launch: for each i, npackets[i] packets of source
 * istar[i] of ctx[i], packet indices packet_offset[i] ... (NULL: 0).
 * Each context accumulates into its own estimators and gets its own
 * counters c[i] (nullable; kernel_ms is the whole launch); the results equal
 * those of n separate mc_transport_run calls.  tuning may be NULL.
 * Replaces: n independent runs of energyPacketDriver (photon_mod.f90:26-284),
 * one per nebula. */
int mc_transport_batch(mc_ctx *const *ctx, int32_t n, const int32_t *istar, const int64_t *npackets,
                       const uint64_t *packet_offset, const mc_batch_tuning *tuning, mc_counters *c,
                       double *kernel_ms);

/* mc_transport_batch without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (serving loops,
 * code under graph capture).  Same arguments and results.  The launch's
 * working block (the models' grids, run arguments and counters, the chunk
 * table) lives in a fixed 8 MiB device array that each GPU the library runs
 * on holds until the process ends; a launch that needs more (at most 256
 * models; 16 bytes per chunk of `chunk` packets) is refused with MC_EINVAL
 * before anything runs.  Every call on a GPU reuses that array: calls that
 * overlap on one GPU race. */
int mc_transport_batch_noalloc(mc_ctx *const *ctx, int32_t n, const int32_t *istar, const int64_t *npackets,
                               const uint64_t *packet_offset, const mc_batch_tuning *tuning, mc_counters *c,
                               double *kernel_ms);

/* Copy the device estimators to the host (layouts above). */
int mc_fetch_estimators(mc_ctx *ctx, double *J_out, double *esc_out);

/* Geometry-only replay of pathSegment (photon_mod.f90:1194-2836) with zero
 * opacity: for each of n packets starting at pos(3,n) with direction
 * dir(3,n) in cell(3,n) (1-based), writes the visited cells
 * cell_seq(3,max_len,n) (1-based x,y,z), the FP32 path length of each
 * segment ds_seq(max_len,n), seq_len(n) and status(n) (1 escaped, 2 dropped
 * by the upper-wall snap return, 3 safe limit, <0 error). */
int mc_walk_replay(mc_ctx *ctx, int32_t n, const float *pos, const float *dir,
                   const int32_t *cell, int32_t *cell_seq, float *ds_seq,
                   int32_t *seq_len, int32_t *status, int32_t max_len);

/* Overwrite the device estimators with host values (layouts above; esc may
 * be NULL).  Row 0 of J is not kept: inactive positions read zero.  For a
 * host that runs the balance on estimators it obtained elsewhere (a restart,
 * another transport).  Replaces: the reads of grid%Jste and
 * grid%escapedPackets by iterateMC and updateCell (iteration_mod.f90:700-722,
 * update_mod.f90:104-154). */
int mc_set_estimators(mc_ctx *ctx, const double *J, const double *esc);

/* Viewing-angle bins of the escapes, as initCartesianGrid sets them from
 * the "inclination" keyword (grid_mod.f90:416-467, set_input_mod.f90:534-556):
 * theta and phi bin counts and widths, the maps from a bin to a viewpoint
 * (0 = none), and the viewpoints' angles in radians (index 0 unused). */
typedef struct {
    int32_t nAngleBins;                 /* 0 (none), 1 or 2 */
    int32_t totAngleBinsTheta, totAngleBinsPhi;   /* 1..64 each */
    float dTheta, dPhi;                 /* Pi/totAngleBinsTheta, twoPi/totAngleBinsPhi (REAL) */
    const int32_t *viewPointPtheta;     /* (0:totAngleBinsTheta), values 0..nAngleBins */
    const int32_t *viewPointPphi;       /* (0:totAngleBinsPhi), values 0..nAngleBins */
    const float *viewPointTheta;        /* (0:nAngleBins) */
    const float *viewPointPhi;          /* (0:nAngleBins), < 0: phi switched off */
} mc_view_angles;

/* Switch the escapes' viewing-angle planes on (v->nAngleBins > 0) or off
 * (v NULL or nAngleBins 0).  esc then holds escapedPackets(0:nCells,
 * 0:nbins, 0:nAngleBins) in mc_transport_run, mc_fetch_estimators and
 * mc_set_estimators, each escape binned by its flight direction
 * (enPacket%direction) as every escape site does (photon_mod.f90:370-465,
 * 1582-1666 and the four alike in pathSegment); the escapes are zeroed.
 * A bin out of range stops the sweep with MC_EPACKET (codes 60, 61), where
 * the reference stops.  Replaces: those sites' viewing-angle branches. */
int mc_set_view_angles(mc_ctx *ctx, const mc_view_angles *v);

/* Planes of esc the context returns and takes: nAngleBins + 1 after
 * mc_set_view_angles, else 1 (0 for a NULL context).  Lets a host check its
 * escapedPackets array before a copy.  No counterpart in the reference. */
int mc_esc_planes(const mc_ctx *ctx);

/* nCells and nbins of the context (both 0 for a NULL context; returns
 * MC_EINVAL then), so that a host can check the extents of the J
 * (0:nCells, 1:nbins) and escapedPackets (0:nCells, 0:nbins, planes) arrays
 * it passes.  No counterpart in the reference, whose arrays share the
 * module globals' extents. */
int mc_ctx_shape(const mc_ctx *ctx, int32_t *nCells, int32_t *nbins);

/* Upload the atomic inputs of the balance (validated; see mc_atomic).
 * Requires H and He switched on, nstages <= 9, nbins <= 2048, at most 16
 * elements, 64 channels, 64 CEL model ions of <= 17 levels and <= 64
 * temperatures.  Replaces: the module data updateCell reads
 * (update_mod.f90:15-2027). */
int mc_balance_setup(mc_ctx *ctx, const mc_atomic *atomic);

/* Temperature solver and warm start of mc_balance_run (no counterpart in
 * the reference; NULL options mean the reference's own schedule). */
#define MC_SOLVER_REFERENCE 0  /* iterateT: false position, |heat-cool| < 0.01 heat, 25 iterations */
#define MC_SOLVER_SECANT    1  /* secant steps kept inside the bracket, |heat-cool| <= tol heat */
typedef struct {
    int32_t solver;             /* MC_SOLVER_* */
    int32_t max_iter;           /* secant: temperature iterations allowed */
    double tol;                 /* secant: relative thermal residual accepted */
    double ion_tol;             /* ion fixed point: relative change of X(H0), X(He0), X(He+)
                                   that ends it (the reference uses 0.01) */
    int32_t warm_start;         /* 1: reached cells start from the network's prediction
                                   (mc_mlp_set first); a cell whose secant solve from there
                                   fails, or has not converged after warm_iter iterations,
                                   restarts from the state passed in (flags bit5) */
    double first_step;          /* secant: relative size of the first trial step when the
                                   start gives no bracket (0: 0.05, or 0.01 with
                                   warm_start: a good warm start wants a step of the
                                   order of its error) */
    double first_slope;         /* secant: that step is the relative residual divided by
                                   this expected |d((heat-cool)/heat)/d ln Te| (0: 1) */
    int32_t warm_iter;          /* warm_start: temperature iterations of the attempt from the
                                   prediction before the restart (0: max_iter).  A small cap
                                   keeps one slow cell from holding the launch's tail where
                                   predictions are good; where cells need many iterations
                                   it wastes the attempt */
} mc_balance_opts;

/* Ionisation and thermal balance of every active cell on the device (FP64),
 * from the device J times jscale (the host's normalisation of Jste, e.g.
 * scale_J).  Hden(0:nCells), Te(0:nCells), Ne(0:nCells) and
 * ionDen(0:nCells, 1:nElemUsed, 1:nstages) (column-major, cell fastest) are
 * copied in and, with flags(0:nCells) and nIter(0:nCells), back out; cells
 * that no packet reached keep their inputs.  flags per cell: bit0 H0
 * fraction converged (lgConverged), bit1 lgBlack, bit2 thermal balance not
 * reached, bit3 ion balance not converged, bit4 not hit, bit5 restarted
 * from the given state after a failed warm-started solve.  nIter:
 * temperature iterations used.  kernel_ms (nullable): device time of the
 * balance (and of the network's inference when warm-starting).
 * opts NULL = MC_SOLVER_REFERENCE, ion_tol 0.01, no warm start.  Refused
 * (MC_ESTATE) when grains with the photoelectric effect are set up: that
 * balance needs Tdust, through mc_update_cells.
 * Replaces: updateCell (update_mod.f90:15-525) with iterateT, ionBalance,
 * thermBalance, calcAlpha, radRecFit, dielectronic and forLines/equilibrium
 * (update_mod.f90:339-1834, emission_mod.f90:1913-2190), called per cell by
 * iterateMC's loop over the grid (iteration_mod.f90:826-838). */
int mc_balance_run(mc_ctx *ctx, double jscale, const mc_balance_opts *opts, const double *Hden,
                   double *Te, double *Ne, double *ionDen, int32_t *flags, int32_t *nIter,
                   double *kernel_ms);

/* ---- grains ---------------------------------------------------------------
 * One dust chemistry: the tables getDustT reads, as the host holds them after
 * dustDriver (dust_mod.f90:14-185) and makeDustXsec (ph_mod.f90:808-1544). */
typedef struct {
    int32_t nSpecies;           /* grain species (nSpeciesPart(1)); nSpecies*nSizes <= 32 */
    int32_t nSpeciesMax;        /* Tdust's first dimension is 0:nSpeciesMax (>= nSpecies) */
    int32_t nSizes;             /* grain sizes */
    int32_t nTemps;             /* emission integrals at 1..nTemps K (<= 8192) */
    int32_t photoelectric;      /* lgPhotoelectric: with gas, the grain charge, photoelectric
                                   heating and gas-grain collisions (update_mod.f90:527-932)
                                   enter the gas balance and the grain temperatures; needs
                                   nbins <= 4096 and the five arrays below */
    const float *absXsec;       /* [nSpecies][nSizes][nbins]: xSecArray(dustAbsXsecP(nS,ai)+i-1) */
    const float *grainWeight;   /* (1:nSizes) */
    const float *grainAbun;     /* (1:nSpecies): grainAbun(1,nS) */
    const float *emIntegral;    /* dustEmIntegral(1:nSpecies, 1:nSizes, 1:nTemps), column-major */
    const float *grainRadius;   /* (1:nSizes) [um]                        photoelectric only */
    const float *grainVn;       /* (1:nSpecies) [Ryd]                      photoelectric only */
    const float *MsurfAtom;     /* (1:nSpecies), as the species file gives it  photoelectric only */
    const float *aWeight;       /* (1:30) atomic weights                   photoelectric only */
    const float *Ndust;         /* (0:nCells) grain number density   photoelectric, opacity */
    const float *scaXsec;       /* [nSpecies][nSizes][nbins]: xSecArray(dustScaXsecP(nS,ai)+i-1);
                                   with TdustSublime and Ndust, the dust add of
                                   mc_opacity_build (NULL: no dust opacity on the device) */
    const float *TdustSublime;  /* (1:nSpecies) sublimation temperatures [K] */
    float XHILimit;             /* grains without gas: |dTdust(0,0)|/Tdust(0,0) <= XHILimit
                                   marks a cell converged (update_mod.f90:316-324) */
} mc_dust_tables;

/* Upload the grain tables (validated).  Replaces: the module data getDustT
 * and the grain charge and photoelectric routines read (update_mod.f90:
 * 527-932, 1836-1945). */
int mc_dust_setup(mc_ctx *ctx, const mc_dust_tables *dust);

/* updateCell over every active cell: the gas balance of mc_balance_run when
 * mc_balance_setup was called, then, when mc_dust_setup was called, the grain
 * temperatures of every reached cell (FP64): the absorption integral
 * sum_i absXsec(i) J(i)/pi of each species and size located in its emission
 * integrals and interpolated linearly, then the size- and species-weighted
 * means Tdust(nS,0,c) and Tdust(0,0,c).  With gas and photoelectric = 1, every
 * temperature iteration of the gas solves each grain's potential
 * (setGrainPotential), adds the photoelectric heating and the gas-grain
 * collision cooling to the thermal balance, and the grain temperatures take
 * the last iteration's collision heating and photoelectric loss
 * (update_mod.f90:367-388, 1252-1255, 1886-1889).  Tdust(0:nSpeciesMax, 0:nSizes,
 * 0:nCells) is copied in and back out (column-major); unreached cells keep
 * it.  The gas arrays may be NULL without gas, Tdust without grains.  flags
 * and nIter as mc_balance_run; grains without gas set flags bit0 from
 * Tdust(0,0) and bit4 for unreached cells.  Replaces: the updateCell loop of
 * iterateMC (iteration_mod.f90:826-838; update_mod.f90:15-335 with getDustT,
 * 1836-1945). */
int mc_update_cells(mc_ctx *ctx, double jscale, const mc_balance_opts *opts, const double *Hden,
                    double *Te, double *Ne, double *ionDen, double *Tdust, int32_t *flags,
                    int32_t *nIter, double *kernel_ms);

/* ---- opacity tables -------------------------------------------------------
 * The photoionisation cross-sections addOpacity adds (ionization_mod.f90:
 * 329-484), in its order: the H, He I and He II ground levels, then every
 * element switched on from Z = 3 by atomic number, ions 1..min(Z, nstages),
 * shells 1..nShells (elementP).  Segment s covers bins seg_meta[s][2] ..
 * seg_meta[s][3] (1-based; the last is max(nuLowP, min(nuHighP, nbins)), the
 * range inOpacity loops over) with those bins' cross-sections in seg_sigma,
 * segment after segment. */
typedef struct {
    int32_t nseg;
    const int32_t *seg_meta;    /* [nseg][4]: elem, ion, first bin, last bin */
    const float *seg_sigma;     /* the cross-sections, in the units xSecArray holds them */
    float brems1;               /* xSecArray(bremsXSecP): free-free factor of bin 1 */
} mc_opacity_tables;

/* Upload the segment tables (validated against the atomic data of
 * mc_balance_setup, which must come first).  Replaces: the module data
 * ionizationDriver reads (ionization_mod.f90:26-484). */
int mc_opacity_setup(mc_ctx *ctx, const mc_opacity_tables *tables);

/* The opacity tables of every active cell on the device, in FP32, written
 * into the transport's own tables (the next sweep uses them; pass opacity =
 * NULL to mc_upload_tables).  The gas part is ionizationDriver per cell:
 * ion densities, the free-free term of bin 1 (Hummer's Gaunt factor at the
 * cell's own Te; the reference reuses one for cells within 10 % of the Te it
 * was computed at) and the photoionisation segments.  With grains set up
 * with scaXsec, the dust add follows: grains below their sublimation
 * temperature add grainAbun grainWeight Ndust sigma to scaOpac and absOpac,
 * and opacity gains both.  The state is Hden, Te, Ne (0:nCells) and
 * ionDen(0:nCells, 1:nElemUsed, 1:nstages) as the host holds them (REAL);
 * all four NULL take the device state of the last mc_update_cells /
 * mc_balance_run, rounded to FP32.  Tdust (0:nSpeciesMax, 0:nSizes,
 * 0:nCells), NULL: the device Tdust of the last mc_update_cells.  opacity
 * and scaOpac (nullable): host copies (0:nCells, 1:nbins).  kernel_ms
 * (nullable): device time.  Replaces: the table step of iterateMC
 * (iteration_mod.f90:117-228) with ionizationDriver (ionization_mod.f90:
 * 26-484). */
int mc_opacity_build(mc_ctx *ctx, const float *Hden, const float *Te, const float *Ne, const float *ionDen,
                     const float *Tdust, float *opacity, float *scaOpac, double *kernel_ms);

/* The module data emissionDriver reads for the gas re-emission (set up by
 * the reference's ph_mod, hydro_mod and readData; emission_mod.f90:32-1311).
 * Bins are 1-based; (a, b) tables are column-major as the host holds them.
 *   tkGamma(ntk), logGammaHI/HeI/HeII(ntk, nHI/nHeI/nHeII) with the edge
 *   bins HINuEdgeP/HeINuEdgeP/HeIINuEdgeP (fb_ff);
 *   gauntFF, gauntFFHeII (nbins);
 *   HIPnuP, HeIPnuP, HeIIPnuP = HlevNuP(1), HeIlevNuP(1), HeIIlevNuP(1), and
 *   the ground-state cross-sections the Milne relation reads, from its first
 *   bin to nbins: xSecHI(i) = xSecArray(i - HIPnuP + HlevXSecP(1)) for i >=
 *   HIPnuP; xSecHeI(i) = xSecArray(i - HeIPnuP + HeISingXSecP(1)) for i >=
 *   max(HIPnuP, HeINuEdgeP(nHeI)); xSecHeII(i) = xSecArray(i - HeIIPnuP +
 *   HeIIXSecP(1)) for i >= max(that, HeIINuEdgeP(nHeII));
 *   the heavy-ion free-bound channels in fb_ff's loop order (elements 3..19
 *   switched on, ion 1..min(Z, nstages-1)): fb_meta[nfb][4] = Z, ion,
 *   IPnuP, highNuP (elementP(Z, ion, outShell, 1:2)), fb_gratio[nfb] =
 *   real(g0)/real(g1) (getOuterShell), fb_sigma = the cross-sections of
 *   bins IPnuP..highNuP-1 of every channel, concatenated;
 *   y_dat, Ay_dat (41), HIRecLineData(3:30, 2:8), HeIIRecLineData(3:30,
 *   2:16), HeIrecLineCoeff(34, 3, 4), HeIILymanData and HeIILymanNu
 *   (nHeIILyman = 4), widFlx (nbins). */
typedef struct {
    int32_t ntk, nHI, nHeI, nHeII;
    const float *tkGamma, *logGammaHI, *logGammaHeI, *logGammaHeII;
    const int32_t *HINuEdgeP, *HeINuEdgeP, *HeIINuEdgeP;
    const float *gauntFF, *gauntFFHeII;
    int32_t HIPnuP, HeIPnuP, HeIIPnuP;
    const float *xSecHI, *xSecHeI, *xSecHeII;
    int32_t nfb;
    const int32_t *fb_meta;
    const float *fb_gratio, *fb_sigma;
    const float *y_dat, *Ay_dat;
    const double *HIRecLineData, *HeIIRecLineData;
    const float *HeIrecLineCoeff;
    int32_t nHeIILyman;
    const float *HeIILymanData, *HeIILymanNu;
    const float *widFlx;
} mc_emission_tables;

/* Upload the re-emission tables (validated; after mc_balance_setup, whose
 * model ions give the forbidden lines).  With grains (mc_dust_setup with
 * TdustSublime), setDiffusePDF's dust term is included.  Replaces: the
 * module data emissionDriver reads (emission_mod.f90:32-1311). */
int mc_emission_setup(mc_ctx *ctx, const mc_emission_tables *tables);

/* The re-emission tables of every active cell on the device, written into
 * the transport's own recPDF and totalLines (the next sweep uses them; pass
 * recPDF = totalLines = NULL to mc_upload_tables).  Per cell, emissionDriver's
 * gas part: free-bound and free-free continua of H I, He I, He II and the
 * heavy ions (fb_ff, free_free), the two-photon continua (twoPhoton), the He
 * I 2s triplet and He II Lyman lines able to ionise H, the recombination and
 * forbidden lines (RecLinesEmission, forLines) and setDiffusePDF's sums; with
 * grains, their thermal emission.  The sums are REALs in the reference's
 * order, so recPDF follows its rounding; the forbidden lines' level
 * populations are solved in FP64, and the grains' total is taken in bin
 * order.  The state is Hden, Te, Ne (0:nCells), ionDen(0:nCells,
 * 1:nElemUsed, 1:nstages) and Tdust (0:nSpeciesMax, 0:nSizes, 0:nCells) as
 * the host holds them (REAL); NULL state (and Tdust) take the device state of
 * the last mc_update_cells / mc_balance_run, rounded to FP32.  recPDF
 * (0:nCells, 1:nbins) and totalLines (0:nCells), nullable: host copies of
 * the active cells' rows (row 0 is left as it is).  Where the reference
 * forms NaN (a cell at a few tens of K, where its REAL sums overflow and its
 * fits give inf*0, or a continuum entirely below its 1e-15 cut), the tables
 * hold what its sampling does with them: getNu2's scan stops at a NaN entry
 * as at an entry of 1, so the row is 1 from its first NaN on, and a NaN
 * totalLines, which random <= totalLines never passes, is 0.  MC_ETABLE
 * names a cell whose state is not finite (or Te <= 0); the tables of the
 * previous build or upload are then kept.  kernel_ms (nullable): device
 * time.  Replaces: emissionDriver per cell in iterateMC (iteration_mod.f90:
 * 229-240; emission_mod.f90:32-173, 253-903, 905-1311). */
int mc_emission_build(mc_ctx *ctx, const float *Hden, const float *Te, const float *Ne, const float *ionDen,
                      const float *Tdust, float *recPDF, float *totalLines, double *kernel_ms);

/* The analytic line luminosities of outputGas for a 3-D grid (output_mod.f90:
 * 302-520), summed over every active cell that is not black: per cell,
 * forLines (1946-2008: each model ion's level populations from equilibrium,
 * emissivity a(j,i) (e(j) - e(i)) n(j) times elemAbun, ionDen and the REAL
 * 1.9865e9) and RecLinesEmission's Hbeta (2065-2079: 2530/Te**0.833 Ne
 * ionDen(H,2) elemAbun(H)), each times Hden and getVolume (grid_mod.f90:
 * 2876-2964).  The state is Hden, Te, Ne (0:nCells) and ionDen(0:nCells,
 * 1:nElemUsed, 1:nstages) as the host holds them, with lgBlack (0:nCells,
 * nullable: no black cell); all four NULL take the device state of the last
 * mc_update_cells / mc_balance_run, rounded to FP32, whose cells without a Te
 * or ion balance are black.  Out: forbVol[nCel][17][17] (model ion in
 * forLines order, upper level, lower level; Fortran forbVol(lower, upper,
 * ion)) and HbetaVol, both before outputGas divides by HbetaVol (line ratios
 * are forbVol / HbetaVol, output_mod.f90:864).  Summed in FP64 in a fixed
 * order (the reference sums REALs).  The level populations are solved in
 * FP64 (the reference evaluates the recombination and collision-strength
 * fits in REAL).  kernel_ms (nullable): device time.  Replaces: the analytic
 * emissivity loop of outputGas (output_mod.f90:302-520, the forbVol and
 * HbetaVol sums). */
int mc_line_luminosities(mc_ctx *ctx, const float *Hden, const float *Te, const float *Ne, const float *ionDen,
                         const int32_t *lgBlack, double *forbVol, double *HbetaVol, double *kernel_ms);

/* ---- grid state files ---------------------------------------------------
 * The state a run leaves in output/ and a warm start reads back, in the
 * reference's layouts: Te, Ne, Hden (0:nCells), ionDen(0:nCells,
 * 1:nElemUsed, 1:nstages), lgConverged and lgBlack (0:nCells) as INTEGER,
 * and with grains Ndust (0:nCells) and Tdust(0:nSpeciesMax, 0:nSizes,
 * 0:nCells).  The geometry and the grid flags come from an mc_desc (as for
 * mc_ctx_create), the elements switched on and nstages from an mc_atomic
 * (gas grids).  Neither call touches the GPU. */
typedef struct {
    float *Te, *Ne, *Hden;      /* gas; NULL without */
    float *ionDen;
    int32_t *lgConverged, *lgBlack;
    float *Ndust, *Tdust;       /* grains; NULL without */
    int32_t nSpeciesMax, nSizes;
    float totalDustMass;        /* [1e45 g], written at the end of dustGrid.out */
} mc_grid_state;

/* Write grid0.out, and grid1.out/grid2.out with gas and dustGrid.out with
 * grains, into directory dir, byte for byte as the reference's writeGrid
 * does (one grid; list-directed records, each REAL as its shortest
 * round-trip decimal).  photoSource.out and grid3.out hold the run's input
 * parameters and stay with the host.  Replaces: the state files of
 * writeGrid (grid_mod.f90:2646-2760). */
int mc_write_grid(const mc_desc *desc, const mc_atomic *atomic, const char *dir, const mc_grid_state *state);

/* Read those files back into the state arrays as resetGrid does: the
 * records of inactive lattice points land in row 0 (the last one wins), and
 * each element's ionDen row takes the whole read buffer, so stages above
 * min(Z+1, nstages) keep the previous element's values.  The files' grid
 * (dimensions, axes, active map) must be the descriptor's (MC_ETABLE
 * otherwise).  Replaces: the state part of resetGrid (grid_mod.f90:
 * 3187-3470), read by mocassinWarm. */
int mc_read_grid(const mc_desc *desc, const mc_atomic *atomic, const char *dir, mc_grid_state *state);

/* ---- warm-start network (no counterpart in the reference) -------------
 * A 72 -> 128 -> 128 -> 128 -> 31 ReLU perceptron with BF16 weights and
 * activations and FP32 accumulation on the matrix cores, trained on the
 * device.  Inputs per cell (72): log10 of J summed over 64 contiguous groups
 * of frequency bins, log10 nH, log10 abundances of H, He, C, N, O, Ne, S
 * (-30 when absent).  Outputs (31): log10 Te, then log10 of the ionic
 * fractions of those elements, stages 1..min(Z+1,5).  Parameters: a flat
 * FP32 vector, per layer W[out][in] then b[out]; layer 0 has 96 inputs (72
 * used, the rest zero) and layer 3 has 32 outputs (31 used).  scale holds
 * x_mean[72], x_std[72], y_mean[31], y_std[31]: the network sees
 * (x - x_mean)/x_std and predicts (y - y_mean)/y_std. */
#define MC_MLP_NIN     72
#define MC_MLP_NOUT    31
#define MC_MLP_NPARAMS 49568
#define MC_MLP_NSCALE  206

/* Set parameters and scaling (resets the optimiser state). */
int mc_mlp_set(mc_ctx *ctx, const float *params, const float *scale);
/* Read them back (either may be NULL). */
int mc_mlp_get(mc_ctx *ctx, float *params, float *scale);
/* Network inputs of every active cell from the device J times jscale:
 * X[nCells][72] (row c-1 for cell c), hit(0:nCells) = 1 where a packet
 * reached the cell.  Needs mc_balance_setup. */
int mc_mlp_features(mc_ctx *ctx, double jscale, const double *Hden, float *X, int32_t *hit);
/* Y[n][31] = network(X[n][72]), unscaled. */
int mc_mlp_predict(mc_ctx *ctx, const float *X, int64_t n, float *Y);
/* Loss L = sum M (y - t)^2 / sum M over scaled outputs y and targets t
 * (T[n][31] unscaled, mask M[n][31]) and its gradient dL/dparams. */
int mc_mlp_loss_grad(mc_ctx *ctx, const float *X, const float *T, const float *M, int64_t n,
                     float *grad, double *loss);
/* Adam (lr, 0.9, 0.999, 1e-8) for `steps` steps of `batch` rows; the rows
 * of step s are order[s*batch .. s*batch+batch) (indices into X, T, M).
 * loss[s] (nullable) receives each step's L; kernel_ms (nullable) the
 * device time. */
int mc_mlp_train(mc_ctx *ctx, const float *X, const float *T, const float *M, int64_t n,
                 const int32_t *order, int32_t steps, int32_t batch, float lr, float *loss,
                 double *kernel_ms);

/* Set / read the transport launch shape (see mc_tuning). */
int mc_set_tuning(mc_ctx *ctx, const mc_tuning *t);
int mc_get_tuning(const mc_ctx *ctx, mc_tuning *t);

/* Throughput probe of the device's atomic adds (no counterpart in the
 * reference; the ceiling the transport's J update is measured against):
 * `requests` FP64 (fp64 != 0) or FP32 atomic adds to uniformly random words
 * of a fresh array of `words` elements (1 .. 2^32-1), issued by wavefronts
 * with `lanes` (1..64) active lanes, 8 workgroups of 256 per CU.  One
 * untimed launch, then one timed: kernel_ms.  Allocates and frees its own
 * array; needs no context. */
int mc_probe_atomics(int32_t device, int64_t words, int64_t requests, int32_t lanes, int32_t fp64,
                     double *kernel_ms);

/* The same probe with the adds of each atomic instruction falling in
 * 64-byte lines (no counterpart in the reference): `requests` FP64 adds
 * (a multiple of 64) into an array of `words` (>= 8) doubles; the 64 lanes
 * form 64 / group sets (group 1, 2, 4 or 8) that add to distinct words of
 * one random line each, the set's lanes adjacent (spread 0) or spread
 * across the wavefront (spread 1).  Measures whether the memory side takes
 * a line's adds as one request.  kernel_ms of the timed launch. */
int mc_probe_atomic_lines(int32_t device, int64_t words, int64_t requests, int32_t group, int32_t spread,
                          double *kernel_ms);

/* Message of the last failure on this context (never NULL). */
const char *mc_ctx_last_error(const mc_ctx *ctx);

/* Release all device memory of the context. */
void mc_ctx_destroy(mc_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* MOCASSIN_AMD_H */

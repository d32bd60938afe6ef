// mc_kernels.h - launch arguments shared by the host library and kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcd {

struct Grid;
struct J3Shape;

// Per-launch arguments of the transport kernel.
struct RunArgs {
    int istar;                          // 1-based source index (random stream id)
    long long npackets;
    unsigned long long packet_offset;
    unsigned long long *next_packet;    // global refill counter (zeroed per launch)
    int *error;                         // first fail-stop code (atomicMax)
    unsigned long long *counts;         // 8 event counters, then walk passes and lanes, emission passes and lanes
    double *qphot;
    float deltaE;
    float px, py, pz;                   // source position [cm]
    int sx, sy, sz;                     // source cell (1-based)
    int emit_min;                       // emission pass when >= this many lanes wait
    int walk_steps;                     // walk passes between emission checks
};

// Chunk table and per-model arguments of a batched launch (device memory)
struct BatchArgs {
    const Grid *G;                      // [models]
    const RunArgs *RA;                  // [models] (istar, packet_offset, source, counters)
    const int *chunk_model;             // [nchunks], non-decreasing
    const long long *chunk_first;       // [nchunks] first packet index within the model
    const int *chunk_count;             // [nchunks]
    int nchunks;
    unsigned *next_chunk;               // zeroed per launch
    int emit_min, walk_steps;
};

}  // namespace mcd

hipError_t mcd_launch_batch(const mcd::BatchArgs &B, size_t lds_bytes, int blocks, hipStream_t s);
size_t mcd_lds_bytes(int nx, int ny, int nz, int nbins);
// The transport kernels: one packet per lane; two (at most 3 wavefronts
// per SIMD, which that kernel's registers allow); one per lane with a pool
// of 32 packets per wavefront in LDS.  The pool kernel needs grids of at
// most 2047 x 2047 x 1023 cells and 65535 bins.
enum { MCD_KERNEL_ONE = 1, MCD_KERNEL_SPARE = 2, MCD_KERNEL_POOL = 3 };
hipError_t mcd_launch_transport(const mcd::Grid &G, const mcd::RunArgs &RA, int blocks, int threads, int kernel,
                                hipStream_t s);
hipError_t mcd_densify(const float *compact, const int *active, const mcd::J3Shape &js, int nc1, int nbins,
                       float *dense, hipStream_t s);
hipError_t mcd_guides(const float *cdf, int rows, int nbins, unsigned short *guide, hipStream_t s);
hipError_t mcd_compact(const double *J3, const int *active, const mcd::J3Shape &js, int nc1, int nbins,
                       double *out, hipStream_t s);
hipError_t mcd_launch_replay(const mcd::Grid &G, int n, const float *pos, const float *dir,
                             const int *cell, int *cell_seq, float *ds_seq, int *seq_len,
                             int *status, int max_len, hipStream_t s);

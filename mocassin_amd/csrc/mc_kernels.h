// mc_kernels.h - launch arguments shared by the host library and kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcd {

struct Grid;

// Per-launch arguments of the transport kernel.
struct RunArgs {
    int istar;                          // 1-based source index (random stream id)
    long long npackets;
    unsigned long long packet_offset;
    unsigned long long *next_packet;    // global refill counter (zeroed per launch)
    int *error;                         // first fail-stop code (atomicMax)
    unsigned long long *counts;         // 8 event counters
    double *qphot;
    float deltaE;
    float px, py, pz;                   // source position [cm]
    int sx, sy, sz;                     // source cell (1-based)
};

// Per-lane event counters, reduced per wavefront at kernel exit.
struct Counts {
    uint32_t seg, abs, sca, esc, lost, line, trapped, flights;
    __device__ void clear() { seg = abs = sca = esc = lost = line = trapped = flights = 0; }
    __device__ void reduce_wave()
    {
        for (int o = 32; o >= 1; o >>= 1) {
            seg += __shfl_xor(seg, o);
            abs += __shfl_xor(abs, o);
            sca += __shfl_xor(sca, o);
            esc += __shfl_xor(esc, o);
            lost += __shfl_xor(lost, o);
            line += __shfl_xor(line, o);
            trapped += __shfl_xor(trapped, o);
            flights += __shfl_xor(flights, o);
        }
    }
};

}  // namespace mcd

hipError_t mcd_launch_transport(const mcd::Grid &G, const mcd::RunArgs &RA, int blocks,
                                int threads, hipStream_t s);
hipError_t mcd_launch_replay(const mcd::Grid &G, int n, const float *pos, const float *dir,
                             const int *cell, int *cell_seq, float *ds_seq, int *seq_len,
                             int *status, int max_len, hipStream_t s);

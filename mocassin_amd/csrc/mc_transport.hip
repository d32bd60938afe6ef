// mc_transport.hip - persistent packet-transport kernel and geometry replay
// kernel for MI355X (gfx950).
//
// Transport: one lane owns one packet at a time.  Lanes refill from a
// global packet counter through one atomic per wavefront (wave-aggregated
// fetch), so a wavefront keeps all 64 lanes busy while packet lifetimes
// differ by orders of magnitude (immediate escapes, line packets, long
// re-emission chains).  Every loop iteration advances each live lane by one
// event: a packet creation (stellar or diffuse re-emission) or one cell
// segment of the walk.  The small per-axis arrays and the source CDF are
// staged in LDS; J_nu and the escape SED are FP64 accumulators updated with
// hardware FP64 atomics (global_atomic_add_f64).
#include "mc_device.h"
#include "mc_kernels.h"

#pragma clang fp contract(off)

namespace mcd {

// LDS layout: axes, mids, dV factors for x,y,z followed by the source CDF
__device__ __forceinline__ void stage_axes(const Grid &G, float *lds, Axes &A, int extra_nu,
                                           const float *extra_src, float **extra_dst)
{
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int na = nx + ny + nz;
    const int nm = (nx - 1) + (ny - 1) + (nz - 1);
    float *ax = lds, *mi = lds + na, *df = lds + na + nm;
    for (int i = threadIdx.x; i < na; i += blockDim.x) { ax[i] = G.axis[i]; df[i] = G.dfac[i]; }
    for (int i = threadIdx.x; i < nm; i += blockDim.x) mi[i] = G.mid[i];
    float *ex = lds + 2 * na + nm;
    for (int i = threadIdx.x; i < extra_nu; i += blockDim.x) ex[i] = extra_src[i];
    __syncthreads();
    A.x = ax; A.y = ax + nx; A.z = ax + nx + ny;
    A.mx = mi; A.my = mi + (nx - 1); A.mz = mi + (nx - 1) + (ny - 1);
    A.dx = df; A.dy = df + nx; A.dz = df + nx + ny;
    if (extra_dst) *extra_dst = ex;
}

// ------------------------------------------------------------ replay kernel
__global__ void __launch_bounds__(256) walk_replay_kernel(Grid G, int n, const float *pos,
                                                          const float *dir, const int *cell,
                                                          int *cell_seq, float *ds_seq, int *seq_len,
                                                          int *status, int max_len)
{
    extern __shared__ float lds[];
    Axes A;
    stage_axes(G, lds, A, 0, nullptr, nullptr);
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    Walker W;
    W.rx = pos[3 * p]; W.ry = pos[3 * p + 1]; W.rz = pos[3 * p + 2];
    W.vx = dir[3 * p]; W.vy = dir[3 * p + 1]; W.vz = dir[3 * p + 2];
    W.xP = cell[3 * p]; W.yP = cell[3 * p + 1]; W.zP = cell[3 * p + 2];
    W.absTau = 0.f; W.passProb = 0.f;
    int len = 0, st = 3;
    for (int i = 1; i <= kSafeLimit; ++i) {
        float kappa, dV;
        int err = 0;
        int r = walk_step<true>(G, A, W, 1, kappa, dV, err);
        if (r == ST_ERROR) { st = -err; break; }
        if (r == ST_LOST) { st = 2; break; }
        if (len < max_len) {
            size_t o = (size_t)p * max_len + len;
            cell_seq[3 * o] = W.xP; cell_seq[3 * o + 1] = W.yP; cell_seq[3 * o + 2] = W.zP;
            ds_seq[o] = W.seg;
        }
        ++len;
        if (r == ST_ESCAPED) { st = 1; break; }
    }
    seq_len[p] = len;
    status[p] = st;
}

// ----------------------------------------------------------- transport kernel
enum Phase { PH_IDLE = 0, PH_STELLAR = 1, PH_WALK = 2, PH_DIFFUSE = 3, PH_DONE = 4 };

struct Lane {
    Rng R;
    Walker W;
    float d0x, d0y, d0z;   // enPacket%direction (start of flight / last scatter)
    int nuP, origin;
    int steps, runs, phase;
};

__device__ __forceinline__ void esc_add(const Grid &G, int origin, int nuP, float dE)
{
    unsafeAtomicAdd(&G.esc[(size_t)nuP * (size_t)(G.nCells + 1) + origin], (double)dE);
}

__device__ __forceinline__ void j_add(const Grid &G, int cell, int nuP, float v)
{
    unsafeAtomicAdd(&G.J[(size_t)(nuP - 1) * (size_t)(G.nCells + 1) + cell], (double)v);
}

__device__ __forceinline__ void begin_flight(Lane &L, const Grid &G, Counts &C)
{
    L.W.absTau = 0.f;
    float random = draw(L.R, G);
    L.W.passProb = -mcf_logf(1.f - random);
    L.steps = 0;
    C.flights++;
}

__global__ void __launch_bounds__(256) transport_kernel(Grid G, RunArgs RA)
{
    extern __shared__ float lds[];
    Axes A;
    float *spec = nullptr;
    stage_axes(G, lds, A, G.nbins, G.specCDF + (size_t)(RA.istar - 1) * G.nbins, &spec);

    const int lane_id = threadIdx.x & 63;
    const bool sym = (G.flags & F_SYM) != 0;
    const bool dust = (G.flags & F_DUST) != 0;
    const bool gas = (G.flags & F_GAS) != 0;
    const bool iso = (G.flags & F_ISO) != 0;
    const bool fold = sym && !(G.flags & F_MULTI);
    const float dE = RA.deltaE;
    const int nbins = G.nbins;

    Lane L;
    L.phase = PH_IDLE;
    L.runs = 0; L.steps = 0; L.nuP = 1; L.origin = 0;
    bool exhausted = false;
    Counts C;
    C.clear();
    double qphot = 0.0;

    for (;;) {
        // ---- wave-aggregated refill
        bool need = (L.phase == PH_IDLE) && !exhausted;
        uint64_t m = __ballot(need);
        if (m) {
            int leader = __ffsll((unsigned long long)m) - 1;
            unsigned long long base = 0;
            if (lane_id == leader)
                base = atomicAdd(RA.next_packet, (unsigned long long)__popcll(m));
            base = __shfl(base, leader);
            if (need) {
                unsigned long long rank = __popcll(m & ((1ull << lane_id) - 1ull));
                unsigned long long id = base + rank;
                if (id < (unsigned long long)RA.npackets) {
                    rng_start(L.R, RA.packet_offset + id, (uint32_t)RA.istar);
                    L.phase = PH_STELLAR;
                    L.runs = 0;
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__any(L.phase != PH_IDLE)) break;

        int err = 0;
        const bool was_live = L.phase != PH_IDLE;
        if (L.phase == PH_STELLAR) {
            // newPhotonPacket("stellar"), photon_mod.f90:800-862
            L.runs = 1;
            int nuP = getnu2(L.R, G, spec, nbins, err);
            if (!err && (nuP < 1 || nuP >= nbins)) err = 11;
            if (!err) {
                L.nuP = nuP;
                new_direction(L.R, G, fold, L.d0x, L.d0y, L.d0z);
                L.W.xP = RA.sx; L.W.yP = RA.sy; L.W.zP = RA.sz;
                L.W.rx = RA.px; L.W.ry = RA.py; L.W.rz = RA.pz;
                L.origin = act(G, RA.sx, RA.sy, RA.sz);
                float nu = G.nuArray[nuP - 1];
                if (nu > 1.f) qphot += (double)(dE / (2.1799153e-11f * nu));
                if (!dust && nu < G.ionEdge1) {          // 370-465
                    esc_add(G, L.origin, nuP, dE); C.esc++;
                    L.phase = PH_IDLE;
                } else {
                    L.W.vx = L.d0x; L.W.vy = L.d0y; L.W.vz = L.d0z;
                    begin_flight(L, G, C);
                    L.phase = PH_WALK;
                }
            }
        } else if (L.phase == PH_DIFFUSE) {
            // newPhotonPacket("diffuse"), photon_mod.f90:907-992
            L.runs++;
            int a = act(G, L.W.xP, L.W.yP, L.W.zP);
            if (!gas) err = 13;
            else if (a <= 0) err = 14;
            if (!err) {
                float random = 1.f - draw(L.R, G);
                if (random <= G.totalLines[a]) {
                    C.line++;                             // line packet: escapes untracked
                    L.phase = PH_IDLE;
                } else {
                    int nuP = getnu2(L.R, G, G.recCDF + (size_t)a * nbins, nbins, err);
                    if (!err && (nuP < 1 || nuP >= nbins)) err = 15;
                    if (!err) {
                        L.nuP = nuP;
                        new_direction(L.R, G, false, L.d0x, L.d0y, L.d0z);
                        L.origin = a;
                        float nu = G.nuArray[nuP - 1];
                        if (!dust && nu < G.ionEdge1) {
                            esc_add(G, L.origin, nuP, dE); C.esc++;
                            L.phase = PH_IDLE;
                        } else {
                            L.W.vx = L.d0x; L.W.vy = L.d0y; L.W.vz = L.d0z;
                            begin_flight(L, G, C);
                            L.phase = PH_WALK;
                        }
                    }
                }
            }
        } else if (L.phase == PH_WALK) {
            L.steps++;
            float kappa, dV;
            int r = walk_step<false>(G, A, L.W, L.nuP, kappa, dV, err);
            if (r == ST_CONTINUE || r == ST_ESCAPED) {
                j_add(G, L.W.segCell, L.nuP, L.W.seg * dE / dV);
                C.seg++;
                if (r == ST_ESCAPED) { esc_add(G, L.origin, L.nuP, dE); C.esc++; L.phase = PH_IDLE; }
            } else if (r == ST_LOST) {
                C.lost++;
                L.phase = PH_IDLE;
            } else if (r == ST_ABSORBED) {
                const int a = L.W.segCell;
                j_add(G, a, L.nuP, L.W.seg * dE / dV);
                C.seg++;
                float rx = L.W.rx / 1.e10f, ry = L.W.ry / 1.e10f, rz = L.W.rz / 1.e10f;
                if (sqrtf(rx * rx + ry * ry + rz * rz) * 1.e10f >= G.R_out && G.R_out > 0.f) {
                    esc_add(G, L.origin, L.nuP, dE); C.esc++;     // 1577-1669
                    L.phase = PH_IDLE;
                } else if (dust) {
                    float probSca = G.scaOpac[(size_t)(L.nuP - 1) * (size_t)(G.nCells + 1) + a] / kappa;
                    float u = 1.f - draw(L.R, G);
                    if (u > probSca) {
                        C.abs++;
                        L.phase = gas ? PH_DIFFUSE : PH_DONE;   // dustEmi needs dust PDFs
                        if (!gas) err = 41;
                    } else if (u <= probSca) {
                        C.sca++;                                  // 1720-1797
                        L.origin = a;
                        if (iso) {
                            new_direction(L.R, G, false, L.d0x, L.d0y, L.d0z);
                        } else {
                            for (int ihg = 1; ihg <= 10; ++ihg)
                                if (hg(L.R, G, L.nuP, L.d0x, L.d0y, L.d0z) == 0) break;
                        }
                        L.W.vx = L.d0x; L.W.vy = L.d0y; L.W.vz = L.d0z;
                        L.W.absTau = 0.f;
                        float r2 = draw(L.R, G);
                        L.W.passProb = -mcf_logf(1.f - r2);
                    } else {
                        err = 39;
                    }
                } else {
                    C.abs++;
                    L.phase = PH_DIFFUSE;
                }
            } else {
                err = err ? err : 50;
            }
            // safe limit and recursion limit (2838-2846, 120-126)
            if (!err && L.phase == PH_WALK && L.steps >= kSafeLimit) { C.lost++; L.phase = PH_IDLE; }
            if (!err && L.phase == PH_DIFFUSE && L.steps >= kSafeLimit) { C.lost++; L.phase = PH_IDLE; }
            // energyPacketDriver abandons the packet after recursionLimit runs
            if (!err && L.phase == PH_DIFFUSE && L.runs >= kRecursionLimit) L.phase = PH_IDLE;
        }
        if (err) {
            atomicMax(RA.error, err);
            L.phase = PH_IDLE;
        } else if (was_live && L.phase == PH_IDLE && L.runs >= kRecursionLimit) {
            C.trapped++;                                   // photon_mod.f90:126
        }
        if (L.phase == PH_DONE) L.phase = PH_IDLE;
    }

    // ---- wave reduction of counters
    C.reduce_wave();
    for (int o = 32; o >= 1; o >>= 1) qphot += __shfl_xor(qphot, o);
    if (lane_id == 0) {
        atomicAdd(&RA.counts[0], C.seg);
        atomicAdd(&RA.counts[1], C.abs);
        atomicAdd(&RA.counts[2], C.sca);
        atomicAdd(&RA.counts[3], C.esc);
        atomicAdd(&RA.counts[4], C.lost);
        atomicAdd(&RA.counts[5], C.line);
        atomicAdd(&RA.counts[6], C.trapped);
        atomicAdd(&RA.counts[7], C.flights);
        unsafeAtomicAdd(RA.qphot, qphot);
    }
}

}  // namespace mcd

// ---------------------------------------------------------------- launchers
hipError_t mcd_launch_transport(const mcd::Grid &G, const mcd::RunArgs &RA, int blocks,
                                int threads, hipStream_t s)
{
    size_t lds = sizeof(float) * (size_t)(2 * (G.nx + G.ny + G.nz) + (G.nx + G.ny + G.nz - 3) + G.nbins);
    hipLaunchKernelGGL(mcd::transport_kernel, dim3(blocks), dim3(threads), lds, s, G, RA);
    return hipGetLastError();
}

hipError_t mcd_launch_replay(const mcd::Grid &G, int n, const float *pos, const float *dir,
                             const int *cell, int *cell_seq, float *ds_seq, int *seq_len,
                             int *status, int max_len, hipStream_t s)
{
    size_t lds = sizeof(float) * (size_t)(2 * (G.nx + G.ny + G.nz) + (G.nx + G.ny + G.nz - 3));
    int threads = 256, blocks = (n + threads - 1) / threads;
    hipLaunchKernelGGL(mcd::walk_replay_kernel, dim3(blocks), dim3(threads), lds, s, G, n, pos,
                       dir, cell, cell_seq, ds_seq, seq_len, status, max_len);
    return hipGetLastError();
}

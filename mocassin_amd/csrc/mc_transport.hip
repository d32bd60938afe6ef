// mc_transport.hip - persistent packet-transport kernels and geometry replay
// kernel for MI355X (gfx950).
//
// Transport: lanes refill from a global packet counter with one atomic per
// wavefront.  Packet events are split in two classes and scheduled per
// wavefront so that SIMD lanes stay coherent:
//   walk      one cell segment of pathSegment (a few FP32 divisions, one
//             opacity gather, one J update): the common event;
//   emission  creation of a stellar packet or re-emission of an absorbed one
//             (Philox draws, CDF search, direction, optical-depth draw):
//             about ten times the instructions of a walk step.
// Three kernels share that code and differ in how a wavefront keeps its
// lanes busy between emission passes: one packet per lane
// (transport_kernel<false>), a spare packet per lane in registers
// (transport_kernel<true>), and a pool of packets per wavefront in LDS
// (pool_kernel).  The per-axis arrays, the source CDF and its guide table
// live in LDS.  J_nu and the opacity tables are stored in 2x2x2 bricks of
// 4x4x4 tiles (mc_device.h); each lane queues its J updates in a ring in
// LDS, and a full ring flushes every lane's ring with FP64 atomics, one
// packet's consecutive updates to an instruction, so that adds sharing a
// 64-byte line travel as one request.
#include "mc_device.h"
#include "mc_kernels.h"

#pragma clang fp contract(off)

namespace mcd {

// LDS layout: axes | mids | dV factors | source CDF | source guide (ushort)
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

__host__ __device__ inline size_t lds_floats(int nx, int ny, int nz, int nbins)
{
    return (size_t)(2 * (nx + ny + nz) + (nx + ny + nz - 3) + nbins) + (kGuide + 2) / 2 + 1;
}

// ------------------------------------------------ J update queue (per wave)
// Each lane keeps its last D updates (J3 point, bin - 1, value) in a ring of
// its own at the start of the wavefront's dynamic LDS: point u32 [64 x
// (D + 1)] | value f32 [...] | bin u16 [...], entry k of lane L at
// L * (D + 1) + k (rows padded by one word against bank conflicts).  When a
// lane's ring is full the wavefront flushes every ring, lane-major: atomic
// instruction j carries the rings of lanes (64 / D) j .. (64 / D)(j + 1) - 1.
// Adds of one instruction that fall in one 64-byte line travel to memory as
// one request, and with J3's 2x2x2 bricks a packet's next segment stays in
// the line after half of its steps, so keeping a packet's updates in one
// instruction turns those pairs into single requests.  The single-model
// kernel keeps 8 per lane; the batched kernel, whose one-wavefront
// workgroups each hold a model's tables too, keeps 4 so that 16 fit a CU.
template <int D>
struct JRing {
    static_assert(D == 4 || D == 8, "rings of 4 or 8 updates");
    static constexpr int kRow = D + 1;
    static constexpr int kEntries = 64 * kRow;
    static constexpr int kFloats = kEntries * 2 + kEntries / 2;
};
constexpr int kJDepth = 8;                            // single-model kernel
constexpr int kJDepthBatch = 4;                       // batched kernel

struct JQ {
    unsigned *pt;
    float *val;
    unsigned short *nu;
    int n;                // this lane's queued updates
};

template <int D>
__device__ __forceinline__ JQ jq_at(float *lds, int wave)
{
    float *b = lds + (size_t)wave * JRing<D>::kFloats;
    JQ Q;
    Q.pt = reinterpret_cast<unsigned *>(b);
    Q.val = b + JRing<D>::kEntries;
    Q.nu = reinterpret_cast<unsigned short *>(b + 2 * JRing<D>::kEntries);
    Q.n = 0;
    return Q;
}

// every queued update: instruction j takes entry lane % D of lane
// (64 / D) j + lane / D's ring
template <int D>
__device__ __forceinline__ void jq_flush(const Grid &G, JQ &Q, int lane)
{
    constexpr int per = 64 / D;                       // rings per instruction
    __builtin_amdgcn_wave_barrier();
    const int k = lane % D;
    for (int j = 0; j < D; ++j) {
        const int owner = per * j + lane / D;
        const int cnt = __shfl(Q.n, owner);
        const bool act = k < cnt;
        if (!__any(act)) continue;
        if (act) {
            const int slot = owner * JRing<D>::kRow + k;
            const unsigned pt = Q.pt[slot];
            unsafeAtomicAdd(&G.J3[j3_index(G.nbins, (int)Q.nu[slot], pt)], (double)Q.val[slot]);
            if (G.cross) atomicAdd(&G.cross[pt], 1u);
        }
    }
    __builtin_amdgcn_wave_barrier();
    Q.n = 0;
}

// the J update of each lane with `has` (path length times deltaE over dV,
// photon_mod.f90:1563-1574, 1822-1833), flushed when a lane's ring is full
template <int D>
__device__ __forceinline__ void jq_push(const Grid &G, JQ &Q, bool has, unsigned pt, int nuP, float v, int lane)
{
    if (has) {
        const int slot = lane * JRing<D>::kRow + Q.n;
        Q.pt[slot] = pt;
        Q.val[slot] = v;
        Q.nu[slot] = (unsigned short)(nuP - 1);
        Q.n += 1;
    }
    if (__any(Q.n == D)) jq_flush<D>(G, Q, lane);
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
            cell_seq[3 * o] = W.segX; cell_seq[3 * o + 1] = W.segY; cell_seq[3 * o + 2] = W.segZ;
            ds_seq[o] = W.seg;
        }
        ++len;
        if (r == ST_ESCAPED) { st = 1; break; }
    }
    seq_len[p] = len;
    status[p] = st;
}

// ----------------------------------------------------------- transport kernel
enum Phase { PH_IDLE = 0, PH_EMIT_STELLAR = 1, PH_EMIT_DIFFUSE = 2, PH_WALK = 3 };

struct Lane {
    Rng R;
    Walker W;
    float d0x, d0y, d0z;   // enPacket%direction (start of flight / last scatter)
    int nuP, origin;
    int steps, runs, phase;
};

// wave-uniform event counters (scalar registers)
struct WaveCounts {
    unsigned long long seg, abs, sca, esc, lost, line, trapped, flights;
    unsigned long long wpass, wlanes, epass, elanes;    // passes and the lanes in them
};

__device__ __forceinline__ unsigned long long wcount(bool p)
{
    return (unsigned long long)__popcll(__ballot(p));
}

__device__ __forceinline__ void esc_plane(const Grid &G, int a, int origin, int nuP, float dE)
{
    const size_t i = ((size_t)a * (size_t)(G.nbins + 1) + (size_t)nuP) * (size_t)(G.nCells + 1) + origin;
    unsafeAtomicAdd(&G.esc[i], (double)dE);
}

// An escape of a packet whose flight direction (enPacket%direction) is d:
// its theta and phi bins, then the viewing-angle planes they feed, as every
// escape site of pathSegment and energyPacketRun does
// (photon_mod.f90:1582-1666, 370-465).  err 60/61: the reference's stops on
// a bin out of range.  The bin tables were range-checked on the host.
__device__ __forceinline__ void esc_add(const Grid &G, int origin, int nuP, float dE, float dx, float dy,
                                        float dz, int &err)
{
    const View *V = G.view;
    if (!V) { esc_plane(G, 0, origin, nuP, dE); return; }
    const bool sym = (G.flags & F_SYM) != 0;
    const float q = mcf_acosf(sym ? fabsf(dz) : dz) / V->dTheta;
    if (isnan(q)) { err = 60; return; }
    int it = (int)q + 1;
    if (it > V->totT) it = V->totT;
    if (it < 1) { err = 60; return; }
    int ip;
    if (fabsf(dx) < 1.e-35f) ip = 0;
    else if (sym) ip = (int)(mcf_atanf(fabsf(dy) / fabsf(dx)) / V->dPhi);
    else ip = (int)(mcf_atanf(dy / dx) / V->dPhi);
    if (ip < 0) ip = V->totP + ip;
    ip = ip + 1;
    if (ip > V->totP) ip = V->totP;
    if (ip < 1) { err = 61; return; }
    const int t = V->Pt[it], ph = V->Pp[ip];
    if (t > 0 && V->vph[t] < 0.f) {
        esc_plane(G, t, origin, nuP, dE);
        esc_plane(G, 0, origin, nuP, dE);
    } else if (t == ph || V->vth[ph] == V->vth[t] || V->vph[t] == V->vph[ph]) {
        esc_plane(G, t, origin, nuP, dE);
        if (t != 0) esc_plane(G, 0, origin, nuP, dE);
    } else {
        esc_plane(G, 0, origin, nuP, dE);
    }
}

// newPhotonPacket + initPhotonPacket + the start of pathSegment
// (photon_mod.f90:768-1059, 491-684, 1176-1184) for stellar packets
// (PH_EMIT_STELLAR) and re-emissions at an absorption point (PH_EMIT_DIFFUSE).
__device__ __forceinline__ void emit(Lane &L, const Grid &G, const RunArgs &RA, const float *spec,
                                     const unsigned short *spec_guide, bool fold, bool dust,
                                     bool gas, int star_cell, double &qphot, bool &line_end,
                                     bool &esc_end, bool &flight, int &err)
{
    const bool stellar = (L.phase == PH_EMIT_STELLAR);
    const float *cdf;
    const unsigned short *guide;
    int a;
    if (stellar) {
        L.runs = 1;
        L.W.xP = RA.sx; L.W.yP = RA.sy; L.W.zP = RA.sz;
        L.W.rx = RA.px; L.W.ry = RA.py; L.W.rz = RA.pz;
        a = star_cell;
        cdf = spec;
        guide = spec_guide;
    } else {
        L.runs++;
        a = act(G, L.W.xP, L.W.yP, L.W.zP);
        if (!gas) { err = 13; return; }
        if (a <= 0) { err = 14; return; }
        float random = 1.f - draw(L.R, G);
        if (random <= G.totalLines[a]) {                 // line packet (923-956)
            line_end = true;
            L.phase = PH_IDLE;
            return;
        }
        cdf = G.recCDF + (size_t)a * G.nbins;
        guide = G.recGuide + (size_t)a * (kGuide + 1);
    }
    int nuP = getnu2(L.R, G, cdf, guide, G.nbins, err);
    if (err) return;
    if (nuP < 1 || nuP >= G.nbins) { err = stellar ? 11 : 15; return; }
    L.nuP = nuP;
    new_direction(L.R, G, stellar && fold, L.d0x, L.d0y, L.d0z);
    L.origin = a;
    const float nu = G.nuArray[nuP - 1];
    if (stellar && nu > 1.f) qphot += (double)(RA.deltaE / (2.1799153e-11f * nu));
    if (!dust && nu < G.ionEdge1) {                      // escapes at birth (370-465)
        esc_add(G, a, nuP, RA.deltaE, L.d0x, L.d0y, L.d0z, err);
        esc_end = true;
        L.phase = PH_IDLE;
        return;
    }
    L.W.vx = L.d0x; L.W.vy = L.d0y; L.W.vz = L.d0z;
    L.W.absTau = 0.f;
    float random = draw(L.R, G);
    L.W.passProb = -mcf_logf(1.f - random);
    L.steps = 0;
    flight = true;
    L.phase = PH_WALK;
}

// Model-wide constants of one grid and source, read by the wavefront loop.
struct Model {
    bool fold, dust, gas, iso;
    float dE;
    int star_cell;
};

__device__ __forceinline__ Model model_of(const Grid &G, const RunArgs &RA)
{
    Model M;
    const bool sym = (G.flags & F_SYM) != 0;
    M.dust = (G.flags & F_DUST) != 0;
    M.gas = (G.flags & F_GAS) != 0;
    M.iso = (G.flags & F_ISO) != 0;
    M.fold = sym && !(G.flags & F_MULTI);
    M.dE = RA.deltaE;
    M.star_cell = act(G, RA.sx, RA.sy, RA.sz);
    return M;
}

// Stage the grid's axes, the source CDF and its guide table in LDS.
__device__ __forceinline__ void stage_model(const Grid &G, const RunArgs &RA, float *lds, Axes &A,
                                            float *&spec, unsigned short *&spec_guide)
{
    stage_axes(G, lds, A, G.nbins, G.specCDF + (size_t)(RA.istar - 1) * G.nbins, &spec);
    spec_guide = (unsigned short *)(spec + G.nbins);
    for (int j = threadIdx.x; j <= kGuide; j += blockDim.x) spec_guide[j] = guide_entry(spec, G.nbins, j);
    __syncthreads();
}

// An emission pass of the wavefront: every lane whose packet E waits for an
// emission runs it together.  Counts events into C.
__device__ __forceinline__ void emission_pass(Lane &E, const Grid &G, const RunArgs &RA, const Model &M,
                                              const float *spec, const unsigned short *spec_guide, WaveCounts &C,
                                              double &qphot, int &err)
{
    const bool pend = (E.phase == PH_EMIT_STELLAR || E.phase == PH_EMIT_DIFFUSE);
    bool line_end = false, esc_end = false, flight = false;
    if (pend) {
        emit(E, G, RA, spec, spec_guide, M.fold, M.dust, M.gas, M.star_cell, qphot, line_end, esc_end, flight,
             err);
        if (err) E.phase = PH_IDLE;
    }
    C.epass += 1;
    C.elanes += wcount(pend);
    C.line += wcount(line_end);
    C.esc += wcount(esc_end);
    C.flights += wcount(flight);
    // energyPacketDriver counts the packet whose last run is the
    // recursionLimit-th as trapped (photon_mod.f90:126)
    C.trapped += wcount(!err && (line_end || esc_end) && E.runs >= kRecursionLimit);
}

// One walk pass of the wavefront: every lane whose packet L walks takes one
// step of pathSegment.  Counts events into C, queues the J updates.
template <int D>
__device__ __forceinline__ void walk_pass(Lane &L, const Grid &G, const Axes &A, const Model &M, WaveCounts &C,
                                          int &err, JQ &Q, int lane)
{
    bool seg = false, escaped = false, lost = false, absorbed = false, scattered = false,
         trapped = false;
    float jv = 0.f;
    unsigned jpt = 0;
    int jnu = 1;
    C.wpass += 1;
    C.wlanes += wcount(L.phase == PH_WALK);
    if (L.phase == PH_WALK) {
        L.steps++;
        float kappa, dV;
        int r = walk_step<false>(G, A, L.W, L.nuP, kappa, dV, err);
        if (r == ST_CONTINUE || r == ST_ESCAPED) {
            jv = L.W.seg * M.dE / dV;
            jpt = L.W.segPt;
            jnu = L.nuP;
            seg = true;
            if (r == ST_ESCAPED) {
                esc_add(G, L.origin, L.nuP, M.dE, L.d0x, L.d0y, L.d0z, err);
                escaped = true;
                L.phase = PH_IDLE;
            }
        } else if (r == ST_LOST) {
            lost = true;
            L.phase = PH_IDLE;
        } else if (r == ST_ABSORBED) {
            jv = L.W.seg * M.dE / dV;
            jpt = L.W.segPt;
            jnu = L.nuP;
            seg = true;
            float rx = L.W.rx / 1.e10f, ry = L.W.ry / 1.e10f, rz = L.W.rz / 1.e10f;
            if (sqrtf(rx * rx + ry * ry + rz * rz) * 1.e10f >= G.R_out && G.R_out > 0.f) {
                esc_add(G, L.origin, L.nuP, M.dE, L.d0x, L.d0y, L.d0z, err);   // 1577-1669
                escaped = true;
                L.phase = PH_IDLE;
            } else if (M.dust) {
                float probSca = G.sca3[j3_index(G.nbins, L.nuP - 1, L.W.segPt)] / kappa;
                float u = 1.f - draw(L.R, G);
                if (u > probSca) {
                    absorbed = true;
                    if (M.gas) L.phase = PH_EMIT_DIFFUSE;
                    else err = 41;                        // dustEmi needs dust PDFs
                } else if (u <= probSca) {
                    scattered = true;                     // 1720-1797
                    L.origin = act(G, L.W.xP, L.W.yP, L.W.zP);
                    if (M.iso) {
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
                absorbed = true;
                L.phase = PH_EMIT_DIFFUSE;
            }
        } else {
            if (!err) err = 50;
        }
        if (!err) {
            // safeLimit (2838-2846); recursionLimit (120-126)
            if ((L.phase == PH_WALK || L.phase == PH_EMIT_DIFFUSE) && L.steps >= kSafeLimit) {
                lost = true;
                L.phase = PH_IDLE;
            }
            if (L.phase == PH_EMIT_DIFFUSE && L.runs >= kRecursionLimit) L.phase = PH_IDLE;
            trapped = (L.phase == PH_IDLE) && L.runs >= kRecursionLimit;
        } else {
            L.phase = PH_IDLE;
        }
    }
    jq_push<D>(G, Q, seg, jpt, jnu, jv, lane);
    C.seg += wcount(seg);
    C.esc += wcount(escaped);
    C.lost += wcount(lost);
    C.abs += wcount(absorbed);
    C.sca += wcount(scattered);
    C.trapped += wcount(trapped);
}

// One scheduling round of a wavefront (the batched kernel): an emission pass
// when at least emit_min lanes (or as many as are walking) wait for one, then
// walk_steps walk passes.  A fail-stop code goes to RA.error.
template <int D>
__device__ __forceinline__ void advance(Lane &L, const Grid &G, const Axes &A, const RunArgs &RA,
                                        const Model &M, const float *spec,
                                        const unsigned short *spec_guide, int emit_min, int walk_steps,
                                        int npend, int nwalk, WaveCounts &C, double &qphot, int &err, JQ &Q,
                                        int lane)
{
    if (npend >= emit_min || npend >= nwalk) emission_pass(L, G, RA, M, spec, spec_guide, C, qphot, err);
    for (int s = 0; s < walk_steps; ++s) walk_pass<D>(L, G, A, M, C, err, Q, lane);
    if (err) {
        atomicMax(RA.error, err);
        err = 0;
    }
}

__device__ __forceinline__ bool pending(const Lane &L)
{
    return L.phase == PH_EMIT_STELLAR || L.phase == PH_EMIT_DIFFUSE;
}

// Exchange the packets of the lanes where c holds (a wavefront-wide vote
// first: nothing to do costs one instruction).
__device__ __forceinline__ void swap_where(bool c, Lane &a, Lane &b)
{
    if (!__any(c)) return;
#define MC_SWAP(f) { const auto t_ = a.f; a.f = c ? b.f : a.f; b.f = c ? t_ : b.f; }
    MC_SWAP(R.plo) MC_SWAP(R.phi) MC_SWAP(R.stream) MC_SWAP(R.k) MC_SWAP(R.blk)
    MC_SWAP(R.w0) MC_SWAP(R.w1) MC_SWAP(R.w2) MC_SWAP(R.w3)
    MC_SWAP(W.rx) MC_SWAP(W.ry) MC_SWAP(W.rz) MC_SWAP(W.vx) MC_SWAP(W.vy) MC_SWAP(W.vz)
    MC_SWAP(W.xP) MC_SWAP(W.yP) MC_SWAP(W.zP) MC_SWAP(W.absTau) MC_SWAP(W.passProb)
    MC_SWAP(d0x) MC_SWAP(d0y) MC_SWAP(d0z) MC_SWAP(nuP) MC_SWAP(origin) MC_SWAP(steps) MC_SWAP(runs)
    MC_SWAP(phase)
#undef MC_SWAP
}

// One scheduling round of a wavefront with two packets per lane (the
// single-model kernel): L walks, S is the spare the emission pass serves.
// A packet that waits for an emission moves to S first; after the emission
// pass, and after every walk pass, a lane whose L does not walk takes S's
// packet when that one does.  A lane then idles in a walk pass only when
// both of its packets wait, so the walk passes stay fuller than with one
// packet per lane; the packets and their streams are the same.
template <int D>
__device__ __forceinline__ void advance_spare(Lane &L, Lane &S, const Grid &G, const Axes &A, const RunArgs &RA,
                                              const Model &M, const float *spec, const unsigned short *spec_guide,
                                              int emit_min, int walk_steps, WaveCounts &C, double &qphot, int &err,
                                              JQ &Q, int lane)
{
    swap_where(pending(L) && !pending(S), L, S);
    const int npend = (int)wcount(pending(S));
    const int nwalk = (int)wcount(L.phase == PH_WALK);
    if (npend > 0 && (npend >= emit_min || npend >= nwalk))
        emission_pass(S, G, RA, M, spec, spec_guide, C, qphot, err);
    swap_where(L.phase != PH_WALK && S.phase == PH_WALK, L, S);
    for (int s = 0; s < walk_steps; ++s) {
        walk_pass<D>(L, G, A, M, C, err, Q, lane);
        swap_where(L.phase != PH_WALK && S.phase == PH_WALK, L, S);
    }
    if (err) {
        atomicMax(RA.error, err);
        err = 0;
    }
}

// ------------------------------------------------ packet pool (per wave, LDS)
// A wavefront keeps a pool of P packets in LDS beside its 64 lanes: packets
// that walk (the walker stack, slots 0 .. nw-1 from the bottom) and packets
// that wait for a diffuse emission (the pending stack, slots P-npd .. P-1
// from the top).  A lane whose flight ends takes a walker from the pool,
// leaving its own packet on the pending stack if it waits for an emission;
// an emission round moves walkers into the pool, so that the emission pass
// runs on a full wavefront, and the walkers come back as flights end.  The
// walk passes then stay nearly full between emissions.  Word k of slot s is
// at pw[k * P + s]; the cell indices are packed in one word (11, 11 and 10
// bits: the host runs this kernel on grids of at most 2047 x 2047 x 1023
// cells) and the frequency bin, run count and phase in another (16, 14 and
// 2 bits).  The Rng's stream is the launch's source, not stored.
constexpr int kPoolWords = 23;
// Slots per wavefront: with 8-deep J rings, 4 workgroups of 256 per CU keep
// 32 slots within the CU's 160 KB of LDS (8.5 KB per wavefront, plus the
// staged tables); 48 slots with 4-deep rings ran slower (the rings merge
// fewer J adds per instruction).
constexpr int kPoolSlots = 32;

template <int P>
__device__ __forceinline__ void pool_put(uint32_t *pw, int s, const Lane &L)
{
    uint32_t *q = pw + s;
    q[0 * P] = L.R.plo; q[1 * P] = L.R.phi; q[2 * P] = L.R.k; q[3 * P] = L.R.blk;
    q[4 * P] = L.R.w0; q[5 * P] = L.R.w1; q[6 * P] = L.R.w2; q[7 * P] = L.R.w3;
    q[8 * P] = __float_as_uint(L.W.rx); q[9 * P] = __float_as_uint(L.W.ry); q[10 * P] = __float_as_uint(L.W.rz);
    q[11 * P] = __float_as_uint(L.W.vx); q[12 * P] = __float_as_uint(L.W.vy); q[13 * P] = __float_as_uint(L.W.vz);
    q[14 * P] = (uint32_t)L.W.xP | ((uint32_t)L.W.yP << 11) | ((uint32_t)L.W.zP << 22);
    q[15 * P] = __float_as_uint(L.W.absTau); q[16 * P] = __float_as_uint(L.W.passProb);
    q[17 * P] = __float_as_uint(L.d0x); q[18 * P] = __float_as_uint(L.d0y); q[19 * P] = __float_as_uint(L.d0z);
    q[20 * P] = (uint32_t)L.nuP | ((uint32_t)L.runs << 16) | ((uint32_t)L.phase << 30);
    q[21 * P] = (uint32_t)L.origin; q[22 * P] = (uint32_t)L.steps;
}

template <int P>
__device__ __forceinline__ void pool_get(const uint32_t *pw, int s, Lane &L, uint32_t stream)
{
    const uint32_t *q = pw + s;
    L.R.plo = q[0 * P]; L.R.phi = q[1 * P]; L.R.k = q[2 * P]; L.R.blk = q[3 * P];
    L.R.w0 = q[4 * P]; L.R.w1 = q[5 * P]; L.R.w2 = q[6 * P]; L.R.w3 = q[7 * P];
    L.R.stream = stream;
    L.W.rx = __uint_as_float(q[8 * P]); L.W.ry = __uint_as_float(q[9 * P]); L.W.rz = __uint_as_float(q[10 * P]);
    L.W.vx = __uint_as_float(q[11 * P]); L.W.vy = __uint_as_float(q[12 * P]); L.W.vz = __uint_as_float(q[13 * P]);
    const uint32_t c = q[14 * P];
    L.W.xP = (int)(c & 2047u); L.W.yP = (int)((c >> 11) & 2047u); L.W.zP = (int)(c >> 22);
    L.W.absTau = __uint_as_float(q[15 * P]); L.W.passProb = __uint_as_float(q[16 * P]);
    L.d0x = __uint_as_float(q[17 * P]); L.d0y = __uint_as_float(q[18 * P]); L.d0z = __uint_as_float(q[19 * P]);
    const uint32_t f = q[20 * P];
    L.nuP = (int)(f & 65535u); L.runs = (int)((f >> 16) & 16383u); L.phase = (int)(f >> 30);
    L.origin = (int)q[21 * P]; L.steps = (int)q[22 * P];
}

// lanes below this one in m
__device__ __forceinline__ int lane_rank(uint64_t m, int lane)
{
    return __popcll(m & ((1ull << lane) - 1ull));
}

// Orders the wavefront's LDS accesses on either side (the compiler may not
// move them across; the LDS serves one wavefront's instructions in order).
__device__ __forceinline__ void wave_lds_order()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Lanes that do not walk take walkers from the pool (the first nw of them in
// lane order); a taker whose packet waits for an emission leaves it on the
// pending stack first, if the pool has a free slot for it.  The pending
// slots written lie in the free region, the walker slots read below it.
template <int P>
__device__ __forceinline__ void pool_settle(Lane &L, uint32_t *pw, int &nw, int &npd, int lane, uint32_t stream)
{
    if (nw == 0) return;
    const bool nonw = L.phase != PH_WALK;
    if (!__any(nonw)) return;
    const int F = P - nw - npd;
    const bool pend = pending(L);
    const uint64_t mp = __ballot(pend);
    const bool elig = nonw && (!pend || lane_rank(mp, lane) < F);
    const uint64_t me = __ballot(elig);
    const int re = lane_rank(me, lane);
    const bool take = elig && re < nw;
    const uint64_t mt = __ballot(take);
    const uint64_t mtp = mt & mp;
    if (mtp) {
        if (take && pend) pool_put<P>(pw, P - npd - 1 - lane_rank(mtp, lane), L);
        npd += __popcll(mtp);
        wave_lds_order();
    }
    if (take) pool_get<P>(pw, nw - 1 - re, L, stream);
    nw -= __popcll(mt);
    wave_lds_order();
}

// An emission round of the pool kernel: walkers move to the pool (as many
// as the candidates for an emission exceed the lanes that do not walk, and
// as fit), idle lanes take packets from the pending stack, then fresh
// packets (one counter atomic per wavefront), and the emission pass runs.
template <int P>
__device__ __forceinline__ void pool_emission_round(Lane &L, uint32_t *pw, int &nw, int &npd, bool &exhausted,
                                                    const Grid &G, const RunArgs &RA, const Model &M,
                                                    const float *spec, const unsigned short *spec_guide,
                                                    WaveCounts &C, double &qphot, int &err, int lane)
{
    const bool walk = L.phase == PH_WALK;
    const uint64_t mw = __ballot(walk);
    const int nwalk = __popcll(mw);
    const int npl = (int)wcount(pending(L));
    const int cand = npl + npd + (exhausted ? 0 : 64);
    const int k = max(0, min(min(nwalk, P - nw - npd), cand - (64 - nwalk)));
    if (k > 0) {
        const int rw = lane_rank(mw, lane);
        if (walk && rw < k) {
            pool_put<P>(pw, nw + rw, L);
            L.phase = PH_IDLE;
        }
        nw += k;
        wave_lds_order();
    }
    if (npd > 0) {
        const bool idle = L.phase == PH_IDLE;
        const uint64_t mi = __ballot(idle);
        const int ri = lane_rank(mi, lane);
        if (idle && ri < npd) pool_get<P>(pw, P - npd + ri, L, (uint32_t)RA.istar);
        npd -= min(__popcll(mi), npd);
        wave_lds_order();
    }
    const bool need = (L.phase == PH_IDLE) && !exhausted;
    const uint64_t m = __ballot(need);
    if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(RA.next_packet, (unsigned long long)__popcll(m));
        base = __shfl(base, leader);
        bool over = false;
        if (need) {
            const unsigned long long id = base + (unsigned long long)lane_rank(m, lane);
            if (id < (unsigned long long)RA.npackets) {
                rng_start(L.R, RA.packet_offset + id, (uint32_t)RA.istar);
                L.phase = PH_EMIT_STELLAR;
            } else {
                over = true;
            }
        }
        exhausted = __any(over);
    }
    if (__any(pending(L))) emission_pass(L, G, RA, M, spec, spec_guide, C, qphot, err);
}

// Add a wavefront's counters to its model's totals and clear them.
__device__ __forceinline__ void flush_counts(const RunArgs &RA, WaveCounts &C, double &qphot)
{
    for (int o = 32; o >= 1; o >>= 1) qphot += __shfl_xor(qphot, o);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&RA.counts[0], C.seg);
        atomicAdd(&RA.counts[1], C.abs);
        atomicAdd(&RA.counts[2], C.sca);
        atomicAdd(&RA.counts[3], C.esc);
        atomicAdd(&RA.counts[4], C.lost);
        atomicAdd(&RA.counts[5], C.line);
        atomicAdd(&RA.counts[6], C.trapped);
        atomicAdd(&RA.counts[7], C.flights);
        atomicAdd(&RA.counts[8], C.wpass);
        atomicAdd(&RA.counts[9], C.wlanes);
        atomicAdd(&RA.counts[10], C.epass);
        atomicAdd(&RA.counts[11], C.elanes);
        unsafeAtomicAdd(RA.qphot, qphot);
    }
    C = WaveCounts{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    qphot = 0.0;
}

// The single-model kernel.  TWO: two packets per lane (advance_spare; 144
// VGPRs, so at most 3 wavefronts per SIMD), else one (advance; 4 per SIMD).
template <bool TWO>
__global__ void __launch_bounds__(256) transport_kernel(Grid G, RunArgs RA)
{
    extern __shared__ float lds[];
    Axes A;
    float *spec = nullptr;
    unsigned short *spec_guide = nullptr;
    const int nwaves = (int)(blockDim.x >> 6);
    stage_model(G, RA, lds + (size_t)nwaves * JRing<kJDepth>::kFloats, A, spec, spec_guide);
    const int lane_id = threadIdx.x & 63;
    JQ Q = jq_at<kJDepth>(lds, (int)(threadIdx.x >> 6));
    const Model M = model_of(G, RA);

    Lane L, S;
    L.phase = PH_IDLE;
    L.runs = 0; L.steps = 0; L.nuP = 1; L.origin = 0;
    S = L;
    bool exhausted = false;
    WaveCounts C = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double qphot = 0.0;
    int err = 0;

    for (;;) {
        // ---- refill idle lanes (their spare with two packets per lane): one
        // atomic per wavefront
        Lane &N = TWO ? S : L;
        bool need = (N.phase == PH_IDLE) && !exhausted;
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
                    rng_start(N.R, RA.packet_offset + id, (uint32_t)RA.istar);
                    N.phase = PH_EMIT_STELLAR;
                } else {
                    exhausted = true;
                }
            }
        }
        if (TWO) {
            if (wcount(L.phase != PH_IDLE || S.phase != PH_IDLE) == 0) break;
            advance_spare<kJDepth>(L, S, G, A, RA, M, spec, spec_guide, RA.emit_min, RA.walk_steps, C, qphot, err,
                                   Q, lane_id);
        } else {
            const int npend = (int)wcount(pending(L));
            const int nwalk = (int)wcount(L.phase == PH_WALK);
            if (npend + nwalk == 0) break;
            advance<kJDepth>(L, G, A, RA, M, spec, spec_guide, RA.emit_min, RA.walk_steps, npend, nwalk, C, qphot,
                             err, Q, lane_id);
        }
    }
    jq_flush<kJDepth>(G, Q, lane_id);
    flush_counts(RA, C, qphot);
}

// The single-model kernel with a packet pool of P slots per wavefront in
// LDS (one packet per lane in registers, J rings of D updates).  An
// emission round runs when at least emit_min lanes do not walk (or none
// does); every walk pass is followed by a settle from the pool.
template <int D, int P>
__global__ void __launch_bounds__(256) pool_kernel(Grid G, RunArgs RA)
{
    extern __shared__ float lds[];
    Axes A;
    float *spec = nullptr;
    unsigned short *spec_guide = nullptr;
    const int nwaves = (int)(blockDim.x >> 6);
    const int wave = (int)(threadIdx.x >> 6);
    uint32_t *pools = reinterpret_cast<uint32_t *>(lds + (size_t)nwaves * JRing<D>::kFloats);
    stage_model(G, RA, reinterpret_cast<float *>(pools + (size_t)nwaves * P * kPoolWords), A, spec, spec_guide);
    uint32_t *pw = pools + (size_t)wave * P * kPoolWords;
    const int lane_id = threadIdx.x & 63;
    JQ Q = jq_at<D>(lds, wave);
    const Model M = model_of(G, RA);
    const uint32_t stream = (uint32_t)RA.istar;

    Lane L;
    L.phase = PH_IDLE;
    L.runs = 0; L.steps = 0; L.nuP = 1; L.origin = 0;
    int nw = 0, npd = 0;
    bool exhausted = false;
    WaveCounts C = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double qphot = 0.0;
    int err = 0;

    for (;;) {
        pool_settle<P>(L, pw, nw, npd, lane_id, stream);
        const int nwalk = (int)wcount(L.phase == PH_WALK);
        if (nwalk == 0 && nw == 0 && npd == 0 && exhausted && !__any(pending(L))) break;
        if (64 - nwalk >= RA.emit_min || nwalk == 0)
            pool_emission_round<P>(L, pw, nw, npd, exhausted, G, RA, M, spec, spec_guide, C, qphot, err, lane_id);
        for (int s = 0; s < RA.walk_steps; ++s) {
            if (!__any(L.phase == PH_WALK)) break;
            walk_pass<D>(L, G, A, M, C, err, Q, lane_id);
            pool_settle<P>(L, pw, nw, npd, lane_id, stream);
        }
        if (err) {
            atomicMax(RA.error, err);
            err = 0;
        }
    }
    jq_flush<D>(G, Q, lane_id);
    flush_counts(RA, C, qphot);
}

// Many models in one persistent launch.  The packets of every model are cut
// into chunks, listed model after model; a one-wavefront workgroup takes
// chunks from a global counter, refills its lanes from the chunk it holds,
// and takes the next chunk when that one is used up.  A chunk of another
// model is taken only after the wavefront's packets of the current model
// have finished: then its grid and source are staged in LDS and its
// counters flushed.  Each packet keeps its own random stream, so results
// equal those of the models' separate runs.  Compiled for 4 wavefronts per
// SIMD (128 VGPRs, a few spilled to scratch): the one-wavefront workgroups
// need the occupancy more than the registers.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) batch_kernel(BatchArgs B)
{
    extern __shared__ float lds[];
    const int lane_id = threadIdx.x;
    JQ Q = jq_at<kJDepthBatch>(lds, 0);
    float *tab = lds + JRing<kJDepthBatch>::kFloats;
    Axes A;
    float *spec = nullptr;
    unsigned short *spec_guide = nullptr;
    // the current model's grid and arguments are read in place (wave-uniform
    // addresses: scalar loads), not copied into registers
    const Grid *Gp = B.G;
    const RunArgs *RAp = B.RA;
    Model M;
    int cur = -1;                       // model of the staged tables
    long long cnext = 0, cend = 0;      // packets left in the held chunk (model cur)
    int pending = -1;                   // chunk of another model, taken after draining
    bool done = false;                  // no chunks left

    Lane L;
    L.phase = PH_IDLE;
    L.runs = 0; L.steps = 0; L.nuP = 1; L.origin = 0;
    WaveCounts C = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double qphot = 0.0;
    int err = 0;

    for (;;) {
        // ---- refill idle lanes from the held chunk, taking chunks of the same model
        uint64_t m = __ballot(L.phase == PH_IDLE);
        while (m && pending < 0 && !done) {
            if (cnext >= cend) {
                int c = 0;
                if (lane_id == 0) c = (int)atomicAdd(B.next_chunk, 1u);
                c = __builtin_amdgcn_readfirstlane(c);     // wave-uniform: scalar registers
                if (c >= B.nchunks) { done = true; break; }
                if (B.chunk_model[c] != cur) { pending = c; break; }
                cnext = B.chunk_first[c];
                cend = cnext + B.chunk_count[c];
                continue;
            }
            const long long avail = cend - cnext;
            const unsigned long long rank = __popcll(m & ((1ull << lane_id) - 1ull));
            const bool mine = ((m >> lane_id) & 1ull) && (long long)rank < avail;
            if (mine) {
                rng_start(L.R, RAp->packet_offset + (unsigned long long)(cnext + (long long)rank),
                          (uint32_t)RAp->istar);
                L.phase = PH_EMIT_STELLAR;
            }
            const uint64_t took = __ballot(mine);
            cnext += __popcll(took);
            m &= ~took;
        }
        const bool pend = (L.phase == PH_EMIT_STELLAR || L.phase == PH_EMIT_DIFFUSE);
        const int npend = (int)wcount(pend);
        const int nwalk = (int)wcount(L.phase == PH_WALK);
        if (npend + nwalk == 0) {
            if (pending >= 0) {
                // all lanes idle: switch to the pending chunk's model
                if (cur >= 0) {
                    jq_flush<kJDepthBatch>(*Gp, Q, lane_id);
                    flush_counts(*RAp, C, qphot);
                }
                cur = __builtin_amdgcn_readfirstlane(B.chunk_model[pending]);
                Gp = B.G + cur;
                RAp = B.RA + cur;
                __syncthreads();
                stage_model(*Gp, *RAp, tab, A, spec, spec_guide);
                M = model_of(*Gp, *RAp);
                cnext = B.chunk_first[pending];
                cend = cnext + B.chunk_count[pending];
                pending = -1;
                continue;
            }
            if (done) break;
            continue;
        }
        advance<kJDepthBatch>(L, *Gp, A, *RAp, M, spec, spec_guide, B.emit_min, B.walk_steps, npend, nwalk, C, qphot, err, Q,
                lane_id);
    }
    if (cur >= 0) {
        jq_flush<kJDepthBatch>(*Gp, Q, lane_id);
        flush_counts(*RAp, C, qphot);
    }
}

// ------------------------------------------------------- table preparation
// opac3 in the J3 layout: the value of lattice point p at bin nu is
// opacity(active(xyz(p)), nu) (reference layout in: [nu][cell]); padding
// points hold 0.  Written in layout order (coalesced stores).
__global__ void densify_kernel(const float *compact, const int *active, J3Shape js, int nc1, int nbins,
                               float *dense)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tot = j3_size(js, nbins);
    for (; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t tile = i / ((size_t)nbins * kJTile);
        const int nu = (int)((i / kJTile) % (size_t)nbins);
        const int xyz = j3_lattice(js, (unsigned)(tile * kJTile + i % kJTile));
        dense[i] = xyz < 0 ? 0.f : compact[(size_t)nu * nc1 + active[xyz]];
    }
}

// guide[row][j] for every row of a [rows][nbins] CDF table
__global__ void guide_kernel(const float *cdf, int rows, int nbins, unsigned short *guide)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t tot = (size_t)rows * (kGuide + 1);
    for (; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        int r = (int)(i / (kGuide + 1)), j = (int)(i % (kGuide + 1));
        guide[i] = guide_entry(cdf + (size_t)r * nbins, nbins, j);
    }
}

// J(cell,nu) = value at the (x,y,z) of that cell; row 0 collects every
// inactive position, as the reference's Jste(0,:) does.  out must be zero.
__global__ void compact_kernel(const double *J3, const int *active, J3Shape js, int nc1, int nbins, double *out)
{
    const int ncell3 = js.nx * js.ny * js.nz;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t tot = (size_t)ncell3 * nbins;
    for (; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        int nu = (int)(i / ncell3), xyz = (int)(i % ncell3);
        double v = J3[j3_index(nbins, nu, j3_point_of(js, xyz))];
        int a = active[xyz];
        if (a > 0) out[(size_t)nu * nc1 + a] = v;
        else if (v != 0.0) unsafeAtomicAdd(&out[(size_t)nu * nc1], v);
    }
}

}  // namespace mcd

// ---------------------------------------------------------------- launchers
hipError_t mcd_launch_batch(const mcd::BatchArgs &B, size_t lds_bytes, int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(mcd::batch_kernel, dim3(blocks), dim3(64), lds_bytes, s, B);
    return hipGetLastError();
}

// dynamic LDS of a batch wavefront: its J queue, then the model's tables
size_t mcd_lds_bytes(int nx, int ny, int nz, int nbins)
{
    return sizeof(float) * (mcd::JRing<mcd::kJDepthBatch>::kFloats + mcd::lds_floats(nx, ny, nz, nbins));
}

template <int D, int P>
static size_t pool_lds(const mcd::Grid &G, int threads)
{
    return sizeof(float) * ((size_t)(threads / 64) * (mcd::JRing<D>::kFloats + (size_t)P * mcd::kPoolWords) +
                            mcd::lds_floats(G.nx, G.ny, G.nz, G.nbins));
}

hipError_t mcd_launch_transport(const mcd::Grid &G, const mcd::RunArgs &RA, int blocks, int threads, int kernel,
                                hipStream_t s)
{
    const size_t lds = sizeof(float) * ((size_t)(threads / 64) * mcd::JRing<mcd::kJDepth>::kFloats +
                                        mcd::lds_floats(G.nx, G.ny, G.nz, G.nbins));
    switch (kernel) {
    case MCD_KERNEL_SPARE:
        hipLaunchKernelGGL(mcd::transport_kernel<true>, dim3(blocks), dim3(threads), lds, s, G, RA);
        break;
    case MCD_KERNEL_POOL: {
        const size_t pl = pool_lds<mcd::kJDepth, mcd::kPoolSlots>(G, threads);
        hipLaunchKernelGGL((mcd::pool_kernel<mcd::kJDepth, mcd::kPoolSlots>), dim3(blocks), dim3(threads), pl, s,
                           G, RA);
        break;
    }
    default:
        hipLaunchKernelGGL(mcd::transport_kernel<false>, dim3(blocks), dim3(threads), lds, s, G, RA);
    }
    return hipGetLastError();
}

hipError_t mcd_launch_replay(const mcd::Grid &G, int n, const float *pos, const float *dir,
                             const int *cell, int *cell_seq, float *ds_seq, int *seq_len,
                             int *status, int max_len, hipStream_t s)
{
    size_t lds = sizeof(float) * mcd::lds_floats(G.nx, G.ny, G.nz, 0);
    int threads = 256, blocks = (n + threads - 1) / threads;
    hipLaunchKernelGGL(mcd::walk_replay_kernel, dim3(blocks), dim3(threads), lds, s, G, n, pos,
                       dir, cell, cell_seq, ds_seq, seq_len, status, max_len);
    return hipGetLastError();
}

hipError_t mcd_densify(const float *compact, const int *active, const mcd::J3Shape &js, int nc1, int nbins,
                       float *dense, hipStream_t s)
{
    hipLaunchKernelGGL(mcd::densify_kernel, dim3(2048), dim3(256), 0, s, compact, active, js, nc1, nbins, dense);
    return hipGetLastError();
}

hipError_t mcd_guides(const float *cdf, int rows, int nbins, unsigned short *guide, hipStream_t s)
{
    hipLaunchKernelGGL(mcd::guide_kernel, dim3(2048), dim3(256), 0, s, cdf, rows, nbins, guide);
    return hipGetLastError();
}

hipError_t mcd_compact(const double *J3, const int *active, const mcd::J3Shape &js, int nc1, int nbins,
                       double *out, hipStream_t s)
{
    hipLaunchKernelGGL(mcd::compact_kernel, dim3(2048), dim3(256), 0, s, J3, active, js, nc1, nbins, out);
    return hipGetLastError();
}

// mc_device.h - device-side view of one transport problem and the pieces of
// MOCASSIN's packet transport that every kernel shares: the Philox stream,
// the source-frequency draw (getNu2), the isotropic and Henyey-Greenstein
// direction draws and one step of the Cartesian cell walk (pathSegment).
//
// FP32 throughout with the reference's expression order; compiled with
// contraction off, correctly rounded division and square root and IEEE
// denormals, so a packet history is bit-identical to the FP32 statement of
// the reference given the same uniform stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mc_fmath.h"

#pragma clang fp contract(off)

namespace mcd {

constexpr int kSafeLimit = 500000;      // photon_mod.f90:1187-1192
constexpr int kRecursionLimit = 5000;   // constants_mod.f90:56

enum Flags { F_SYM = 1, F_DUST = 2, F_GAS = 4, F_ISO = 8, F_MULTI = 16 };

// Per-problem constants and device tables.
// Viewing-angle bins of the escapes as initCartesianGrid sets them
// (grid_mod.f90:416-467): theta and phi bin widths and the maps from a bin
// to a viewpoint (0: none); the viewpoints' angles (0:nAngle).
constexpr int kViewMax = 64;
struct View {
    int nAngle, totT, totP;
    float dTheta, dPhi;
    int Pt[kViewMax + 1], Pp[kViewMax + 1];
    float vth[3], vph[3];
};

// J3 layout.  The lattice is cut into tiles of 4x4x4 points (each axis
// padded to a multiple of 4) and a tile holds [nu][64 points].  Inside a
// tile the points run [2x2x2 brick][2x2x2 point], so the 8 doubles of one
// 64-byte line are a 2x2x2 brick at one frequency.  Atomic adds of one
// instruction that fall in one line travel to memory as one request, and a
// packet's next segment (same frequency, a neighbouring cell) stays in the
// brick after half of its steps, against 7/24 for 8 points along x.  One
// wavefront reading a tile's spectra (64 lanes, one point each) streams
// contiguous memory.  The walk's opacity tables use the same layout in
// FP32: a 64-byte line is a 4x2x2 block of points at one frequency, so a
// packet's next opacity gather falls in the line of its last one after
// more than half of its steps (a third with the lattice x-major).
constexpr int kJTile = 64;
struct J3Shape {
    int nx, ny, nz;          // lattice
    int tx, ty, tz;          // tiles per axis
};
__host__ __device__ __forceinline__ J3Shape j3_shape(int nx, int ny, int nz)
{
    return J3Shape{ nx, ny, nz, (nx + 3) / 4, (ny + 3) / 4, (nz + 3) / 4 };
}
__host__ __device__ __forceinline__ size_t j3_points(const J3Shape &s)
{
    return (size_t)s.tx * (size_t)s.ty * (size_t)s.tz * kJTile;
}
__host__ __device__ __forceinline__ size_t j3_size(const J3Shape &s, int nb)
{
    return j3_points(s) * (size_t)nb;
}
// the point of lattice cell (x, y, z), 0-based
__host__ __device__ __forceinline__ unsigned j3_point(const J3Shape &s, int x, int y, int z)
{
    const unsigned tile = ((unsigned)(z >> 2) * (unsigned)s.ty + (unsigned)(y >> 2)) * (unsigned)s.tx +
                          (unsigned)(x >> 2);
    const unsigned slot = ((unsigned)(z & 2) << 4) | ((unsigned)(y & 2) << 3) | ((unsigned)(x & 2) << 2) |
                          ((unsigned)(z & 1) << 2) | ((unsigned)(y & 1) << 1) | (unsigned)(x & 1);
    return tile * kJTile + slot;
}
// the point of dense lattice index xyz = x + nx*(y + ny*z)
__host__ __device__ __forceinline__ unsigned j3_point_of(const J3Shape &s, int xyz)
{
    return j3_point(s, xyz % s.nx, (xyz / s.nx) % s.ny, xyz / (s.nx * s.ny));
}
// the dense lattice index of point p, or -1 for a point of the padding
__host__ __device__ __forceinline__ int j3_lattice(const J3Shape &s, unsigned p)
{
    const unsigned tile = p / kJTile, slot = p % kJTile;
    const int x = (int)(tile % (unsigned)s.tx) * 4 + (int)((slot >> 2) & 2) + (int)(slot & 1);
    const int y = (int)((tile / (unsigned)s.tx) % (unsigned)s.ty) * 4 + (int)((slot >> 3) & 2) +
                  (int)((slot >> 1) & 1);
    const int z = (int)(tile / ((unsigned)s.tx * (unsigned)s.ty)) * 4 + (int)((slot >> 4) & 2) +
                  (int)((slot >> 2) & 1);
    if (x >= s.nx || y >= s.ny || z >= s.nz) return -1;
    return x + s.nx * (y + s.ny * z);
}
// the word of bin nu0 (0-based) of point p
__host__ __device__ __forceinline__ size_t j3_index(int nb, int nu0, unsigned p)
{
    return ((size_t)(p / kJTile) * (unsigned)nb + (unsigned)nu0) * kJTile + (size_t)(p % kJTile);
}

struct Grid {
    int nx, ny, nz, nCells, nbins, flags, nStars;
    float R_out, geoCorrX, geoCorrY, geoCorrZ, ionEdge1;
    const float *axis;       // x(1:nx), y(1:ny), z(1:nz) concatenated
    const float *mid;        // (a(i)+a(i+1))/2, i=1..n-1 per axis, concatenated
    const float *dfac;       // per-axis cell widths /1e15 (getVolume factors)
    int ncell3;              // nx*ny*nz
    const int *active;       // active(x,y,z), x fastest
    const float *opac3;      // J3 layout, fp32: opacity(active(x,y,z),nu); inactive points 0
    const float *sca3;       // J3 layout, fp32: scaOpac, or null
    const float *recCDF;     // [cell][nu]   (one row per cell)
    const unsigned short *recGuide;  // [cell][kGuide+1] guide table of recCDF rows
    const float *totalLines; // [cell]
    const float *specCDF;    // [star][nu]
    const float *nuArray;    // [nu]
    const float *gSca;       // [nu] or null
    double *J3;              // FP64 path-length estimator, tiled: see j3_index
    J3Shape js;              // its tiling
    double *esc;             // [plane 0..nAngle][nu 0..nbins][cell]
    const View *view;        // viewing-angle bins, or null (plane 0 only)
    unsigned *cross;         // [J3 point] segments walked in each cell, or null
    uint32_t seed_lo, seed_hi;
};

// Guide table (Chen & Asau 1974, "On generating random variates from an
// empirical distribution"): g[j] = #{i : cdf[i] <= j/K}.  For u in
// [j/K, (j+1)/K) the count of entries <= u lies in [g[j], g[j+1]], so a
// search restricted to that range returns exactly the full binary search.
constexpr int kGuide = 32;

// LDS copy of the small per-axis arrays (axis, midpoints, volume factors).
struct Axes {
    const float *x, *y, *z;      // 1-based through the accessors below
    const float *mx, *my, *mz;
    const float *dx, *dy, *dz;
};

__device__ __forceinline__ float AX(const float *a, int i) { return a[i - 1]; }

// --------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon et al., SC'11).  Counter = (draw/4, packet lo,
// packet hi, stream), key = seed; uniform = (word >> 8) * 2^-24 in [0,1).
struct Rng {
    uint32_t plo, phi, stream, k, blk;
    uint32_t w0, w1, w2, w3;
};

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, uint32_t &o0, uint32_t &o1,
                                       uint32_t &o2, uint32_t &o3)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

__device__ __forceinline__ void rng_start(Rng &R, uint64_t packet, uint32_t stream)
{
    R.plo = (uint32_t)packet; R.phi = (uint32_t)(packet >> 32);
    R.stream = stream; R.k = 0; R.blk = 0xffffffffu;
}

__device__ __forceinline__ float draw(Rng &R, const Grid &G)
{
    uint32_t b = R.k >> 2;
    if (b != R.blk) {
        philox(b, R.plo, R.phi, R.stream, G.seed_lo, G.seed_hi, R.w0, R.w1, R.w2, R.w3);
        R.blk = b;
    }
    uint32_t s = R.k & 3u;
    uint32_t w = s == 0 ? R.w0 : (s == 1 ? R.w1 : (s == 2 ? R.w2 : R.w3));
    R.k++;
    return (float)(w >> 8) * 0x1p-24f;
}

// --------------------------------------------------------------- sampling
// getNu2 (photon_mod.f90:720-764): k = number of leading CDF entries <= u,
// nuP = max(k,1), then +1 unless nuP >= nbins-1.  The linear scan becomes a
// binary search over the (validated monotone) CDF row; identical result.
__device__ __forceinline__ int getnu2(Rng &R, const Grid &G, const float *cdf,
                                      const unsigned short *guide, int nbins, int &err)
{
    float u = draw(R, G);
    int i;
    for (i = 1; i <= 10000; ++i) {
        if (u == 0.f || u == 1.f || u == 0.9999999f) u = draw(R, G);
        else break;
    }
    if (i >= 10000) { err = 10; return -1; }
    const int j = (int)(u * (float)kGuide);          // exact: u has 24 bits, K = 2^5
    int lo = guide[j], hi = guide[j + 1];            // count of entries <= u in [lo, hi]
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (u >= cdf[m]) lo = m + 1; else hi = m;
    }
    int nuP = lo < 1 ? 1 : lo;
    if (nuP < nbins - 1) nuP = nuP + 1;
    return nuP;
}

// g[j] = #{i : cdf[i] <= j/kGuide}, j = 0..kGuide, for a monotone row
__device__ __forceinline__ unsigned short guide_entry(const float *cdf, int nbins, int j)
{
    const float t = (float)j / (float)kGuide;
    int lo = 0, hi = nbins;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (t >= cdf[m]) lo = m + 1; else hi = m;
    }
    return (unsigned short)lo;
}

// randomUnitVector (vector_mod.f90:303-314)
__device__ __forceinline__ void random_unit_vector(Rng &R, const Grid &G, float &vx, float &vy, float &vz)
{
    float r1 = draw(R, G);
    float w = 2.f * r1 - 1.f;
    float t = sqrtf(1.f - w * w);
    float r2 = draw(R, G);
    float ang = 3.141592654f * (2.f * r2 - 1.f);
    float s, c;
    mcf_sincosf(ang, &s, &c);
    vx = t * c; vy = t * s; vz = w;
}

// direction loop of initPhotonPacket (photon_mod.f90:650-675)
__device__ __forceinline__ void new_direction(Rng &R, const Grid &G, bool fold,
                                              float &vx, float &vy, float &vz)
{
    for (int irep = 1; irep <= 1000000; ++irep) {
        random_unit_vector(R, G, vx, vy, vz);
        if (vx != 0.f && vy != 0.f && vz != 0.f) break;
    }
    if (fold) {
        if (vx < 0.f) vx = -vx;
        if (vy < 0.f) vy = -vy;
        if (vz < 0.f) vz = -vz;
    }
}

// hg (photon_mod.f90:2875-2972); returns ierr, updates d* only on success
__device__ __forceinline__ int hg(Rng &R, const Grid &G, int nuP, float &dx, float &dy, float &dz)
{
    float v0 = dx, v1 = dy, v2 = dz;
    float hgg = G.gSca[nuP - 1];
    float random0 = draw(R, G);
    float s = 2.f * random0 - 1.f;
    float cost, sint;
    if (hgg >= 0.0001f) {
        float q = (1.f - hgg * hgg) / (1.f + hgg * s);
        cost = 0.5f / hgg * ((1.f + hgg * hgg) - q * q);
    } else {
        cost = s;
    }
    if (cost >= 1.0f) { cost = 1.0f; sint = 0.f; }
    else if (cost < -1.0f) { cost = -1.0f; sint = 0.f; }
    else sint = sqrtf(1.f - cost * cost);
    float random = draw(R, G);
    float phi = 6.28318548f * random;
    float sinp, cosp;
    mcf_sincosf(phi, &sinp, &cosp);
    float denom = sqrtf(1.f - v2 * v2);
    float o0, o1, o2;
    if (denom > 0.001f) {
        o0 = sint / denom * (v0 * v2 * cosp - v1 * sinp) + v0 * cost;
        o1 = sint / denom * (v1 * v2 * cosp + v0 * sinp) + v1 * cost;
        o2 = -(sint * cosp * denom) + v2 * cost;
    } else {
        o0 = sint * cosp;
        o1 = sint * sinp;
        o2 = (v2 >= 0.f) ? cost : -cost;
    }
    bool finite = !isnan(o0) && !isnan(o1) && !isnan(o2);
    if (fabsf(o0) <= 1.f && fabsf(o1) <= 1.f && fabsf(o2) <= 1.f && finite) {
        dx = o0; dy = o1; dz = o2;
        return 0;
    }
    // out-of-range components: the reference normalises a copy it discards
    if (fabsf(o0) >= 1.f || fabsf(o1) >= 1.f || (fabsf(o2) >= 1.f && finite)) return 0;
    return 1;
}

// --------------------------------------------------------------- walk step
enum StepResult { ST_CONTINUE = 0, ST_ESCAPED = 1, ST_LOST = 2, ST_ABSORBED = 3, ST_ERROR = 4 };

struct Walker {
    float rx, ry, rz;      // rVec
    float vx, vy, vz;      // vHat
    int xP, yP, zP;
    float absTau, passProb;
    // outputs of the last step
    float seg;             // dS or dlLoc of the segment just walked
    unsigned segPt;        // J3 lattice point of its cell (j3_point)
    int segX, segY, segZ;  // its cell indices (1-based)
};

__device__ __forceinline__ int act(const Grid &G, int x, int y, int z)
{
    return G.active[(x - 1) + G.nx * ((y - 1) + G.ny * (z - 1))];
}

// Distance to the next wall along one axis (photon_mod.f90:1263-1298 for
// x; y and z alike).  Moving up (v > 1e-10) the wall is the midpoint above
// cell P, or the outer edge ax(n) in the last cell; moving down it is the
// midpoint below, or ax(1) in the first cell; otherwise 1e35.  A packet
// within 1e-10 of an inner wall steps over it (r = wall, P +- 1); at the
// outer edge moving up it is lost (lost = true, r = the edge); at ax(1)
// moving down it is put on the edge.  1 <= P <= n on entry.
__device__ __forceinline__ float wall_distance(const float *ax, const float *mid, int n, float v, float &r,
                                               int &P, bool &lost)
{
    const bool pos = v > 1.e-10f, neg = v < -1.e-10f;
    const bool inner = pos ? (P < n) : (P > 1);
    const float m = inner ? mid[pos ? P - 1 : P - 2] : ax[pos ? n - 1 : 0];
    const float d = (m - r) / v;
    lost = false;
    if (!(pos || neg)) return 1.e35f;
    if (fabsf(d) < 1.e-10f) {
        if (inner) { r = m; P += pos ? 1 : -1; }
        else if (pos) { r = m; lost = true; }
        else r = m;
    }
    return d;
}

// One iteration of pathSegment's "do i" loop (photon_mod.f90:1194-2836) on
// the mother grid.  ZERO_OPACITY replays geometry only.  kappa is looked up
// here; on ST_ABSORBED the interaction point is in rVec and (xP,yP,zP).
// err receives the fail-stop code of the reference's "insanity" stops.
template <bool ZERO_OPACITY>
__device__ __forceinline__ int walk_step(const Grid &G, const Axes &A, Walker &W, int nuP,
                                         float &kappa_out, float &dV_out, int &err)
{
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const bool sym = (G.flags & F_SYM) != 0;
    float dSx, dSy, dSz, dS;
    int xP = W.xP, yP = W.yP, zP = W.zP;
    float rx = W.rx, ry = W.ry, rz = W.rz, vx = W.vx, vy = W.vy, vz = W.vz;

    if (xP > nx || xP < 1 || yP > ny || yP < 1 || zP > nz || zP < 1) { err = 32; return ST_ERROR; }
    // the cell's opacity, issued first so that its gather overlaps the wall
    // distances (the indices were just checked); a packet on a wall is
    // stepped into the next cell there, and then the gather is repeated
    const unsigned p0 = j3_point(G.js, xP - 1, yP - 1, zP - 1);
    float kappa = 0.f;
    if (!ZERO_OPACITY) kappa = G.opac3[j3_index(G.nbins, nuP - 1, p0)];

    if (sym) {                                                   // 1248-1261
        if (rx <= AX(A.x, 1)) { if (vx < 0.f) vx = -vx; rx = AX(A.x, 1); }
        if (ry <= AX(A.y, 1)) { if (vy < 0.f) vy = -vy; ry = AX(A.y, 1); }
        if (rz <= AX(A.z, 1)) { if (vz < 0.f) vz = -vz; rz = AX(A.z, 1); }
    }

    // wall distances, 1263-1379: one division per axis whatever the signs
    // of the wavefront's directions (the reference's branches select the
    // wall; they are evaluated here as selects, in the same order)
    bool lostX, lostY, lostZ;
    dSx = wall_distance(A.x, A.mx, nx, vx, rx, xP, lostX);
    dSy = wall_distance(A.y, A.my, ny, vy, ry, yP, lostY);
    dSz = wall_distance(A.z, A.mz, nz, vz, rz, zP, lostZ);
    if (lostX) { W.rx = rx; return ST_LOST; }
    if (isnan(dSx)) { err = 34; return ST_ERROR; }
    if (lostY) { W.ry = ry; return ST_LOST; }
    if (isnan(dSy)) { err = 34; return ST_ERROR; }
    if (lostZ) { W.rz = rz; return ST_LOST; }
    if (isnan(dSz)) { err = 34; return ST_ERROR; }

    // cells on the cell wall, 1382-1384
    if (fabsf(dSx) < 1.e-10f) dSx = AX(A.x, nx);
    if (fabsf(dSy) < 1.e-10f) dSy = AX(A.y, ny);
    if (fabsf(dSz) < 1.e-10f) dSz = AX(A.z, nz);
    dSx = fabsf(dSx); dSy = fabsf(dSy); dSz = fabsf(dSz);
    if (dSx <= 0.f) dS = fminf(dSy, dSz);
    else if (dSy <= 0.f) dS = fminf(dSx, dSz);
    else if (dSz <= 0.f) dS = fminf(dSx, dSy);
    else { dS = fminf(dSx, dSy); dS = fminf(dS, dSz); }
    if (dS <= 0.f) { err = 36; return ST_ERROR; }

    // dense opacity: inactive cells hold row 0 (= 0), so "active > 0" of the
    // interaction test (1517) holds whenever tau can exceed passProb
    const unsigned p = j3_point(G.js, xP - 1, yP - 1, zP - 1);
    if (!ZERO_OPACITY && p != p0) kappa = G.opac3[j3_index(G.nbins, nuP - 1, p)];
    const float tauCell = dS * kappa;
    const float dV = AX(A.dx, xP) * AX(A.dy, yP) * AX(A.dz, zP);  // 1442-1514
    kappa_out = kappa;
    dV_out = dV;
    W.segPt = p;
    W.segX = xP; W.segY = yP; W.segZ = zP;

    if (!ZERO_OPACITY && (W.absTau + tauCell > W.passProb)) {
        // interaction, 1517-1580
        float dlLoc = (W.passProb - W.absTau) / kappa;
        rx = rx + dlLoc * vx;
        ry = ry + dlLoc * vy;
        rz = rz + dlLoc * vz;
        if (isnan(rx) || isnan(ry) || isnan(rz)) { err = 37; return ST_ERROR; }
        if (sym) {
            if (rx <= AX(A.x, 1)) { if (vx < 0.f) vx = -vx; rx = AX(A.x, 1); }
            if (ry <= AX(A.y, 1)) { if (vy < 0.f) vy = -vy; ry = AX(A.y, 1); }
            if (rz <= AX(A.z, 1)) { if (vz < 0.f) vz = -vz; rz = AX(A.z, 1); }
        }
        W.seg = dlLoc;
        W.rx = rx; W.ry = ry; W.rz = rz; W.vx = vx; W.vy = vy; W.vz = vz;
        W.xP = xP; W.yP = yP; W.zP = zP;
        return ST_ABSORBED;   // caller: J, R_out escape test, scatter/absorb decision
    }

    // free flight to the wall, 1818-1976
    W.seg = dS;
    W.absTau = W.absTau + tauCell;
    rx = rx + dS * vx;
    ry = ry + dS * vy;
    rz = rz + dS * vz;
    if (dS == dSx && vx > 0.f) xP = xP + 1;
    else if (dS == dSx && vx < 0.f) xP = xP - 1;
    else if (dS == dSy && vy > 0.f) yP = yP + 1;
    else if (dS == dSy && vy < 0.f) yP = yP - 1;
    else if (dS == dSz && vz > 0.f) zP = zP + 1;
    else if (dS == dSz && vz < 0.f) zP = zP - 1;

    int res = ST_CONTINUE;
    if (!sym) {                                                  // 1986-2194
        bool ret = false;
        if (ry <= AX(A.y, 1) - G.geoCorrY || yP < 1) { yP = 1; ret = true; }
        if (ry > AX(A.y, ny) + G.geoCorrY || yP > ny) { yP = ny; ret = true; }
        if (rx <= AX(A.x, 1) - G.geoCorrX || xP < 1) { xP = 1; ret = true; }
        if (rx >= AX(A.x, nx) + G.geoCorrX || xP > nx) { xP = nx; ret = true; }
        if (rz <= AX(A.z, 1) - G.geoCorrZ || zP < 1) { zP = 1; ret = true; }
        if (rz >= AX(A.z, nz) + G.geoCorrZ || zP > nz) { zP = nz; ret = true; }
        if (ret) res = ST_ESCAPED;
    }
    if (res == ST_CONTINUE) {                                    // 2421-2554
        if ((!sym && (rx <= AX(A.x, 1) - G.geoCorrX || ry <= AX(A.y, 1) - G.geoCorrY ||
                      rz <= AX(A.z, 1) - G.geoCorrZ)) ||
            rx >= AX(A.x, nx) + G.geoCorrX || ry >= AX(A.y, ny) + G.geoCorrY ||
            rz >= AX(A.z, nz) + G.geoCorrZ || xP > nx || yP > ny || zP > nz)
            res = ST_ESCAPED;
    }
    if (res == ST_CONTINUE && sym) {                             // 2674-2699
        if (rx <= AX(A.x, 1) || xP < 1) { if (vx < 0.f) vx = -vx; xP = 1; rx = AX(A.x, 1); }
        if (ry <= AX(A.y, 1) || yP < 1) { if (vy < 0.f) vy = -vy; yP = 1; ry = AX(A.y, 1); }
        if (rz <= AX(A.z, 1) || zP < 1) { if (vz < 0.f) vz = -vz; zP = 1; rz = AX(A.z, 1); }
    }
    if (res == ST_CONTINUE && (xP > nx || yP > ny || zP > nz)) res = ST_ESCAPED;  // 2728
    W.rx = rx; W.ry = ry; W.rz = rz; W.vx = vx; W.vy = vy; W.vz = vz;
    W.xP = xP; W.yP = yP; W.zP = zP;
    return res;
}

}  // namespace mcd

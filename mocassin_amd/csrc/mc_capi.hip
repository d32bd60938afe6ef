// mc_capi.hip - C ABI (include/mocassin_amd.h) over the HIP kernels.
//
// A context owns one device copy of the grid and tables, the FP64
// accumulators and one HIP stream.  No global mutable state: contexts are
// independent, so several nebulae can live side by side in one process.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/mocassin_amd.h"
#include "mc_balance.h"
#include "mc_device.h"
#include "mc_dust.h"
#include "mc_emission.h"
#include "mc_kernels.h"
#include "mc_mlp.h"
#include "mc_opacity.h"

#pragma clang fp contract(off)

// Device event counters of one launch, after the packet counter: the eight
// event counts of mc_counters (n_seg .. n_flights), then the wavefront walk
// passes, the lanes that walked in them, the emission passes and the lanes
// that emitted in them.
constexpr int kCounts = 12;

static void fill_counters(mc_counters &k, const unsigned long long *h, double q, float ms)
{
    k.n_seg = (double)h[1]; k.n_abs = (double)h[2]; k.n_sca = (double)h[3];
    k.n_esc = (double)h[4]; k.n_lost = (double)h[5]; k.n_line = (double)h[6];
    k.n_trapped = (double)h[7]; k.n_flights = (double)h[8];
    k.qphot = q; k.kernel_ms = ms;
    k.n_walk_passes = (double)h[9]; k.n_walk_lanes = (double)h[10];
    k.n_emit_passes = (double)h[11]; k.n_emit_lanes = (double)h[12];
}

struct mc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    mcd::Grid G{};
    int nStars = 0;
    std::vector<float> starPos, deltaE;
    std::vector<int> starIdx;
    bool tables = false;
    int cus = 256;
    mc_tuning tune = { 4, 256, 16, 16, MC_KERNEL_AUTO };   // the pool kernel where the grid allows    // the autotuner's best on the octant; within 1.3% on the dust and clumpy grids
    // device allocations
    float *d_axis = nullptr, *d_mid = nullptr, *d_dfac = nullptr;
    int *d_active = nullptr;
    float *d_opac3 = nullptr, *d_sca3 = nullptr, *d_recCDF = nullptr;
    unsigned short *d_recGuide = nullptr;
    float *d_totalLines = nullptr, *d_specCDF = nullptr, *d_nuArray = nullptr, *d_gSca = nullptr;
    double *d_J3 = nullptr, *d_J = nullptr, *d_esc = nullptr;
    mcd::View *d_view = nullptr;            // viewing-angle bins (null: escapes in plane 0 only)
    int esc_planes = 1;                      // nAngleBins + 1
    unsigned long long *d_scratch = nullptr;  // next_packet, counts[kCounts]
    int *d_error = nullptr;
    double *d_qphot = nullptr;
    // per-cell balance
    std::vector<int> active, cell_xyz;      // host copies: active map, xyz of each cell
    int *d_cell_xyz = nullptr;
    bool balance = false;
    mcb::Prepared bal;
    mcb::Atomic *d_atomic = nullptr;
    double *d_bal_tab = nullptr;             // sigma, nu, CEL arrays (one allocation)
    double *d_Hden = nullptr, *d_Te = nullptr, *d_Ne = nullptr, *d_ion = nullptr;
    int *d_flags = nullptr, *d_nIter = nullptr;
    double *d_cold = nullptr;                // Te0, Ne0, ion0 of a warm-started run
    // launch order of the balance: cells by their temperature iterations in
    // the last balance, most first (d_order[b] = cell of block b)
    std::vector<int32_t> last_nit, order_h;
    int *d_order = nullptr;
    double *d_ion_host = nullptr;            // ionDen in the host's layout, staged for the layout kernels
    double *d_Rint = nullptr, *d_Hint = nullptr;  // [nCells+1][nch] channel integrals
    unsigned *d_cross = nullptr;             // segment counts per lattice point (when enabled)
    // warm-start network
    bool mlp = false;
    int mlp_step = 0;
    float *d_mlp = nullptr;                  // P | grad | m | v (kParams each) | scale
    uint16_t *d_mlpW = nullptr;              // bf16 copy of P
    float *d_mX = nullptr, *d_mY = nullptr;  // inference rows [nCells][72], [nCells][32]
    uint16_t *d_mX0 = nullptr;
    int *d_hit = nullptr;
    // grains
    bool dust = false;
    mcg::DustArgs dust_args{};
    double *d_dust_tab = nullptr, *d_Tdust = nullptr;
    int *d_dust_mono = nullptr;
    bool dust_pe = false;                    // grain charge and photoelectric terms (with gas)
    double *d_pe_tab = nullptr;              // Qa, Yn, Ndust, gasDustColl_d, photoelHeat_d
    double *pe_gdc = nullptr, *pe_peh = nullptr;
    mcb::Grains *d_grains = nullptr;
    std::vector<float> nuArray;
    // opacity tables on the device
    bool opac = false, opac_built = false, opac_sca_built = false;
    bool state_dev = false, tdust_dev = false;   // device state written by a balance / grain run
    bool dust_opac = false;                      // grains carry the dust-add tables
    int dust_rows = 0;
    float *d_op_dust = nullptr;                  // dsca | dabs | dw | tsub | Ndust
    int *d_op_i = nullptr;                       // bin_start | slot_e | slot_stage
    float *d_op_f = nullptr;                     // slot_abun | slot_efac
    mco::Ent *d_op_ent = nullptr;                // cross-section entries of every bin
    float *d_op_st = nullptr, *d_op_Td = nullptr, *d_op_out = nullptr;
    mco::OpacArgs op{};
    int *op_slot_e = nullptr, *op_slot_st = nullptr;   // into d_op_i
    std::vector<int> op_sl_e, op_sl_st;                // slot -> element index, stage
    // line luminosities: per-block partial sums, a host state, the result
    double *d_ln_part = nullptr, *d_ln_st = nullptr, *d_ln_out = nullptr;
    size_t ln_part_n = 0;
    int *d_ln_black = nullptr;
    // re-emission tables on the device
    bool emis = false, rec_built = false;
    // mc_emission_build writes into these and swaps them with d_recCDF,
    // d_recGuide and d_totalLines only when every cell's row is good
    float *d_recCDF_b = nullptr, *d_totalLines_b = nullptr;
    unsigned short *d_recGuide_b = nullptr;
    int *d_em_i = nullptr, *d_em_err = nullptr;
    float *d_em_f = nullptr;
    double *d_em_d = nullptr, *d_em_st = nullptr, *d_em_Td = nullptr;
    mce::EmArgs em{};
    char err[512] = "";
};

static thread_local char g_create_err[512] = "";

static int fail(mc_ctx *c, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c ? c->err : g_create_err, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(c, call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((c), e_ == hipErrorOutOfMemory ? MC_ENOMEM : MC_ENODEV, "%s: %s", \
                        #call, hipGetErrorString(e_));                                    \
    } while (0)

template <class T>
static hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, n * sizeof(T) ? n * sizeof(T) : sizeof(T));
}

// Every copy and fill of a context goes on its stream, in order with its
// kernels: the context's stream is non-blocking, so a blocking copy or fill
// on the null stream is not ordered with its work (a fill of a fresh buffer
// could land after the kernel that writes it).  A copy returns when it is
// done, so its host buffer may be a temporary.
static hipError_t copy_on(hipStream_t s, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}
static hipError_t fill_on(hipStream_t s, void *dst, int value, size_t bytes)
{
    return hipMemsetAsync(dst, value, bytes, s);
}

static int ctx_init(mc_ctx *c, const mc_desc *d);
static int noalloc_scratch_init(mc_ctx *c);

extern "C" int mc_ctx_create(const mc_desc *d, mc_ctx **out)
{
    if (!out) return fail(nullptr, MC_EINVAL, "null output pointer");
    *out = nullptr;
    if (!d || d->abi_version != MC_ABI_VERSION)
        return fail(nullptr, MC_EINVAL, "descriptor missing or ABI version mismatch");
    if (d->nx < 2 || d->ny < 2 || d->nz < 2 || d->nCells < 1 || d->nbins < 3 || d->nStars < 1 ||
        !d->xAxis || !d->yAxis || !d->zAxis || !d->active || !d->nuArray || !d->starPos ||
        !d->starIdx || !d->deltaE)
        return fail(nullptr, MC_EINVAL, "descriptor has empty sizes or null arrays");
    if (d->nbins > 65536)
        return fail(nullptr, MC_EUNSUP, "nbins %d: the transport's J queue indexes at most 65536 bins", d->nbins);
    if (mcd::j3_points(mcd::j3_shape(d->nx, d->ny, d->nz)) > 0x7fffffffu)
        return fail(nullptr, MC_EUNSUP, "lattice %d x %d x %d: at most 2^31 points (padded to multiples of 4)",
                    d->nx, d->ny, d->nz);
    // the energy of a packet of each source (iteration_mod.f90:477-493
    // divides the luminosity by the packet count): finite and positive
    for (int s = 0; s < d->nStars; ++s)
        if (!(d->deltaE[s] > 0.f) || std::isinf(d->deltaE[s]))
            return fail(nullptr, MC_EINVAL, "deltaE of source %d is %g: must be finite and > 0", s + 1,
                        (double)d->deltaE[s]);
    mc_ctx *c = new mc_ctx();
    int rc = ctx_init(c, d);
    if (rc) {
        memcpy(g_create_err, c->err, sizeof g_create_err);
        mc_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return MC_OK;
}

static int ctx_init(mc_ctx *c, const mc_desc *d)
{
    c->device = d->device;
    HIPCHK(c, hipSetDevice(d->device));
    {
        const int rs = noalloc_scratch_init(c);
        if (rs) return rs;
    }
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreate(&c->ev0));
    HIPCHK(c, hipEventCreate(&c->ev1));
    const int nx = d->nx, ny = d->ny, nz = d->nz, nb = d->nbins, nc = d->nCells;
    const int na = nx + ny + nz, nm = na - 3;
    // axes, midpoints and per-axis volume factors in FP32 exactly as the
    // reference evaluates them inline (photon_mod.f90:1266, 1448-1500)
    std::vector<float> axis(na), mid(nm), dfac(na);
    const float *ax[3] = { d->xAxis, d->yAxis, d->zAxis };
    const int n3[3] = { nx, ny, nz };
    const bool sym = (d->flags & MC_SYMMETRIC) != 0;
    for (int a = 0, oa = 0, om = 0; a < 3; ++a) {
        const float *v = ax[a];
        const int n = n3[a];
        for (int i = 0; i < n; ++i) {
            axis[oa + i] = v[i];
            if (std::isnan(v[i])) return fail(c, MC_EINVAL, "NaN in axis %d", a);
        }
        for (int i = 0; i + 1 < n; ++i) {
            mid[om + i] = (v[i + 1] + v[i]) / 2.f;
            if (!(v[i + 1] > v[i])) return fail(c, MC_EINVAL, "axis %d not increasing", a);
        }
        for (int i = 1; i <= n; ++i) {
            float w;
            if (i > 1 && i < n) w = std::fabs(v[i] - v[i - 2]) / 2.f;
            else if (i == 1) w = sym ? std::fabs(v[1] - v[0]) / 2.f : std::fabs(v[1] - v[0]);
            else w = std::fabs(v[n - 1] - v[n - 2]);
            dfac[oa + i - 1] = w / 1.e15f;
        }
        oa += n;
        om += n - 1;
    }
    // active map: values in [0, nCells]; sub-grids (negative) are outside the path
    for (size_t i = 0; i < (size_t)nx * ny * nz; ++i)
        if (d->active[i] < 0 || d->active[i] > nc)
            return fail(c, d->active[i] < 0 ? MC_EUNSUP : MC_EINVAL,
                        "active map entry %zu = %d outside [0,%d]", i, d->active[i], nc);
    c->active.assign(d->active, d->active + (size_t)nx * ny * nz);
    c->cell_xyz.assign((size_t)nc + 1, -1);
    for (size_t i = 0; i < (size_t)nx * ny * nz; ++i)
        if (d->active[i] > 0) c->cell_xyz[d->active[i]] = (int)i;
    c->nuArray.assign(d->nuArray, d->nuArray + d->nbins);
    c->nStars = d->nStars;
    c->starPos.assign(d->starPos, d->starPos + 3 * d->nStars);
    c->starIdx.assign(d->starIdx, d->starIdx + 3 * d->nStars);
    c->deltaE.assign(d->deltaE, d->deltaE + d->nStars);
    for (int s = 0; s < d->nStars; ++s) {
        const int *si = &c->starIdx[3 * s];
        if (si[0] < 1 || si[0] > nx || si[1] < 1 || si[1] > ny || si[2] < 1 || si[2] > nz)
            return fail(c, MC_EINVAL, "source %d cell outside the grid", s + 1);
    }

    HIPCHK(c, dalloc(&c->d_axis, na));
    HIPCHK(c, dalloc(&c->d_mid, nm));
    HIPCHK(c, dalloc(&c->d_dfac, na));
    HIPCHK(c, dalloc(&c->d_active, (size_t)nx * ny * nz));
    HIPCHK(c, dalloc(&c->d_nuArray, nb));
    const size_t ncell3 = (size_t)nx * ny * nz;
    HIPCHK(c, dalloc(&c->d_J3, mcd::j3_size(mcd::j3_shape(nx, ny, nz), nb)));
    HIPCHK(c, dalloc(&c->d_esc, (size_t)(nc + 1) * (nb + 1)));
    HIPCHK(c, dalloc(&c->d_scratch, 1 + kCounts));
    HIPCHK(c, dalloc(&c->d_error, 1));
    HIPCHK(c, dalloc(&c->d_qphot, 1));
    HIPCHK(c, copy_on(c->stream, c->d_axis, axis.data(), na * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c->stream, c->d_mid, mid.data(), nm * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c->stream, c->d_dfac, dfac.data(), na * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c->stream, c->d_active, d->active, (size_t)nx * ny * nz * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c->stream, c->d_nuArray, d->nuArray, nb * sizeof(float), hipMemcpyHostToDevice));

    mcd::Grid &G = c->G;
    G.nx = nx; G.ny = ny; G.nz = nz; G.nCells = nc; G.nbins = nb; G.flags = d->flags;
    G.nStars = d->nStars;
    G.R_out = d->R_out;
    G.geoCorrX = d->geoCorr[0]; G.geoCorrY = d->geoCorr[1]; G.geoCorrZ = d->geoCorr[2];
    G.ionEdge1 = d->ionEdge1;
    G.axis = c->d_axis; G.mid = c->d_mid; G.dfac = c->d_dfac; G.active = c->d_active;
    G.nuArray = c->d_nuArray;
    G.ncell3 = (int)ncell3;
    G.js = mcd::j3_shape(nx, ny, nz);
    G.J3 = c->d_J3; G.esc = c->d_esc;
    G.seed_lo = (uint32_t)d->seed; G.seed_hi = (uint32_t)(d->seed >> 32);

    int cus = 0;
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->device));
    c->cus = cus > 0 ? cus : 256;
    int rc = mc_zero_estimators(c);
    if (rc) return rc;
    return MC_OK;
}

// monotone non-decreasing, finite, within [0, 1 + tiny]
static bool cdf_ok(const float *row, size_t stride, int nb)
{
    float prev = 0.f;
    for (int i = 0; i < nb; ++i) {
        float v = row[(size_t)i * stride];
        if (!(v >= prev) || !(v <= 1.0001f)) return false;
        prev = v;
    }
    return true;
}

extern "C" int mc_upload_tables(mc_ctx *c, const float *opacity, const float *scaOpac,
                                const float *recPDF, const float *totalLines,
                                const float *specCDF, const float *gSca)
{
    if (!c) return MC_EINVAL;
    mcd::Grid &G = c->G;
    const int nc = G.nCells, nb = G.nbins;
    const size_t nt = (size_t)(nc + 1) * nb;
    const bool dust = (G.flags & MC_DUST) != 0, gas = (G.flags & MC_GAS) != 0;
    if (!specCDF) return fail(c, MC_EINVAL, "specCDF is required");
    if (!opacity && !c->opac_built)
        return fail(c, MC_EINVAL, "opacity is required (no tables on the device yet)");
    if (gas && (!recPDF || !totalLines) && !(!recPDF && !totalLines && c->rec_built))
        return fail(c, MC_EINVAL, "gas needs recPDF and totalLines (no re-emission tables on the device yet)");
    const bool rec_host = gas && recPDF;
    if (dust && !scaOpac && !c->opac_sca_built)
        return fail(c, MC_EINVAL, "dust needs scaOpac (no scattering table on the device yet)");
    if (dust && !(G.flags & MC_ISOTROPIC) && !gSca) return fail(c, MC_EINVAL, "HG scattering needs gSca");
    HIPCHK(c, hipSetDevice(c->device));
    if (opacity)
        for (size_t i = 0; i < nt; ++i)
            if (!(opacity[i] >= 0.f) || std::isinf(opacity[i]))
                return fail(c, MC_ETABLE, "opacity entry %zu not finite and >= 0", i);
    for (int s = 0; s < c->nStars; ++s)
        if (!cdf_ok(specCDF + (size_t)s * nb, 1, nb))
            return fail(c, MC_ETABLE, "specCDF of source %d not a monotone CDF", s + 1);
    std::vector<float> rec;
    if (rec_host) {
        rec.resize(nt);
        for (int cell = 0; cell <= nc; ++cell) {
            if (cell > 0 && !cdf_ok(recPDF + cell, (size_t)(nc + 1), nb))
                return fail(c, MC_ETABLE, "recPDF row of cell %d not a monotone CDF", cell);
            for (int nu = 0; nu < nb; ++nu) rec[(size_t)cell * nb + nu] = recPDF[(size_t)nu * (nc + 1) + cell];
        }
        for (int cell = 0; cell <= nc; ++cell)
            if (std::isnan(totalLines[cell])) return fail(c, MC_ETABLE, "NaN in totalLines");
    }
    if (opacity)
        for (int nu = 0; nu < nb; ++nu)
            if (opacity[(size_t)nu * (nc + 1)] != 0.f)
                return fail(c, MC_ETABLE, "opacity row 0 (inactive cells) must be zero (bin %d)", nu + 1);
    float *tmp = nullptr;
    HIPCHK(c, dalloc(&tmp, nt));
    int rc = MC_OK;
    hipError_t e = hipSuccess;
    if (opacity) {
        e = copy_on(c->stream, tmp, opacity, nt * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess && !c->d_opac3) e = dalloc(&c->d_opac3, mcd::j3_size(G.js, nb));
        if (e == hipSuccess) e = mcd_densify(tmp, c->d_active, G.js, nc + 1, nb, c->d_opac3, c->stream);
        c->opac_built = e == hipSuccess;
    }
    if (e == hipSuccess && dust && scaOpac) {
        e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = copy_on(c->stream, tmp, scaOpac, nt * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess && !c->d_sca3) e = dalloc(&c->d_sca3, mcd::j3_size(G.js, nb));
        if (e == hipSuccess) e = mcd_densify(tmp, c->d_active, G.js, nc + 1, nb, c->d_sca3, c->stream);
        c->opac_sca_built = e == hipSuccess;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? MC_ENOMEM : MC_ENODEV, "table upload: %s", hipGetErrorString(e));
    (void)rc;
    if (!c->d_specCDF) HIPCHK(c, dalloc(&c->d_specCDF, (size_t)nb * c->nStars));
    HIPCHK(c, copy_on(c->stream, c->d_specCDF, specCDF, (size_t)nb * c->nStars * sizeof(float), hipMemcpyHostToDevice));
    if (rec_host) {
        if (!c->d_recCDF) HIPCHK(c, dalloc(&c->d_recCDF, nt));
        if (!c->d_recGuide) HIPCHK(c, dalloc(&c->d_recGuide, (size_t)(nc + 1) * (mcd::kGuide + 1)));
        if (!c->d_totalLines) HIPCHK(c, dalloc(&c->d_totalLines, (size_t)nc + 1));
        HIPCHK(c, copy_on(c->stream, c->d_recCDF, rec.data(), nt * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(c, copy_on(c->stream, c->d_totalLines, totalLines, (nc + 1) * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(c, mcd_guides(c->d_recCDF, nc + 1, nb, c->d_recGuide, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (dust && gSca) {
        if (!c->d_gSca) HIPCHK(c, dalloc(&c->d_gSca, nb));
        HIPCHK(c, copy_on(c->stream, c->d_gSca, gSca, nb * sizeof(float), hipMemcpyHostToDevice));
    }
    G.opac3 = c->d_opac3; G.sca3 = c->d_sca3; G.specCDF = c->d_specCDF;
    G.recCDF = c->d_recCDF; G.recGuide = c->d_recGuide; G.totalLines = c->d_totalLines;
    G.gSca = c->d_gSca;
    c->tables = true;
    return MC_OK;
}

extern "C" int mc_zero_estimators(mc_ctx *c)
{
    if (!c) return MC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nc1 = (size_t)c->G.nCells + 1;
    HIPCHK(c, hipMemsetAsync(c->d_J3, 0, mcd::j3_size(c->G.js, c->G.nbins) * sizeof(double), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_esc, 0, nc1 * (c->G.nbins + 1) * c->esc_planes * sizeof(double), c->stream));
    if (c->G.cross) HIPCHK(c, hipMemsetAsync(c->G.cross, 0, mcd::j3_points(c->G.js) * sizeof(unsigned), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

extern "C" int mc_count_crossings(mc_ctx *c, int32_t enable)
{
    if (!c) return MC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    if (enable && !c->d_cross) {
        HIPCHK(c, dalloc(&c->d_cross, mcd::j3_points(c->G.js)));
        HIPCHK(c, fill_on(c->stream, c->d_cross, 0, mcd::j3_points(c->G.js) * sizeof(unsigned)));
    }
    c->G.cross = enable ? c->d_cross : nullptr;
    return MC_OK;
}

extern "C" int mc_fetch_crossings(mc_ctx *c, uint32_t *out)
{
    if (!c) return MC_EINVAL;
    if (!out) return fail(c, MC_EINVAL, "null output");
    if (!c->d_cross) return fail(c, MC_ESTATE, "crossing counts are not enabled");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<uint32_t> pts(mcd::j3_points(c->G.js));
    HIPCHK(c, copy_on(c->stream, pts.data(), c->d_cross, pts.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int cell = 0; cell <= c->G.nCells; ++cell) out[cell] = 0;
    for (int i = 0; i < c->G.ncell3; ++i) out[c->active[i]] += pts[mcd::j3_point_of(c->G.js, i)];
    return MC_OK;
}

extern "C" int mc_fetch_estimators(mc_ctx *c, double *J_out, double *esc_out)
{
    if (!c) return MC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nc1 = (size_t)c->G.nCells + 1;
    if (J_out) {
        if (!c->d_J) HIPCHK(c, dalloc(&c->d_J, nc1 * c->G.nbins));
        HIPCHK(c, hipMemsetAsync(c->d_J, 0, nc1 * c->G.nbins * sizeof(double), c->stream));
        HIPCHK(c, mcd_compact(c->d_J3, c->d_active, c->G.js, (int)nc1, c->G.nbins, c->d_J, c->stream));
        HIPCHK(c, hipMemcpyAsync(J_out, c->d_J, nc1 * c->G.nbins * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    }
    if (esc_out)
        HIPCHK(c, hipMemcpyAsync(esc_out, c->d_esc, nc1 * (c->G.nbins + 1) * c->esc_planes * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

extern "C" int mc_set_estimators(mc_ctx *c, const double *J, const double *esc)
{
    if (!c) return MC_EINVAL;
    if (!J) return fail(c, MC_EINVAL, "J is required");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nc1 = (size_t)c->G.nCells + 1, nb = c->G.nbins, n3 = c->G.ncell3;
    std::vector<double> dense(mcd::j3_size(c->G.js, (int)nb), 0.0);
    for (size_t i = 0; i < n3; ++i) {
        const int a = c->active[i];
        if (a <= 0) continue;
        const unsigned pt = mcd::j3_point_of(c->G.js, (int)i);
        for (size_t nu = 0; nu < nb; ++nu) dense[mcd::j3_index((int)nb, (int)nu, pt)] = J[nu * nc1 + a];
    }
    HIPCHK(c, copy_on(c->stream, c->d_J3, dense.data(), dense.size() * sizeof(double), hipMemcpyHostToDevice));
    if (esc) HIPCHK(c, copy_on(c->stream, c->d_esc, esc, nc1 * (nb + 1) * c->esc_planes * sizeof(double), hipMemcpyHostToDevice));
    return MC_OK;
}

extern "C" int mc_esc_planes(const mc_ctx *c)
{
    return c ? c->esc_planes : 0;
}

extern "C" int mc_ctx_shape(const mc_ctx *c, int32_t *nCells, int32_t *nbins)
{
    if (nCells) *nCells = c ? c->G.nCells : 0;
    if (nbins) *nbins = c ? c->G.nbins : 0;
    return c ? MC_OK : MC_EINVAL;
}

extern "C" int mc_set_view_angles(mc_ctx *c, const mc_view_angles *v)
{
    if (!c) return MC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const int n = v ? v->nAngleBins : 0;
    mcd::View h{};
    if (n > 0) {
        if (n > 2) return fail(c, MC_EINVAL, "nAngleBins %d: at most two viewing angles", n);
        if (v->totAngleBinsTheta < 1 || v->totAngleBinsTheta > mcd::kViewMax || v->totAngleBinsPhi < 1 ||
            v->totAngleBinsPhi > mcd::kViewMax)
            return fail(c, MC_EINVAL, "angle bins %d x %d outside 1..%d", v->totAngleBinsTheta,
                        v->totAngleBinsPhi, mcd::kViewMax);
        if (!(v->dTheta > 0.f) || !(v->dPhi > 0.f) || !std::isfinite(v->dTheta) || !std::isfinite(v->dPhi))
            return fail(c, MC_EINVAL, "dTheta and dPhi must be finite and positive");
        if (!v->viewPointPtheta || !v->viewPointPphi || !v->viewPointTheta || !v->viewPointPhi)
            return fail(c, MC_EINVAL, "viewing-angle tables missing");
        h.nAngle = n; h.totT = v->totAngleBinsTheta; h.totP = v->totAngleBinsPhi;
        h.dTheta = v->dTheta; h.dPhi = v->dPhi;
        for (int i = 0; i <= h.totT; ++i) {
            h.Pt[i] = v->viewPointPtheta[i];
            if (h.Pt[i] < 0 || h.Pt[i] > n) return fail(c, MC_EINVAL, "viewPointPtheta(%d) = %d", i, h.Pt[i]);
        }
        for (int i = 0; i <= h.totP; ++i) {
            h.Pp[i] = v->viewPointPphi[i];
            if (h.Pp[i] < 0 || h.Pp[i] > n) return fail(c, MC_EINVAL, "viewPointPphi(%d) = %d", i, h.Pp[i]);
        }
        for (int i = 0; i <= n; ++i) { h.vth[i] = v->viewPointTheta[i]; h.vph[i] = v->viewPointPhi[i]; }
    }
    const size_t nc1 = (size_t)c->G.nCells + 1, nb1 = (size_t)c->G.nbins + 1;
    if (n + 1 != c->esc_planes) {
        double *e = nullptr;
        HIPCHK(c, dalloc(&e, nc1 * nb1 * (size_t)(n + 1)));
        (void)hipFree(c->d_esc);
        c->d_esc = e;
        c->G.esc = e;
        c->esc_planes = n + 1;
    }
    HIPCHK(c, fill_on(c->stream, c->d_esc, 0, nc1 * nb1 * c->esc_planes * sizeof(double)));
    if (n > 0) {
        if (!c->d_view) HIPCHK(c, hipMalloc((void **)&c->d_view, sizeof(mcd::View)));
        HIPCHK(c, copy_on(c->stream, c->d_view, &h, sizeof h, hipMemcpyHostToDevice));
        c->G.view = c->d_view;
    } else {
        c->G.view = nullptr;
    }
    return MC_OK;
}

// the lattice point of every cell on the device (per-cell kernels that read
// the cell's spectrum from J3)
static int cell_positions(mc_ctx *c)
{
    for (int cell = 1; cell <= c->G.nCells; ++cell)
        if (c->cell_xyz[cell] < 0) return fail(c, MC_EINVAL, "active map has no position for cell %d", cell);
    if (!c->d_cell_xyz) {
        const size_t nc1 = (size_t)c->G.nCells + 1;
        HIPCHK(c, dalloc(&c->d_cell_xyz, nc1));
        HIPCHK(c, copy_on(c->stream, c->d_cell_xyz, c->cell_xyz.data(), nc1 * sizeof(int), hipMemcpyHostToDevice));
    }
    return MC_OK;
}

extern "C" int mc_balance_setup(mc_ctx *c, const mc_atomic *at)
{
    if (!c) return MC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    {
        int rc = cell_positions(c);
        if (rc) return rc;
    }
    c->balance = false;
    std::vector<float> nu(c->G.nbins);
    HIPCHK(c, copy_on(c->stream, nu.data(), c->d_nuArray, nu.size() * sizeof(float), hipMemcpyDeviceToHost));
    mcb::Prepared &P = c->bal;
    c->err[0] = 0;
    int rc = mcb::prepare(at, c->G.nbins, nu.data(), P, c->err, sizeof c->err);
    if (rc) return rc;
    if (at->nElemUsed != P.A.nE) return fail(c, MC_ETABLE, "nElemUsed %d but %d elements on", at->nElemUsed, P.A.nE);
    // one device block for every table, in this order
    const std::vector<double> *parts[] = { &P.sigma, &P.nu, &P.cel_e, &P.cel_a, &P.cel_q,
                                           &P.cel_y2, &P.cel_t, &P.cel_g, &P.cR, &P.cH };
    size_t tot = 0;
    for (auto *v : parts) tot += v->size();
    if (c->d_bal_tab) { (void)hipFree(c->d_bal_tab); c->d_bal_tab = nullptr; }
    HIPCHK(c, dalloc(&c->d_bal_tab, tot));
    const double *ptr[10];
    size_t off = 0;
    for (int k = 0; k < 10; ++k) {
        ptr[k] = c->d_bal_tab + off;
        if (!parts[k]->empty())
            HIPCHK(c, copy_on(c->stream, c->d_bal_tab + off, parts[k]->data(), parts[k]->size() * sizeof(double),
                                hipMemcpyHostToDevice));
        off += parts[k]->size();
    }
    mcb::Atomic A = P.A;
    A.sigma = ptr[0]; A.nu = ptr[1]; A.cel_e = ptr[2]; A.cel_a = ptr[3]; A.cel_q = ptr[4];
    A.cel_y2 = ptr[5]; A.cel_t = ptr[6]; A.cel_g = ptr[7]; A.cR = ptr[8]; A.cH = ptr[9];
    if (!c->d_atomic) HIPCHK(c, dalloc(&c->d_atomic, 1));
    HIPCHK(c, copy_on(c->stream, c->d_atomic, &A, sizeof A, hipMemcpyHostToDevice));
    const size_t nc1 = (size_t)c->G.nCells + 1;
    const size_t nion = nc1 * P.A.nE * (P.A.nstages + 1);
    if (!c->d_Hden) {
        HIPCHK(c, dalloc(&c->d_Hden, nc1));
        HIPCHK(c, dalloc(&c->d_Te, nc1));
        HIPCHK(c, dalloc(&c->d_Ne, nc1));
    }
    if (!c->d_flags) {
        HIPCHK(c, dalloc(&c->d_flags, nc1));
        HIPCHK(c, dalloc(&c->d_nIter, nc1));
    }
    if (c->d_ion) { (void)hipFree(c->d_ion); c->d_ion = nullptr; }
    HIPCHK(c, dalloc(&c->d_ion, nion));
    if (c->d_Rint) { (void)hipFree(c->d_Rint); c->d_Rint = nullptr; }
    if (c->d_Hint) { (void)hipFree(c->d_Hint); c->d_Hint = nullptr; }
    HIPCHK(c, dalloc(&c->d_Rint, nc1 * (P.A.nch > 0 ? P.A.nch : 1)));
    HIPCHK(c, dalloc(&c->d_Hint, nc1 * (P.A.nch > 0 ? P.A.nch : 1)));
    if (!c->d_hit) HIPCHK(c, dalloc(&c->d_hit, nc1));
    c->balance = true;
    return MC_OK;
}

static mcm::Net mlp_net(mc_ctx *c) { return mcm::Net{ c->d_mlpW, c->d_mlp }; }
static mcm::Scale mlp_scale(mc_ctx *c)
{
    const float *sc = c->d_mlp + 4 * (size_t)mcm::kParams;
    return mcm::Scale{ sc, sc + mcm::kIn, sc + 2 * mcm::kIn, sc + 2 * mcm::kIn + mcm::kOut };
}

// inference buffers for one row per active cell
static int mlp_rows(mc_ctx *c)
{
    if (c->d_mX) return MC_OK;
    const size_t nc = (size_t)c->G.nCells;
    HIPCHK(c, dalloc(&c->d_mX, nc * mcm::kIn));
    HIPCHK(c, dalloc(&c->d_mY, nc * mcm::kOutP));
    HIPCHK(c, dalloc(&c->d_mX0, nc * mcm::kInP));
    return MC_OK;
}

static mcb::CellArgs cell_args(mc_ctx *c, double jscale)
{
    mcb::CellArgs a{};
    a.ncells = c->G.nCells;
    a.jscale = jscale;
    a.Rint = c->d_Rint;
    a.Hint = c->d_Hint;
    a.hit = c->d_hit;
    a.Hden = c->d_Hden; a.Te = c->d_Te; a.Ne = c->d_Ne; a.ion = c->d_ion;
    a.flags = c->d_flags; a.nIter = c->d_nIter;
    a.solver = MC_SOLVER_REFERENCE; a.max_iter = 25; a.tol = 0.01; a.ion_tol = 0.01; a.step0 = 0.05;
    a.slope0 = 1.0;
    a.warm_iter = a.max_iter;
    return a;
}

// channel integrals, hit flags and (X non-null) network inputs of every cell
static int integrals(mc_ctx *c, const mcb::CellArgs &a, float *X)
{
    HIPCHK(c, hipMemsetAsync(c->d_hit, 0, ((size_t)c->G.nCells + 1) * sizeof(int), c->stream));
    HIPCHK(c, mcb_launch_integrals(c->d_atomic, a, c->d_J3, c->d_active, c->G.nx, c->G.ny, c->G.nz, c->d_Rint, c->d_Hint,
                                   c->d_hit, X, c->stream));
    return MC_OK;
}

// grains: tables of getDustT (rows per (species, size)) and the state buffer
extern "C" int mc_dust_setup(mc_ctx *c, const mc_dust_tables *d)
{
    if (!c) return MC_EINVAL;
    if (!d) return fail(c, MC_EINVAL, "null dust description");
    if (d->nSpecies < 1 || d->nSizes < 1 || d->nSpecies * d->nSizes > mcg::kMaxRows)
        return fail(c, MC_EUNSUP, "nSpecies * nSizes must be 1..%d", mcg::kMaxRows);
    if (d->nSpeciesMax < d->nSpecies) return fail(c, MC_EINVAL, "nSpeciesMax < nSpecies");
    if (d->nTemps < 2 || d->nTemps > mcg::kMaxTemps) return fail(c, MC_EUNSUP, "nTemps must be 2..%d", mcg::kMaxTemps);
    if (!d->absXsec || !d->grainWeight || !d->grainAbun || !d->emIntegral)
        return fail(c, MC_EINVAL, "null dust array");
    const bool pe = d->photoelectric != 0;
    if (pe) {
        if (!d->grainRadius || !d->grainVn || !d->MsurfAtom || !d->aWeight || !d->Ndust)
            return fail(c, MC_EINVAL, "null photoelectric array");
        if (c->G.nbins > 4096) return fail(c, MC_EUNSUP, "the photoelectric balance needs nbins <= 4096");
        for (int i = 1; i < c->G.nbins; ++i)
            if (!(c->nuArray[i] > c->nuArray[i - 1]))
                return fail(c, MC_ETABLE, "nuArray not increasing at bin %d", i + 1);
    }
    if (!(d->XHILimit > 0.0)) return fail(c, MC_EINVAL, "XHILimit must be > 0");
    const int nS = d->nSpecies, nSz = d->nSizes, nT = d->nTemps, nb = c->G.nbins, K = nS * nSz;
    // one FP64 allocation: absx rows | emission rows | weight | abun
    std::vector<double> tab((size_t)K * nb + (size_t)K * nT + nSz + nS);
    std::vector<int> mono(K, 1);
    double *absx = tab.data(), *em = absx + (size_t)K * nb, *w = em + (size_t)K * nT, *ab = w + nSz;
    for (size_t i = 0; i < (size_t)K * nb; ++i) {
        if (!std::isfinite(d->absXsec[i])) return fail(c, MC_ETABLE, "absXsec entry %zu not finite", i);
        absx[i] = d->absXsec[i];
    }
    for (int s = 0; s < nS; ++s)
        for (int ai = 0; ai < nSz; ++ai) {
            const int k = s * nSz + ai;
            for (int t = 0; t < nT; ++t) {
                const float v = d->emIntegral[s + (size_t)nS * (ai + (size_t)nSz * t)];
                if (!std::isfinite(v)) return fail(c, MC_ETABLE, "emIntegral(%d,%d,%d) not finite", s + 1, ai + 1, t + 1);
                em[(size_t)k * nT + t] = v;
                if (t > 0 && em[(size_t)k * nT + t] < em[(size_t)k * nT + t - 1]) mono[k] = 0;
            }
        }
    for (int ai = 0; ai < nSz; ++ai) w[ai] = d->grainWeight[ai];
    for (int s = 0; s < nS; ++s) ab[s] = d->grainAbun[s];
    // photoelectric tables: Qa = absx / (1e-8 a^2) per row and bin, Yn per
    // species and bin, as getGrainEmission and setPhotoelHeatCool form them
    std::vector<double> pet;
    mcb::Grains GH{};
    const size_t nc1 = (size_t)c->G.nCells + 1;
    if (pe) {
        pet.assign((size_t)K * nb + (size_t)nS * nb + nc1, 0.0);
        double *qa = pet.data(), *yn = qa + (size_t)K * nb, *nd = yn + (size_t)nS * nb;
        for (int s = 0; s < nS; ++s) {
            const double vn = d->grainVn[s];
            if (!std::isfinite(vn) || !std::isfinite(d->MsurfAtom[s]))
                return fail(c, MC_ETABLE, "grainVn/MsurfAtom of species %d not finite", s + 1);
            for (int i = 0; i < nb; ++i) {
                const double v = 0.5 * (1. - vn / (double)c->nuArray[i]);
                yn[(size_t)s * nb + i] = v < 0.2 ? v : 0.2;
            }
            GH.vn[s] = vn;
            GH.msurf[s] = d->MsurfAtom[s];
            GH.gabun[s] = d->grainAbun[s];
        }
        for (int ai = 0; ai < nSz; ++ai) {
            const double r = d->grainRadius[ai];
            if (!(r > 0.0) || !std::isfinite(r)) return fail(c, MC_ETABLE, "grainRadius(%d) must be > 0", ai + 1);
            GH.radius[ai] = r;
            GH.weight[ai] = d->grainWeight[ai];
            for (int s = 0; s < nS; ++s)
                for (int i = 0; i < nb; ++i) {
                    const size_t q = ((size_t)s * nSz + ai) * nb + i;
                    qa[q] = absx[q] / (1.e-8 * (r * r));
                }
        }
        for (int z = 0; z < 30; ++z) {
            if (!(d->aWeight[z] > 0.0f)) return fail(c, MC_ETABLE, "aWeight(%d) must be > 0", z + 1);
            GH.aweight[z] = d->aWeight[z];
        }
        for (size_t i = 0; i < nc1; ++i) {
            if (!std::isfinite(d->Ndust[i]) || d->Ndust[i] < 0.0f) return fail(c, MC_ETABLE, "Ndust(%zu) invalid", i);
            nd[i] = d->Ndust[i];
        }
        GH.nS = nS; GH.nSizes = nSz; GH.nsMax = d->nSpeciesMax; GH.K = K;
    }
    HIPCHK(c, hipSetDevice(c->device));
    if (c->d_dust_tab) { (void)hipFree(c->d_dust_tab); c->d_dust_tab = nullptr; }
    if (c->d_dust_mono) { (void)hipFree(c->d_dust_mono); c->d_dust_mono = nullptr; }
    if (c->d_Tdust) { (void)hipFree(c->d_Tdust); c->d_Tdust = nullptr; }
    if (c->d_pe_tab) { (void)hipFree(c->d_pe_tab); c->d_pe_tab = nullptr; }
    if (c->d_grains) { (void)hipFree(c->d_grains); c->d_grains = nullptr; }
    c->dust = false;
    HIPCHK(c, dalloc(&c->d_dust_tab, tab.size()));
    HIPCHK(c, dalloc(&c->d_dust_mono, (size_t)K));
    HIPCHK(c, dalloc(&c->d_Tdust, nc1 * (d->nSpeciesMax + 1) * (nSz + 1)));
    if (pe) {
        // qa | yn | Ndust | gdc | peh
        HIPCHK(c, dalloc(&c->d_pe_tab, pet.size() + 2 * nc1 * K));
        HIPCHK(c, copy_on(c->stream, c->d_pe_tab, pet.data(), pet.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, fill_on(c->stream, c->d_pe_tab + pet.size(), 0, 2 * nc1 * K * sizeof(double)));
        int rc = cell_positions(c);
        if (rc) return rc;
        GH.absx = c->d_dust_tab;
        GH.qa = c->d_pe_tab;
        GH.yn = GH.qa + (size_t)K * nb;
        GH.ndust = GH.yn + (size_t)nS * nb;
        GH.gdc = c->d_pe_tab + pet.size();
        GH.peh = GH.gdc + nc1 * K;
        GH.Tdust = c->d_Tdust;
        GH.J3 = c->d_J3;
        GH.cell_xyz = c->d_cell_xyz;
        GH.nx = c->G.nx; GH.ny = c->G.ny; GH.nz = c->G.nz;
        HIPCHK(c, dalloc(&c->d_grains, 1));
        HIPCHK(c, copy_on(c->stream, c->d_grains, &GH, sizeof GH, hipMemcpyHostToDevice));
        c->pe_gdc = GH.gdc;
        c->pe_peh = GH.peh;
    }
    HIPCHK(c, copy_on(c->stream, c->d_dust_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c->stream, c->d_dust_mono, mono.data(), K * sizeof(int), hipMemcpyHostToDevice));
    if (!c->d_flags) {
        HIPCHK(c, dalloc(&c->d_flags, nc1));
        HIPCHK(c, dalloc(&c->d_nIter, nc1));
    }
    mcg::DustArgs &a = c->dust_args;
    a = mcg::DustArgs{};
    a.nx = c->G.nx; a.ny = c->G.ny; a.nz = c->G.nz; a.nb = nb; a.nc = c->G.nCells;
    a.nS = nS; a.nsMax = d->nSpeciesMax; a.nSizes = nSz; a.nTemps = nT;
    a.absx = c->d_dust_tab; a.em = c->d_dust_tab + (size_t)K * nb;
    a.weight = a.em + (size_t)K * nT;
    a.abun = a.weight + nSz;
    a.mono = c->d_dust_mono;
    a.Tdust = c->d_Tdust;
    a.XHILimit = d->XHILimit;
    // the dust add of the opacity build: per row the scattering and
    // absorption cross-sections, grainAbun * grainWeight (FP32, as the
    // reference multiplies them), the sublimation temperature; Ndust
    c->dust_opac = false;
    if (c->d_op_dust) { (void)hipFree(c->d_op_dust); c->d_op_dust = nullptr; }
    if (d->scaXsec) {
        if (!d->TdustSublime || !d->Ndust) return fail(c, MC_EINVAL, "scaXsec needs TdustSublime and Ndust");
        std::vector<float> t((size_t)2 * K * nb + 2 * K + nc1 + nSz + nS);
        float *sc = t.data(), *ab_ = sc + (size_t)K * nb, *dw = ab_ + (size_t)K * nb, *ts = dw + K, *nd = ts + K;
        float *gw = nd + nc1, *ga = gw + nSz;       // grainWeight, grainAbun (re-emission)
        for (int ai = 0; ai < nSz; ++ai) gw[ai] = d->grainWeight[ai];
        for (int s_ = 0; s_ < nS; ++s_) ga[s_] = d->grainAbun[s_];
        for (size_t i = 0; i < (size_t)K * nb; ++i) {
            if (!std::isfinite(d->scaXsec[i])) return fail(c, MC_ETABLE, "scaXsec entry %zu not finite", i);
            sc[i] = d->scaXsec[i];
            ab_[i] = d->absXsec[i];
        }
        for (int s_ = 0; s_ < nS; ++s_)
            for (int ai = 0; ai < nSz; ++ai) {
                dw[s_ * nSz + ai] = d->grainAbun[s_] * d->grainWeight[ai];
                ts[s_ * nSz + ai] = d->TdustSublime[s_];
            }
        for (size_t i = 0; i < nc1; ++i) {
            if (!std::isfinite(d->Ndust[i]) || d->Ndust[i] < 0.0f) return fail(c, MC_ETABLE, "Ndust(%zu) invalid", i);
            nd[i] = d->Ndust[i];
        }
        HIPCHK(c, dalloc(&c->d_op_dust, t.size()));
        HIPCHK(c, copy_on(c->stream, c->d_op_dust, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
        c->dust_opac = true;
        c->dust_rows = K;
    }
    c->dust = true;
    c->dust_pe = pe;
    return MC_OK;
}

// updateCell over every cell: the gas balance when mc_balance_setup was
// called, then the grain temperatures when mc_dust_setup was called
constexpr int kOrderBins = 1024;   // iteration counts above this share the first bin

// ionDen as the host holds it, (0:nCells, xref 1..nE, stage 1..ns) column
// major, to and from the balance's [cell][element index][stage 0..ns] rows
// (stage 0 unused, kept 0).  xr[e] = the host position of element index e.
struct IonMap {
    int xr[mcb::kMaxElem];              // balance_setup allows at most kMaxElem elements
};

static unsigned ion_blocks(size_t n)
{
    return (unsigned)std::min<size_t>((n + 255) / 256, 65536);
}

__global__ void ion_to_rows_kernel(const double *__restrict__ host, double *__restrict__ rows, long long nc1,
                                   int nE, int ns, IonMap m)
{
    const long long n = nc1 * nE;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long cell = t % nc1;
        const int e = (int)(t / nc1);
        double *r = rows + (cell * nE + e) * (ns + 1);
        r[0] = 0.0;
        for (int st = 1; st <= ns; ++st) r[st] = host[cell + nc1 * (m.xr[e] + (long long)nE * (st - 1))];
    }
}

__global__ void ion_from_rows_kernel(const double *__restrict__ rows, double *__restrict__ host, long long nc1,
                                     int nE, int ns, IonMap m)
{
    const long long n = nc1 * nE;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long cell = t % nc1;
        const int e = (int)(t / nc1);
        const double *r = rows + (cell * nE + e) * (ns + 1);
        for (int st = 1; st <= ns; ++st) host[cell + nc1 * (m.xr[e] + (long long)nE * (st - 1))] = r[st];
    }
}

static int update_cells(mc_ctx *c, double jscale, const mc_balance_opts *opts, const double *Hden, double *Te,
                        double *Ne, double *ionDen, double *Tdust, int32_t *flags, int32_t *nIter,
                        double *kernel_ms)
{
    const bool gas = c->balance, dust = c->dust;
    if (!gas && !dust) return fail(c, MC_ESTATE, "neither mc_balance_setup nor mc_dust_setup has been called");
    if (gas && (!Hden || !Te || !Ne || !ionDen)) return fail(c, MC_EINVAL, "null gas array");
    if (dust && !Tdust) return fail(c, MC_EINVAL, "null Tdust");
    if (!flags || !nIter) return fail(c, MC_EINVAL, "null array");
    if (!(jscale > 0.0) || std::isinf(jscale)) return fail(c, MC_EINVAL, "jscale must be finite and > 0");
    const int nc = c->G.nCells;
    const size_t nc1 = (size_t)nc + 1;
    mcb::CellArgs a{};
    bool warm = false;
    int nE = 0, ns = 0;
    IonMap imap{};
    if (gas) {
        a = cell_args(c, jscale);
        if (opts) {
            if (opts->solver != MC_SOLVER_REFERENCE && opts->solver != MC_SOLVER_SECANT)
                return fail(c, MC_EINVAL, "unknown solver %d", opts->solver);
            if (!(opts->ion_tol > 0.0 && opts->ion_tol < 1.0)) return fail(c, MC_EINVAL, "ion_tol outside (0,1)");
            if (opts->solver == MC_SOLVER_SECANT &&
                (!(opts->tol > 0.0 && opts->tol < 1.0) || opts->max_iter < 1 || opts->max_iter > 1000))
                return fail(c, MC_EINVAL, "secant solver needs 0 < tol < 1 and 1 <= max_iter <= 1000");
            a.solver = opts->solver;
            a.ion_tol = opts->ion_tol;
            if (opts->solver == MC_SOLVER_SECANT) {
                a.tol = opts->tol;
                a.max_iter = opts->max_iter;
                if (opts->warm_iter < 0 || opts->warm_iter > 1000)
                    return fail(c, MC_EINVAL, "warm_iter outside [0,1000]");
                a.warm_iter = opts->warm_iter ? opts->warm_iter : opts->max_iter;
                if (opts->first_step != 0.0) {
                    if (!(opts->first_step > 0.0 && opts->first_step <= 1.0))
                        return fail(c, MC_EINVAL, "first_step outside (0,1]");
                    a.step0 = opts->first_step;
                }
                if (opts->first_slope != 0.0) {
                    if (!(opts->first_slope >= 0.1 && opts->first_slope <= 100.0))
                        return fail(c, MC_EINVAL, "first_slope outside [0.1,100]");
                    a.slope0 = opts->first_slope;
                }
            }
            warm = opts->warm_start != 0;
            if (warm && !c->mlp) return fail(c, MC_ESTATE, "warm start needs mc_mlp_set");
            // a warm start is about a per cent off: its first step is that size
            if (warm && opts->solver == MC_SOLVER_SECANT && opts->first_step == 0.0) a.step0 = 0.01;
        }
        nE = c->bal.A.nE;
        ns = c->bal.A.nstages;
        for (int cell = 1; cell <= nc; ++cell)
            if (!(Te[cell] > 0.0) || !(Hden[cell] > 0.0) || !(Ne[cell] > 0.0))
                return fail(c, MC_EINVAL, "cell %d: Te, Ne and Hden must be > 0", cell);
    }
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nion = nc1 * nE * (ns + 1);
    const mcg::DustArgs &da0 = c->dust_args;
    const size_t ntd = dust ? nc1 * (da0.nsMax + 1) * (da0.nSizes + 1) : 0;
    if (gas) {
        // ionDen(0:nCells, xref, stage) goes up as the host holds it; a
        // kernel lays it out as [cell][element index][stage 0..ns]
        for (int ei = 0; ei < nE; ++ei) imap.xr[ei] = c->bal.eidx_xref[ei] - 1;
        if (!c->d_ion_host) HIPCHK(c, dalloc(&c->d_ion_host, nc1 * nE * ns));
        HIPCHK(c, hipMemcpyAsync(c->d_Hden, Hden, nc1 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_Te, Te, nc1 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_Ne, Ne, nc1 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_ion_host, ionDen, nc1 * nE * ns * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        hipLaunchKernelGGL(ion_to_rows_kernel, dim3(ion_blocks(nc1 * nE)), dim3(256), 0, c->stream, c->d_ion_host,
                           c->d_ion, (long long)nc1, nE, ns, imap);
        HIPCHK(c, hipGetLastError());
    }
    if (dust) HIPCHK(c, hipMemcpyAsync(c->d_Tdust, Tdust, ntd * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_flags, 0, nc1 * sizeof(int), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_nIter, 0, nc1 * sizeof(int), c->stream));
    if (warm) {
        if (!c->d_cold) HIPCHK(c, dalloc(&c->d_cold, 2 * nc1 + nion));
        int rc = mlp_rows(c);
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(c->d_cold, c->d_Te, nc1 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_cold + nc1, c->d_Ne, nc1 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_cold + 2 * nc1, c->d_ion, nion * sizeof(double), hipMemcpyDeviceToDevice,
                                 c->stream));
        a.Te0 = c->d_cold;
        a.Ne0 = c->d_cold + nc1;
        a.ion0 = c->d_cold + 2 * nc1;
    }
    if (gas && c->last_nit.size() == nc1) {
        // slowest first: a cell that took many iterations last time starts
        // early instead of holding the launch's tail (counting sort)
        int cnt[kOrderBins + 1] = {0};
        for (int cell = 1; cell <= nc; ++cell) ++cnt[std::min(std::max(c->last_nit[cell], 0), kOrderBins)];
        int pos[kOrderBins + 1];
        for (int k = kOrderBins, p = 0; k >= 0; --k) { pos[k] = p; p += cnt[k]; }
        c->order_h.resize(nc);
        for (int cell = 1; cell <= nc; ++cell)
            c->order_h[pos[std::min(std::max(c->last_nit[cell], 0), kOrderBins)]++] = cell;
        if (!c->d_order) HIPCHK(c, dalloc(&c->d_order, (size_t)nc));
        HIPCHK(c, hipMemcpyAsync(c->d_order, c->order_h.data(), (size_t)nc * sizeof(int), hipMemcpyHostToDevice,
                                 c->stream));
        a.order = c->d_order;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (gas) {
        int rc = integrals(c, a, warm ? c->d_mX : nullptr);
        if (rc) return rc;
        if (warm) {
            HIPCHK(c, mcm_standardise(c->d_mX, nc, mlp_scale(c), c->d_mX0, c->stream));
            HIPCHK(c, mcm_forward(c->d_mX0, nc, mlp_net(c), c->d_mY, mcm::kOutP, nullptr, 2 * c->cus, c->stream));
            HIPCHK(c, mcm_unscale(c->d_mY, nc, mcm::kOutP, mlp_scale(c), c->stream));
            HIPCHK(c, mcb_launch_warm(c->d_atomic, a, c->d_mY, mcm::kOutP, c->d_hit, c->stream));
        }
        if (dust && c->dust_pe) {
            a.grains = c->d_grains;
            a.grains_nb = c->G.nbins;
        }
        HIPCHK(c, mcb_launch_balance(c->d_atomic, a, c->stream));
        c->state_dev = true;
    }
    if (dust) {
        mcg::DustArgs da = da0;
        da.jscale = jscale;
        da.flags = gas ? nullptr : c->d_flags;      // grains alone: Tdust(0,0) decides convergence
        if (gas && c->dust_pe) {                    // the gas balance left gasDustColl_d, photoelHeat_d
            da.gdc = c->pe_gdc;
            da.peh = c->pe_peh;
        }
        HIPCHK(c, mcg_launch_dust(da, c->d_J3, c->d_active, c->stream));
        c->tdust_dev = true;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if (gas) {
        HIPCHK(c, hipMemcpyAsync(Te, c->d_Te, nc1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(Ne, c->d_Ne, nc1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        hipLaunchKernelGGL(ion_from_rows_kernel, dim3(ion_blocks(nc1 * nE)), dim3(256), 0, c->stream, c->d_ion,
                           c->d_ion_host, (long long)nc1, nE, ns, imap);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(ionDen, c->d_ion_host, nc1 * nE * ns * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
    }
    if (dust) HIPCHK(c, hipMemcpyAsync(Tdust, c->d_Tdust, ntd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(flags, c->d_flags, nc1 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(nIter, c->d_nIter, nc1 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (kernel_ms) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *kernel_ms = ms;
    }
    if (gas) c->last_nit.assign(nIter, nIter + nc1);
    return MC_OK;
}

extern "C" int mc_balance_run(mc_ctx *c, double jscale, const mc_balance_opts *opts, const double *Hden,
                              double *Te, double *Ne, double *ionDen, int32_t *flags, int32_t *nIter,
                              double *kernel_ms)
{
    if (!c) return MC_EINVAL;
    if (!c->balance) return fail(c, MC_ESTATE, "mc_balance_setup has not been called");
    if (c->dust && c->dust_pe)
        return fail(c, MC_ESTATE, "grains with the photoelectric effect are set up: the gas balance needs "
                                  "Tdust, call mc_update_cells");
    const bool dust = c->dust;
    c->dust = false;                                 // the gas alone
    const int rc = update_cells(c, jscale, opts, Hden, Te, Ne, ionDen, nullptr, flags, nIter, kernel_ms);
    c->dust = dust;
    return rc;
}

extern "C" int mc_update_cells(mc_ctx *c, double jscale, const mc_balance_opts *opts, const double *Hden,
                               double *Te, double *Ne, double *ionDen, double *Tdust, int32_t *flags,
                               int32_t *nIter, double *kernel_ms)
{
    if (!c) return MC_EINVAL;
    return update_cells(c, jscale, opts, Hden, Te, Ne, ionDen, Tdust, flags, nIter, kernel_ms);
}

// --------------------------------------------------------- opacity tables
extern "C" int mc_opacity_setup(mc_ctx *c, const mc_opacity_tables *t)
{
    if (!c) return MC_EINVAL;
    if (!t || !t->seg_meta || !t->seg_sigma || t->nseg < 1) return fail(c, MC_EINVAL, "empty segment tables");
    if (!c->balance) return fail(c, MC_ESTATE, "mc_balance_setup has not been called");
    if (!std::isfinite(t->brems1)) return fail(c, MC_ETABLE, "brems1 not finite");
    const mcb::Atomic &A = c->bal.A;
    const int nb = c->G.nbins, ns = A.nstages;
    // slots: (element, ion 1..min(Z, nstages)) in atomic-number order
    std::vector<int> kind, sl_e, sl_st, slot_of(31 * 10, -1);
    std::vector<float> sabun, efac;
    int he2 = -1;
    for (int e = 0; e < A.nE; ++e) {
        const int z = A.el[e].z;
        for (int st = 1; st <= std::min(z, ns); ++st) {
            slot_of[z * 10 + st] = (int)kind.size();
            if (z == 2 && st == 2) he2 = (int)kind.size();
            kind.push_back(z > 2 ? 1 : 0);
            sl_e.push_back(e);
            sl_st.push_back(st);
            sabun.push_back((float)A.el[e].abun);
            efac.push_back(z > 2 && st >= 2 ? (float)((st - 1) * (st - 1)) : 0.f);
        }
    }
    const int nslot = (int)kind.size();
    if (nslot > mco::kMaxSlots) return fail(c, MC_EUNSUP, "more than %d ion slots", mco::kMaxSlots);
    // entries of every bin, in segment order
    std::vector<std::vector<std::pair<int, float>>> bins(nb);
    size_t off = 0;
    for (int s = 0; s < t->nseg; ++s) {
        const int32_t *m = t->seg_meta + 4 * s;
        const int z = m[0], ion = m[1], lo = m[2], hi = m[3];
        if (z < 1 || z > 30 || ion < 1 || ion > 9 || slot_of[z * 10 + ion] < 0)
            return fail(c, MC_ETABLE, "segment %d: (Z=%d, ion %d) not among the ions switched on", s, z, ion);
        if (lo < 1 || hi < lo || hi > nb) return fail(c, MC_ETABLE, "segment %d: bins %d..%d outside 1..%d", s, lo, hi, nb);
        for (int i = lo; i <= hi; ++i) {
            const float v = t->seg_sigma[off + (size_t)(i - lo)];
            if (!std::isfinite(v)) return fail(c, MC_ETABLE, "segment %d: cross-section not finite", s);
            bins[i - 1].push_back({ slot_of[z * 10 + ion], v });
        }
        off += (size_t)(hi - lo + 1);
    }
    std::vector<int> bstart(nb + 1, 0);
    std::vector<mco::Ent> ent;
    for (int i = 0; i < nb; ++i) {
        bstart[i + 1] = bstart[i] + (int)bins[i].size();
        for (auto &p : bins[i]) ent.push_back(mco::Ent{ p.first * 2 + kind[p.first], p.second });
    }
    std::vector<int> iv;
    iv.insert(iv.end(), bstart.begin(), bstart.end());
    iv.insert(iv.end(), sl_e.begin(), sl_e.end());
    iv.insert(iv.end(), sl_st.begin(), sl_st.end());
    std::vector<float> fv;
    fv.insert(fv.end(), sabun.begin(), sabun.end());
    fv.insert(fv.end(), efac.begin(), efac.end());
    HIPCHK(c, hipSetDevice(c->device));
    c->opac = false;
    if (c->d_op_i) { (void)hipFree(c->d_op_i); c->d_op_i = nullptr; }
    if (c->d_op_f) { (void)hipFree(c->d_op_f); c->d_op_f = nullptr; }
    if (c->d_op_ent) { (void)hipFree(c->d_op_ent); c->d_op_ent = nullptr; }
    if (c->d_op_st) { (void)hipFree(c->d_op_st); c->d_op_st = nullptr; }
    HIPCHK(c, dalloc(&c->d_op_i, iv.size()));
    HIPCHK(c, dalloc(&c->d_op_f, fv.size()));
    HIPCHK(c, dalloc(&c->d_op_ent, ent.size()));
    HIPCHK(c, copy_on(c->stream, c->d_op_i, iv.data(), iv.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c->stream, c->d_op_f, fv.data(), fv.size() * sizeof(float), hipMemcpyHostToDevice));
    if (!ent.empty())
        HIPCHK(c, copy_on(c->stream, c->d_op_ent, ent.data(), ent.size() * sizeof(mco::Ent), hipMemcpyHostToDevice));
    const size_t nc1 = (size_t)c->G.nCells + 1;
    const int ldst = 3 + nslot;
    HIPCHK(c, dalloc(&c->d_op_st, nc1 * ldst));
    mco::OpacArgs &o = c->op;
    o = mco::OpacArgs{};
    o.ncell3 = c->G.ncell3; o.nb = nb; o.nc = c->G.nCells;
    o.js = c->G.js;
    o.active = c->d_active;
    o.st = c->d_op_st; o.ldst = ldst; o.nslot = nslot;
    o.bin_start = c->d_op_i;
    o.ent = c->d_op_ent;
    o.slot_abun = c->d_op_f;
    o.slot_efac = c->d_op_f + nslot;
    o.slot_he2 = he2;
    o.brems1 = t->brems1;
    o.nu1 = c->nuArray[0];
    o.xlf1 = log10f(c->nuArray[0]);
    c->op_slot_e = c->d_op_i + (nb + 1);
    c->op_slot_st = c->op_slot_e + nslot;
    c->op_sl_e = sl_e;
    c->op_sl_st = sl_st;
    c->opac = true;
    return MC_OK;
}

extern "C" int mc_opacity_build(mc_ctx *c, const float *Hden, const float *Te, const float *Ne, const float *ionDen,
                                const float *Tdust, float *opacity, float *scaOpac, double *kernel_ms)
{
    if (!c) return MC_EINVAL;
    if (!c->opac) return fail(c, MC_ESTATE, "mc_opacity_setup has not been called");
    const int nc = c->G.nCells, nb = c->G.nbins;
    const size_t nc1 = (size_t)nc + 1;
    mco::OpacArgs o = c->op;
    const int nslot = o.nslot, ldst = o.ldst;
    const bool host_state = Hden || Te || Ne || ionDen;
    if (host_state && !(Hden && Te && Ne && ionDen))
        return fail(c, MC_EINVAL, "Hden, Te, Ne and ionDen go together (or all NULL for the device state)");
    if (!host_state && !c->state_dev)
        return fail(c, MC_ESTATE, "no device state yet: pass the state or run mc_update_cells first");
    const bool dust = c->dust && c->dust_opac;
    if (dust && !Tdust && !c->tdust_dev)
        return fail(c, MC_ESTATE, "no device Tdust yet: pass Tdust or run mc_update_cells first");
    HIPCHK(c, hipSetDevice(c->device));
    if (host_state) {
        const int nE = c->bal.A.nE;
        std::vector<float> st(nc1 * ldst, 0.f);
        const std::vector<int> &sl_e = c->op_sl_e, &sl_st = c->op_sl_st;
        for (int cell = 1; cell <= nc; ++cell) {
            if (!(Te[cell] > 0.f) || !(Hden[cell] >= 0.f) || !std::isfinite(Ne[cell]))
                return fail(c, MC_EINVAL, "cell %d: Te must be > 0, Hden >= 0, Ne finite", cell);
        }
        for (size_t cell = 0; cell < nc1; ++cell) {
            float *r = st.data() + cell * ldst;
            r[0] = Te[cell]; r[1] = Ne[cell]; r[2] = Hden[cell];
            for (int k = 0; k < nslot; ++k) {
                const size_t xr = (size_t)c->bal.eidx_xref[sl_e[k]] - 1;
                r[3 + k] = ionDen[cell + nc1 * (xr + (size_t)nE * (sl_st[k] - 1))];
            }
        }
        HIPCHK(c, hipMemcpyAsync(c->d_op_st, st.data(), st.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (dust) {
        const mcg::DustArgs &da = c->dust_args;
        const size_t ntd = nc1 * (da.nsMax + 1) * (da.nSizes + 1);
        if (!c->d_op_Td) HIPCHK(c, dalloc(&c->d_op_Td, ntd));
        if (Tdust) HIPCHK(c, copy_on(c->stream, c->d_op_Td, Tdust, ntd * sizeof(float), hipMemcpyHostToDevice));
        const int K = c->dust_rows;
        o.ndrow = K; o.nSizes = da.nSizes; o.nsMax = da.nsMax;
        o.dsca = c->d_op_dust;
        o.dabs = c->d_op_dust + (size_t)K * nb;
        o.dw = c->d_op_dust + 2 * (size_t)K * nb;
        o.tsub = o.dw + K;
        o.ndust = o.tsub + K;
        o.Tdust = c->d_op_Td;
        if (!c->d_sca3) HIPCHK(c, dalloc(&c->d_sca3, mcd::j3_size(c->G.js, nb)));
        o.sca3 = c->d_sca3;
    }
    if (!c->d_opac3) HIPCHK(c, dalloc(&c->d_opac3, mcd::j3_size(c->G.js, nb)));
    o.opac3 = c->d_opac3;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (!host_state)
        HIPCHK(c, mco_launch_state(c->d_Te, c->d_Ne, c->d_Hden, c->d_ion, c->bal.A.nE, c->bal.A.nstages + 1,
                                   c->op_slot_e, c->op_slot_st, nslot, nc, c->d_op_st, ldst, c->stream));
    if (dust && !Tdust) {
        const mcg::DustArgs &da = c->dust_args;
        HIPCHK(c, mco_launch_f32(c->d_Tdust, nc1 * (da.nsMax + 1) * (da.nSizes + 1), c->d_op_Td, c->stream));
    }
    const size_t lds = (size_t)(nslot + o.ndrow) * 64 * sizeof(float);
    HIPCHK(c, mco_launch_opacity(o, lds, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if (opacity || scaOpac) {
        if (!c->d_op_out) HIPCHK(c, dalloc(&c->d_op_out, nc1 * nb));
        if (opacity) {
            HIPCHK(c, mco_launch_gather(c->d_opac3, c->G.js, c->d_cell_xyz, nc, nb, c->d_op_out, c->stream));
            HIPCHK(c, hipMemcpyAsync(opacity, c->d_op_out, nc1 * nb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        if (scaOpac) {
            if (dust) HIPCHK(c, mco_launch_gather(c->d_sca3, c->G.js, c->d_cell_xyz, nc, nb, c->d_op_out, c->stream));
            else HIPCHK(c, hipMemsetAsync(c->d_op_out, 0, nc1 * nb * sizeof(float), c->stream));
            HIPCHK(c, hipMemcpyAsync(scaOpac, c->d_op_out, nc1 * nb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (kernel_ms) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *kernel_ms = ms;
    }
    c->opac_built = true;
    c->opac_sca_built = dust;
    return MC_OK;
}

// ------------------------------------------------------- re-emission tables

// locate (interpolation_mod.f90:48-81) in the increasing frequency mesh
static int locate_nu(const std::vector<float> &xa, float x)
{
    const int n = (int)xa.size();
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    int p = 0;
    float bd = 0.f;
    for (int i = 0; i < n; ++i) {
        const float d = xa[i] - x;
        if (d > 0.f && (p == 0 || d < bd)) { p = i + 1; bd = d; }
    }
    return std::max(p - 1, 1);
}

extern "C" int mc_emission_setup(mc_ctx *c, const mc_emission_tables *t)
{
    if (!c) return MC_EINVAL;
    if (!t) return fail(c, MC_EINVAL, "no tables");
    if (!c->balance) return fail(c, MC_ESTATE, "mc_balance_setup has not been called");
    const mcb::Atomic &A = c->bal.A;
    const int nb = c->G.nbins;
    int eH = -1, eHe = -1;
    for (int e = 0; e < A.nE; ++e) {
        if (A.el[e].z == 1) eH = e;
        if (A.el[e].z == 2) eHe = e;
    }
    if (eH != 0 || eHe < 0 || A.nstages < 3) return fail(c, MC_EUNSUP, "re-emission needs H and He (3 stages)");
    if (mce_lds_bytes(nb, true) > 48 * 1024)
        return fail(c, MC_EUNSUP, "re-emission keeps 4 x nbins partial sums in LDS: %d bins is too many", nb);
    if (!t->tkGamma || !t->logGammaHI || !t->logGammaHeI || !t->logGammaHeII || !t->HINuEdgeP || !t->HeINuEdgeP ||
        !t->HeIINuEdgeP || !t->gauntFF || !t->gauntFFHeII || !t->xSecHI || !t->xSecHeI || !t->xSecHeII ||
        !t->y_dat || !t->Ay_dat || !t->HIRecLineData || !t->HeIIRecLineData || !t->HeIrecLineCoeff ||
        !t->HeIILymanData || !t->HeIILymanNu || !t->widFlx || (t->nfb > 0 && (!t->fb_meta || !t->fb_gratio || !t->fb_sigma)))
        return fail(c, MC_EINVAL, "a required table is NULL");
    const int nl3[3] = { t->nHI, t->nHeI, t->nHeII };
    if (t->ntk < 1) return fail(c, MC_ETABLE, "no gamma temperatures");
    for (int i = 0; i + 1 < t->ntk; ++i)
        if (!(t->tkGamma[i + 1] > t->tkGamma[i])) return fail(c, MC_ETABLE, "tkGamma not increasing");
    const int32_t *edges[3] = { t->HINuEdgeP, t->HeINuEdgeP, t->HeIINuEdgeP };
    const float *lg[3] = { t->logGammaHI, t->logGammaHeI, t->logGammaHeII };
    std::vector<int> pairs((size_t)3 * nb, -1);
    for (int s_ = 0; s_ < 3; ++s_) {
        if (nl3[s_] < 2 || nl3[s_] > mce::kMaxLim || nl3[s_] % 2) return fail(c, MC_ETABLE, "gamma table %d: %d edges", s_, nl3[s_]);
        for (size_t i = 0; i < (size_t)t->ntk * nl3[s_]; ++i)
            if (!std::isfinite(lg[s_][i])) return fail(c, MC_ETABLE, "log gamma table %d not finite", s_);
        for (int k = 0; k < nl3[s_] / 2; ++k) {
            const int lo = edges[s_][2 * k], hi = edges[s_][2 * k + 1];
            if (lo < 1 || hi > nb || hi < lo) return fail(c, MC_ETABLE, "gamma edge pair %d of series %d outside 1..%d", k, s_, nb);
            if (hi == lo) return fail(c, MC_ETABLE, "frequency grid too coarse: gamma edges %d of series %d coincide", k, s_);
            for (int j = lo; j <= hi - 1; ++j) pairs[(size_t)s_ * nb + j - 1] = k;
        }
    }
    const int ip[3] = { t->HIPnuP, t->HeIPnuP, t->HeIIPnuP };
    for (int s_ = 0; s_ < 3; ++s_)
        if (ip[s_] < 1 || ip[s_] > nb) return fail(c, MC_ETABLE, "ionisation bin %d of series %d outside 1..%d", ip[s_], s_, nb);
    const int i0[3] = { t->HIPnuP, std::max(t->HIPnuP, t->HeINuEdgeP[t->nHeI - 1]),
                        std::max(std::max(t->HIPnuP, t->HeINuEdgeP[t->nHeI - 1]), t->HeIINuEdgeP[t->nHeII - 1]) };
    const float *xs[3] = { t->xSecHI, t->xSecHeI, t->xSecHeII };
    // heavy-ion free-bound channels
    std::vector<int> fbi, fboff;
    std::vector<float> fbf;
    size_t nsig = 0;
    for (int k = 0; k < t->nfb; ++k) {
        const int32_t *m = t->fb_meta + 4 * k;
        const int z = m[0], ion = m[1], ipk = m[2], hik = m[3];
        int e = -1;
        for (int q = 0; q < A.nE; ++q)
            if (A.el[q].z == z) e = q;
        if (e < 0 || ion < 1 || ion + 1 > A.nstages) return fail(c, MC_ETABLE, "free-bound channel %d: (Z=%d, ion %d) not switched on", k, z, ion);
        if (ipk < 1 || hik > nb || hik < ipk) return fail(c, MC_ETABLE, "free-bound channel %d: bins %d..%d outside 1..%d", k, ipk, hik, nb);
        fbi.insert(fbi.end(), { e, ion + 1, ipk, hik });
        fboff.push_back((int)nsig);
        fbf.push_back(t->fb_gratio[k]);
        fbf.push_back((float)A.el[e].abun);
        nsig += (size_t)(hik - ipk);
    }
    for (size_t i = 0; i < nsig; ++i)
        if (!std::isfinite(t->fb_sigma[i])) return fail(c, MC_ETABLE, "free-bound cross-section %zu not finite", i);
    if (t->nHeIILyman < 4 || t->nHeIILyman > mce::kMaxLy) return fail(c, MC_ETABLE, "nHeIILyman %d outside 4..%d", t->nHeIILyman, mce::kMaxLy);
    // He I two-photon segments per bin (HeI2photSub's running index)
    std::vector<int> jHeI(nb, 0);
    {
        int j = 1;
        for (int i = 0; i < nb; ++i) {
            const float y = c->nuArray[i] / 1.514f;
            if (y > t->y_dat[j - 1]) j = j + 1;
            if (j >= 41) break;
            jHeI[i] = j;
        }
    }
    const int j2TsP = locate_nu(c->nuArray, 1.45673f);
    if (j2TsP < 1 || j2TsP > nb) return fail(c, MC_ETABLE, "1.45673 Ryd outside the frequency mesh");
    int lyP[mce::kMaxLy];
    for (int k = 0; k < t->nHeIILyman; ++k) {
        lyP[k] = locate_nu(c->nuArray, t->HeIILymanNu[k]);
        if (lyP[k] < 1 || lyP[k] > nb) return fail(c, MC_ETABLE, "He II Lyman line %d outside the frequency mesh", k + 1);
    }
    // device tables
    std::vector<int> iv(pairs);
    std::vector<size_t> io(8);
    io[0] = iv.size(); for (int s_ = 0; s_ < 3; ++s_) iv.insert(iv.end(), edges[s_], edges[s_] + nl3[s_]);
    io[1] = iv.size(); iv.insert(iv.end(), fbi.begin(), fbi.end());
    io[2] = iv.size(); iv.insert(iv.end(), fboff.begin(), fboff.end());
    io[3] = iv.size(); iv.insert(iv.end(), jHeI.begin(), jHeI.end());
    std::vector<float> fv;
    std::vector<size_t> fo(16);
    fo[0] = fv.size(); fv.insert(fv.end(), t->tkGamma, t->tkGamma + t->ntk);
    for (int s_ = 0; s_ < 3; ++s_) { fo[1 + s_] = fv.size(); fv.insert(fv.end(), lg[s_], lg[s_] + (size_t)t->ntk * nl3[s_]); }
    fo[4] = fv.size(); fv.insert(fv.end(), t->gauntFF, t->gauntFF + nb);
    fo[5] = fv.size(); fv.insert(fv.end(), t->gauntFFHeII, t->gauntFFHeII + nb);
    for (int s_ = 0; s_ < 3; ++s_) { fo[6 + s_] = fv.size(); fv.insert(fv.end(), xs[s_], xs[s_] + (nb - i0[s_] + 1)); }
    fo[9] = fv.size(); fv.insert(fv.end(), fbf.begin(), fbf.end());
    fo[10] = fv.size(); fv.insert(fv.end(), t->fb_sigma, t->fb_sigma + nsig);
    fo[11] = fv.size(); fv.insert(fv.end(), t->y_dat, t->y_dat + 41);
    fo[12] = fv.size(); fv.insert(fv.end(), t->Ay_dat, t->Ay_dat + 41);
    fo[13] = fv.size(); fv.insert(fv.end(), t->HeIrecLineCoeff, t->HeIrecLineCoeff + 34 * 3 * 4);
    fo[14] = fv.size(); fv.insert(fv.end(), t->widFlx, t->widFlx + nb);
    for (float v : fv)
        if (!std::isfinite(v)) return fail(c, MC_ETABLE, "a re-emission table holds a value that is not finite");
    std::vector<double> dv(t->HIRecLineData, t->HIRecLineData + 28 * 7);
    dv.insert(dv.end(), t->HeIIRecLineData, t->HeIIRecLineData + 28 * 15);
    HIPCHK(c, hipSetDevice(c->device));
    c->emis = false;
    if (c->d_em_i) { (void)hipFree(c->d_em_i); c->d_em_i = nullptr; }
    if (c->d_em_f) { (void)hipFree(c->d_em_f); c->d_em_f = nullptr; }
    if (c->d_em_d) { (void)hipFree(c->d_em_d); c->d_em_d = nullptr; }
    HIPCHK(c, dalloc(&c->d_em_i, iv.size()));
    HIPCHK(c, dalloc(&c->d_em_f, fv.size()));
    HIPCHK(c, dalloc(&c->d_em_d, dv.size()));
    HIPCHK(c, copy_on(c->stream, c->d_em_i, iv.data(), iv.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c->stream, c->d_em_f, fv.data(), fv.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c->stream, c->d_em_d, dv.data(), dv.size() * sizeof(double), hipMemcpyHostToDevice));
    mce::EmArgs &a = c->em;
    a = mce::EmArgs{};
    a.nb = nb; a.nc = c->G.nCells;
    a.nE = A.nE; a.ns1 = A.nstages + 1; a.eH = eH; a.eHe = eHe;
    a.abH = (float)A.el[eH].abun; a.abHe = (float)A.el[eHe].abun;
    a.nu = c->d_nuArray; a.widFlx = c->d_em_f + fo[14];
    a.ntk = t->ntk; a.tk = c->d_em_f + fo[0];
    size_t eo = io[0];
    for (int s_ = 0; s_ < 3; ++s_) {
        a.nlim[s_] = nl3[s_];
        a.lg[s_] = c->d_em_f + fo[1 + s_];
        a.edge[s_] = c->d_em_i + eo;
        eo += nl3[s_];
        a.pair[s_] = c->d_em_i + (size_t)s_ * nb;
        a.ip[s_] = ip[s_];
        a.i0[s_] = i0[s_];
        a.sig[s_] = c->d_em_f + fo[6 + s_];
    }
    a.gff = c->d_em_f + fo[4]; a.gffHeII = c->d_em_f + fo[5];
    a.nfb = t->nfb; a.fb_i = c->d_em_i + io[1]; a.fb_off = c->d_em_i + io[2];
    a.fb_f = c->d_em_f + fo[9]; a.fb_sig = c->d_em_f + fo[10];
    a.jHeI = c->d_em_i + io[3];
    a.ydat = c->d_em_f + fo[11]; a.Aydat = c->d_em_f + fo[12];
    a.j2TsP = j2TsP; a.nly = t->nHeIILyman;
    for (int k = 0; k < t->nHeIILyman; ++k) { a.lyP[k] = lyP[k]; a.lyData[k] = t->HeIILymanData[k]; }
    a.HIrec = c->d_em_d; a.HeIIrec = c->d_em_d + 28 * 7;
    a.HeIc = c->d_em_f + fo[13];
    c->emis = true;
    return MC_OK;
}

extern "C" int mc_emission_build(mc_ctx *c, const float *Hden, const float *Te, const float *Ne,
                                 const float *ionDen, const float *Tdust, float *recPDF, float *totalLines,
                                 double *kernel_ms)
{
    if (!c) return MC_EINVAL;
    if (!c->emis) return fail(c, MC_ESTATE, "mc_emission_setup has not been called");
    const int nc = c->G.nCells, nb = c->G.nbins;
    const size_t nc1 = (size_t)nc + 1, nt = nc1 * nb;
    const bool host_state = Hden || Te || Ne || ionDen;
    if (host_state && !(Hden && Te && Ne && ionDen))
        return fail(c, MC_EINVAL, "Hden, Te, Ne and ionDen go together (or all NULL for the device state)");
    if (!host_state && !c->state_dev)
        return fail(c, MC_ESTATE, "no device state yet: pass the state or run mc_update_cells first");
    if (c->dust && !c->dust_opac)
        return fail(c, MC_EUNSUP, "grains need TdustSublime and Ndust (mc_dust_tables.scaXsec) for their re-emission");
    const bool dust = c->dust && c->dust_opac;
    if (dust && !Tdust && !c->tdust_dev)
        return fail(c, MC_ESTATE, "no device Tdust yet: pass Tdust or run mc_update_cells first");
    HIPCHK(c, hipSetDevice(c->device));
    mce::EmArgs a = c->em;
    const mcb::Atomic &A = c->bal.A;
    const int nE = A.nE, ns = A.nstages;
    const size_t nion = nc1 * nE * (ns + 1);
    if (host_state) {
        for (int cell = 1; cell <= nc; ++cell)
            if (!(Te[cell] > 0.f) || !(Hden[cell] > 0.f) || !std::isfinite(Ne[cell]))
                return fail(c, MC_EINVAL, "cell %d: Te and Hden must be > 0, Ne finite", cell);
        std::vector<double> st(3 * nc1 + nion, 0.0);
        double *ion = st.data() + 3 * nc1;
        for (size_t cell = 0; cell < nc1; ++cell) {
            st[cell] = Te[cell];
            st[nc1 + cell] = Ne[cell];
            st[2 * nc1 + cell] = Hden[cell];
            for (int e = 0; e < nE; ++e) {
                const size_t xr = (size_t)c->bal.eidx_xref[e] - 1;
                for (int k = 1; k <= ns; ++k)
                    ion[(cell * nE + e) * (ns + 1) + k] = ionDen[cell + nc1 * (xr + (size_t)nE * (k - 1))];
            }
        }
        if (!c->d_em_st) HIPCHK(c, dalloc(&c->d_em_st, st.size()));
        HIPCHK(c, copy_on(c->stream, c->d_em_st, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice));
        a.Te = c->d_em_st; a.Ne = c->d_em_st + nc1; a.Hden = c->d_em_st + 2 * nc1; a.ion = c->d_em_st + 3 * nc1;
    } else {
        a.Te = c->d_Te; a.Ne = c->d_Ne; a.Hden = c->d_Hden; a.ion = c->d_ion;
    }
    if (dust) {
        const mcg::DustArgs &da = c->dust_args;
        const int K = c->dust_rows;
        a.dS = da.nS; a.dSizes = da.nSizes; a.dnsMax = da.nsMax;
        a.dabs = c->d_op_dust + (size_t)K * nb;
        a.dtsub = c->d_op_dust + 2 * (size_t)K * nb + K;
        a.ndust = a.dtsub + K;
        a.dweight = a.ndust + nc1;
        a.dgabun = a.dweight + da.nSizes;
        const size_t ntd = nc1 * (da.nsMax + 1) * (da.nSizes + 1);
        if (Tdust && !c->d_em_Td) HIPCHK(c, dalloc(&c->d_em_Td, ntd));
        if (Tdust) {
            std::vector<double> td(ntd);
            for (size_t i = 0; i < ntd; ++i) td[i] = Tdust[i];
            HIPCHK(c, copy_on(c->stream, c->d_em_Td, td.data(), ntd * sizeof(double), hipMemcpyHostToDevice));
            a.Tdust = c->d_em_Td;
        } else {
            a.Tdust = c->d_Tdust;
        }
    }
    // the rows are written into the back buffers; the tables the transport
    // samples change only when every row is good
    if (!c->d_recCDF_b) {
        HIPCHK(c, dalloc(&c->d_recCDF_b, nt));
        HIPCHK(c, fill_on(c->stream, c->d_recCDF_b, 0, nt * sizeof(float)));
    }
    if (!c->d_recGuide_b) HIPCHK(c, dalloc(&c->d_recGuide_b, nc1 * (mcd::kGuide + 1)));
    if (!c->d_totalLines_b) {
        HIPCHK(c, dalloc(&c->d_totalLines_b, nc1));
        HIPCHK(c, fill_on(c->stream, c->d_totalLines_b, 0, nc1 * sizeof(float)));
    }
    if (!c->d_em_err) HIPCHK(c, dalloc(&c->d_em_err, nc1));
    a.rec = c->d_recCDF_b; a.totalLines = c->d_totalLines_b; a.err = c->d_em_err;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, mce_launch(c->d_atomic, a, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, mcd_guides(c->d_recCDF_b, nc + 1, nb, c->d_recGuide_b, c->stream));
    std::vector<int> err(nc1);
    HIPCHK(c, hipMemcpyAsync(err.data(), c->d_em_err, nc1 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int cell = 1; cell <= nc; ++cell)
        if (err[cell] == 1)
            return fail(c, MC_ETABLE, "cell %d: the state is not finite (or Te <= 0); the previous "
                        "re-emission tables are kept", cell);
    std::swap(c->d_recCDF, c->d_recCDF_b);
    std::swap(c->d_recGuide, c->d_recGuide_b);
    std::swap(c->d_totalLines, c->d_totalLines_b);
    c->G.recCDF = c->d_recCDF; c->G.recGuide = c->d_recGuide; c->G.totalLines = c->d_totalLines;
    c->rec_built = true;
    if (kernel_ms) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *kernel_ms = ms;
    }
    if (recPDF) {
        std::vector<float> rec(nt);
        HIPCHK(c, copy_on(c->stream, rec.data(), c->d_recCDF, nt * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t cell = 1; cell < nc1; ++cell)
            for (int i = 0; i < nb; ++i) recPDF[cell + nc1 * i] = rec[cell * nb + i];
    }
    if (totalLines) {
        std::vector<float> tl(nc1);
        HIPCHK(c, copy_on(c->stream, tl.data(), c->d_totalLines, nc1 * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t cell = 1; cell < nc1; ++cell) totalLines[cell] = tl[cell];
    }
    return MC_OK;
}

// ------------------------------------------------------- line luminosities

extern "C" int mc_line_luminosities(mc_ctx *c, const float *Hden, const float *Te, const float *Ne,
                                    const float *ionDen, const int32_t *lgBlack, double *forbVol,
                                    double *HbetaVol, double *kernel_ms)
{
    if (!c) return MC_EINVAL;
    if (!c->balance) return fail(c, MC_ESTATE, "mc_balance_setup has not been called");
    if (!forbVol || !HbetaVol) return fail(c, MC_EINVAL, "forbVol and HbetaVol are required");
    const bool host_state = Hden || Te || Ne || ionDen;
    if (host_state && !(Hden && Te && Ne && ionDen))
        return fail(c, MC_EINVAL, "Hden, Te, Ne and ionDen go together (or all NULL for the device state)");
    if (!host_state && lgBlack)
        return fail(c, MC_EINVAL, "lgBlack goes with a host state (the device state uses its last update's flags)");
    if (!host_state && !c->state_dev)
        return fail(c, MC_ESTATE, "no device state yet: pass the state or run mc_update_cells first");
    const mcb::Atomic &A = c->bal.A;
    const int nc = c->G.nCells, nE = A.nE, ns = A.nstages, ncel = A.ncel;
    const size_t nc1 = (size_t)nc + 1, nion = nc1 * nE * (ns + 1);
    if (A.el[0].z != 1) return fail(c, MC_EUNSUP, "Hbeta needs hydrogen among the elements");
    int rc = cell_positions(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    mcb::LineArgs a{};
    a.ncells = nc;
    a.cells_per_blk = std::max(1, (nc + 4095) / 4096);
    const int nblk = (nc + a.cells_per_blk - 1) / a.cells_per_blk;
    a.ldp = ncel * mcb::kLineEnt + 1;
    a.cell_xyz = c->d_cell_xyz;
    a.nx = c->G.nx;
    a.ny = c->G.ny;
    a.dfac = c->d_dfac;
    if (host_state) {
        std::vector<double> st(3 * nc1 + nion, 0.0);
        for (int cell = 1; cell <= nc; ++cell)
            if (!(Te[cell] > 0.f) || !(Hden[cell] >= 0.f) || !std::isfinite(Ne[cell]))
                return fail(c, MC_EINVAL, "cell %d: Te must be > 0, Hden >= 0, Ne finite", cell);
        double *ion = st.data() + 3 * nc1;
        for (size_t cell = 0; cell < nc1; ++cell) {
            st[cell] = Te[cell];
            st[nc1 + cell] = Ne[cell];
            st[2 * nc1 + cell] = Hden[cell];
            for (int e = 0; e < nE; ++e) {
                const size_t xr = (size_t)c->bal.eidx_xref[e] - 1;
                for (int k = 1; k <= ns; ++k)
                    ion[(cell * nE + e) * (ns + 1) + k] = ionDen[cell + nc1 * (xr + (size_t)nE * (k - 1))];
            }
        }
        if (!c->d_ln_st) HIPCHK(c, dalloc(&c->d_ln_st, st.size()));
        HIPCHK(c, hipMemcpyAsync(c->d_ln_st, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        a.Te = c->d_ln_st;
        a.Ne = c->d_ln_st + nc1;
        a.Hden = c->d_ln_st + 2 * nc1;
        a.ion = c->d_ln_st + 3 * nc1;
        if (lgBlack) {
            std::vector<int> bl(nc1);
            for (size_t i = 0; i < nc1; ++i) bl[i] = lgBlack[i] >= 1 ? 1 : 0;
            if (!c->d_ln_black) HIPCHK(c, dalloc(&c->d_ln_black, nc1));
            HIPCHK(c, hipMemcpyAsync(c->d_ln_black, bl.data(), nc1 * sizeof(int), hipMemcpyHostToDevice, c->stream));
            a.black = c->d_ln_black;
            a.black_mask = 1;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        a.Te = c->d_Te;
        a.Ne = c->d_Ne;
        a.Hden = c->d_Hden;
        a.ion = c->d_ion;
        a.black = c->d_flags;             // bit 1: the last update found no Te or ion balance
        a.black_mask = 2;
    }
    const size_t npart = (size_t)nblk * a.ldp;
    if (c->ln_part_n < npart) {
        if (c->d_ln_part) { (void)hipFree(c->d_ln_part); c->d_ln_part = nullptr; }
        if (c->d_ln_out) { (void)hipFree(c->d_ln_out); c->d_ln_out = nullptr; }
        c->ln_part_n = 0;
        HIPCHK(c, dalloc(&c->d_ln_part, npart));
        HIPCHK(c, dalloc(&c->d_ln_out, (size_t)a.ldp));
        c->ln_part_n = npart;
    }
    a.part = c->d_ln_part;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, mcb_launch_lines(c->d_atomic, a, nblk, ncel, c->d_ln_out, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipMemcpyAsync(forbVol, c->d_ln_out, (size_t)ncel * mcb::kLineEnt * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(HbetaVol, c->d_ln_out + (size_t)ncel * mcb::kLineEnt, sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (kernel_ms) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *kernel_ms = ms;
    }
    return MC_OK;
}

// ------------------------------------------------------- warm-start network
extern "C" int mc_mlp_set(mc_ctx *c, const float *params, const float *scale)
{
    if (!c) return MC_EINVAL;
    if (!params || !scale) return fail(c, MC_EINVAL, "params and scale are required");
    for (int i = 0; i < mcm::kParams; ++i)
        if (!std::isfinite(params[i])) return fail(c, MC_EINVAL, "parameter %d not finite", i);
    for (int i = 0; i < MC_MLP_NSCALE; ++i)
        if (!std::isfinite(scale[i])) return fail(c, MC_EINVAL, "scale entry %d not finite", i);
    for (int i = 0; i < mcm::kIn; ++i)
        if (!(scale[mcm::kIn + i] > 0.f)) return fail(c, MC_EINVAL, "x_std[%d] must be > 0", i);
    for (int i = 0; i < mcm::kOut; ++i)
        if (!(scale[2 * mcm::kIn + mcm::kOut + i] > 0.f))
            return fail(c, MC_EINVAL, "y_std[%d] must be > 0", i);
    HIPCHK(c, hipSetDevice(c->device));
    const size_t P = mcm::kParams;
    if (!c->d_mlp) {
        HIPCHK(c, dalloc(&c->d_mlp, 4 * P + MC_MLP_NSCALE));
        HIPCHK(c, dalloc(&c->d_mlpW, P));
    }
    HIPCHK(c, copy_on(c->stream, c->d_mlp, params, P * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, fill_on(c->stream, c->d_mlp + P, 0, 3 * P * sizeof(float)));
    HIPCHK(c, copy_on(c->stream, c->d_mlp + 4 * P, scale, MC_MLP_NSCALE * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, mcm_publish(c->d_mlp, c->d_mlpW, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->mlp = true;
    c->mlp_step = 0;
    return MC_OK;
}

extern "C" int mc_mlp_get(mc_ctx *c, float *params, float *scale)
{
    if (!c) return MC_EINVAL;
    if (!c->mlp) return fail(c, MC_ESTATE, "mc_mlp_set has not been called");
    HIPCHK(c, hipSetDevice(c->device));
    if (params) HIPCHK(c, copy_on(c->stream, params, c->d_mlp, mcm::kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (scale)
        HIPCHK(c, copy_on(c->stream, scale, c->d_mlp + 4 * (size_t)mcm::kParams, MC_MLP_NSCALE * sizeof(float),
                            hipMemcpyDeviceToHost));
    return MC_OK;
}

extern "C" int mc_mlp_features(mc_ctx *c, double jscale, const double *Hden, float *X, int32_t *hit)
{
    if (!c) return MC_EINVAL;
    if (!c->balance) return fail(c, MC_ESTATE, "mc_balance_setup has not been called");
    if (!Hden || !X || !hit) return fail(c, MC_EINVAL, "null array");
    if (!(jscale > 0.0) || std::isinf(jscale)) return fail(c, MC_EINVAL, "jscale must be finite and > 0");
    const int nc = c->G.nCells;
    for (int cell = 1; cell <= nc; ++cell)
        if (!(Hden[cell] > 0.0)) return fail(c, MC_EINVAL, "cell %d: Hden must be > 0", cell);
    HIPCHK(c, hipSetDevice(c->device));
    int rc = mlp_rows(c);
    if (rc) return rc;
    const size_t nc1 = (size_t)nc + 1;
    HIPCHK(c, hipMemcpyAsync(c->d_Hden, Hden, nc1 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    mcb::CellArgs a = cell_args(c, jscale);
    rc = integrals(c, a, c->d_mX);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(X, c->d_mX, (size_t)nc * mcm::kIn * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hit, c->d_hit, nc1 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

// device copies of a host data set and the batch buffers of a training call
struct MlpBatch {
    float *X = nullptr, *T = nullptr, *M = nullptr, *Yb = nullptr, *Tb = nullptr, *Mb = nullptr;
    uint16_t *X0 = nullptr, *H = nullptr, *dA = nullptr, *dB = nullptr;
    int *idx = nullptr;
    double *sums = nullptr;
    void release()
    {
        void *b[] = { X, T, M, Yb, Tb, Mb, X0, H, dA, dB, idx, sums };
        for (void *p : b)
            if (p) (void)hipFree(p);
    }
};

static hipError_t batch_alloc(MlpBatch &B, int64_t n, int batch, int64_t nidx, int nsums)
{
    hipError_t e;
    if ((e = dalloc(&B.X, (size_t)n * mcm::kIn)) != hipSuccess) return e;
    if ((e = dalloc(&B.T, (size_t)n * mcm::kOut)) != hipSuccess) return e;
    if ((e = dalloc(&B.M, (size_t)n * mcm::kOut)) != hipSuccess) return e;
    if ((e = dalloc(&B.Yb, (size_t)batch * mcm::kOutP)) != hipSuccess) return e;
    if ((e = dalloc(&B.Tb, (size_t)batch * mcm::kOut)) != hipSuccess) return e;
    if ((e = dalloc(&B.Mb, (size_t)batch * mcm::kOut)) != hipSuccess) return e;
    if ((e = dalloc(&B.X0, (size_t)batch * mcm::kInP)) != hipSuccess) return e;
    if ((e = dalloc(&B.H, (size_t)3 * batch * mcm::kHid)) != hipSuccess) return e;
    if ((e = dalloc(&B.dA, (size_t)batch * mcm::kHid)) != hipSuccess) return e;
    if ((e = dalloc(&B.dB, (size_t)batch * mcm::kHid)) != hipSuccess) return e;
    if ((e = dalloc(&B.idx, (size_t)nidx)) != hipSuccess) return e;
    return dalloc(&B.sums, (size_t)nsums);
}

static int check_data(mc_ctx *c, const float *X, const float *T, const float *M, int64_t n)
{
    if (!c->mlp) return fail(c, MC_ESTATE, "mc_mlp_set has not been called");
    if (!X || !T || !M || n < 1 || n > (1LL << 26)) return fail(c, MC_EINVAL, "bad data set (n = %lld)", (long long)n);
    for (int64_t i = 0; i < n * mcm::kIn; ++i)
        if (!std::isfinite(X[i])) return fail(c, MC_EINVAL, "X entry %lld not finite", (long long)i);
    for (int64_t i = 0; i < n * mcm::kOut; ++i)
        if (!std::isfinite(T[i]) || !std::isfinite(M[i]) || M[i] < 0.f)
            return fail(c, MC_EINVAL, "T or M entry %lld not finite (or M < 0)", (long long)i);
    return MC_OK;
}

static hipError_t upload_data(mc_ctx *c, MlpBatch &B, const float *X, const float *T, const float *M, int64_t n)
{
    hipError_t e = hipMemcpyAsync(B.X, X, (size_t)n * mcm::kIn * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(B.T, T, (size_t)n * mcm::kOut * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(B.M, M, (size_t)n * mcm::kOut * sizeof(float), hipMemcpyHostToDevice, c->stream);
    return e;
}

// forward + backward of one batch (rows idx[0..nb)); sums must be zero
static hipError_t train_pass(mc_ctx *c, MlpBatch &B, const int *idx, int nb, double *sums)
{
    hipError_t e = mcm_gather(B.X, B.T, B.M, idx, nb, mlp_scale(c), B.X0, B.Tb, B.Mb, c->stream);
    if (e == hipSuccess) e = mcm_forward(B.X0, nb, mlp_net(c), B.Yb, mcm::kOutP, B.H, 2 * c->cus, c->stream);
    if (e == hipSuccess)
        e = mcm_backward(B.X0, B.H, B.Yb, B.Tb, B.Mb, nb, mlp_net(c), B.dA, B.dB,
                         c->d_mlp + mcm::kParams, sums, c->stream);
    return e;
}

extern "C" int mc_mlp_predict(mc_ctx *c, const float *X, int64_t n, float *Y)
{
    if (!c) return MC_EINVAL;
    if (!c->mlp) return fail(c, MC_ESTATE, "mc_mlp_set has not been called");
    if (!X || !Y || n < 1 || n > (1LL << 26)) return fail(c, MC_EINVAL, "bad rows (n = %lld)", (long long)n);
    HIPCHK(c, hipSetDevice(c->device));
    float *dX = nullptr, *dY = nullptr;
    uint16_t *dX0 = nullptr;
    hipError_t e = dalloc(&dX, (size_t)n * mcm::kIn);
    if (e == hipSuccess) e = dalloc(&dY, (size_t)n * mcm::kOut);
    if (e == hipSuccess) e = dalloc(&dX0, (size_t)n * mcm::kInP);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dX, X, (size_t)n * mcm::kIn * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = mcm_standardise(dX, (int)n, mlp_scale(c), dX0, c->stream);
    if (e == hipSuccess) e = mcm_forward(dX0, (int)n, mlp_net(c), dY, mcm::kOut, nullptr, 2 * c->cus, c->stream);
    if (e == hipSuccess) e = mcm_unscale(dY, (int)n, mcm::kOut, mlp_scale(c), c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(Y, dY, (size_t)n * mcm::kOut * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dX); (void)hipFree(dY); (void)hipFree(dX0);
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? MC_ENOMEM : MC_ENODEV, "mlp predict: %s", hipGetErrorString(e));
    return MC_OK;
}

extern "C" int mc_mlp_loss_grad(mc_ctx *c, const float *X, const float *T, const float *M, int64_t n,
                                float *grad, double *loss)
{
    if (!c) return MC_EINVAL;
    int rc = check_data(c, X, T, M, n);
    if (rc) return rc;
    if (!grad || n > 65536) return fail(c, MC_EINVAL, "gradient needs grad and at most 65536 rows");
    HIPCHK(c, hipSetDevice(c->device));
    MlpBatch B;
    std::vector<int> idx(n);
    for (int64_t i = 0; i < n; ++i) idx[i] = (int)i;
    double sums[2] = { 0.0, 0.0 };
    hipError_t e = batch_alloc(B, n, (int)n, n, 2);
    if (e == hipSuccess) e = upload_data(c, B, X, T, M, n);
    if (e == hipSuccess) e = hipMemcpyAsync(B.idx, idx.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(B.sums, 0, 2 * sizeof(double), c->stream);
    if (e == hipSuccess) e = train_pass(c, B, B.idx, (int)n, B.sums);
    if (e == hipSuccess)
        e = hipMemcpyAsync(grad, c->d_mlp + mcm::kParams, mcm::kParams * sizeof(float), hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sums, B.sums, sizeof sums, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    B.release();
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? MC_ENOMEM : MC_ENODEV, "mlp gradient: %s", hipGetErrorString(e));
    const double inv = sums[1] > 0.0 ? 1.0 / sums[1] : 0.0;
    for (int i = 0; i < mcm::kParams; ++i) grad[i] = (float)(grad[i] * inv);
    if (loss) *loss = sums[0] * inv;
    return MC_OK;
}

extern "C" int mc_mlp_train(mc_ctx *c, const float *X, const float *T, const float *M, int64_t n,
                            const int32_t *order, int32_t steps, int32_t batch, float lr, float *loss,
                            double *kernel_ms)
{
    if (!c) return MC_EINVAL;
    int rc = check_data(c, X, T, M, n);
    if (rc) return rc;
    if (!order || steps < 1 || batch < 1 || batch > 65536 || !(lr > 0.f) || (int64_t)steps * batch > (1LL << 28))
        return fail(c, MC_EINVAL, "bad training schedule");
    for (int64_t i = 0; i < (int64_t)steps * batch; ++i)
        if (order[i] < 0 || order[i] >= n) return fail(c, MC_EINVAL, "order[%lld] outside the data set", (long long)i);
    HIPCHK(c, hipSetDevice(c->device));
    MlpBatch B;
    const int64_t nidx = (int64_t)steps * batch;
    hipError_t e = batch_alloc(B, n, batch, nidx, 2 * steps);
    if (e == hipSuccess) e = upload_data(c, B, X, T, M, n);
    if (e == hipSuccess) e = hipMemcpyAsync(B.idx, order, nidx * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(B.sums, 0, 2 * steps * sizeof(double), c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    float *P = c->d_mlp, *G = P + mcm::kParams, *m1 = G + mcm::kParams, *m2 = m1 + mcm::kParams;
    for (int s = 0; s < steps && e == hipSuccess; ++s) {
        e = train_pass(c, B, B.idx + (size_t)s * batch, batch, B.sums + 2 * s);
        if (e == hipSuccess) e = mcm_adam(P, G, m1, m2, B.sums + 2 * s, ++c->mlp_step, lr, c->d_mlpW, c->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    std::vector<double> sums(2 * (size_t)steps);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sums.data(), B.sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    B.release();
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? MC_ENOMEM : MC_ENODEV, "mlp training: %s", hipGetErrorString(e));
    if (loss)
        for (int s = 0; s < steps; ++s) loss[s] = (float)(sums[2 * s + 1] > 0.0 ? sums[2 * s] / sums[2 * s + 1] : 0.0);
    if (kernel_ms) *kernel_ms = ms;
    return MC_OK;
}

// Grids the pool kernel takes (its packed cell indices and bins).
static bool pool_fits(const mcd::Grid &G)
{
    return G.nx <= 2047 && G.ny <= 2047 && G.nz <= 1023 && G.nbins <= 65535;
}

// The kernel a launch shape runs (MC_KERNEL_AUTO: two packets per lane at
// most 3 workgroups of 256 per CU, else the pool kernel where the grid
// allows it, else one packet per lane).
static int transport_kernel_of(const mc_tuning &t, const mcd::Grid &G)
{
    switch (t.kernel) {
    case MC_KERNEL_ONE: return MCD_KERNEL_ONE;
    case MC_KERNEL_SPARE: return MCD_KERNEL_SPARE;
    case MC_KERNEL_POOL: return MCD_KERNEL_POOL;
    default: return t.blocks_per_cu <= 3 ? MCD_KERNEL_SPARE : pool_fits(G) ? MCD_KERNEL_POOL : MCD_KERNEL_ONE;
    }
}

extern "C" int mc_set_tuning(mc_ctx *c, const mc_tuning *t)
{
    if (!c || !t) return MC_EINVAL;
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 32 || t->threads < 64 || t->threads > 256 ||
        t->threads % 64 || t->emit_min < 1 || t->emit_min > 64 || t->walk_steps < 1 ||
        t->walk_steps > 64 || t->kernel < MC_KERNEL_AUTO || t->kernel > MC_KERNEL_POOL)
        return fail(c, MC_EINVAL, "tuning out of range");
    if (t->kernel == MC_KERNEL_POOL && !pool_fits(c->G))
        return fail(c, MC_EUNSUP, "the pool kernel takes grids of at most 2047 x 2047 x 1023 cells and 65535 bins");
    c->tune = *t;
    return MC_OK;
}

extern "C" int mc_get_tuning(const mc_ctx *c, mc_tuning *t)
{
    if (!c || !t) return MC_EINVAL;
    *t = c->tune;
    return MC_OK;
}

static const char *packet_error_text(int code)
{
    switch (code) {
    case 10: return "getNu2: problem with random number generator";
    case 11: return "newPhotonPacket: insanity in stellar photon nuP assignment";
    case 13: return "newPhotonPacket: diffuse packet cannot be created in a no gas grid";
    case 14: return "newPhotonPacket: position of the new diffuse photon packet is inside the inner region";
    case 15: return "newPhotonPacket: insanity in diffuse photon nuP assignment";
    case 32: case 35: return "pathSegment: insanity (cell index outside the grid)";
    case 34: return "pathSegment: insane wall distance (NaN)";
    case 36: return "pathSegment: dS <= 0";
    case 37: return "pathSegment: insane position (NaN)";
    case 39: return "pathSegment: insanity at the scattering/absorption decision";
    case 41: return "dust-only re-emission (dustEmi) is outside the supported path";
    case 60: return "pathSegment: error in theta direction cosine assignment";
    case 61: return "pathSegment: error in Phi direction cosine assignment";
    default: return "packet fail-stop";
    }
}

extern "C" int mc_transport_run(mc_ctx *c, int32_t istar, int64_t npackets, uint64_t packet_offset,
                                double *J_out, double *esc_out, mc_counters *cnt)
{
    if (!c) return MC_EINVAL;
    if (!c->tables) return fail(c, MC_ESTATE, "mc_upload_tables has not been called");
    if (istar < 1 || istar > c->nStars) return fail(c, MC_EINVAL, "istar %d outside 1..%d", istar, c->nStars);
    if (npackets < 0) return fail(c, MC_EINVAL, "negative packet count");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemsetAsync(c->d_scratch, 0, (1 + kCounts) * sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_error, 0, sizeof(int), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_qphot, 0, sizeof(double), c->stream));
    mcd::RunArgs RA;
    RA.istar = istar;
    RA.npackets = npackets;
    RA.packet_offset = packet_offset;
    RA.next_packet = c->d_scratch;
    RA.counts = c->d_scratch + 1;
    RA.error = c->d_error;
    RA.qphot = c->d_qphot;
    RA.deltaE = c->deltaE[istar - 1];
    RA.px = c->starPos[3 * (istar - 1)]; RA.py = c->starPos[3 * (istar - 1) + 1];
    RA.pz = c->starPos[3 * (istar - 1) + 2];
    RA.sx = c->starIdx[3 * (istar - 1)]; RA.sy = c->starIdx[3 * (istar - 1) + 1];
    RA.sz = c->starIdx[3 * (istar - 1) + 2];
    RA.emit_min = c->tune.emit_min;
    RA.walk_steps = c->tune.walk_steps;
    const int blocks = c->tune.blocks_per_cu * c->cus * (256 / c->tune.threads);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (npackets > 0)
        HIPCHK(c, mcd_launch_transport(c->G, RA, blocks, c->tune.threads, transport_kernel_of(c->tune, c->G), c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long h[1 + kCounts];
    int herr = 0;
    double q = 0.0;
    HIPCHK(c, hipMemcpyAsync(h, c->d_scratch, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&herr, c->d_error, sizeof herr, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&q, c->d_qphot, sizeof q, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (cnt) fill_counters(*cnt, h, q, ms);
    if (herr) return fail(c, MC_EPACKET, "%s (code %d)", packet_error_text(herr), herr);
    if (J_out || esc_out) return mc_fetch_estimators(c, J_out, esc_out);
    return MC_OK;
}

// Arguments of a batched launch: the launch shape (tu), the LDS it needs and
// its chunk count.  Failures are reported on ctx[0].
static int batch_check(mc_ctx *const *ctx, int32_t n, const int32_t *istar, const int64_t *npackets,
                       const mc_batch_tuning *tuning, mc_batch_tuning &tu, size_t &lds, size_t &nchunks)
{
    if (!ctx || n < 1 || !ctx[0]) return fail(nullptr, MC_EINVAL, "no contexts");
    mc_ctx *c0 = ctx[0];
    if (!istar || !npackets) return fail(c0, MC_EINVAL, "istar and npackets are required");
    tu = { 16, 32, 8, 16384 };
    if (tuning) {
        if (tuning->waves_per_cu < 1 || tuning->waves_per_cu > 32 || tuning->emit_min < 1 ||
            tuning->emit_min > 64 || tuning->walk_steps < 1 || tuning->walk_steps > 64 || tuning->chunk < 64 ||
            tuning->chunk > (1LL << 24))
            return fail(c0, MC_EINVAL, "batch tuning out of range");
        tu = *tuning;
    }
    lds = 0;
    nchunks = 0;
    for (int i = 0; i < n; ++i) {
        mc_ctx *c = ctx[i];
        if (!c) return fail(c0, MC_EINVAL, "context %d is null", i);
        if (c->device != c0->device) return fail(c0, MC_EINVAL, "context %d is on another device", i);
        if (!c->tables) return fail(c0, MC_ESTATE, "context %d has no tables", i);
        if (istar[i] < 1 || istar[i] > c->nStars) return fail(c0, MC_EINVAL, "context %d: istar %d", i, istar[i]);
        if (npackets[i] < 0) return fail(c0, MC_EINVAL, "context %d: negative packet count", i);
        for (int j = 0; j < i; ++j)
            if (ctx[j] == c) return fail(c0, MC_EINVAL, "context %d appears twice", i);
        lds = std::max(lds, mcd_lds_bytes(c->G.nx, c->G.ny, c->G.nz, c->G.nbins));
        nchunks += (size_t)((npackets[i] + tu.chunk - 1) / tu.chunk);
    }
    if (nchunks > (1u << 30)) return fail(c0, MC_EINVAL, "too many chunks");
    return MC_OK;
}

// The run arguments of model i of a batched launch (counters set by the caller).
static void batch_run_args(const mc_ctx *c, int32_t istar, int64_t npackets, uint64_t offset,
                           const mc_batch_tuning &tu, mcd::RunArgs &R)
{
    const int s = istar - 1;
    R.istar = istar;
    R.npackets = npackets;
    R.packet_offset = offset;
    R.deltaE = c->deltaE[s];
    R.px = c->starPos[3 * s]; R.py = c->starPos[3 * s + 1]; R.pz = c->starPos[3 * s + 2];
    R.sx = c->starIdx[3 * s]; R.sy = c->starIdx[3 * s + 1]; R.sz = c->starIdx[3 * s + 2];
    R.emit_min = tu.emit_min;
    R.walk_steps = tu.walk_steps;
}

extern "C" int mc_transport_batch(mc_ctx *const *ctx, int32_t n, const int32_t *istar, const int64_t *npackets,
                                  const uint64_t *packet_offset, const mc_batch_tuning *tuning, mc_counters *cnt,
                                  double *kernel_ms)
{
    mc_batch_tuning tu;
    size_t lds = 0, nchunks = 0;
    int rc0 = batch_check(ctx, n, istar, npackets, tuning, tu, lds, nchunks);
    if (rc0) return rc0;
    mc_ctx *c0 = ctx[0];
    HIPCHK(c0, hipSetDevice(c0->device));
    // order every context's earlier work before the launch
    for (int i = 0; i < n; ++i) HIPCHK(c0, hipStreamSynchronize(ctx[i]->stream));
    std::vector<mcd::Grid> G(n);
    std::vector<mcd::RunArgs> RA(n);
    std::vector<int> cm;
    std::vector<long long> cf;
    std::vector<int> cc;
    cm.reserve(nchunks); cf.reserve(nchunks); cc.reserve(nchunks);
    for (int i = 0; i < n; ++i) {
        mc_ctx *c = ctx[i];
        HIPCHK(c0, hipMemsetAsync(c->d_scratch, 0, (1 + kCounts) * sizeof(unsigned long long), c0->stream));
        HIPCHK(c0, hipMemsetAsync(c->d_error, 0, sizeof(int), c0->stream));
        HIPCHK(c0, hipMemsetAsync(c->d_qphot, 0, sizeof(double), c0->stream));
        G[i] = c->G;
        mcd::RunArgs &R = RA[i];
        batch_run_args(c, istar[i], npackets[i], packet_offset ? packet_offset[i] : 0, tu, R);
        R.next_packet = c->d_scratch;
        R.counts = c->d_scratch + 1;
        R.error = c->d_error;
        R.qphot = c->d_qphot;
        for (long long f = 0; f < npackets[i]; f += tu.chunk) {
            cm.push_back(i);
            cf.push_back(f);
            cc.push_back((int)std::min<long long>(tu.chunk, npackets[i] - f));
        }
    }
    // one device block: grids | run arguments | chunk table | chunk counter
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t oG = 0, oR = up(n * sizeof(mcd::Grid)), oM = oR + up(n * sizeof(mcd::RunArgs));
    const size_t oF = oM + up(cm.size() * sizeof(int)), oC = oF + up(cf.size() * sizeof(long long));
    const size_t oN = oC + up(cc.size() * sizeof(int)), tot = oN + 256;
    char *buf = nullptr;
    HIPCHK(c0, dalloc(&buf, tot));
    hipError_t e = hipMemcpyAsync(buf + oG, G.data(), n * sizeof(mcd::Grid), hipMemcpyHostToDevice, c0->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(buf + oR, RA.data(), n * sizeof(mcd::RunArgs), hipMemcpyHostToDevice, c0->stream);
    if (e == hipSuccess && !cm.empty())
        e = hipMemcpyAsync(buf + oM, cm.data(), cm.size() * sizeof(int), hipMemcpyHostToDevice, c0->stream);
    if (e == hipSuccess && !cf.empty())
        e = hipMemcpyAsync(buf + oF, cf.data(), cf.size() * sizeof(long long), hipMemcpyHostToDevice, c0->stream);
    if (e == hipSuccess && !cc.empty())
        e = hipMemcpyAsync(buf + oC, cc.data(), cc.size() * sizeof(int), hipMemcpyHostToDevice, c0->stream);
    if (e == hipSuccess) e = hipMemsetAsync(buf + oN, 0, sizeof(unsigned), c0->stream);
    mcd::BatchArgs B;
    B.G = reinterpret_cast<const mcd::Grid *>(buf + oG);
    B.RA = reinterpret_cast<const mcd::RunArgs *>(buf + oR);
    B.chunk_model = reinterpret_cast<const int *>(buf + oM);
    B.chunk_first = reinterpret_cast<const long long *>(buf + oF);
    B.chunk_count = reinterpret_cast<const int *>(buf + oC);
    B.nchunks = (int)cm.size();
    B.next_chunk = reinterpret_cast<unsigned *>(buf + oN);
    B.emit_min = tu.emit_min;
    B.walk_steps = tu.walk_steps;
    if (e == hipSuccess) e = hipEventRecord(c0->ev0, c0->stream);
    if (e == hipSuccess && B.nchunks > 0) e = mcd_launch_batch(B, lds, tu.waves_per_cu * c0->cus, c0->stream);
    if (e == hipSuccess) e = hipEventRecord(c0->ev1, c0->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c0->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c0->ev0, c0->ev1);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return fail(c0, e == hipErrorOutOfMemory ? MC_ENOMEM : MC_ENODEV, "batch: %s", hipGetErrorString(e));
    if (kernel_ms) *kernel_ms = ms;
    int rc = MC_OK;
    for (int i = 0; i < n; ++i) {
        mc_ctx *c = ctx[i];
        unsigned long long h[1 + kCounts];
        int herr = 0;
        double q = 0.0;
        HIPCHK(c0, copy_on(c0->stream, h, c->d_scratch, sizeof h, hipMemcpyDeviceToHost));
        HIPCHK(c0, copy_on(c0->stream, &herr, c->d_error, sizeof herr, hipMemcpyDeviceToHost));
        HIPCHK(c0, copy_on(c0->stream, &q, c->d_qphot, sizeof q, hipMemcpyDeviceToHost));
        if (cnt) fill_counters(cnt[i], h, q, ms);
        if (herr && rc == MC_OK) {
            fail(c, MC_EPACKET, "%s (code %d)", packet_error_text(herr), herr);
            rc = fail(c0, MC_EPACKET, "context %d: %s (code %d)", i, packet_error_text(herr), herr);
        }
    }
    return rc;
}

// ------------------------------------------ batched launch, no allocation
// mc_transport_batch_noalloc: mc_transport_batch for callers that cannot
// allocate device memory.  The launch's working block (the models' grids,
// run arguments and counters, the chunk table and the chunk counter) lives
// in the file-scope device array `scratch` (kNoAllocBytes on every GPU the
// library runs on, held until the process ends) instead of a per-call
// allocation.  Its device address is taken once per device when the first
// context on that device is created (noalloc_scratch_init) and only that
// address is handed to kernels and copies.  Every call on one GPU reuses the
// array, which keeps the last call's contents: the counters and the chunk
// counter are cleared by a kernel at the start of each call, and calls that
// overlap on one GPU race.
constexpr size_t kNoAllocBytes = 8u << 20;
constexpr int kNoAllocModels = 256;
constexpr int kNoAllocDevices = 64;
constexpr int kCounterWords = 16;       // next_packet, counts[kCounts], error, qphot per model
static_assert(kCounterWords >= 1 + kCounts + 2, "counter block too small");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kNoAllocDevices];

// With c's device current: record the device address of `scratch` on it.
static int noalloc_scratch_init(mc_ctx *c)
{
    if (c->device < 0 || c->device >= kNoAllocDevices) return MC_OK;   // the launch refuses such a device
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return MC_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(c, MC_ENODEV, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[c->device] = static_cast<unsigned char *>(p);
    return MC_OK;
}

namespace {

// chunk k of a batched launch: its model (the last i with cstart[i] <= k),
// first packet and count, from the models' chunk prefix sums
__global__ void noalloc_chunks_kernel(int n, const long long *cstart, const long long *npk, long long chunk,
                                      int *cm, long long *cf, int *cc, long long nchunks)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nchunks) return;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cstart[mid] <= k) lo = mid;
        else hi = mid;
    }
    const long long f = (k - cstart[lo]) * chunk;
    const long long left = npk[lo] - f;
    cm[k] = lo;
    cf[k] = f;
    cc[k] = (int)(left < chunk ? left : chunk);
}

__global__ void noalloc_clear_kernel(unsigned long long *w, long long nwords)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
         i += (long long)gridDim.x * blockDim.x)
        w[i] = 0ull;
}

}  // namespace

extern "C" int mc_transport_batch_noalloc(mc_ctx *const *ctx, int32_t n, const int32_t *istar,
                                          const int64_t *npackets, const uint64_t *packet_offset,
                                          const mc_batch_tuning *tuning, mc_counters *cnt, double *kernel_ms)
{
    mc_batch_tuning tu;
    size_t lds = 0, nchunks = 0;
    int rc0 = batch_check(ctx, n, istar, npackets, tuning, tu, lds, nchunks);
    if (rc0) return rc0;
    mc_ctx *c0 = ctx[0];
    if (n > kNoAllocModels)
        return fail(c0, MC_EINVAL, "%d models: the allocation-free launch takes at most %d", n, kNoAllocModels);
    if (c0->device < 0 || c0->device >= kNoAllocDevices)
        return fail(c0, MC_EUNSUP, "device %d: the allocation-free launch serves devices 0..%d", c0->device,
                    kNoAllocDevices - 1);
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch_ptr[c0->device];
    }
    if (!base) return fail(c0, MC_ESTATE, "no scratch address for device %d", c0->device);
    // layout: grids | run arguments | chunk prefix | packets | counters | chunk counter | chunk table
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t oG = 0, oR = up(n * sizeof(mcd::Grid)), oS = oR + up(n * sizeof(mcd::RunArgs));
    const size_t oP = oS + up((n + 1) * sizeof(long long)), oK = oP + up(n * sizeof(long long));
    const size_t oN = oK + up((size_t)n * kCounterWords * sizeof(unsigned long long));
    const size_t oM = oN + 256, oF = oM + up(nchunks * sizeof(int)), oC = oF + up(nchunks * sizeof(long long));
    const size_t tot = oC + up(nchunks * sizeof(int));
    if (tot > kNoAllocBytes)
        return fail(c0, MC_EINVAL, "the launch needs %zu bytes of scratch (%zu chunks of %lld packets); the "
                    "allocation-free launch holds %zu: use larger chunks", tot, nchunks, (long long)tu.chunk,
                    kNoAllocBytes);
    HIPCHK(c0, hipSetDevice(c0->device));
    for (int i = 0; i < n; ++i) HIPCHK(c0, hipStreamSynchronize(ctx[i]->stream));
    // host staging in thread-local arrays (no heap)
    static thread_local mcd::Grid tG[kNoAllocModels];
    static thread_local mcd::RunArgs tR[kNoAllocModels];
    static thread_local long long tS[kNoAllocModels + 1], tP[kNoAllocModels];
    static thread_local unsigned long long tK[kNoAllocModels * kCounterWords];
    unsigned long long *dK = reinterpret_cast<unsigned long long *>(base + oK);
    tS[0] = 0;
    for (int i = 0; i < n; ++i) {
        mc_ctx *c = ctx[i];
        tG[i] = c->G;
        mcd::RunArgs &R = tR[i];
        batch_run_args(c, istar[i], npackets[i], packet_offset ? packet_offset[i] : 0, tu, R);
        unsigned long long *k = dK + (size_t)i * kCounterWords;
        R.next_packet = k;
        R.counts = k + 1;
        R.error = reinterpret_cast<int *>(k + 1 + kCounts);
        R.qphot = reinterpret_cast<double *>(k + 2 + kCounts);
        tP[i] = npackets[i];
        tS[i + 1] = tS[i] + (npackets[i] + tu.chunk - 1) / tu.chunk;
    }
    hipStream_t st = c0->stream;
    HIPCHK(c0, hipMemcpyAsync(base + oG, tG, n * sizeof(mcd::Grid), hipMemcpyHostToDevice, st));
    HIPCHK(c0, hipMemcpyAsync(base + oR, tR, n * sizeof(mcd::RunArgs), hipMemcpyHostToDevice, st));
    HIPCHK(c0, hipMemcpyAsync(base + oS, tS, (n + 1) * sizeof(long long), hipMemcpyHostToDevice, st));
    HIPCHK(c0, hipMemcpyAsync(base + oP, tP, n * sizeof(long long), hipMemcpyHostToDevice, st));
    // counters and the chunk counter start at zero (a kernel, not a memset)
    const long long nclear = (long long)(oM - oK) / (long long)sizeof(unsigned long long);
    hipLaunchKernelGGL(noalloc_clear_kernel, dim3(64), dim3(256), 0, st,
                       reinterpret_cast<unsigned long long *>(base + oK), nclear);
    HIPCHK(c0, hipGetLastError());
    int *cm = reinterpret_cast<int *>(base + oM);
    long long *cf = reinterpret_cast<long long *>(base + oF);
    int *cc = reinterpret_cast<int *>(base + oC);
    if (nchunks > 0) {
        const unsigned blocks = (unsigned)((nchunks + 255) / 256);
        hipLaunchKernelGGL(noalloc_chunks_kernel, dim3(blocks), dim3(256), 0, st, (int)n,
                           reinterpret_cast<const long long *>(base + oS), reinterpret_cast<const long long *>(base + oP),
                           (long long)tu.chunk, cm, cf, cc, (long long)nchunks);
        HIPCHK(c0, hipGetLastError());
    }
    mcd::BatchArgs B;
    B.G = reinterpret_cast<const mcd::Grid *>(base + oG);
    B.RA = reinterpret_cast<const mcd::RunArgs *>(base + oR);
    B.chunk_model = cm;
    B.chunk_first = cf;
    B.chunk_count = cc;
    B.nchunks = (int)nchunks;
    B.next_chunk = reinterpret_cast<unsigned *>(base + oN);
    B.emit_min = tu.emit_min;
    B.walk_steps = tu.walk_steps;
    HIPCHK(c0, hipEventRecord(c0->ev0, st));
    if (B.nchunks > 0) HIPCHK(c0, mcd_launch_batch(B, lds, tu.waves_per_cu * c0->cus, st));
    HIPCHK(c0, hipEventRecord(c0->ev1, st));
    HIPCHK(c0, hipMemcpyAsync(tK, dK, (size_t)n * kCounterWords * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, st));
    HIPCHK(c0, hipStreamSynchronize(st));
    float ms = 0.f;
    HIPCHK(c0, hipEventElapsedTime(&ms, c0->ev0, c0->ev1));
    if (kernel_ms) *kernel_ms = ms;
    int rc = MC_OK;
    for (int i = 0; i < n; ++i) {
        const unsigned long long *h = tK + (size_t)i * kCounterWords;
        int herr = 0;
        double q = 0.0;
        std::memcpy(&herr, h + 1 + kCounts, sizeof herr);
        std::memcpy(&q, h + 2 + kCounts, sizeof q);
        if (cnt) fill_counters(cnt[i], h, q, ms);
        if (herr && rc == MC_OK) {
            fail(ctx[i], MC_EPACKET, "%s (code %d)", packet_error_text(herr), herr);
            rc = fail(c0, MC_EPACKET, "context %d: %s (code %d)", i, packet_error_text(herr), herr);
        }
    }
    return rc;
}

extern "C" int mc_walk_replay(mc_ctx *c, int32_t n, const float *pos, const float *dir,
                              const int32_t *cell, int32_t *cell_seq, float *ds_seq,
                              int32_t *seq_len, int32_t *status, int32_t max_len)
{
    if (!c) return MC_EINVAL;
    if (n < 0 || max_len < 1 || (n > 0 && (!pos || !dir || !cell || !cell_seq || !ds_seq || !seq_len || !status)))
        return fail(c, MC_EINVAL, "bad replay arguments");
    if (n == 0) return MC_OK;
    for (int i = 0; i < n; ++i) {
        const int *ce = cell + 3 * i;
        if (ce[0] < 1 || ce[0] > c->G.nx || ce[1] < 1 || ce[1] > c->G.ny || ce[2] < 1 || ce[2] > c->G.nz)
            return fail(c, MC_EINVAL, "replay packet %d starts outside the grid", i);
    }
    HIPCHK(c, hipSetDevice(c->device));
    float *dp = nullptr, *dd = nullptr, *dds = nullptr;
    int *dc = nullptr, *dseq = nullptr, *dlen = nullptr, *dst = nullptr;
    const size_t tot = (size_t)n * max_len;
    int rc = MC_OK;
    hipError_t e = hipSuccess;
    if ((e = dalloc(&dp, 3 * (size_t)n)) == hipSuccess && (e = dalloc(&dd, 3 * (size_t)n)) == hipSuccess &&
        (e = dalloc(&dc, 3 * (size_t)n)) == hipSuccess && (e = dalloc(&dseq, 3 * tot)) == hipSuccess &&
        (e = dalloc(&dds, tot)) == hipSuccess && (e = dalloc(&dlen, (size_t)n)) == hipSuccess &&
        (e = dalloc(&dst, (size_t)n)) == hipSuccess) {
        e = copy_on(c->stream, dp, pos, 3 * n * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = copy_on(c->stream, dd, dir, 3 * n * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = copy_on(c->stream, dc, cell, 3 * n * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemsetAsync(dseq, 0, 3 * tot * sizeof(int), c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(dds, 0, tot * sizeof(float), c->stream);
        if (e == hipSuccess)
            e = mcd_launch_replay(c->G, n, dp, dd, dc, dseq, dds, dlen, dst, max_len, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = copy_on(c->stream, cell_seq, dseq, 3 * tot * sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = copy_on(c->stream, ds_seq, dds, tot * sizeof(float), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = copy_on(c->stream, seq_len, dlen, n * sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = copy_on(c->stream, status, dst, n * sizeof(int), hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) rc = fail(c, e == hipErrorOutOfMemory ? MC_ENOMEM : MC_ENODEV, "replay: %s", hipGetErrorString(e));
    (void)hipFree(dp); (void)hipFree(dd); (void)hipFree(dc); (void)hipFree(dseq);
    (void)hipFree(dds); (void)hipFree(dlen); (void)hipFree(dst);
    return rc;
}

// the message of a call that has no context (grid files)
__attribute__((visibility("hidden"))) int mc_fail_global(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_create_err, 512, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *mc_ctx_last_error(const mc_ctx *c)
{
    return c ? c->err : g_create_err;
}

extern "C" void mc_ctx_destroy(mc_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    void *bufs[] = { c->d_axis, c->d_mid, c->d_dfac, c->d_active, c->d_opac3, c->d_sca3,
                     c->d_recCDF, c->d_recGuide, c->d_totalLines, c->d_specCDF, c->d_nuArray,
                     c->d_gSca, c->d_J3, c->d_J, c->d_esc, c->d_view, c->d_scratch, c->d_error, c->d_qphot,
                     c->d_cell_xyz, c->d_atomic, c->d_bal_tab, c->d_Hden, c->d_Te, c->d_Ne,
                     c->d_ion, c->d_flags, c->d_nIter, c->d_cold, c->d_mlp, c->d_mlpW, c->d_mX,
                     c->d_mY, c->d_mX0, c->d_hit, c->d_cross, c->d_Rint, c->d_Hint,
                     c->d_dust_tab, c->d_Tdust, c->d_dust_mono, c->d_pe_tab, c->d_grains, c->d_op_dust,
                     c->d_op_i, c->d_op_f, c->d_op_st, c->d_op_Td, c->d_op_out, c->d_op_ent,
                     c->d_ln_part, c->d_ln_st, c->d_ln_out, c->d_ln_black, c->d_em_i, c->d_em_err,
                     c->d_em_f, c->d_em_d, c->d_em_st, c->d_em_Td, c->d_recCDF_b, c->d_recGuide_b,
                     c->d_totalLines_b, c->d_order, c->d_ion_host };
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

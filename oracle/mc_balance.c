/*
 * mc_balance.c - CPU restatement of MOCASSIN's per-cell balance (updateCell).
 * TEST INFRASTRUCTURE ONLY: built into oracle/liboracle_mc.so next to the
 * transport restatement and used as the checker of the GPU balance kernel.
 * mc_balance_impl.h holds the restatement; it is compiled twice, in the
 * reference's own FP32 (mco_balance_cell_f32) and in FP64
 * (mco_balance_cell_f64, the FP64 oracle).  Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "mc_oracle.h"

int mco_atomic_init(mco_atomic *A)
{
    memset(A->elemOn, 0, sizeof A->elemOn);
    for (int z = 1; z <= 30; ++z) A->elemOn[z] = A->elem_on_in[z - 1];
    for (int z = 1; z <= 30; ++z) A->xref[z] = A->xref_in[z - 1];
    for (int z = 1; z <= 30; ++z) A->abun[z] = A->abun_in[z - 1];
    memset(A->chan_g, 0, sizeof A->chan_g);
    memset(A->chan_ip, 0, sizeof A->chan_ip);
    for (int k = 0; k < A->nch; ++k) {
        const int32_t *m = A->ch_meta + 8 * k;
        if (m[0] < 1 || m[0] > 30 || m[1] < 1 || m[1] > 9) return 1;
        A->chan_g[2 * (m[0] * 10 + m[1])] = m[2];
        A->chan_g[2 * (m[0] * 10 + m[1]) + 1] = m[3];
        A->chan_ip[m[0] * 10 + m[1]] = m[4];
    }
    if (A->ncel > MCO_MAX_CEL) return 2;
    size_t od = 0, of = 0;
    for (int k = 0; k < A->ncel; ++k) {
        const int32_t *m = A->cel_meta + 6 * k;
        const int nl = m[4], nt = m[5];
        if (nl < 1 || nl > 17 || nt < 1 || nt > 64) return 3;
        A->cel_dbl_off[k] = od;
        A->cel_flt_off[k] = of;
        od += (size_t)nl + (size_t)nl * nl + (size_t)nt * nl * nl;
        of += (size_t)nt + (size_t)nl + 2 + 4 + 5;
    }
    return 0;
}

/* ---- FP32: the reference's precision and C library float functions ---- */
#define R float
#define L(x) x##f
#define FN(n) n##_f32
#define EXPR expf
#define LOG10R log10f
#define SQRTR sqrtf
#define POWR powf
#define FABSR fabsf
#define QCAST float   /* qq(iT) = real(qom(...)), emission_mod.f90:2008 */
#define GC_SAVED_SLOPE 1   /* setGrainPotential's SAVEd slope carries over between calls */
static R pow10m12_f32(void) { return 1.e-12f; }
#include "mc_balance_impl.h"
#include "mc_emission_impl.h"
#undef R
#undef L
#undef FN
#undef EXPR
#undef LOG10R
#undef SQRTR
#undef POWR
#undef FABSR
#undef QCAST
#undef NFOR
#undef GC_SAVED_SLOPE

/* ---- FP64 oracle ---- */
#define R double
#define L(x) x
#define FN(n) n##_f64
#define EXPR exp
#define LOG10R log10
#define SQRTR sqrt
#define POWR pow
#define FABSR fabs
#define QCAST double
#define GC_SAVED_SLOPE 0   /* each grain-potential solve starts from slope 0 */
static R pow10m12_f64(void) { return 1.e-12; }
#include "mc_balance_impl.h"
#include "mc_emission_impl.h"

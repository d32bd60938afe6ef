/*
 * mc_balance_impl.h - restatement of MOCASSIN's per-cell balance
 * (updateCell, update_mod.f90:15-2027, with equilibrium/luSlv from
 * emission_mod.f90:1913-2190 and spline/splint from interpolation_mod.f90:
 * 182-316), included twice by mc_balance.c:
 *   R = float : the reference's own precision.  Variables the reference
 *               declares REAL are float, DOUBLE PRECISION ones double, with
 *               the reference's expression order and the C library's float
 *               functions (what the compiled reference calls);
 *   R = double: the FP64 oracle of the GPU balance kernel.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Macros supplied by the includer: R, L(x) (literal of type R), FN(name),
 * EXPR, LOG10R, SQRTR, POWR, FABSR.
 */

/* charge-exchange fit coefficients chex(elem,ion,1:4) and deltaE/k of the
 * reverse reaction, as ionBalance sets them (update_mod.f90:1389-1446);
 * a later assignment to the same (elem,ion) overrides an earlier one
 * exactly as in the reference (Ar II/III entries at 1436-1439). */
static void FN(chex_table)(R chex[31][10][4], R dEk[31][10])
{
    memset(chex, 0, sizeof(R) * 31 * 10 * 4);
    memset(dEk, 0, sizeof(R) * 31 * 10);
#define CX(e, i, a, b, c, d) do { chex[e][i][0] = L(a); chex[e][i][1] = L(b); chex[e][i][2] = L(c); chex[e][i][3] = L(d); } while (0)
    CX(2, 1, 7.47e-6, 2.06, 9.93, -3.89);
    CX(2, 2, 1.e-5, 0., 0., 0.);
    CX(3, 2, 1.26, 0.96, 3.02, -0.65);
    CX(3, 3, 1.e-5, 0., 0., 0.);
    CX(4, 2, 1.e-5, 0., 0., 0.);
    CX(4, 3, 1.e-5, 0., 0., 0.);
    CX(4, 4, 5.17, 0.82, -.69, -1.12);
    CX(5, 2, 2.e-2, 0., 0., 0.);
    CX(5, 3, 1.e-5, 0., 0., 0.);
    CX(5, 4, 5.27e-1, 0.76, -0.63, -1.17);
    CX(6, 1, 1.76e-9, 8.33, 4278.78, -6.41);
    CX(6, 2, 1.67e-4, 2.79, 304.72, -4.07);
    CX(6, 3, 3.25, 0.21, 0.19, -3.29);
    CX(6, 4, 332.46, -0.11, -0.995, -1.58e-3);
    CX(7, 1, 1.01e-3, -0.29, -0.92, -8.38);
    CX(7, 2, 3.05e-1, 0.60, 2.65, -0.93);
    CX(7, 3, 4.54, 0.57, -0.65, -0.89);
    CX(7, 4, 3.28, 0.52, -0.52, -0.19);
    CX(8, 1, 1.04, 3.15e-2, -0.61, -9.73);
    CX(8, 2, 1.04, 0.27, 2.02, -5.92);
    CX(8, 3, 3.98, 0.26, 0.56, -2.62);
    CX(8, 4, 2.52e-1, 0.63, 2.08, -4.16);
    CX(9, 2, 1.e-5, 0., 0., 0.);
    CX(9, 3, 9.86, 0.29, -0.21, -1.15);
    CX(9, 4, 7.15e-1, 1.21, -0.70, -0.85);
    CX(10, 2, 1.e-5, 0., 0., 0.);
    CX(10, 3, 14.73, 4.52e-2, -0.84, -0.31);
    CX(10, 4, 6.47, 0.54, 3.59, -5.22);
    CX(11, 2, 1.e-5, 0., 0., 0.);
    CX(11, 3, 1.33, 1.15, 1.20, -0.32);
    CX(11, 4, 1.01e-1, 1.34, 10.05, -6.41);
    CX(12, 2, 8.58e-5, 2.49e-3, 2.93e-2, -4.33);
    CX(12, 3, 6.49, 0.53, 2.82, -7.63);
    CX(12, 4, 6.36, 0.55, 3.86, -5.19);
    CX(13, 2, 1.e-5, 0., 0., 0.);
    CX(13, 3, 7.11e-5, 4.12, 1.72e4, -22.24);
    CX(13, 4, 7.52e-1, 0.77, 6.24, -5.67);
    CX(14, 2, 1.23, 0.24, 3.17, 4.18e-3);
    CX(14, 3, 4.900e-1, -8.74e-2, -0.36, -0.79);
    CX(14, 4, 7.58, 0.37, 1.06, -4.09);
    CX(16, 1, 3.82e-7, 11.10, 2.57e4, -8.22);
    CX(16, 2, 1.e-5, 0., 0., 0.);
    CX(16, 3, 2.29, 4.02e-2, 1.59, -6.06);
    CX(16, 4, 6.44, 0.13, 2.69, -5.69);
    CX(18, 2, 1.e-5, 0., 0., 0.);
    CX(18, 3, 4.57, 0.27, -0.18, -1.57);
    CX(18, 4, 6.37, 2.12, 10.21, -6.22);
    CX(18, 3, 3.17e-2, 2.12, 12.06, -0.40);
    CX(18, 4, 2.68, 0.69, -0.68, -4.47);
    CX(26, 2, 1.26, 7.72e-2, -0.41, -7.31);
    CX(26, 3, 3.42, 0.51, -2.06, -8.99);
#undef CX
    dEk[7][1] = L(10863.);
    dEk[8][1] = L(2205.);
    for (int e = 1; e <= 30; ++e)
        for (int i = 1; i < 10; ++i) chex[e][i][0] = chex[e][i][0] * L(1.e-9);
}

/* radRecFit, update_mod.f90:1674-1780 */
static R FN(rad_rec_fit)(const mco_atomic *A, int z, int n, R Te)
{
    const float *rrec = A->rrec, *rnew = A->rnew, *fe = A->fe;
#define RREC(i, z_, n_) ((R)rrec[((z_) - 1) * 60 + ((n_) - 1) * 2 + (i) - 1])
#define RNEW(i, z_, n_) ((R)rnew[((z_) - 1) * 120 + ((n_) - 1) * 4 + (i) - 1])
    if (n <= 3 || n == 11 || (z > 5 && z < 9) || z == 10 || (z == 26 && n > 11)) {
        R tt = SQRTR(Te / RNEW(3, z, n));
        return RNEW(1, z, n) / (tt * POWR(tt + L(1.), L(1.) - RNEW(2, z, n)) *
                                POWR(L(1.) + SQRTR(Te / RNEW(4, z, n)), L(1.) + RNEW(2, z, n)));
    }
    R tt = Te * L(1.e-4);
    if (z == 26 && n <= 13) {
        const float *f = fe + (n - 4) * 3;
        return (R)f[0] / POWR(tt, (R)f[1] + (R)f[2] * LOG10R(tt));
    }
    return RREC(1, z, n) / POWR(tt, RREC(2, z, n));
#undef RREC
#undef RNEW
}

/* dielectronic, update_mod.f90:1782-1834 */
static void FN(dielectronic)(const mco_atomic *A, int nstages, R TeUsed, R diRec[31][10])
{
    R aldro[31][10];
    memset(aldro, 0, sizeof aldro);
    R t = TeUsed / L(10000.);
    const R e12 = FN(pow10m12)();
    for (int i = 0; i < 25; ++i) {
        int elem = A->dr_int[3 * i], n = A->dr_int[3 * i + 1], g = A->dr_int[3 * i + 2];
        int ion = elem + 1 - n;
        if (elem < 1 || ion > nstages || ion < 1) continue;
        const float *c = A->dr_real + 5 * i;
        if (g == 0 || (g == 1 && TeUsed < L(20000.)) || (g == 2 && TeUsed >= L(20000.)))
            diRec[elem][ion] = e12 * ((R)c[0] / t + (R)c[1] + (R)c[2] * t + (R)c[3] * (t * t)) *
                               POWR(t, L(-1.5)) * EXPR(-(R)c[4] / t);
    }
    t = TeUsed;
    for (int i = 0; i < 167; ++i) {
        int elem = A->ap_int[2 * i], n = A->ap_int[2 * i + 1];
        int ion = elem + 1 - n;
        const float *c = A->ap_real + 4 * i;
        R alpha = (R)c[0] * POWR(t, L(-1.5)) * EXPR(-(R)c[2] / t) * (L(1.) + (R)c[1] * EXPR(-(R)c[3] / t));
        if (elem >= 1 && ion >= 1 && ion <= nstages) aldro[elem][ion] = alpha;
    }
    for (int e = 1; e <= 30; ++e)
        for (int i = 1; i <= nstages && i < 10; ++i)
            if (TeUsed > L(60000.) || diRec[e][i] == L(0.)) diRec[e][i] = aldro[e][i];
}

/* calcAlpha, update_mod.f90:1624-1656 */
static void FN(calc_alpha)(const mco_atomic *A, int nstages, R TeUsed, R alphaTot[31][10])
{
    R diRec[31][10];
    memset(diRec, 0, sizeof diRec);
    memset(alphaTot, 0, sizeof(R) * 31 * 10);
    for (int elem = 1; elem <= 30; ++elem)
        for (int ion = 1; ion <= (nstages - 1 < elem ? nstages - 1 : elem); ++ion)
            alphaTot[elem][ion] = A->elemOn[elem] ? FN(rad_rec_fit)(A, elem, elem - ion + 1, TeUsed) : L(0.);
    FN(dielectronic)(A, nstages, TeUsed, diRec);
    for (int elem = 3; elem <= 30; ++elem)
        for (int ion = 1; ion <= (nstages - 1 < elem ? nstages - 1 : elem); ++ion) {
            if (diRec[elem][ion] == L(0.)) diRec[elem][ion] = diRec[8][ion];
            alphaTot[elem][ion] = alphaTot[elem][ion] + diRec[elem][ion];
        }
}

/* spline / splint, interpolation_mod.f90:182-316 (natural spline), REAL */
static void FN(spline_nat)(const float *xa, const float *ya, int n, float *y2a)
{
    float u[64];
    y2a[0] = 0.f;
    u[0] = 0.f;
    for (int i = 1; i < n - 1; ++i) {
        float sig = (xa[i] - xa[i - 1]) / (xa[i + 1] - xa[i - 1]);
        float p = sig * y2a[i - 1] + 2.f;
        y2a[i] = (sig - 1.f) / p;
        float v = ya[i + 1];
        v = v - ya[i];
        v = v / (xa[i + 1] - xa[i]);
        v = v - (ya[i] - ya[i - 1]) / (xa[i] - xa[i - 1]);
        u[i] = (6.f * v / (xa[i + 1] - xa[i - 1]) - sig * u[i - 1]) / p;
    }
    float qn = 0.f, un = 0.f;
    y2a[n - 1] = (un - qn * u[n - 2]) / (qn * y2a[n - 2] + 1.f);
    for (int k = n - 2; k >= 0; --k) y2a[k] = y2a[k] * y2a[k + 1] + u[k];
}

static float FN(splint_f)(const float *xa, const float *ya, const float *y2a, int n, float x)
{
    if (x <= xa[0]) return ya[0];
    if (x >= xa[n - 1]) return ya[n - 1];
    int klo = 1, khi = n;
    for (int i = 1; i <= 1000; ++i) {
        if (khi - klo <= 1) break;
        int k = (khi + klo) / 2;
        if (xa[k - 1] > x) khi = k; else klo = k;
    }
    float h = xa[khi - 1] - xa[klo - 1];
    float a = (xa[khi - 1] - x) / h;
    float b = (x - xa[klo - 1]) / h;
    return a * ya[klo - 1] + b * ya[khi - 1] +
           ((a * a * a - a) * y2a[klo - 1] + (b * b * b - b) * y2a[khi - 1]) * (h * h) / 6.f;
}

/* spline/splint in the working precision (the FP64 oracle evaluates the
 * collision strengths in double) */
static void FN(spline_R)(const R *xa, const R *ya, int n, R *y2a)
{
    R u[64];
    y2a[0] = L(0.);
    u[0] = L(0.);
    for (int i = 1; i < n - 1; ++i) {
        R sig = (xa[i] - xa[i - 1]) / (xa[i + 1] - xa[i - 1]);
        R p = sig * y2a[i - 1] + L(2.);
        y2a[i] = (sig - L(1.)) / p;
        R v = ya[i + 1];
        v = v - ya[i];
        v = v / (xa[i + 1] - xa[i]);
        v = v - (ya[i] - ya[i - 1]) / (xa[i] - xa[i - 1]);
        u[i] = (L(6.) * v / (xa[i + 1] - xa[i - 1]) - sig * u[i - 1]) / p;
    }
    y2a[n - 1] = L(0.);
    for (int k = n - 2; k >= 0; --k) y2a[k] = y2a[k] * y2a[k + 1] + u[k];
}

static R FN(splint_R)(const R *xa, const R *ya, const R *y2a, int n, R x)
{
    if (x <= xa[0]) return ya[0];
    if (x >= xa[n - 1]) return ya[n - 1];
    int klo = 1, khi = n;
    while (khi - klo > 1) {
        int k = (khi + klo) / 2;
        if (xa[k - 1] > x) khi = k; else klo = k;
    }
    R h = xa[khi - 1] - xa[klo - 1];
    R a = (xa[khi - 1] - x) / h;
    R b = (x - xa[klo - 1]) / h;
    return a * ya[klo - 1] + b * ya[khi - 1] +
           ((a * a * a - a) * y2a[klo - 1] + (b * b * b - b) * y2a[khi - 1]) * (h * h) / L(6.);
}

/* equilibrium, emission_mod.f90:1913-2114, with luSlv/lured/reslv
 * (2119-2190): level populations of one model ion and the emissivity
 * fLineEm(i,j) = A(j,i) E_ji n_j (cm^-1 s^-1 per ion); returns 0, or 1 when
 * ionDenUp is NaN (all emissivities zero).  fl is (nFor x nFor),
 * column-major, zeroed by the caller. */
#define NFOR 17
static void FN(equilibrium)(const mco_atomic *A, int k, R ionDenUp, R Te, R Ne, double *fl)
{
    const int32_t *m = A->cel_meta + 6 * k;
    const int nl = m[4], nt = m[5];
    const double *e = A->cel_dbl + A->cel_dbl_off[k];
    const double *aa = e + nl;                 /* a(nl,nl) column-major */
    const double *qom = aa + nl * nl;          /* qom(nt,nl,nl) */
    const float *logT = A->cel_flt + A->cel_flt_off[k];
    const float *gf = logT + nt;
    const float br = gf[nl], z = gf[nl + 1];
    const float *a_r = gf + nl + 2, *a_d = gf + nl + 6;
#define AA(i, j) aa[((i) - 1) + nl * ((j) - 1)]
#define QOM(t, i, j) qom[((t) - 1) + nt * (((i) - 1) + nl * ((j) - 1))]
#define FL(i, j) fl[((i) - 1) + NFOR * ((j) - 1)]
    if (ionDenUp != ionDenUp) return;
    const double sqrTe = (double)SQRTR(Te);
    const R log10Te = LOG10R(Te);
    R alphaTotal[NFOR + 1];
    for (int j = 0; j <= NFOR; ++j) alphaTotal[j] = L(0.);
    for (int j = 2; j <= nl; ++j) {
        if (br <= 0.f) break;
        alphaTotal[1] = L(1.);
        R t4 = Te * L(1.e-4);
        R zz = (R)z;
        R x1 = t4 / (zz * zz);
        alphaTotal[j] = L(1.e-13) * (R)br * zz * (R)a_r[0] * POWR(x1, (R)a_r[1]) /
                            (L(1.) + (R)a_r[2] * POWR(t4 / (zz * zz), (R)a_r[3])) +
                        L(1.0e-12) * (R)br *
                            ((R)a_d[0] / t4 + (R)a_d[1] + (R)a_d[2] * Te * L(1.e-4) + (R)a_d[3] * (t4 * t4)) *
                            POWR(t4, L(-1.5)) * EXPR(-((R)a_d[4] * L(1.e4) / Te));
    }
    double qeff[NFOR + 1][NFOR + 1];
    for (int i = 2; i <= nl; ++i) {
        for (int j = i; j <= nl; ++j) {
            R qomInt;
            if (nt == 1) {
                qomInt = (R)(QCAST)QOM(1, i - 1, j);
            } else if (nt == 2) {
                R q1 = (R)(QCAST)QOM(1, i - 1, j), q2 = (R)(QCAST)QOM(2, i - 1, j);
                qomInt = q1 + (q2 - q1) / ((R)logT[1] - (R)logT[0]) * (log10Te - (R)logT[0]);
            } else if (sizeof(R) == sizeof(float)) {
                float qq[64], qq2[64];
                for (int t = 1; t <= nt; ++t) qq[t - 1] = (float)QOM(t, i - 1, j);
                FN(spline_nat)(logT, qq, nt, qq2);
                qomInt = (R)FN(splint_f)(logT, qq, qq2, nt, (float)log10Te);
            } else {
                R xa[64], qq[64], qq2[64];
                for (int t = 1; t <= nt; ++t) { xa[t - 1] = (R)logT[t - 1]; qq[t - 1] = (R)QOM(t, i - 1, j); }
                FN(spline_R)(xa, qq, nt, qq2);
                qomInt = FN(splint_R)(xa, qq, qq2, nt, log10Te);
            }
            const double cs = (double)qomInt;
            const double delTeK = (e[i - 2] - e[j - 1]) * 1.4388463;
            const double expFac = exp(delTeK / (double)Te);
            qeff[i - 1][j] = 8.63e-6 * cs * expFac / ((double)gf[i - 2] * sqrTe);
            qeff[j][i - 1] = 8.63e-6 * cs / ((double)gf[j - 1] * sqrTe);
        }
    }
    double x[NFOR + 1][NFOR + 1], y[NFOR + 1];
    memset(x, 0, sizeof x);
    memset(y, 0, sizeof y);
    const double dNe = (double)Ne;
    for (int i = 2; i <= nl; ++i) {
        for (int j = 1; j <= nl; ++j) {
            for (int jj = 1; jj <= nl; ++jj) x[1][jj] = 1.0;
            y[1] = 1.0;
            if (j != i) {
                x[i][j] = x[i][j] + dNe * qeff[j][i];
                x[i][i] = x[i][i] - dNe * qeff[i][j];
                if (j > i) x[i][j] = x[i][j] + AA(j, i);
                else x[i][i] = x[i][i] - AA(i, j);
            }
        }
        y[i] = (double)(-(Ne * ionDenUp * alphaTotal[i]));
    }
    /* lured */
    for (int i = 1; i <= nl - 1; ++i)
        for (int kk = i + 1; kk <= nl; ++kk) {
            double factor = x[kk][i] / x[i][i];
            for (int j = i + 1; j <= nl; ++j) x[kk][j] = x[kk][j] - x[i][j] * factor;
        }
    /* reslv */
    for (int i = 1; i <= nl - 1; ++i)
        for (int j = i + 1; j <= nl; ++j) y[j] = y[j] - y[i] * x[j][i] / x[i][i];
    y[nl] = y[nl] / x[nl][nl];
    for (int i = 1; i <= nl - 1; ++i) {
        int kk = nl - i;
        for (int j = kk + 1; j <= nl; ++j) y[kk] = y[kk] - y[j] * x[kk][j];
        y[kk] = y[kk] / x[kk][kk];
    }
    double sumN = 0.0;
    for (int i = 1; i <= nl; ++i) sumN = sumN + y[i];
    for (int i = 1; i <= nl; ++i) y[i] = y[i] / sumN;
    for (int i = 1; i <= nl - 1; ++i)
        for (int j = i + 1; j <= nl; ++j)
            if (AA(j, i) != 0.0) {
                double Eji = e[j - 1] - e[i - 1];
                FL(i, j) = AA(j, i) * Eji * y[j];
            }
#undef AA
#undef QOM
#undef FL
}

typedef struct {
    R Te, Ne, Hden;
    R ion[31][10];          /* ionDenUsed(xref(elem), stage) by atomic number */
    R nPhotoSte[31][10];
    R heatNu[31][10];       /* frequency integral of the heating of each ion */
    R alphaTot[31][10];
    int nIterateX;
} FN(cell_t);

/* ionBalance, update_mod.f90:1320-1622 (recursion written as a loop) */
static int FN(ion_balance)(const mco_atomic *A, FN(cell_t) *c, R chex0[31][10][4], R dEk[31][10])
{
    const int ns = A->nstages;
    for (;;) {
        int lgConv = 1;
        double ionRatio[31][10], ionProd[31][10], denom[31];
        for (int e = 0; e <= 30; ++e) {
            denom[e] = 1.0;
            for (int i = 0; i < 10; ++i) { ionRatio[e][i] = 0.0; ionProd[e][i] = 1.0; }
        }
        R expFact = L(157893.94) / c->Te;
        if (expFact > L(75.)) expFact = L(75.);
        R collIon = L(2.75E-16) * c->Te * SQRTR(c->Te) * (L(157893.94) / c->Te + L(2.)) * EXPR(-expFact);
        const R HIOld = c->ion[1][1], HeIOld = c->ion[2][1], HeIIOld = c->ion[2][2];
        R chex[31][10];
        const R t4 = c->Te / L(10000.);
        const R xH0 = c->ion[1][1];
        for (int elem = 1; elem <= 30; ++elem) {
            for (int ion = 1; ion <= (elem < ns - 1 ? elem : ns - 1); ++ion) {
                if (!A->elemOn[elem]) break;
                R cx = chex0[elem][ion][0] * POWR(t4, chex0[elem][ion][1]) *
                       (L(1.) + chex0[elem][ion][2] * EXPR(chex0[elem][ion][3] * t4));
                if (cx < L(0.)) cx = L(0.);
                if (c->Te < L(6000.) || c->Te > L(5.e4)) cx = L(0.);
                chex[elem][ion] = cx;
                R revRate = L(0.);
                if (dEk[elem][ion] > L(1.)) {
                    const int32_t *g = A->chan_g + 2 * (elem * 10 + ion);
                    revRate = cx * (L(1.) - xH0) * c->Hden * L(2.) * EXPR(-dEk[elem][ion] / c->Te) /
                              ((R)g[0] / (R)g[1]);
                }
                if (elem > 1 || ion > 1) collIon = L(0.);
                R r = (c->nPhotoSte[elem][ion] + collIon + revRate) /
                      (c->Ne * c->alphaTot[elem][ion] + cx * c->Hden * xH0);
                ionRatio[elem][ion] = (double)r;
            }
        }
        for (int elem = 1; elem <= 30; ++elem)
            for (int ion = 1; ion <= (elem < ns - 1 ? elem : ns - 1); ++ion) {
                if (!A->elemOn[elem]) break;
                for (int i = 1; i <= ion; ++i) ionProd[elem][ion] = ionProd[elem][ion] * ionRatio[elem][i];
            }
        for (int elem = 1; elem <= 30; ++elem)
            for (int ion = 1; ion <= (elem < ns - 1 ? elem : ns - 1); ++ion) {
                if (!A->elemOn[elem]) break;
                denom[elem] = denom[elem] + ionProd[elem][ion];
            }
        for (int elem = 1; elem <= 30; ++elem)
            for (int ion = 1; ion <= (elem + 1 < ns ? elem + 1 : ns); ++ion) {
                if (!A->elemOn[elem]) break;
                R v = (ion == 1) ? (R)(1.0 / denom[elem]) : (R)(ionProd[elem][ion - 1] / denom[elem]);
                if (v > L(0.99999)) v = L(0.99999);
                if (v < L(1.e-20)) v = L(1.e-20);
                c->ion[elem][ion] = v;
            }
        R Ne = L(0.);
        for (int elem = 1; elem <= 30; ++elem)
            for (int ion = 2; ion <= (elem + 1 < ns ? elem + 1 : ns); ++ion)
                if (A->elemOn[elem] && c->ion[elem][ion] >= L(1.e-10))
                    Ne = Ne + (R)(ion - 1) * (R)A->abun[elem] * c->ion[elem][ion];
        Ne = Ne * c->Hden;
        if (Ne == L(0.)) Ne = L(1.);
        c->Ne = Ne;
        const R dHI = (c->ion[1][1] - HIOld) / HIOld;
        const R dHeI = (c->ion[2][1] - HeIOld) / HeIOld;
        const R dHeII = (c->ion[2][2] - HeIIOld) / HeIIOld;
        const int big = FABSR(dHI) > L(0.01) || FABSR(dHeI) > L(0.01) || FABSR(dHeII) > L(0.01);
        if (big && c->nIterateX < 25) { c->nIterateX++; continue; }
        if (big && c->nIterateX == 25) lgConv = 0;
        return lgConv;
    }
}

/* thermBalance, update_mod.f90:933-1258 (gas only) */
static void FN(therm_balance)(const mco_atomic *A, FN(cell_t) *c, R *heatInt, R *coolInt, double *fl)
{
    const R T = c->Te, Ne = c->Ne;
    const R kB = L(1.380658e-16), hcRyd = L(2.1799153e-11);
    const R lg = LOG10R(T);
    const R Te4 = T / L(1.e4);
    R Np = c->ion[1][2] * (R)A->abun[1];
    R betaFF = L(1.0108464E-11) + L(9.7930778E-13) * lg - L(6.6433144E-13) * lg * lg +
               L(2.4793747E-13) * lg * lg * lg - L(2.3938215E-14) * lg * lg * lg * lg;
    R coolFF = Np * Ne * betaFF * kB * T / SQRTR(T);
    R betaRec = L(9.4255985E-11) - L(4.04794384E-12) * lg - L(1.0055237E-11) * lg * lg +
                L(1.99266862E-12) * lg * lg * lg - L(1.06681387E-13) * lg * lg * lg * lg;
    R coolRec = Np * Ne * betaRec * kB * T / SQRTR(T);
    Np = c->ion[2][3] * (R)A->abun[2];
    const R ls = LOG10R(T / L(4.));
    betaFF = L(2.) * (L(1.0108464E-11) + L(9.7930778E-13) * ls - L(6.6433144E-13) * ls * ls +
                      L(2.4793747E-13) * ls * ls * ls - L(2.3938215E-14) * ls * ls * ls * ls);
    coolFF = coolFF + Np * Ne * betaFF * kB * T / SQRTR(T / L(4.));
    betaRec = L(2.) * (L(9.4255985E-11) - L(4.04794384E-12) * ls - L(1.0055237E-11) * ls * ls +
                       L(1.99266862E-12) * ls * ls * ls - L(1.06681387E-13) * ls * ls * ls * ls);
    coolRec = coolRec + Np * Ne * betaRec * kB * T / SQRTR(T / L(4.));
    Np = c->ion[2][2] * (R)A->abun[2];
    betaFF = L(1.070073e-11) - L(2.5730207e-13) * lg + L(2.109134e-13) * lg * lg;
    coolFF = coolFF + Np * Ne * betaFF * kB * T / SQRTR(T);
    betaRec = L(9.4255985E-11) - L(4.04794384E-12) * lg - L(1.0055237E-11) * lg * lg +
              L(1.99266862E-12) * lg * lg * lg - L(1.06681387E-13) * lg * lg * lg * lg;
    coolRec = coolRec + Np * Ne * betaRec * kB * T / SQRTR(T);
    R coolColl;
    if (T > L(5000.))
        coolColl = (L(2.47e-8) * EXPR(-L(118338.) / T) * POWR(Te4, L(-0.228)) +
                    L(1.32e-8) * EXPR(-L(140252.) / T) * POWR(Te4, L(-0.460))) *
                   hcRyd * c->ion[1][1] * Ne;
    else
        coolColl = L(0.);
    /* forLines (update_mod.f90:1263-1318) and the CEL cooling sum, which
     * adds levels 1..10 only (1099-1113) */
    for (int k = 0; k < A->ncel; ++k) {
        const int elem = A->cel_meta[6 * k], ion = A->cel_meta[6 * k + 1];
        double *f = fl + (size_t)k * NFOR * NFOR;
        for (int i = 0; i < NFOR * NFOR; ++i) f[i] = 0.0;
        R up = (ion < A->nstages) ? c->ion[elem][ion + 1] / c->ion[elem][ion] : L(0.);
        FN(equilibrium)(A, k, up, T, Ne, f);
        const double s1 = (double)(R)A->abun[elem], s2 = (double)c->ion[elem][ion];
        for (int i = 0; i < NFOR * NFOR; ++i) f[i] = f[i] * s1 * s2;
    }
    const double kE = (double)(R)L(1.9865e-16);
    for (int k = 0; k < A->ncel; ++k) {
        double *f = fl + (size_t)k * NFOR * NFOR;
        for (int j = 1; j <= 10; ++j)
            for (int kk = 1; kk <= 10; ++kk)
                coolColl = coolColl + (R)(f[(j - 1) + NFOR * (kk - 1)] * kE);
    }
    R heatSte = L(0.);
    for (int elem = 1; elem <= 30; ++elem)
        for (int ion = 1; ion <= (elem < A->nstages - 1 ? elem : A->nstages - 1); ++ion) {
            if (!A->elemOn[elem]) break;
            R h = c->heatNu[elem][ion] * c->ion[elem][ion] * (R)A->abun[elem];
            heatSte = heatSte + h;
        }
    *heatInt = heatSte;
    *coolInt = coolFF + coolRec + coolColl;
}

/* outputGas's analytic line luminosities for a 3-D grid (output_mod.f90:
 * 302-520): per cell in the caller's order, forLines (1946-2008: equilibrium
 * of every model ion, times elemAbun * ionDen, times 1.9865e9, the REAL
 * literal) and RecLinesEmission's Hbeta (2065-2079: 2530/Te**0.833 * Ne *
 * ionDen(H,2) * elemAbun(H); HIRecLines(4,2) = (fh/hb) * Hbeta with fh = hb),
 * each times Hden * dV and added to forbVol and HbetaVol, which the reference
 * holds as REAL: with R = float each cell's term is rounded to float and
 * summed in float, with R = double everything stays double.  State per cell c:
 * Hden/Te/Ne/dV[c], ionDen[c * nstages * nElemUsed + (s-1)*nElemUsed + xref-1].
 * forb[(k * 17 + up - 1) * 17 + lo - 1] is forbVol of model ion k (forLines
 * order), line up -> lo; both outputs are accumulated into (not zeroed). */
void FN(mco_line_lum)(const mco_atomic *A, int n, const R *Hden, const R *Te, const R *Ne, const R *ionDen,
                      const R *dV, R *forb, R *hbeta)
{
    const int ns = A->nstages, nE = A->nElemUsed;
    double fl[NFOR * NFOR];
    const double kL = (double)(R)L(1.9865e9);
    for (int c = 0; c < n; ++c) {
        const R *ion = ionDen + (size_t)c * ns * nE;
        for (int k = 0; k < A->ncel; ++k) {
            const int elem = A->cel_meta[6 * k], st = A->cel_meta[6 * k + 1];
            const int xr = A->xref[elem] - 1;
            const R x = ion[(st - 1) * nE + xr];
            const R up = (st < ns) ? ion[st * nE + xr] / x : L(0.);
            for (int i = 0; i < NFOR * NFOR; ++i) fl[i] = 0.0;
            FN(equilibrium)(A, k, up, Te[c], Ne[c], fl);
            for (int lo = 1; lo <= NFOR; ++lo)
                for (int hi = 1; hi <= NFOR; ++hi) {
                    double f = fl[(lo - 1) + NFOR * (hi - 1)];
                    f = f * (double)(R)A->abun[elem] * (double)x;
                    f = f * kL;
                    R *o = forb + ((size_t)k * NFOR + hi - 1) * NFOR + lo - 1;
                    *o = *o + (R)(f * (double)Hden[c] * (double)dV[c]);
                }
        }
        R hb = L(2530.) / POWR(Te[c], L(0.833));
        hb = hb * Ne[c] * ion[1 * nE + A->xref[1] - 1] * (R)A->abun[1];
        *hbeta = *hbeta + (R)((double)hb * (double)Hden[c] * (double)dV[c]);
    }
}

/* ---- grain charge, photoelectric heating and gas-grain collisions ------
 * (update_mod.f90:527-932), one dust chemistry.  Constants as the reference
 * declares them: Pi a REAL (constants_mod.f90:32), Y0 = 0.5, Y1 = 0.2,
 * errorLim 0.005, dm 0.05, maxIterateGC 100 (update_mod.f90:20,81,533-535). */
typedef struct {
    const mco_grains *G;
    const R *J;              /* the cell's scaled Jste */
    R ndust;                 /* Ndust(cell) */
    const R *Tdust;          /* Tdust(0:nsMax, 0:nSizes, cell) before the update */
    int ldT;                 /* nsMax + 1 */
    R pot[MCO_MAX_GROWS];    /* grainPot(isp, ai) at (isp-1)*nSizes + ai-1 */
    R gdc[MCO_MAX_GROWS], peh[MCO_MAX_GROWS];   /* gasDustColl_d, photoelHeat_d */
    R gdcg, pehg;            /* gasDustColl_g, photoelHeat_g */
} FN(grain_t);

#define PI_R ((R)3.141592654f)
#define FOURPI_R ((R)(4.f * 3.141592654f))
#define MINR(a, b) ((a) < (b) ? (a) : (b))
#define MAXR(a, b) ((a) > (b) ? (a) : (b))

/* locate(nuArray, x), interpolation_mod.f90:48-81, for the increasing
 * frequency mesh: n when x > nu(n), 0 when x < nu(1), else the last i with
 * nu(i) <= x (at least 1); x == nu(n) gives 1 */
static int FN(locate_nu)(const float *nu, int n, R x)
{
    if (x > (R)nu[n - 1]) return n;
    if (x < (R)nu[0]) return 0;
    int p = 0;
    R bd = L(0.);
    for (int i = 0; i < n; ++i) {
        const R d = (R)nu[i] - x;
        if (d > L(0.) && (p == 0 || d < bd)) { p = i + 1; bd = d; }
    }
    return p - 1 > 1 ? p - 1 : 1;
}

/* getGrainEmission, update_mod.f90:601-651 */
static R FN(grain_emission)(const mco_atomic *A, const FN(grain_t) *g, R Vg, int isp, int ai)
{
    const mco_grains *G = g->G;
    const int nb = A->nbins;
    const R hcRyd = L(2.1799153e-11);
    const R vn = (R)G->vn[isp - 1];
    const R thres = MAXR(vn + Vg, vn);
    const int ip = FN(locate_nu)(A->nuArray, nb, thres) + 1;
    const R rad = (R)G->radius[ai - 1];
    const R qdiv = L(1.e-8) * (rad * rad);
    const float *ax = G->absx + ((size_t)(isp - 1) * G->nSizes + (ai - 1)) * nb;
    R sum = L(0.);
    for (int f = ip; f <= nb; ++f) {
        const R nu = (R)A->nuArray[f - 1];
        const R Yn = MINR(L(0.5) * (L(1.) - vn / nu), L(0.2));
        const R Yhat = Yn * MINR(L(1.), MAXR(L(0.), L(1.) - Vg / (nu - vn)));
        R photFlux = g->J[f - 1] / (hcRyd * nu);
        photFlux = photFlux * FOURPI_R;
        const R Qa = (R)ax[f - 1] / qdiv;
        sum = sum + Yhat * photFlux * Qa;
    }
    return sum;
}

/* getGrainRecombination, update_mod.f90:656-733 */
static R FN(grain_recombination)(const mco_atomic *A, const FN(cell_t) *c, const FN(grain_t) *g, R Vg)
{
    const R kT = L(6.336e-6) * c->Te;
    const R e8 = L(1.1045e-15) * c->Te / PI_R;
    R vmean = SQRTR(e8 / L(9.109382e-28));
    const R S = L(1.);
    R eta = -(Vg / kT);
    if (eta <= L(0.)) eta = L(1.) - eta;
    else eta = EXPR(-eta);
    R rec = c->Ne * vmean * S * eta;
    for (int elem = 1; elem <= 30; ++elem) {
        if (!A->elemOn[elem]) continue;
        const int top = elem + 1 < A->nstages ? elem + 1 : A->nstages;
        for (int ist = 2; ist <= top; ++ist) {
            const R cpDen = c->ion[elem][ist] * (R)A->abun[elem] * c->Hden;
            const R mcp = (R)g->G->aweight[elem - 1] * L(1.66053e-24);
            vmean = SQRTR(e8 / mcp);
            const R Z = (R)(ist - 1);
            eta = Z * Vg / kT;
            if (eta <= L(0.)) eta = L(1.) - eta;
            else eta = EXPR(-eta);
            rec = rec - cpDen * vmean * S * eta;
        }
    }
    return rec;
}

#if GC_SAVED_SLOPE
static R FN(gc_slope) = L(0.);
#endif

/* setGrainPotential, update_mod.f90:527-597 (its recursion as a loop) */
static R FN(grain_potential)(const mco_atomic *A, const FN(cell_t) *c, const FN(grain_t) *g, int isp, int ai)
{
#if GC_SAVED_SLOPE
    R slope = FN(gc_slope);
#else
    R slope = L(0.);
#endif
    const R vn = (R)g->G->vn[isp - 1];
    R potOld = L(0.);
    R emiOld = FN(grain_emission)(A, g, L(0.), isp, ai);
    R recOld = FN(grain_recombination)(A, c, g, L(0.));
    R pot = potOld + L(0.05);
    for (int n = 1;; ++n) {
        const R threshold = MAXR(vn + pot, vn);
        const R emi = FN(grain_emission)(A, g, pot, isp, ai);
        const R rec = FN(grain_recombination)(A, c, g, pot);
        R delta = emi - rec;
        R delta1 = FABSR(delta / (L(0.5) * MAXR(L(1.e-35), emi + rec)));
        if (delta1 < L(0.005)) break;
        if (pot != potOld) {
            const R fac = (emi - emiOld) - (rec - recOld);
            if (fac != L(0.)) slope = fac / (pot - potOld);
        }
        potOld = pot;
        recOld = rec;
        emiOld = emi;
        delta1 = -(delta / slope);
        delta = FABSR(delta1);
        delta = MINR(delta, L(0.05) * threshold);
        delta = signbit(delta1) ? -FABSR(delta) : FABSR(delta);
        pot = pot + delta;
        if (n >= 100) break;
    }
#if GC_SAVED_SLOPE
    FN(gc_slope) = slope;
#endif
    return pot;
}

/* setPhotoelHeatCool, update_mod.f90:736-804 */
static void FN(photoel_heat_cool)(const mco_atomic *A, const FN(cell_t) *c, FN(grain_t) *g)
{
    const mco_grains *G = g->G;
    const int nb = A->nbins;
    const R hcRyd = L(2.1799153e-11);
    g->pehg = L(0.);
    for (int k = 0; k < G->nS * G->nSizes; ++k) g->peh[k] = L(0.);
    for (int ns = 1; ns <= G->nS; ++ns)
        for (int na = 1; na <= G->nSizes; ++na) {
            const int k = (ns - 1) * G->nSizes + na - 1;
            const R pot = g->pot[k], vn = (R)G->vn[ns - 1];
            const R th = MAXR(vn + pot, vn);
            int thP = FN(locate_nu)(A->nuArray, nb, th);
            if (thP <= 0) thP = 1;
            const float *ax = G->absx + (size_t)k * nb;
            for (int f = thP; f <= nb; ++f) {
                const R nu = (R)A->nuArray[f - 1];
                const R Yn = MINR(L(0.5) * (L(1.) - vn / nu), L(0.2));
                const R Yhat = Yn * MINR(L(1.), MAXR(L(0.), L(1.) - pot / (nu - vn)));
                const R photFlux = g->J[f - 1] / (hcRyd * nu);
                const R Qa = (R)ax[f - 1];
                const R d = nu - vn;
                const R EY = Yn * L(0.5) * MINR(d, MAXR(L(0.), (d * d - pot * pot) / d));
                g->peh[k] = g->peh[k] + Qa * photFlux * EY * hcRyd / PI_R;
                g->pehg = g->pehg + Qa * photFlux * (EY - Yhat * pot) * (R)G->gabun[ns - 1] * (R)G->weight[na - 1];
            }
        }
    g->pehg = g->pehg * g->ndust * hcRyd / c->Hden;
}

/* setDustGasCollHeatCool, update_mod.f90:806-930.  IPerg is assigned for
 * ionised stages only and keeps its last value for the neutral stage that
 * follows (the reference's own order); it starts at 0. */
static void FN(dust_gas_coll)(const mco_atomic *A, const FN(cell_t) *c, FN(grain_t) *g)
{
    const mco_grains *G = g->G;
    const R Ryd2erg = L(2.1799e-11), kB = L(1.380658e-16);
    const R kT = L(6.336e-6) * c->Te;
    const R e8 = L(1.1045e-15) * c->Te / PI_R;
    R IPerg = L(0.);
    g->gdcg = L(0.);
    for (int k = 0; k < G->nS * G->nSizes; ++k) g->gdc[k] = L(0.);
    for (int elem = 1; elem <= 30; ++elem) {
        if (!A->elemOn[elem]) continue;
        const int top = A->nstages < elem + 1 ? A->nstages : elem + 1;
        for (int ist = 1; ist <= top; ++ist) {
            const R Z = (R)(ist - 1);
            const R mcp = (R)G->aweight[elem - 1] * L(1.66053e-24);
            const R vmean = SQRTR(e8 / mcp);
            if (ist > 1) IPerg = (R)A->nuArray[A->chan_ip[elem * 10 + ist - 1] - 1] * Ryd2erg;
            for (int ns = 1; ns <= G->nS; ++ns) {
                const R ms = (R)G->msurf[ns - 1];
                const R S = (ist == 1) ? L(2.) * mcp * ms / ((mcp + ms) * (mcp + ms)) : L(1.);
                for (int na = 1; na <= G->nSizes; ++na) {
                    const int k = (ns - 1) * G->nSizes + na - 1;
                    const R pot = g->pot[k], rad = (R)G->radius[na - 1];
                    const R Td = g->Tdust[ns + g->ldT * na];
                    const R psi = Z * pot / kT;
                    R eta, xi;
                    if (psi <= L(0.)) { eta = L(1.) - psi; xi = L(1.) - psi / L(2.); }
                    else { eta = EXPR(-psi); xi = (L(1.) + psi / L(2.)) * eta; }
                    const R x = c->ion[elem][ist];
                    g->gdc[k] = g->gdc[k] + x * (R)A->abun[elem] * c->Hden * PI_R * rad * rad * L(1.e-8) * S *
                                vmean * (L(2.) * kT * Ryd2erg * xi - eta * (Z * pot * Ryd2erg - IPerg + L(2.) * kB * Td));
                    g->gdcg = g->gdcg + x * (R)G->weight[na - 1] * (R)G->gabun[ns - 1] * g->ndust * (R)A->abun[elem] *
                              PI_R * rad * rad * L(1.e-8) * S * vmean *
                              (L(2.) * kT * Ryd2erg * xi - eta * L(2.) * kB * Td);
                }
            }
        }
    }
    const R vme = SQRTR(e8 / L(9.109382e-28));
    for (int ns = 1; ns <= G->nS; ++ns) {
        const R S = L(1.), Z = L(-1.);
        for (int na = 1; na <= G->nSizes; ++na) {
            const int k = (ns - 1) * G->nSizes + na - 1;
            const R pot = g->pot[k], rad = (R)G->radius[na - 1];
            const R psi = Z * pot / kT;
            R eta, xi;
            if (psi <= L(0.)) { eta = L(1.) - psi; xi = L(1.) - psi / L(2.); }
            else { eta = EXPR(-psi); xi = (L(1.) + psi / L(2.)) * eta; }
            g->gdc[k] = g->gdc[k] + c->Ne * PI_R * rad * rad * L(1.e-8) * S * vme *
                        (L(2.) * kT * Ryd2erg * xi - eta * (Z * pot * Ryd2erg));
            g->gdcg = g->gdcg + c->Ne * (R)G->weight[na - 1] * (R)G->gabun[ns - 1] * g->ndust * PI_R * rad * rad *
                      L(1.e-8) * S * vme * (L(2.) * kT * Ryd2erg * xi) / c->Hden;
        }
    }
    for (int k = 0; k < G->nS * G->nSizes; ++k) g->gdc[k] = g->gdc[k] / FOURPI_R;
    g->gdcg = g->gdcg / FOURPI_R;
}

/* the grain block of iterateT after ionBalance (update_mod.f90:367-388) */
static void FN(grain_terms)(const mco_atomic *A, const FN(cell_t) *c, FN(grain_t) *g)
{
    const mco_grains *G = g->G;
    for (int isp = 1; isp <= G->nS; ++isp)
        for (int ai = 1; ai <= G->nSizes; ++ai)
            g->pot[(isp - 1) * G->nSizes + ai - 1] = FN(grain_potential)(A, c, g, isp, ai);
    FN(photoel_heat_cool)(A, c, g);
    FN(dust_gas_coll)(A, c, g);
}

/* updateCell + iterateT for one active cell (update_mod.f90:15-525).
 * J: the cell's scaled Jste(cell,1:nbins).  ionDen: [stage-1][xref-1]
 * slice of ionDen(cell,:,:) (in/out).  Returns flags: bit0 converged,
 * bit1 black, bit2 Te not balanced, bit3 ion balance not converged,
 * bit4 not hit; *nT = T-iterations used. */
static int FN(balance_cell)(const mco_atomic *A, const R *J, R Hden, R *Te, R *Ne, R *ionDen, int *nT,
                            FN(grain_t) *g)
{
    const int ns = A->nstages, nE = A->nElemUsed, nb = A->nbins;
    int hit = 0;
    for (int i = 0; i < nb; ++i) if (J[i] > L(0.)) { hit = 1; break; }
    if (!hit) { if (nT) *nT = 0; return 16; }
    FN(cell_t) c;
    memset(&c, 0, sizeof c);
    c.Te = *Te; c.Ne = *Ne; c.Hden = Hden;
    for (int elem = 1; elem <= 30; ++elem)
        if (A->elemOn[elem])
            for (int s = 1; s <= ns; ++s) c.ion[elem][s] = ionDen[(s - 1) * nE + A->xref[elem] - 1];
    const R XOldHI = c.ion[1][1];
    const R hcRyd = L(2.1799153e-11);
    for (int e = 0; e <= 30; ++e)
        for (int i = 0; i < 10; ++i) c.nPhotoSte[e][i] = L(1.e-20);
    for (int k = 0; k < A->nch; ++k) {
        const int32_t *m = A->ch_meta + 8 * k;
        const int elem = m[0], ion = m[1], ip = m[4], hiR = m[5], hiH = m[6];
        const float *sig = A->ch_sigma + (size_t)k * nb;
        R acc = c.nPhotoSte[elem][ion];
        for (int j = ip; j <= hiR; ++j) {
            R ph = (R)sig[j - 1];
            if (ph < L(1.e-35)) ph = L(0.);
            if (J[j - 1] > L(0.)) acc = acc + J[j - 1] * ph / (hcRyd * (R)A->nuArray[j - 1]);
        }
        c.nPhotoSte[elem][ion] = acc;
        R h = L(0.);
        for (int j = ip; j <= hiH; ++j) {
            R ph = (R)sig[j - 1];
            if (ph < L(1.e-35)) break;
            if (J[j - 1] > L(0.))
                h = h + ph * J[j - 1] * ((R)A->nuArray[j - 1] - (R)A->nuArray[ip - 1]) / (R)A->nuArray[j - 1];
        }
        c.heatNu[elem][ion] = h;
    }
    R chex0[31][10][4], dEk[31][10];
    FN(chex_table)(chex0, dEk);
    double *fl = (double *)malloc(sizeof(double) * NFOR * NFOR * (A->ncel > 0 ? A->ncel : 1));
    int nIterateT = 0, flags = 0, lgIBConv = 1;
    R thResidualHigh = L(0.), thResidualLow = L(0.), Thigh = L(0.), Tlow = L(0.);
    for (;;) {
        nIterateT++;
        FN(calc_alpha)(A, ns, c.Te, c.alphaTot);
        c.nIterateX = 1;
        lgIBConv = FN(ion_balance)(A, &c, chex0, dEk);
        if (g) FN(grain_terms)(A, &c, g);
        R heatInt, coolInt;
        FN(therm_balance)(A, &c, &heatInt, &coolInt, fl);
        if (g) {
            /* thermBalance's dust terms (update_mod.f90:1252-1255) */
            coolInt = coolInt + g->gdcg;
            heatInt = heatInt + g->pehg;
        }
        R thResidual = heatInt - coolInt;
        if (FABSR(thResidual) >= L(0.01) * heatInt) {
            if (nIterateT < 25) {
                if (thResidual < L(0.)) {
                    thResidualLow = thResidual;
                    Tlow = c.Te;
                    if (Thigh != L(0.)) c.Te = Tlow - (Thigh - Tlow) * thResidualLow / (thResidualHigh - thResidualLow);
                    else c.Te = c.Te / L(1.2);
                } else {
                    thResidualHigh = thResidual;
                    Thigh = c.Te;
                    if (Tlow != L(0.)) c.Te = Tlow - (Thigh - Tlow) * thResidualLow / (thResidualHigh - thResidualLow);
                    else c.Te = c.Te * L(1.2);
                }
                if (nIterateT >= 6 && FABSR(Thigh - Tlow) <= L(0.002) * (Thigh + Tlow)) {
                    if (thResidual < L(0.)) { c.Te = Thigh - L(100.); Thigh = L(0.); }
                    else { c.Te = Tlow + L(100.); Tlow = L(0.); }
                }
                if (c.Te <= L(0.)) c.Te = L(1.);
                continue;
            }
            if (Thigh == L(0.)) c.Te = Tlow;
            else if (Tlow == L(0.)) c.Te = Thigh;
            else c.Te = (Thigh + Tlow) * L(0.5);
            flags |= 2 | 4;
        }
        break;
    }
    free(fl);
    if (!lgIBConv) flags |= 2 | 8;
    const R dXHI = (c.ion[1][1] - XOldHI) / XOldHI;
    if (FABSR(dXHI) <= (R)A->XHILimit) flags |= 1;
    *Te = c.Te;
    *Ne = c.Ne;
    for (int elem = 1; elem <= 30; ++elem)
        if (A->elemOn[elem])
            for (int s = 1; s <= ns; ++s) ionDen[(s - 1) * nE + A->xref[elem] - 1] = c.ion[elem][s];
    if (nT) *nT = nIterateT;
    return flags;
}

int FN(mco_balance_cell)(const mco_atomic *A, const R *J, R Hden, R *Te, R *Ne, R *ionDen, int *nT)
{
    return FN(balance_cell)(A, J, Hden, Te, Ne, ionDen, nT, NULL);
}

int FN(mco_balance_cell_pe)(const mco_atomic *A, const mco_grains *G, int nsMax, const R *J, R Hden, R *Te,
                            R *Ne, R *ionDen, int *nT, R ndust, const R *Tdust, R *gdc, R *peh, R *pot)
{
    if (G->nS * G->nSizes > MCO_MAX_GROWS) return -1;
    FN(grain_t) g;
    memset(&g, 0, sizeof g);
    g.G = G;
    g.J = J;
    g.ndust = ndust;
    g.Tdust = Tdust;
    g.ldT = nsMax + 1;
    const int fl = FN(balance_cell)(A, J, Hden, Te, Ne, ionDen, nT, &g);
    for (int k = 0; k < G->nS * G->nSizes; ++k) {
        gdc[k] = g.gdc[k];
        peh[k] = g.peh[k];
        if (pot) pot[k] = g.pot[k];
    }
    return fl;
}

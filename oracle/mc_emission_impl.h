/*
 * mc_emission_impl.h - restatement of the gas part of MOCASSIN's
 * emissionDriver for one cell (emission_mod.f90:32-173): fb_ff (253-531),
 * twoPhoton/hydro2phot/HeI2photSub (533-705), free_free/ffCoeff (709-742),
 * RecLinesEmission (744-832), forLines (837-903) and setDiffusePDF
 * (905-1311), giving the cell's recPDF row and totalLines.
 * TEST INFRASTRUCTURE ONLY.  Included by mc_balance.c after
 * mc_balance_impl.h, once per precision R.  Everything but the level
 * populations follows the reference's own mixed precision (REAL variables
 * are float, DOUBLE PRECISION ones double, in its expression order) whatever
 * R is; R selects the precision of equilibrium: float is the reference's,
 * double the GPU kernel's.
 */

#ifndef MCO_EMISSION_CONSTS
#define MCO_EMISSION_CONSTS
#define EM_HCRYD_K 157893.94f       /* continuum_mod.f90:10 */
#define EM_HCRYD 2.1799153e-11f     /* continuum_mod.f90:11 */
#define EM_CRYD 3.2898423e15f       /* continuum_mod.f90:12 */
#define EM_HPLANCK 6.6262e-27f      /* constants_mod.f90:14 */
#define EM_FOURPI (4.f * 3.141592654f)

/* locate (interpolation_mod.f90:48-81) of a float x in an increasing table */
static int em_locate(const float *xa, int n, float x)
{
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    int p = 0;
    float bd = 0.f;
    for (int i = 0; i < n; ++i) {
        const float d = xa[i] - x;
        if (d > 0.f && (p == 0 || d < bd)) { p = i + 1; bd = d; }
    }
    return p - 1 > 1 ? p - 1 : 1;
}

/* hydro2phot (emission_mod.f90:553-663) */
static double em_hydro2phot(float nu, int Z, float Te, float Ne)
{
    float nu0, y, alpha, Ay, gNu, factor;
    if (Z == 1) {
        nu0 = 0.7496f;
        y = nu / nu0;
        if (nu >= nu0) return 0.0;
        alpha = 0.8368f * powf(Te * 1.e-4f, -0.723f);
        Ay = 202.0f * (y * (1.f - y) * (1.f - powf(4.f * y * (1.f - y), 0.8f)) +
                       0.88f * powf(y * (1.f - y), 1.53f) * powf(4.f * y * (1.f - y), 0.8f));
        gNu = EM_HPLANCK * y * Ay / 8.2249f;
        gNu = gNu * 1e27f;
        float Q;
        if (Te <= 10000.f) Q = 5.31e-4f;
        else if (Te >= 20000.f) Q = 4.71e-4f;
        else {
            factor = log10f(4.71e-4f / 5.31e-4f) / log10f(2.f);
            Q = powf(10.f, log10f(5.31e-4f) + factor * log10f(Te / 10000.f));
        }
        return (double)(alpha * gNu / (1.f + (Ne * Q) / 8.23f));
    }
    nu0 = 3.00f;
    y = nu / nu0;
    if (nu >= nu0) return 0.0;
    if (Te <= 5000.f) alpha = 6.161f;
    else if (Te >= 30000.f) alpha = 2.035f;
    else if (Te > 5000.f && Te <= 10000.f) {
        factor = log10f(4.091f / 6.161f) / log10f(2.f);
        alpha = powf(10.f, log10f(6.161f) + factor * log10f(Te / 5000.f));
    } else if (Te > 10000.f && Te <= 15000.f) {
        factor = log10f(3.189f / 4.091f) / log10f(1.5f);
        alpha = powf(10.f, log10f(4.091f) + factor * log10f(Te / 10000.f));
    } else {
        factor = log10f(2.035f / 3.189f) / log10f(2.f);
        alpha = powf(10.f, log10f(3.189f) + factor * log10f(Te / 15000.f));
    }
    const float z6 = (float)(Z * Z * Z * Z * Z * Z);
    Ay = z6 * 0.9994667f * 202.0f *
         (y * (1.f - y) * (1.f - powf(4.f * y * (1.f - y), 0.8f)) +
          0.88f * powf(y * (1.f - y), 1.53f) * powf(4.f * y * (1.f - y), 0.8f));
    const float A2qZ = 8.226f * z6;
    gNu = EM_HPLANCK * y * Ay / A2qZ;
    gNu = gNu * 1e27f;
    return (double)(alpha * gNu);
}

/* getFlux's blackbody (continuum_mod.f90:359-400), the Planck function over h */
static float em_bb(float energy, float T)
{
    const float constant = 0.5250229f / EM_HPLANCK;
    if (EM_HCRYD_K * energy / T > 86.f)
        return (float)((double)(constant * energy * energy * energy) * exp((double)(-EM_HCRYD_K * energy / T)));
    const float den = expf(EM_HCRYD_K * energy / T) - 1.f;
    if (den <= 0.f) return 3.32154e-6f * energy * energy * T / EM_HPLANCK;
    return constant * energy * energy * energy / den;
}

/* ffCoeff (emission_mod.f90:726-742) */
static double em_ffcoeff(float nu, int Z, float g, float Te, float sqrTe)
{
    const double expF = exp((double)(-nu * EM_HCRYD_K / Te));
    return (double)(EM_FOURPI * 54.43f * (float)Z * (float)Z * g) * expF / (double)sqrTe;
}

/* the log10 gamma table of one series at Te (fb_ff, emission_mod.f90:
 * 306-336); g is (ntk, nlim) column-major */
static void em_gamma_at(const mco_emission *E, const float *g, int nlim, float Te, float *loc)
{
    const int ntk = E->ntk;
    if (Te <= E->tk[0]) {
        for (int i = 0; i < nlim; ++i) loc[i] = g[0 + ntk * i];
    } else if (Te >= E->tk[ntk - 1]) {
        for (int i = 0; i < nlim; ++i) loc[i] = g[(ntk - 1) + ntk * i];
    } else {
        const int t = em_locate(E->tk, ntk, Te);
        const double factor = (double)log10f(Te / E->tk[t - 1]);
        for (int i = 0; i < nlim; ++i) {
            const double aFit = (double)((g[t + ntk * i] - g[(t - 1) + ntk * i]) / log10f(E->tk[t] / E->tk[t - 1]));
            loc[i] = g[(t - 1) + ntk * i] + (float)(aFit * factor);
        }
    }
}

/* gamma between the series' edges, interpolated linearly in log (fb_ff,
 * emission_mod.f90:341-426); returns nonzero on coincident edges */
static int em_edges(const mco_emission *E, const int *edge, int nlim, const float *loc, double *gam)
{
    for (int k = 1; k <= nlim / 2; ++k) {
        const int iup = edge[2 * k - 1], ilow = edge[2 * k - 2];
        if (iup == ilow) return 1;
        const double aFit = (double)((loc[2 * k - 1] - loc[2 * k - 2]) / (E->nu[iup - 1] - E->nu[ilow - 1]));
        for (int j = ilow; j <= iup - 1; ++j) {
            const double v = (double)loc[2 * k - 1] + aFit * (double)(E->nu[j - 1] - E->nu[iup - 1]);
            gam[j - 1] = pow(10.0, v);
        }
    }
    return 0;
}
#endif

/* One cell: ion[(stage-1) * nElemUsed + xref-1] (REAL state), D the
 * cell's grains or NULL; out: the cell's recPDF row rec[nbins] and
 * *totalLines.  Returns 0, or 1 on a coincident gamma edge. */
int FN(mco_emission_cell)(const mco_atomic *A, const mco_emission *E, const mco_emission_dust *D, float Te,
                          float NeIn, const float *ion, float *rec, float *totalLines)
{
    const int nb = E->nb, nE = A->nElemUsed, ns = A->nstages;
    const float *nu = E->nu;
#define EION(z, st) ion[((st) - 1) * nE + A->xref[z] - 1]
    float NeUsed = NeIn;
    const float sqrTe = sqrtf(Te);
    const float log10Te = log10f(Te);
    if (NeUsed < 1.e-5f) NeUsed = 1.e-5f;
    const float abH = A->abun[1], abHe = A->abun[2];
    double *gHI = (double *)calloc((size_t)nb * 10, sizeof(double));
    double *gHeI = gHI + nb, *gHeII = gHI + 2 * nb, *gHv = gHI + 3 * nb;
    double *tHI = gHI + 4 * nb, *tHeI = gHI + 5 * nb, *tHeII = gHI + 6 * nb;
    double *eHI = gHI + 7 * nb, *eHeI = gHI + 8 * nb, *eHeII = gHI + 9 * nb;
    float *sHI = (float *)calloc((size_t)nb * 3, sizeof(float));
    float *sHeI = sHI + nb, *sHeII = sHI + 2 * nb;
    float lHI[64], lHeI[64], lHeII[64];
    int bad = 0;

    /* fb_ff: edges, then the Milne relation above the H, He I, He II edges */
    em_gamma_at(E, E->gHI, E->nHI, Te, lHI);
    em_gamma_at(E, E->gHeI, E->nHeI, Te, lHeI);
    em_gamma_at(E, E->gHeII, E->nHeII, Te, lHeII);
    bad |= em_edges(E, E->eHI, E->nHI, lHI, gHI);
    bad |= em_edges(E, E->eHeI, E->nHeI, lHeI, gHeI);
    bad |= em_edges(E, E->eHeII, E->nHeII, lHeII, gHeII);
    const double constant = (double)4.9874105e-6f;
    const float statW[3] = { 2.f, 0.5f, 2.f };
    for (int i = E->ipHI; i <= nb; ++i) {
        const double factor = (double)((nu[i - 1] * nu[i - 1] * nu[i - 1]) / (Te * sqrTe));
        double expF = exp((double)((-nu[i - 1] + nu[E->ipHI - 1]) * EM_HCRYD_K / Te));
        gHI[i - 1] = (double)EM_FOURPI * (double)E->sHI[i - E->ipHI] * (double)statW[0] * (double)EM_HCRYD * constant *
                     factor * expF * (double)1.e20f * (double)1.e20f;
        gHI[i - 1] = gHI[i - 1] + em_ffcoeff(nu[i - 1], 1, E->gff[i - 1], Te, sqrTe);
        if (i >= E->i0HeI) {
            expF = exp((double)((-nu[i - 1] + nu[E->ipHeI - 1]) * EM_HCRYD_K / Te));
            gHeI[i - 1] = (double)EM_FOURPI * (double)E->sHeI[i - E->i0HeI] * (double)statW[1] * (double)EM_HCRYD *
                          constant * factor * expF * (double)1.e20f * (double)1.e20f;
            gHeI[i - 1] = gHeI[i - 1] + em_ffcoeff(nu[i - 1], 1, E->gff[i - 1], Te, sqrTe);
            if (i >= E->i0HeII) {
                expF = exp((double)((-nu[i - 1] + nu[E->ipHeII - 1]) * EM_HCRYD_K / Te));
                gHeII[i - 1] = (double)EM_FOURPI * (double)E->sHeII[i - E->i0HeII] * (double)statW[2] *
                               (double)EM_HCRYD * constant * factor * expF * (double)1.e20f * (double)1.e20f;
                gHeII[i - 1] = gHeII[i - 1] + em_ffcoeff(nu[i - 1], 2, E->gffHeII[i - 1], Te, sqrTe);
            }
        }
    }
    for (int i = 1; i <= nb; ++i) {
        double g = 0.0;
        size_t off = 0;
        for (int c = 0; c < E->nfb; ++c) {
            const int *m = E->fbm + 4 * c;
            const int z = m[0], st = m[1], ip = m[2], hi = m[3];
            if (i >= ip && i < hi) {
                const double expF = exp((double)((-nu[i - 1] + nu[ip - 1]) * EM_HCRYD_K / Te));
                g = g + expF * (double)E->fbs[off + (size_t)(i - ip)] * (double)E->fbg[c] * (double)EION(z, st + 1) *
                            (double)A->abun[z];
            }
            off += (size_t)(hi - ip);
        }
        const float factor = nu[i - 1] * nu[i - 1] * nu[i - 1] / (Te * sqrTe);
        gHv[i - 1] = g * (double)EM_FOURPI * constant * (double)EM_HCRYD * (double)factor * (double)1.e20f *
                     (double)1.e20f;
    }
    /* twoPhoton */
    for (int i = 0; i < nb; ++i) {
        tHI[i] = em_hydro2phot(nu[i], 1, Te, NeUsed);
        tHeII[i] = em_hydro2phot(nu[i], 2, Te, NeUsed);
    }
    {
        const float a21 = 6.23f * powf(Te / 10000.f, -0.827f);
        int j = 1;
        for (int i = 0; i < nb; ++i) {
            const float y = nu[i] / 1.514f;
            if (y > E->ydat[j - 1]) j = j + 1;
            if (j >= 41) break;
            const float fit1 = (E->Aydat[j] - E->Aydat[j - 1]) / (E->ydat[j] - E->ydat[j - 1]);
            const float fit2 = E->Aydat[j - 1] - E->ydat[j - 1] * fit1;
            float Ay = fit1 * y + fit2;
            if (Ay < 9.2163086E-03f) Ay = 0.f;
            tHeI[i] = (double)(a21 * Ay * 0.66262f * y / 51.3f);
        }
    }
    /* emission coefficients and the partial sums of setDiffusePDF */
    float normHI = 0.f, normHeI = 0.f, normHeII = 0.f;
    const float xHp = EION(1, 2), xHe1 = EION(2, 2), xHe2 = EION(2, 3);
    for (int i = 0; i < nb; ++i) {
        eHI[i] = (gHI[i] + tHI[i]) * (double)abH * (double)xHp * (double)NeUsed + gHv[i] * (double)NeUsed;
        eHeI[i] = (gHeI[i] + tHeI[i]) * (double)abHe * (double)xHe1 * (double)NeUsed;
        eHeII[i] = (gHeII[i] + tHeII[i]) * (double)abHe * (double)xHe2 * (double)NeUsed;
        if (eHI[i] < (double)1.e-15f) eHI[i] = 0.0;
        if (eHeI[i] < (double)1.e-15f) eHeI[i] = 0.0;
        if (eHeII[i] < (double)1.e-15f) eHeII[i] = 0.0;
        const float cw = EM_CRYD * E->widFlx[i];
        sHI[i] = (float)((double)cw * eHI[i] / (double)1.e15f);
        normHI = normHI + sHI[i];
        sHeI[i] = (float)((double)cw * eHeI[i] / (double)1.e15f);
        normHeI = normHeI + sHeI[i];
        sHeII[i] = (float)((double)cw * eHeII[i] / (double)1.e15f);
        normHeII = normHeII + sHeII[i];
    }
    /* He I 2s triplet and He II Lyman lines able to ionise H */
    {
        const float T4 = Te * 1.e-4f;
        const float a2tS = 27.2f * powf(T4, -0.678f);
        const float corr = a2tS * NeUsed * abHe * xHe1 * 1.45673f * EM_HCRYD * 1e11f;
        sHeI[E->j2TsP - 1] = sHeI[E->j2TsP - 1] + corr;
        normHeI = normHeI + corr;
        float ly[16];
        for (int k = 0; k < E->nly; ++k) ly[k] = E->lyData[k];
        ly[3] = powf(10.f, -0.792f * log10f(Te) + 6.01f);
        ly[3] = ly[3] * NeUsed * xHe2 * abHe;
        for (int k = 0; k < E->nly - 1; ++k) ly[k] = ly[k] * ly[3];
        for (int k = 0; k < E->nly; ++k) {
            const int p = E->lyP[k];
            sHeII[p - 1] = sHeII[p - 1] + ly[k];
            normHeII = normHeII + ly[k];
        }
    }
    /* grains below their sublimation temperature (setDiffusePDF,
     * emission_mod.f90:1027-1097; no quantum heating) */
    float *sD = NULL, normDust = 0.f;
    if (D) {
        sD = (float *)calloc((size_t)nb, sizeof(float));
        for (int ai = 1; ai <= D->nSizes; ++ai)
            for (int s = 1; s <= D->nS; ++s) {
                const float td = D->Tdust[s + (D->nsMax + 1) * ai];
                if (!(td < D->tsub[s - 1])) continue;
                const float *xs = D->absx + ((size_t)(s - 1) * D->nSizes + (ai - 1)) * nb;
                for (int f = 0; f < nb; ++f) {
                    const float bb = em_bb(nu[f], td);
                    sD[f] = sD[f] + xs[f] * bb * E->widFlx[f] * D->weight[ai - 1] * D->gabun[s - 1];
                    normDust = normDust + xs[f] * bb * E->widFlx[f] * D->weight[ai - 1] * D->gabun[s - 1];
                }
            }
        const float cst = EM_HPLANCK * 4.f * 3.141592654f * 1.e25f * EM_CRYD;
        for (int f = 0; f < nb; ++f) sD[f] = sD[f] * cst * D->Ndust / D->Hden;
        normDust = normDust * cst * D->Ndust / D->Hden;
    }
    float normalize = normHI + normHeI + normHeII;
    if (D) normalize = normalize + normDust;
    /* recombination lines (RecLinesEmission, emission_mod.f90:744-832) */
    float normRec = 0.f;
    {
        float Hbeta = powf(10.f, -0.870f * log10Te + 3.57f);
        Hbeta = Hbeta * NeUsed * xHp * abH;
        const float Lalpha = powf(10.f, -0.897f * log10Te + 5.05f);
#define HIR(u, l) E->HIrec[((u) - 3) + 28 * ((l) - 2)]
#define HEIIR(u, l) E->HeIIrec[((u) - 3) + 28 * ((l) - 2)]
        double hi[31][9];
        for (int u = 3; u <= 15; ++u)
            for (int l = 2; l <= (u - 1 < 8 ? u - 1 : 8); ++l) hi[u][l] = HIR(u, l) * (double)Hbeta;
        hi[15][8] = hi[15][8] + (double)(abH * xHp * NeUsed * Lalpha);
        for (int u = 3; u <= 15; ++u)
            for (int l = 2; l <= (u - 1 < 8 ? u - 1 : 8); ++l) normRec = normRec + (float)hi[u][l];
        float h4686 = powf(10.f, -.997f * log10f(Te) + 5.16f);
        h4686 = h4686 * NeUsed * abHe * xHe2;
        for (int u = 3; u <= 30; ++u)
            for (int l = 2; l <= (u - 1 < 16 ? u - 1 : 16); ++l) normRec = normRec + (float)(HEIIR(u, l) * (double)h4686);
        const float T4 = Te / 10000.f;
        int denint;
        if (NeIn <= 100.f) denint = 0;
        else if (NeIn <= 1.e4f) denint = 1;
        else if (NeIn <= 1.e6f) denint = 2;
        else denint = 3;
#define HEIC(i, d, k) E->HeIc[((i) - 1) + 34 * (((d) - 1) + 3 * ((k) - 1))]
        const float p100[4] = { 1.f, 100.f, 10000.f, 1.e6f };
        for (int i = 1; i <= 34; ++i) {
            double h;
            if (denint > 0 && denint < 3) {
                const float x1 = HEIC(i, denint, 1) * powf(T4, HEIC(i, denint, 2)) * expf(HEIC(i, denint, 3) / T4);
                const float x2 =
                    HEIC(i, denint + 1, 1) * powf(T4, HEIC(i, denint + 1, 2)) * expf(HEIC(i, denint + 1, 3) / T4);
                h = (double)(x1 + ((x2 - x1) * (NeUsed - p100[denint]) / (p100[denint + 1] - p100[denint])));
            } else {
                const int d = denint == 0 ? 1 : 3;
                h = (double)(HEIC(i, d, 1) * powf(T4, HEIC(i, d, 2)) * expf(HEIC(i, d, 3) / T4));
            }
            h = h * (double)NeUsed * (double)abHe * (double)xHe1;
            normRec = normRec + (float)h;
        }
#undef HIR
#undef HEIIR
#undef HEIC
    }
    /* forbidden lines (forLines), summed in setDiffusePDF's order */
    float normFor = 0.f;
    {
        double fl[NFOR * NFOR];
        const double kL = (double)1.9865e9f;
        for (int k = 0; k < A->ncel; ++k) {
            const int elem = A->cel_meta[6 * k], st = A->cel_meta[6 * k + 1];
            const float x = EION(elem, st);
            const float up = (st < ns) ? EION(elem, st + 1) / x : 0.f;
            for (int i = 0; i < NFOR * NFOR; ++i) fl[i] = 0.0;
            FN(equilibrium)(A, k, (R)up, (R)Te, (R)NeUsed, fl);
            for (int lo = 1; lo <= NFOR; ++lo)
                for (int hi = 1; hi <= NFOR; ++hi) {
                    double v = fl[(lo - 1) + NFOR * (hi - 1)];
                    v = v * (double)A->abun[elem] * (double)x;
                    v = v * kL;
                    if (v > (double)1.e-35f) normFor = normFor + (float)v;
                }
        }
    }
    float tl = normRec + normFor;
    /* recPDF: the cumulative sum, normalised by its last entry */
    if (D) {
        for (int i = 0; i < nb; ++i)
            rec[i] = i == 0 ? sHI[i] + sHeI[i] + sHeII[i] + sD[i] : rec[i - 1] + sHI[i] + sHeI[i] + sHeII[i] + sD[i];
        free(sD);
    } else {
        for (int i = 0; i < nb; ++i) {
            const float t = sHI[i] + sHeI[i] + sHeII[i];
            rec[i] = i == 0 ? t : rec[i - 1] + t;
        }
    }
    const float last = rec[nb - 1];
    for (int i = 0; i < nb; ++i) rec[i] = rec[i] / last;
    rec[nb - 1] = 1.f;
    for (int i = 0; i < nb; ++i)
        if (rec[i] > 0.999998f) rec[i] = 1.f;
    normalize = normalize + tl;
    *totalLines = tl / normalize;
    free(gHI);
    free(sHI);
#undef EION
    return bad;
}

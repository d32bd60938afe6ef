// mc_gridio.hip - the grid state files of MOCASSIN's writeGrid and
// resetGrid (grid_mod.f90:2646-2760, 2967-3568): grid0.out (geometry,
// active map, lgConverged, lgBlack), grid1.out (Te, Ne, Hden), grid2.out
// (ionic fractions of the elements switched on) and dustGrid.out (Ndust,
// Tdust).  Host code only.
//
// Writing follows list-directed output as the reference's compiler emits it:
// a record starts with a blank, items are separated by one blank, a REAL is
// a shortest round-trip decimal chosen as that runtime chooses it, in F form
// when the value rounded to one digit lies in [0.1, 1e7) (no leading zero:
// "1000.", ".5", ".0993065") and 1P E form otherwise ("2.5E+18", "3.E+18"),
// so that the files are byte for byte the reference's.  Reading
// follows list-directed input as resetGrid does it: every READ starts a new
// record and takes values across records as needed; REAL flags are rounded
// to the nearest integer (nint).
#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mocassin_amd.h"

// the context-free error message of mc_ctx_last_error(NULL) (mc_capi.hip)
__attribute__((visibility("hidden"))) int mc_fail_global(int code, const char *fmt, ...);

namespace {

// Unsigned integers of a few hundred bits: the exact decimal values of a
// float and its neighbours.
struct Big {
    std::vector<uint32_t> d;                 // little-endian base 2^32
    explicit Big(uint64_t v = 0)
    {
        if (v) d.push_back((uint32_t)v);
        if (v >> 32) d.push_back((uint32_t)(v >> 32));
    }
    void trim() { while (!d.empty() && d.back() == 0) d.pop_back(); }
    bool zero() const { return d.empty(); }
    void mul(uint32_t m)
    {
        uint64_t c = 0;
        for (uint32_t &w : d) { const uint64_t t = (uint64_t)w * m + c; w = (uint32_t)t; c = t >> 32; }
        if (c) d.push_back((uint32_t)c);
    }
    void shl(int n) { while (n-- > 0) mul(2); }
    void add(const Big &o)
    {
        if (d.size() < o.d.size()) d.resize(o.d.size(), 0);
        uint64_t c = 0;
        for (size_t i = 0; i < d.size(); ++i) {
            const uint64_t t = (uint64_t)d[i] + (i < o.d.size() ? o.d[i] : 0) + c;
            d[i] = (uint32_t)t;
            c = t >> 32;
        }
        if (c) d.push_back((uint32_t)c);
    }
    void inc()
    {
        for (uint32_t &w : d) if (++w != 0) return;
        d.push_back(1);
    }
    void dec()                               // *this > 0
    {
        for (uint32_t &w : d) if (w-- != 0) break;
        trim();
    }
    uint32_t divmod(uint32_t m)              // *this /= m, the remainder
    {
        uint64_t r = 0;
        for (size_t i = d.size(); i-- > 0;) {
            const uint64_t t = (r << 32) | d[i];
            d[i] = (uint32_t)(t / m);
            r = t % m;
        }
        trim();
        return (uint32_t)r;
    }
    bool odd() const { return !d.empty() && (d[0] & 1); }
    bool operator==(const Big &o) const { return d == o.d; }
    uint64_t u64() const { return (d.size() > 0 ? d[0] : 0) | ((uint64_t)(d.size() > 1 ? d[1] : 0) << 32); }
};

// n * 2^f * 10^K as an integer (K >= -f)
Big scaled(uint64_t n, int f, int K)
{
    Big b(n);
    for (int i = 0; i < K; ++i) b.mul(5);
    b.shl(f + K);
    return b;
}

// The shortest decimal the reference's runtime writes for a positive float
// (its "minimize" step): the decimals of x and its two neighbours at their
// common scale, the bounds ceil((prev + x)/2) and the midpoint to next (less
// one when exact), then the shortest prefix on which the bounds differ in
// their last digit only, whose last digit is taken midway between theirs
// (rounded towards the upper bound).  Returns the digits and expo with
// value = 0.digits x 10^expo.
void shortest(float ax, uint64_t &digits, int &expo)
{
    int e2;
    const float fr = std::frexp(ax, &e2);
    uint64_t m = (uint64_t)std::ldexp(fr, 24);
    int ex = e2 - 24;
    if (ex < -149) { m >>= (-149 - ex); ex = -149; }
    // x = m 2^ex; next = (m+1) 2^ex; prev = (m-1) 2^ex, or (2m-1) 2^(ex-1)
    // at the bottom of a binade of normal numbers
    uint64_t nx_[3] = { m, m + 1, m - 1 };
    int fx[3] = { ex, ex, ex };
    if (m == (1u << 23) && ex > -149) { nx_[2] = 2 * m - 1; fx[2] = ex - 1; }
    int K = 0;
    for (int i = 0; i < 3; ++i) {
        while (nx_[i] && !(nx_[i] & 1)) { nx_[i] >>= 1; ++fx[i]; }
        K = std::max(K, -fx[i]);
    }
    Big X = scaled(nx_[0], fx[0], K), N = scaled(nx_[1], fx[1], K), P = scaled(nx_[2], fx[2], K);
    Big lo = P;
    lo.add(X);
    const bool lodd = lo.odd();
    lo.divmod(2);
    if (lodd) lo.inc();
    Big hi = N;
    hi.add(X);
    const bool hodd = hi.odd();
    hi.divmod(2);
    if (!hodd) hi.dec();
    int t = 0;
    // skip nine digits at a time while the bounds differ above them
    for (;;) {
        Big a = lo, b = hi;
        a.divmod(1000000000u);
        a.divmod(10);
        b.divmod(1000000000u);
        b.divmod(10);
        if (a == b) break;
        lo.divmod(1000000000u);
        hi.divmod(1000000000u);
        t += 9;
    }
    uint64_t res;
    for (;;) {
        Big q = hi, lq = lo;
        const uint32_t r = q.divmod(10), lr = lq.divmod(10);
        if (r != 0 && q == lq) {
            res = hi.u64() - ((r - lr) >> 1);
            break;
        }
        hi = q;
        lo = lq;
        ++t;
    }
    int z = 0;
    while (res && res % 10 == 0) { res /= 10; ++z; }
    int nd = 0;
    for (uint64_t v = res; v; v /= 10) ++nd;
    digits = res;
    expo = t - K + z + nd;
}

// one REAL as list-directed output writes it: the shortest decimal; F form
// when x rounded to one significant digit lies in [0.1, 1e7), else 1P E
void put_real(std::string &out, float x)
{
    out += ' ';
    if (std::isnan(x)) { out += "NaN"; return; }
    if (std::isinf(x)) { out += x < 0 ? "-Inf" : "Inf"; return; }
    if (x == 0.f) { out += std::signbit(x) ? "-0." : "0."; return; }
    const float ax = std::fabs(x);
    uint64_t dig;
    int expo;
    shortest(ax, dig, expo);
    const std::string ds = std::to_string(dig);
    const double a = ax;
    int e = (int)std::floor(std::log10(a));
    if (std::pow(10.0, e + 1) <= a) ++e;
    if (std::pow(10.0, e) > a) --e;
    const int expo1 = a / std::pow(10.0, e) >= 9.5 ? e + 2 : e + 1;
    if (x < 0) out += '-';
    if (expo1 >= 0 && expo1 <= 7) {
        if (expo <= 0) {
            out += '.';
            out.append((size_t)(-expo), '0');
            out += ds;
        } else if ((size_t)expo >= ds.size()) {
            out += ds;
            out.append((size_t)expo - ds.size(), '0');
            out += '.';
        } else {
            out.append(ds, 0, (size_t)expo);
            out += '.';
            out.append(ds, (size_t)expo, std::string::npos);
        }
    } else {
        const int e10 = expo - 1;
        out += ds[0];
        out += '.';
        out.append(ds, 1, std::string::npos);
        char eb[16];
        std::snprintf(eb, sizeof eb, "E%c%02d", e10 < 0 ? '-' : '+', e10 < 0 ? -e10 : e10);
        out += eb;
    }
}

void put_int(std::string &out, long v)
{
    out += ' ';
    out += std::to_string(v);
}

bool write_file(const std::string &path, const std::string &text)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    return std::fclose(f) == 0 && ok;
}

// list-directed input: each read() starts at the next record
struct Reader {
    FILE *f = nullptr;
    std::vector<std::string> tok;
    size_t pos = 0;
    bool open(const std::string &path)
    {
        f = std::fopen(path.c_str(), "rb");
        return f != nullptr;
    }
    ~Reader()
    {
        if (f) std::fclose(f);
    }
    bool next_record()
    {
        tok.clear();
        pos = 0;
        std::string line;
        int ch;
        bool any = false;
        while ((ch = std::fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            line += (char)ch;
        }
        if (!any) return false;
        size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == ',' || line[i] == '\r')) ++i;
            size_t j = i;
            while (j < line.size() && !(line[j] == ' ' || line[j] == '\t' || line[j] == ',' || line[j] == '\r')) ++j;
            if (j > i) tok.emplace_back(line, i, j - i);
            i = j;
        }
        return true;
    }
    // a READ of n values: starts at a new record, continues across records
    bool read(std::vector<std::string> &out, size_t n)
    {
        out.clear();
        if (!next_record()) return false;
        while (out.size() < n) {
            if (pos >= tok.size()) {
                if (!next_record()) return false;
                continue;
            }
            out.push_back(tok[pos++]);
        }
        return true;
    }
};

bool to_real(const std::string &s, float &v)
{
    std::string t(s);
    for (char &ch : t)
        if (ch == 'D' || ch == 'd') ch = 'E';
    char *end = nullptr;
    v = std::strtof(t.c_str(), &end);
    return end && *end == 0 && end != t.c_str();
}

bool to_int(const std::string &s, long &v)
{
    char *end = nullptr;
    v = std::strtol(s.c_str(), &end, 10);
    return end && *end == 0 && end != s.c_str();
}

int nint(float x) { return (int)std::lround(x); }

struct Elements {
    int nst = 0, nE = 0;
    std::vector<int> z, xref, nval;       // elements switched on, by atomic number
};

int elements(const mc_atomic *at, Elements &E)
{
    if (!at || !at->elemOn || !at->xref) return mc_fail_global(MC_EINVAL, "gas grids need the element list (mc_atomic)");
    if (at->nstages < 1 || at->nstages > 9 || at->nElemUsed < 1)
        return mc_fail_global(MC_EINVAL, "nstages %d or nElemUsed %d out of range", at->nstages, at->nElemUsed);
    E.nst = at->nstages;
    E.nE = at->nElemUsed;
    for (int zz = 1; zz <= 30; ++zz)
        if (at->elemOn[zz - 1]) {
            if (at->xref[zz - 1] < 1 || at->xref[zz - 1] > at->nElemUsed)
                return mc_fail_global(MC_EINVAL, "elementXref(%d) = %d outside 1..%d", zz, at->xref[zz - 1], at->nElemUsed);
            E.z.push_back(zz);
            E.xref.push_back(at->xref[zz - 1]);
            E.nval.push_back(std::min(zz + 1, at->nstages));
        }
    return MC_OK;
}

int check_desc(const mc_desc *d)
{
    if (!d || d->nx < 1 || d->ny < 1 || d->nz < 1 || d->nCells < 0 || !d->xAxis || !d->yAxis || !d->zAxis || !d->active)
        return mc_fail_global(MC_EINVAL, "incomplete grid descriptor");
    return MC_OK;
}

}  // namespace

extern "C" int mc_write_grid(const mc_desc *d, const mc_atomic *at, const char *dir, const mc_grid_state *st)
{
    int rc = check_desc(d);
    if (rc) return rc;
    if (!dir || !st || !st->lgConverged || !st->lgBlack) return mc_fail_global(MC_EINVAL, "dir, state and flags are required");
    const bool gas = (d->flags & MC_GAS) != 0, dust = (d->flags & MC_DUST) != 0;
    const size_t nc1 = (size_t)d->nCells + 1;
    Elements E;
    if (gas) {
        if ((rc = elements(at, E))) return rc;
        if (!st->Te || !st->Ne || !st->Hden || !st->ionDen) return mc_fail_global(MC_EINVAL, "gas grids need Te, Ne, Hden, ionDen");
    }
    if (dust && (!st->Ndust || !st->Tdust || st->nSpeciesMax < 0 || st->nSizes < 0))
        return mc_fail_global(MC_EINVAL, "grains need Ndust, Tdust and its shape");
    const std::string base = std::string(dir) + "/";
    std::string g0, g1, g2, gd;
    put_int(g0, 1);
    g0 += '\n';
    put_int(g0, d->nx); put_int(g0, d->ny); put_int(g0, d->nz); put_int(g0, d->nCells); put_int(g0, 0);
    put_real(g0, d->R_out);
    g0 += '\n';
    for (int i = 0; i < d->nx; ++i) { put_real(g0, d->xAxis[i]); g0 += '\n'; }
    for (int i = 0; i < d->ny; ++i) { put_real(g0, d->yAxis[i]); g0 += '\n'; }
    for (int i = 0; i < d->nz; ++i) { put_real(g0, d->zAxis[i]); g0 += '\n'; }
    const int nsm1 = st->nSpeciesMax + 1;
    for (int i = 0; i < d->nx; ++i)
        for (int j = 0; j < d->ny; ++j)
            for (int k = 0; k < d->nz; ++k) {
                const int act = d->active[i + (size_t)d->nx * (j + (size_t)d->ny * k)];
                if (act > d->nCells) return mc_fail_global(MC_EINVAL, "active entry %d beyond nCells", act);
                const size_t c = act < 0 ? 0 : (size_t)act;
                put_int(g0, act); put_int(g0, st->lgConverged[c]); put_int(g0, st->lgBlack[c]);
                g0 += '\n';
                if (gas) {
                    put_real(g1, st->Te[c]); put_real(g1, st->Ne[c]); put_real(g1, st->Hden[c]);
                    g1 += '\n';
                    for (size_t e = 0; e < E.z.size(); ++e) {
                        for (int s = 1; s <= E.nval[e]; ++s)
                            put_real(g2, st->ionDen[c + nc1 * ((size_t)(E.xref[e] - 1) + (size_t)E.nE * (s - 1))]);
                        g2 += '\n';
                    }
                }
                if (dust) {
                    put_real(gd, st->Ndust[c]);
                    gd += '\n';
                    for (int ai = 0; ai <= st->nSizes; ++ai) {
                        for (int s = 0; s <= st->nSpeciesMax; ++s) {
                            put_real(gd, st->Tdust[s + (size_t)nsm1 * (ai + (size_t)(st->nSizes + 1) * c)]);
                            gd += "  ";
                        }
                        gd += '\n';
                    }
                }
            }
    if (dust) {
        gd += "  \n";
        gd += " Total dust mass [1.e45 g]: ";
        put_real(gd, st->totalDustMass);
        gd += "\n Total dust mass [Msol]: ";
        put_real(gd, st->totalDustMass * 5.028e11f);
        gd += '\n';
    }
    if (!write_file(base + "grid0.out", g0)) return mc_fail_global(MC_EINVAL, "cannot write %sgrid0.out", base.c_str());
    if (gas && (!write_file(base + "grid1.out", g1) || !write_file(base + "grid2.out", g2)))
        return mc_fail_global(MC_EINVAL, "cannot write %sgrid1.out / grid2.out", base.c_str());
    if (dust && !write_file(base + "dustGrid.out", gd))
        return mc_fail_global(MC_EINVAL, "cannot write %sdustGrid.out", base.c_str());
    return MC_OK;
}

extern "C" int mc_read_grid(const mc_desc *d, const mc_atomic *at, const char *dir, mc_grid_state *st)
{
    int rc = check_desc(d);
    if (rc) return rc;
    if (!dir || !st || !st->lgConverged || !st->lgBlack) return mc_fail_global(MC_EINVAL, "dir, state and flags are required");
    const bool gas = (d->flags & MC_GAS) != 0, dust = (d->flags & MC_DUST) != 0;
    const size_t nc1 = (size_t)d->nCells + 1;
    Elements E;
    if (gas) {
        if ((rc = elements(at, E))) return rc;
        if (!st->Te || !st->Ne || !st->Hden || !st->ionDen) return mc_fail_global(MC_EINVAL, "gas grids need Te, Ne, Hden, ionDen");
    }
    if (dust && (!st->Ndust || !st->Tdust || st->nSpeciesMax < 0 || st->nSizes < 0))
        return mc_fail_global(MC_EINVAL, "grains need Ndust, Tdust and its shape");
    const std::string base = std::string(dir) + "/";
    Reader r0, r1, r2, rd;
    if (!r0.open(base + "grid0.out")) return mc_fail_global(MC_EINVAL, "cannot open %sgrid0.out", base.c_str());
    if (gas && (!r1.open(base + "grid1.out") || !r2.open(base + "grid2.out")))
        return mc_fail_global(MC_EINVAL, "cannot open %sgrid1.out / grid2.out", base.c_str());
    if (dust && !rd.open(base + "dustGrid.out")) return mc_fail_global(MC_EINVAL, "cannot open %sdustGrid.out", base.c_str());
    std::vector<std::string> t;
    long iv;
    float fv;
#define BAD(what) return mc_fail_global(MC_ETABLE, "%s: %s", base.c_str(), what)
    if (!r0.read(t, 1) || !to_int(t[0], iv)) BAD("grid0.out: no grid count");
    if (iv != 1) BAD("grid0.out: more than one grid (sub-grids are outside this path)");
    if (!r0.read(t, 6)) BAD("grid0.out: no grid header");
    long h[5];
    for (int q = 0; q < 5; ++q)
        if (!to_int(t[q], h[q])) BAD("grid0.out: bad grid header");
    if (h[0] != d->nx || h[1] != d->ny || h[2] != d->nz || h[3] != d->nCells)
        return mc_fail_global(MC_ETABLE, "%sgrid0.out: grid %ldx%ldx%ld with %ld cells, the descriptor %dx%dx%d with %d",
                              base.c_str(), h[0], h[1], h[2], h[3], d->nx, d->ny, d->nz, d->nCells);
    const float *ax[3] = { d->xAxis, d->yAxis, d->zAxis };
    const int n3[3] = { d->nx, d->ny, d->nz };
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < n3[a]; ++i) {
            if (!r0.read(t, 1) || !to_real(t[0], fv)) BAD("grid0.out: short axis");
            if (fv != ax[a][i]) return mc_fail_global(MC_ETABLE, "%sgrid0.out: axis %d entry %d differs from the descriptor", base.c_str(), a, i + 1);
        }
    std::vector<float> p((size_t)std::max(E.nst, 1), 0.f);
    const int nsm1 = st->nSpeciesMax + 1;
    for (int i = 0; i < d->nx; ++i)
        for (int j = 0; j < d->ny; ++j)
            for (int k = 0; k < d->nz; ++k) {
                long act;
                float p0, p00;
                if (!r0.read(t, 3) || !to_int(t[0], act) || !to_real(t[1], p0) || !to_real(t[2], p00))
                    BAD("grid0.out: short cell records");
                if (act != d->active[i + (size_t)d->nx * (j + (size_t)d->ny * k)])
                    return mc_fail_global(MC_ETABLE, "%sgrid0.out: active(%d,%d,%d) differs from the descriptor", base.c_str(), i + 1, j + 1, k + 1);
                const size_t c = act < 0 ? 0 : (size_t)act;
                st->lgConverged[c] = nint(p0);
                st->lgBlack[c] = nint(p00);
                if (gas) {
                    float v[3];
                    if (!r1.read(t, 3)) BAD("grid1.out: short");
                    for (int q = 0; q < 3; ++q)
                        if (!to_real(t[q], v[q])) BAD("grid1.out: bad value");
                    st->Te[c] = v[0]; st->Ne[c] = v[1]; st->Hden[c] = v[2];
                    // every stage takes p: stages beyond the element's count
                    // keep the previous element's values, as resetGrid's
                    // whole-array assignment does
                    for (size_t e = 0; e < E.z.size(); ++e) {
                        if (!r2.read(t, (size_t)E.nval[e])) BAD("grid2.out: short");
                        for (int s = 0; s < E.nval[e]; ++s)
                            if (!to_real(t[s], p[s])) BAD("grid2.out: bad value");
                        for (int s = 0; s < E.nst; ++s)
                            st->ionDen[c + nc1 * ((size_t)(E.xref[e] - 1) + (size_t)E.nE * s)] = p[s];
                    }
                }
                if (dust) {
                    if (!rd.read(t, 1) || !to_real(t[0], fv)) BAD("dustGrid.out: short");
                    for (int ai = 0; ai <= st->nSizes; ++ai) {
                        if (!rd.read(t, (size_t)nsm1)) BAD("dustGrid.out: short Tdust");
                        for (int s = 0; s < nsm1; ++s) {
                            float tv;
                            if (!to_real(t[s], tv)) BAD("dustGrid.out: bad Tdust");
                            st->Tdust[s + (size_t)nsm1 * (ai + (size_t)(st->nSizes + 1) * c)] = tv;
                        }
                    }
                    st->Ndust[c] = fv;
                }
            }
#undef BAD
    return MC_OK;
}

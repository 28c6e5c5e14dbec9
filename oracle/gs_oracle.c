/*
 * gs_oracle.c -- CPU restatement of the Gramsci counting path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see gs_oracle.h).
 * Built with -O2 -ffp-contract=off so that every FP64 expression is evaluated
 * exactly as written, in the reference's operand order.
 *
 * Pinned by: tests/test_oracle.py against the golden fixtures written by the
 * reference binary (oracle/_ref/gramsci, built from the reference sources by
 * oracle/Makefile), against scipy.spatial.cKDTree pair counts, and against
 * small-case brute force written independently in numpy.
 */
#include "gs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { OR_OK = 0, OR_EINVAL = -1, OR_ENOMEM = -2 };

/* ---------------------------------------------------------------- bins -- */

void or_radial_bins(const or_params* p, double* edges)
{
    /* gramsci.F90:88-96.  float(nbins) is a default REAL, exact for nbins<=127. */
    for (int i = 1; i <= p->nbins + 1; ++i) {
        int j = i - 1;
        double v;
        if (p->log_bins) {
            double lo = log10(p->rmin), hi = log10(p->rmax);
            v = pow(10.0, lo + ((double)j * (hi - lo)) / (double)(float)p->nbins);
        } else {
            v = p->rmin + ((double)j * (p->rmax - p->rmin)) / (double)(float)p->nbins;
        }
        edges[i - 1] = v;
    }
}

void or_normalise(const double* w, int64_t n_data, int64_t n_rand, double* wn)
{
    /* gramsci.F90:1390-1391 */
    double sd = 0.0, sr = 0.0;
    for (int64_t i = 0; i < n_data; ++i) sd += w ? w[i] : 1.0;
    for (int64_t i = n_data; i < n_data + n_rand; ++i) sr += w ? w[i] : 1.0;
    for (int64_t i = 0; i < n_data; ++i) wn[i] = (w ? w[i] : 1.0) / sd;
    for (int64_t i = n_data; i < n_data + n_rand; ++i) wn[i] = (-1.0 * (w ? w[i] : 1.0)) / sr;
}

/* s = (xj-xi)^2 + (yj-yi)^2 + (zj-zi)^2, left to right: kdtree2.F90:1686-1690
 * (sd starts at 0) and dist() gramsci.F90:1532 give the same value. */
static inline double pair_s(const double* xi, const double* xj)
{
    double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
    return dx * dx + dy * dy + dz * dz;
}

/* bin from the distance d, gramsci.F90:136-145 and :334-338 */
static inline int rbin_of_d(const or_params* p, double d)
{
    double delta_r, inv_delta_r, log_rmin;
    if (p->log_bins) {
        delta_r = (log10(p->rmax) - log10(p->rmin)) / (double)p->nbins;
        inv_delta_r = 1.0 / delta_r;
        log_rmin = log10(p->rmin);
        return (int)floor((log10(d) - log_rmin) * inv_delta_r) + 1;
    }
    delta_r = (p->rmax - p->rmin) / (double)p->nbins;
    inv_delta_r = 1.0 / delta_r;
    return (int)floor((d - p->rmin) * inv_delta_r) + 1;
}

int or_pair_rbin(const or_params* p, const double* xi, const double* xj)
{
    double s = pair_s(xi, xj);
    if (!(s <= p->rmax * p->rmax)) return 0;       /* r-ball, inclusive */
    double d = sqrt(s);
    if (d <= p->rmin) return 0;                    /* gramsci.F90:316 */
    if (d >= p->rmax) return 0;                    /* gramsci.F90:317 */
    return rbin_of_d(p, d);
}

int or_pair_mubin(const or_params* p, const double* ci, const double* nj)
{
    /* gramsci.F90:320-328; mu_scale = 0.5*nmu (:145) */
    if (p->nmu < 2) return 1;
    double mu_scale = (double)(0.5f * (float)p->nmu);
    double m1 = 0.5 * (ci[0] + nj[0]), m2 = 0.5 * (ci[1] + nj[1]), m3 = 0.5 * (ci[2] + nj[2]);
    double w1 = ci[0] - nj[0], w2 = ci[1] - nj[1], w3 = ci[2] - nj[2];
    double theta = m1 * w1 + m2 * w2 + m3 * w3;
    theta = theta / (sqrt(m1 * m1 + m2 * m2 + m3 * m3) * sqrt(w1 * w1 + w2 * w2 + w3 * w3));
    int mu = (int)floor((theta + 1.0) * mu_scale) + 1;
    mu = (int)(int8_t)mu;                          /* mu1 is integer(int8) (:25) */
    if (mu < 1) mu = 1;
    if (mu > p->nmu) mu = p->nmu;
    return mu;
}

static int check_params(const or_params* p)
{
    if (!p || p->nbins < 1 || p->nbins > 127 || p->nmu < 1 || p->nmu > 127) return OR_EINVAL;
    if (!(p->rmax > p->rmin) || p->rmin < 0.0) return OR_EINVAL;
    if (p->log_bins && !(p->rmin > 0.0)) return OR_EINVAL;
    return OR_OK;
}

/* ---------------------------------------------------------------- grid -- */
/* A plain uniform grid of cell size rmax, used only to enumerate candidate
 * neighbours; the kept set is decided by or_pair_rbin alone. */
typedef struct {
    double lo[3], h;
    int64_t dim[3];
    int64_t* start;  /* ncell+1 */
    int64_t* idx;    /* n */
} grid_t;

static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

static int grid_build(grid_t* g, const double* xyz, int64_t n, double h)
{
    memset(g, 0, sizeof(*g));
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (xyz[3 * i + k] < lo[k]) lo[k] = xyz[3 * i + k];
            if (xyz[3 * i + k] > hi[k]) hi[k] = xyz[3 * i + k];
        }
    if (n == 0) for (int k = 0; k < 3; ++k) lo[k] = hi[k] = 0.0;
    g->h = h;
    int64_t ncell = 1;
    for (int k = 0; k < 3; ++k) {
        g->lo[k] = lo[k];
        int64_t d = (int64_t)((hi[k] - lo[k]) / h) + 1;
        if (d > 512) d = 512;
        g->dim[k] = d;
        ncell *= d;
    }
    g->start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    g->idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!g->start || !g->idx || !cell) { free(cell); return OR_ENOMEM; }
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3];
        for (int k = 0; k < 3; ++k)
            c[k] = clampi((int64_t)((xyz[3 * i + k] - g->lo[k]) / h), 0, g->dim[k] - 1);
        cell[i] = (c[2] * g->dim[1] + c[1]) * g->dim[0] + c[0];
        g->start[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) g->start[c + 1] += g->start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    if (!fill) { free(cell); return OR_ENOMEM; }
    memcpy(fill, g->start, sizeof(int64_t) * (size_t)ncell);
    for (int64_t i = 0; i < n; ++i) g->idx[fill[cell[i]]++] = i;   /* ascending i within a cell */
    free(fill);
    free(cell);
    return OR_OK;
}

static void grid_free(grid_t* g) { free(g->start); free(g->idx); }

/* Visit every j != i whose cell is within one cell of i's (covers d <= h). */
#define GRID_FOR_NEIGHBOURS(g, xyz, i, j, ...)                                           \
    do {                                                                                 \
        int64_t c_[3];                                                                   \
        for (int k_ = 0; k_ < 3; ++k_)                                                   \
            c_[k_] = clampi((int64_t)(((xyz)[3 * (i) + k_] - (g)->lo[k_]) / (g)->h), 0,   \
                            (g)->dim[k_] - 1);                                           \
        for (int64_t z_ = c_[2] - 1; z_ <= c_[2] + 1; ++z_) {                            \
            if (z_ < 0 || z_ >= (g)->dim[2]) continue;                                   \
            for (int64_t y_ = c_[1] - 1; y_ <= c_[1] + 1; ++y_) {                        \
                if (y_ < 0 || y_ >= (g)->dim[1]) continue;                               \
                for (int64_t x_ = c_[0] - 1; x_ <= c_[0] + 1; ++x_) {                    \
                    if (x_ < 0 || x_ >= (g)->dim[0]) continue;                           \
                    int64_t cc_ = (z_ * (g)->dim[1] + y_) * (g)->dim[0] + x_;            \
                    for (int64_t q_ = (g)->start[cc_]; q_ < (g)->start[cc_ + 1]; ++q_) { \
                        int64_t j = (g)->idx[q_];                                        \
                        if (j == (i)) continue;                                          \
                        __VA_ARGS__                                                      \
                    }                                                                    \
                }                                                                        \
            }                                                                            \
        }                                                                                \
    } while (0)

/* Cell size for candidate enumeration: rmax, widened so the grid stays <= 512^3. */
static double grid_h(const double* xyz, int64_t n, double rmax)
{
    double lo = 1e300, hi = -1e300;
    for (int64_t i = 0; i < 3 * n; ++i) { if (xyz[i] < lo) lo = xyz[i]; if (xyz[i] > hi) hi = xyz[i]; }
    double h = rmax;
    if (n > 0 && (hi - lo) / h > 500.0) h = (hi - lo) / 500.0;
    return h * (1.0 + 1e-12) + 1e-300;
}

/* ---------------------------------------------------------------- 2PCF -- */

int or_2pcf(const double* xyz, const double* w, const int32_t* buf,
            int64_t n_data, int64_t n_rand, const or_params* p,
            uint64_t* cnt, double* wsum, double* n2, double* n3, uint64_t* n_overflow)
{
    int rc = check_params(p);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    const int nb = p->nbins, nm = p->nmu, nbm = nb * nm;
    double* wn = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    long double* acc = (long double*)calloc((size_t)nbm * 5, sizeof(long double)); /* 3 wsum + n2 + n3 */
    uint64_t* c = (uint64_t*)calloc((size_t)nbm * 3, sizeof(uint64_t));
    if (!wn || !acc || !c) { free(wn); free(acc); free(c); return OR_ENOMEM; }
    or_normalise(w, n_data, n_rand, wn);
    grid_t g;
    if ((rc = grid_build(&g, xyz, n, grid_h(xyz, n, p->rmax)))) { free(wn); free(acc); free(c); grid_free(&g); return rc; }
    uint64_t over = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (buf && buf[i] == 1) continue;             /* gramsci.F90:363 */
        const int ri = i >= n_data;
        GRID_FOR_NEIGHBOURS(&g, xyz, i, j, {
            int b = or_pair_rbin(p, &xyz[3 * i], &xyz[3 * j]);
            if (b == 0) continue;
            if (b > nb) { over++; continue; }
            int m = or_pair_mubin(p, &xyz[3 * i], &xyz[3 * j]);
            const int rj = j >= n_data;
            const int cls = ri + rj;
            const int64_t o = (int64_t)(b - 1) * nm + (m - 1);
            c[cls * nbm + o]++;
            double wi = w ? w[i] : 1.0, wj = w ? w[j] : 1.0;
            acc[cls * nbm + o] += (long double)(wi * wj);
            double prod = wn[i] * wn[j];                 /* gramsci.F90:374,376 */
            acc[3 * nbm + o] += (long double)prod;
            if (ri && rj) acc[4 * nbm + o] += (long double)prod;
        });
    }
    /* output layout [cls][bin][mu] */
    for (int64_t k = 0; k < 3 * nbm; ++k) {
        if (cnt) cnt[k] = c[k];
        if (wsum) wsum[k] = (double)acc[k];
    }
    for (int64_t k = 0; k < nbm; ++k) {
        if (n2) n2[k] = (double)acc[3 * nbm + k];
        if (n3) n3[k] = (double)acc[4 * nbm + k];
    }
    if (n_overflow) *n_overflow = over;
    grid_free(&g);
    free(wn); free(acc); free(c);
    return OR_OK;
}

/* ------------------------------------------------------------- configs -- */

int or_n3cfg(int nb) { return nb * (nb + 1) * (nb + 2) / 6; }
int or_n4cfg(int nb) { return nb * (nb + 1) * (nb + 2) * (nb + 3) / 24; }

int or_bintable3(int nb, int a, int b, int c)
{
    /* gramsci.F90:1459-1474: index of the sorted triple in the i<=j<=k enumeration */
    int t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    int cfg = 0;
    for (int i = 1; i <= nb; ++i)
        for (int j = i; j <= nb; ++j)
            for (int k = j; k <= nb; ++k) {
                ++cfg;
                if (i == a && j == b && k == c) return cfg;
            }
    return 0;
}

int or_bintable4(int nb, int a, int b, int c, int d)
{
    /* gramsci.F90:1402-1420: only non-decreasing quadruples are filled */
    if (!(a <= b && b <= c && c <= d)) return 0;
    int cfg = 0;
    for (int i = 1; i <= nb; ++i)
        for (int j = i; j <= nb; ++j)
            for (int k = j; k <= nb; ++k)
                for (int l = k; l <= nb; ++l) {
                    ++cfg;
                    if (i == a && j == b && k == c && l == d) return cfg;
                }
    return 0;
}

/* --------------------------------------------------------------- graph -- */
/* The reference's neighbour graph (gramsci.F90:288-348): for each point the
 * kept neighbours, sorted by id (sorting_module.F90:36-47), with their bin
 * and their mu bin (output(i)%mu, 1 when nmu < 2). */
typedef struct {
    int64_t* off;   /* n+1 */
    int64_t* id;
    int8_t*  bin;
    int8_t*  mu;
} graph_t;

static int cmp_pair(const void* a, const void* b)
{
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

static int graph_build(graph_t* G, const double* xyz, int64_t n, const or_params* p)
{
    memset(G, 0, sizeof(*G));
    grid_t g;
    int rc = grid_build(&g, xyz, n, grid_h(xyz, n, p->rmax));
    if (rc) { grid_free(&g); return rc; }
    G->off = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    if (!G->off) { grid_free(&g); return OR_ENOMEM; }
    for (int64_t i = 0; i < n; ++i) {
        int64_t k = 0;
        GRID_FOR_NEIGHBOURS(&g, xyz, i, j, {
            int b = or_pair_rbin(p, &xyz[3 * i], &xyz[3 * j]);
            if (b >= 1 && b <= p->nbins) ++k;
        });
        G->off[i + 1] = G->off[i] + k;
    }
    int64_t tot = G->off[n];
    int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)(tot > 0 ? tot : 1));
    G->id = (int64_t*)malloc(sizeof(int64_t) * (size_t)(tot > 0 ? tot : 1));
    G->bin = (int8_t*)malloc((size_t)(tot > 0 ? tot : 1));
    G->mu = (int8_t*)malloc((size_t)(tot > 0 ? tot : 1));
    if (!tmp || !G->id || !G->bin || !G->mu) { free(tmp); grid_free(&g); return OR_ENOMEM; }
    for (int64_t i = 0; i < n; ++i) {
        int64_t k = 0;
        int64_t* t = tmp + 2 * G->off[i];
        GRID_FOR_NEIGHBOURS(&g, xyz, i, j, {
            int b = or_pair_rbin(p, &xyz[3 * i], &xyz[3 * j]);
            if (b >= 1 && b <= p->nbins) {
                t[2 * k] = j;
                t[2 * k + 1] = b | ((int64_t)or_pair_mubin(p, &xyz[3 * i], &xyz[3 * j]) << 8);
                ++k;
            }
        });
        qsort(t, (size_t)k, 2 * sizeof(int64_t), cmp_pair);
        for (int64_t q = 0; q < k; ++q) {
            G->id[G->off[i] + q] = t[2 * q];
            G->bin[G->off[i] + q] = (int8_t)(t[2 * q + 1] & 255);
            G->mu[G->off[i] + q] = (int8_t)(t[2 * q + 1] >> 8);
        }
    }
    free(tmp);
    grid_free(&g);
    return OR_OK;
}

static void graph_free(graph_t* G) { free(G->off); free(G->id); free(G->bin); free(G->mu); }

/* find_dist (gramsci.F90:1579-1609): bin of id2 in id1's list, 0 if absent */
static inline int find_dist(const graph_t* G, int64_t id1, int64_t id2)
{
    int64_t lo = G->off[id1], hi = G->off[id1 + 1] - 1;
    while (lo <= hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (G->id[mid] == id2) return G->bin[mid];
        if (G->id[mid] < id2) lo = mid + 1; else hi = mid - 1;
    }
    return 0;
}

/* ---------------------------------------------------------------- 3PCF -- */

/* find_normal (gramsci.F90:1611-1625) in its own precision: mu11, mu22 are
 * default reals (the subtraction mu1-0.5 is single precision, the division by
 * the double mu_scale and the "- 1.0" double, rounded to single on
 * assignment); the rest of the expression is single precision, evaluated left
 * to right; the result is stored in an integer(int8) before the clamp.  A
 * negative radicand gives NaN, whose floor is not representable: any integer
 * conversion of it ends in bin 1 after the int8 store and the clamp. */
int or_normal_bin(int nmu, int mu1, int mu2)
{
    const double mu_scale = (double)(0.5f * (float)nmu);
    const float a = (float)((double)((float)mu1 - 0.5f) / mu_scale - 1.0);
    const float b = (float)((double)((float)mu2 - 0.5f) / mu_scale - 1.0);
    volatile float x = 0.75f - a * a;
    x = x - b * b;
    x = x + a * b;
    if (!(x >= 0.0f)) return 1;
    volatile float y = 1.1547f * sqrtf(x);
    y = y * (float)nmu;
    int v = (int)floorf(y) + 1;
    v = (int)(int8_t)v;
    if (v < 1) v = 1;
    if (v > nmu) v = nmu;
    return v;
}

/* 3PCF, all configurations (gramsci.F90:463-545); with nmu >= 2 (RSD, :1219)
 * the triplet goes to mu bin find_normal(mu(k1), mu(k2)) (:495-500). */
int or_3pcf(const double* xyz, const double* w, const int32_t* buf,
            int64_t n_data, int64_t n_rand, const or_params* p,
            uint64_t* cnt, double* n2, double* n3, uint64_t* n_cand)
{
    int rc = check_params(p);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    const int nb = p->nbins, ncfg = or_n3cfg(nb), nm = p->nmu;
    const int64_t nh = (int64_t)ncfg * nm;
    int* tab = (int*)malloc(sizeof(int) * (size_t)(nb * nb * nb));
    int* ntab = (int*)malloc(sizeof(int) * (size_t)(nm * nm));
    double* wn = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    long double* acc = (long double*)calloc((size_t)nh * 2, sizeof(long double));
    uint64_t* c = (uint64_t*)calloc((size_t)nh * 4, sizeof(uint64_t));
    graph_t G;
    memset(&G, 0, sizeof(G));
    if (!tab || !ntab || !wn || !acc || !c) { rc = OR_ENOMEM; goto out; }
    for (int a = 1; a <= nb; ++a)
        for (int b = 1; b <= nb; ++b)
            for (int d = 1; d <= nb; ++d) tab[((a - 1) * nb + (b - 1)) * nb + (d - 1)] = or_bintable3(nb, a, b, d);
    for (int a = 1; a <= nm; ++a)
        for (int b = 1; b <= nm; ++b) ntab[(a - 1) * nm + (b - 1)] = nm >= 2 ? or_normal_bin(nm, a, b) : 1;
    or_normalise(w, n_data, n_rand, wn);
    if ((rc = graph_build(&G, xyz, n, p))) goto out;
    uint64_t cand = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (buf && buf[i] == 1) continue;
        const int64_t o = G.off[i], k = G.off[i + 1] - o;
        cand += (uint64_t)(k * (k - 1) / 2);
        const int ri = i >= n_data;
        for (int64_t k1 = 0; k1 < k; ++k1) {
            const int64_t id1 = G.id[o + k1];
            const int ind1 = G.bin[o + k1];
            const double wi_w1 = wn[i] * wn[id1];           /* gramsci.F90:482 */
            for (int64_t k2 = k1 + 1; k2 < k; ++k2) {
                const int64_t id2 = G.id[o + k2];
                const int ind2 = G.bin[o + k2];
                const int ind3 = find_dist(&G, id1, id2);
                if (ind3 == 0) continue;
                const int cfg = tab[((ind1 - 1) * nb + (ind2 - 1)) * nb + (ind3 - 1)];
                const int mu = ntab[(G.mu[o + k1] - 1) * nm + (G.mu[o + k2] - 1)];
                const int64_t h = (int64_t)(cfg - 1) * nm + (mu - 1);
                const int nr = ri + (id1 >= n_data) + (id2 >= n_data);
                c[h * 4 + nr]++;
                const double prod = wi_w1 * wn[id2];
                acc[h] += (long double)prod;                    /* N2 (:501, :507) */
                if (nr == 3) acc[nh + h] -= (long double)prod;  /* N3 (:498, :505) */
            }
        }
    }
    for (int64_t q = 0; q < nh; ++q) {
        if (n2) n2[q] = (double)acc[q];
        if (n3) n3[q] = (double)acc[nh + q];
    }
    if (cnt) memcpy(cnt, c, sizeof(uint64_t) * (size_t)nh * 4);
    if (n_cand) *n_cand = cand;
out:
    graph_free(&G);
    free(tab); free(ntab); free(wn); free(acc); free(c);
    return rc;
}

/* Equilateral 3PCF (gramsci.F90:402-461, -equi): triangles whose three sides
 * share one bin.  With nmu >= 2 only N2 is accumulated, at mu bin
 * find_normal(mu(k1), mu(k2)) (:429-431); N3 stays 0. */
int or_3pcf_equi(const double* xyz, const double* w, const int32_t* buf,
                 int64_t n_data, int64_t n_rand, const or_params* p,
                 uint64_t* cnt, double* n2, double* n3)
{
    int rc = check_params(p);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    const int nb = p->nbins, nm = p->nmu;
    const int64_t nh = (int64_t)nb * nm;
    int* ntab = (int*)malloc(sizeof(int) * (size_t)(nm * nm));
    double* wn = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    long double* acc = (long double*)calloc((size_t)nh * 2, sizeof(long double));
    uint64_t* c = (uint64_t*)calloc((size_t)nh * 4, sizeof(uint64_t));
    graph_t G;
    memset(&G, 0, sizeof(G));
    if (!ntab || !wn || !acc || !c) { rc = OR_ENOMEM; goto out; }
    for (int a = 1; a <= nm; ++a)
        for (int b = 1; b <= nm; ++b) ntab[(a - 1) * nm + (b - 1)] = nm >= 2 ? or_normal_bin(nm, a, b) : 1;
    or_normalise(w, n_data, n_rand, wn);
    if ((rc = graph_build(&G, xyz, n, p))) goto out;
    for (int64_t i = 0; i < n; ++i) {
        if (buf && buf[i] == 1) continue;                       /* :409 */
        const int64_t o = G.off[i], k = G.off[i + 1] - o;
        const int ri = i >= n_data;
        for (int64_t k1 = 0; k1 < k; ++k1) {
            const int ind1 = G.bin[o + k1];
            const int64_t id1 = G.id[o + k1];
            for (int64_t k2 = k1 + 1; k2 < k; ++k2) {           /* :419-420 */
                if (G.bin[o + k2] != ind1) continue;            /* :422 */
                const int64_t id2 = G.id[o + k2];
                if (find_dist(&G, id1, id2) != ind1) continue;  /* :425-426 */
                const int mu = ntab[(G.mu[o + k1] - 1) * nm + (G.mu[o + k2] - 1)];
                const int64_t h = (int64_t)(ind1 - 1) * nm + (mu - 1);
                const int nr = ri + (id1 >= n_data) + (id2 >= n_data);
                c[h * 4 + nr]++;
                const double prod = wn[i] * wn[id1] * wn[id2];
                acc[h] += (long double)prod;                                /* :431, :437 */
                if (nm < 2 && nr == 3) acc[nh + h] -= (long double)prod;   /* :434-435 */
            }
        }
    }
    for (int64_t q = 0; q < nh; ++q) {
        if (n2) n2[q] = (double)acc[q];
        if (n3) n3[q] = (double)acc[nh + q];
    }
    if (cnt) memcpy(cnt, c, sizeof(uint64_t) * (size_t)nh * 4);
out:
    graph_free(&G);
    free(ntab); free(wn); free(acc); free(c);
    return rc;
}

/* ---------------------------------------------------------------- 4PCF -- */

int or_4pcf(const double* xyz, const double* w, const int32_t* buf,
            int64_t n_data, int64_t n_rand, const or_params* p,
            uint64_t* cnt, double* n2, double* n3, double* disc_n, double* disc_r,
            uint64_t* n_quads)
{
    int rc = check_params(p);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    const int nb = p->nbins, ncfg = or_n4cfg(nb);
    const int64_t ncell = (int64_t)ncfg * nb * nb;
    int* tab = (int*)malloc(sizeof(int) * (size_t)(nb * nb * nb * nb));
    double* wn = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    long double* acc = (long double*)calloc((size_t)ncell * 14, sizeof(long double));
    uint64_t* c = (uint64_t*)calloc((size_t)ncell * 16, sizeof(uint64_t));
    graph_t G;
    memset(&G, 0, sizeof(G));
    if (!tab || !wn || !acc || !c) { rc = OR_ENOMEM; goto out; }
    for (int a = 1; a <= nb; ++a)
        for (int b = 1; b <= nb; ++b)
            for (int d = 1; d <= nb; ++d)
                for (int e = 1; e <= nb; ++e)
                    tab[(((a - 1) * nb + (b - 1)) * nb + (d - 1)) * nb + (e - 1)] = or_bintable4(nb, a, b, d, e);
    or_normalise(w, n_data, n_rand, wn);
    if ((rc = graph_build(&G, xyz, n, p))) goto out;
    uint64_t quads = 0;
    long double* N2 = acc;
    long double* N3 = acc + ncell;
    long double* DN = acc + 2 * ncell;      /* [6][ncell] */
    long double* DR = acc + 8 * ncell;      /* [6][ncell] */
    for (int64_t i = 0; i < n; ++i) {
        if (buf && buf[i] == 1) continue;
        const int64_t o = G.off[i], k = G.off[i + 1] - o;
        if (k <= 3) continue;                                   /* gramsci.F90:675 */
        const int ri = i >= n_data;
        const double wi = wn[i];
        for (int64_t k1 = 0; k1 < k; ++k1) {
            const int ind1 = G.bin[o + k1];
            const int64_t id1 = G.id[o + k1];
            const int r1 = id1 >= n_data;
            const double w1 = wn[id1];
            for (int64_t k2 = k1 + 1; k2 < k; ++k2) {
                const int ind2 = G.bin[o + k2];
                const int64_t id2 = G.id[o + k2];
                const int ind4 = find_dist(&G, id1, id2);
                if (ind4 == 0) continue;
                const int r2 = id2 >= n_data;
                const double w2 = wn[id2];
                quads += (uint64_t)(k - k2 - 1);
                for (int64_t k3 = k2 + 1; k3 < k; ++k3) {
                    const int ind3 = G.bin[o + k3];
                    const int64_t id3 = G.id[o + k3];
                    const int ind5 = find_dist(&G, id2, id3);
                    if (ind5 == 0) continue;
                    const int ind6 = find_dist(&G, id1, id3);
                    if (ind6 == 0) continue;
                    const int bin = tab[(((ind1 - 1) * nb + (ind4 - 1)) * nb + (ind5 - 1)) * nb + (ind3 - 1)];
                    if (bin <= 0) continue;
                    const int r3 = id3 >= n_data;
                    const double w3 = wn[id3];
                    const int64_t cell = ((int64_t)(bin - 1) * nb + (ind2 - 1)) * nb + (ind6 - 1);
                    c[cell * 16 + (ri | r1 << 1 | r2 << 2 | r3 << 3)]++;
                    const double prod = wi * w1 * w2 * w3;        /* gramsci.F90:710 */
                    N2[cell] += (long double)prod;
                    if (ri && r1 && r2 && r3) N3[cell] += (long double)prod;
                    /* disc terms, gramsci.F90:754-818 (index t-1 <-> discN(:,:,:,t)) */
                    if (ri) {
                        DN[3 * ncell + cell] += (long double)(w1 * w2);
                        if (r1 && r2) DR[3 * ncell + cell] += (long double)(w1 * w2);
                        DN[4 * ncell + cell] += (long double)(w3 * w2);
                        if (r3 && r2) DR[4 * ncell + cell] += (long double)(w3 * w2);
                        DN[5 * ncell + cell] += (long double)(w3 * w1);
                        if (r3 && r1) DR[5 * ncell + cell] += (long double)(w3 * w1);
                    }
                    if (r3) {
                        DN[1 * ncell + cell] += (long double)(wi * w2);
                        if (ri && r2) DR[1 * ncell + cell] += (long double)(wi * w2);
                        DN[0 * ncell + cell] += (long double)(wi * w1);
                        if (ri && r1) DR[0 * ncell + cell] += (long double)(wi * w1);
                    }
                    if (r1) {
                        DN[2 * ncell + cell] += (long double)(wi * w3);
                        if (ri && r3) DR[2 * ncell + cell] += (long double)(wi * w3);
                    }
                }
            }
        }
    }
    for (int64_t q = 0; q < ncell; ++q) {
        if (n2) n2[q] = (double)N2[q];
        if (n3) n3[q] = (double)N3[q];
    }
    for (int64_t q = 0; q < 6 * ncell; ++q) {
        if (disc_n) disc_n[q] = (double)DN[q];
        if (disc_r) disc_r[q] = (double)DR[q];
    }
    if (cnt) memcpy(cnt, c, sizeof(uint64_t) * (size_t)ncell * 16);
    if (n_quads) *n_quads = quads;
out:
    graph_free(&G);
    free(tab); free(wn); free(acc); free(c);
    return rc;
}

/* Regular-tetrahedron 4PCF (gramsci.F90:548-660, -tetra): apex i and
 * neighbours k1 < k2 < k3 (id order) whose six sides all share one bin b.
 * No minimum neighbour count (unlike query_4pcf_all2).  Per bin:
 *   cnt[b][16]  counts by random pattern (bit0 apex, bit1 k1, bit2 k2, bit3 k3)
 *   n2[b], n3[b]      N2(b,1,3) and N3(b,1,3) (N3: all four random, positive)
 *   disc_n, disc_r [6][b]  discN(b,1,1,t), discR(b,1,1,t): triple products
 *                           (:590-631). */
int or_4pcf_tetra(const double* xyz, const double* w, const int32_t* buf,
                  int64_t n_data, int64_t n_rand, const or_params* p,
                  uint64_t* cnt, double* n2, double* n3, double* disc_n, double* disc_r)
{
    int rc = check_params(p);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    const int nb = p->nbins;
    double* wn = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    long double* acc = (long double*)calloc((size_t)nb * 14, sizeof(long double));
    uint64_t* c = (uint64_t*)calloc((size_t)nb * 16, sizeof(uint64_t));
    graph_t G;
    memset(&G, 0, sizeof(G));
    if (!wn || !acc || !c) { rc = OR_ENOMEM; goto out; }
    or_normalise(w, n_data, n_rand, wn);
    if ((rc = graph_build(&G, xyz, n, p))) goto out;
    long double* N2 = acc;
    long double* N3 = acc + nb;
    long double* DN = acc + 2 * nb;     /* [6][nb] */
    long double* DR = acc + 8 * nb;     /* [6][nb] */
    for (int64_t i = 0; i < n; ++i) {
        if (buf && buf[i] == 1) continue;
        const int64_t o = G.off[i], k = G.off[i + 1] - o;
        const int ri = i >= n_data;
        const double wi = wn[i];
        for (int64_t k1 = 0; k1 < k; ++k1) {
            const int b = G.bin[o + k1];
            const int64_t id1 = G.id[o + k1];
            const int r1 = id1 >= n_data;
            const double w1 = wn[id1];
            for (int64_t k2 = k1 + 1; k2 < k; ++k2) {
                if (G.bin[o + k2] != b) continue;                       /* :564 */
                const int64_t id2 = G.id[o + k2];
                if (find_dist(&G, id1, id2) != b) continue;             /* :566-567 */
                const int r2 = id2 >= n_data;
                const double w2 = wn[id2];
                for (int64_t k3 = k2 + 1; k3 < k; ++k3) {
                    if (G.bin[o + k3] != b) continue;                   /* :571 */
                    const int64_t id3 = G.id[o + k3];
                    if (find_dist(&G, id2, id3) != b) continue;         /* :575-576 */
                    if (find_dist(&G, id1, id3) != b) continue;         /* :578-579 */
                    const int r3 = id3 >= n_data;
                    const double w3 = wn[id3];
                    const int q = b - 1;
                    c[q * 16 + (ri | r1 << 1 | r2 << 2 | r3 << 3)]++;
                    if (ri) {
                        DN[5 * nb + q] += (long double)(wi * w3 * w1);
                        if (r3 && r1) DR[5 * nb + q] += (long double)(wi * w3 * w1);
                    }
                    if (r3) {
                        DN[0 * nb + q] += (long double)(w3 * wi * w1);
                        if (ri && r1) DR[0 * nb + q] += (long double)(w3 * wi * w1);
                    }
                    if (ri) {
                        DN[3 * nb + q] += (long double)(wi * w1 * w2);
                        if (r1 && r2) DR[3 * nb + q] += (long double)(wi * w1 * w2);
                    }
                    if (r3) {
                        DN[1 * nb + q] += (long double)(w3 * wi * w2);
                        if (ri && r2) DR[1 * nb + q] += (long double)(w3 * wi * w2);
                    }
                    if (r1) {
                        DN[2 * nb + q] += (long double)(w1 * wi * w3);
                        if (ri && r3) DR[2 * nb + q] += (long double)(w1 * wi * w3);
                    }
                    if (ri) {
                        DN[4 * nb + q] += (long double)(wi * w3 * w2);
                        if (r3 && r2) DR[4 * nb + q] += (long double)(wi * w3 * w2);
                    }
                    const double prod = wi * w1 * w2 * w3;               /* :633 */
                    N2[q] += (long double)prod;
                    if (ri && r1 && r2 && r3) N3[q] += (long double)prod; /* :634-636 */
                }
            }
        }
    }
    for (int q = 0; q < nb; ++q) {
        if (n2) n2[q] = (double)N2[q];
        if (n3) n3[q] = (double)N3[q];
    }
    for (int q = 0; q < 6 * nb; ++q) {
        if (disc_n) disc_n[q] = (double)DN[q];
        if (disc_r) disc_r[q] = (double)DR[q];
    }
    if (cnt) memcpy(cnt, c, sizeof(uint64_t) * (size_t)nb * 16);
out:
    graph_free(&G);
    free(wn); free(acc); free(c);
    return rc;
}

/* -------------------------------------------------------------- adjoint -- */

int or_adjoint_2pcf(const double* xyz, const double* w, const int32_t* buf,
                    int64_t n_data, int64_t n_rand, const or_params* p,
                    const double* gN2, const double* gN3, double* dldw)
{
    /* N2 = sum over kept ordered pairs (i apex) of wn_i wn_j, N3 the same over RR
     * pairs (gramsci.F90:373-376).  dL/dwn_k collects every ordered pair that
     * contains k, as apex or as neighbour; the raw-weight gradient follows from
     * wn = w/Wd (data) and wn = -w/Wr (randoms). */
    int rc = check_params(p);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    const int nm = p->nmu;
    double* wn = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    long double* G = (long double*)calloc((size_t)(n > 0 ? n : 1), sizeof(long double));
    if (!wn || !G) { free(wn); free(G); return OR_ENOMEM; }
    or_normalise(w, n_data, n_rand, wn);
    grid_t g;
    if ((rc = grid_build(&g, xyz, n, grid_h(xyz, n, p->rmax)))) { free(wn); free(G); grid_free(&g); return rc; }
    for (int64_t i = 0; i < n; ++i) {
        if (buf && buf[i] == 1) continue;
        const int ri = i >= n_data;
        GRID_FOR_NEIGHBOURS(&g, xyz, i, j, {
            int b = or_pair_rbin(p, &xyz[3 * i], &xyz[3 * j]);
            if (b == 0 || b > p->nbins) continue;
            int m = or_pair_mubin(p, &xyz[3 * i], &xyz[3 * j]);
            const int rj = j >= n_data;
            const int64_t o = (int64_t)(b - 1) * nm + (m - 1);
            long double c = gN2[o];
            if (ri && rj) c += gN3[o];
            /* d(wn_i wn_j)/dwn_i = wn_j, d/dwn_j = wn_i */
            G[i] += c * wn[j];
            G[j] += c * wn[i];
        });
    }
    long double sd = 0, sr = 0, Wd = 0, Wr = 0;
    for (int64_t k = 0; k < n; ++k) {
        const double wk = w ? w[k] : 1.0;
        if (k < n_data) { sd += (long double)wn[k] * G[k]; Wd += wk; }
        else { sr += (long double)wn[k] * G[k]; Wr += wk; }
    }
    for (int64_t k = 0; k < n; ++k)
        dldw[k] = k < n_data ? (double)((G[k] - sd) / Wd) : (double)(-(G[k] + sr) / Wr);
    grid_free(&g);
    free(wn); free(G);
    return OR_OK;
}

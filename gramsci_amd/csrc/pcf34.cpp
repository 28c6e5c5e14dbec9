// pcf34.cpp -- C ABI for the 3PCF and 4PCF paths (gs_count_3pcf, gs_count_3pcf_equi,
// gs_count_4pcf, gs_count_4pcf_tetra):
// grid, stencil, configuration tables, kernel launch and the host-side
// combination of exact class counts / sums into the reference's N2, N3,
// discN, discR.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gs_api_internal.h"
#include "gs_kernels.h"

using namespace gs;

namespace gs {

void extra_release(GsExtra& x)
{
    x.params.release();
    x.out.release();
    x.queue.release();
    x.dbl.release();
    x.i32.release();
    x.big.release();
}

// bintable for the 3PCF (gramsci.F90:1459-1474), 0-based configuration of a
// sorted triple in the i <= j <= k enumeration, for every permutation.
std::vector<uint16_t> table3(int nb)
{
    std::vector<uint16_t> t((size_t)nb * nb * nb, 0);
    int cfg = 0;
    for (int i = 0; i < nb; ++i)
        for (int j = i; j < nb; ++j)
            for (int k = j; k < nb; ++k) {
                const int p[6][3] = {{i, j, k}, {i, k, j}, {j, i, k}, {j, k, i}, {k, i, j}, {k, j, i}};
                for (auto& q : p) t[((size_t)q[0] * nb + q[1]) * nb + q[2]] = (uint16_t)cfg;
                ++cfg;
            }
    return t;
}

// The triplet kernel's table: 4 * configuration (the histogram index of the configuration's class 0).
// Empty when 4 * configurations does not fit 16 bits (then the table would not fit in LDS either).
std::vector<uint16_t> table3_hist(int nb)
{
    const long ncfg = (long)nb * (nb + 1) * (nb + 2) / 6;
    if (4 * ncfg > 65535) return {};
    std::vector<uint16_t> t = table3(nb);
    for (auto& x : t) x = (uint16_t)(4 * x);
    return t;
}

// bintable for the 4PCF (gramsci.F90:1402-1420): 1-based configuration for
// non-decreasing quadruples only, 0 elsewhere.
static std::vector<uint16_t> table4(int nb)
{
    std::vector<uint16_t> t((size_t)nb * nb * nb * nb, 0);
    int cfg = 0;
    for (int i = 0; i < nb; ++i)
        for (int j = i; j < nb; ++j)
            for (int k = j; k < nb; ++k)
                for (int l = k; l < nb; ++l) t[(((size_t)i * nb + j) * nb + k) * nb + l] = (uint16_t)(++cfg);
    return t;
}

static void fill_stats(gs_stats* st, hipEvent_t* ev, const unsigned long long* s, int m)
{
    if (!st) return;
    memset(st, 0, sizeof(*st));
    float ms_grid = 0, ms_count = 0;
    (void)hipEventElapsedTime(&ms_grid, ev[0], ev[1]);
    (void)hipEventElapsedTime(&ms_count, ev[2], ev[3]);
    st->ms_grid = ms_grid;
    st->ms_count = ms_count;
    st->ms_kernel_avg = ms_count;
    st->n_candidates = s[0];
    st->n_recheck = s[1];
    st->n_overflow = s[2];
    st->n_units = s[3];
    st->n_binned = s[4];
    st->n_evaluated = s[3];
    st->cell_div = m;
}

}  // namespace gs

// ====================================================================== 3PCF

namespace gs {

// find_normal (gramsci.F90:1611-1625) tabulated over its two int8 arguments,
// 0-based: mu11/mu22 are default reals (mu1-0.5 in single precision, the
// division by the double mu_scale and "- 1.0" in double, rounded to single on
// assignment), the rest of the expression is single precision left to right,
// and floor()+1 is stored in an int8 before the clamp.  A negative radicand
// (NaN) ends in bin 1 whatever integer its floor converts to.
std::vector<uint8_t> normal_table(int nm)
{
    std::vector<uint8_t> t((size_t)nm * nm, 0);
    const double mu_scale = (double)(0.5f * (float)nm);
    for (int a = 1; a <= nm; ++a)
        for (int b = 1; b <= nm; ++b) {
            const float x = (float)((double)((float)a - 0.5f) / mu_scale - 1.0);
            const float y = (float)((double)((float)b - 0.5f) / mu_scale - 1.0);
            volatile float r = 0.75f - x * x;
            r = r - y * y;
            r = r + x * y;
            int v = 1;
            if (r >= 0.0f) {
                volatile float q = 1.1547f * std::sqrt((float)r);
                q = q * (float)nm;
                v = (int)(int8_t)((int)std::floor((float)q) + 1);
                v = v < 1 ? 1 : (v > nm ? nm : v);
            }
            t[(size_t)(a - 1) * nm + (b - 1)] = (uint8_t)(v - 1);
        }
    return t;
}

// LDS plan of the triplet kernel for a.ncfg/a.nm: histogram replicas and
// neighbour capacity within 64 KiB; a histogram too large for LDS goes to
// global atomics instead.
int triplet_wg(double kmean)
{
    const char* ev = getenv("GS_T3_WG");        // tuning override (unset or empty: automatic)
    if (ev && *ev) {
        const int w = atoi(ev);
        if (w == 64 || w == 128 || w == 256) return w;
    }
    // 256 threads per apex measured fastest also for short neighbour lists (64 batched regions,
    // mean ~100 neighbours: 108 ms per pass with 256 threads, 153 ms with 64), once work items are
    // fetched in batches; smaller work-groups stay available for other shapes
    (void)kmean;
    return 256;
}

int triplet_batch(int64_t n, int nblocks)
{
    const char* ev = getenv("GS_APEX_BATCH");   // tuning override (unset or empty: automatic)
    if (ev && *ev) return std::max(1, std::min(64, atoi(ev)));
    return (int)std::max<int64_t>(1, std::min<int64_t>(16, n / (128 * (int64_t)std::max(nblocks, 1))));
}

int plan_triplet_lds(TripArgs& a, int nb, bool fp64, bool weighted)
{
    const size_t nh = (size_t)a.ncfg * a.nm * 4;
    const size_t budget = 64 * 1024;
    auto lds = [&]() { return triplet3_lds_bytes(a, nb, fp64, weighted); };
    a.lds_hist = 1;
    a.reps = 1;
    while (a.reps < 32 && nh * a.reps * 2 * 4 <= 16 * 1024) a.reps *= 2;
    a.kmax = 4096;
    while (a.kmax > 512 && lds() > budget) a.kmax -= 256;
    while (a.reps > 1 && lds() > budget) a.reps /= 2;
    if (lds() > budget) {
        a.lds_hist = 0;
        a.reps = 1;
        a.kmax = 4096;
        while (a.kmax > 256 && lds() > budget) a.kmax -= 256;
    }
    if (lds() > budget) return fail(GS_ELIMIT, "3PCF tables do not fit in LDS");
    return GS_OK;
}

// The last pass of the 3PCF: apexes whose neighbours outnumber every LDS list were queued
// (d_list, count d_cnt[0], largest neighbour count d_cnt[1]); they run here with their
// neighbour lists in device memory, one list per work-group, and the histogram in global
// memory.  Reads the count (synchronising the stream) and launches nothing if it is 0.
// `base` carries the pass's tables and outputs; d_queue is 8 work counters it may clear.
int triplet_glist_pass(const TripArgs& base, int nb, bool fp64, bool weighted, const int32_t* d_list,
                       const int32_t* d_cnt, int32_t* d_queue, DevBuf<unsigned char>& scratch, int n_cu,
                       hipStream_t st, int* n_apex)
{
    int32_t h[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(h, d_cnt, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "3PCF queued apex count");
    if (n_apex) *n_apex = h[0];
    if (h[0] <= 0) return GS_OK;
    TripArgs a = base;
    a.items = d_list;
    a.n_items = h[0];
    a.n_items_dev = nullptr;
    a.ovf = nullptr;
    a.ovf_n = nullptr;
    a.ovf_kmax = nullptr;
    a.kmax = (std::max(h[1], 2) + 3) / 4 * 4;       // every list holds its apex's neighbours
    a.wg = kWG;
    a.batch = 1;
    a.queue = d_queue;
    const size_t stride = (a.kmax * triplet3_glist_entry_bytes(a, fp64, weighted) + 255) / 256 * 256;
    a.glist = (void*)d_list;        // selects the device-memory variant (LDS sizes and occupancy query)
    // the LDS holds no list here: the histogram stays in LDS (replicated) where it fits
    a.lds_hist = 1;
    a.reps = 1;
    const size_t nh = (size_t)a.ncfg * (a.nm > 1 ? a.nm : 1) * 4;
    while (a.reps < 32 && nh * a.reps * 2 * 4 <= 16 * 1024) a.reps *= 2;
    while (a.reps > 1 && triplet3_lds_bytes(a, nb, fp64, weighted) > 64 * 1024) a.reps /= 2;
    if (triplet3_lds_bytes(a, nb, fp64, weighted) > 64 * 1024) a.lds_hist = 0;
    const int occ = triplet3_occupancy(a, nb, fp64, weighted);
    const size_t budget = (size_t)1 << 30;           // device memory for the lists of concurrent apexes
    const int nblk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)h[0], (size_t)n_cu * occ, budget / stride}));
    int rc = scratch.reserve((size_t)nblk * stride);
    if (rc) return rc;
    a.glist = scratch.ptr;
    a.glist_stride = (int64_t)stride;
    e = hipMemsetAsync(d_queue, 0, 8 * sizeof(int32_t), st);
    if (e == hipSuccess) e = launch_triplet3(a, nb, fp64, weighted, nblk, st);
    if (e != hipSuccess) return hip_fail(e, "3PCF device-memory pass");
    return GS_OK;
}

// N2/N3 of the 3PCF from the class sums S[q][4] (q = configuration x mu bin):
// N2 = sum over classes of (-1)^r S_r / (Wd^(3-r) Wr^r) (gramsci.F90:501, :507, :431, :437);
// N3 = S_3 / Wr^3 (:498, :505, :434) unless with_n3 is false (the equilateral RSD
// branch accumulates no N3, :429-431)
void combine3(const long double* S, int nq, long double Wd, long double Wr, bool with_n3, double* n2, double* n3)
{
    for (int q = 0; q < nq; ++q) {
        long double v2 = 0, v3 = 0;
        for (int r = 0; r <= 3; ++r) {
            const long double sv = S[q * 4 + r];
            if (sv == 0) continue;
            long double den = 1;
            for (int t = 0; t < 3 - r; ++t) den *= Wd;
            for (int t = 0; t < r; ++t) den *= Wr;
            v2 += ((r & 1) ? -sv : sv) / den;
        }
        if (with_n3 && Wr > 0) v3 = S[q * 4 + 3] / (Wr * Wr * Wr);
        if (n2) n2[q] = (double)v2;
        if (n3) n3[q] = (double)v3;
    }
}

// Shared body of gs_count_3pcf (all configurations) and gs_count_3pcf_equi.
static int count3(gs_ctx* c, const gs_params* p, bool equi, uint64_t* cnt, double* wsum, double* n2, double* n3,
                  gs_stats* st)
{
    if (!c) return fail(GS_EINVAL, "NULL context");
    int rc = check_params(p, 3);
    if (rc) return rc;
    if ((rc = ctx_set_device(c))) return rc;
    const int nb = p->nbins, nm = p->nmu;
    const int ncfg = equi ? nb : nb * (nb + 1) * (nb + 2) / 6;
    const int nh = ncfg * nm * 4;
    const bool fp64 = p->precision == GS_FP64;
    CtxView v = view(c);
    const bool weighted = v.has_w;
    const int m = p->cell_div > 0 ? p->cell_div : 2;
    const int sx = apex_cell_sx();

    (void)hipEventRecord(v.ev[0], v.stream);
    double h;
    if ((rc = ctx_build_grid(c, p->rmax, m, sx, &h))) return rc;
    (void)hipEventRecord(v.ev[1], v.stream);
    v = view(c);
    const int mst = (int)std::ceil(p->rmax / h - 1e-12);
    std::vector<Row> rows = full_stencil(h, h / sx, p->rmax, mst);
    BinDesc B;
    if ((rc = fill_bins(*p, 1.5 * p->rmax, ctx_xmax(c), v.wmax_d, v.wmax_r, B))) return rc;
    std::vector<uint16_t> tab = equi ? std::vector<uint16_t>() : table3_hist(nb);
    if (!equi && tab.empty()) return fail(GS_ELIMIT, "3PCF with %d bins: configuration table too large", nb);
    std::vector<uint8_t> ntab = nm > 1 ? normal_table(nm) : std::vector<uint8_t>();

    TripArgs a{};
    a.ncfg = ncfg;
    a.nm = nm;
    a.log_bins = p->log_bins != 0;
    a.equi = equi ? 1 : 0;
    a.nrows = (int)rows.size();      // the LDS plan holds the stencil rows' table
    if ((rc = plan_triplet_lds(a, nb, fp64, weighted))) return rc;

    GsExtra& X = *v.extra;
    const size_t tab_off = ((sizeof(BinDesc) + 15) / 16) * 16;
    const size_t ntab_off = tab_off + ((tab.size() * 2 + 15) / 16) * 16;
    const size_t rows_off = ntab_off + ((ntab.size() + 15) / 16) * 16;
    const size_t nout = (size_t)nh + 2 * (size_t)nh + 8;
    if ((rc = X.params.reserve(rows_off + rows.size() * sizeof(Row) + 16)) || (rc = X.out.reserve(nout)) ||
        (rc = X.queue.reserve(16)))
        return rc;
    hipError_t e = hipMemcpyAsync(X.params.ptr, &B, sizeof(B), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && !tab.empty())
        e = hipMemcpyAsync(X.params.ptr + tab_off, tab.data(), tab.size() * 2, hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && !ntab.empty())
        e = hipMemcpyAsync(X.params.ptr + ntab_off, ntab.data(), ntab.size(), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && !rows.empty())
        e = hipMemcpyAsync(X.params.ptr + rows_off, rows.data(), rows.size() * sizeof(Row), hipMemcpyHostToDevice,
                           v.stream);
    if (e == hipSuccess) e = hipMemsetAsync(X.out.ptr, 0, nout * sizeof(unsigned long long), v.stream);
    if (e == hipSuccess) e = hipMemsetAsync(X.queue.ptr, 0, 16 * sizeof(int32_t), v.stream);
    if (e != hipSuccess) return hip_fail(e, "3PCF parameter upload");

    a.pw = v.pw;
    a.flags = v.flags;
    a.orig = v.orig;
    a.cell_start = v.cell_start;
    a.rows = (const Row*)(X.params.ptr + rows_off);
    a.bins = (const BinDesc*)X.params.ptr;
    a.cfg_tab = tab.empty() ? nullptr : (const uint16_t*)(X.params.ptr + tab_off);
    a.ntab = ntab.empty() ? nullptr : (const uint8_t*)(X.params.ptr + ntab_off);
    a.queue = X.queue.ptr;
    a.err = X.queue.ptr + 8;
    a.out_cnt = X.out.ptr;
    a.out_w = X.out.ptr + nh;
    a.out_stats = X.out.ptr + 3 * nh;
    a.g = v.g;
    a.g.m = mst;
    a.nrows = (int)rows.size();
    const int n = (int)(v.n_data + v.n_rand);
    a.n_items = n;

    // Two passes.  The neighbour list in LDS bounds the work-groups a CU can hold, so the
    // first pass sizes it for the 99.5th percentile of the neighbour counts of ~8192
    // sampled apexes (+5%) and hands every apex with more neighbours to a second pass over
    // just those apexes, with the largest list that fits.  An apex is counted in exactly
    // one pass, so the sums are those of one pass.
    // An apex whose neighbours outnumber even the second pass's list is queued once more, for a
    // last pass with its list in device memory (triplet_glist_pass): no neighbour count is refused.
    const int kmax_full = a.kmax;
    int kmax1 = kmax_full;
    const int nsample = std::min(n, 8192);
    const int stride = nsample > 0 ? n / nsample : 1;
    // i32: [0] second-pass count, [1, n] second-pass apexes, [n+1] last-pass count, [n+2] its largest
    // neighbour count, [n+3, 2n+2] last-pass apexes, then the sampled neighbour counts
    if ((rc = X.i32.reserve(2 * (size_t)n + 3 + nsample))) return rc;
    int32_t* d_list2 = X.i32.ptr + 1;
    int32_t* d_cnt3 = X.i32.ptr + n + 1;
    int32_t* d_list3 = X.i32.ptr + n + 3;
    if (nsample > 0) {
        int32_t* d_cnt = X.i32.ptr + 2 * n + 3;
        std::vector<int32_t> kc(nsample);
        e = launch_nbr_count(a, stride, nsample, nm > 1 ? 0 : 1, d_cnt, v.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(kc.data(), d_cnt, nsample * sizeof(int32_t), hipMemcpyDeviceToHost, v.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
        if (e != hipSuccess) return hip_fail(e, "3PCF neighbour sample");
        std::vector<int32_t> ks;
        double ksum = 0.0;
        for (int32_t k : kc)
            if (k >= 0) { ks.push_back(k); ksum += k; }
        if (!ks.empty()) {
            a.wg = triplet_wg(ksum / (double)ks.size());
            const size_t q = std::min(ks.size() - 1, (size_t)std::ceil(0.995 * (ks.size() - 1)));
            std::nth_element(ks.begin(), ks.begin() + q, ks.end());
            const int k995 = ks[q];
            kmax1 = std::min(kmax_full, std::max(128, (int)std::ceil(1.05 * k995 / 64.0) * 64));
        }
    }
    {
        const char* ev = getenv("GS_T3_KMAX");      // tuning override (unset or empty: automatic)
        if (ev && *ev) kmax1 = std::max(64, std::min(kmax_full, atoi(ev)));
    }
    {
        const char* ev = getenv("GS_T3_KMAX_LDS");  // test override: a shorter second-pass list
        if (ev && *ev) {
            a.kmax = std::max(64, std::min(kmax_full, atoi(ev)));
            kmax1 = std::min(kmax1, a.kmax);
        }
    }
    const bool two_pass = kmax1 < a.kmax;
    (void)hipEventRecord(v.ev[2], v.stream);
    if (n > 0) {
        e = hipMemsetAsync(X.i32.ptr, 0, sizeof(int32_t), v.stream);
        if (e == hipSuccess) e = hipMemsetAsync(d_cnt3, 0, 2 * sizeof(int32_t), v.stream);
        if (e != hipSuccess) return hip_fail(e, "3PCF overflow counters");
        TripArgs a1 = a;
        a1.ovf_kmax = d_cnt3 + 1;
        if (two_pass) {
            a1.kmax = kmax1;
            a1.ovf = d_list2;
            a1.ovf_n = X.i32.ptr;
        } else {
            a1.ovf = d_list3;
            a1.ovf_n = d_cnt3;
        }
        const int occ = triplet3_occupancy(a1, nb, fp64, weighted);
        const int nblk1 = std::max(1, std::min(v.n_cu * occ, n));
        a1.batch = triplet_batch(n, nblk1);
        e = launch_triplet3(a1, nb, fp64, weighted, nblk1, v.stream);
        if (e != hipSuccess) return hip_fail(e, "3PCF kernel launch");
        if (two_pass) {
            int32_t nov = 0;
            e = hipMemcpyAsync(&nov, X.i32.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, v.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
            if (e != hipSuccess) return hip_fail(e, "3PCF first pass");
            if (nov > 0) {
                TripArgs a2 = a;
                a2.items = d_list2;
                a2.n_items = nov;
                a2.wg = kWG;            // the apexes with the most neighbours, one per fetch
                a2.batch = 1;
                a2.ovf = d_list3;
                a2.ovf_n = d_cnt3;
                a2.ovf_kmax = d_cnt3 + 1;
                e = hipMemsetAsync(X.queue.ptr, 0, 8 * sizeof(int32_t), v.stream);   // work counters only
                const int occ2 = triplet3_occupancy(a2, nb, fp64, weighted);
                if (e == hipSuccess)
                    e = launch_triplet3(a2, nb, fp64, weighted, std::max(1, std::min(v.n_cu * occ2, nov)), v.stream);
                if (e != hipSuccess) return hip_fail(e, "3PCF second pass");
            }
        }
        if ((rc = triplet_glist_pass(a, nb, fp64, weighted, d_list3, d_cnt3, X.queue.ptr, X.big, v.n_cu, v.stream,
                                     nullptr)))
            return rc;
    }
    (void)hipEventRecord(v.ev[3], v.stream);
    std::vector<unsigned long long> h_out(nout);
    e = hipMemcpyAsync(h_out.data(), X.out.ptr, nout * sizeof(unsigned long long), hipMemcpyDeviceToHost, v.stream);
    int32_t err = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&err, X.queue.ptr + 8, sizeof(int32_t), hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) return hip_fail(e, "3PCF kernel");
    if (err) return fail(GS_EHIP, "3PCF: an apex's neighbour list overflowed with no later pass (internal error)");

    std::vector<long double> S(nh);
    for (int k = 0; k < nh; ++k) {
        if (cnt) cnt[k] = h_out[k];
        S[k] = weighted ? fx_value(h_out[nh + 2 * k], h_out[nh + 2 * k + 1], B.fx_shift3) : (long double)h_out[k];
        if (wsum) wsum[k] = (double)S[k];
    }
    combine3(S.data(), ncfg * nm, v.wsum_d, v.wsum_r, !(equi && nm > 1), n2, n3);
    fill_stats(st, v.ev, h_out.data() + 3 * nh, mst);
    if (st) {
        st->n_evaluated = h_out[3 * nh + 4];
        st->n_binned = 0;           // the binned triplets are the counts' total
        for (int k = 0; k < nh; ++k) st->n_binned += h_out[k];
    }
    return GS_OK;
}

}  // namespace gs

extern "C" int gs_count_3pcf(gs_ctx* c, const gs_params* p, uint64_t* cnt, double* wsum, double* n2, double* n3,
                             gs_stats* st)
{
    return count3(c, p, false, cnt, wsum, n2, n3, st);
}

extern "C" int gs_count_3pcf_equi(gs_ctx* c, const gs_params* p, uint64_t* cnt, double* wsum, double* n2,
                                  double* n3, gs_stats* st)
{
    return count3(c, p, true, cnt, wsum, n2, n3, st);
}

// ====================================================================== 4PCF

namespace gs {

// Shared body of gs_count_4pcf (perimeter configurations) and gs_count_4pcf_tetra.
static int count4(gs_ctx* c, const gs_params* p, bool tetra, uint64_t* cnt, double* n2, double* n3, double* disc_n,
                  double* disc_r, gs_stats* st)
{
    if (!c) return fail(GS_EINVAL, "NULL context");
    int rc = check_params(p, 4);
    if (rc) return rc;
    if ((rc = ctx_set_device(c))) return rc;
    const int nb = p->nbins;
    if (!tetra && nb > 16) return fail(GS_ELIMIT, "4PCF supports nbins <= 16 (got %d)", nb);
    const int ncfg = tetra ? nb : nb * (nb + 1) * (nb + 2) * (nb + 3) / 24;
    const int ncell = tetra ? nb : ncfg * nb * nb;
    CtxView v = view(c);
    const bool weighted = v.has_w;
    const int m = p->cell_div > 0 ? p->cell_div : 2;
    const int sx = apex_cell_sx();

    (void)hipEventRecord(v.ev[0], v.stream);
    double h;
    if ((rc = ctx_build_grid(c, p->rmax, m, sx, &h))) return rc;
    (void)hipEventRecord(v.ev[1], v.stream);
    v = view(c);
    const int mst = (int)std::ceil(p->rmax / h - 1e-12);
    std::vector<Row> rows = full_stencil(h, h / sx, p->rmax, mst);
    BinDesc B;
    if ((rc = fill_bins(*p, 1.5 * p->rmax, ctx_xmax(c), v.wmax_d, v.wmax_r, B))) return rc;
    std::vector<uint16_t> tab = tetra ? std::vector<uint16_t>() : table4(nb);

    // LDS plan of the second pass: the longest list (a power of two, 64..256, with its K x K link
    // matrix) that fits gfx950's 160 KiB of LDS per work-group beside the tables and, where it
    // fits in 16 KiB, the count histogram.  Apexes with more neighbours go to a last pass with
    // device-memory lists, so no neighbour count is refused.
    const size_t budget = 160 * 1024;
    int kmax = 256;
    static_assert(kWG >= 256, "the LDS-list sort of k_quad4 takes one thread per list entry");
    bool lds_hist = (size_t)ncell * 64 <= 16 * 1024;
    while (kmax > 64 && quad4_lds_bytes(nb, ncell, kmax, lds_hist, tetra, (int)rows.size()) > budget) kmax /= 2;
    if (quad4_lds_bytes(nb, ncell, kmax, lds_hist, tetra, (int)rows.size()) > budget) {
        lds_hist = false;
        kmax = 256;
        while (kmax > 64 && quad4_lds_bytes(nb, ncell, kmax, lds_hist, tetra, (int)rows.size()) > budget) kmax /= 2;
    }
    if (quad4_lds_bytes(nb, ncell, kmax, lds_hist, tetra, (int)rows.size()) > budget)
        return fail(GS_ELIMIT, "4PCF tables do not fit in LDS");
    {
        const char* ev = getenv("GS_T4_KMAX_LDS");  // test override: a shorter second-pass list
        if (ev && *ev)
            while (kmax > 64 && kmax > atoi(ev)) kmax /= 2;
    }

    GsExtra& X = *v.extra;
    const size_t tab_off = ((sizeof(BinDesc) + 15) / 16) * 16;
    const size_t rows_off = tab_off + ((tab.size() * 2 + 15) / 16) * 16;
    const size_t ncnt = (size_t)ncell * 16;
    const size_t nout = ncnt + 8;
    const size_t nw = weighted ? (size_t)ncell * 16 + (size_t)6 * ncell * 4 : 0;
    if ((rc = X.params.reserve(rows_off + rows.size() * sizeof(Row) + 16)) || (rc = X.out.reserve(nout)) ||
        (rc = X.queue.reserve(16)) || (rc = X.dbl.reserve(std::max<size_t>(nw, 1))))
        return rc;
    hipError_t e = hipMemcpyAsync(X.params.ptr, &B, sizeof(B), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && !tab.empty())
        e = hipMemcpyAsync(X.params.ptr + tab_off, tab.data(), tab.size() * 2, hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && !rows.empty())
        e = hipMemcpyAsync(X.params.ptr + rows_off, rows.data(), rows.size() * sizeof(Row), hipMemcpyHostToDevice,
                           v.stream);
    if (e == hipSuccess) e = hipMemsetAsync(X.out.ptr, 0, nout * sizeof(unsigned long long), v.stream);
    if (e == hipSuccess && nw) e = hipMemsetAsync(X.dbl.ptr, 0, nw * sizeof(double), v.stream);
    if (e == hipSuccess) e = hipMemsetAsync(X.queue.ptr, 0, 16 * sizeof(int32_t), v.stream);
    if (e != hipSuccess) return hip_fail(e, "4PCF parameter upload");

    QuadArgs a{};
    a.pw = v.pw;
    a.flags = v.flags;
    a.orig = v.orig;
    a.cell_start = v.cell_start;
    a.rows = (const Row*)(X.params.ptr + rows_off);
    a.bins = (const BinDesc*)X.params.ptr;
    a.cfg_tab = tab.empty() ? nullptr : (const uint16_t*)(X.params.ptr + tab_off);
    a.queue = X.queue.ptr;
    a.err = X.queue.ptr + 8;
    a.out_cnt = X.out.ptr;
    a.out_w = X.dbl.ptr;
    a.out_stats = X.out.ptr + ncnt;
    a.g = v.g;
    a.g.m = mst;
    a.nrows = (int)rows.size();
    a.n_items = (int)(v.n_data + v.n_rand);
    a.ncell = ncell;
    a.kmax = kmax;
    a.lds_hist = lds_hist ? 1 : 0;
    a.tetra = tetra ? 1 : 0;
    // Two passes, as for the 3PCF: the first pass's neighbour list (and its K x K link matrix)
    // is sized for the 99.5th percentile of ~8192 sampled neighbour counts (+5%, a power of two
    // for the bitonic sort, at least 64 for the sort's staging area); apexes with more
    // neighbours are queued for a second pass with the full list.
    const int n = a.n_items;
    int kmax1 = kmax;
    const int nsample = std::min(n, 8192);
    const int stride = nsample > 0 ? n / nsample : 1;
    // i32 as for the 3PCF: [0] second-pass count, [1, n] its apexes, [n+1] last-pass count, [n+2]
    // its largest neighbour count, [n+3, 2n+2] its apexes, then the sampled neighbour counts
    if ((rc = X.i32.reserve(2 * (size_t)n + 3 + nsample))) return rc;
    int32_t* d_list2 = X.i32.ptr + 1;
    int32_t* d_cnt3 = X.i32.ptr + n + 1;
    int32_t* d_list3 = X.i32.ptr + n + 3;
    if (nsample > 0 && kmax > 64) {
        TripArgs ts{};
        ts.pw = a.pw;
        ts.flags = a.flags;
        ts.cell_start = a.cell_start;
        ts.rows = a.rows;
        ts.bins = a.bins;
        ts.g = a.g;
        ts.nrows = a.nrows;
        int32_t* d_cnt = X.i32.ptr + 2 * n + 3;
        std::vector<int32_t> kc(nsample);
        e = launch_nbr_count(ts, stride, nsample, 0, d_cnt, v.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(kc.data(), d_cnt, nsample * sizeof(int32_t), hipMemcpyDeviceToHost, v.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
        if (e != hipSuccess) return hip_fail(e, "4PCF neighbour sample");
        std::vector<int32_t> ks;
        for (int32_t k : kc)
            if (k >= 0) ks.push_back(k);
        if (!ks.empty()) {
            const size_t q = std::min(ks.size() - 1, (size_t)std::ceil(0.995 * (ks.size() - 1)));
            std::nth_element(ks.begin(), ks.begin() + q, ks.end());
            const double want = 1.05 * ks[q];
            kmax1 = 64;
            while (kmax1 < kmax && kmax1 < want) kmax1 *= 2;
        }
    }
    {
        const char* ev = getenv("GS_T4_KMAX");      // tuning override (unset or empty: automatic)
        if (ev && *ev) {
            kmax1 = 64;
            while (kmax1 < kmax && kmax1 < atoi(ev)) kmax1 *= 2;
        }
    }
    const bool two_pass = kmax1 < kmax;
    (void)hipEventRecord(v.ev[2], v.stream);
    if (n > 0) {
        e = hipMemsetAsync(X.i32.ptr, 0, sizeof(int32_t), v.stream);
        if (e == hipSuccess) e = hipMemsetAsync(d_cnt3, 0, 2 * sizeof(int32_t), v.stream);
        if (e != hipSuccess) return hip_fail(e, "4PCF overflow counters");
        QuadArgs a1 = a;
        a1.ovf_kmax = d_cnt3 + 1;
        if (two_pass) {
            a1.kmax = kmax1;
            a1.ovf_n = X.i32.ptr;
            a1.ovf = d_list2;
        } else {
            a1.ovf_n = d_cnt3;
            a1.ovf = d_list3;
        }
        const int occ = quad4_occupancy(a1, nb, weighted);
        const int nblk1 = std::max(1, std::min(v.n_cu * occ, n));
        a1.batch = triplet_batch(n, nblk1);
        e = launch_quad4(a1, nb, weighted, nblk1, v.stream);
        if (e != hipSuccess) return hip_fail(e, "4PCF kernel launch");
        if (two_pass) {
            int32_t nov = 0;
            e = hipMemcpyAsync(&nov, X.i32.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, v.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
            if (e != hipSuccess) return hip_fail(e, "4PCF first pass");
            if (nov > 0) {
                QuadArgs a2 = a;
                a2.items = d_list2;
                a2.n_items = nov;
                a2.batch = 1;
                a2.ovf = d_list3;
                a2.ovf_n = d_cnt3;
                a2.ovf_kmax = d_cnt3 + 1;
                e = hipMemsetAsync(X.queue.ptr, 0, 8 * sizeof(int32_t), v.stream);   // work counters only
                const int occ2 = quad4_occupancy(a2, nb, weighted);
                if (e == hipSuccess)
                    e = launch_quad4(a2, nb, weighted, std::max(1, std::min(v.n_cu * occ2, nov)), v.stream);
                if (e != hipSuccess) return hip_fail(e, "4PCF second pass");
            }
        }
        // last pass: the apexes whose neighbours outnumber the second pass's list, with their lists
        // in device memory (one per work-group; no link matrix)
        int32_t h3[2] = {0, 0};
        e = hipMemcpyAsync(h3, d_cnt3, sizeof(h3), hipMemcpyDeviceToHost, v.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
        if (e != hipSuccess) return hip_fail(e, "4PCF queued apex count");
        if (h3[0] > 0) {
            QuadArgs a3 = a;
            a3.items = d_list3;
            a3.n_items = h3[0];
            a3.batch = 1;
            a3.kmax = 64;
            while (a3.kmax < h3[1]) a3.kmax *= 2;           // a power of two for the bitonic sort
            const size_t gstride = (a3.kmax * kQuadGlistEntryBytes + 255) / 256 * 256;
            a3.glist = (void*)d_list3;      // selects the device-memory variant for the occupancy query
            const int occ3 = quad4_occupancy(a3, nb, weighted);
            const size_t gbudget = (size_t)1 << 30;
            const int nblk3 = (int)std::max<size_t>(1, std::min<size_t>({(size_t)h3[0], (size_t)v.n_cu * occ3,
                                                                      gbudget / gstride}));
            if ((rc = X.big.reserve((size_t)nblk3 * gstride))) return rc;
            a3.glist = X.big.ptr;
            a3.glist_stride = (int64_t)gstride;
            e = hipMemsetAsync(X.queue.ptr, 0, 8 * sizeof(int32_t), v.stream);
            if (e == hipSuccess) e = launch_quad4(a3, nb, weighted, nblk3, v.stream);
            if (e != hipSuccess) return hip_fail(e, "4PCF device-memory pass");
        }
    }
    (void)hipEventRecord(v.ev[3], v.stream);
    std::vector<unsigned long long> h_out(nout);
    std::vector<double> h_w(nw);
    e = hipMemcpyAsync(h_out.data(), X.out.ptr, nout * sizeof(unsigned long long), hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess && nw)
        e = hipMemcpyAsync(h_w.data(), X.dbl.ptr, nw * sizeof(double), hipMemcpyDeviceToHost, v.stream);
    int32_t err = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&err, X.queue.ptr + 8, sizeof(int32_t), hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) return hip_fail(e, "4PCF kernel");
    if (err) return fail(GS_EHIP, "4PCF: an apex's neighbour list overflowed with no later pass (internal error)");

    if (cnt) memcpy(cnt, h_out.data(), ncnt * sizeof(uint64_t));
    // normalised weight of one point: +w/Wd (data) or -w/Wr (random) (gramsci.F90:1390-1391)
    const long double Wd = v.wsum_d, Wr = v.wsum_r;
    auto norm = [&](int r) -> long double { return r ? -1.0L / Wr : 1.0L / Wd; };
    for (int q = 0; q < ncell; ++q) {
        long double v2 = 0, v3 = 0;
        for (int pat = 0; pat < 16; ++pat) {
            const long double s = weighted ? (long double)h_w[(size_t)q * 16 + pat] : (long double)h_out[(size_t)q * 16 + pat];
            if (s == 0) continue;
            const long double f = norm(pat & 1) * norm((pat >> 1) & 1) * norm((pat >> 2) & 1) * norm((pat >> 3) & 1);
            v2 += s * f;
            if (pat == 15) v3 += s * f;
        }
        if (n2) n2[q] = (double)v2;
        if (n3) n3[q] = (double)v3;
    }
    // disc terms: t -> (points a, b of the pair product; point c that must be random)
    //   t0: (i,k1 | k3)  t1: (i,k2 | k3)  t2: (i,k3 | k1)  t3: (k1,k2 | i)  t4: (k3,k2 | i)  t5: (k3,k1 | i)
    // the tetrahedron's terms are triple products that also carry c's normalised
    // weight, -w_c/Wr (gramsci.F90:590-631)
    static const int ta[6] = {0, 0, 0, 1, 3, 3}, tb[6] = {1, 2, 3, 2, 2, 1}, tc[6] = {3, 3, 1, 0, 0, 0};
    const long double fc = tetra ? norm(1) : 1.0L;
    for (int t = 0; t < 6; ++t)
        for (int q = 0; q < ncell; ++q) {
            long double dn = 0, dr = 0;
            for (int ab = 0; ab < 4; ++ab) {
                const int ra = ab & 1, rb = (ab >> 1) & 1;
                long double s;
                if (weighted) {
                    s = h_w[(size_t)ncell * 16 + ((size_t)t * ncell + q) * 4 + ab];
                } else {
                    s = 0;
                    for (int pat = 0; pat < 16; ++pat)
                        if (((pat >> tc[t]) & 1) && ((pat >> ta[t]) & 1) == ra && ((pat >> tb[t]) & 1) == rb)
                            s += (long double)h_out[(size_t)q * 16 + pat];
                }
                if (s == 0) continue;
                const long double term = s * norm(ra) * norm(rb) * fc;
                dn += term;
                if (ra && rb) dr += term;
            }
            if (disc_n) disc_n[(size_t)t * ncell + q] = (double)dn;
            if (disc_r) disc_r[(size_t)t * ncell + q] = (double)dr;
        }
    std::vector<unsigned long long> sts(5, 0);
    for (int k = 0; k < 5; ++k) sts[k] = h_out[ncnt + k];
    fill_stats(st, v.ev, sts.data(), mst);
    return GS_OK;
}

}  // namespace gs

extern "C" int gs_count_4pcf(gs_ctx* c, const gs_params* p, uint64_t* cnt, double* n2, double* n3, double* disc_n,
                             double* disc_r, gs_stats* st)
{
    return count4(c, p, false, cnt, n2, n3, disc_n, disc_r, st);
}

extern "C" int gs_count_4pcf_tetra(gs_ctx* c, const gs_params* p, uint64_t* cnt, double* n2, double* n3,
                                   double* disc_n, double* disc_r, gs_stats* st)
{
    return count4(c, p, true, cnt, n2, n3, disc_n, disc_r, st);
}

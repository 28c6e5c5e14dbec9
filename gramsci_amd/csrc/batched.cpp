// batched.cpp -- many independent regions in one persistent launch.
//
// The reference's scaling route is offline domain decomposition: the volume
// is cut into regions with an rmax-wide buffer shell (domain_decomposition.f90
// :49-74), each region is run separately in -cut mode with the shell points
// flagged as never-apex (gramsci.F90:1224-1231, :363), and the per-region
// outputs are summed afterwards.  Here all regions are counted by ONE launch
// of the persistent pair kernel: every region keeps its own grid, points and
// output histogram, the work queue spans the x-row chunks of all regions (so
// uneven regions balance automatically), and the whole pipeline - grid build
// (counting sort + scans) and the pair kernel - is captured once in a
// hipGraph and replayed.  Buffers are cleared by fill kernels (no memset
// nodes in the graph).  The cells-per-rmax factor is chosen by timing the
// captured graph for each candidate on the actual batch (cell_div = 0), and
// the choice is cached per parameter set.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gs_api_internal.h"
#include "gs_kernels.h"

using namespace gs;

namespace gs {

void batch_release(GsBatch& b)
{
    for (auto& pl : b.plans) {
        if (pl.exec) (void)hipGraphExecDestroy(pl.exec);
        if (pl.graph) (void)hipGraphDestroy(pl.graph);
        pl.blob.release();
    }
    b.plans.clear();
    b.d_off.release(); b.d_row_base.release(); b.cell_of.release(); b.count.release(); b.cell_start.release();
    b.nch.release(); b.row_chunk.release(); b.scan_tmp.release(); b.queue.release(); b.orig.release();
    b.d_regions.release(); b.params.release(); b.pw.release(); b.ws.release(); b.pl.release(); b.flags.release();
    b.out.release();
    b.xyz.release(); b.w.release(); b.buf.release(); b.ovf.release(); b.big.release();
    b.valid = false;
}

static std::string plan_key(const gs_params& p, int kind)
{
    char k[160];
    snprintf(k, sizeof(k), "%d:%.17g/%.17g/%d/%d/%d/%d", kind, p.rmin, p.rmax, p.nbins, p.nmu, p.log_bins,
             p.precision);
    return k;
}

static bool same_params(const gs_params& a, const gs_params& b)
{
    return a.rmin == b.rmin && a.rmax == b.rmax && a.nbins == b.nbins && a.nmu == b.nmu &&
           a.log_bins == b.log_bins && a.precision == b.precision;
}

// Builds (or finds) the captured pipeline for parameters p with launch parameters tu (tu.m cells per
// rmax), ending in the pair kernel (kind 2) or the triplet kernel (kind 3).
static int get_plan(gs_ctx* c, const gs_params& p, const BatchTune& tu, int kind, GsBatch::Plan** out)
{
    CtxView v = view(c);
    GsBatch& B = *v.batch;
    const int m = tu.m;
    const int sx_want = tu.sx > 0 ? tu.sx : (kind == 2 ? pair_cell_sx() : apex_cell_sx());
    for (auto& pl : B.plans)
        if (pl.tune == tu && pl.sx == sx_want && pl.kind == kind && same_params(pl.p, p)) {
            *out = &pl;
            return GS_OK;
        }
    const int nreg = B.nreg;
    const int64_t n = B.off[nreg];   // the batch's own copy of its catalogue (input order)
    const int nb = p.nbins, nm = p.nmu;
    const bool rsd = nm >= 2, fp64 = p.precision == GS_FP64, weighted = B.has_w;
    const int nbt = 3 * nb * nm;

    // cell edges (x cut finer for the pair kernel) and per-region grids
    const double h = round11(p.rmax / m);
    const int sx = sx_want;
    const double hx = h / sx;
    std::vector<RegionDesc> reg(nreg);
    std::vector<int32_t> row_base(nreg + 1, 0);
    int64_t cells = 0, rows_tot = 0;
    for (int r = 0; r < nreg; ++r) {
        RegionDesc& d = reg[r];
        memset(&d, 0, sizeof(d));
        int64_t nc = 1;
        for (int k = 0; k < 3; ++k) {
            d.lo[k] = B.bbox[6 * r + k];
            d.dim[k] = (int32_t)std::floor((B.bbox[6 * r + 3 + k] - B.bbox[6 * r + k]) / (k == 0 ? hx : h)) + 1;
            nc *= d.dim[k];
        }
        d.cell_base = (int32_t)cells;
        d.row_base = (int32_t)rows_tot;
        d.pt_base = (int32_t)B.off[r];
        d.n_data = (int32_t)B.ndata[r];
        // the scale a launch over this region alone would use, so that sums match it bit for bit
        d.fx_shift = kind == 3 ? fx_shift_triples(B.wmr[r]) : fx_shift_point(B.wmr_d[r]);
        d.fx_shift_rand = kind == 3 ? d.fx_shift : fx_shift_point(B.wmr_r[r]);
        row_base[r] = (int32_t)rows_tot;
        cells += nc;
        rows_tot += (int64_t)d.dim[1] * d.dim[2];
        if (cells > (1 << 26)) return fail(GS_ELIMIT, "batched grid too large (%lld cells)", (long long)cells);
    }
    row_base[nreg] = (int32_t)rows_tot;
    const int mst = (int)std::ceil(p.rmax / h - 1e-12);
    std::vector<Row> rows = kind == 3 ? full_stencil(h, hx, p.rmax, mst) : half_stencil(h, hx, p.rmax, mst);
    if ((int)rows.size() > kMaxRows) return fail(GS_ELIMIT, "stencil too large");
    BinDesc bins;
    int rc;
    if ((rc = fill_bins(p, kind == 3 ? 1.5 * p.rmax : (mst + 1) * h, B.xmax, B.wmax_d, B.wmax_r, bins))) return rc;

    // triplet kernel: configuration and find_normal tables, LDS plan
    std::vector<uint16_t> tab3;
    std::vector<uint8_t> ntab;
    TripArgs ta{};
    int nh = 0;
    if (kind == 3) {
        tab3 = table3_hist(nb);
        if (tab3.empty()) return fail(GS_ELIMIT, "3PCF with %d bins: configuration table too large", nb);
        if (nm > 1) ntab = normal_table(nm);
        ta.ncfg = nb * (nb + 1) * (nb + 2) / 6;
        ta.nm = nm;
        ta.log_bins = p.log_bins != 0;
        ta.nrows = (int)rows.size();     // the LDS plan holds the stencil rows' table
        if ((rc = plan_triplet_lds(ta, nb, fp64, weighted))) return rc;
        nh = ta.ncfg * nm * 4;
    }
    const size_t nout = kind == 3 ? (size_t)nreg * 3 * nh + 8 : (size_t)nreg * 3 * nbt + 4;
    const int64_t scan_n = std::max<int64_t>(cells + 1, rows_tot + 1);
    // captured graphs hold raw pointers to the shared work buffers: if any of
    // them has to grow, every existing plan is destroyed before reallocating
    const bool grow = (size_t)n > B.cell_of.cap || (size_t)(cells + 1) > B.count.cap ||
                      (size_t)(cells + 1) > B.cell_start.cap || (size_t)(rows_tot + 1) > B.nch.cap ||
                      (size_t)(rows_tot + 1) > B.row_chunk.cap || (size_t)scan_scratch_size(scan_n) > B.scan_tmp.cap ||
                      nout > B.out.cap || (kind == 3 && (size_t)(2 * n + 3) > B.ovf.cap) || B.queue.cap < 32;
    if (grow) {
        for (auto& pl : B.plans) {
            if (pl.exec) (void)hipGraphExecDestroy(pl.exec);
            if (pl.graph) (void)hipGraphDestroy(pl.graph);
            pl.blob.release();
        }
        B.plans.clear();
    }
    if ((rc = B.cell_of.reserve(n)) || (rc = B.count.reserve(cells + 1)) || (rc = B.cell_start.reserve(cells + 1)) ||
        (rc = B.nch.reserve(rows_tot + 1)) || (rc = B.row_chunk.reserve(rows_tot + 1)) ||
        (rc = B.scan_tmp.reserve(scan_scratch_size(scan_n))) || (rc = B.queue.reserve(32)) ||
        (kind == 3 && (rc = B.ovf.reserve(2 * n + 3))) ||
        (rc = B.orig.reserve(n)) || (rc = B.pw.reserve(n)) || (rc = B.ws.reserve(n)) || (rc = B.pl.reserve(n)) ||
        (rc = B.flags.reserve(n)) ||
        (rc = B.out.reserve(nout)))
        return rc;
    // per-plan parameter blob: regions, bins, stencil rows (kept for the plan's lifetime)
    const size_t reg_bytes = ((nreg * sizeof(RegionDesc) + 15) / 16) * 16;
    const size_t bin_off = reg_bytes, rows_off = bin_off + ((sizeof(BinDesc) + 15) / 16) * 16;
    const size_t rb_off = rows_off + ((rows.size() * sizeof(Row) + 15) / 16) * 16;
    const size_t tab_off = rb_off + (((nreg + 1) * sizeof(int32_t) + 15) / 16) * 16;
    const size_t ntab_off = tab_off + ((tab3.size() * 2 + 15) / 16) * 16;
    const size_t blob = ntab_off + ntab.size() + 16;
    // each plan owns its blob so that graphs of several parameter sets can coexist
    B.plans.emplace_back();
    GsBatch::Plan& P = B.plans.back();
    DevBuf<unsigned char>* pblob = &P.blob;
    if ((rc = pblob->reserve(blob))) { B.plans.pop_back(); return rc; }
    hipError_t e = hipMemcpyAsync(pblob->ptr, reg.data(), nreg * sizeof(RegionDesc), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pblob->ptr + bin_off, &bins, sizeof(bins), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && !rows.empty())
        e = hipMemcpyAsync(pblob->ptr + rows_off, rows.data(), rows.size() * sizeof(Row), hipMemcpyHostToDevice,
                           v.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pblob->ptr + rb_off, row_base.data(), (nreg + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                           v.stream);
    if (e == hipSuccess && !tab3.empty())
        e = hipMemcpyAsync(pblob->ptr + tab_off, tab3.data(), tab3.size() * 2, hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && !ntab.empty())
        e = hipMemcpyAsync(pblob->ptr + ntab_off, ntab.data(), ntab.size(), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) { pblob->release(); B.plans.pop_back(); return hip_fail(e, "batch upload"); }
    const RegionDesc* d_reg = (const RegionDesc*)pblob->ptr;

    PairArgs a{};
    a.pw = B.pw.ptr;
    a.ws = B.ws.ptr;
    a.pl = B.pl.ptr;
    a.flags = B.flags.ptr;
    a.cell_start = B.cell_start.ptr;
    a.row_chunk = B.row_chunk.ptr;
    a.rows = (const Row*)(pblob->ptr + rows_off);
    a.bins = (const BinDesc*)(pblob->ptr + bin_off);
    a.queue = B.queue.ptr;
    a.out_cnt = B.out.ptr;
    a.out_w = B.out.ptr + (size_t)nreg * nbt;
    a.out_stats = B.out.ptr + (size_t)nreg * 3 * nbt;
    a.g.h = h;
    a.g.h32 = (float)h;
    a.g.hx = hx;
    a.g.hx32 = (float)hx;
    a.g.m = mst;
    a.nrows = (int)rows.size();
    a.reps = 1;
    a.mirror = (rsd && !B.has_buf) ? 1 : 0;
    a.regions = d_reg;
    a.row_base = (const int32_t*)(pblob->ptr + rb_off);
    a.n_chunks_dev = B.row_chunk.ptr + rows_tot;
    a.nreg = nreg;
    if (kind == 2 && pair2_lds_bytes(nb, rsd ? nm : 1, 1, fp64, weighted, true) > 64 * 1024) {
        pblob->release(); B.plans.pop_back();
        return fail(GS_ELIMIT, "per-wave histograms of %d bins do not fit in LDS", nbt);
    }
    if (kind == 3) {
        ta.pw = B.pw.ptr;
        ta.flags = B.flags.ptr;
        ta.orig = B.orig.ptr;
        ta.cell_start = B.cell_start.ptr;
        ta.rows = (const Row*)(pblob->ptr + rows_off);
        ta.bins = (const BinDesc*)(pblob->ptr + bin_off);
        ta.cfg_tab = (const uint16_t*)(pblob->ptr + tab_off);
        ta.ntab = ntab.empty() ? nullptr : (const uint8_t*)(pblob->ptr + ntab_off);
        ta.queue = B.queue.ptr;
        ta.err = B.queue.ptr + 8;
        ta.out_cnt = B.out.ptr;
        ta.out_w = B.out.ptr + (size_t)nreg * nh;
        ta.out_stats = B.out.ptr + (size_t)nreg * 3 * nh;
        ta.g.h = h;
        ta.g.h32 = (float)h;
        ta.g.hx = hx;
        ta.g.hx32 = (float)hx;
        ta.g.m = mst;
        ta.nrows = (int)rows.size();
        ta.n_items = (int)n;
        ta.equi = 0;
        ta.nreg = nreg;
        ta.regions = d_reg;
    }
    // 3PCF in two captured passes: the first with a neighbour list sized for twice the largest
    // mean neighbour count of any region's box (apexes with more neighbours are queued), the
    // second over the queued apexes with the largest list that fits (count read on the device).
    // Apexes with more neighbours than that are queued again (B.ovf[n+1] count, [n+2] largest
    // neighbour count, apexes from n+3) for a pass with device-memory lists after the replay.
    TripArgs ta1 = ta, ta2 = ta;
    int nblocks2 = 0;
    if (kind == 3) {
        ta1.ovf_n = B.ovf.ptr + n + 1;
        ta1.ovf = B.ovf.ptr + n + 3;
        ta1.ovf_kmax = B.ovf.ptr + n + 2;
        ta2.ovf_n = ta1.ovf_n;
        ta2.ovf = ta1.ovf;
        ta2.ovf_kmax = ta1.ovf_kmax;
        const double shell = 4.0 / 3.0 * M_PI * (std::pow(p.rmax, 3) - std::pow(std::max(p.rmin, 0.0), 3));
        double kmean = 0.0;
        for (int r = 0; r < nreg; ++r) {
            double vol = 1.0;
            for (int k = 0; k < 3; ++k) vol *= std::max(B.bbox[6 * r + 3 + k] - B.bbox[6 * r + k], p.rmax);
            kmean = std::max(kmean, (double)(B.off[r + 1] - B.off[r]) * shell / vol);
        }
        ta1.wg = tu.wg > 0 ? tu.wg : triplet_wg(kmean);
        const char* evl = getenv("GS_T3_KMAX_LDS");     // test override: a shorter second-pass list
        if (evl && *evl) ta.kmax = ta1.kmax = ta2.kmax = std::max(64, std::min(ta.kmax, atoi(evl)));
        int kmax1 = std::min(ta.kmax, std::max(128, (int)std::ceil(2.0 * kmean / 64.0) * 64));
        const char* ev = getenv("GS_T3_KMAX");      // tuning override (unset or empty: automatic)
        if (tu.kmax1 > 0) kmax1 = std::max(64, std::min(ta.kmax, tu.kmax1));
        else if (ev && *ev) kmax1 = std::max(64, std::min(ta.kmax, atoi(ev)));
        if (kmax1 < ta.kmax) {
            ta1.kmax = kmax1;
            ta1.ovf_n = B.ovf.ptr;
            ta1.ovf = B.ovf.ptr + 1;
            ta2.items = B.ovf.ptr + 1;
            ta2.n_items_dev = B.ovf.ptr;
            ta2.queue = B.queue.ptr + 16;
            ta2.wg = kWG;           // the apexes with the most neighbours, one per fetch
            ta2.batch = 1;
            nblocks2 = v.n_cu * triplet3_occupancy(ta2, nb, fp64, weighted);
        }
    }
    const int occ = kind == 3 ? triplet3_occupancy(ta1, nb, fp64, weighted)
                              : std::max(1, pair2_occupancy(nb, nm, 1, fp64, rsd, weighted, p.log_bins != 0, true));
    const int nblocks = kind == 3 ? std::max(1, (int)std::min<int64_t>((int64_t)v.n_cu * occ, std::max<int64_t>(n, 1)))
                                  : v.n_cu * occ;
    if (kind == 3) ta1.batch = triplet_batch(n, nblocks);

    // capture: clear, count sort into cells, scans, scatter, row chunks, pair or triplet kernel
    e = hipStreamBeginCapture(v.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) e = launch_fill_u32(B.count.ptr, cells + 1, 0, v.stream);
    if (e == hipSuccess) e = launch_fill_u32(B.queue.ptr, 32, 0, v.stream);
    if (e == hipSuccess && nblocks2 > 0) e = launch_fill_u32(B.ovf.ptr, 1, 0, v.stream);
    if (e == hipSuccess && kind == 3) e = launch_fill_u32(B.ovf.ptr + n + 1, 2, 0, v.stream);
    if (e == hipSuccess) e = launch_fill_u32(B.out.ptr, (int64_t)nout * 2, 0, v.stream);
    if (e == hipSuccess)
        e = launch_cell_count_b(B.xyz.ptr, n, B.d_off.ptr, nreg, d_reg, h, hx, B.cell_of.ptr, B.count.ptr, v.stream);
    if (e == hipSuccess) e = exclusive_scan(B.count.ptr, B.cell_start.ptr, cells + 1, B.scan_tmp.ptr, v.stream);
    if (e == hipSuccess) e = launch_fill_u32(B.count.ptr, cells + 1, 0, v.stream);
    if (e == hipSuccess)
        e = launch_scatter_b(B.xyz.ptr, B.has_w ? B.w.ptr : nullptr, B.has_buf ? B.buf.ptr : nullptr, n, B.d_off.ptr,
                             nreg, d_reg, h, hx, B.cell_of.ptr, B.cell_start.ptr,
                             B.count.ptr, B.pw.ptr, B.ws.ptr, B.pl.ptr, B.flags.ptr, B.orig.ptr, v.stream);
    if (kind == 2) {
        if (e == hipSuccess)
            e = launch_row_chunks_b(B.cell_start.ptr, d_reg, a.row_base, nreg, (int)rows_tot, B.nch.ptr, v.stream);
        if (e == hipSuccess) e = exclusive_scan(B.nch.ptr, B.row_chunk.ptr, rows_tot + 1, B.scan_tmp.ptr, v.stream);
        if (e == hipSuccess) e = launch_pair2(a, nb, nm, fp64, rsd, weighted, p.log_bins != 0, nblocks, kWG, v.stream);
    } else if (e == hipSuccess && n > 0) {
        e = launch_triplet3(ta1, nb, fp64, weighted, nblocks, v.stream);
        if (e == hipSuccess && nblocks2 > 0) e = launch_triplet3(ta2, nb, fp64, weighted, nblocks2, v.stream);
    }
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(v.stream, &graph);
    if (e == hipSuccess) e = e2;
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        pblob->release(); B.plans.pop_back();
        return hip_fail(e, "batched graph capture");
    }
    P.p = p;
    P.tune = tu;
    P.m = m;
    P.sx = sx;
    P.graph = graph;
    P.exec = exec;
    P.nbt = nbt;
    P.fx_shift = kind == 3 ? bins.fx_shift3 : bins.fx_sd;
    P.fx_shift_r.resize(nreg);
    P.fx_sd_r.resize(nreg);
    P.fx_sr_r.resize(nreg);
    for (int r = 0; r < nreg; ++r) {
        P.fx_shift_r[r] = reg[r].fx_shift;
        P.fx_sd_r[r] = reg[r].fx_shift;
        P.fx_sr_r[r] = reg[r].fx_shift_rand;
    }
    P.mst = mst;
    P.kind = kind;
    P.nh = nh;
    if (kind == 3) {
        P.kmax = ta.kmax;
        P.wg1 = ta1.wg;
        P.kmax1 = ta1.kmax;
        P.t3 = ta;              // tables, outputs and regions of the device-memory pass
        P.t3_list = B.ovf.ptr + n + 3;
        P.t3_cnt = B.ovf.ptr + n + 1;
        P.nb = nb;
        P.fp64 = fp64;
        P.weighted = weighted;
    }
    *out = &P;
    return GS_OK;
}

static int run_plan(gs_ctx* c, GsBatch::Plan& P, float* ms)
{
    CtxView v = view(c);
    GsBatch& B = *v.batch;
    (void)hipEventRecord(v.ev[2], v.stream);
    hipError_t e = hipGraphLaunch(P.exec, v.stream);
    if (e != hipSuccess) return hip_fail(e, "batched graph launch");
    if (P.kind == 3) {
        // apexes with more neighbours than any LDS list: lists in device memory (not captured)
        int rc = triplet_glist_pass(P.t3, P.nb, P.fp64, P.weighted, P.t3_list, P.t3_cnt, B.queue.ptr + 24, B.big,
                                    v.n_cu, v.stream, nullptr);
        if (rc) return rc;
    }
    (void)hipEventRecord(v.ev[3], v.stream);
    e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) return hip_fail(e, "batched graph launch");
    if (ms) (void)hipEventElapsedTime(ms, v.ev[2], v.ev[3]);
    return GS_OK;
}

}  // namespace gs

extern "C" int gs_set_catalogue_batched(gs_ctx* c, const double* xyz, const double* w, const int32_t* buf,
                                        const int64_t* off, const int64_t* ndata, int nreg)
{
    if (!c || !off || !ndata || nreg < 1 || nreg > 65536) return fail(GS_EINVAL, "bad batched catalogue arguments");
    const int64_t n = off[nreg];
    if (off[0] != 0 || n < 0 || n > (int64_t)INT32_MAX - 1024 || (!xyz && n > 0))
        return fail(GS_EINVAL, "region offsets must start at 0 and be non-decreasing");
    for (int r = 0; r < nreg; ++r)
        if (off[r + 1] < off[r] || ndata[r] < 0 || ndata[r] > off[r + 1] - off[r])
            return fail(GS_EINVAL, "bad region %d (offsets or data count)", r);
    int rc = ctx_set_device(c);
    if (rc) return rc;
    CtxView v = view(c);
    GsBatch& B = *v.batch;
    batch_release(B);
    B.nreg = nreg;
    B.off.assign(off, off + nreg + 1);
    B.ndata.assign(ndata, ndata + nreg);
    B.bbox.assign((size_t)6 * nreg, 0.0);
    B.wd.assign(nreg, 0.0L);
    B.wr.assign(nreg, 0.0L);
    B.wmr.assign(nreg, 1.0);
    B.wmr_d.assign(nreg, 1.0);
    B.wmr_r.assign(nreg, 1.0);
    B.wmax = B.wmax_d = B.wmax_r = 1.0;
    B.xmax = 0.0;
    double wm = 0.0, wmd = 0.0, wmrr = 0.0;
    for (int r = 0; r < nreg; ++r) {
        double* bb = &B.bbox[6 * r];
        for (int k = 0; k < 3; ++k) { bb[k] = 1e308; bb[3 + k] = -1e308; }
        if (off[r + 1] == off[r])
            for (int k = 0; k < 6; ++k) bb[k] = 0.0;
        double wmr = 0.0, wmr_c[2] = {0.0, 0.0};
        for (int64_t i = off[r]; i < off[r + 1]; ++i) {
            for (int k = 0; k < 3; ++k) {
                const double x = xyz[3 * i + k];
                if (!std::isfinite(x)) return fail(GS_EINVAL, "non-finite coordinate at point %lld", (long long)i);
                bb[k] = std::min(bb[k], x);
                bb[3 + k] = std::max(bb[3 + k], x);
                B.xmax = std::max(B.xmax, std::fabs(x));
            }
            const double wi = w ? w[i] : 1.0;
            if (!(wi >= 0.0) || !std::isfinite(wi)) return fail(GS_EINVAL, "weights must be finite and >= 0");
            wm = std::max(wm, wi);
            wmr = std::max(wmr, wi);
            double& mc = wmr_c[i - off[r] < ndata[r] ? 0 : 1];
            mc = std::max(mc, wi);
        }
        wmd = std::max(wmd, wmr_c[0]);
        wmrr = std::max(wmrr, wmr_c[1]);
        // the region's sums exactly as a single-catalogue upload of the region computes them
        if (w) {
            B.wmr[r] = wmr > 0 ? wmr : 1.0;
            B.wmr_d[r] = wmr_c[0] > 0 ? wmr_c[0] : 1.0;
            B.wmr_r[r] = wmr_c[1] > 0 ? wmr_c[1] : 1.0;
            const int k = wsum_scale(wmr);
            B.wd[r] = fixed_sum(w + off[r], ndata[r], k);
            B.wr[r] = fixed_sum(w + off[r] + ndata[r], off[r + 1] - off[r] - ndata[r], k);
        } else {
            B.wd[r] = (long double)ndata[r];
            B.wr[r] = (long double)(off[r + 1] - off[r] - ndata[r]);
        }
    }
    if (w) {
        B.wmax = wm > 0 ? wm : 1.0;
        B.wmax_d = wmd > 0 ? wmd : 1.0;
        B.wmax_r = wmrr > 0 ? wmrr : 1.0;
    }
    // the batch keeps its own copy of the catalogue, so that later single-catalogue
    // uploads on this context cannot change what the captured pipelines read
    B.has_w = w != nullptr;
    B.has_buf = buf != nullptr;
    if ((rc = B.xyz.reserve(3 * std::max<int64_t>(n, 1))) || (B.has_w && (rc = B.w.reserve(std::max<int64_t>(n, 1)))) ||
        (B.has_buf && (rc = B.buf.reserve(std::max<int64_t>(n, 1)))))
        return rc;
    hipError_t eu = n > 0 ? hipMemcpy(B.xyz.ptr, xyz, (size_t)n * 24, hipMemcpyHostToDevice) : hipSuccess;
    if (eu == hipSuccess && B.has_w && n > 0) eu = hipMemcpy(B.w.ptr, w, (size_t)n * 8, hipMemcpyHostToDevice);
    if (eu == hipSuccess && B.has_buf && n > 0) eu = hipMemcpy(B.buf.ptr, buf, (size_t)n * 4, hipMemcpyHostToDevice);
    if (eu != hipSuccess) return hip_fail(eu, "batched catalogue upload");
    std::vector<int32_t> o32(nreg + 1);
    for (int r = 0; r <= nreg; ++r) o32[r] = (int32_t)off[r];
    if ((rc = B.d_off.reserve(nreg + 1))) return rc;
    hipError_t e = hipMemcpy(B.d_off.ptr, o32.data(), (nreg + 1) * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "region offsets upload");
    B.valid = true;
    return GS_OK;
}

// Times a plan on this batch: one replay to warm up, one timed.
static int time_plan(gs_ctx* c, const gs_params& p, const BatchTune& tu, int kind, float* ms, GsBatch::Plan** out)
{
    GsBatch::Plan* P;
    int rc;
    if ((rc = get_plan(c, p, tu, kind, &P))) return rc;
    if ((rc = run_plan(c, *P, nullptr)) || (rc = run_plan(c, *P, ms))) return rc;
    *out = P;
    return GS_OK;
}

// Launch parameters for parameter set p on this batch: the cached winner, or else the autotuner's.  The
// tuner times the captured pipeline by coordinate descent: cells per rmax (pair 2..6, triplet 1..3),
// then the x cells per y/z cell (pair 1, 2, 4, 8, 16; triplet 1, 2, 4, 8), and for the triplet kernel
// the threads per work-group (64, 128, 256) and then the first-pass list length (half and twice the
// automatic one).  Plans of the losing candidates are released; the winner's is kept.
static int batch_tune(gs_ctx* c, const gs_params& p, int kind, BatchTune* out)
{
    GsBatch& B = *view(c).batch;
    const std::string key = plan_key(p, kind);
    for (auto& kv : B.tuned)
        if (kv.first == key) { *out = kv.second; return GS_OK; }
    BatchTune best{};
    float tbest = 1e30f;
    GsBatch::Plan* pbest = nullptr;
    const char* ev_log = getenv("GS_TUNE_LOG");      // diagnostics: each candidate's time on stderr
    const bool log = ev_log && *ev_log && *ev_log != '0';
    auto trial = [&](const BatchTune& t) -> int {
        for (auto& pl : B.plans)        // already timed
            if (pl.kind == kind && same_params(pl.p, p) && pl.tune == t) return GS_OK;
        float ms = 0;
        GsBatch::Plan* P = nullptr;
        int rc = time_plan(c, p, t, kind, &ms, &P);
        if (rc == GS_ELIMIT) {
            if (log) fprintf(stderr, "[gs tune] %dPCF m %d sx %d: %s (skipped)\n", kind, t.m, t.sx, gs_last_error());
            return GS_OK;           // a candidate beyond a capacity limit (e.g. too many cells) is skipped
        }
        if (rc) return rc;
        if (log)
            fprintf(stderr, "[gs tune] %dPCF m %d sx %d wg %d kmax1 %d: %.3f ms\n", kind, P->m, P->sx, P->wg1,
                    P->kmax1, ms);
        if (ms < tbest) { tbest = ms; best = t; }
        return GS_OK;
    };
    auto best_plan = [&]() -> GsBatch::Plan* {
        for (auto& pl : B.plans)
            if (pl.kind == kind && same_params(pl.p, p) && pl.tune == best) return &pl;
        return nullptr;
    };
    int rc;
    for (int m = kind == 2 ? 2 : 1; m <= (kind == 2 ? 6 : 3); ++m) {
        BatchTune t{};
        t.m = m;
        if ((rc = trial(t))) return rc;
    }
    if (best.m == 0) return fail(GS_ELIMIT, "batched grid too large at every tuned cell size");
    if (kind == 2) {
        const BatchTune b0 = best;
        for (int sx : {1, 2, 4, 8, 16}) {
            if (sx == pair_cell_sx()) continue;     // the automatic choice, timed above
            BatchTune t = b0;
            t.sx = sx;
            if ((rc = trial(t))) return rc;
        }
    } else {
        BatchTune b0 = best;
        for (int sx : {1, 2, 4, 8}) {
            if (sx == apex_cell_sx()) continue;     // the automatic choice, timed above
            BatchTune t = b0;
            t.sx = sx;
            if ((rc = trial(t))) return rc;
        }
        b0 = best;
        pbest = best_plan();
        const int wg_auto = pbest ? pbest->wg1 : kWG;
        for (int wg : {64, 128, 256}) {
            if (wg == wg_auto) continue;
            BatchTune t = b0;
            t.wg = wg;
            if ((rc = trial(t))) return rc;
        }
        b0 = best;
        pbest = best_plan();
        if (pbest) {
            const int k1 = pbest->kmax1, kx = pbest->kmax;
            for (int cand : {k1 / 2 / 64 * 64, std::min(kx, k1 * 2)}) {
                if (cand < 64 || cand == k1) continue;
                BatchTune t = b0;
                t.kmax1 = cand;
                if ((rc = trial(t))) return rc;
            }
        }
    }
    // keep only the winner's plan among this parameter set's
    for (size_t k = 0; k < B.plans.size();) {
        GsBatch::Plan& pl = B.plans[k];
        if (pl.kind == kind && same_params(pl.p, p) && !(pl.tune == best)) {
            if (pl.exec) (void)hipGraphExecDestroy(pl.exec);
            if (pl.graph) (void)hipGraphDestroy(pl.graph);
            pl.blob.release();
            B.plans.erase(B.plans.begin() + k);
        } else {
            ++k;
        }
    }
    B.tuned.emplace_back(key, best);
    *out = best;
    return GS_OK;
}

extern "C" int gs_count_2pcf_batched(gs_ctx* c, const gs_params* p, uint64_t* cnt, double* wsum, double* n2,
                                     double* n3, gs_stats* st)
{
    if (!c) return fail(GS_EINVAL, "NULL context");
    int rc = check_params(p, 2);
    if (rc) return rc;
    if ((rc = ctx_set_device(c))) return rc;
    CtxView v = view(c);
    GsBatch& B = *v.batch;
    if (!B.valid) return fail(GS_EINVAL, "no batched catalogue (gs_set_catalogue_batched)");
    const int nreg = B.nreg, nb = p->nbins, nm = p->nmu, nbt = 3 * nb * nm;
    const bool weighted = B.has_w;

    BatchTune tu{};
    tu.m = p->cell_div;
    if (tu.m == 0 && (rc = batch_tune(c, *p, 2, &tu))) return rc;
    GsBatch::Plan* P;
    if ((rc = get_plan(c, *p, tu, 2, &P))) return rc;
    float ms = 0;
    if ((rc = run_plan(c, *P, &ms))) return rc;
    const size_t nout = (size_t)nreg * 3 * nbt + 4;
    std::vector<unsigned long long> h(nout);
    hipError_t e = hipMemcpy(h.data(), B.out.ptr, nout * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "batched result copy");
    uint64_t total = 0;
    for (int r = 0; r < nreg; ++r) {
        const unsigned long long* hc = h.data() + (size_t)r * nbt;
        const unsigned long long* hw = h.data() + (size_t)nreg * nbt + (size_t)r * 2 * nbt;
        std::vector<long double> S(nbt);
        for (int k = 0; k < nbt; ++k) {
            if (cnt) cnt[(size_t)r * nbt + k] = hc[k];
            total += hc[k];
            // class sums carry the region's scale of the class: DD 2 sd, DR sd + sr, RR 2 sr
            const int cls = k / (nb * nm);
            const int sh = cls == 0 ? 2 * P->fx_sd_r[r] : cls == 1 ? P->fx_sd_r[r] + P->fx_sr_r[r] : 2 * P->fx_sr_r[r];
            S[k] = weighted ? fx_value(hw[2 * k], hw[2 * k + 1], sh) : (long double)hc[k];
            if (wsum) wsum[(size_t)r * nbt + k] = (double)S[k];
        }
        const long double Wd = B.wd[r], Wr = B.wr[r];
        for (int b = 0; b < nb * nm; ++b) {
            const long double dd = S[b], dr = S[nb * nm + b], rr = S[2 * nb * nm + b];
            long double v2 = 0, v3 = 0;
            if (Wd > 0) v2 += dd / (Wd * Wd);
            if (Wd > 0 && Wr > 0) v2 -= dr / (Wd * Wr);
            if (Wr > 0) { v2 += rr / (Wr * Wr); v3 = rr / (Wr * Wr); }
            if (n2) n2[(size_t)r * nb * nm + b] = (double)v2;
            if (n3) n3[(size_t)r * nb * nm + b] = (double)v3;
        }
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->ms_count = ms;
        st->ms_kernel_avg = ms;
        st->n_units = total;
        st->n_binned = total;
        const unsigned long long* s = h.data() + (size_t)nreg * 3 * nbt;
        st->n_candidates = s[0];
        st->n_evaluated = s[0];
        st->n_recheck = s[1];
        st->n_overflow = s[2];
        st->cell_div = P->mst;
    }
    return GS_OK;
}

extern "C" int gs_count_3pcf_batched(gs_ctx* c, const gs_params* p, uint64_t* cnt, double* wsum, double* n2,
                                     double* n3, gs_stats* st)
{
    if (!c) return fail(GS_EINVAL, "NULL context");
    int rc = check_params(p, 3);
    if (rc) return rc;
    if ((rc = ctx_set_device(c))) return rc;
    CtxView v = view(c);
    GsBatch& B = *v.batch;
    if (!B.valid) return fail(GS_EINVAL, "no batched catalogue (gs_set_catalogue_batched)");
    const int nreg = B.nreg, nb = p->nbins, nm = p->nmu;
    const int ncfg = nb * (nb + 1) * (nb + 2) / 6;
    const bool weighted = B.has_w;

    BatchTune tu{};
    tu.m = p->cell_div;
    if (tu.m == 0 && (rc = batch_tune(c, *p, 3, &tu))) return rc;
    GsBatch::Plan* P;
    if ((rc = get_plan(c, *p, tu, 3, &P))) return rc;
    float ms = 0;
    if ((rc = run_plan(c, *P, &ms))) return rc;
    const int nh = P->nh;
    const size_t nout = (size_t)nreg * 3 * nh + 8;
    std::vector<unsigned long long> h(nout);
    int32_t err = 0;
    hipError_t e = hipMemcpy(h.data(), B.out.ptr, nout * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&err, B.queue.ptr + 8, sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "batched result copy");
    if (err) return fail(GS_EHIP, "batched 3PCF: an apex's neighbour list overflowed with no later pass (internal error)");
    std::vector<long double> S(nh);
    for (int r = 0; r < nreg; ++r) {
        const unsigned long long* hc = h.data() + (size_t)r * nh;
        const unsigned long long* hw = h.data() + (size_t)nreg * nh + (size_t)r * 2 * nh;
        for (int k = 0; k < nh; ++k) {
            if (cnt) cnt[(size_t)r * nh + k] = hc[k];
            S[k] = weighted ? fx_value(hw[2 * k], hw[2 * k + 1], P->fx_shift_r[r]) : (long double)hc[k];
            if (wsum) wsum[(size_t)r * nh + k] = (double)S[k];
        }
        combine3(S.data(), ncfg * nm, B.wd[r], B.wr[r], true, n2 ? n2 + (size_t)r * ncfg * nm : nullptr,
                 n3 ? n3 + (size_t)r * ncfg * nm : nullptr);
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->ms_count = ms;
        st->ms_kernel_avg = ms;
        const unsigned long long* s = h.data() + (size_t)nreg * 3 * nh;
        st->n_candidates = s[0];
        st->n_recheck = s[1];
        st->n_overflow = s[2];
        st->n_units = s[3];
        st->n_evaluated = s[4];
        st->n_binned = 0;           // the binned triplets are the counts' total
        for (size_t k = 0; k < (size_t)nreg * nh; ++k) st->n_binned += h[k];
        st->cell_div = P->mst;
    }
    return GS_OK;
}

// adjoint_api.cpp -- C ABI of the 2PCF adjoint (gs_adjoint_2pcf, gs_adjoint_2pcf_device).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gs_api_internal.h"
#include "gs_kernels.h"

using namespace gs;

static int adjoint_impl(gs_ctx* c, const gs_params* p, const double* gN2, const double* gN3, double* out_host,
                        double* out_dev, gs_stats* st)
{
    if (!c) return fail(GS_EINVAL, "NULL context");
    int rc = check_params(p, 2);
    if (rc) return rc;
    if (!gN2 || !gN3) return fail(GS_EINVAL, "dL/dN2 and dL/dN3 are required");
    if ((rc = ctx_set_device(c))) return rc;
    const int nb = p->nbins, nm = p->nmu;
    if (nb * nm > 2048) return fail(GS_ELIMIT, "adjoint supports nbins*nmu <= 2048");
    const bool rsd = nm >= 2, fp64 = p->precision == GS_FP64;
    CtxView v = view(c);
    const int64_t n = v.n_data + v.n_rand;
    const int m = p->cell_div > 0 ? p->cell_div : default_cell_div(c, p->rmax, 2);
    const int sx = pair_cell_sx();

    (void)hipEventRecord(v.ev[0], v.stream);
    double h;
    if ((rc = ctx_build_grid(c, p->rmax, m, sx, &h))) return rc;
    (void)hipEventRecord(v.ev[1], v.stream);
    v = view(c);
    const int mst = (int)std::ceil(p->rmax / h - 1e-12);
    BinDesc B;
    if ((rc = fill_bins(*p, (mst + 1) * h, ctx_xmax(c), v.wmax_d, v.wmax_r, B))) return rc;
    if (!(v.wsum_d > 0) && v.n_data > 0) return fail(GS_EINVAL, "data weights sum to zero");
    if (!(v.wsum_r > 0) && v.n_rand > 0) return fail(GS_EINVAL, "random weights sum to zero");
    const double inv_wd = v.n_data > 0 ? (double)(1.0L / v.wsum_d) : 0.0;
    const double inv_wr = v.n_rand > 0 ? (double)(1.0L / v.wsum_r) : 0.0;
    // FP32: the half-stencil kernel with exact fixed-point sums at the scale 2^F, chosen so that no pair
    // term c * wn exceeds 2^50 (|c| <= 2 gmax, |wn| <= wmax / min(Wd, Wr)); split into a scale of the
    // dL/dN tables (2^Fa, bringing gmax to [1, 2)) and one of the weights (2^Fb).  Non-finite dL/dN
    // (or scales out of range) take the full-stencil kernel, which sums in FP64.
    bool half = !fp64;
    int F = 0, Fa = 0, Fb = 0;
    {
        double gmax = 0.0;
        for (int k = 0; k < nb * nm; ++k) {
            const double a2 = gN2[k], a23 = gN2[k] + gN3[k];
            if (!std::isfinite(a2) || !std::isfinite(a23)) half = false;
            else gmax = std::max(gmax, std::max(std::fabs(a2), std::fabs(a23)));
        }
        const double wn_max = v.wmax * std::max(inv_wd, inv_wr);
        if (half && gmax > 0.0 && wn_max > 0.0) {
            const double T = 2.0 * gmax * wn_max;
            F = 50 - (int)std::ceil(std::log2(T));
            Fa = -std::ilogb(gmax);
            Fb = F - Fa;
            if (!std::isfinite(T) || T <= 0.0 || std::abs(Fa) > 1000 || std::abs(Fb) > 1000 || std::abs(F) > 1000)
                half = false;
        }
    }
    std::vector<Row> rows = half ? half_stencil(h, h / sx, p->rmax, mst) : full_stencil(h, h / sx, p->rmax, mst);

    GsExtra& X = *v.extra;
    const size_t g_off = ((sizeof(BinDesc) + 15) / 16) * 16;
    const size_t rows_off = g_off + ((size_t)nb * nm * 16 + 15) / 16 * 16;
    const int nblk_red = 256;
    if ((rc = X.params.reserve(rows_off + rows.size() * sizeof(Row) + 16)) || (rc = X.out.reserve(8)) ||
        (rc = X.queue.reserve(16)) ||
        (rc = X.dbl.reserve(2 * (size_t)n + 2 * nblk_red + 8 + 2 * (size_t)std::max<int64_t>(n, 1))))
        return rc;
    double* G = X.dbl.ptr;
    double* outd = out_dev ? out_dev : X.dbl.ptr + n;
    double* part = X.dbl.ptr + 2 * n;
    unsigned long long* Gq = (unsigned long long*)(X.dbl.ptr + 2 * n + 2 * nblk_red + 8);   // [max(n, 1)][2]
    hipError_t e = hipMemcpyAsync(X.params.ptr, &B, sizeof(B), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(X.params.ptr + g_off, gN2, (size_t)nb * nm * 8, hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(X.params.ptr + g_off + (size_t)nb * nm * 8, gN3, (size_t)nb * nm * 8,
                           hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && !rows.empty())
        e = hipMemcpyAsync(X.params.ptr + rows_off, rows.data(), rows.size() * sizeof(Row), hipMemcpyHostToDevice,
                           v.stream);
    if (e == hipSuccess) e = hipMemsetAsync(X.queue.ptr, 0, 16 * sizeof(int32_t), v.stream);
    if (e == hipSuccess) e = hipMemsetAsync(X.out.ptr, 0, 8 * sizeof(unsigned long long), v.stream);
    if (e == hipSuccess) e = hipMemsetAsync(G, 0, (size_t)std::max<int64_t>(n, 1) * 8, v.stream);
    if (e == hipSuccess && half) e = hipMemsetAsync(Gq, 0, (size_t)std::max<int64_t>(n, 1) * 16, v.stream);
    if (e != hipSuccess) return hip_fail(e, "adjoint parameter upload");

    AdjArgs a{};
    a.pw = v.pw;
    a.ws = v.ws;
    a.pl = v.pl;
    a.flags = v.flags;
    a.orig = v.orig;
    a.cell_start = v.cell_start;
    a.row_chunk = v.row_chunk;
    a.rows = (const Row*)(X.params.ptr + rows_off);
    a.bins = (const BinDesc*)X.params.ptr;
    a.g2 = (const double*)(X.params.ptr + g_off);
    a.g3 = a.g2 + nb * nm;
    a.queue = X.queue.ptr;
    a.G = G;
    a.out_stats = X.out.ptr;
    a.g = v.g;
    a.g.m = mst;
    a.nrows = (int)rows.size();
    a.n_chunks = v.n_chunks;
    a.inv_wd = inv_wd;
    a.inv_wr = inv_wr;
    a.Gq = Gq;
    a.g_scale = std::ldexp(1.0, Fa);
    a.w_scale = std::ldexp(1.0, Fb);
    const int occ = half ? adjoint2h_occupancy(nb, nm, rsd) : adjoint2_occupancy(nb, nm, fp64, rsd);
    int nblocks = std::max(1, std::min(v.n_cu * occ, std::max(v.n_chunks, 1)));
    (void)hipEventRecord(v.ev[2], v.stream);
    if (v.n_chunks > 0) {
        e = half ? launch_adjoint2h(a, nb, nm, rsd, nblocks, v.stream)
                 : launch_adjoint2(a, nb, nm, fp64, rsd, nblocks, v.stream);
        if (e != hipSuccess) return hip_fail(e, "adjoint kernel launch");
        if (half && (e = launch_adj_fx(Gq, v.orig, n, F, G, v.stream)) != hipSuccess)
            return hip_fail(e, "adjoint fixed-point conversion");
    }
    (void)hipEventRecord(v.ev[3], v.stream);
    e = launch_adj_partials(G, v.w, n, v.n_data, inv_wd, inv_wr, part, nblk_red, v.stream);
    std::vector<double> hp(2 * nblk_red);
    unsigned long long hs[2] = {0, 0};
    if (e == hipSuccess)
        e = hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hs, X.out.ptr, 16, hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) return hip_fail(e, "adjoint kernel");
    long double sd = 0, sr = 0;
    for (int b = 0; b < nblk_red; ++b) { sd += hp[2 * b]; sr += hp[2 * b + 1]; }
    e = launch_adj_finish(G, n, v.n_data, (double)sd, (double)sr, inv_wd, inv_wr, outd, v.stream);
    if (e == hipSuccess && out_host) e = hipMemcpyAsync(out_host, outd, (size_t)n * 8, hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) return hip_fail(e, "adjoint finish");
    if (st) {
        memset(st, 0, sizeof(*st));
        float ms_grid = 0, ms_count = 0;
        (void)hipEventElapsedTime(&ms_grid, v.ev[0], v.ev[1]);
        (void)hipEventElapsedTime(&ms_count, v.ev[2], v.ev[3]);
        st->ms_grid = ms_grid;
        st->ms_count = ms_count;
        st->ms_kernel_avg = ms_count;
        st->n_candidates = hs[0];
        st->n_evaluated = hs[0];
        st->n_recheck = hs[1];
        st->cell_div = mst;
    }
    return GS_OK;
}

extern "C" int gs_adjoint_2pcf(gs_ctx* c, const gs_params* p, const double* dl_dn2, const double* dl_dn3,
                               double* dl_dw, gs_stats* st)
{
    if (!dl_dw) return fail(GS_EINVAL, "dl_dw is NULL");
    return adjoint_impl(c, p, dl_dn2, dl_dn3, dl_dw, nullptr, st);
}

extern "C" int gs_adjoint_2pcf_device(gs_ctx* c, const gs_params* p, const double* dl_dn2, const double* dl_dn3,
                                      double* d_dl_dw, gs_stats* st)
{
    if (!d_dl_dw) return fail(GS_EINVAL, "d_dl_dw is NULL");
    return adjoint_impl(c, p, dl_dn2, dl_dn3, nullptr, d_dl_dw, st);
}

// api.cpp -- the C ABI (include/gramsci_hip.h): contexts, catalogue upload,
// grid build, exact bin thresholds, kernel launches and the host-side
// combination of exact class sums into the reference's N2 / N3.
#include "../../include/gramsci_hip.h"

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gs_kernels.h"
#include "gs_api_internal.h"

using namespace gs;

namespace gs {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(GS_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// ------------------------------------------------------------ device buffers
template <class T>
int DevBuf<T>::reserve(size_t count)
{
    if (count <= cap) return GS_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t c = std::max<size_t>(count, 1);
    hipError_t e = hipMalloc((void**)&ptr, c * sizeof(T));
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(GS_ENOMEM, "hipMalloc(%zu bytes): %s", c * sizeof(T), hipGetErrorString(e));
    }
    cap = c;
    return GS_OK;
}

template <class T>
void DevBuf<T>::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

template struct DevBuf<double>;
template struct DevBuf<int32_t>;
template struct DevBuf<uint32_t>;
template struct DevBuf<double4>;
template struct DevBuf<float4>;
template struct DevBuf<unsigned long long>;
template struct DevBuf<unsigned char>;
template struct DevBuf<int64_t>;
template struct DevBuf<float>;

// ------------------------------------------------------------ thresholds
// The reference's radial bin of a squared distance s (gramsci.F90:136-145,
// :334-338), 1-based, evaluated with the host libm exactly as written.
static int ref_bin(const gs_params& p, double s)
{
    const double d = std::sqrt(s);
    if (p.log_bins) {
        const double delta_r = (std::log10(p.rmax) - std::log10(p.rmin)) / (double)p.nbins;
        const double inv_delta_r = 1.0 / delta_r;
        const double log_rmin = std::log10(p.rmin);
        return (int)std::floor((std::log10(d) - log_rmin) * inv_delta_r) + 1;
    }
    const double delta_r = (p.rmax - p.rmin) / (double)p.nbins;
    const double inv_delta_r = 1.0 / delta_r;
    return (int)std::floor((d - p.rmin) * inv_delta_r) + 1;
}

static inline uint64_t dbits(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}
static inline double bitsd(uint64_t u)
{
    double x;
    memcpy(&x, &u, 8);
    return x;
}

// smallest non-negative double s in [lo, hi] (bit patterns) with pred(s) true; pred monotone false->true.
template <class F>
static uint64_t first_true(uint64_t lo, uint64_t hi, F pred)
{
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (pred(bitsd(mid))) hi = mid; else lo = mid + 1;
    }
    return lo;
}

int check_params(const gs_params* p, int mode)
{
    if (!p) return fail(GS_EINVAL, "params is NULL");
    if (p->nbins < 1 || p->nbins > kMaxBins) return fail(GS_EINVAL, "nbins must be in 1..127 (got %d)", p->nbins);
    if (p->nmu < 1 || p->nmu > kMaxBins) return fail(GS_EINVAL, "nmu must be in 1..127 (got %d)", p->nmu);
    if (!(p->rmin >= 0.0) || !(p->rmax > p->rmin) || !std::isfinite(p->rmax))
        return fail(GS_EINVAL, "need 0 <= rmin < rmax (got rmin=%g rmax=%g)", p->rmin, p->rmax);
    if (p->log_bins && !(p->rmin > 0.0)) return fail(GS_EINVAL, "log bins need rmin > 0");
    if (p->precision != GS_FP64 && p->precision != GS_FP32_GUARD)
        return fail(GS_EINVAL, "precision must be GS_FP64 or GS_FP32_GUARD");
    if (p->cell_div < 0 || p->cell_div > 16) return fail(GS_EINVAL, "cell_div must be in 0..16");
    if (mode >= 4 && p->nmu != 1) return fail(GS_EINVAL, "the 4PCF has no mu binning: nmu must be 1");
    return GS_OK;
}

}  // namespace gs

extern "C" int gs_bin_thresholds(const gs_params* p, double* t)
{
    int rc = check_params(p, 2);
    if (rc) return rc;
    if (!t) return fail(GS_EINVAL, "t is NULL");
    const gs_params P = *p;
    const int nb = P.nbins;
    const double rmax2 = P.rmax * P.rmax;
    const uint64_t top = dbits(rmax2 * 4.0 + 1.0);
    // t[0]: largest s with sqrt(s) <= rmin
    uint64_t u = first_true(0, top, [&](double s) { return std::sqrt(s) > P.rmin; });
    t[0] = bitsd(u - (u > 0 ? 1 : 0));
    if (std::sqrt(t[0]) > P.rmin) t[0] = 0.0;
    // t[nb+1]: smallest s not kept by the r-ball (s <= rmax^2) and the strict d < rmax
    const uint64_t hi = first_true(dbits(t[0]), top, [&](double s) { return !(s <= rmax2 && std::sqrt(s) < P.rmax); });
    t[nb + 1] = bitsd(hi);
    const uint64_t lo = dbits(t[0]) + 1;
    for (int k = 1; k <= nb; ++k) {
        uint64_t v = first_true(lo, hi, [&](double s) { return ref_bin(P, s) >= k + 1; });
        t[k] = bitsd(v);
    }
    // the bin must be a non-decreasing step function of s near each edge, with
    // bin >= 1 just above rmin; verify in a window of ulps around every threshold
    const int win = 256;
    if (lo < hi && ref_bin(P, bitsd(lo)) < 1) return fail(GS_EINVAL, "bin formula gives 0 just above rmin");
    for (int k = 0; k <= nb + 1; ++k) {
        const uint64_t c = dbits(t[k]);
        int prev = -1000000;
        for (int64_t d = -win; d <= win; ++d) {
            const int64_t b = (int64_t)c + d;
            if (b <= (int64_t)lo || b >= (int64_t)hi) continue;
            const int v = ref_bin(P, bitsd((uint64_t)b));
            if (v < prev) return fail(GS_EINVAL, "bin formula not monotone near threshold %d", k);
            prev = v;
        }
    }
    return GS_OK;
}

namespace gs {

// ------------------------------------------------------------ FP32 guard band
// Upper bound of |s32 - s| for a pair whose cell-relative FP32 coordinates are
// bounded by rrel per component (the squared distance s near t).
static double s_guard(double t, double ew)
{
    const double root = std::sqrt(std::max(t, 0.0));
    return 2.0 * (2.0 * std::sqrt(3.0) * root * ew + 3.0 * ew * ew + std::ldexp(t, -22));
}

static float f_up(double x)
{
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}
static float f_down(double x)
{
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}

int fill_bins(const gs_params& p, double rrel, double xmax, double wmax_d, double wmax_r, BinDesc& B)
{
    memset(&B, 0, sizeof(B));
    int rc = gs_bin_thresholds(&p, B.t64);
    if (rc) return rc;
    const int nb = p.nbins;
    B.nb = nb;
    B.nm = p.nmu;
    B.log_bins = p.log_bins;
    B.rsd = p.nmu >= 2;
    B.mu_scale = (double)(0.5f * (float)p.nmu);
    B.ms32 = (float)B.mu_scale;
    // error of one cell-relative coordinate difference (FP64 -> FP32 rounding of both ends + the FP32 subtract)
    const double ew = std::ldexp(2.0 * rrel + p.rmax, -24) * 1.01;
    for (int b = 0; b < nb; ++b) {
        const double lo = b == 0 ? B.t64[0] : B.t64[b];
        const double hi = b == nb - 1 ? std::min(B.t64[nb], B.t64[nb + 1]) : B.t64[b + 1];
        B.ok_lo[b] = f_up(lo + s_guard(lo, ew));
        B.ok_hi[b] = f_down(hi - s_guard(hi, ew));
    }
    B.rej_lo = f_down(B.t64[0] - s_guard(B.t64[0], ew));
    B.rej_hi = f_up(B.t64[nb + 1] + s_guard(B.t64[nb + 1], ew));
    if (p.log_bins) {
        const double delta = (std::log10(p.rmax) - std::log10(p.rmin)) / nb;
        const double inv = 1.0 / delta;
        B.ua = (float)(0.5 * std::log10(2.0) * inv);
        B.ub = (float)(-std::log10(p.rmin) * inv);
    } else {
        const double inv = 1.0 / ((p.rmax - p.rmin) / nb);
        B.ua = (float)inv;
        B.ub = (float)(-p.rmin * inv);
    }
    // mu guard: per-component errors of sep (ew) and of the absolute midpoint (em)
    const double em = std::ldexp(2.0 * xmax + 3.0 * rrel, -24) * 1.01;
    B.mu_ew = (float)(2.0 * std::sqrt(3.0) * ew * 4.0);
    B.mu_em = (float)(2.0 * std::sqrt(3.0) * em * 4.0);
    B.mu_c = (float)std::ldexp(1.0, -18);
    B.fx_sd = fx_shift_point(wmax_d);
    B.fx_sr = fx_shift_point(wmax_r);
    B.fx_shift3 = fx_shift_triples(std::max(wmax_d, wmax_r));
    return GS_OK;
}

// Fixed-point scales 2^F of the weights and products.  Pairs: each weight is scaled by its
// class's 2^F, below 2^26, so the largest pair product of every class stays below 2^52 (the
// pair kernel rounds it to an integer with one FP64 addition of 2^52 and sums runs of up to
// 33 in one 64-bit register) and each class sum is exact to 2^-52 of its own largest product.
// Triples: the largest triple product stays below 2^61.
int fx_shift_point(double wmax)
{
    const double wm = std::max(wmax, 1e-300);
    return std::min(500, 26 - (int)std::ceil(std::log2(wm * (1.0 + 1e-12))));
}

int fx_shift_triples(double wmax)
{
    const double wm = std::max(wmax, 1e-300);
    return std::min(1000, 61 - (int)std::ceil(std::log2(wm * wm * wm * (1.0 + 1e-12))));
}

// Half stencil of x-rows for cells of edge h in y/z and hx in x, search radius rmax, m = ceil(rmax/h).
// Row (dy, dz) takes x-offsets dx with gap(dy)^2 h^2 + gap(dz)^2 h^2 + gap(dx)^2 hx^2 <= rmax^2 (plus a
// relative margin of 1e-6), gap(d) = max(|d| - 1, 0) being the fewest whole cells between two cells d
// apart: no pair within rmax is left out, and a fine hx cuts each row close to the sphere.
std::vector<Row> half_stencil(double h, double hx, double rmax, int m)
{
    std::vector<Row> rows;
    auto gap = [](int d) { return (double)std::max(std::abs(d) - 1, 0); };
    const double lim = rmax * (1.0 + 1e-6) + 1e-9 * h;
    const double lim2 = lim * lim;
    const int mxmax = (int)std::ceil(lim / hx) + 1;
    for (int dz = 0; dz <= m; ++dz)
        for (int dy = -m; dy <= m; ++dy) {
            if (dz == 0 && dy < 0) continue;
            const double gyz = (gap(dy) * gap(dy) + gap(dz) * gap(dz)) * h * h;
            if (gyz > lim2) continue;
            int mx = 0;
            for (int dx = 1; dx <= mxmax; ++dx)
                if (gyz + gap(dx) * gap(dx) * hx * hx <= lim2) mx = dx;
            Row r;
            r.dy = (int16_t)dy;
            r.dz = (int16_t)dz;
            r.x0 = (int16_t)((dz == 0 && dy == 0) ? 0 : -mx);
            r.x1 = (int16_t)mx;
            rows.push_back(r);
        }
    return rows;
}

// Full stencil (both half spaces) as x-rows, for apex-centred kernels.
std::vector<Row> full_stencil(double h, double hx, double rmax, int m)
{
    std::vector<Row> half = half_stencil(h, hx, rmax, m);
    std::vector<Row> rows;
    for (const Row& r : half) {
        if (r.dy == 0 && r.dz == 0) {
            Row c = r;
            c.x0 = (int16_t)(-r.x1);
            rows.push_back(c);
        } else {
            rows.push_back(r);
            Row n = r;
            n.dy = (int16_t)(-r.dy);
            n.dz = (int16_t)(-r.dz);
            rows.push_back(n);
        }
    }
    return rows;
}

// 128-bit fixed point (lo, hi) scaled by 2^-shift -> long double
long double fx_value(unsigned long long lo, unsigned long long hi, int shift)
{
    long double v = (long double)hi * 18446744073709551616.0L + (long double)lo;
    return std::ldexp(v, -shift);
}

}  // namespace gs

// ====================================================================== context

struct gs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int n_cu = 256;
    // catalogue (original order)
    int64_t n_data = 0, n_rand = 0;
    bool has_w = false, has_buf = false;
    long double wsum_d = 0, wsum_r = 0;
    double wmax = 1.0, wmax_d = 1.0, wmax_r = 1.0;
    double bbox[6] = {0, 0, 0, 0, 0, 0};
    DevBuf<double> xyz, w;
    DevBuf<int32_t> buf;
    // grid
    GridDesc g{};
    bool grid_valid = false;
    double grid_h = 0.0;
    int grid_sx = 1;
    DevBuf<int32_t> cell_of, cell_start, fill, chunk_nc, chunk_off, scan_tmp;
    DevBuf<double4> pw;
    DevBuf<double> ws;          // sorted raw weights
    DevBuf<float4> pl;
    DevBuf<uint32_t> flags;
    DevBuf<int32_t> orig;
    int32_t n_chunks = 0;
    // launch scratch
    DevBuf<unsigned char> params;
    DevBuf<unsigned long long> out;
    DevBuf<int32_t> queue;
    DevBuf<double> part;
    DevBuf<unsigned long long> wpart;
    GsExtra extra;
    GsBatch batch;
};

using namespace gs;

static int set_device(gs_ctx* c)
{
    hipError_t e = hipSetDevice(c->device);
    return e == hipSuccess ? GS_OK : hip_fail(e, "hipSetDevice");
}

extern "C" const char* gs_last_error(void) { return gs::g_err.c_str(); }
extern "C" const char* gs_version(void) { return "gramsci_hip 0.1 (gfx950)"; }

extern "C" int gs_create(gs_ctx** out, int device)
{
    if (!out) return fail(GS_EINVAL, "ctx is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= device || device < 0) return fail(GS_ENODEV, "no HIP device %d", device);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(GS_ENODEV, "device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
    gs_ctx* c = new gs_ctx();
    c->device = device;
    c->n_cu = prop.multiProcessorCount;
    if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreate(&c->stream)) != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    for (auto& ev : c->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) {
            delete c;
            return hip_fail(e, "hipEventCreate");
        }
    *out = c;
    return GS_OK;
}

extern "C" int gs_destroy(gs_ctx* c)
{
    if (!c) return GS_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->xyz.release(); c->w.release(); c->buf.release();
    c->cell_of.release(); c->cell_start.release(); c->fill.release(); c->chunk_nc.release();
    c->chunk_off.release(); c->scan_tmp.release(); c->pw.release(); c->ws.release(); c->pl.release(); c->flags.release();
    c->orig.release();
    c->params.release(); c->out.release(); c->queue.release(); c->part.release(); c->wpart.release();
    extra_release(c->extra);
    batch_release(c->batch);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GS_OK;
}

static int finish_catalogue(gs_ctx* c)
{
    // bounding box on the device
    const int64_t n = c->n_data + c->n_rand;
    int rc;
    const int nblk = 256;
    if ((rc = c->part.reserve(nblk * 6))) return rc;
    hipError_t e = launch_bbox(c->xyz.ptr, n, c->part.ptr, nblk, c->stream);
    if (e != hipSuccess) return hip_fail(e, "bbox kernel");
    std::vector<double> part(nblk * 6);
    e = hipMemcpyAsync(part.data(), c->part.ptr, nblk * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "bbox copy");
    for (int k = 0; k < 3; ++k) { c->bbox[k] = 1e308; c->bbox[3 + k] = -1e308; }
    for (int b = 0; b < nblk; ++b)
        for (int k = 0; k < 3; ++k) {
            c->bbox[k] = std::min(c->bbox[k], part[b * 6 + k]);
            c->bbox[3 + k] = std::max(c->bbox[3 + k], part[b * 6 + 3 + k]);
        }
    if (n == 0)
        for (int k = 0; k < 6; ++k) c->bbox[k] = 0.0;
    for (int k = 0; k < 6; ++k)
        if (!std::isfinite(c->bbox[k])) return fail(GS_EINVAL, "catalogue has non-finite coordinates");
    c->grid_valid = false;
    return GS_OK;
}

long double gs::fixed_sum(const double* w, int64_t n, int k)
{
    unsigned __int128 s = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t lo, hi;
        wsum_fixed(w[i], k, lo, hi);
        s += ((unsigned __int128)hi << 64) | lo;
    }
    return std::ldexp((long double)s, -k);
}

static int host_weight_stats(gs_ctx* c, const double* w)
{
    c->wsum_d = 0;
    c->wsum_r = 0;
    c->wmax = c->wmax_d = c->wmax_r = 1.0;
    const int64_t n = c->n_data + c->n_rand;
    if (!w) {
        c->wsum_d = (long double)c->n_data;
        c->wsum_r = (long double)c->n_rand;
        return GS_OK;
    }
    double wm[2] = {0.0, 0.0};
    for (int64_t i = 0; i < n; ++i) {
        if (!(w[i] >= 0.0) || !std::isfinite(w[i])) return fail(GS_EINVAL, "weights must be finite and >= 0 (w[%lld]=%g)", (long long)i, w[i]);
        double& m = wm[i < c->n_data ? 0 : 1];
        m = std::max(m, w[i]);
    }
    const double wall = std::max(wm[0], wm[1]);
    const int k = wsum_scale(wall);
    c->wsum_d = fixed_sum(w, c->n_data, k);
    c->wsum_r = fixed_sum(w + c->n_data, c->n_rand, k);
    c->wmax = wall > 0 ? wall : 1.0;
    c->wmax_d = wm[0] > 0 ? wm[0] : 1.0;
    c->wmax_r = wm[1] > 0 ? wm[1] : 1.0;
    return GS_OK;
}

// The same statistics for weights already on the device (c->w): the check, the maximum and the
// fixed-point sums run as reductions on the GPU; only per-block partials come back.
static int device_weight_stats(gs_ctx* c)
{
    const int64_t n = c->n_data + c->n_rand;
    const int nblk = 256;
    int rc;
    if ((rc = c->part.reserve(nblk * 6)) || (rc = c->wpart.reserve(nblk * 4 + 1))) return rc;
    unsigned long long* bad = c->wpart.ptr + nblk * 4;
    hipError_t e = launch_fill_u32(bad, 2, 0xffffffffu, c->stream);
    if (e == hipSuccess) e = launch_wcheck(c->w.ptr, n, c->n_data, c->part.ptr, bad, nblk, c->stream);
    std::vector<double> pmax(2 * nblk);
    unsigned long long first_bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(pmax.data(), c->part.ptr, 2 * nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&first_bad, bad, sizeof(first_bad), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "weight check");
    if (first_bad != ~0ull) {
        double v = 0.0;
        (void)hipMemcpy(&v, c->w.ptr + first_bad, sizeof(double), hipMemcpyDeviceToHost);
        return fail(GS_EINVAL, "weights must be finite and >= 0 (w[%llu]=%g)", first_bad, v);
    }
    double wmc[2] = {0.0, 0.0};
    for (int b = 0; b < nblk; ++b)
        for (int q = 0; q < 2; ++q) wmc[q] = std::max(wmc[q], pmax[2 * b + q]);
    const double wm = std::max(wmc[0], wmc[1]);
    const int k = wsum_scale(wm);
    std::vector<unsigned long long> ps(nblk * 4);
    e = launch_wsum(c->w.ptr, n, c->n_data, k, c->wpart.ptr, nblk, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ps.data(), c->wpart.ptr, ps.size() * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "weight sums");
    unsigned __int128 sd = 0, sr = 0;
    for (int b = 0; b < nblk; ++b) {
        sd += ((unsigned __int128)ps[4 * b + 1] << 64) | ps[4 * b];
        sr += ((unsigned __int128)ps[4 * b + 3] << 64) | ps[4 * b + 2];
    }
    c->wsum_d = std::ldexp((long double)sd, -k);
    c->wsum_r = std::ldexp((long double)sr, -k);
    c->wmax = wm > 0 ? wm : 1.0;
    c->wmax_d = wmc[0] > 0 ? wmc[0] : 1.0;
    c->wmax_r = wmc[1] > 0 ? wmc[1] : 1.0;
    return GS_OK;
}

extern "C" int gs_set_catalogue(gs_ctx* c, const double* xyz, const double* w, const int32_t* buf, int64_t n_data,
                                int64_t n_rand)
{
    if (!c || (!xyz && n_data + n_rand > 0)) return fail(GS_EINVAL, "NULL context or coordinates");
    if (n_data < 0 || n_rand < 0 || n_data + n_rand > (int64_t)INT32_MAX - 1024)
        return fail(GS_EINVAL, "bad catalogue size (n_data=%lld n_rand=%lld)", (long long)n_data, (long long)n_rand);
    int rc = set_device(c);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    c->n_data = n_data;
    c->n_rand = n_rand;
    if ((rc = host_weight_stats(c, w))) return rc;
    c->has_w = w != nullptr;
    c->has_buf = buf != nullptr;
    if ((rc = c->xyz.reserve(3 * n))) return rc;
    hipError_t e = hipMemcpyAsync(c->xyz.ptr, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && w) {
        if ((rc = c->w.reserve(n))) return rc;
        e = hipMemcpyAsync(c->w.ptr, w, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    }
    if (e == hipSuccess && buf) {
        if ((rc = c->buf.reserve(n))) return rc;
        e = hipMemcpyAsync(c->buf.ptr, buf, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    }
    if (e != hipSuccess) return hip_fail(e, "catalogue upload");
    return finish_catalogue(c);
}

extern "C" int gs_set_catalogue_device(gs_ctx* c, const double* d_xyz, const double* d_w, const int32_t* d_buf,
                                       int64_t n_data, int64_t n_rand)
{
    if (!c || (!d_xyz && n_data + n_rand > 0)) return fail(GS_EINVAL, "NULL context or coordinates");
    if (n_data < 0 || n_rand < 0 || n_data + n_rand > (int64_t)INT32_MAX - 1024)
        return fail(GS_EINVAL, "bad catalogue size");
    int rc = set_device(c);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    c->n_data = n_data;
    c->n_rand = n_rand;
    c->has_w = d_w != nullptr;
    c->has_buf = d_buf != nullptr;
    if ((rc = c->xyz.reserve(3 * n))) return rc;
    hipError_t e = hipMemcpyAsync(c->xyz.ptr, d_xyz, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess && d_w) {
        if ((rc = c->w.reserve(n))) return rc;
        e = hipMemcpyAsync(c->w.ptr, d_w, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    }
    if (e == hipSuccess && d_buf) {
        if ((rc = c->buf.reserve(n))) return rc;
        e = hipMemcpyAsync(c->buf.ptr, d_buf, n * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream);
    }
    if (e != hipSuccess) return hip_fail(e, "catalogue copy");
    if ((rc = d_w && n > 0 ? device_weight_stats(c) : host_weight_stats(c, nullptr))) {
        c->n_data = c->n_rand = 0;      // no half-set catalogue stays behind
        c->grid_valid = false;
        return rc;
    }
    return finish_catalogue(c);
}

namespace gs {

// h rounded up to 11 significant bits: then k*h is exact in FP32 for |k| < 2^13
double round11(double h)
{
    int ex;
    const double fr = std::frexp(h, &ex);
    return std::ldexp(std::ceil(std::ldexp(fr, 11)), ex - 11);
}

// Picks the y/z cell edge for a catalogue with bounding box bbox: rmax / m rounded up to 11 significant
// bits (then k*h is exact in FP32 for |k| < 2^13), widened while the grid, with x cells of edge h / sx,
// would exceed max_cells.
double choose_h_bbox(const double* bbox, double rmax, int m, int64_t max_cells, int sx)
{
    int ex;
    double fr = std::frexp(rmax / m, &ex);
    double h = std::ldexp(std::ceil(std::ldexp(fr, 11)), ex - 11);
    for (;;) {
        int64_t nc = 1;
        for (int k = 0; k < 3; ++k) nc *= (int64_t)std::floor((bbox[3 + k] - bbox[k]) / (k == 0 ? h / sx : h)) + 1;
        if (nc <= max_cells) break;
        h = std::ldexp(std::ceil(std::ldexp(std::frexp(h * 1.25, &ex), 11)), ex - 11);
    }
    return h;
}

static double choose_h(const gs_ctx* c, double rmax, int m, int sx)
{
    return choose_h_bbox(c->bbox, rmax, m, (int64_t)1 << 24, sx);
}

// Grid of a catalogue (input-order xyz/w/buf in device memory) over its bounding box with cell edges
// h (y, z) and h / sx (x), built into caller-provided device buffers; *nch receives the number of row
// chunks.
int grid_build_raw(const double* xyz, const double* w, const int32_t* buf, int64_t n, int64_t n_data,
                   const double* bbox, double h, int sx, const GridBufs& b, int64_t max_cells, GridDesc* g_out,
                   int32_t* nch, hipStream_t st)
{
    GridDesc g{};
    g.h = h;
    g.hx = h / sx;
    int64_t ncell = 1;
    for (int k = 0; k < 3; ++k) {
        g.lo[k] = bbox[k];
        g.dim[k] = (int32_t)std::floor((bbox[3 + k] - bbox[k]) / (k == 0 ? g.hx : h)) + 1;
        ncell *= g.dim[k];
    }
    if (ncell > max_cells) return fail(GS_ELIMIT, "grid too large (%lld cells)", (long long)ncell);
    g.ncell = (int32_t)ncell;
    g.h32 = (float)h;
    g.hx32 = (float)g.hx;
    hipError_t e = launch_fill_u32(b.fill, ncell + 1, 0u, st);
    if (e == hipSuccess) e = launch_cell_count(xyz, n, g, b.cell_of, b.fill, st);
    if (e == hipSuccess) e = exclusive_scan(b.fill, b.cell_start, ncell + 1, b.scan_tmp, st);
    if (e == hipSuccess) e = launch_fill_u32(b.fill, ncell + 1, 0u, st);
    if (e == hipSuccess)
        e = launch_scatter(xyz, w, buf, n, n_data, g, b.cell_of, b.cell_start, b.fill, b.pw, b.ws, b.pl, b.flags, b.orig,
                           st);
    const int nrow = g.dim[1] * g.dim[2];
    if (e == hipSuccess) e = launch_row_chunks(b.cell_start, g.dim[0], nrow, b.chunk_nc, st);
    if (e == hipSuccess) e = exclusive_scan(b.chunk_nc, b.chunk_off, nrow + 1, b.scan_tmp, st);
    int32_t nc = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nc, b.chunk_off + nrow, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "grid build");
    *g_out = g;
    *nch = nc;
    return GS_OK;
}

int build_grid(gs_ctx* c, double h, int sx)
{
    if (c->grid_valid && c->grid_h == h && c->grid_sx == sx) return GS_OK;
    const int64_t n = c->n_data + c->n_rand;
    int64_t ncell = 1;
    for (int k = 0; k < 3; ++k) ncell *= (int64_t)std::floor((c->bbox[3 + k] - c->bbox[k]) / (k == 0 ? h / sx : h)) + 1;
    if (ncell > (1 << 26)) return fail(GS_ELIMIT, "grid too large (%lld cells)", (long long)ncell);
    int rc;
    if ((rc = c->cell_of.reserve(n)) || (rc = c->cell_start.reserve(ncell + 1)) || (rc = c->fill.reserve(ncell + 1)) ||
        (rc = c->chunk_nc.reserve(ncell + 1)) || (rc = c->chunk_off.reserve(ncell + 1)) ||
        (rc = c->scan_tmp.reserve(scan_scratch_size(ncell + 1))) || (rc = c->pw.reserve(n)) || (rc = c->ws.reserve(n)) ||
        (rc = c->pl.reserve(n)) || (rc = c->flags.reserve(n)) || (rc = c->orig.reserve(n)))
        return rc;
    GridBufs b{c->cell_of.ptr, c->cell_start.ptr, c->fill.ptr, c->chunk_nc.ptr, c->chunk_off.ptr, c->scan_tmp.ptr,
               c->pw.ptr, c->ws.ptr, c->pl.ptr, c->flags.ptr, c->orig.ptr};
    GridDesc g;
    int32_t nch = 0;
    if ((rc = grid_build_raw(c->xyz.ptr, c->has_w ? c->w.ptr : nullptr, c->has_buf ? c->buf.ptr : nullptr, n,
                             c->n_data, c->bbox, h, sx, b, (int64_t)1 << 26, &g, &nch, c->stream)))
        return rc;
    c->g = g;
    c->grid_h = h;
    c->grid_sx = sx;
    c->n_chunks = nch;
    c->grid_valid = true;
    return GS_OK;
}

// Cells per rmax of the pair kernel for n points in bounding box bbox: the finest grid (m <= 8)
// whose cells still hold ~32 points or more.
int pair_cell_div(int64_t n, const double* bbox, double rmax)
{
    double vol = 1.0;
    for (int k = 0; k < 3; ++k) vol *= std::max(bbox[3 + k] - bbox[k], rmax * 1e-3);
    const double dens = n / vol;
    int best = 1;
    for (int m = 1; m <= 8; ++m) {
        const double h = rmax / m;
        if (dens * h * h * h >= 32.0) best = m;
    }
    return best;
}

// x cells per y/z cell of the pair kernels' grid: finer x cells cut each stencil row to the x-range that
// can reach rmax from the work item's apexes, at the cost of one more entry per cell in the cell index.
static int cell_sx_env(const char* name, int dflt)
{
    const char* ev = getenv(name);                  // tuning override (unset or empty: automatic)
    if (ev && *ev) {
        int v = atoi(ev), sx = 1;
        while (sx < 64 && sx * 2 <= v) sx *= 2;
        return sx;
    }
    return dflt;
}
int pair_cell_sx() { return cell_sx_env("GS_CELL_SX", 8); }

// x cells per y/z cell of the triplet and quad kernels' grid: their gathers scan the apex's stencil, whose
// x-range finer x cells cut close to the sphere.
int apex_cell_sx() { return cell_sx_env("GS_T3_CELL_SX", 4); }

int default_cell_div(const gs_ctx* c, double rmax, int mode)
{
    if (mode != 2) return 1;
    return pair_cell_div(c->n_data + c->n_rand, c->bbox, rmax);
}

}  // namespace gs

// ====================================================================== 2PCF

namespace gs {

// Launch shape of the pair kernel: work-group size and LDS histogram replicas (spreading same-bin
// atomics).  256-thread work-groups (four waves per shared histogram) are kept when they reach 16
// resident waves per CU; otherwise larger work-groups share one histogram among more waves to raise
// residency.  Among shapes with the same residency, more replicas win.  Cached per variant.
int pair2_shape(int nb, int nm, bool fp64, bool weighted, bool logb, int* reps_out, int* wg_out, int* waves_out)
{
    const bool rsd = nm >= 2;
    int reps = 1, wg = kWG, best_waves = 0;
    const char* ev_reps = getenv("GS_PAIR_REPS");    // tuning overrides (unset or empty: automatic)
    const char* ev_wg = getenv("GS_PAIR_WG");
    if (ev_reps && !*ev_reps) ev_reps = nullptr;
    if (ev_wg && !*ev_wg) ev_wg = nullptr;
    char key[96];
    snprintf(key, sizeof(key), "%d/%d/%d/%d/%d/%d/%s/%s", nb, nm, (int)fp64, (int)rsd, (int)weighted, (int)logb,
             ev_reps ? ev_reps : "", ev_wg ? ev_wg : "");
    static std::mutex mu;
    static std::map<std::string, std::array<int, 3>> shapes;
    std::lock_guard<std::mutex> lock(mu);
    auto it = shapes.find(key);
    if (it != shapes.end()) {
        reps = it->second[0]; wg = it->second[1]; best_waves = it->second[2];
    } else {
        for (int pass = 0; pass < 2 && (pass == 0 || best_waves < 16); ++pass) {
            for (int r = 1; r <= 32; r *= 2) {
                for (int w = pass == 0 ? 256 : 384; w <= (pass == 0 ? 256 : 1024); w += 128) {
                    if (ev_reps && atoi(ev_reps) != r) continue;
                    if (ev_wg && atoi(ev_wg) != w) continue;
                    if (pair2_lds_bytes(nb, rsd ? nm : 1, r, fp64, weighted, false, w) > 160 * 1024) continue;
                    const int waves = pair2_occupancy(nb, nm, r, fp64, rsd, weighted, logb, false, w) * (w / kWave);
                    if (waves > best_waves || (waves == best_waves && r > reps && w == wg)) {
                        best_waves = waves;
                        reps = r;
                        wg = w;
                    }
                }
            }
        }
        shapes[key] = {reps, wg, best_waves};
    }
    if (best_waves == 0) return fail(GS_ELIMIT, "histogram of %d bins does not fit in LDS", 3 * nb * nm);
    *reps_out = reps;
    *wg_out = wg;
    *waves_out = best_waves;
    return GS_OK;
}

// Pair kernel outputs (counts [3][nb][nm], 128-bit class sums, statistics) -> the C ABI's outputs:
// cnt, the class sums of raw-weight products, N2 = DD/Wd^2 - DR/(Wd Wr) + RR/Wr^2 and N3 = RR/Wr^2
// (gramsci.F90:373-379 over ordered pairs), and the pair statistics of st (times left to the caller).
void combine2(const unsigned long long* h_out, int nb, int nm, const BinDesc& B, bool weighted, long double Wd,
              long double Wr, uint64_t* cnt, double* wsum, double* n2, double* n3, gs_stats* st)
{
    const int nbt = 3 * nb * nm;
    uint64_t total = 0;
    for (int k = 0; k < nbt; ++k) {
        if (cnt) cnt[k] = h_out[k];
        total += h_out[k];
    }
    std::vector<long double> S(nbt);
    for (int k = 0; k < nbt; ++k)
        S[k] = weighted ? fx_value(h_out[nbt + 2 * k], h_out[nbt + 2 * k + 1], fx_pair_class_shift(B, k / (nb * nm)))
                        : (long double)h_out[k];
    if (wsum)
        for (int k = 0; k < nbt; ++k) wsum[k] = (double)S[k];
    for (int b = 0; b < nb * nm; ++b) {
        const long double dd = S[b], dr = S[nb * nm + b], rr = S[2 * nb * nm + b];
        long double v2 = 0, v3 = 0;
        if (Wd > 0) v2 += dd / (Wd * Wd);
        if (Wd > 0 && Wr > 0) v2 -= dr / (Wd * Wr);
        if (Wr > 0) { v2 += rr / (Wr * Wr); v3 = rr / (Wr * Wr); }
        if (n2) n2[b] = (double)v2;
        if (n3) n3[b] = (double)v3;
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->n_units = total;
        st->n_binned = total;
        st->n_candidates = h_out[3 * nbt + 0];
        st->n_evaluated = st->n_candidates;
        st->n_recheck = h_out[3 * nbt + 1];
        st->n_overflow = h_out[3 * nbt + 2];
    }
}

}  // namespace gs


extern "C" int gs_count_2pcf(gs_ctx* c, const gs_params* p, uint64_t* cnt, double* wsum, double* n2, double* n3,
                             gs_stats* st)
{
    if (!c) return fail(GS_EINVAL, "NULL context");
    int rc = check_params(p, 2);
    if (rc) return rc;
    if ((rc = set_device(c))) return rc;
    const int nb = p->nbins, nm = p->nmu;
    const bool rsd = nm >= 2, fp64 = p->precision == GS_FP64, weighted = c->has_w;
    const int nbt = 3 * nb * nm;
    const int m = p->cell_div > 0 ? p->cell_div : default_cell_div(c, p->rmax, 2);
    const int sx = pair_cell_sx();
    const double h = choose_h(c, p->rmax, m, sx);
    const int mst = (int)std::ceil(p->rmax / h - 1e-12);
    std::vector<Row> rows = half_stencil(h, h / sx, p->rmax, mst);
    if ((int)rows.size() > kMaxRows) return fail(GS_ELIMIT, "stencil too large");

    (void)hipEventRecord(c->ev[0], c->stream);
    if ((rc = build_grid(c, h, sx))) return rc;
    (void)hipEventRecord(c->ev[1], c->stream);

    double xmax = 0;
    for (int k = 0; k < 6; ++k) xmax = std::max(xmax, std::fabs(c->bbox[k]));
    BinDesc B;
    if ((rc = fill_bins(*p, (mst + 1) * h, xmax, c->wmax_d, c->wmax_r, B))) return rc;

    int reps = 1, wg = kWG, best_waves = 0;
    if ((rc = pair2_shape(nb, nm, fp64, weighted, p->log_bins != 0, &reps, &wg, &best_waves))) return rc;

    const size_t nout = (size_t)nbt + 2 * (size_t)nbt + 4;
    if ((rc = c->params.reserve(sizeof(BinDesc) + rows.size() * sizeof(Row) + 64)) || (rc = c->out.reserve(nout)) ||
        (rc = c->queue.reserve(8)))
        return rc;
    unsigned char* dB = c->params.ptr;
    Row* dRows = (Row*)(c->params.ptr + ((sizeof(BinDesc) + 15) / 16) * 16);
    hipError_t e = hipMemcpyAsync(dB, &B, sizeof(B), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && !rows.empty())
        e = hipMemcpyAsync(dRows, rows.data(), rows.size() * sizeof(Row), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->out.ptr, 0, nout * sizeof(unsigned long long), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->queue.ptr, 0, 8 * sizeof(int32_t), c->stream);
    if (e != hipSuccess) return hip_fail(e, "parameter upload");

    PairArgs a{};
    a.pw = c->pw.ptr;
    a.ws = c->ws.ptr;
    a.pl = c->pl.ptr;
    a.flags = c->flags.ptr;
    a.cell_start = c->cell_start.ptr;
    a.row_chunk = c->chunk_off.ptr;
    a.rows = dRows;
    a.bins = (const BinDesc*)dB;
    a.queue = c->queue.ptr;
    a.out_cnt = c->out.ptr;
    a.out_w = c->out.ptr + nbt;
    a.out_stats = c->out.ptr + 3 * nbt;
    a.g = c->g;
    a.g.m = mst;
    a.nrows = (int)rows.size();
    a.n_chunks = c->n_chunks;
    a.reps = reps;
    a.mirror = (rsd && !c->has_buf) ? 1 : 0;
    const int occ = best_waves / (wg / kWave);
    int nblocks = c->n_cu * occ;
    if (nblocks > c->n_chunks) nblocks = std::max(c->n_chunks, 1);
    (void)hipEventRecord(c->ev[2], c->stream);
    if (c->n_chunks > 0) {
        e = launch_pair2(a, nb, nm, fp64, rsd, weighted, p->log_bins != 0, nblocks, wg, c->stream);
        if (e != hipSuccess) return hip_fail(e, "pair kernel launch");
    }
    (void)hipEventRecord(c->ev[3], c->stream);
    std::vector<unsigned long long> h_out(nout);
    e = hipMemcpyAsync(h_out.data(), c->out.ptr, nout * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "pair kernel");

    combine2(h_out.data(), nb, nm, B, weighted, c->wsum_d, c->wsum_r, cnt, wsum, n2, n3, st);
    if (st) {
        float ms_grid = 0, ms_count = 0;
        (void)hipEventElapsedTime(&ms_grid, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms_count, c->ev[2], c->ev[3]);
        st->ms_grid = ms_grid;
        st->ms_count = ms_count;
        st->ms_kernel_avg = ms_count;
        st->cell_div = mst;
    }
    return GS_OK;
}

// ====================================================================== one-shot

static gs_ctx* g_ctx = nullptr;

static int oneshot_ctx(gs_ctx** c)
{
    if (!g_ctx) {
        int rc = gs_create(&g_ctx, 0);
        if (rc) return rc;
    }
    *c = g_ctx;
    return GS_OK;
}

extern "C" int gs_pcf2(const double* xyz, const double* w, const int32_t* buf, int64_t n_data, int64_t n_rand,
                       const gs_params* p, uint64_t* cnt, double* n2, double* n3, gs_stats* st)
{
    gs_ctx* c;
    int rc = oneshot_ctx(&c);
    if (rc) return rc;
    if ((rc = gs_set_catalogue(c, xyz, w, buf, n_data, n_rand))) return rc;
    return gs_count_2pcf(c, p, cnt, nullptr, n2, n3, st);
}

extern "C" int gs_pcf3(const double* xyz, const double* w, const int32_t* buf, int64_t n_data, int64_t n_rand,
                       const gs_params* p, uint64_t* cnt, double* n2, double* n3, gs_stats* st)
{
    gs_ctx* c;
    int rc = oneshot_ctx(&c);
    if (rc) return rc;
    if ((rc = gs_set_catalogue(c, xyz, w, buf, n_data, n_rand))) return rc;
    return gs_count_3pcf(c, p, cnt, nullptr, n2, n3, st);
}

extern "C" int gs_pcf4(const double* xyz, const double* w, const int32_t* buf, int64_t n_data, int64_t n_rand,
                       const gs_params* p, uint64_t* cnt, double* n2, double* n3, double* disc_n, double* disc_r,
                       gs_stats* st)
{
    gs_ctx* c;
    int rc = oneshot_ctx(&c);
    if (rc) return rc;
    if ((rc = gs_set_catalogue(c, xyz, w, buf, n_data, n_rand))) return rc;
    return gs_count_4pcf(c, p, cnt, n2, n3, disc_n, disc_r, st);
}

// Internal accessors for the other translation units (3PCF/4PCF/adjoint).
namespace gs {
CtxView view(gs_ctx* c)
{
    CtxView v;
    v.stream = c->stream;
    v.ev = c->ev;
    v.n_cu = c->n_cu;
    v.n_data = c->n_data;
    v.n_rand = c->n_rand;
    v.has_w = c->has_w;
    v.has_buf = c->has_buf;
    v.wsum_d = c->wsum_d;
    v.wsum_r = c->wsum_r;
    v.wmax = c->wmax;
    v.wmax_d = c->wmax_d;
    v.wmax_r = c->wmax_r;
    for (int k = 0; k < 6; ++k) v.bbox[k] = c->bbox[k];
    v.g = c->g;
    v.pw = c->pw.ptr;
    v.ws = c->ws.ptr;
    v.flags = c->flags.ptr;
    v.orig = c->orig.ptr;
    v.cell_start = c->cell_start.ptr;
    v.pl = c->pl.ptr;
    v.row_chunk = c->chunk_off.ptr;
    v.w = c->has_w ? c->w.ptr : nullptr;
    v.n_chunks = c->n_chunks;
    v.extra = &c->extra;
    v.batch = &c->batch;
    v.xyz = c->xyz.ptr;
    v.buf = c->has_buf ? c->buf.ptr : nullptr;
    return v;
}
int ctx_build_grid(gs_ctx* c, double rmax, int m, int sx, double* h_out)
{
    double h = choose_h(c, rmax, m, sx);
    *h_out = h;
    return build_grid(c, h, sx);
}
double ctx_xmax(gs_ctx* c)
{
    double xmax = 0;
    for (int k = 0; k < 6; ++k) xmax = std::max(xmax, std::fabs(c->bbox[k]));
    return xmax;
}
int ctx_set_device(gs_ctx* c) { return set_device(c); }
int ctx_upload(gs_ctx* c, const double* xyz, const double* w, const int32_t* buf, int64_t n)
{
    int rc;
    c->n_data = n;
    c->n_rand = 0;
    c->has_w = w != nullptr;
    c->has_buf = buf != nullptr;
    c->grid_valid = false;
    if ((rc = c->xyz.reserve(3 * n))) return rc;
    hipError_t e = hipMemcpyAsync(c->xyz.ptr, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && w) {
        if ((rc = c->w.reserve(n))) return rc;
        e = hipMemcpyAsync(c->w.ptr, w, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    }
    if (e == hipSuccess && buf) {
        if ((rc = c->buf.reserve(n))) return rc;
        e = hipMemcpyAsync(c->buf.ptr, buf, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    }
    if (e != hipSuccess) return hip_fail(e, "catalogue upload");
    if ((rc = host_weight_stats(c, w))) return rc;
    return finish_catalogue(c);
}
}  // namespace gs

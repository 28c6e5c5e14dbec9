// decompose_api.cpp -- C ABI of the GPU domain decomposition (gs_decompose).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "gs_api_internal.h"
#include "gs_kernels.h"

using namespace gs;

namespace {

// Device buffers of one call, released on every return path.
struct Scratch {
    DevBuf<float> xyz, box;
    DevBuf<int32_t> reg, keys, flag, pos, scan, idx, buf;
    ~Scratch()
    {
        xyz.release(); box.release(); reg.release(); keys.release(); flag.release(); pos.release();
        scan.release(); idx.release(); buf.release();
    }
};

float key_to_float(int32_t k)
{
    const int32_t i = k >= 0 ? k : (k ^ 0x7fffffff);
    float f;
    memcpy(&f, &i, 4);
    return f;
}

// x - r and x + r rounded to single precision, as the tool's default reals are
float fsub(float x, float r)
{
    volatile float v = x - r;
    return v;
}
float fadd(float x, float r)
{
    volatile float v = x + r;
    return v;
}

}  // namespace

extern "C" int gs_decompose(gs_ctx* c, const float* xyz, const int32_t* region, int64_t n, float rmax, int nreg,
                            int64_t* off, int32_t* idx, int32_t* buf, int64_t cap, float* box)
{
    if (!c || !off || nreg < 1 || nreg > 1000000 || n < 0 || n > (int64_t)INT32_MAX - 1 ||
        ((!xyz || !region) && n > 0))
        return fail(GS_EINVAL, "bad decomposition arguments");
    if (!std::isfinite(rmax) || rmax < 0.0f) return fail(GS_EINVAL, "rmax must be finite and >= 0");
    if ((idx == nullptr) != (buf == nullptr)) return fail(GS_EINVAL, "idx and buf must both be given or both NULL");
    for (int64_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(xyz[i])) return fail(GS_EINVAL, "non-finite coordinate at point %lld", (long long)(i / 3));
    int rc = ctx_set_device(c);
    if (rc) return rc;
    CtxView v = view(c);
    hipStream_t st = v.stream;

    Scratch S;
    const int64_t nn = std::max<int64_t>(n, 1);
    if ((rc = S.xyz.reserve(3 * nn)) || (rc = S.reg.reserve(nn)) || (rc = S.keys.reserve((size_t)6 * nreg)) ||
        (rc = S.box.reserve(12)) || (rc = S.flag.reserve(nn + 1)) || (rc = S.pos.reserve(nn + 1)) ||
        (rc = S.scan.reserve(scan_scratch_size(n + 1))))
        return rc;
    hipError_t e = hipMemcpyAsync(S.xyz.ptr, xyz, (size_t)n * 12, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(S.reg.ptr, region, (size_t)n * 4, hipMemcpyHostToDevice, st);
    // min keys start at INT_MAX, max keys at INT_MIN
    std::vector<int32_t> k0((size_t)6 * nreg);
    for (int r = 0; r < nreg; ++r)
        for (int k = 0; k < 6; ++k) k0[(size_t)r * 6 + k] = k < 3 ? INT32_MAX : INT32_MIN;
    if (e == hipSuccess) e = hipMemcpyAsync(S.keys.ptr, k0.data(), k0.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_region_box(S.xyz.ptr, S.reg.ptr, n, nreg, S.keys.ptr, st);
    if (e == hipSuccess) e = hipMemcpyAsync(k0.data(), S.keys.ptr, k0.size() * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "region boxes");

    // boxes; a label with no point keeps minval/maxval's empty results, +huge / -huge
    std::vector<float> bx((size_t)6 * nreg);
    for (int r = 0; r < nreg; ++r) {
        const bool empty = k0[(size_t)r * 6] == INT32_MAX;
        for (int k = 0; k < 6; ++k)
            bx[(size_t)r * 6 + k] = empty ? (k < 3 ? FLT_MAX : -FLT_MAX) : key_to_float(k0[(size_t)r * 6 + k]);
    }
    if (box) memcpy(box, bx.data(), bx.size() * sizeof(float));

    // per region: flags, exclusive scan (pos[n] = member count), scatter
    std::vector<int32_t> h_idx, h_buf;
    if (idx && (rc = S.idx.reserve(std::max<int64_t>(cap, 1)))) return rc;
    if (idx && (rc = S.buf.reserve(std::max<int64_t>(cap, 1)))) return rc;
    off[0] = 0;
    if ((e = launch_fill_u32(S.flag.ptr + n, 1, 0u, st)) != hipSuccess) return hip_fail(e, "decomposition");
    for (int r = 0; r < nreg; ++r) {
        const float* b = &bx[(size_t)r * 6];
        float be[12];
        for (int k = 0; k < 6; ++k) be[k] = b[k];
        for (int k = 0; k < 3; ++k) {
            be[6 + k] = fsub(b[k], rmax);            // xmin2 = xmin - rmax (domain_decomposition.f90:48-53)
            be[9 + k] = fadd(b[3 + k], rmax);        // xmax2 = xmax + rmax
        }
        int32_t cnt = 0;
        e = hipMemcpyAsync(S.box.ptr, be, sizeof(be), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_region_flag(S.xyz.ptr, n, S.box.ptr + 6, S.flag.ptr, st);
        if (e == hipSuccess) e = exclusive_scan(S.flag.ptr, S.pos.ptr, n + 1, S.scan.ptr, st);
        if (e == hipSuccess) e = hipMemcpyAsync(&cnt, S.pos.ptr + n, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "region membership");
        off[r + 1] = off[r] + cnt;
        if (idx && off[r + 1] <= cap && cnt > 0) {
            // positions are region-local: scatter into the region's slice of the output
            e = launch_region_scatter(S.xyz.ptr, n, S.box.ptr, S.flag.ptr, S.pos.ptr, S.idx.ptr + off[r],
                                      S.buf.ptr + off[r], st);
            if (e != hipSuccess) return hip_fail(e, "region scatter");
            // the next region overwrites box/flag/pos: finish this one first
            if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "region scatter");
        }
    }
    if (idx) {
        if (off[nreg] > cap)
            return fail(GS_ELIMIT, "membership lists need %lld entries, cap is %lld", (long long)off[nreg],
                        (long long)cap);
        e = hipMemcpyAsync(idx, S.idx.ptr, (size_t)off[nreg] * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(buf, S.buf.ptr, (size_t)off[nreg] * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "membership download");
    }
    return GS_OK;
}

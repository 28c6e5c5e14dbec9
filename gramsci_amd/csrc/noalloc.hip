// noalloc.hip -- gs_count_2pcf_noalloc: the 2PCF count of a catalogue already in
// device memory, with every device buffer it needs (the sorted copy of the
// catalogue, the cell grid, scan scratch, kernel parameters, histogram
// accumulators, work counters, reduction partials) carved out of one
// file-scope device array.  The call allocates no device memory, for callers
// that cannot allocate (code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc).  The kernels are those of gs_count_2pcf.
//
// Limits: at most kNoAllocMaxPoints points and kNoAllocMaxCells grid cells
// (the cell edge is widened until the grid fits).  Each GPU that runs the
// library holds the array (about 54 MB) until the process ends.  Every call on
// a GPU uses that GPU's one array: calls that overlap on one GPU race.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "gs_api_internal.h"
#include "gs_kernels.h"

namespace gs {
namespace {

constexpr int64_t kNoAllocMaxPoints = (int64_t)1 << 19;
constexpr int64_t kNoAllocMaxCells = (int64_t)1 << 20;
constexpr int64_t kScanTmpInts = 16384;            // >= scan_scratch_size(kNoAllocMaxCells + 1)
constexpr int kRedBlocks = 256;                    // blocks of the bbox / weight reductions
constexpr int64_t kMaxHist = 3 * (int64_t)kMaxBins * kMaxBins;   // histogram entries [3][nb][nm]

constexpr size_t al(size_t b) { return (b + 255) / 256 * 256; }

// byte offsets of the parts of the array
constexpr size_t kOffPw = 0;
constexpr size_t kOffWs = kOffPw + al(kNoAllocMaxPoints * sizeof(double4));
constexpr size_t kOffPl = kOffWs + al(kNoAllocMaxPoints * sizeof(double));
constexpr size_t kOffFlags = kOffPl + al(kNoAllocMaxPoints * sizeof(float4));
constexpr size_t kOffOrig = kOffFlags + al(kNoAllocMaxPoints * 4);
constexpr size_t kOffCellOf = kOffOrig + al(kNoAllocMaxPoints * 4);
constexpr size_t kOffCellStart = kOffCellOf + al(kNoAllocMaxPoints * 4);
constexpr size_t kOffFill = kOffCellStart + al((kNoAllocMaxCells + 1) * 4);
constexpr size_t kOffChunkNc = kOffFill + al((kNoAllocMaxCells + 1) * 4);
constexpr size_t kOffChunkOff = kOffChunkNc + al((kNoAllocMaxCells + 1) * 4);
constexpr size_t kOffScan = kOffChunkOff + al((kNoAllocMaxCells + 1) * 4);
constexpr size_t kOffParams = kOffScan + al(kScanTmpInts * 4);
constexpr size_t kOffOut = kOffParams + al(sizeof(BinDesc) + 16 + kMaxRows * sizeof(Row));
constexpr size_t kOffQueue = kOffOut + al((3 * kMaxHist + 4) * 8);
constexpr size_t kOffPart = kOffQueue + al(8 * 4);
constexpr size_t kOffWpart = kOffPart + al(kRedBlocks * 6 * 8);
constexpr size_t kNoAllocBytes = kOffWpart + al((kRedBlocks * 4 + 2) * 8);

// The array.  In host code its name stands for the symbol, not for a device address: host code
// uses the address hipGetSymbolAddress returns for the current device (scratch_base below).
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_base[kMaxDevices] = {};

// The array's address on `device` (which must be current), looked up once per device.
int scratch_base(int device, unsigned char** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(GS_ENODEV, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_base[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_base[device] = (unsigned char*)ptr;
    }
    *out = g_base[device];
    return GS_OK;
}

}  // namespace
}  // namespace gs

using namespace gs;

extern "C" int gs_count_2pcf_noalloc(gs_ctx* c, const gs_params* p, const double* d_xyz, const double* d_w,
                                     const int32_t* d_buf, int64_t n_data, int64_t n_rand, uint64_t* cnt,
                                     double* wsum, double* n2, double* n3, gs_stats* st)
{
    if (!c) return fail(GS_EINVAL, "NULL context");
    int rc = check_params(p, 2);
    if (rc) return rc;
    const int64_t n = n_data + n_rand;
    if (n_data < 0 || n_rand < 0 || (!d_xyz && n > 0)) return fail(GS_EINVAL, "bad catalogue (NULL or sizes)");
    if (n > kNoAllocMaxPoints)
        return fail(GS_ELIMIT, "gs_count_2pcf_noalloc holds at most %lld points (got %lld)",
                    (long long)kNoAllocMaxPoints, (long long)n);
    if ((rc = ctx_set_device(c))) return rc;
    CtxView v = view(c);
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    unsigned char* base;
    if ((rc = scratch_base(device, &base))) return rc;
    const hipStream_t s = v.stream;
    double* part = (double*)(base + kOffPart);
    unsigned long long* wpart = (unsigned long long*)(base + kOffWpart);

    // ---- weights: check, largest weight per class, exact fixed-point class sums
    long double Wd = (long double)n_data, Wr = (long double)n_rand;
    double wmax_d = 1.0, wmax_r = 1.0;
    if (d_w && n > 0) {
        unsigned long long* bad = wpart + kRedBlocks * 4;
        double pmax[2 * kRedBlocks];
        unsigned long long first_bad = 0;
        e = launch_fill_u32(bad, 2, 0xffffffffu, s);
        if (e == hipSuccess) e = launch_wcheck(d_w, n, n_data, part, bad, kRedBlocks, s);
        if (e == hipSuccess) e = hipMemcpyAsync(pmax, part, sizeof(pmax), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&first_bad, bad, sizeof(first_bad), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "weight check");
        if (first_bad != ~0ull) return fail(GS_EINVAL, "weights must be finite and >= 0 (w[%llu])", first_bad);
        double wm[2] = {0.0, 0.0};
        for (int b = 0; b < kRedBlocks; ++b)
            for (int q = 0; q < 2; ++q) wm[q] = std::max(wm[q], pmax[2 * b + q]);
        const int k = wsum_scale(std::max(wm[0], wm[1]));
        unsigned long long ps[4 * kRedBlocks];
        e = launch_wsum(d_w, n, n_data, k, wpart, kRedBlocks, s);
        if (e == hipSuccess) e = hipMemcpyAsync(ps, wpart, sizeof(ps), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "weight sums");
        unsigned __int128 sd = 0, sr = 0;
        for (int b = 0; b < kRedBlocks; ++b) {
            sd += ((unsigned __int128)ps[4 * b + 1] << 64) | ps[4 * b];
            sr += ((unsigned __int128)ps[4 * b + 3] << 64) | ps[4 * b + 2];
        }
        Wd = std::ldexp((long double)sd, -k);
        Wr = std::ldexp((long double)sr, -k);
        wmax_d = wm[0] > 0 ? wm[0] : 1.0;
        wmax_r = wm[1] > 0 ? wm[1] : 1.0;
    }

    // ---- bounding box
    double bbox[6] = {0, 0, 0, 0, 0, 0};
    if (n > 0) {
        double pb[6 * kRedBlocks];
        e = launch_bbox(d_xyz, n, part, kRedBlocks, s);
        if (e == hipSuccess) e = hipMemcpyAsync(pb, part, sizeof(pb), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "bbox");
        for (int k = 0; k < 3; ++k) { bbox[k] = 1e308; bbox[3 + k] = -1e308; }
        for (int b = 0; b < kRedBlocks; ++b)
            for (int k = 0; k < 3; ++k) {
                bbox[k] = std::min(bbox[k], pb[b * 6 + k]);
                bbox[3 + k] = std::max(bbox[3 + k], pb[b * 6 + 3 + k]);
            }
        for (int k = 0; k < 6; ++k)
            if (!std::isfinite(bbox[k])) return fail(GS_EINVAL, "catalogue has non-finite coordinates");
    }

    // ---- grid (as gs_count_2pcf, with the cell edge widened to fit kNoAllocMaxCells)
    const int nb = p->nbins, nm = p->nmu;
    const bool rsd = nm >= 2, fp64 = p->precision == GS_FP64, weighted = d_w != nullptr;
    const int nbt = 3 * nb * nm;
    const int m = p->cell_div > 0 ? p->cell_div : pair_cell_div(n, bbox, p->rmax);
    const int sx = pair_cell_sx();
    const double h = choose_h_bbox(bbox, p->rmax, m, kNoAllocMaxCells, sx);
    const int mst = (int)std::ceil(p->rmax / h - 1e-12);
    std::vector<Row> rows = half_stencil(h, h / sx, p->rmax, mst);
    if ((int)rows.size() > kMaxRows) return fail(GS_ELIMIT, "stencil too large");
    GridBufs gb{(int32_t*)(base + kOffCellOf), (int32_t*)(base + kOffCellStart), (int32_t*)(base + kOffFill),
                (int32_t*)(base + kOffChunkNc), (int32_t*)(base + kOffChunkOff), (int32_t*)(base + kOffScan),
                (double4*)(base + kOffPw), (double*)(base + kOffWs), (float4*)(base + kOffPl),
                (uint32_t*)(base + kOffFlags), (int32_t*)(base + kOffOrig)};
    (void)hipEventRecord(v.ev[0], s);
    GridDesc g{};
    int32_t n_chunks = 0;
    if (n > 0) {
        int64_t ncell = 1;
        for (int k = 0; k < 3; ++k) ncell *= (int64_t)std::floor((bbox[3 + k] - bbox[k]) / (k == 0 ? h / sx : h)) + 1;
        if (scan_scratch_size(ncell + 1) > kScanTmpInts) return fail(GS_ELIMIT, "scan scratch too small");
        if ((rc = grid_build_raw(d_xyz, d_w, d_buf, n, n_data, bbox, h, sx, gb, kNoAllocMaxCells, &g, &n_chunks, s)))
            return rc;
    }
    (void)hipEventRecord(v.ev[1], s);

    // ---- bins, launch shape, parameters, cleared accumulators
    double xmax = 0;
    for (int k = 0; k < 6; ++k) xmax = std::max(xmax, std::fabs(bbox[k]));
    BinDesc B;
    if ((rc = fill_bins(*p, (mst + 1) * h, xmax, wmax_d, wmax_r, B))) return rc;
    int reps = 1, wg = kWG, best_waves = 0;
    if ((rc = pair2_shape(nb, nm, fp64, weighted, p->log_bins != 0, &reps, &wg, &best_waves))) return rc;
    unsigned char* dB = base + kOffParams;
    Row* dRows = (Row*)(dB + ((sizeof(BinDesc) + 15) / 16) * 16);
    unsigned long long* out = (unsigned long long*)(base + kOffOut);
    int32_t* queue = (int32_t*)(base + kOffQueue);
    const size_t nout = (size_t)3 * nbt + 4;
    e = hipMemcpyAsync(dB, &B, sizeof(B), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !rows.empty())
        e = hipMemcpyAsync(dRows, rows.data(), rows.size() * sizeof(Row), hipMemcpyHostToDevice, s);
    // the array keeps the previous call's contents: clear what starts at zero (with fill kernels)
    if (e == hipSuccess) e = launch_fill_u32(out, (int64_t)nout * 2, 0u, s);
    if (e == hipSuccess) e = launch_fill_u32(queue, 8, 0u, s);
    if (e != hipSuccess) return hip_fail(e, "parameter upload");

    PairArgs a{};
    a.pw = gb.pw;
    a.ws = gb.ws;
    a.pl = gb.pl;
    a.flags = gb.flags;
    a.cell_start = gb.cell_start;
    a.row_chunk = gb.chunk_off;
    a.rows = dRows;
    a.bins = (const BinDesc*)dB;
    a.queue = queue;
    a.out_cnt = out;
    a.out_w = out + nbt;
    a.out_stats = out + 3 * nbt;
    a.g = g;
    a.g.m = mst;
    a.nrows = (int)rows.size();
    a.n_chunks = n_chunks;
    a.reps = reps;
    a.mirror = (rsd && !d_buf) ? 1 : 0;
    const int occ = best_waves / (wg / kWave);
    int nblocks = v.n_cu * occ;
    if (nblocks > n_chunks) nblocks = std::max(n_chunks, 1);
    (void)hipEventRecord(v.ev[2], s);
    if (n_chunks > 0) {
        e = launch_pair2(a, nb, nm, fp64, rsd, weighted, p->log_bins != 0, nblocks, wg, s);
        if (e != hipSuccess) return hip_fail(e, "pair kernel launch");
    }
    (void)hipEventRecord(v.ev[3], s);
    std::vector<unsigned long long> h_out(nout);
    e = hipMemcpyAsync(h_out.data(), out, nout * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "pair kernel");
    combine2(h_out.data(), nb, nm, B, weighted, Wd, Wr, cnt, wsum, n2, n3, st);
    if (st) {
        float ms_grid = 0, ms_count = 0;
        (void)hipEventElapsedTime(&ms_grid, v.ev[0], v.ev[1]);
        (void)hipEventElapsedTime(&ms_count, v.ev[2], v.ev[3]);
        st->ms_grid = ms_grid;
        st->ms_count = ms_count;
        st->ms_kernel_avg = ms_count;
        st->cell_div = mst;
    }
    return GS_OK;
}

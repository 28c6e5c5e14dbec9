// grid.hip -- GPU cell grid: bounding box, counting sort of points by cell,
// cell start offsets.  Replaces the kd-tree build (kdtree2.F90:611-1012) and
// its r-ball queries: a pair kernel reads each neighbour cell row as one
// contiguous range of the sorted points.
#include "gs_kernels.h"

namespace gs {

// ---------------------------------------------------------------- bounding box
// Per-block min/max of the three coordinates; the host reduces the partials.
__global__ __launch_bounds__(256) void k_bbox(const double* __restrict__ xyz, int64_t n,
                                              double* __restrict__ part /*[gridDim][6]*/)
{
    __shared__ double red[6][256];
    double mn[3] = {1e308, 1e308, 1e308}, mx[3] = {-1e308, -1e308, -1e308};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            double v = xyz[3 * i + k];
            mn[k] = fmin(mn[k], v);
            mx[k] = fmax(mx[k], v);
        }
    }
    for (int k = 0; k < 3; ++k) { red[k][threadIdx.x] = mn[k]; red[3 + k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------- weights
// per-block maxima of the data weights (i < n_data) and of the random weights: part_max[2 * block + class]
__global__ __launch_bounds__(256) void k_wcheck(const double* __restrict__ w, int64_t n, int64_t n_data,
                                                double* __restrict__ part_max, unsigned long long* __restrict__ bad)
{
    __shared__ double red[2][256];
    double mx[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = w[i];
        if (!(v >= 0.0) || !isfinite(v)) atomicMin(bad, (unsigned long long)i);
        else if (i < n_data) mx[0] = fmax(mx[0], v);
        else mx[1] = fmax(mx[1], v);
    }
    red[0][threadIdx.x] = mx[0];
    red[1][threadIdx.x] = mx[1];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 2; ++q) red[q][threadIdx.x] = fmax(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 2) part_max[2 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ void add128(unsigned long long& lo, unsigned long long& hi, unsigned long long l,
                                       unsigned long long h)
{
    lo += l;
    hi += h + (lo < l ? 1ull : 0ull);
}

__global__ __launch_bounds__(256) void k_wsum(const double* __restrict__ w, int64_t n, int64_t n_data, int k,
                                              unsigned long long* __restrict__ part)
{
    __shared__ unsigned long long red[4][256];
    unsigned long long a[4] = {0, 0, 0, 0};     // data lo, hi, randoms lo, hi
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t lo, hi;
        wsum_fixed(w[i], k, lo, hi);
        const int o = i < n_data ? 0 : 2;
        add128(a[o], a[o + 1], lo, hi);
    }
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const int t = threadIdx.x;
            unsigned long long l0 = red[0][t], h0 = red[1][t], l1 = red[2][t], h1 = red[3][t];
            add128(l0, h0, red[0][t + s], red[1][t + s]);
            add128(l1, h1, red[2][t + s], red[3][t + s]);
            red[0][t] = l0; red[1][t] = h0; red[2][t] = l1; red[3][t] = h1;
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------- cell of point
__device__ __forceinline__ int cell_coord(double v, double lo, double inv_h, int dim)
{
    int c = (int)floor((v - lo) * inv_h);
    return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

__global__ __launch_bounds__(256) void k_cell_count(const double* __restrict__ xyz, int64_t n, GridDesc g,
                                                    int32_t* __restrict__ cell_of, int32_t* __restrict__ count)
{
    const double inv_h = 1.0 / g.h, inv_hx = 1.0 / g.hx;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int cx = cell_coord(xyz[3 * i + 0], g.lo[0], inv_hx, g.dim[0]);
        int cy = cell_coord(xyz[3 * i + 1], g.lo[1], inv_h, g.dim[1]);
        int cz = cell_coord(xyz[3 * i + 2], g.lo[2], inv_h, g.dim[2]);
        int c = (cz * g.dim[1] + cy) * g.dim[0] + cx;
        cell_of[i] = c;
        atomicAdd(&count[c], 1);
    }
}

// ---------------------------------------------------------------- exclusive scan
// Three-phase scan of int32 counts: each block scans 1024 entries in LDS and
// emits its total; the totals are scanned recursively; offsets are added back.
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(256) void k_scan_block(const int32_t* __restrict__ in, int64_t n,
                                                    int32_t* __restrict__ out, int32_t* __restrict__ sums)
{
    __shared__ int32_t s[kScanBlock];
    __shared__ int32_t wsum[4];
    const int64_t base = (int64_t)blockIdx.x * kScanBlock;
    const int t = threadIdx.x;
    int v[4];
    int run = 0;
    for (int k = 0; k < 4; ++k) {
        int64_t i = base + t * 4 + k;
        v[k] = i < n ? in[i] : 0;
        run += v[k];
    }
    // inclusive scan of per-thread totals across the block (wave64 shuffles, then across 4 waves)
    int x = run;
    const int lane = t & 63, w = t >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int wpre = 0;
    for (int k = 0; k < w; ++k) wpre += wsum[k];
    int excl = wpre + x - run;
    for (int k = 0; k < 4; ++k) {
        s[t * 4 + k] = excl;
        excl += v[k];
    }
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
        int64_t i = base + t * 4 + k;
        if (i < n) out[i] = s[t * 4 + k];
    }
    if (t == 255) sums[blockIdx.x] = excl;
}

__global__ __launch_bounds__(256) void k_scan_add(int32_t* __restrict__ out, int64_t n, const int32_t* __restrict__ offs)
{
    const int64_t base = (int64_t)blockIdx.x * kScanBlock;
    const int32_t o = offs[blockIdx.x];
    for (int k = threadIdx.x; k < kScanBlock; k += blockDim.x) {
        int64_t i = base + k;
        if (i < n) out[i] += o;
    }
}

hipError_t exclusive_scan(const int32_t* in, int32_t* out, int64_t n, int32_t* scratch, hipStream_t st)
{
    // scratch must hold scan_scratch_size(n) int32
    if (n <= 0) return hipSuccess;
    int64_t nb = (n + kScanBlock - 1) / kScanBlock;
    int32_t* sums = scratch;
    int32_t* sums_scanned = scratch + nb;
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(256), 0, st, in, n, out, sums);
    if (nb > 1) {
        hipError_t e = exclusive_scan(sums, sums_scanned, nb, scratch + 2 * nb, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(256), 0, st, out, n, sums_scanned);
    }
    return hipGetLastError();
}

int64_t scan_scratch_size(int64_t n)
{
    int64_t tot = 0;
    while (n > 1) {
        int64_t nb = (n + kScanBlock - 1) / kScanBlock;
        tot += 2 * nb;
        n = nb;
    }
    return tot + 2;
}

// ---------------------------------------------------------------- scatter
// Writes the points in cell order, a cell's data points before its randoms
// (consecutive neighbours then mostly share a class).  Order inside a cell otherwise follows the atomic
// slot order; every count and fixed-point sum computed from the sorted arrays
// is independent of it.
__global__ __launch_bounds__(256) void k_scatter(const double* __restrict__ xyz, const double* __restrict__ w,
                                                 const int32_t* __restrict__ buf, int64_t n, int64_t n_data, GridDesc g,
                                                 const int32_t* __restrict__ cell_of, const int32_t* __restrict__ start,
                                                 int32_t* __restrict__ fill, double4* __restrict__ pw,
                                                 double* __restrict__ ws, float4* __restrict__ pl,
                                                 uint32_t* __restrict__ flags, int32_t* __restrict__ orig, int pass)
{
    const int nx = g.dim[0], ny = g.dim[1];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if ((i >= n_data) != (pass == 1)) continue;    // pass 0: data points, pass 1: randoms
        const int c = cell_of[i];
        const int pos = start[c] + atomicAdd(&fill[c], 1);
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const double wi = w ? w[i] : 1.0;
        pw[pos] = make_double4(x, y, z, wi);
        ws[pos] = wi;
        const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        // cell-local FP32 coordinates; .w carries the cell's x index (exact in FP32)
        pl[pos] = make_float4((float)(x - (g.lo[0] + cx * g.hx)), (float)(y - (g.lo[1] + cy * g.h)),
                              (float)(z - (g.lo[2] + cz * g.h)), (float)cx);
        const uint32_t f = (i >= n_data ? 1u : 0u) | ((buf && buf[i] == 1) ? 2u : 0u);
        flags[pos] = f;
        orig[pos] = (int32_t)i;
    }
}

// Work chunks of the pair kernel: kChunk consecutive points of one x-row of
// cells (fixed cy, cz).  nch[r] = chunks of row r; nch[nrow] = 0.
__global__ __launch_bounds__(256) void k_row_chunks(const int32_t* __restrict__ start, int nx, int nrow,
                                                    int32_t* __restrict__ nch)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r <= nrow; r += gridDim.x * blockDim.x) {
        int v = 0;
        if (r < nrow) v = (start[(r + 1) * nx] - start[r * nx] + kChunk - 1) / kChunk;
        nch[r] = v;
    }
}

hipError_t launch_row_chunks(const int32_t* start, int nx, int nrow, int32_t* nch, hipStream_t st)
{
    int nblk = (nrow + 256) / 256;
    if (nblk > 65535) nblk = 65535;
    hipLaunchKernelGGL(k_row_chunks, dim3(nblk), dim3(256), 0, st, start, nx, nrow, nch);
    return hipGetLastError();
}

hipError_t launch_bbox(const double* xyz, int64_t n, double* part, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_bbox, dim3(nblk), dim3(256), 0, st, xyz, n, part);
    return hipGetLastError();
}

hipError_t launch_wcheck(const double* w, int64_t n, int64_t n_data, double* part_max, unsigned long long* bad, int nblk,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_wcheck, dim3(nblk), dim3(256), 0, st, w, n, n_data, part_max, bad);
    return hipGetLastError();
}

hipError_t launch_wsum(const double* w, int64_t n, int64_t n_data, int k, unsigned long long* part, int nblk,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_wsum, dim3(nblk), dim3(256), 0, st, w, n, n_data, k, part);
    return hipGetLastError();
}

hipError_t launch_cell_count(const double* xyz, int64_t n, const GridDesc& g, int32_t* cell_of, int32_t* count,
                             hipStream_t st)
{
    int nblk = (int)((n + 255) / 256);
    if (nblk > 65535) nblk = 65535;
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL(k_cell_count, dim3(nblk), dim3(256), 0, st, xyz, n, g, cell_of, count);
    return hipGetLastError();
}

hipError_t launch_scatter(const double* xyz, const double* w, const int32_t* buf, int64_t n, int64_t n_data,
                          const GridDesc& g, const int32_t* cell_of, const int32_t* start, int32_t* fill,
                          double4* pw, double* ws, float4* pl, uint32_t* flags, int32_t* orig, hipStream_t st)
{
    int nblk = (int)((n + 255) / 256);
    if (nblk > 65535) nblk = 65535;
    if (nblk < 1) nblk = 1;
    // two passes on one stream: a cell's data points take its first slots, its randoms the rest
    for (int pass = 0; pass < 2; ++pass)
        hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(256), 0, st, xyz, w, buf, n, n_data, g, cell_of, start, fill,
                           pw, ws, pl, flags, orig, pass);
    return hipGetLastError();
}

// ---------------------------------------------------------------- batched regions
__device__ __forceinline__ int region_of(const int32_t* off, int nreg, int64_t i)
{
    int lo = 0, hi = nreg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_cell_count_b(const double* __restrict__ xyz, int64_t n,
                                                      const int32_t* __restrict__ off, int nreg,
                                                      const RegionDesc* __restrict__ reg, double h, double hx,
                                                      int32_t* __restrict__ cell_of, int32_t* __restrict__ count)
{
    const double inv_h = 1.0 / h, inv_hx = 1.0 / hx;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const RegionDesc R = reg[region_of(off, nreg, i)];
        int cx = cell_coord(xyz[3 * i + 0], R.lo[0], inv_hx, R.dim[0]);
        int cy = cell_coord(xyz[3 * i + 1], R.lo[1], inv_h, R.dim[1]);
        int cz = cell_coord(xyz[3 * i + 2], R.lo[2], inv_h, R.dim[2]);
        int c = R.cell_base + (cz * R.dim[1] + cy) * R.dim[0] + cx;
        cell_of[i] = c;
        atomicAdd(&count[c], 1);
    }
}

__global__ __launch_bounds__(256) void k_scatter_b(const double* __restrict__ xyz, const double* __restrict__ w,
                                                   const int32_t* __restrict__ buf, int64_t n,
                                                   const int32_t* __restrict__ off, int nreg,
                                                   const RegionDesc* __restrict__ reg, double h, double hx,
                                                   const int32_t* __restrict__ cell_of,
                                                   const int32_t* __restrict__ start, int32_t* __restrict__ fill,
                                                   double4* __restrict__ pw, double* __restrict__ ws,
                                                   float4* __restrict__ pl, uint32_t* __restrict__ flags,
                                                   int32_t* __restrict__ orig, int pass)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = region_of(off, nreg, i);
        const RegionDesc R = reg[r];
        if (((i - R.pt_base) >= R.n_data) != (pass == 1)) continue;   // pass 0: data points, pass 1: randoms
        const int c = cell_of[i];
        const int pos = start[c] + atomicAdd(&fill[c], 1);
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const double wi = w ? w[i] : 1.0;
        pw[pos] = make_double4(x, y, z, wi);
        ws[pos] = wi;
        const int lc = c - R.cell_base;
        const int cx = lc % R.dim[0], cy = (lc / R.dim[0]) % R.dim[1], cz = lc / (R.dim[0] * R.dim[1]);
        pl[pos] = make_float4((float)(x - (R.lo[0] + cx * hx)), (float)(y - (R.lo[1] + cy * h)),
                              (float)(z - (R.lo[2] + cz * h)), (float)cx);
        const uint32_t f = ((i - R.pt_base) >= R.n_data ? 1u : 0u) | ((buf && buf[i] == 1) ? 2u : 0u);
        flags[pos] = f;
        orig[pos] = (int32_t)i;
    }
}

// chunks of each global x-row; nch[total_rows] = 0
__global__ __launch_bounds__(256) void k_row_chunks_b(const int32_t* __restrict__ start,
                                                      const RegionDesc* __restrict__ reg,
                                                      const int32_t* __restrict__ row_base, int nreg, int nrow_tot,
                                                      int32_t* __restrict__ nch)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r <= nrow_tot; r += gridDim.x * blockDim.x) {
        int v = 0;
        if (r < nrow_tot) {
            const RegionDesc R = reg[region_of(row_base, nreg, r)];
            const int lr = r - R.row_base;
            const int c0 = R.cell_base + lr * R.dim[0];
            v = (start[c0 + R.dim[0]] - start[c0] + kChunk - 1) / kChunk;
        }
        nch[r] = v;
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static int blocks_for(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 65535) b = 65535;
    return (int)(b < 1 ? 1 : b);
}

hipError_t launch_fill_u32(void* p, int64_t n, uint32_t v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n)), dim3(256), 0, st, (uint32_t*)p, n, v);
    return hipGetLastError();
}

hipError_t launch_cell_count_b(const double* xyz, int64_t n, const int32_t* off, int nreg, const RegionDesc* reg,
                               double h, double hx, int32_t* cell_of, int32_t* count, hipStream_t st)
{
    hipLaunchKernelGGL(k_cell_count_b, dim3(blocks_for(n)), dim3(256), 0, st, xyz, n, off, nreg, reg, h, hx, cell_of,
                       count);
    return hipGetLastError();
}

hipError_t launch_scatter_b(const double* xyz, const double* w, const int32_t* buf, int64_t n, const int32_t* off,
                            int nreg, const RegionDesc* reg, double h, double hx, const int32_t* cell_of,
                            const int32_t* start,
                            int32_t* fill, double4* pw, double* ws, float4* pl, uint32_t* flags, int32_t* orig,
                            hipStream_t st)
{
    for (int pass = 0; pass < 2; ++pass)
        hipLaunchKernelGGL(k_scatter_b, dim3(blocks_for(n)), dim3(256), 0, st, xyz, w, buf, n, off, nreg, reg, h, hx,
                           cell_of, start, fill, pw, ws, pl, flags, orig, pass);
    return hipGetLastError();
}

hipError_t launch_row_chunks_b(const int32_t* start, const RegionDesc* reg, const int32_t* row_base, int nreg,
                               int nrow_tot, int32_t* nch, hipStream_t st)
{
    hipLaunchKernelGGL(k_row_chunks_b, dim3(blocks_for(nrow_tot + 1)), dim3(256), 0, st, start, reg, row_base, nreg,
                       nrow_tot, nch);
    return hipGetLastError();
}

}  // namespace gs


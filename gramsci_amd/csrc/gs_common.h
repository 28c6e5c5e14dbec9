// gs_common.h -- structures shared by the host side (api.cpp) and the HIP
// kernels of the N-point correlation engine.  gfx950 (MI355X) only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gs {

constexpr int kMaxBins = 127;   // int8 bin storage in the reference (node_module.F90:7)
constexpr int kWG = 256;        // work-group size of the counting kernels
constexpr int kWave = 64;       // CDNA wavefront
constexpr int kChunk = 64;      // points per work item of the pair kernel (one per lane)

// Cell grid over the catalogue's bounding box.  Points are stored sorted by
// cell id = (cz*ny + cy)*nx + cx (x fastest), so one x-row of cells is one
// contiguous range of points.  Cells have edge h in y and z and hx = h / sx in
// x (sx a power of two): a row cut finely in x lets the pair kernel take, per
// stencil row, just the x-range that can reach rmax from its chunk of apexes.
struct GridDesc {
    double lo[3];
    double h;                // y/z cell edge (a float with <= 11 significant bits, so k*h is exact in FP32)
    double hx;               // x cell edge, h / sx (exact)
    int32_t dim[3];          // dim[0] counts cells of edge hx
    int32_t ncell;
    int32_t m;               // y/z cells per rmax
    float h32;               // h as float (exact)
    float hx32;              // hx as float (exact)
};

// Binning constants.  t64 are the exact FP64 thresholds on s = |xj - xi|^2
// (see gs_bin_thresholds): a pair is kept iff t64[0] < s < t64[nb+1]; its
// 0-based bin is #{k in 1..nb : s >= t64[k]} (== nb means the reference's
// out-of-range bin nbins+1).
struct BinDesc {
    int32_t nb, nm, log_bins, rsd;
    double  t64[kMaxBins + 2];
    double  mu_scale;        // 0.5*nmu (gramsci.F90:145)
    // FP32 guard-band data
    float   ok_lo[kMaxBins + 1];   // bin b is certain if ok_lo[b] <= s32 < ok_hi[b]
    float   ok_hi[kMaxBins + 1];
    float   rej_lo, rej_hi;        // s32 <= rej_lo or s32 >= rej_hi: certainly not kept
    float   ua, ub;                // approximate bin: u = f(s)*ua + ub, f = sqrt (linear) or log2 (log)
    float   ms32;                  // mu_scale in FP32
    float   mu_ew, mu_em, mu_c;    // mu guard: dt = ms*(mu_ew/|sep| + mu_em/|mid| + mu_c)
    // weighted fixed point.  Pairs: each point's weight is scaled by 2^fx_sd (data) or 2^fx_sr
    // (randoms), so that every scaled weight of its class is below 2^26, and a pair contributes
    // q = rint(w_i' * w_j') < 2^52: the DD, DR and RR sums carry the scales 2^(2 fx_sd),
    // 2^(fx_sd + fx_sr) and 2^(2 fx_sr), each set by its own class's largest weights.
    // Triples: q = rint(w_i*w_j*w_k * 2^fx_shift3) with one scale.
    int32_t fx_sd, fx_sr, fx_shift3;
};

// One independent region of a batched run (a domain-decomposed sub-volume
// with its buffer shell, or a jackknife region): its own grid, its own points
// (data first, then randoms) and its own output histogram.
struct RegionDesc {
    double lo[3];
    int32_t dim[3];
    int32_t cell_base;       // first global cell id of the region
    int32_t row_base;        // first global x-row of the region
    int32_t pt_base;         // first point (input order and cell-sorted order)
    int32_t n_data;          // data points of the region (the rest are randoms)
    // the region's fixed-point scales, as its own launch uses them: triples fx_shift; pairs fx_shift
    // for data points and fx_shift_rand for randoms (BinDesc::fx_sd, fx_sr)
    int32_t fx_shift;
    int32_t fx_shift_rand;
};

// Exact, order-independent weight sums (the normalisations sum w over data
// and over randoms, gramsci.F90:1390-1391).  Each weight is rounded to the
// integer q = rint(w * 2^k) < 2^95, with k = wsum_scale(wmax); a sum of up to
// 2^31 such integers fits in 128 bits, so the host loops and the device
// reduction give identical sums in any order.  Relative to wmax, a weight
// loses at most 2^-96 in the rounding.
inline int wsum_scale(double wmax)
{
    int e = 0;
    if (wmax > 0) (void)frexp(wmax, &e);   // wmax < 2^e
    return 95 - e;
}

__host__ __device__ inline void wsum_fixed(double w, int k, uint64_t& lo, uint64_t& hi)
{
    const double x = ldexp(w, k);                   // < 2^95, exact
    const double h = floor(ldexp(x, -64));          // < 2^31
    lo = (uint64_t)rint(x - ldexp(h, 64));          // the difference is exact and < 2^64
    hi = (uint64_t)h;
}

// Stencil rows: cells (ax+dx, ay+dy, az+dz) with dx in [x0, x1] (x in cells of edge hx).
struct Row {
    int16_t dy, dz, x0, x1;
};
constexpr int kMaxRows = 2048;

}  // namespace gs

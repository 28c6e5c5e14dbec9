// pair2.hip -- persistent 2PCF pair kernel (binned (r, mu) DD/DR/RR counts).
//
// Replaces create_graph (gramsci.F90:288-348) + query_graph_2pcf (:352-399):
// no neighbour graph is materialised.  Each unordered pair {i, j} is visited
// once; it contributes the reference's two ordered pairs i->j and j->i (each
// only if its apex is not a buffer point, :363), with mu bins m and nmu-1-m
// (theta(j->i) = -theta(i->j) exactly).
//
// Work decomposition: a work item is a chunk of <= 256 points of one home cell
// A (one point per lane).  The item walks the half stencil of A as x-rows of
// cells; an x-row is one contiguous range of the cell-sorted points, staged
// through LDS in tiles of 256 and broadcast to every lane.  Work-groups are
// persistent and pull items from 8 queues, one per XCD, each covering a
// contiguous z-slab of the grid (so an XCD's L2 sees one slab), then steal.
//
// Precision variants:
//   FP64        s = |xj-xi|^2 in FP64 exactly as the reference; bins from the
//               exact thresholds; mu with IEEE sqrt/div.
//   FP32 guard  coordinates relative to A's cell origin in FP32; a pair is
//               binned directly only when s32 lies inside the bin's certain
//               interval [ok_lo, ok_hi) and mu is farther than a rigorous
//               error bound from a mu edge; otherwise it is re-evaluated with
//               the FP64 path.  Both variants produce identical counts.
//
// Histograms live in LDS: u32 counts replicated `reps` times (replica =
// lane % reps) to spread same-bin atomics over banks, and one 128-bit fixed
// point weighted sum per bin.  They are flushed to global u64 / 128-bit
// accumulators when the work-group ends (or before a u32 could overflow).
#include "gs_kernels.h"

namespace gs {

namespace {

struct Stats {
    unsigned long long cand, rechk, over;
};

__device__ __forceinline__ void lds_add128(unsigned long long* slot, unsigned long long q)
{
    unsigned long long old = atomicAdd(slot, q);
    if (old + q < old) atomicAdd(slot + 1, 1ull);
}

__device__ __forceinline__ void glb_add128(unsigned long long* slot, unsigned long long lo, unsigned long long hi)
{
    unsigned long long old = atomicAdd(slot, lo);
    if (old + lo < old) hi += 1ull;
    if (hi) atomicAdd(slot + 1, hi);
}

// Reference arithmetic in FP64, no contraction: s (kdtree2.F90:1689, gramsci.F90:1532).
__device__ __forceinline__ double exact_s(const double4& a, const double4& b)
{
    double dx = __dsub_rn(b.x, a.x), dy = __dsub_rn(b.y, a.y), dz = __dsub_rn(b.z, a.z);
    return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// 0-based bin of s: #{k in 1..nb : s >= t[k]}; nb means "out of range" (reference bin nbins+1).
__device__ __forceinline__ int exact_bin(double s, const double* t, int nb)
{
    int lo = 0, hi = nb;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (t[mid] <= s) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// mu bin of the ordered pair apex a -> neighbour b (gramsci.F90:320-328), 0-based,
// and of the reverse pair b -> a.  The reference stores floor(.)+1 in an int8
// before clamping; that wrap is reproduced.
__device__ __forceinline__ int mu_from_t(double t, int nm)
{
    int v = (int)floor(t) + 1;
    v = (int)(int8_t)v;
    v = v < 1 ? 1 : (v > nm ? nm : v);
    return v - 1;
}

__device__ __forceinline__ void exact_mu(const double4& a, const double4& b, double ms, int nm, int& m, int& mr)
{
    double m1 = __dmul_rn(0.5, __dadd_rn(a.x, b.x));
    double m2 = __dmul_rn(0.5, __dadd_rn(a.y, b.y));
    double m3 = __dmul_rn(0.5, __dadd_rn(a.z, b.z));
    double w1 = __dsub_rn(a.x, b.x), w2 = __dsub_rn(a.y, b.y), w3 = __dsub_rn(a.z, b.z);
    double th = __dadd_rn(__dadd_rn(__dmul_rn(m1, w1), __dmul_rn(m2, w2)), __dmul_rn(m3, w3));
    double mm = __dadd_rn(__dadd_rn(__dmul_rn(m1, m1), __dmul_rn(m2, m2)), __dmul_rn(m3, m3));
    double ww = __dadd_rn(__dadd_rn(__dmul_rn(w1, w1), __dmul_rn(w2, w2)), __dmul_rn(w3, w3));
    double den = __dmul_rn(__dsqrt_rn(mm), __dsqrt_rn(ww));
    th = __ddiv_rn(th, den);
    m = mu_from_t(__dmul_rn(__dadd_rn(th, 1.0), ms), nm);
    mr = mu_from_t(__dmul_rn(__dadd_rn(-th, 1.0), ms), nm);
}

__device__ __forceinline__ unsigned long long quantise(double p, int shift)
{
    return __double2ull_rn(ldexp(p, shift));
}

template <bool RSD, bool W>
__device__ __forceinline__ void accumulate(uint32_t* hc, unsigned long long* hw, int reps, int rep, int nb, int nm,
                                           int cls, int b, int m, int mr, int ci, int cj, double p, int shift)
{
    const int base = (cls * nb + b) * nm;
    if constexpr (!RSD) {
        const int c = ci + cj;
        if (c) {
            atomicAdd(&hc[base * reps + rep], (uint32_t)c);
            if constexpr (W) lds_add128(&hw[2 * base], quantise(p, shift) * (unsigned long long)c);
        }
    } else {
        unsigned long long q = 0;
        if constexpr (W) q = quantise(p, shift);
        if (ci) {
            atomicAdd(&hc[(base + m) * reps + rep], 1u);
            if constexpr (W) lds_add128(&hw[2 * (base + m)], q);
        }
        if (cj) {
            atomicAdd(&hc[(base + mr) * reps + rep], 1u);
            if constexpr (W) lds_add128(&hw[2 * (base + mr)], q);
        }
    }
}

__device__ void flush(uint32_t* hc, unsigned long long* hw, int nbt, int reps, bool weighted,
                      unsigned long long* out_cnt, unsigned long long* out_w)
{
    for (int k = threadIdx.x; k < nbt; k += blockDim.x) {
        unsigned long long s = 0;
        for (int r = 0; r < reps; ++r) {
            s += hc[k * reps + r];
            hc[k * reps + r] = 0;
        }
        if (s) atomicAdd(&out_cnt[k], s);
        if (weighted) {
            unsigned long long lo = hw[2 * k], hi = hw[2 * k + 1];
            if (lo | hi) glb_add128(&out_w[2 * k], lo, hi);
            hw[2 * k] = 0;
            hw[2 * k + 1] = 0;
        }
    }
}

__device__ __forceinline__ int find_cell(const int32_t* chunk_off, int ncell, int chunk)
{
    // largest A with chunk_off[A] <= chunk
    int lo = 0, hi = ncell - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (chunk_off[mid] <= chunk) lo = mid; else hi = mid - 1;
    }
    return lo;
}

}  // namespace

size_t pair2_lds_bytes(int nb, int nm, int reps, bool fp64, bool weighted)
{
    const size_t nbt = (size_t)3 * nb * nm;
    size_t b = 0;
    b += ((nbt * reps * 4 + 15) / 16) * 16;                  // counts
    if (weighted) b += nbt * 16;                              // 128-bit sums
    b += ((size_t)(kMaxBins + 1) * 8);                        // ok_lo/ok_hi as float2
    b += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;         // t64
    if (fp64) b += (size_t)kWG * (24 + 8 + 4 + 4);            // x, y, z, w, flags, pad
    else b += (size_t)kWG * (16 + 8);                         // float4, w
    return b;
}

template <bool FP64, bool RSD, bool W>
__global__ __launch_bounds__(kWG) void k_pair2(PairArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BinDesc& B = *a.bins;
    const int nb = B.nb;
    const int nm = RSD ? B.nm : 1;
    const int nbt = 3 * nb * nm;
    const int reps = a.reps;
    const int tid = threadIdx.x;
    const int rep = tid & (reps - 1);

    unsigned char* p = smem;
    uint32_t* hc = (uint32_t*)p;
    p += ((size_t)nbt * reps * 4 + 15) / 16 * 16;
    unsigned long long* hw = (unsigned long long*)p;
    if (W) p += (size_t)nbt * 16;
    float2* okb = (float2*)p;
    p += (size_t)(kMaxBins + 1) * 8;
    double* t64 = (double*)p;
    p += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;
    // tiles
    float4* tile32 = (float4*)p;
    double* tx = (double*)p;
    double* ty = tx + kWG;
    double* tz = ty + kWG;
    double* tw = FP64 ? (tz + kWG) : (double*)(tile32 + kWG);
    uint32_t* tf = (uint32_t*)(tw + kWG);

    for (int k = tid; k < nbt * reps; k += kWG) hc[k] = 0;
    if (W)
        for (int k = tid; k < 2 * nbt; k += kWG) hw[k] = 0;
    for (int k = tid; k < nb; k += kWG) okb[k] = make_float2(B.ok_lo[k], B.ok_hi[k]);
    for (int k = tid; k < nb + 2; k += kWG) t64[k] = B.t64[k];
    __shared__ int s_item;
    __syncthreads();

    const float rej_lo = B.rej_lo, rej_hi = B.rej_hi, ua = B.ua, ub = B.ub;
    const bool logb = B.log_bins != 0;
    const double t_lo = B.t64[0], t_hi = B.t64[nb + 1];
    const double ms = B.mu_scale;
    const float ms32 = B.ms32, mu_ew = B.mu_ew, mu_em = B.mu_em, mu_c = B.mu_c;
    const int shift = B.fx_shift2;
    const int nx = a.g.dim[0], ny = a.g.dim[1], nz = a.g.dim[2];

    Stats st{0, 0, 0};
    unsigned long long bound = 0;   // upper bound of u32 increments since the last flush
    const int xcd = blockIdx.x & 7;
    int tried = 0;

    while (true) {
        if (tid == 0) {
            int it = -1;
            while (tried < 8) {
                const int sidx = (xcd + tried) & 7;
                const int lo = (int)(((long long)a.n_chunks * sidx) >> 3);
                const int hi = (int)(((long long)a.n_chunks * (sidx + 1)) >> 3);
                const int k = atomicAdd(&a.queue[sidx], 1);
                if (lo + k < hi) { it = lo + k; break; }
                ++tried;
            }
            s_item = it;
        }
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item < 0) break;

        const int A = find_cell(a.chunk_off, a.g.ncell, item);
        const int q = item - a.chunk_off[A];
        const int cs = a.cell_start[A], ce = a.cell_start[A + 1];
        const int i0 = cs + q * kWG;
        const int i = i0 + tid;
        const bool vi = i < ce;
        const int nvalid = min(kWG, ce - i0);
        const int ax = A % nx, ay = (A / nx) % ny, az = A / (nx * ny);
        const double ox = a.g.lo[0] + ax * a.g.h, oy = a.g.lo[1] + ay * a.g.h, oz = a.g.lo[2] + az * a.g.h;

        double4 Pi = make_double4(0, 0, 0, 0);
        uint32_t fi = 2u;
        if (vi) { Pi = a.pw[i]; fi = a.flags[i]; }
        const float xi = (float)(Pi.x - ox), yi = (float)(Pi.y - oy), zi = (float)(Pi.z - oz);
        const float ox32 = (float)ox, oy32 = (float)oy, oz32 = (float)oz;
        const int ri = fi & 1, ci = (fi & 2) ? 0 : 1;

        for (int r = 0; r < a.nrows; ++r) {
            const Row row = a.rows[r];
            const int by = ay + row.dy, bz = az + row.dz;
            if (by < 0 || by >= ny || bz < 0 || bz >= nz) continue;
            const int bx0 = max(ax + row.x0, 0), bx1 = min(ax + row.x1, nx - 1);
            if (bx0 > bx1) continue;
            const bool self = (row.dy == 0 && row.dz == 0);
            const int rb = (bz * ny + by) * nx;
            int js = a.cell_start[rb + bx0];
            const int je = a.cell_start[rb + bx1 + 1];
            if (self) js = max(js, i0 + 1);
            for (int jt = js; jt < je; jt += kWG) {
                const int nt = min(kWG, je - jt);
                __syncthreads();
                if (tid < nt) {
                    const double4 Pj = a.pw[jt + tid];
                    const uint32_t fj = a.flags[jt + tid];
                    if constexpr (FP64) {
                        tx[tid] = Pj.x; ty[tid] = Pj.y; tz[tid] = Pj.z;
                        tf[tid] = fj;
                    } else {
                        tile32[tid] = make_float4((float)(Pj.x - ox), (float)(Pj.y - oy), (float)(Pj.z - oz),
                                                  __uint_as_float(fj));
                    }
                    if (W) tw[tid] = Pj.w;
                }
                __syncthreads();
                bound += (unsigned long long)nvalid * nt * 2ull;
                int kbeg = 0;
                if (self) kbeg = max(0, i + 1 - jt);
                if (!vi) kbeg = nt;
                st.cand += (unsigned long long)max(0, nt - kbeg);
                for (int k = kbeg; k < nt; ++k) {
                    int b, m = 0, mr = 0;
                    uint32_t fj;
                    if constexpr (FP64) {
                        const double4 Pj = make_double4(tx[k], ty[k], tz[k], 0.0);
                        const double s = exact_s(Pi, Pj);
                        if (!(s > t_lo && s < t_hi)) continue;
                        const float sf = (float)s;
                        float u = logb ? fmaf(__builtin_amdgcn_logf(sf), ua, ub)
                                       : fmaf(__builtin_amdgcn_sqrtf(sf), ua, ub);
                        b = (int)u;
                        b = b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
                        while (b > 0 && s < t64[b]) --b;
                        while (b < nb && s >= t64[b + 1]) ++b;
                        if (b == nb) { st.over += (uint64_t)(ci + (((tf[k] & 2) == 0) ? 1 : 0)); continue; }
                        if constexpr (RSD) exact_mu(Pi, Pj, ms, nm, m, mr);
                        fj = tf[k];
                    } else {
                        const float4 tj = tile32[k];
                        const float dx = tj.x - xi, dy = tj.y - yi, dz = tj.z - zi;
                        const float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        if (!(s > rej_lo && s < rej_hi)) continue;
                        float u = logb ? fmaf(__builtin_amdgcn_logf(s), ua, ub)
                                       : fmaf(__builtin_amdgcn_sqrtf(s), ua, ub);
                        b = (int)u;
                        b = b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
                        const float2 ok = okb[b];
                        bool safe = (s >= ok.x) && (s < ok.y);
                        if constexpr (RSD) {
                            const float mx = fmaf(0.5f, xi + tj.x, ox32);
                            const float my = fmaf(0.5f, yi + tj.y, oy32);
                            const float mz = fmaf(0.5f, zi + tj.z, oz32);
                            // sep = i - j = -(dx, dy, dz)
                            const float D = -fmaf(mz, dz, fmaf(my, dy, mx * dx));
                            const float M2 = fmaf(mz, mz, fmaf(my, my, mx * mx));
                            const float rm = __builtin_amdgcn_rsqf(M2), rs = __builtin_amdgcn_rsqf(s);
                            const float th = D * rm * rs;
                            const float t = fmaf(th, ms32, ms32);
                            const float tfl = floorf(t);
                            const float fr = t - tfl;
                            const float dt = ms32 * fmaf(mu_ew, rs, fmaf(mu_em, rm, mu_c));
                            const int mi = (int)tfl;
                            const bool near_lo = (fr < dt) && (mi >= 1) && (mi <= nm - 1);
                            const bool near_hi = (fr > 1.0f - dt) && (mi + 1 >= 1) && (mi + 1 <= nm - 1);
                            // dt is not finite when mid or sep is ~0: always re-evaluate
                            safe = safe && !near_lo && !near_hi && (dt < 0.25f);
                            m = mi < 0 ? 0 : (mi > nm - 1 ? nm - 1 : mi);
                            mr = nm - 1 - m;
                        }
                        fj = __float_as_uint(tj.w);
                        if (!safe) {
                            ++st.rechk;
                            const double4 Pj = a.pw[jt + k];
                            const double se = exact_s(Pi, Pj);
                            if (!(se > t_lo && se < t_hi)) continue;
                            b = exact_bin(se, t64, nb);
                            if (b == nb) { st.over += (uint64_t)(ci + (((fj & 2) == 0) ? 1 : 0)); continue; }
                            if constexpr (RSD) exact_mu(Pi, Pj, ms, nm, m, mr);
                        }
                    }
                    const int rj = fj & 1, cj = (fj & 2) ? 0 : 1;
                    double pwt = 0.0;
                    if constexpr (W) pwt = Pi.w * tw[k];
                    accumulate<RSD, W>(hc, hw, reps, rep, nb, nm, ri + rj, b, m, mr, ci, cj, pwt, shift);
                }
                // keep every u32 LDS counter below 2^32 (bound is uniform over the work-group)
                if (bound > (1ull << 31)) {
                    __syncthreads();
                    flush(hc, hw, nbt, reps, W, a.out_cnt, a.out_w);
                    __syncthreads();
                    bound = 0;
                }
            }
        }
    }
    __syncthreads();
    flush(hc, hw, nbt, reps, W, a.out_cnt, a.out_w);

    // wave-reduce the statistics
    for (int off = 32; off > 0; off >>= 1) {
        st.cand += __shfl_down(st.cand, off, 64);
        st.rechk += __shfl_down(st.rechk, off, 64);
        st.over += __shfl_down(st.over, off, 64);
    }
    if ((tid & 63) == 0) {
        if (st.cand) atomicAdd(&a.out_stats[0], st.cand);
        if (st.rechk) atomicAdd(&a.out_stats[1], st.rechk);
        if (st.over) atomicAdd(&a.out_stats[2], st.over);
    }
}

template <bool FP64, bool RSD, bool W>
static hipError_t launch_t(const PairArgs& a, size_t lds, int nblocks, hipStream_t st)
{
    auto kern = k_pair2<FP64, RSD, W>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(kWG), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_pair2(const PairArgs& a, int nb, int nm, bool fp64, bool rsd, bool weighted, int nblocks,
                        hipStream_t st)
{
    const size_t lds = pair2_lds_bytes(nb, rsd ? nm : 1, a.reps, fp64, weighted);
    if (fp64) {
        if (rsd) return weighted ? launch_t<true, true, true>(a, lds, nblocks, st) : launch_t<true, true, false>(a, lds, nblocks, st);
        return weighted ? launch_t<true, false, true>(a, lds, nblocks, st) : launch_t<true, false, false>(a, lds, nblocks, st);
    }
    if (rsd) return weighted ? launch_t<false, true, true>(a, lds, nblocks, st) : launch_t<false, true, false>(a, lds, nblocks, st);
    return weighted ? launch_t<false, false, true>(a, lds, nblocks, st) : launch_t<false, false, false>(a, lds, nblocks, st);
}

template <bool FP64, bool RSD, bool W>
static int occ_t(size_t lds)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_pair2<FP64, RSD, W>, kWG, lds) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

int pair2_occupancy(int nb, int nm, int reps, bool fp64, bool rsd, bool weighted)
{
    const size_t lds = pair2_lds_bytes(nb, rsd ? nm : 1, reps, fp64, weighted);
    if (fp64) {
        if (rsd) return weighted ? occ_t<true, true, true>(lds) : occ_t<true, true, false>(lds);
        return weighted ? occ_t<true, false, true>(lds) : occ_t<true, false, false>(lds);
    }
    if (rsd) return weighted ? occ_t<false, true, true>(lds) : occ_t<false, true, false>(lds);
    return weighted ? occ_t<false, false, true>(lds) : occ_t<false, false, false>(lds);
}

}  // namespace gs

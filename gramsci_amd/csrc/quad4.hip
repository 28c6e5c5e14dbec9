// quad4.hip -- persistent 4PCF quad kernel.
//
// Replaces create_graph (gramsci.F90:288-348) + query_4pcf_all2 (:662-861).
// For every apex i that is not a buffer point and has more than 3 kept
// neighbours (:675), and every k1 < k2 < k3 of its neighbours in ascending
// ORIGINAL index (the reference's neighbour lists are sorted by id,
// sorting_module.F90:36-47), with all three cross sides kept
// (find_dist :685,692,694), the quad is binned at
//     cell = (bintable(ind1, ind4, ind5, ind3), ind2, ind6)
// where ind1..3 are the apex sides, ind4 = (k1,k2), ind5 = (k2,k3) and
// ind6 = (k1,k3); bintable is only filled for non-decreasing perimeters
// (:1402-1420), so other orientations are dropped (:701).
//
// One work-group per apex, persistent over apexes:
//   1. gather the kept neighbours (exact FP64 bins) into LDS;
//   2. sort them by original index (a rank sort in LDS; bitonic for device-memory lists);
//   3. tabulate every neighbour-neighbour side bin in a K x K LDS matrix;
//   4. enumerate linked (k1, k2) pairs over the lanes, each lane scanning k3.
// Counts go to an LDS histogram [cell][16 random patterns] (or straight to
// global memory when it does not fit); weighted runs add FP64 sums of the
// quad products per pattern and of the six disc pair products (:754-818).
//
// With `tetra` the kernel is query_graph_tetrahedron (:548-660) instead: no
// minimum neighbour count, only quads whose six sides share one bin b
// (cell = b), and the disc sums are triple products that include the weight
// of the point whose random status selects the term (:590-631).
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

using namespace dev;

size_t quad4_lds_bytes(int nb, int ncell, int kmax, bool lds_hist, bool tetra, int nrows, bool glist)
{
    size_t b = 0;
    if (lds_hist) b += (size_t)ncell * 16 * 4;
    b += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;            // t64
    if (!tetra) b += ((size_t)nb * nb * nb * nb * 2 + 15) / 16 * 16;   // perimeter table (u16)
    b += ((size_t)nrows * 12 + 15) / 16 * 16;                    // stencil rows: first point, length, offset
    if (glist) return b;                                         // neighbour list in device memory
    b += (size_t)kmax * (24 + 8 + 4 + 4 + 4);                     // x, y, z, w, key, slot/packed, sorted j
    b += ((size_t)kmax * kmax + 15) / 16 * 16;                    // link matrix (u8: bin+1, 0 = not kept)
    return b;
}

// BIG: the apex's neighbour list in device memory (a.glist), for apexes with more neighbours
// than an LDS list holds; the neighbour-neighbour bins are evaluated where needed (no K x K
// matrix), and the quads run in segments with the LDS counts drained after each.
template <bool W, bool BIG>
__global__ __launch_bounds__(kWG) void k_quad4(QuadArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BinDesc& B = *a.bins;
    const int nb = B.nb;
    const int kmax = a.kmax;
    const int ncell = a.ncell;
    const bool lh = a.lds_hist != 0;
    const bool tetra = a.tetra != 0;
    const int tid = threadIdx.x;

    unsigned char* p = smem;
    uint32_t* hc = (uint32_t*)p;
    if (lh) p += (size_t)ncell * 16 * 4;
    double* t64 = (double*)p;
    p += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;
    uint16_t* tab = (uint16_t*)p;
    if (!tetra) p += ((size_t)nb * nb * nb * nb * 2 + 15) / 16 * 16;
    // the apex's stencil rows: first point, number of points, and candidate number of the first point
    int32_t* rw_js = (int32_t*)p;
    int32_t* rw_len = rw_js + a.nrows;
    int32_t* rw_off = rw_len + a.nrows;
    p += ((size_t)a.nrows * 12 + 15) / 16 * 16;
    if constexpr (BIG) p = (unsigned char*)a.glist + (size_t)blockIdx.x * (size_t)a.glist_stride;
    double* nx_ = (double*)p;
    double* ny_ = nx_ + kmax;
    double* nz_ = ny_ + kmax;
    double* nw_ = nz_ + kmax;
    int32_t* key = (int32_t*)(nw_ + kmax);   // original index (sort key)
    int32_t* pk = key + kmax;                // sort permutation, then bin | rand << 8 in sorted order
    int32_t* jl = pk + kmax;                 // bin | rand << 8 in gather order
    uint8_t* L = (uint8_t*)(jl + kmax);      // L[u * kmax + v] (LDS list; BIG: the sort's staging area)

    __shared__ int s_item, s_k, s_over, s_ncand;

    if (lh)
        for (int k = tid; k < ncell * 16; k += kWG) hc[k] = 0;
    for (int k = tid; k < nb + 2; k += kWG) t64[k] = B.t64[k];
    if (!tetra)
        for (int k = tid; k < nb * nb * nb * nb; k += kWG) tab[k] = a.cfg_tab[k];
    __syncthreads();

    const int nxg = a.g.dim[0], nyg = a.g.dim[1], nzg = a.g.dim[2];
    const double inv_h = 1.0 / a.g.h, inv_hx = 1.0 / a.g.hx;
    unsigned long long units = 0, binned = 0, over = 0, cand = 0;
    const int xcd = blockIdx.x & 7;
    int tried = 0;
    const int nb2 = nb * nb;

    // adds the LDS counts to the global counts and clears them (all threads of the work-group)
    auto drain_counts = [&]() {
        for (int k = tid; k < ncell * 16; k += kWG) {
            const uint32_t v = hc[k];
            if (v) {
                atomicAdd(&a.out_cnt[k], (unsigned long long)v);
                hc[k] = 0;
            }
        }
    };
    unsigned long long bound = 0;   // LDS count increments since the last drain (work-group uniform)

    const int bsz = max(a.batch, 1);
    int b_next = 0, b_end = 0;      // thread 0: the work-group's current batch of apexes
    while (true) {
        if (tid == 0) {
            if (b_next >= b_end) b_next = next_items(a.queue, a.n_items, xcd, tried, bsz, b_end);
            s_item = b_next;        // -1 at the end
            if (b_next >= 0) ++b_next;
            s_k = 0;
            s_over = 0;
        }
        __syncthreads();
        const int item = s_item;
        if (item < 0) break;
        const int apex = a.items ? a.items[item] : item;
        const uint32_t fa = a.flags[apex];
        if (fa & 2u) {
            __syncthreads();
            continue;
        }
        const double4 Pa = a.pw[apex];
        const int cx = min(max((int)floor((Pa.x - a.g.lo[0]) * inv_hx), 0), nxg - 1);
        const int cy = min(max((int)floor((Pa.y - a.g.lo[1]) * inv_h), 0), nyg - 1);
        const int cz = min(max((int)floor((Pa.z - a.g.lo[2]) * inv_h), 0), nzg - 1);

        const unsigned long long cand0 = cand, over0 = over;   // restored if the apex is queued
        // ---- 1. gather: the stencil rows' point ranges are read in parallel (one row per thread) into
        // LDS and numbered by a prefix sum over the rows; the candidates of all rows are then numbered
        // consecutively and thread t takes numbers t, t + 256, ... (every thread busy even where rows
        // are short)
        for (int r = tid; r < a.nrows; r += kWG) {
            const Row rw = a.rows[r];
            const int by = cy + rw.dy, bz = cz + rw.dz;
            const int bx0 = max(cx + rw.x0, 0), bx1 = min(cx + rw.x1, nxg - 1);
            int js = 0, len = 0;
            if (by >= 0 && by < nyg && bz >= 0 && bz < nzg && bx0 <= bx1) {
                const int rb = (bz * nyg + by) * nxg;
                js = a.cell_start[rb + bx0];
                len = a.cell_start[rb + bx1 + 1] - js;
            }
            rw_js[r] = js;
            rw_len[r] = len;
        }
        __syncthreads();
        if (tid < 64) {         // exclusive prefix of the row lengths, 64 rows at a time (wave 0)
            int run = 0;
            for (int b0 = 0; b0 < a.nrows; b0 += 64) {
                const int x = b0 + tid < a.nrows ? rw_len[b0 + tid] : 0;
                int inc = x;
                for (int off = 1; off < 64; off <<= 1) {
                    const int y = __shfl_up(inc, off, 64);
                    if (tid >= off) inc += y;
                }
                if (b0 + tid < a.nrows) rw_off[b0 + tid] = run + inc - x;
                run += __shfl(inc, 63, 64);
            }
            if (tid == 0) s_ncand = run;
        }
        __syncthreads();
        const int ncand = s_ncand;
        int row = 0, row_end = a.nrows > 0 ? rw_len[0] : 0, j0 = a.nrows > 0 ? rw_js[0] : 0;
        {
            for (int next = tid; next < ncand; next += kWG) {
                while (next >= row_end) {           // the row holding candidate number next
                    ++row;
                    const int row_beg = rw_off[row];
                    row_end = row_beg + rw_len[row];
                    j0 = rw_js[row] - row_beg;
                }
                const int j = j0 + next;
                ++cand;
                const double4 Pj = a.pw[j];
                const int b = bin_of(exact_s(Pa.x, Pa.y, Pa.z, Pj.x, Pj.y, Pj.z), t64, nb);
                if (b < 0) continue;
                if (b == nb) { ++over; continue; }
                const int pos = atomicAdd(&s_k, 1);
                if (pos >= kmax) { s_over = 1; continue; }
                nx_[pos] = Pj.x; ny_[pos] = Pj.y; nz_[pos] = Pj.z; nw_[pos] = Pj.w;
                key[pos] = a.orig[j];
                jl[pos] = (int)(b | ((a.flags[j] & 1u) << 8));   // packed, moved by the sort below
            }
        }
        __syncthreads();
        const int K = min(s_k, kmax);
        if (s_over) {
            if (a.ovf) {
                // left to a later pass with room for all its neighbours (s_k counted every one)
                if (tid == 0) {
                    a.ovf[atomicAdd(a.ovf_n, 1)] = apex;
                    if (a.ovf_kmax) atomicMax(a.ovf_kmax, s_k);
                }
                cand = cand0;       // the later pass counts this apex's candidates
                over = over0;
                __syncthreads();
                continue;
            }
            if (tid == 0) atomicExch(a.err, 1);
        }
        if (K <= (tetra ? 2 : 3)) {        // gramsci.F90:675 (the tetrahedron query has no such skip)
            __syncthreads();
            continue;
        }

        // ---- 2. sort by original index
        if constexpr (!BIG) {
            // K <= kmax <= kWG: one thread per entry.  The original indices are distinct, so entry u goes
            // to the number of keys below its own (all threads read key[j] together: LDS broadcasts); the
            // entry is held in registers across the barrier and written to its sorted place
            const bool own = tid < K;
            int r = 0, vj = 0;
            double vx = 0.0, vy = 0.0, vz = 0.0, vw = 0.0;
            if (own) {
                const int ku = key[tid];
                for (int j = 0; j < K; ++j) r += key[j] < ku ? 1 : 0;
                vx = nx_[tid]; vy = ny_[tid]; vz = nz_[tid]; vw = nw_[tid];
                vj = jl[tid];
            }
            __syncthreads();
            if (own) {
                nx_[r] = vx; ny_[r] = vy; nz_[r] = vz; nw_[r] = vw;
                pk[r] = vj;
            }
            __syncthreads();
        } else {
            // bitonic sort of (key, slot), then gather (slot pk[u] -> position u) through the device-memory
            // list's staging area
            double* tmp = (double*)L;
            int np2 = 1;
            while (np2 < K) np2 <<= 1;
            for (int k = tid; k < np2; k += kWG) pk[k] = k;           // slot permutation
            __syncthreads();
            for (int size = 2; size <= np2; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int t = tid; t < np2 / 2; t += kWG) {
                        const int lo = 2 * t - (t & (stride - 1));
                        const int hi = lo + stride;
                        const bool up = ((lo & size) == 0);
                        const int sa = pk[lo], sb = pk[hi];
                        const int ka = sa < K ? key[sa] : 0x7fffffff;
                        const int kb = sb < K ? key[sb] : 0x7fffffff;
                        if ((ka > kb) == up) { pk[lo] = sb; pk[hi] = sa; }
                    }
                    __syncthreads();
                }
            }
            for (int u = tid; u < K; u += kWG) {
                const int sl = pk[u];
                tmp[u] = nx_[sl];
                tmp[K + u] = ny_[sl];
                tmp[2 * K + u] = nz_[sl];
                tmp[3 * K + u] = nw_[sl];
                ((int*)(tmp + 4 * K))[u] = jl[sl];
            }
            __syncthreads();
            for (int u = tid; u < K; u += kWG) {
                nx_[u] = tmp[u]; ny_[u] = tmp[K + u]; nz_[u] = tmp[2 * K + u]; nw_[u] = tmp[3 * K + u];
                pk[u] = ((int*)(tmp + 4 * K))[u];
            }
            __syncthreads();
        }

        // ---- 3. link matrix of neighbour-neighbour side bins (find_dist); a device-memory list
        // evaluates them where they are needed instead
        auto link = [&](int x, int y) -> int {       // x < y: bin + 1 of side (x, y), 0 if not kept
            if constexpr (BIG) {
                const int b = bin_of(exact_s(nx_[x], ny_[x], nz_[x], nx_[y], ny_[y], nz_[y]), t64, nb);
                return (b >= 0 && b < nb) ? b + 1 : 0;
            } else {
                return L[x * K + y];
            }
        };
        if (!BIG)
        for (int e = tid; e < K * K; e += kWG) {
            const int u = e / K, v = e - u * K;
            uint8_t val = 0;
            if (u < v) {
                const int b = bin_of(exact_s(nx_[u], ny_[u], nz_[u], nx_[v], ny_[v], nz_[v]), t64, nb);
                val = (b >= 0 && b < nb) ? (uint8_t)(b + 1) : 0;
            }
            L[u * K + v] = val;
        }
        __syncthreads();

        // ---- 4. quads: linked (k1, k2) over the lanes, k3 > k2 scanned by the lane
        const int ri = (int)(fa & 1u);
        const double wi = Pa.w;
        const long long P = (long long)K * (K - 1) / 2;
        int u = 0;
        long long rem = tid;
        while (u < K - 1 && rem >= K - 1 - u) { rem -= K - 1 - u; ++u; }
        // one segment for an LDS list; for a device-memory list segments of at most 2^30 / K pairs
        // (each adds fewer than K counts), with the u32 LDS counts drained after each
        const long long seg = BIG ? max((long long)kWG, ((1ll << 30) / max(K, 1)) / kWG * kWG) : max(P, 1ll);
        long long pi = tid;
        for (long long s0 = 0; s0 < P; s0 += seg) {
        if (BIG && s0 > 0) {
            __syncthreads();
            if (lh) drain_counts();
            __syncthreads();
        }
        const long long s1 = min(P, s0 + seg);
        for (; pi < s1; pi += kWG) {
            const int v = u + 1 + (int)rem;
            const int ind4 = link(u, v);
            const int p1 = pk[u], p2 = pk[v];
            const int ind1 = (p1 & 255) + 1, ind2 = (p2 & 255) + 1;
            if (ind4 && (!tetra || (ind2 == ind1 && ind4 == ind1))) {     // tetra: gramsci.F90:564-567
                units += (unsigned long long)(K - 1 - v);
                const int r1 = (p1 >> 8) & 1, r2 = (p2 >> 8) & 1;
                for (int k3 = v + 1; k3 < K; ++k3) {
                    const int ind5 = link(v, k3);
                    if (!ind5) continue;
                    const int ind6 = link(u, k3);
                    if (!ind6) continue;
                    const int p3 = pk[k3];
                    const int ind3 = (p3 & 255) + 1;
                    int cell;
                    if (tetra) {
                        if (ind3 != ind1 || ind5 != ind1 || ind6 != ind1) continue;   // :571-579
                        cell = ind1 - 1;
                    } else {
                        const int cfg = tab[(((ind1 - 1) * nb + (ind4 - 1)) * nb + (ind5 - 1)) * nb + (ind3 - 1)];
                        if (cfg == 0) continue;
                        cell = (cfg - 1) * nb2 + (ind2 - 1) * nb + (ind6 - 1);
                    }
                    const int r3 = (p3 >> 8) & 1;
                    const int pat = ri | (r1 << 1) | (r2 << 2) | (r3 << 3);
                    ++binned;
                    if (lh) atomicAdd(&hc[cell * 16 + pat], 1u);
                    else atomicAdd(&a.out_cnt[(size_t)cell * 16 + pat], 1ull);
                    if constexpr (W) {
                        const double w1 = nw_[u], w2 = nw_[v], w3 = nw_[k3];
                        double* ow = a.out_w;
                        atomicAdd(&ow[(size_t)cell * 16 + pat], wi * w1 * w2 * w3);   // gramsci.F90:710
                        double* od = ow + (size_t)ncell * 16;                       // [6][ncell][4]
                        // disc pair products and their conditions, gramsci.F90:754-818; the
                        // tetrahedron's terms also carry the conditioning point's weight (:590-631)
                        const double fi = tetra ? wi : 1.0, f1 = tetra ? w1 : 1.0, f3 = tetra ? w3 : 1.0;
                        if (ri) {
                            atomicAdd(&od[((size_t)3 * ncell + cell) * 4 + (r1 | r2 << 1)], w1 * w2 * fi);
                            atomicAdd(&od[((size_t)4 * ncell + cell) * 4 + (r3 | r2 << 1)], w3 * w2 * fi);
                            atomicAdd(&od[((size_t)5 * ncell + cell) * 4 + (r3 | r1 << 1)], w3 * w1 * fi);
                        }
                        if (r3) {
                            atomicAdd(&od[((size_t)1 * ncell + cell) * 4 + (ri | r2 << 1)], wi * w2 * f3);
                            atomicAdd(&od[((size_t)0 * ncell + cell) * 4 + (ri | r1 << 1)], wi * w1 * f3);
                        }
                        if (r1) atomicAdd(&od[((size_t)2 * ncell + cell) * 4 + (ri | r3 << 1)], wi * w3 * f1);
                    }
                }
            }
            rem += kWG;
            while (u < K - 1 && rem >= K - 1 - u) { rem -= K - 1 - u; ++u; }
        }
        }
        __syncthreads();
        // The u32 LDS counts take at most C(K,3) increments per apex (K <= 256 in LDS, so fewer
        // than 2^26); the work-group drains them into the 64-bit global counts before the
        // increments since the last drain could pass 2^31 (a persistent work-group may count many
        // apexes of hundreds of neighbours).  A device-memory list drains after every apex.
        if (lh) {
            bound += BIG ? (1ull << 31) : (unsigned long long)K * (K - 1) * (K - 2) / 6;
            if (bound > (1ull << 31) - (1ull << 26)) {
                drain_counts();
                bound = 0;
                __syncthreads();
            }
        }
    }
    __syncthreads();
    if (lh) drain_counts();
    for (int off = 32; off > 0; off >>= 1) {
        units += __shfl_down(units, off, 64);
        binned += __shfl_down(binned, off, 64);
        over += __shfl_down(over, off, 64);
        cand += __shfl_down(cand, off, 64);
    }
    if ((tid & 63) == 0) {
        if (cand) atomicAdd(&a.out_stats[0], cand);
        if (over) atomicAdd(&a.out_stats[2], over);
        if (units) atomicAdd(&a.out_stats[3], units);
        if (binned) atomicAdd(&a.out_stats[4], binned);
    }
}

template <bool W, bool BIG>
static hipError_t launch_q(const QuadArgs& a, size_t lds, int nblocks, hipStream_t st)
{
    auto kern = k_quad4<W, BIG>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(kWG), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_quad4(const QuadArgs& a, int nb, bool weighted, int nblocks, hipStream_t st)
{
    const bool big = a.glist != nullptr;
    const size_t lds = quad4_lds_bytes(nb, a.ncell, a.kmax, a.lds_hist != 0, a.tetra != 0, a.nrows, big);
    if (big) return weighted ? launch_q<true, true>(a, lds, nblocks, st) : launch_q<false, true>(a, lds, nblocks, st);
    return weighted ? launch_q<true, false>(a, lds, nblocks, st) : launch_q<false, false>(a, lds, nblocks, st);
}

int quad4_occupancy(const QuadArgs& a, int nb, bool weighted)
{
    const bool big = a.glist != nullptr;
    const size_t lds = quad4_lds_bytes(nb, a.ncell, a.kmax, a.lds_hist != 0, a.tetra != 0, a.nrows, big);
    void (*kern)(QuadArgs) = big ? (weighted ? k_quad4<true, true> : k_quad4<false, true>)
                                 : (weighted ? k_quad4<true, false> : k_quad4<false, false>);
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return 1;
    int n = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, kWG, lds);
    return (e == hipSuccess && n > 0) ? n : 1;
}

}  // namespace gs

// triplet3.hip -- persistent 3PCF triplet kernel.
//
// Replaces create_graph (gramsci.F90:288-348) + query_graph_3pcf_all
// (:463-545) with find_dist (:1579-1609).  For every apex i that is not a
// buffer point, and every unordered pair {k1, k2} of its kept neighbours
// whose third side (k1, k2) is also kept, the triangle is counted in the
// configuration of its sorted side bins (bintable, :1459-1474) and in the
// class given by the number of randoms among (i, k1, k2).  Each triangle is
// thus counted once per apex (3 times in total), as in the reference.
//
// Variants selected at run time:
//   nm >= 2  RSD 3PCF (:495-500): the triangle also goes to the mu bin
//            find_normal(mu(k1), mu(k2)) (:1611-1625), tabulated on the host;
//            mu(k) is the apex -> k mu bin of create_graph (:320-328), k1 the
//            neighbour with the smaller original index;
//   equi     query_graph_equilateral_triangle (:402-461): only triangles whose
//            three sides share one bin b, counted at configuration b.
//
// One work-group per apex, persistent over apexes in cell order (8 XCD
// queues as in the pair kernel):
//   1. gather: the work-group scans the apex's full stencil of cell rows and
//      appends every neighbour with rmin < d < rmax to an LDS list, with its
//      exact FP64 bin (side 1 / side 2 of the reference);
//   2. pairs: the K(K-1)/2 neighbour pairs, folded into rows of K, go to the
//      waves row by row or, where rows would leave lanes idle, as one flat
//      sequence interleaved over the work-group; the third side is
//      evaluated directly from the neighbours' coordinates - in FP64, or in
//      FP32 relative to the apex with the same guard band and FP64 recheck as
//      the pair kernel - instead of a binary search in a neighbour list.
#include <type_traits>

#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

using namespace dev;

size_t triplet3_glist_entry_bytes(const TripArgs& a, bool fp64, bool weighted)
{
    return (fp64 ? 32 : 16) + 4 + (weighted ? 8 : 0) + ((a.nm > 1 || weighted) ? 4 : 0);
}

size_t triplet3_lds_bytes(const TripArgs& a, int nb, bool fp64, bool weighted)
{
    const size_t nh = (size_t)a.ncfg * a.nm * 4;
    const size_t kmax = (size_t)a.kmax;
    size_t b = 0;
    if (a.lds_hist) {
        b += (nh * a.reps * 4 + 15) / 16 * 16;                  // counts
        if (weighted) b += nh * 16;                             // split fixed-point sums
    }
    b += ((size_t)(kMaxBins + 1) * 8);                          // ok_lo/ok_hi
    b += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;           // t64
    b += ((size_t)nb * nb * nb * 2 + 15) / 16 * 16;             // config table (u16)
    if (a.nm > 1) b += ((size_t)a.nm * a.nm + 15) / 16 * 16;    // find_normal table (u8)
    b += ((size_t)a.nrows * 12 + 15) / 16 * 16;                 // stencil rows: first point, length, offset
    // neighbour list: coords + packed bin/class/mu, sorted index, weight, orig (in LDS unless glist)
    if (!a.glist) b += kmax * triplet3_glist_entry_bytes(a, fp64, weighted);
    return b;
}

// record packing: bits 0..14 2 * bin * nb (0-based bin, pre-multiplied for the byte offset of the
// configuration table entry), byte 2 the mu bin (0-based, RSD) or the buffer flag (0/1, otherwise),
// byte 3 the class (1: random).  The configuration table holds 4 * configuration (the histogram index
// of class 0).
//
// The variant is a template parameter so that each one pays only for its own
// registers: MODE bit 0 RSD (nm >= 2), bit 1 equilateral (0 = plain); GEN 1 for
// a batched launch or a histogram in global memory (the plain launch with an
// LDS histogram, the common case, runs with GEN 0), GEN 2 for neighbour lists in
// device memory (apexes with more neighbours than an LDS list holds).
enum { kT3Rsd = 1, kT3Equi = 2 };

template <bool FP64, bool W, int MODE, int GEN, bool LOGB>
__global__ __launch_bounds__(kWG) void k_triplet3(TripArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BinDesc& B = *a.bins;
    const int nb = B.nb;
    const int nm = (MODE & kT3Rsd) ? a.nm : 1;
    const int nh = a.ncfg * nm * 4;
    const int reps = a.reps;
    const int kmax = a.kmax;
    const bool lh = !GEN || a.lds_hist != 0;
    constexpr bool equi = (MODE & kT3Equi) != 0;
    // Isotropic modes visit each triangle once: a triangle's configuration (its sorted side bins) and
    // class (its randoms) do not depend on which vertex is the apex, so the reference's count at each
    // vertex that is not a buffer point is added at once by the first vertex in the order "points that
    // are not buffer points, by index, then buffer points", whose list keeps only the neighbours after
    // it.  (The RSD mode bins by the apex's mu values and visits every apex's pairs.)
    constexpr bool SYM = (MODE & kT3Rsd) == 0;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;              // threads per work-group: 256, or 64 for short neighbour lists
    const int lane = tid & 63;
    const int rep = lane & (reps - 1);
    const int rsh = 31 - __clz(reps);       // reps is a power of two

    unsigned char* p = smem;
    uint32_t* hc = (uint32_t*)p;
    unsigned long long* hw = nullptr;
    if (lh) {
        p += ((size_t)nh * reps * 4 + 15) / 16 * 16;
        hw = (unsigned long long*)p;
        if (W) p += (size_t)nh * 16;
    }
    float2* okb = (float2*)p;
    p += (size_t)(kMaxBins + 1) * 8;
    double* t64 = (double*)p;
    p += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;
    uint16_t* tab = (uint16_t*)p;
    p += ((size_t)nb * nb * nb * 2 + 15) / 16 * 16;
    uint8_t* ntab = (uint8_t*)p;
    if (nm > 1) p += ((size_t)nm * nm + 15) / 16 * 16;
    // the apex's stencil rows: first point, number of points, and candidate number of the first point
    int32_t* rw_js = (int32_t*)p;
    int32_t* rw_len = rw_js + a.nrows;
    int32_t* rw_off = rw_len + a.nrows;
    p += ((size_t)a.nrows * 12 + 15) / 16 * 16;
    // the neighbour list: in LDS, or (GEN 2) in this work-group's part of device memory
    if constexpr (GEN == 2) p = (unsigned char*)a.glist + (size_t)blockIdx.x * (size_t)a.glist_stride;
    float4* rec32 = (float4*)p;          // FP32: x, y, z relative to the apex, packed bin/class/mu
    double4* rec64 = (double4*)p;        // FP64: absolute x, y, z, packed bin/class/mu (as a double)
    p += (size_t)kmax * (FP64 ? 32 : 16);
    int32_t* jl = (int32_t*)p;
    p += (size_t)kmax * 4;
    double* wl = (double*)p;             // neighbour weights * 2^shift3
    if (W) p += (size_t)kmax * 8;
    int32_t* ol = (int32_t*)p;           // RSD or weighted: original index of each neighbour

    __shared__ int s_item, s_k, s_over, s_ncand, s_kall;

    if (lh) {
        for (int k = tid; k < nh * reps; k += nt) hc[k] = 0;
        if (W)
            for (int k = tid; k < 2 * nh; k += nt) hw[k] = 0;
    }
    for (int k = tid; k < nb; k += nt) okb[k] = make_float2(B.ok_lo[k], B.ok_hi[k]);
    for (int k = tid; k < nb + 2; k += nt) t64[k] = B.t64[k];
    if (!equi)
        for (int k = tid; k < nb * nb * nb; k += nt) tab[k] = a.cfg_tab[k];
    if (nm > 1)
        for (int k = tid; k < nm * nm; k += nt) ntab[k] = a.ntab[k];
    __syncthreads();
    const double ms = B.mu_scale;

    const float ua = B.ua, ub = B.ub;
    // ub in a VGPR for the pair loop's bin estimate (an FMA reads at most one scalar operand)
    float ubv = ub;
    asm volatile("" : "+v"(ubv));
    const float nbm1f = (float)(nb - 1);
    const double t_lo = B.t64[0], t_hi = B.t64[nb + 1];
    const float rej_lo = B.rej_lo, rej_hi = B.rej_hi;
    int shift = B.fx_shift3;         // batched: the apex's region's scale
    const double inv_h = 1.0 / a.g.h, inv_hx = 1.0 / a.g.hx;
    const bool batched = GEN && a.nreg > 0;
    int cur_r = batched ? -1 : 0;   // region whose histogram the LDS holds (work-group uniform)

    // adds the LDS histogram to region r's output and clears it (all threads of the work-group)
    auto drain = [&](int r) {
        unsigned long long* oc = a.out_cnt + (size_t)r * nh;
        unsigned long long* ow = a.out_w + (size_t)r * 2 * nh;
        for (int k = tid; k < nh; k += nt) {
            unsigned long long sum = 0;
            for (int q = 0; q < reps; ++q) { sum += hc[k * reps + q]; hc[k * reps + q] = 0; }
            if (sum) atomicAdd(&oc[k], sum);
            if (W) {
                const unsigned long long lo = hw[2 * k], hi = hw[2 * k + 1];
                hw[2 * k] = 0; hw[2 * k + 1] = 0;
                if (lo | hi) {
                    const unsigned long long L = (hi << 32) + lo;
                    const unsigned long long H = (hi >> 32) + (L < lo ? 1ull : 0ull);
                    glb_add128(&ow[2 * k], L, H);
                }
            }
        }
    };

    unsigned long long units = 0, rechk = 0, over = 0, cand = 0;   // binned triplets: the histogram's total
    unsigned long long evals = 0;   // neighbour pairs whose third side this kernel evaluated
    unsigned long long bound = 0;   // u32 increments since the last drain (uniform)
    const int xcd = blockIdx.x & 7;
    int tried = 0;

    const int n_items = a.n_items_dev ? *a.n_items_dev : a.n_items;
    const int bsz = max(a.batch, 1);
    int b_next = 0, b_end = 0;      // thread 0: the work-group's current batch of apexes
    while (true) {
        if (tid == 0) {
            if (b_next >= b_end) b_next = next_items(a.queue, n_items, xcd, tried, bsz, b_end);
            s_item = b_next;        // -1 at the end
            if (b_next >= 0) ++b_next;
            s_k = 0;
            s_kall = 0;
            s_over = 0;
        }
        __syncthreads();
        const int item = s_item;
        if (item < 0) break;
        const int apex = a.items ? a.items[item] : item;
        const uint32_t fa = a.flags[apex];
        if (fa & 2u) {          // buffer point: never an apex (gramsci.F90:476)
            __syncthreads();
            continue;
        }
        // the apex's grid: the whole catalogue's, or its region's in a batch
        int nx = a.g.dim[0], ny = a.g.dim[1], nz = a.g.dim[2], cb = 0;
        double lx = a.g.lo[0], ly = a.g.lo[1], lz = a.g.lo[2];
        if (batched) {
            int lo = 0, hi = a.nreg - 1;       // last region with pt_base <= apex
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (a.regions[mid].pt_base <= apex) lo = mid; else hi = mid - 1;
            }
            const RegionDesc& rd = a.regions[lo];
            nx = rd.dim[0]; ny = rd.dim[1]; nz = rd.dim[2];
            cb = rd.cell_base;
            shift = rd.fx_shift;
            lx = rd.lo[0]; ly = rd.lo[1]; lz = rd.lo[2];
            if (lo != cur_r) {
                if (lh && cur_r >= 0) {
                    drain(cur_r);
                    bound = 0;
                    __syncthreads();
                }
                cur_r = lo;
            }
        }
        const double4 Pa = a.pw[apex];
        const int cx = min(max((int)floor((Pa.x - lx) * inv_hx), 0), nx - 1);
        const int cy = min(max((int)floor((Pa.y - ly) * inv_h), 0), ny - 1);
        const int cz = min(max((int)floor((Pa.z - lz) * inv_h), 0), nz - 1);

        const unsigned long long cand0 = cand, over0 = over;   // restored if the apex is queued
        // ---- 1. gather the kept neighbours with their exact bin.  The stencil rows' point ranges are
        // read in parallel (one row per thread) into LDS and numbered by a prefix sum over the rows, so
        // the gather waits for one round of index loads, not one per row; the candidates of all rows
        // are then numbered consecutively and thread t takes numbers t, t + nt, ..., so every thread is
        // busy even where a row holds fewer points than the work-group has threads.
        for (int r = tid; r < a.nrows; r += nt) {
            const Row rw = a.rows[r];
            const int by = cy + rw.dy, bz = cz + rw.dz;
            const int bx0 = max(cx + rw.x0, 0), bx1 = min(cx + rw.x1, nx - 1);
            int js = 0, len = 0;
            if (by >= 0 && by < ny && bz >= 0 && bz < nz && bx0 <= bx1) {
                const int rb = cb + (bz * ny + by) * nx;
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
        int row = 0, row_beg = 0, row_end = a.nrows > 0 ? rw_len[0] : 0, j0 = a.nrows > 0 ? rw_js[0] : 0;
        {
            for (int next = tid; next < ncand; next += nt) {
                while (next >= row_end) {           // the row holding candidate number next
                    ++row;
                    row_beg = rw_off[row];
                    row_end = row_beg + rw_len[row];
                    j0 = rw_js[row] - row_beg;
                }
                const int j = j0 + next;
                ++cand;
                const double4 Pj = a.pw[j];
                const double s = exact_s(Pa.x, Pa.y, Pa.z, Pj.x, Pj.y, Pj.z);
                if (!(s > t_lo && s < t_hi)) continue;
                const float sf = (float)s;
                int b = (int)__builtin_amdgcn_fmed3f(bin_estimate(sf, LOGB, ua, ub), 0.0f, nbm1f);
                while (b > 0 && s < t64[b]) --b;
                while (b < nb && s >= t64[b + 1]) ++b;
                if (b == nb) { ++over; continue; }
                const uint32_t fj = a.flags[j];
                if constexpr (SYM) {
                    // every kept neighbour counts for the apex's C(k, 2) candidates; only those after the
                    // apex in the visiting order (buffer points, then larger index) enter its list
                    atomicAdd(&s_kall, 1);
                    if (!((fj & 2u) || j > apex)) continue;
                }
                const int pos = atomicAdd(&s_k, 1);
                if (pos >= kmax) { s_over = 1; continue; }
                // whole bytes for the flags let the pair loop add and subtract them with byte selects
                // (SDWA) instead of extracting bits
                uint32_t pk = __umul24((uint32_t)b, 2u * (uint32_t)nb) | ((fj & 1u) << 24);
                if constexpr (SYM) pk |= (fj & 2u) << 15;       // byte 2 = 1: buffer point (never an apex)
                if (nm > 1) {
                    int m, mr;
                    exact_mu(Pa.x, Pa.y, Pa.z, Pj.x, Pj.y, Pj.z, ms, nm, m, mr);   // apex -> j
                    pk |= (uint32_t)m << 16;
                }
                if (nm > 1 || W) ol[pos] = a.orig[j];
                if constexpr (FP64) {
                    rec64[pos] = make_double4(Pj.x, Pj.y, Pj.z, (double)pk);
                } else {
                    rec32[pos] = make_float4((float)(Pj.x - Pa.x), (float)(Pj.y - Pa.y), (float)(Pj.z - Pa.z),
                                             __uint_as_float(pk));
                }
                jl[pos] = j;
                if constexpr (W) wl[pos] = ldexp(Pj.w, shift);
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
        // pair index: K <= 4096 in LDS, so P < 2^31; a device-memory list may hold more
        using pidx_t = std::conditional_t<GEN == 2, long long, int>;
        const pidx_t P = (pidx_t)K * (K - 1) / 2;
        if (tid == 0) {
            const pidx_t kall = SYM ? (pidx_t)s_kall : (pidx_t)K;
            units += (unsigned long long)(kall * (kall - 1) / 2);   // the reference's apex-pair candidates
            evals += (unsigned long long)P;
        }
        const int oa = SYM && W ? a.orig[apex] : 0;     // the apex's original index
        const int ra = (int)(fa & 1u);
        const double wa = Pa.w;

        // ---- 2. neighbour pairs
        // The triangle u < v is folded into rows: folded row r (0 <= r < K/2) holds the K pairs
        // (r, r+1+c) for c < K-1-r and (K-2-r, c) for K-1-r <= c < K (row K-2-r's pairs); for
        // even K the last folded row is row r alone (K/2 pairs).  Neighbouring lanes read
        // neighbouring entries.
        // The u side of a pair (its record, original index and weight) is loaded by the caller, so
        // that a run of pairs sharing u (a row segment) loads it once and sweeps the v entries.
        using rec_t = std::conditional_t<FP64, double4, float4>;
        struct USide {
            rec_t R;
            int o;          // original index (RSD or weighted)
            double w;       // weight * 2^shift3 (weighted)
            uint32_t m3;    // SYM: 3 - (u is a buffer point), the multiplicity before v's flag
        };
        auto load_u = [&](const int u) {
            USide U;
            if constexpr (FP64) U.R = rec64[u]; else U.R = rec32[u];
            U.o = (nm > 1 || W) ? ol[u] : 0;
            U.w = W ? wl[u] : 0.0;
            if constexpr (FP64) U.m3 = 3u - (((uint32_t)U.R.w >> 16) & 0xffu);
            else U.m3 = 3u - ((__float_as_uint(U.R.w) >> 16) & 0xffu);
            return U;
        };
        // histogram update of the triangle (apex, u, v) whose third side has bin b3
        auto update = [&](const USide& U, const int v, const uint32_t pu, const uint32_t pv, const int b3) {
            if (equi && __umul24((uint32_t)b3, 2u * (uint32_t)nb) != (pu & 0x7fffu)) return;
            // the triple product in the reference's order, (w_i w_k1) w_k2 with k1 the neighbour of
            // smaller original index (gramsci.F90:482, :501), so that sums do not depend on the
            // order in which neighbours were gathered
            auto triple = [&]() {
                const double wv_ = wl[v];
                const bool uf = U.o < ol[v];
                return (wa * ldexp(uf ? U.w : wv_, -shift)) * (uf ? wv_ : U.w);
            };
            // 4 * configuration, from the table at byte offset 2 (bin_u nb^2 + bin_v nb + b3)
            const int cfg4 = equi ? 4 * b3
                                  : (int)*(const uint16_t*)((const unsigned char*)tab +
                                                            (__umul24(pu & 0x7fffu, (uint32_t)nb) + (pv & 0x7fffu) +
                                                             2u * (uint32_t)b3));
            int mu = 0;
            if (nm > 1) {
                const int mu_u = (int)((pu >> 16) & 127u), mu_v = (int)((pv >> 16) & 127u);
                mu = U.o < ol[v] ? ntab[mu_u * nm + mu_v] : ntab[mu_v * nm + mu_u];
            }
            const int cls = ra + (int)(pu >> 24) + (int)(pv >> 24);
            const int hk = cfg4 * nm + 4 * mu + cls;
            // the reference's count at each vertex that is not a buffer point (the apex always)
            const uint32_t bu = SYM ? (pu >> 16) & 0xffu : 0u, bv = SYM ? (pv >> 16) & 0xffu : 0u;
            const uint32_t mult = SYM ? U.m3 - bv : 1u;
            unsigned long long qs = 0;      // fixed-point sum of those vertices' triple products
            if constexpr (W) {
                auto q64 = [](double x) {
                    uint32_t lo, hi;
                    quantise(x, lo, hi);
                    return ((unsigned long long)hi << 32) + lo;
                };
                qs = q64(triple());
                if constexpr (SYM) {
                    // vertex p's product in the reference's order, (w_p w_x) w_y with x the other
                    // vertex of smaller original index; scaled weights W = w 2^shift, as for the apex
                    const double Wa = ldexp(wa, shift), Wv = wl[v];
                    const int ov = ol[v], ou = U.o;
                    if (!bu) qs += q64(oa < ov ? (ldexp(U.w, -shift) * wa) * Wv
                                               : (ldexp(U.w, -shift) * ldexp(Wv, -shift)) * Wa);
                    if (!bv) qs += q64(oa < ou ? (ldexp(Wv, -shift) * wa) * U.w
                                               : (ldexp(Wv, -shift) * ldexp(U.w, -shift)) * Wa);
                }
            }
            if (lh) {
                atomicAdd(&hc[(hk << rsh) + rep], mult);
                if constexpr (W) {
                    atomicAdd(&hw[2 * hk], qs & 0xffffffffull);
                    atomicAdd(&hw[2 * hk + 1], qs >> 32);
                }
            } else {
                const size_t go = (size_t)cur_r * nh + hk;
                atomicAdd(&a.out_cnt[go], (unsigned long long)mult);
                if constexpr (W) glb_add128(&a.out_w[2 * go], qs, 0ull);
            }
        };
        // the FP32 path calls update where the bin is known (the certain interval, or after the FP64
        // recheck), so no "no bin" sentinel is kept and tested at a join
        auto visit = [&](const USide& U, const int u, const int v) {
            uint32_t pu, pv;
            if constexpr (FP64) {
                int b3 = -1;
                const double4 Ru = U.R, Rv = rec64[v];
                pu = (uint32_t)Ru.w;
                pv = (uint32_t)Rv.w;
                if (!equi || ((pu ^ pv) & 0x7fffu) == 0) {
                    const double s = exact_s(Ru.x, Ru.y, Ru.z, Rv.x, Rv.y, Rv.z);
                    if (s > t_lo && s < t_hi) {
                        const float sf = (float)s;
                        int b = (int)__builtin_amdgcn_fmed3f(bin_estimate(sf, LOGB, ua, ubv), 0.0f, nbm1f);
                        while (b > 0 && s < t64[b]) --b;
                        while (b < nb && s >= t64[b + 1]) ++b;
                        if (b == nb) ++over; else b3 = b;
                    }
                }
                if (b3 >= 0) update(U, v, pu, pv, b3);
            } else {
                const float4 Ru = U.R, Rv = rec32[v];
                pu = __float_as_uint(Ru.w);
                pv = __float_as_uint(Rv.w);
                const float dx = Rv.x - Ru.x, dy = Rv.y - Ru.y, dz = Rv.z - Ru.z;
                const float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                if ((!equi || ((pu ^ pv) & 0x7fffu) == 0) && s > rej_lo && s < rej_hi) {
                    const int b = (int)__builtin_amdgcn_fmed3f(bin_estimate(s, LOGB, ua, ubv), 0.0f, nbm1f);
                    const float2 ok = okb[b];
                    if ((s >= ok.x) & (s < ok.y)) {
                        update(U, v, pu, pv, b);
                    } else {
                        ++rechk;
                        const double4 Pu = a.pw[jl[u]], Pv = a.pw[jl[v]];
                        const double se = exact_s(Pu.x, Pu.y, Pu.z, Pv.x, Pv.y, Pv.z);
                        if (se > t_lo && se < t_hi) {
                            const int be = exact_bin(se, t64, nb);
                            if (be == nb) ++over; else update(U, v, pu, pv, be);
                        }
                    }
                }
            }
        };
        const int R = K / 2;
        const int wv = tid >> 6, ln = tid & 63, nw = nt >> 6;
        if constexpr (GEN != 2) {
            // Two schedules, chosen per apex (K is uniform in the work-group):
            //  * rows, when they keep at least 7/8 of the lanes busy: wave w takes folded rows (groups of
            //    G = 64 / K rows for K <= 64) w, w + nw, ...; its lanes take columns l, l + 64, ...
            //    (lane l: row l / K, column l mod K for K <= 64).  Row interleaving balances the waves,
            //    and the index arithmetic is one add and one compare per step;
            //  * otherwise the flat schedule: the pair index pi = r K + c (0 <= pi < K(K-1)/2) starts at
            //    the thread index and advances by the work-group size, i.e. c by that mod K and r by its
            //    quotient, with at most one carry: no lane idles before the last step.
            // In both, neighbouring lanes read neighbouring LDS entries.
            const int G = (K > 0 && K <= 64) ? 64 / K : 1;
            const int passes = K > 64 ? (K + 63) >> 6 : 1;
            const bool by_rows = K > 64 ? K * 8 >= passes * 64 * 7 : G * K * 8 >= 64 * 7;
            if (by_rows) {
                const int rr = K <= 64 ? ln / K : 0;
                const int c0 = K <= 64 ? ln - rr * K : ln;
                const int grp = nw * G;
                if (rr < G)
                    for (int r = wv * G + rr; r < R; r += grp) {
                        // the row's two segments, each with its u loaded once: (r, r+1+c) for c < t,
                        // then (K-2-r, c) for t <= c < K (absent in an even K's last row)
                        const int t = K - 1 - r;
                        int c = c0;
                        if (c < t) {
                            const USide U = load_u(r);
                            for (; c < t; c += 64) visit(U, r, r + 1 + c);
                        }
                        if (c < K && 2 * r + 2 != K) {
                            const int u2 = K - 2 - r;
                            const USide U = load_u(u2);
                            for (; c < K; c += 64) visit(U, u2, c);
                        }
                    }
            } else {
                const int P = K * (K - 1) / 2;
                const int dq = nt / K, dr = nt - dq * K;
                int r = tid / K, c = tid - r * K;
                for (int pi = tid; pi < P; pi += nt) {
                    const bool first = c < K - 1 - r;
                    const int u = first ? r : K - 2 - r;
                    visit(load_u(u), u, first ? r + 1 + c : c);
                    c += dr;
                    const bool carry = c >= K;
                    c -= carry ? K : 0;
                    r += dq + (carry ? 1 : 0);
                }
            }
        } else {
            // Device-memory list (K beyond any LDS list): rows go to the waves in turn, wave w taking
            // folded rows w, w + nw, ..., its lanes columns l, l + 64, ... of the row.  So many pairs
            // alone could overflow a u32 LDS count: the rows run in blocks of fewer than 2^30 pairs,
            // with the LDS histogram drained after each block (all waves), and after the apex.
            const int blk = max(nw, (int)std::min<long long>((1ll << 30) / max(K, 1), (long long)R + nw) / nw * nw);
            for (int b0 = 0; b0 < R; b0 += blk) {
                if (b0 > 0) {
                    __syncthreads();
                    if (lh) drain(cur_r);
                    bound = 0;
                    __syncthreads();
                }
                const int b1 = min(R, b0 + blk);
                for (int r = b0 + wv; r < b1; r += nw) {
                    const int t = K - 1 - r;                // as in the LDS-list row schedule
                    int c = ln;
                    if (c < t) {
                        const USide U = load_u(r);
                        for (; c < t; c += 64) visit(U, r, r + 1 + c);
                    }
                    if (c < K && 2 * r + 2 != K) {
                        const int u2 = K - 2 - r;
                        const USide U = load_u(u2);
                        for (; c < K; c += 64) visit(U, u2, c);
                    }
                }
            }
        }
        bound += GEN == 2 ? (1ull << 31) + 1 : (unsigned long long)P * (SYM ? 3 : 1);   // LDS increments since the last drain
        __syncthreads();
        if (lh && bound > (1ull << 31)) {
            drain(cur_r);
            bound = 0;
            __syncthreads();
        }
    }
    __syncthreads();
    if (lh && cur_r >= 0) drain(cur_r);
    for (int off = 32; off > 0; off >>= 1) {
        units += __shfl_down(units, off, 64);
        evals += __shfl_down(evals, off, 64);
        rechk += __shfl_down(rechk, off, 64);
        over += __shfl_down(over, off, 64);
        cand += __shfl_down(cand, off, 64);
    }
    if (lane == 0) {
        if (cand) atomicAdd(&a.out_stats[0], cand);
        if (rechk) atomicAdd(&a.out_stats[1], rechk);
        if (over) atomicAdd(&a.out_stats[2], over);
        if (units) atomicAdd(&a.out_stats[3], units);
        if (evals) atomicAdd(&a.out_stats[4], evals);
    }
}

using T3Kernel = void (*)(TripArgs);

template <bool FP64, bool W, int MODE, bool LOGB>
static T3Kernel pick_t3m(int gen)
{
    return gen == 2 ? k_triplet3<FP64, W, MODE, 2, LOGB>
                    : (gen ? k_triplet3<FP64, W, MODE, 1, LOGB> : k_triplet3<FP64, W, MODE, 0, LOGB>);
}

template <bool FP64, bool W, bool LOGB>
static T3Kernel pick_t3w(const TripArgs& a)
{
    const int gen = a.glist ? 2 : ((a.nreg > 0 || !a.lds_hist) ? 1 : 0);
    if (a.equi)
        return a.nm > 1 ? pick_t3m<FP64, W, kT3Equi | kT3Rsd, LOGB>(gen) : pick_t3m<FP64, W, kT3Equi, LOGB>(gen);
    return a.nm > 1 ? pick_t3m<FP64, W, kT3Rsd, LOGB>(gen) : pick_t3m<FP64, W, 0, LOGB>(gen);
}

template <bool LOGB>
static T3Kernel pick_t3l(const TripArgs& a, bool fp64, bool weighted)
{
    if (fp64) return weighted ? pick_t3w<true, true, LOGB>(a) : pick_t3w<true, false, LOGB>(a);
    return weighted ? pick_t3w<false, true, LOGB>(a) : pick_t3w<false, false, LOGB>(a);
}

static T3Kernel pick_t3(const TripArgs& a, bool fp64, bool weighted)
{
    return a.log_bins ? pick_t3l<true>(a, fp64, weighted) : pick_t3l<false>(a, fp64, weighted);
}

static hipError_t launch_t3(T3Kernel kern, const TripArgs& a, size_t lds, int nblocks, hipStream_t st)
{
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(a.wg > 0 ? a.wg : kWG), lds, st, a);
    return hipGetLastError();
}

// Neighbour counts (rmin < d < rmax, exact FP64 distances) of the sampled apexes
// t * stride, t < nsample, over the full stencil of the catalogue's grid; -1 for a
// buffer point.  Sizes the first pass's neighbour list (one thread per apex).
// Kept-neighbour counts of sampled apexes (every stride-th point; -1 for a buffer point), for sizing the
// first pass's lists: one wave per sampled apex, its lanes over the apex's stencil candidates.
// after_only: count only the neighbours after the apex in the isotropic 3PCF's visiting order (buffer
// points, or a larger index), which are the ones its list holds.
__global__ __launch_bounds__(256) void k_nbr_count(TripArgs a, int stride, int nsample, int after_only,
                                                   int32_t* __restrict__ out)
{
    const int t = (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6);   // wave-uniform
    const int lane = threadIdx.x & 63;
    if (t >= nsample) return;
    const int apex = t * stride;
    if (a.flags[apex] & 2u) {
        if (lane == 0) out[t] = -1;
        return;
    }
    const BinDesc& B = *a.bins;
    const double t_lo = B.t64[0], t_hi = B.t64[B.nb + 1];
    const int nx = a.g.dim[0], ny = a.g.dim[1], nz = a.g.dim[2];
    const double inv_h = 1.0 / a.g.h, inv_hx = 1.0 / a.g.hx;
    const double4 Pa = a.pw[apex];
    const int cx = min(max((int)floor((Pa.x - a.g.lo[0]) * inv_hx), 0), nx - 1);
    const int cy = min(max((int)floor((Pa.y - a.g.lo[1]) * inv_h), 0), ny - 1);
    const int cz = min(max((int)floor((Pa.z - a.g.lo[2]) * inv_h), 0), nz - 1);
    int cnt = 0;
    for (int r = 0; r < a.nrows; ++r) {
        const Row rw = a.rows[r];
        const int by = cy + rw.dy, bz = cz + rw.dz;
        if (by < 0 || by >= ny || bz < 0 || bz >= nz) continue;
        const int bx0 = max(cx + rw.x0, 0), bx1 = min(cx + rw.x1, nx - 1);
        if (bx0 > bx1) continue;
        const int rb = (bz * ny + by) * nx;
        const int js = a.cell_start[rb + bx0], je = a.cell_start[rb + bx1 + 1];
        for (int j = js + lane; j < je; j += 64) {
            const double4 Pj = a.pw[j];
            const double s = exact_s(Pa.x, Pa.y, Pa.z, Pj.x, Pj.y, Pj.z);
            cnt += (s > t_lo && s < t_hi && (!after_only || j > apex || (a.flags[j] & 2u))) ? 1 : 0;
        }
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
    if (lane == 0) out[t] = cnt;
}

hipError_t launch_nbr_count(const TripArgs& a, int stride, int nsample, int after_only, int32_t* out, hipStream_t st)
{
    if (nsample <= 0) return hipSuccess;
    const long long threads = (long long)nsample * 64;
    hipLaunchKernelGGL(k_nbr_count, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, a, stride, nsample,
                       after_only, out);
    return hipGetLastError();
}

hipError_t launch_triplet3(const TripArgs& a, int nb, bool fp64, bool weighted, int nblocks, hipStream_t st)
{
    return launch_t3(pick_t3(a, fp64, weighted), a, triplet3_lds_bytes(a, nb, fp64, weighted), nblocks, st);
}

int triplet3_occupancy(const TripArgs& a, int nb, bool fp64, bool weighted)
{
    const size_t lds = triplet3_lds_bytes(a, nb, fp64, weighted);
    T3Kernel kern = pick_t3(a, fp64, weighted);
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return 1;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, a.wg > 0 ? a.wg : kWG, lds) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

}  // namespace gs

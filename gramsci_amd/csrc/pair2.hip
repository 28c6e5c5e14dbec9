// pair2.hip -- persistent 2PCF pair kernel (binned (r, mu) DD/DR/RR counts).
//
// Replaces create_graph (gramsci.F90:288-348) + query_graph_2pcf (:352-399):
// no neighbour graph is materialised.  Each unordered pair {i, j} is visited
// once; it contributes the reference's two ordered pairs i->j and j->i (each
// only if its apex is not a buffer point, :363), with mu bins m and nmu-1-m
// (theta(j->i) = -theta(i->j) exactly).
//
// Work decomposition.  Points are sorted by cell with x fastest, so an x-row
// of cells (fixed cy, cz) is one contiguous range.  A work item is a chunk of
// 64 consecutive points of one x-row, one point per lane of a wavefront.  The
// wave walks the half stencil of x-rows (dy, dz); for each it streams the
// contiguous range of cells [cx_first + x0, cx_last + x1] in tiles of 64
// points: in FP32 a register tile that rotates through the wave by DPP, in
// FP64 a wave-private LDS tile read in a rotated order.
// Waves are persistent and independent (no work-group barriers in the main
// loop); they pull items from 8 queues, one per XCD, each covering a
// contiguous z-slab of the grid so that an XCD's L2 sees one slab, and steal
// from the other queues when their own is empty.
//
// Precision variants (identical counts):
//   FP64        s = |xj-xi|^2 in FP64 exactly as the reference; bins from the
//               exact thresholds; mu with IEEE sqrt/div.
//   FP32 guard  d = (xj_l - xi_l) + (cx_j - cx_i) * h from cell-local FP32
//               coordinates (h has <= 11 significant bits, so the cell offset
//               is exact).  A pair is binned directly only when s32 lies in the
//               bin's certain interval [ok_lo, ok_hi) and its mu is farther
//               than a rigorous error bound from a mu edge; otherwise it is
//               re-evaluated with the FP64 path.
//
// (r, mu) without buffer points: every pair is counted in both directions and
// theta(j->i) = -theta(i->j) exactly, so one update at mu bin m stands for
// both ordered pairs (m and nm-1-m) and the drain adds it to both bins.  Only
// pairs exactly on a mu edge (where the two directions are not mirror images)
// are added explicitly.
//
// Histograms live in LDS, shared by the work-group's waves: u32 counts
// replicated `reps` times (replica = lane % reps, one plane of nbt counts per
// replica) to spread same-bin atomics,
// and per bin two u64 sums of the weight product's fixed-point value split at
// bit 32 (no carries, so no returning atomics).  A wave that has issued 2^29
// increments drains the histogram with atomic exchanges into the global
// 128-bit accumulators; the rest is drained when the work-group ends.
#include <cstdlib>

#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

namespace {

using namespace dev;

struct Stats {
    unsigned long long cand, rechk, over;
};

// One pair's histogram update.  base = class offset + b*nm (flat bin index
// without the mu bin); counts live at hcr[base + m], hcr = this lane's replica plane.  mirrored:
// mr == nm-1-m (always so for the FP32 mu bins; tested after an FP64 recheck).
// Returns the number of LDS bin increments made.
template <bool RSD, bool W>
__device__ __forceinline__ int accumulate(uint32_t* hcr, unsigned long long* hw, int nm,
                                          bool mirror, int base, int m, int mr, bool mirrored, int ci, int cj,
                                          double x, unsigned long long* out_cnt, unsigned long long* out_w)
{
    uint32_t qlo = 0, qhi = 0;
    if constexpr (W) {
        const unsigned long long q = quantise_rn(x);
        qlo = (uint32_t)q;
        qhi = (uint32_t)(q >> 32);
    }
    auto add = [&](int k, uint32_t c) {
        atomicAdd(&hcr[k], c);
        if constexpr (W) {
            atomicAdd(&hw[2 * k], (unsigned long long)qlo * c);
            atomicAdd(&hw[2 * k + 1], (unsigned long long)qhi * c);
        }
    };
    if constexpr (!RSD) {
        const int c = ci + cj;
        if (c) add(base, (uint32_t)c);
        return c;
    } else if (mirror && mirrored) {
        add(base + m, 1u);    // mirror-implied: stands for i->j at m and j->i at nm-1-m
        return 1;
    } else if (mirror) {
        // a pair exactly on a mu edge: both ordered pairs explicitly, straight to global memory (rare)
        const unsigned long long L = ((unsigned long long)qhi << 32) + qlo;
        atomicAdd(&out_cnt[base + m], 1ull);
        atomicAdd(&out_cnt[base + mr], 1ull);
        if constexpr (W) {
            glb_add128(&out_w[2 * (base + m)], L, 0ull);
            glb_add128(&out_w[2 * (base + mr)], L, 0ull);
        }
        return 0;
    } else {
        if (ci) add(base + m, 1u);
        if (cj) add(base + mr, 1u);
        return ci + cj;
    }
}

// Drains the LDS histogram into the global accumulators with atomic
// exchanges, so any wave may do it while the others keep adding.  Called by
// all lanes of one wave (stride 64) or of the whole work-group.
// With `mirror`, entry (cls, b, m) also stands for (cls, b, nm-1-m).
__device__ void drain(uint32_t* hc, unsigned long long* hw, int nbt, int nm, int reps, bool weighted, bool mirror,
                      int t0, int stride, unsigned long long* out_cnt, unsigned long long* out_w)
{
    for (int k = t0; k < nbt; k += stride) {
        const int km = mirror ? (k - k % nm) + (nm - 1 - k % nm) : k;
        unsigned long long s = 0;
        for (int r = 0; r < reps; ++r) s += atomicExch(&hc[k + r * nbt], 0u);
        if (s) {
            atomicAdd(&out_cnt[k], s);
            if (mirror) atomicAdd(&out_cnt[km], s);
        }
        if (weighted) {
            const unsigned long long lo = atomicExch(&hw[2 * k], 0ull);
            const unsigned long long hi = atomicExch(&hw[2 * k + 1], 0ull);
            if (lo | hi) {
                // value = hi * 2^32 + lo as a 128-bit number
                const unsigned long long L = (hi << 32) + lo;
                const unsigned long long H = (hi >> 32) + (L < lo ? 1ull : 0ull);
                glb_add128(&out_w[2 * k], L, H);
                if (mirror) glb_add128(&out_w[2 * km], L, H);
            }
        }
    }
}


}  // namespace

size_t pair2_lds_bytes(int nb, int nm, int reps, bool fp64, bool weighted, bool batched, int wg)
{
    const size_t nbt = (size_t)3 * nb * nm;
    const size_t copies = batched ? (size_t)(wg / kWave) : 1;   // batched: one histogram per wave
    size_t b = 0;
    b += ((nbt * reps * 4 * copies + 15) / 16) * 16;          // counts
    if (weighted) b += nbt * 16 * copies;                     // split fixed-point sums
    b += ((size_t)(kMaxBins + 1) * 8);                        // ok_lo/ok_hi as float2
    b += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;         // t64
    const int waves = wg / kWave;
    if (fp64) b += (size_t)waves * kWave * (24 + 8 + 4 + 4);  // LDS tile: x, y, z, w, flags, pad
    return b;
}

// a lane's run of histogram increments to one bin, merged in registers
struct RunCache {
    int ck;
    uint32_t cc;
    unsigned long long cq;
};

// The unweighted FP32 kernels are compiled for six waves per SIMD (at most 80 VGPRs), the occupancy
// their LDS use allows; the weighted ones are held to fewer waves by their LDS sums anyway.
// DUAL (weighted FP32 only): each lane also holds the chunk point 32 lanes away, and a full tile is
// done in 32 steps of two pairs each (below).  It needs about 127 VGPRs, so it is launched only where
// that costs no resident waves.
template <bool FP64, bool RSD, bool W, bool LOGB, bool DUAL>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu((!FP64 && !W) ? 6 : 1))) void k_pair2(PairArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BinDesc& B = *a.bins;
    const int nb = B.nb;
    const int nm = RSD ? B.nm : 1;
    const int nbt = 3 * nb * nm;
    const bool batched = a.nreg > 0;
    const int reps = batched ? 1 : a.reps;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid / kWave;
    const int rep = lane & (reps - 1);
    const bool mirror = RSD && a.mirror;
    const int nthr = blockDim.x;        // a multiple of 64, at most 1024
    const int copies = batched ? nthr / kWave : 1;

    unsigned char* p = smem;
    uint32_t* hc0 = (uint32_t*)p;
    p += ((size_t)nbt * reps * 4 * copies + 15) / 16 * 16;
    unsigned long long* hw0 = (unsigned long long*)p;
    if (W) p += (size_t)nbt * 16 * copies;
    // the histogram this wave adds to: shared by the work-group, or wave-private in batched mode
    uint32_t* hc = batched ? hc0 + (size_t)wv * nbt : hc0;
    unsigned long long* hw = batched ? hw0 + (size_t)wv * 2 * nbt : hw0;
    uint32_t* hcr = hc + rep * nbt;      // this lane's replica plane of the counts
    float2* okb = (float2*)p;
    p += (size_t)(kMaxBins + 1) * 8;
    double* t64 = (double*)p;
    p += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;
    // FP64: wave-private LDS tile x[64], y[64], z[64], w[64], flags[64]
    unsigned char* tb = p + (size_t)wv * kWave * 40;
    double* tx = (double*)tb;
    double* ty = tx + kWave;
    double* tz = ty + kWave;
    double* tw = tz + kWave;
    uint32_t* tf = (uint32_t*)(tw + kWave);

    for (int k = tid; k < nbt * reps * copies; k += nthr) hc0[k] = 0;
    if (W)
        for (int k = tid; k < 2 * nbt * copies; k += nthr) hw0[k] = 0;
    for (int k = tid; k < nb; k += nthr) okb[k] = make_float2(B.ok_lo[k], B.ok_hi[k]);
    for (int k = tid; k < nb + 2; k += nthr) t64[k] = B.t64[k];
    __syncthreads();

    const float rej_lo = B.rej_lo, rej_hi = B.rej_hi, ua = B.ua, ub = B.ub;
    const double t_lo = B.t64[0], t_hi = B.t64[nb + 1];
    const double ms = B.mu_scale;
    const int nbm = nb * nm;
    const float ms32 = B.ms32, mu_ew = B.mu_ew, mu_em = B.mu_em, mu_c = B.mu_c;
    const float nbm1f = (float)(nb - 1), nmm1f = (float)(nm - 1);
    // mu edges are the integers 1..nm-1 of t; with nm = 127 also 0 and 127, where the reference's
    // int8 store of floor(t)+1 wraps (gramsci.F90:326-328), so pairs there get the FP64 recheck
    const float e_lo = nm == 127 ? 0.0f : 1.0f, e_hi = nm == 127 ? 127.0f : nmm1f;
    // mu error bound dt = ms * (mu_ew / |sep| + mu_em / |mid| + mu_c), with ms folded into the
    // constants (their 4x margin dwarfs the reassociation).  In the weighted kernel, whose
    // residency LDS limits, dt_c and ub are kept in VGPRs so that each FMA of the loop reads at
    // most one scalar operand; elsewhere the extra VGPRs would cost residency.
    const float dt_ew = ms32 * mu_ew, dt_em = ms32 * mu_em;
    float dt_c = ms32 * mu_c, ubv = ub;
    if constexpr (W) asm volatile("" : "+v"(dt_c), "+v"(ubv));
    // fixed-point scales of data and random weights (batched: the current region's)
    int sd = B.fx_sd, sr = B.fx_sr;
    const int nrow_all = batched ? 0 : a.g.dim[1] * a.g.dim[2];
    const int n_chunks = batched ? *a.n_chunks_dev : a.n_chunks;
    const float h32 = a.g.h32, hx32 = a.g.hx32;    // cell edges in y/z and in x
    int cur_r = -1;                 // region of this wave's private histogram (batched mode)

    Stats st{0, 0, 0};
    // FP32 weighted path: per-lane run cache of the last histogram bin.  A lane's consecutive
    // neighbours (adjacent points of one cell, data before randoms) often fall in the same
    // bin; their updates are merged in registers and reach the LDS histogram with one set of
    // atomics when the bin changes, which relieves same-address LDS atomic conflicts.
    // The run's sum is one 64-bit integer: products are below 2^52 (fx_sd, fx_sr) and a run ends
    // after at most 33 increments, so it cannot overflow; the flush splits it at bit 32 into
    // the two LDS halves, which keeps value = hi * 2^32 + lo.  The update is written with
    // selects, not branches, so that no register copies are needed where paths merge.  An empty
    // run (cc = 0) stays attached to its last bin (bin 0 at the start): flushing it adds zeros, so
    // the flush needs no test for a missing run.
    static_assert(!DUAL || (W && !FP64), "two apexes per lane: weighted FP32 kernels only");
    // with DUAL each lane's second apex has a run cache of its own
    RunCache rc0{0, 0u, 0ull}, rc1{0, 0u, 0ull};
    auto flush_one = [&](RunCache& c) {
        atomicAdd(&hcr[c.ck], c.cc);
        if constexpr (W) {
            atomicAdd(&hw[2 * c.ck], c.cq & 0xffffffffull);
            atomicAdd(&hw[2 * c.ck + 1], c.cq >> 32);
        }
        c.cc = 0;
        c.cq = 0;
    };
    auto flush = [&]() {
        flush_one(rc0);
        if constexpr (DUAL) flush_one(rc1);
    };
    auto put = [&](RunCache& rc, int k, uint32_t c, unsigned long long q) {   // c in {1, 2}
        const bool nw = (k != rc.ck) || (rc.cc >= 32u);
        if (nw) flush_one(rc);
        rc.cc = (nw ? 0u : rc.cc) + c;
        if constexpr (W) rc.cq = (nw ? 0ull : rc.cq) + (q << (c - 1));
        rc.ck = k;
    };
    unsigned long long bound = 0;   // LDS increments issued by this wave since its last drain
    // LDS bin increments between drains, over all waves sharing the histogram, stay below 2^31,
    // which keeps every u32 count and every split sum from overflowing
    const unsigned long long drain_at = (1ull << 31) / (unsigned long long)(batched ? 1 : nthr / kWave);
    const int xcd = blockIdx.x & 7;
    int tried = 0;

    while (true) {
        int item = -1;
        if (lane == 0) item = next_item(a.queue, n_chunks, xcd, tried);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item < 0) break;

        int nx, ny, nz, cb, row;
        const int grow = upper_index(a.row_chunk, batched ? a.row_base[a.nreg] : nrow_all, item);
        const int q = item - a.row_chunk[grow];
        if (batched) {
            const int R = upper_index(a.row_base, a.nreg, grow);
            const RegionDesc& rd = a.regions[R];
            nx = rd.dim[0]; ny = rd.dim[1]; nz = rd.dim[2];
            cb = rd.cell_base;
            row = grow - rd.row_base;
            sd = rd.fx_shift;
            sr = rd.fx_shift_rand;
            if (R != cur_r) {
                if constexpr (W) flush();
                if (cur_r >= 0)
                    drain(hc, hw, nbt, nm, reps, W, mirror, lane, kWave, a.out_cnt + (size_t)cur_r * nbt,
                          a.out_w + (size_t)cur_r * 2 * nbt);
                cur_r = R;
                bound = 0;
            }
        } else {
            nx = a.g.dim[0]; ny = a.g.dim[1]; nz = a.g.dim[2];
            cb = 0;
            row = grow;
        }
        const int cy = row % ny, cz = row / ny;
        const int rs = a.cell_start[cb + row * nx], re = a.cell_start[cb + (row + 1) * nx];
        const int i0 = rs + q * kChunk;
        const int n_i = min(kChunk, re - i0);
        const int i = i0 + lane;
        const bool vi = lane < n_i;
        const int cx_first = (int)a.pl[i0].w;
        const int cx_last = (int)a.pl[i0 + n_i - 1].w;

        double4 Pi = make_double4(0, 0, 0, 0);
        float4 Li = make_float4(0, 0, 0, 0);
        uint32_t fi = 2u;
        if (vi) {
            Pi = a.pw[i];
            Li = a.pl[i];
            fi = a.flags[i];
        }
        const int ri = fi & 1, ci = (fi & 2) ? 0 : 1;
        // the lane's second apex: the chunk point 32 lanes away
        const int i2 = i0 + (lane ^ 32);
        const bool vi2 = DUAL && (lane ^ 32) < n_i;
        double4 Pi2 = make_double4(0, 0, 0, 0);
        float4 Li2 = make_float4(0, 0, 0, 0);
        uint32_t fi2 = 2u;
        if (vi2) {
            Pi2 = a.pw[i2];
            Li2 = a.pl[i2];
            fi2 = a.flags[i2];
        }
        const int ri2 = fi2 & 1, ci2 = (fi2 & 2) ? 0 : 1;
        // a chunk of buffer points only (never apexes) gains nothing from a tile of buffer points only
        const bool chunk_buf = __ballot(ci != 0) == 0;
        // the apex weight scaled by its class's 2^F (exact): a pair's product w_i' w_j' then carries
        // the scale of its class (DD, DR or RR)
        const double wis = ldexp(Pi.w, ri ? sr : sd);
        const float ax32 = (float)Pi.x, ay32 = (float)Pi.y, az32 = (float)Pi.z;   // absolute, for mu
        const double wis2 = DUAL ? ldexp(Pi2.w, ri2 ? sr : sd) : 0.0;
        const float ax2 = (float)Pi2.x, ay2 = (float)Pi2.y, az2 = (float)Pi2.z;
        const int cb02 = ri2 ? nbm : 0;
        // histogram class offsets of this lane's pairs with a data / a random neighbour
        const int cb0 = ri ? nbm : 0, cb1 = cb0 + nbm;

        for (int r = 0; r < a.nrows; ++r) {
            const Row rw = a.rows[r];
            const int by = cy + rw.dy, bz = cz + rw.dz;
            if (by < 0 || by >= ny || bz < 0 || bz >= nz) continue;
            const int bx0 = max(cx_first + rw.x0, 0), bx1 = min(cx_last + rw.x1, nx - 1);
            if (bx0 > bx1) continue;
            const bool self = (rw.dy == 0 && rw.dz == 0);
            const int rb = cb + (bz * ny + by) * nx;
            int js = a.cell_start[rb + bx0];
            const int je = a.cell_start[rb + bx1 + 1];
            if (self) js = max(js, i0 + 1);
            // lane position relative to the neighbour row's cell origin in y and z
            const float yi_r = Li.y - (float)rw.dy * h32;
            const float zi_r = Li.z - (float)rw.dz * h32;
            const float yi_r2 = Li2.y - (float)rw.dy * h32;
            const float zi_r2 = Li2.z - (float)rw.dz * h32;
            for (int jt = js; jt < je; jt += kWave) {
                const int nt = min(kWave, je - jt);
                const int jmin = max(self ? i + 1 - jt : 0, 0);   // self row: only j > i
                const int jmin2 = max(self ? i2 + 1 - jt : 0, 0);
                if (vi) st.cand += (unsigned long long)(nt - min(jmin, nt));
                bound += (unsigned long long)nt * 2ull * kWave;
                // histogram update of one kept pair with neighbour flags fj, bin b, mu bins m / mr
                auto finish = [&](uint32_t fj, double wj, int b, int m, int mr) {
                    const int cj = (fj & 2) ? 0 : 1;
                    const int base = ((fj & 1) ? cb1 : cb0) + __umul24(b, nm);
                    double x = 0.0;
                    if constexpr (W) x = wis * wj;
                    accumulate<RSD, W>(hcr, hw, nm, mirror, base, m, mr, mr == nm - 1 - m, ci, cj, x,
                                       a.out_cnt, a.out_w);
                };
                if constexpr (FP64) {
                    wave_sync();
                    uint32_t fl = 2u;
                    if (lane < nt) {
                        const double4 Pj = a.pw[jt + lane];
                        fl = a.flags[jt + lane];
                        tx[lane] = Pj.x; ty[lane] = Pj.y; tz[lane] = Pj.z;
                        if (W) tw[lane] = ldexp(Pj.w, (fl & 1u) ? sr : sd);
                        tf[lane] = fl;
                    }
                    const bool skip = chunk_buf && __ballot((fl & 2u) == 0) == 0;
                    wave_sync();
                    // rotated schedule: at step k lane L takes tile entry (k + L) mod nt, so the
                    // lanes of a wave pair with different neighbours (spreading their histogram
                    // updates over many bins) while reading consecutive, conflict-free LDS words
                    const int kend = (vi && !skip) ? nt : 0;
                    const int lm = lane % nt;
                    for (int k = 0; k < kend; ++k) {
                        int kk = k + lm;
                        kk = kk >= nt ? kk - nt : kk;
                        if (kk < jmin) continue;
                        const uint32_t fj = tf[kk];
                        const double jx = tx[kk], jy = ty[kk], jz = tz[kk];
                        const double s = exact_s(Pi.x, Pi.y, Pi.z, jx, jy, jz);
                        if (!(s > t_lo && s < t_hi)) continue;
                        const int b = adjust_bin(s, bin_estimate((float)s, LOGB, ua, ub), t64, nb);
                        if (b == nb) { st.over += (uint64_t)(ci + ((fj & 2) ? 0 : 1)); continue; }
                        int m = 0, mr = 0;
                        if constexpr (RSD) exact_mu(Pi.x, Pi.y, Pi.z, jx, jy, jz, ms, nm, m, mr);
                        finish(fj, W ? tw[kk] : 0.0, b, m, mr);
                    }
                } else {
                    {
                        // Register tile: lane L loads entry L; the entries then rotate through the
                        // wave one lane per step (DPP), so after 64 steps every lane has met every
                        // entry (after P steps for a tile of nt <= P entries, below).  Consecutive
                        // lanes meet different neighbours at each step, which spreads their histogram
                        // updates over many bins, and no LDS is read for the neighbours.
                        // A partial tile is loaded with period P (the power of two >= nt): lane L
                        // holds entry L mod P, a full-wave rotation keeps that pattern, and P steps
                        // bring every entry to every lane.
                        const int P = nt <= 1 ? 1 : 1 << (32 - __clz(nt - 1));
                        const int le = lane & (P - 1);
                        const bool has = le < nt;
                        const int jl = jt + (has ? le : 0);
                        uint32_t ef = a.flags[jl];
                        const bool skip = chunk_buf && __ballot(has && (ef & 2u) == 0) == 0;
                        float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
                        double ewj = 0.0;
                        if (!skip) {
                            e = a.pl[jl];
                            if constexpr (W) ewj = ldexp(a.ws[jl], (ef & 1u) ? sr : sd);
                        }
                        // entry index, class and buffer flag in one rotated register (-1: no entry):
                        // ep = eidx << 16 | (class offset of a random entry, nb*nm, or 0) << 1 | buffer,
                        // so the self-row test eidx >= jmin is ep >= jmin << 16 and the pair's class
                        // offset is cb0 + a field of ep (nb*nm <= 127*127 fits in 15 bits)
                        int ep = has ? (le << 16) | ((ef & 1u) ? nbm << 1 : 0) | (int)((ef >> 1) & 1u) : -1;
                        if constexpr (!DUAL) {
                            const int jmin16 = jmin << 16;
                            const int kend = __builtin_amdgcn_readfirstlane(skip ? 0 : P);
                            for (int k = 0; k < kend; ++k) {
                                if (vi && ep >= jmin16) do {
                                    const float dx = fmaf(e.w - Li.w, hx32, e.x - Li.x);
                                    const float dy = e.y - yi_r, dz = e.z - zi_r;
                                    const float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                                    if (!(s > rej_lo && s < rej_hi)) break;
                                    int b = (int)__builtin_amdgcn_fmed3f(bin_estimate(s, LOGB, ua, ubv), 0.0f, nbm1f);
                                    const float2 ok = okb[b];
                                    bool safe = (s >= ok.x) && (s < ok.y);
                                    int m = 0, mr;      // mr is read only after an FP64 recheck has set it
                                    if constexpr (!RSD) mr = 0;
                                    if constexpr (RSD) {
                                        const float r = __builtin_amdgcn_rsqf(s);
                                        // mid = x_i + d/2 (absolute), sep = x_i - x_j = -d
                                        const float mx = fmaf(0.5f, dx, ax32), my = fmaf(0.5f, dy, ay32), mz = fmaf(0.5f, dz, az32);
                                        const float D = -fmaf(mz, dz, fmaf(my, dy, mx * dx));
                                        const float M2 = fmaf(mz, mz, fmaf(my, my, mx * mx));
                                        const float rm = __builtin_amdgcn_rsqf(M2);
                                        const float t = fmaf(D * rm * r, ms32, ms32);     // (theta + 1) * mu_scale
                                        const float dt = fmaf(dt_ew, r, fmaf(dt_em, rm, dt_c));
                                        // mu edges sit at the integers 1..nm-1 of t: the bin is certain when the
                                        // nearest edge, clamp(rint(t), 1, nm-1), is farther than the error bound
                                        // dt < 1/2 (a nearest integer of 0 or nm is more than 1/2 from any edge)
                                        const float edge = __builtin_amdgcn_fmed3f(rintf(t), e_lo, e_hi);
                                        safe = safe && (dt < 0.5f) && (fabsf(t - edge) >= dt);
                                        m = min((int)t, nm - 1);       // t >= 0 up to rounding; truncation = floor there
                                    }
                                    bool mirrored = true;     // the FP32 mu bins are mirror images by construction
                                    if (!safe) {
                                        ++st.rechk;
                                        const double4 Pj = a.pw[jt + (ep >> 16)];
                                        Pi = a.pw[i];
                                        const double se = exact_s(Pi.x, Pi.y, Pi.z, Pj.x, Pj.y, Pj.z);
                                        if (!(se > t_lo && se < t_hi)) break;
                                        b = exact_bin(se, t64, nb);
                                        if (b == nb) { st.over += (uint64_t)(ci + (ep & 1 ? 0 : 1)); break; }
                                        if constexpr (RSD) {
                                            exact_mu(Pi.x, Pi.y, Pi.z, Pj.x, Pj.y, Pj.z, ms, nm, m, mr);
                                            mirrored = mr == nm - 1 - m;
                                        }
                                    }
                                    {
                                        const int cj = (ep & 1) ? 0 : 1;
                                        const int base = cb0 + ((ep >> 1) & 0x7fff) + __umul24(b, nm);
                                        unsigned long long q = 0;
                                        if constexpr (W) q = quantise_rn(wis * ewj);
                                        if constexpr (W && !RSD) {
                                            if (ci + cj) put(rc0, base, (uint32_t)(ci + cj), q);
                                        } else if (W && mirror && mirrored) {
                                            put(rc0, base + m, 1u, q);   // stands for i->j at m and j->i at nm-1-m
                                        } else {
                                            // mr: the mirror image of m unless the recheck found otherwise
                                            accumulate<RSD, W>(hcr, hw, nm, mirror, base, m,
                                                               mirrored ? nm - 1 - m : mr, mirrored, ci, cj,
                                                               wis * ewj, a.out_cnt, a.out_w);
                                        }
                                    }
                                } while (false);
                                // rotate the register tile by one lane (all lanes active here)
                                e.x = rot_lane(e.x); e.y = rot_lane(e.y); e.z = rot_lane(e.z); e.w = rot_lane(e.w);
                                if constexpr (W) ewj = rot_lane(ewj);
                                ep = rot_lane(ep);
                            }
                        } else {
                            const int kend = __builtin_amdgcn_readfirstlane(skip ? 0 : P);
                            // one apex's pair with the lane's current tile entry
                            auto visit = [&](const bool va, const int ia, const float4 La, const float yr, const float zr,
                                             const float axa, const float aya, const float aza, const int cia,
                                             const int cba, const double wisa, const int jm16,
                                             RunCache& rc) __attribute__((always_inline)) {
                                if (va && ep >= jm16) do {
                                    const float dx = fmaf(e.w - La.w, hx32, e.x - La.x);
                                    const float dy = e.y - yr, dz = e.z - zr;
                                    const float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                                    if (!(s > rej_lo && s < rej_hi)) break;
                                    int b = (int)__builtin_amdgcn_fmed3f(bin_estimate(s, LOGB, ua, ubv), 0.0f, nbm1f);
                                    const float2 ok = okb[b];
                                    bool safe = (s >= ok.x) && (s < ok.y);
                                    int m = 0, mr;      // mr is read only after an FP64 recheck has set it
                                    if constexpr (!RSD) mr = 0;
                                    if constexpr (RSD) {
                                        const float r = __builtin_amdgcn_rsqf(s);
                                        // mid = x_i + d/2 (absolute), sep = x_i - x_j = -d
                                        const float mx = fmaf(0.5f, dx, axa), my = fmaf(0.5f, dy, aya), mz = fmaf(0.5f, dz, aza);
                                        const float D = -fmaf(mz, dz, fmaf(my, dy, mx * dx));
                                        const float M2 = fmaf(mz, mz, fmaf(my, my, mx * mx));
                                        const float rm = __builtin_amdgcn_rsqf(M2);
                                        const float t = fmaf(D * rm * r, ms32, ms32);     // (theta + 1) * mu_scale
                                        const float dt = fmaf(dt_ew, r, fmaf(dt_em, rm, dt_c));
                                        // mu edges sit at the integers 1..nm-1 of t: the bin is certain when the
                                        // nearest edge, clamp(rint(t), 1, nm-1), is farther than the error bound
                                        // dt < 1/2 (a nearest integer of 0 or nm is more than 1/2 from any edge)
                                        const float edge = __builtin_amdgcn_fmed3f(rintf(t), e_lo, e_hi);
                                        safe = safe && (dt < 0.5f) && (fabsf(t - edge) >= dt);
                                        m = min((int)t, nm - 1);       // t >= 0 up to rounding; truncation = floor there
                                    }
                                    bool mirrored = true;     // the FP32 mu bins are mirror images by construction
                                    if (!safe) {
                                        ++st.rechk;
                                        const double4 Pj = a.pw[jt + (ep >> 16)];
                                        const double4 Pa = a.pw[ia];
                                        const double se = exact_s(Pa.x, Pa.y, Pa.z, Pj.x, Pj.y, Pj.z);
                                        if (!(se > t_lo && se < t_hi)) break;
                                        b = exact_bin(se, t64, nb);
                                        if (b == nb) { st.over += (uint64_t)(cia + (ep & 1 ? 0 : 1)); break; }
                                        if constexpr (RSD) {
                                            exact_mu(Pa.x, Pa.y, Pa.z, Pj.x, Pj.y, Pj.z, ms, nm, m, mr);
                                            mirrored = mr == nm - 1 - m;
                                        }
                                    }
                                    {
                                        const int cj = (ep & 1) ? 0 : 1;
                                        const int base = cba + ((ep >> 1) & 0x7fff) + __umul24(b, nm);
                                        unsigned long long q = 0;
                                        if constexpr (W) q = quantise_rn(wisa * ewj);
                                        if constexpr (W && !RSD) {
                                            if (cia + cj) put(rc, base, (uint32_t)(cia + cj), q);
                                        } else if (W && mirror && mirrored) {
                                            put(rc, base + m, 1u, q);   // stands for i->j at m and j->i at nm-1-m
                                        } else {
                                            // mr: the mirror image of m unless the recheck found otherwise
                                            accumulate<RSD, W>(hcr, hw, nm, mirror, base, m,
                                                               mirrored ? nm - 1 - m : mr, mirrored, cia, cj,
                                                               wisa * ewj, a.out_cnt, a.out_w);
                                        }
                                    }
                                } while (false);
                            };
                            auto rotate = [&]() __attribute__((always_inline)) {
                                // rotate the register tile by one lane (all lanes active here)
                                e.x = rot_lane(e.x); e.y = rot_lane(e.y); e.z = rot_lane(e.z); e.w = rot_lane(e.w);
                                if constexpr (W) ewj = rot_lane(ewj);
                                ep = rot_lane(ep);
                            };
                            if (kend == kWave) {
                                // a full tile in 32 steps: lane L meets 32 cyclically consecutive entries from
                                // its own, lane L ^ 32 the other 32, and each pairs them with both apexes,
                                // points L and L ^ 32, so every (apex, entry) pair is visited exactly once
                                // and each rotation serves two pairs
                                for (int k = 0; k < kWave / 2; ++k) {
                                    visit(vi, i, Li, yi_r, zi_r, ax32, ay32, az32, ci, cb0, wis, jmin << 16, rc0);
                                    visit(vi2, i2, Li2, yi_r2, zi_r2, ax2, ay2, az2, ci2, cb02, wis2, jmin2 << 16, rc1);
                                    rotate();
                                }
                            } else {
                                for (int k = 0; k < kend; ++k) {
                                    visit(vi, i, Li, yi_r, zi_r, ax32, ay32, az32, ci, cb0, wis, jmin << 16, rc0);
                                    rotate();
                                }
                            }
                        }
                    }
                }
                // each wave drains after drain_at of its own possible increments (every candidate
                // counted twice; the increments held in the lanes' run caches included)
                if (bound > drain_at) {
                    if constexpr (W) flush();
                    const size_t ro = batched ? (size_t)cur_r : 0;
                    drain(hc, hw, nbt, nm, reps, W, mirror, lane, kWave, a.out_cnt + ro * nbt,
                          a.out_w + ro * 2 * nbt);
                    bound = 0;
                }
            }
        }
    }
    if constexpr (W) flush();
    if (batched) {
        if (cur_r >= 0)
            drain(hc, hw, nbt, nm, reps, W, mirror, lane, kWave, a.out_cnt + (size_t)cur_r * nbt,
                  a.out_w + (size_t)cur_r * 2 * nbt);
    } else {
        __syncthreads();
        drain(hc, hw, nbt, nm, reps, W, mirror, tid, nthr, a.out_cnt, a.out_w);
    }

    // wave-reduce the statistics
    for (int off = 32; off > 0; off >>= 1) {
        st.cand += __shfl_down(st.cand, off, 64);
        st.rechk += __shfl_down(st.rechk, off, 64);
        st.over += __shfl_down(st.over, off, 64);
    }
    if (lane == 0) {
        if (st.cand) atomicAdd(&a.out_stats[0], st.cand);
        if (st.rechk) atomicAdd(&a.out_stats[1], st.rechk);
        if (st.over) atomicAdd(&a.out_stats[2], st.over);
    }
}

template <bool FP64, bool RSD, bool W, bool LOGB, bool DUAL>
static int occ_k(size_t lds, int wg)
{
    int nb = 0;
    auto kern = k_pair2<FP64, RSD, W, LOGB, DUAL>;
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, wg, lds) != hipSuccess)
        return 1;
    return nb;
}

// Whether the weighted FP32 kernel runs with two apexes per lane: where the LDS histogram already limits
// the resident work-groups to what the larger register file allows (many bins), else not.
// GS_PAIR_DUAL=0 / 1 forces it off / on (tests, tuning).
template <bool RSD, bool LOGB>
static bool use_dual(size_t lds, int wg)
{
    const char* ev = getenv("GS_PAIR_DUAL");
    if (ev && *ev) return atoi(ev) != 0;
    return occ_k<false, RSD, true, LOGB, true>(lds, wg) >= occ_k<false, RSD, true, LOGB, false>(lds, wg);
}

template <bool FP64, bool RSD, bool W, bool LOGB>
static hipError_t launch_t(const PairArgs& a, size_t lds, int nblocks, int wg, hipStream_t st)
{
    auto kern = k_pair2<FP64, RSD, W, LOGB, false>;
    if constexpr (W && !FP64)
        if (use_dual<RSD, LOGB>(lds, wg)) kern = k_pair2<FP64, RSD, W, LOGB, true>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(wg), lds, st, a);
    return hipGetLastError();
}

template <bool FP64, bool RSD, bool W, bool LOGB>
static int occ_t(size_t lds, int wg)
{
    return occ_k<FP64, RSD, W, LOGB, false>(lds, wg);
}

// Runs F<FP64, RSD, W, LOGB>(args...) for the runtime flags.
template <template <bool, bool, bool, bool> class F, class... Args>
static auto dispatch(bool fp64, bool rsd, bool w, bool logb, Args... args)
{
#define GS_D4(A, B, C, D) \
    if (fp64 == A && rsd == B && w == C && logb == D) return F<A, B, C, D>::run(args...);
    GS_D4(true, true, true, true) GS_D4(true, true, true, false) GS_D4(true, true, false, true)
    GS_D4(true, true, false, false) GS_D4(true, false, true, true) GS_D4(true, false, true, false)
    GS_D4(true, false, false, true) GS_D4(true, false, false, false) GS_D4(false, true, true, true)
    GS_D4(false, true, true, false) GS_D4(false, true, false, true) GS_D4(false, true, false, false)
    GS_D4(false, false, true, true) GS_D4(false, false, true, false) GS_D4(false, false, false, true)
#undef GS_D4
    return F<false, false, false, false>::run(args...);
}

template <bool A, bool B, bool C, bool D>
struct Launch {
    static hipError_t run(const PairArgs& a, size_t lds, int nblocks, int wg, hipStream_t st)
    {
        return launch_t<A, B, C, D>(a, lds, nblocks, wg, st);
    }
};
template <bool A, bool B, bool C, bool D>
struct Occ {
    static int run(size_t lds, int wg) { return occ_t<A, B, C, D>(lds, wg); }
};

hipError_t launch_pair2(const PairArgs& a, int nb, int nm, bool fp64, bool rsd, bool weighted, bool logb,
                        int nblocks, int wg, hipStream_t st)
{
    const size_t lds = pair2_lds_bytes(nb, rsd ? nm : 1, a.reps, fp64, weighted, a.nreg > 0, wg);
    return dispatch<Launch>(fp64, rsd, weighted, logb, a, lds, nblocks, wg, st);
}

int pair2_occupancy(int nb, int nm, int reps, bool fp64, bool rsd, bool weighted, bool logb, bool batched, int wg)
{
    const size_t lds = pair2_lds_bytes(nb, rsd ? nm : 1, reps, fp64, weighted, batched, wg);
    if (lds > 160 * 1024) return 0;
    return dispatch<Occ>(fp64, rsd, weighted, logb, lds, wg);
}

}  // namespace gs

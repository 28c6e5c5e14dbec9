// adjoint.hip -- adjoint of the weighted 2PCF (r, mu) counts with respect to
// the per-point weights.
//
// Forward (gramsci.F90:352-380): N2[b,m] = sum over kept ordered pairs i->j of
// wn_i wn_j, N3[b,m] the same over random-random pairs, with the normalised
// weights wn = w/Wd (data) and -w/Wr (randoms) of gramsci.F90:1390-1391.
// Given g2 = dL/dN2 and g3 = dL/dN3 this computes, per point k,
//     G_k = dL/dwn_k = sum_j wn_j [c_k (g2 + [k,j random] g3)(b, m_kj)
//                                 + c_j (g2 + [k,j random] g3)(b, m_jk)]
// (c = 1 if the point may be an apex), and then the raw-weight gradient
//     dL/dw_k = (G_k - S_D)/Wd (data),   -(G_k + S_R)/Wr (randoms),
// S_D = sum_{data} wn G, S_R = sum_{randoms} wn G.
//
// Two kernels enumerate the pairs as the pair kernel does (wave-level x-row
// chunks, FP32 geometry with the same guard band and FP64 recheck, or plain
// FP64): k_adjoint2 over the FULL stencil, so that every ordered pair is seen
// from its apex's lane and G_k accumulates in an FP64 register without
// atomics (FP64 precision, and non-finite dL/dN); k_adjoint2h (FP32) over the
// half stencil, each unordered pair once, with exact fixed-point sums.
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

using namespace dev;

size_t adjoint2_lds_bytes(int nb, int nm, bool fp64)
{
    size_t b = 0;
    b += (size_t)nb * nm * 8 * 2;                              // g2, g2 + g3
    b += ((size_t)(kMaxBins + 1) * 8);                          // ok_lo/ok_hi
    b += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;           // t64
    const int waves = kWG / kWave;
    b += (size_t)waves * kWave * (fp64 ? 40 : 32);              // tiles
    return b;
}

template <bool FP64, bool RSD>
__global__ __launch_bounds__(kWG) void k_adjoint2(AdjArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BinDesc& B = *a.bins;
    const int nb = B.nb;
    const int nm = RSD ? B.nm : 1;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid / kWave;

    unsigned char* p = smem;
    double* g2 = (double*)p;                // dL/dN2 per (bin, mu)
    double* g23 = g2 + nb * nm;             // dL/dN2 + dL/dN3 (a random-random pair feeds both)
    p += (size_t)nb * nm * 16;
    float2* okb = (float2*)p;
    p += (size_t)(kMaxBins + 1) * 8;
    double* t64 = (double*)p;
    p += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;
    constexpr int kTileBytes = FP64 ? kWave * 40 : kWave * 32;
    unsigned char* tb = p + (size_t)wv * kTileBytes;
    double* tx = (double*)tb;
    double* ty = tx + kWave;
    double* tz = ty + kWave;
    float4* tl = (float4*)tb;
    double* tw = FP64 ? (tz + kWave) : (double*)(tl + kWave);
    uint32_t* tf = (uint32_t*)(tw + kWave);

    for (int k = tid; k < nb * nm; k += kWG) { g2[k] = a.g2[k]; g23[k] = a.g2[k] + a.g3[k]; }
    for (int k = tid; k < nb; k += kWG) okb[k] = make_float2(B.ok_lo[k], B.ok_hi[k]);
    for (int k = tid; k < nb + 2; k += kWG) t64[k] = B.t64[k];
    __syncthreads();

    const float rej_lo = B.rej_lo, rej_hi = B.rej_hi, ua = B.ua, ub = B.ub;
    const bool logb = B.log_bins != 0;
    const double t_lo = B.t64[0], t_hi = B.t64[nb + 1];
    const double ms = B.mu_scale;
    const float ms32 = B.ms32;
    const float dt_ew = ms32 * B.mu_ew, dt_em = ms32 * B.mu_em, dt_c = ms32 * B.mu_c;   // mu error bound
    const float nbm1f = (float)(nb - 1), nmm1f = (float)(nm - 1);
    // mu edges are the integers 1..nm-1 of t; with nm = 127 also 0 and 127, where the reference's
    // int8 store of floor(t)+1 wraps (gramsci.F90:326-328), so pairs there get the FP64 recheck
    const float e_lo = nm == 127 ? 0.0f : 1.0f, e_hi = nm == 127 ? 127.0f : nmm1f;
    const int nx = a.g.dim[0], ny = a.g.dim[1], nz = a.g.dim[2];
    const int nrow = ny * nz;
    const float h32 = a.g.h32, hx32 = a.g.hx32;    // cell edges in y/z and in x
    const double inv_wd = a.inv_wd, inv_wr = a.inv_wr;

    unsigned long long cand = 0, rechk = 0;
    const int xcd = blockIdx.x & 7;
    int tried = 0;

    while (true) {
        int item = -1;
        if (lane == 0) item = next_item(a.queue, a.n_chunks, xcd, tried);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item < 0) break;

        const int row = upper_index(a.row_chunk, nrow, item);
        const int q = item - a.row_chunk[row];
        const int cy = row % ny, cz = row / ny;
        const int rs = a.cell_start[row * nx], re = a.cell_start[(row + 1) * nx];
        const int i0 = rs + q * kChunk;
        const int n_i = min(kChunk, re - i0);
        const int i = i0 + lane;
        const bool vi = lane < n_i;
        const int cx_first = (int)a.pl[i0].w;
        const int cx_last = (int)a.pl[i0 + n_i - 1].w;

        double4 Pi = make_double4(0, 0, 0, 0);
        float4 Li = make_float4(0, 0, 0, 0);
        uint32_t fi = 2u;
        if (vi) { Pi = a.pw[i]; Li = a.pl[i]; fi = a.flags[i]; }
        const int ri = fi & 1, ci = (fi & 2) ? 0 : 1;
        const float ax32 = (float)Pi.x, ay32 = (float)Pi.y, az32 = (float)Pi.z;
        double G = 0.0;

        for (int r = 0; r < a.nrows; ++r) {
            const Row rw = a.rows[r];
            const int by = cy + rw.dy, bz = cz + rw.dz;
            if (by < 0 || by >= ny || bz < 0 || bz >= nz) continue;
            const int bx0 = max(cx_first + rw.x0, 0), bx1 = min(cx_last + rw.x1, nx - 1);
            if (bx0 > bx1) continue;
            const int rb = (bz * ny + by) * nx;
            const int js = a.cell_start[rb + bx0], je = a.cell_start[rb + bx1 + 1];
            const float yi_r = Li.y - (float)rw.dy * h32;
            const float zi_r = Li.z - (float)rw.dz * h32;
            for (int jt = js; jt < je; jt += kWave) {
                const int nt = min(kWave, je - jt);
                wave_sync();
                if (lane < nt) {
                    const int j = jt + lane;
                    const double4 Pj = a.pw[j];
                    const uint32_t fj = a.flags[j];
                    if constexpr (FP64) {
                        tx[lane] = Pj.x; ty[lane] = Pj.y; tz[lane] = Pj.z;
                    } else {
                        tl[lane] = a.pl[j];
                    }
                    tw[lane] = (fj & 1u) ? -Pj.w * inv_wr : Pj.w * inv_wd;   // normalised weight
                    tf[lane] = fj;
                }
                wave_sync();
                const int kend = vi ? nt : 0;
                cand += (unsigned long long)kend;
                for (int k = 0; k < kend; ++k) {
                    int b, m = 0, mr = 0;
                    const uint32_t fj = tf[k];
                    if constexpr (FP64) {
                        const double jx = tx[k], jy = ty[k], jz = tz[k];
                        const double s = exact_s(Pi.x, Pi.y, Pi.z, jx, jy, jz);
                        if (!(s > t_lo && s < t_hi)) continue;
                        b = adjust_bin(s, bin_estimate((float)s, logb, ua, ub), t64, nb);
                        if (b == nb) continue;
                        if constexpr (RSD) exact_mu(Pi.x, Pi.y, Pi.z, jx, jy, jz, ms, nm, m, mr);
                    } else {
                        const float4 tj = tl[k];
                        const float dx = fmaf(tj.w - Li.w, hx32, tj.x - Li.x);
                        const float dy = tj.y - yi_r, dz = tj.z - zi_r;
                        const float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        if (!(s > rej_lo && s < rej_hi)) continue;
                        b = (int)__builtin_amdgcn_fmed3f(bin_estimate(s, logb, ua, ub), 0.0f, nbm1f);
                        const float2 ok = okb[b];
                        bool safe = (s >= ok.x) && (s < ok.y);
                        if constexpr (RSD) {
                            const float mx = fmaf(0.5f, dx, ax32), my = fmaf(0.5f, dy, ay32), mz = fmaf(0.5f, dz, az32);
                            const float D = -fmaf(mz, dz, fmaf(my, dy, mx * dx));
                            const float M2 = fmaf(mz, mz, fmaf(my, my, mx * mx));
                            const float rm = __builtin_amdgcn_rsqf(M2), rsq = __builtin_amdgcn_rsqf(s);
                            const float t = fmaf(D * rm * rsq, ms32, ms32);       // (theta + 1) * mu_scale
                            const float dt = fmaf(dt_ew, rsq, fmaf(dt_em, rm, dt_c));
                            // certain when the nearest mu edge (an integer 1..nm-1 of t) is farther than dt
                            const float edge = __builtin_amdgcn_fmed3f(rintf(t), e_lo, e_hi);
                            safe = safe && (dt < 0.5f) && (fabsf(t - edge) >= dt);
                            m = min((int)t, nm - 1);       // t >= 0 up to rounding; truncation = floor there
                            mr = nm - 1 - m;
                        }
                        if (!safe) {
                            ++rechk;
                            const double4 Pj = a.pw[jt + k];
                            const double se = exact_s(Pi.x, Pi.y, Pi.z, Pj.x, Pj.y, Pj.z);
                            if (!(se > t_lo && se < t_hi)) continue;
                            b = exact_bin(se, t64, nb);
                            if (b == nb) continue;
                            if constexpr (RSD) exact_mu(Pi.x, Pi.y, Pi.z, Pj.x, Pj.y, Pj.z, ms, nm, m, mr);
                        }
                    }
                    const int cj = (fj & 2) ? 0 : 1;
                    const bool rr = ri && (fj & 1);
                    const int o = b * nm;
                    const double* gt = rr ? g23 : g2;
                    double c = 0.0;
                    if (ci) c += gt[o + m];
                    if (cj) c += gt[o + mr];
                    G = fma(c, tw[k], G);
                }
            }
        }
        if (vi) a.G[a.orig[i]] = G;
    }
    for (int off = 32; off > 0; off >>= 1) {
        cand += __shfl_down(cand, off, 64);
        rechk += __shfl_down(rechk, off, 64);
    }
    if (lane == 0) {
        if (cand) atomicAdd(&a.out_stats[0], cand);
        if (rechk) atomicAdd(&a.out_stats[1], rechk);
    }
}

// Half-stencil FP32 adjoint.  The pair (i, j) contributes the same coefficient
//     c = c_i g(b, m_ij) + c_j g(b, m_ji)     (g = dL/dN2, plus dL/dN3 for a random-random pair)
// to both ends: G_i += c wn_j and G_j += c wn_i.  So every unordered pair is evaluated once, from
// the lane of its first point in the pair kernel's half-stencil order (x-rows with dz > 0, or dz = 0
// and dy >= 0; in the own row only j > i), with the pair kernel's FP32 geometry, guard band and FP64
// recheck.  The neighbour tile is a register tile rotated one lane per step (DPP); the j-side sum
// rides along as one more rotated register and is added to the neighbour's slot when the tile ends.
// Terms are rounded to integers at the scale 2^F (|term| * 2^F <= 2^50) and summed exactly in 64-bit
// registers (at most 64 terms per tile) and 128-bit global accumulators, so the result does not
// depend on the order of the additions and is reproducible bit for bit.
size_t adjoint2h_lds_bytes(int nb, int nm)
{
    size_t b = (size_t)nb * nm * 8 * 2;                        // g2, g2 + g3 (scaled)
    b += ((size_t)(kMaxBins + 1) * 8);                          // ok_lo/ok_hi
    b += ((size_t)(kMaxBins + 2) * 8 + 15) / 16 * 16;           // t64
    return b;
}

template <bool RSD>
__global__ __launch_bounds__(kWG) void k_adjoint2h(AdjArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const BinDesc& B = *a.bins;
    const int nb = B.nb;
    const int nm = RSD ? B.nm : 1;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);

    unsigned char* p = smem;
    double* g2 = (double*)p;
    double* g23 = g2 + nb * nm;
    p += (size_t)nb * nm * 16;
    float2* okb = (float2*)p;
    p += (size_t)(kMaxBins + 1) * 8;
    double* t64 = (double*)p;

    const double gsc = a.g_scale;
    for (int k = tid; k < nb * nm; k += kWG) { g2[k] = a.g2[k] * gsc; g23[k] = (a.g2[k] + a.g3[k]) * gsc; }
    for (int k = tid; k < nb; k += kWG) okb[k] = make_float2(B.ok_lo[k], B.ok_hi[k]);
    for (int k = tid; k < nb + 2; k += kWG) t64[k] = B.t64[k];
    __syncthreads();

    const float rej_lo = B.rej_lo, rej_hi = B.rej_hi, ua = B.ua, ub = B.ub;
    const bool logb = B.log_bins != 0;
    const double t_lo = B.t64[0], t_hi = B.t64[nb + 1];
    const double ms = B.mu_scale;
    const float ms32 = B.ms32;
    const float dt_ew = ms32 * B.mu_ew, dt_em = ms32 * B.mu_em, dt_c = ms32 * B.mu_c;
    const float nbm1f = (float)(nb - 1), nmm1f = (float)(nm - 1);
    const float e_lo = nm == 127 ? 0.0f : 1.0f, e_hi = nm == 127 ? 127.0f : nmm1f;
    const int nx = a.g.dim[0], ny = a.g.dim[1], nz = a.g.dim[2];
    const int nrow = ny * nz;
    const float h32 = a.g.h32, hx32 = a.g.hx32;    // cell edges in y/z and in x
    // normalised weights times w_scale: w / Wd for data, -w / Wr for randoms
    const double wsd = a.inv_wd * a.w_scale, wsr = -a.inv_wr * a.w_scale;

    unsigned long long cand = 0, rechk = 0;
    const int xcd = blockIdx.x & 7;
    int tried = 0;

    while (true) {
        int item = -1;
        if (lane == 0) item = next_item(a.queue, a.n_chunks, xcd, tried);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item < 0) break;

        const int row = upper_index(a.row_chunk, nrow, item);
        const int q = item - a.row_chunk[row];
        const int cy = row % ny, cz = row / ny;
        const int rs = a.cell_start[row * nx], re = a.cell_start[(row + 1) * nx];
        const int i0 = rs + q * kChunk;
        const int n_i = min(kChunk, re - i0);
        const int i = i0 + lane;
        const bool vi = lane < n_i;
        const int cx_first = (int)a.pl[i0].w;
        const int cx_last = (int)a.pl[i0 + n_i - 1].w;

        double4 Pi = make_double4(0, 0, 0, 0);
        float4 Li = make_float4(0, 0, 0, 0);
        uint32_t fi = 2u;
        if (vi) { Pi = a.pw[i]; Li = a.pl[i]; fi = a.flags[i]; }
        const int ri = fi & 1, ci = (fi & 2) ? 0 : 1;
        const bool chunk_buf = __ballot(ci != 0) == 0;
        const double wiF = vi ? Pi.w * (ri ? wsr : wsd) : 0.0;
        const float ax32 = (float)Pi.x, ay32 = (float)Pi.y, az32 = (float)Pi.z;
        unsigned long long Ilo = 0, Ihi = 0;        // this lane's 128-bit i-side sum

        for (int r = 0; r < a.nrows; ++r) {
            const Row rw = a.rows[r];
            const int by = cy + rw.dy, bz = cz + rw.dz;
            if (by < 0 || by >= ny || bz < 0 || bz >= nz) continue;
            const int bx0 = max(cx_first + rw.x0, 0), bx1 = min(cx_last + rw.x1, nx - 1);
            if (bx0 > bx1) continue;
            const bool self = (rw.dy == 0 && rw.dz == 0);
            const int rb = (bz * ny + by) * nx;
            int js = a.cell_start[rb + bx0];
            const int je = a.cell_start[rb + bx1 + 1];
            if (self) js = max(js, i0 + 1);
            const float yi_r = Li.y - (float)rw.dy * h32;
            const float zi_r = Li.z - (float)rw.dz * h32;
            for (int jt = js; jt < je; jt += kWave) {
                const int nt = min(kWave, je - jt);
                const int jmin = max(self ? i + 1 - jt : 0, 0);   // own row: only j > i
                if (vi) cand += (unsigned long long)(nt - min(jmin, nt));
                // tile of nt entries with period P (the power of two >= nt): lane L holds entry L mod P
                const int P = nt <= 1 ? 1 : 1 << (32 - __clz(nt - 1));
                const int le = lane & (P - 1);
                const bool has = le < nt;
                const int jl = jt + (has ? le : 0);
                const uint32_t ef = a.flags[jl];
                const bool skip = chunk_buf && __ballot(has && (ef & 2u) == 0) == 0;
                float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
                double ewj = 0.0;
                if (!skip) {
                    e = a.pl[jl];
                    const double wj = a.ws[jl];
                    ewj = wj * ((ef & 1u) ? wsr : wsd);
                }
                int ep = has ? (le << 2) | (int)(ef & 3u) : -1;   // entry index << 2 | flags, -1: none
                const int jmin4 = jmin << 2;
                long long qi = 0, qj = 0;
                const int kend = __builtin_amdgcn_readfirstlane(skip ? 0 : P);     // wave-uniform
                for (int k = 0; k < kend; ++k) {
                    if (vi && ep >= jmin4) do {
                        const float dx = fmaf(e.w - Li.w, hx32, e.x - Li.x);
                        const float dy = e.y - yi_r, dz = e.z - zi_r;
                        const float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        if (!(s > rej_lo && s < rej_hi)) break;
                        int b = (int)__builtin_amdgcn_fmed3f(bin_estimate(s, logb, ua, ub), 0.0f, nbm1f);
                        const float2 ok = okb[b];
                        bool safe = (s >= ok.x) && (s < ok.y);
                        int m = 0, mr = 0;
                        if constexpr (RSD) {
                            const float rq = __builtin_amdgcn_rsqf(s);
                            const float mx = fmaf(0.5f, dx, ax32), my = fmaf(0.5f, dy, ay32), mz = fmaf(0.5f, dz, az32);
                            const float D = -fmaf(mz, dz, fmaf(my, dy, mx * dx));
                            const float M2 = fmaf(mz, mz, fmaf(my, my, mx * mx));
                            const float rm = __builtin_amdgcn_rsqf(M2);
                            const float t = fmaf(D * rm * rq, ms32, ms32);        // (theta + 1) * mu_scale
                            const float dt = fmaf(dt_ew, rq, fmaf(dt_em, rm, dt_c));
                            const float edge = __builtin_amdgcn_fmed3f(rintf(t), e_lo, e_hi);
                            safe = safe && (dt < 0.5f) && (fabsf(t - edge) >= dt);
                            m = min((int)t, nm - 1);       // t >= 0 up to rounding; truncation = floor there
                            mr = nm - 1 - m;
                        }
                        if (!safe) {
                            ++rechk;
                            const double4 Pj = a.pw[jt + (ep >> 2)];
                            const double4 Pq = a.pw[i];
                            const double se = exact_s(Pq.x, Pq.y, Pq.z, Pj.x, Pj.y, Pj.z);
                            if (!(se > t_lo && se < t_hi)) break;
                            b = exact_bin(se, t64, nb);
                            if (b == nb) break;
                            if constexpr (RSD) exact_mu(Pq.x, Pq.y, Pq.z, Pj.x, Pj.y, Pj.z, ms, nm, m, mr);
                        }
                        const int cj = (ep & 2) ? 0 : 1;
                        const double* gt = (ri && (ep & 1)) ? g23 : g2;
                        const int o = __umul24(b, nm);
                        double c = 0.0;
                        if (ci) c = gt[o + m];
                        if (cj) c += gt[o + mr];
                        qi += quantise_rn_signed(c * ewj);
                        qj += quantise_rn_signed(c * wiF);
                    } while (false);
                    // rotate the register tile and the j-side sums by one lane (all lanes active here)
                    e.x = rot_lane(e.x); e.y = rot_lane(e.y); e.z = rot_lane(e.z); e.w = rot_lane(e.w);
                    ewj = rot_lane(ewj);
                    ep = rot_lane(ep);
                    qj = rot_lane(qj);
                }
                // after P steps lane L again holds a copy of entry L mod P: the copies together have met
                // every lane once, so each adds its sum to the entry's slot
                if (has && qj != 0)
                    glb_add128(a.Gq + 2 * (size_t)(jt + le), (unsigned long long)qj, qj < 0 ? ~0ull : 0ull);
                const unsigned long long lo = Ilo + (unsigned long long)qi;
                Ihi += (lo < Ilo ? 1ull : 0ull) + (qi < 0 ? ~0ull : 0ull);
                Ilo = lo;
            }
        }
        if (vi && (Ilo | Ihi)) glb_add128(a.Gq + 2 * (size_t)i, Ilo, Ihi);
    }
    for (int off = 32; off > 0; off >>= 1) {
        cand += __shfl_down(cand, off, 64);
        rechk += __shfl_down(rechk, off, 64);
    }
    if (lane == 0) {
        if (cand) atomicAdd(&a.out_stats[0], cand);
        if (rechk) atomicAdd(&a.out_stats[1], rechk);
    }
}

// G[orig[k]] = (hi * 2^64 + lo) * 2^-F for the two's-complement 128-bit sum of sorted point k,
// converted through its magnitude (both words non-negative, so no cancellation)
__global__ __launch_bounds__(256) void k_adj_fx(const unsigned long long* __restrict__ Gq,
                                                const int32_t* __restrict__ orig, int64_t n, int F,
                                                double* __restrict__ G)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long lo = Gq[2 * k], hi = Gq[2 * k + 1];
        const bool neg = (long long)hi < 0;
        if (neg) {
            lo = ~lo + 1ull;
            hi = ~hi + (lo == 0ull ? 1ull : 0ull);
        }
        const double v = fma((double)hi, 0x1p64, (double)lo);
        G[orig[k]] = ldexp(neg ? -v : v, -F);
    }
}

// per-block partial sums of wn*G over data (slot 0) and randoms (slot 1)
__global__ __launch_bounds__(256) void k_adj_partials(const double* __restrict__ G, const double* __restrict__ w,
                                                      int64_t n, int64_t n_data, double inv_wd, double inv_wr,
                                                      double* __restrict__ part)
{
    __shared__ double sd[256], sr[256];
    double a = 0.0, b = 0.0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const double wk = w ? w[k] : 1.0;
        if (k < n_data) a += wk * inv_wd * G[k];
        else b += -wk * inv_wr * G[k];
    }
    sd[threadIdx.x] = a;
    sr[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sd[threadIdx.x] += sd[threadIdx.x + s];
            sr[threadIdx.x] += sr[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sd[0];
        part[2 * blockIdx.x + 1] = sr[0];
    }
}

__global__ __launch_bounds__(256) void k_adj_finish(const double* __restrict__ G, int64_t n, int64_t n_data,
                                                    double s_d, double s_r, double inv_wd, double inv_wr,
                                                    double* __restrict__ out)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        out[k] = k < n_data ? (G[k] - s_d) * inv_wd : -(G[k] + s_r) * inv_wr;
}

template <bool RSD>
static hipError_t launch_h(const AdjArgs& a, size_t lds, int nblocks, hipStream_t st)
{
    auto kern = k_adjoint2h<RSD>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(kWG), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_adjoint2h(const AdjArgs& a, int nb, int nm, bool rsd, int nblocks, hipStream_t st)
{
    const size_t lds = adjoint2h_lds_bytes(nb, rsd ? nm : 1);
    return rsd ? launch_h<true>(a, lds, nblocks, st) : launch_h<false>(a, lds, nblocks, st);
}

int adjoint2h_occupancy(int nb, int nm, bool rsd)
{
    const size_t lds = adjoint2h_lds_bytes(nb, rsd ? nm : 1);
    int n = 0;
    const hipError_t e = rsd ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjoint2h<true>, kWG, lds)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjoint2h<false>, kWG, lds);
    return (e == hipSuccess && n > 0) ? n : 1;
}

hipError_t launch_adj_fx(const unsigned long long* Gq, const int32_t* orig, int64_t n, int F, double* G, hipStream_t st)
{
    int nblk = (int)std::min<int64_t>((n + 255) / 256, 65535);
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL(k_adj_fx, dim3(nblk), dim3(256), 0, st, Gq, orig, n, F, G);
    return hipGetLastError();
}

template <bool FP64, bool RSD>
static hipError_t launch_a(const AdjArgs& a, size_t lds, int nblocks, hipStream_t st)
{
    auto kern = k_adjoint2<FP64, RSD>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(kWG), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_adjoint2(const AdjArgs& a, int nb, int nm, bool fp64, bool rsd, int nblocks, hipStream_t st)
{
    const size_t lds = adjoint2_lds_bytes(nb, rsd ? nm : 1, fp64);
    if (fp64) return rsd ? launch_a<true, true>(a, lds, nblocks, st) : launch_a<true, false>(a, lds, nblocks, st);
    return rsd ? launch_a<false, true>(a, lds, nblocks, st) : launch_a<false, false>(a, lds, nblocks, st);
}

int adjoint2_occupancy(int nb, int nm, bool fp64, bool rsd)
{
    const size_t lds = adjoint2_lds_bytes(nb, rsd ? nm : 1, fp64);
    int n = 0;
    hipError_t e;
    if (fp64) e = rsd ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjoint2<true, true>, kWG, lds)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjoint2<true, false>, kWG, lds);
    else e = rsd ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjoint2<false, true>, kWG, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjoint2<false, false>, kWG, lds);
    return (e == hipSuccess && n > 0) ? n : 1;
}

hipError_t launch_adj_partials(const double* G, const double* w, int64_t n, int64_t n_data, double inv_wd,
                               double inv_wr, double* part, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_partials, dim3(nblk), dim3(256), 0, st, G, w, n, n_data, inv_wd, inv_wr, part);
    return hipGetLastError();
}

hipError_t launch_adj_finish(const double* G, int64_t n, int64_t n_data, double s_d, double s_r, double inv_wd,
                             double inv_wr, double* out, hipStream_t st)
{
    int nblk = (int)std::min<int64_t>((n + 255) / 256, 65535);
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL(k_adj_finish, dim3(nblk), dim3(256), 0, st, G, n, n_data, s_d, s_r, inv_wd, inv_wr, out);
    return hipGetLastError();
}

}  // namespace gs

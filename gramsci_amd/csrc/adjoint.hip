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
// The pair enumeration is the pair kernel's (wave-level x-row chunks, LDS
// tiles, FP32 geometry with the same guard band and FP64 recheck, or plain
// FP64), over the FULL stencil so that every ordered pair is seen from its
// apex's lane and G_k accumulates in a register without atomics.
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
    const float h32 = a.g.h32;
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
                        const float dx = fmaf(tj.w - Li.w, h32, tj.x - Li.x);
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
                            m = (int)__builtin_amdgcn_fmed3f(floorf(t), 0.0f, nmm1f);
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

// gs_device.h -- device helpers shared by the counting kernels: the
// reference's FP64 pair arithmetic, exact binning against the thresholds,
// fixed-point splitting and small synchronisation/accumulation utilities.
#pragma once

#include "gs_common.h"

namespace gs {
namespace dev {

__device__ __forceinline__ void glb_add128(unsigned long long* slot, unsigned long long lo, unsigned long long hi)
{
    unsigned long long old = atomicAdd(slot, lo);
    if (old + lo < old) hi += 1ull;
    if (hi) atomicAdd(slot + 1, hi);
}

// s = (bx-ax)^2 + (by-ay)^2 + (bz-az)^2 in FP64, evaluated left to right with
// no contraction, as the reference forms it (kdtree2.F90:1689, gramsci.F90:1532).
__device__ __forceinline__ double exact_s(double ax, double ay, double az, double bx, double by, double bz)
{
    double dx = __dsub_rn(bx, ax), dy = __dsub_rn(by, ay), dz = __dsub_rn(bz, az);
    return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// 0-based bin of a kept s: #{k in 1..nb : s >= t[k]}; nb means the reference's
// out-of-range bin nbins+1.
__device__ __forceinline__ int exact_bin(double s, const double* t, int nb)
{
    int lo = 0, hi = nb;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (t[mid] <= s) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// exact 0-based bin, -1 if not kept, nb on overflow
__device__ __forceinline__ int bin_of(double s, const double* t, int nb)
{
    if (!(s > t[0] && s < t[nb + 1])) return -1;
    return exact_bin(s, t, nb);
}

// exact bin of a kept s, starting from the FP32 estimate u and correcting it
// against the FP64 thresholds (0 or 1 steps in practice)
__device__ __forceinline__ int adjust_bin(double s, float u, const double* t, int nb)
{
    int b = (int)u;
    b = b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
    while (b > 0 && s < t[b]) --b;
    while (b < nb && s >= t[b + 1]) ++b;
    return b;
}

// mu bin from t = (theta+1)*mu_scale (gramsci.F90:326-328), 0-based.  The
// reference stores floor(t)+1 in an int8 before clamping; that wrap is kept.
__device__ __forceinline__ int mu_from_t(double t, int nm)
{
    int v = (int)floor(t) + 1;
    v = (int)(int8_t)v;
    v = v < 1 ? 1 : (v > nm ? nm : v);
    return v - 1;
}

// mu bins of apex a -> neighbour b and of b -> a (gramsci.F90:320-325),
// in the reference's operation order with IEEE sqrt and division.
__device__ __forceinline__ void exact_mu(double ax, double ay, double az, double bx, double by, double bz, double ms,
                                         int nm, int& m, int& mr)
{
    double m1 = __dmul_rn(0.5, __dadd_rn(ax, bx));
    double m2 = __dmul_rn(0.5, __dadd_rn(ay, by));
    double m3 = __dmul_rn(0.5, __dadd_rn(az, bz));
    double w1 = __dsub_rn(ax, bx), w2 = __dsub_rn(ay, by), w3 = __dsub_rn(az, bz);
    double th = __dadd_rn(__dadd_rn(__dmul_rn(m1, w1), __dmul_rn(m2, w2)), __dmul_rn(m3, w3));
    double mm = __dadd_rn(__dadd_rn(__dmul_rn(m1, m1), __dmul_rn(m2, m2)), __dmul_rn(m3, m3));
    double ww = __dadd_rn(__dadd_rn(__dmul_rn(w1, w1), __dmul_rn(w2, w2)), __dmul_rn(w3, w3));
    double den = __dmul_rn(__dsqrt_rn(mm), __dsqrt_rn(ww));
    th = __ddiv_rn(th, den);
    m = mu_from_t(__dmul_rn(__dadd_rn(th, 1.0), ms), nm);
    mr = mu_from_t(__dmul_rn(__dadd_rn(-th, 1.0), ms), nm);
}

// FP32 estimate of the radial bin coordinate of s (exactness comes from the
// thresholds, not from this estimate)
__device__ __forceinline__ float bin_estimate(float s, bool logb, float ua, float ub)
{
    return logb ? fmaf(__builtin_amdgcn_logf(s), ua, ub) : fmaf(__builtin_amdgcn_sqrtf(s), ua, ub);
}

// floor(x) split at bit 32 for 0 <= x < 2^61: q = hi * 2^32 + lo.
__device__ __forceinline__ void quantise(double x, uint32_t& lo, uint32_t& hi)
{
    const double h = floor(x * 0x1p-32);
    hi = (uint32_t)h;
    lo = (uint32_t)fma(-h, 0x1p32, x);
}

// rint(x) (round to nearest even) as a 64-bit integer, for 0 <= x < 2^52: x + 2^52
// lies in [2^52, 2^53], where the spacing of doubles is 1, so the addition rounds x to an
// integer and the bit pattern of the sum is that of 2^52 plus it (also at 2^53)
__device__ __forceinline__ unsigned long long quantise_rn(double x)
{
    return (unsigned long long)(__double_as_longlong(x + 0x1p52) - 0x4330000000000000ll);
}

// Rotation of a register across the 64 lanes of a wavefront by one lane (DPP
// wave_rol:1); every lane must be active.
// The register itself is passed as the unused "old" operand (a full rotation
// writes every lane), so the move can be done in place: one v_mov_b32_dpp per
// 32-bit register and step, with no zeroed temporary and no copy back.
__device__ __forceinline__ int rot_lane(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x134, 0xf, 0xf, false); }
__device__ __forceinline__ float rot_lane(float v) { return __int_as_float(rot_lane(__float_as_int(v))); }
__device__ __forceinline__ long long rot_lane(long long b)
{
    const int lo = rot_lane((int)(b & 0xffffffffll)), hi = rot_lane((int)(b >> 32));
    return ((long long)hi << 32) | (long long)(uint32_t)lo;
}
__device__ __forceinline__ double rot_lane(double v) { return __longlong_as_double(rot_lane(__double_as_longlong(v))); }

// rint(x) as a signed 64-bit integer, for |x| < 2^51: x + 1.5 * 2^52 lies in
// [2^52, 2^53), where the spacing of doubles is 1, so the addition rounds x to an
// integer and the bit pattern of the sum is that of 1.5 * 2^52 plus it
__device__ __forceinline__ long long quantise_rn_signed(double x)
{
    return __double_as_longlong(x + 0x1.8p52) - 0x4338000000000000ll;
}

// largest r with off[r] <= v over off[0..n-1] (off non-decreasing)
__device__ __forceinline__ int upper_index(const int32_t* off, int n, int v)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Next batch of up to bsz work items from the 8 queues of next_item below: returns
// the first item and sets `end` (exclusive), or returns -1 at the end.  One atomic
// per batch: with millions of short work items (apexes with few neighbours) one
// atomic per item would queue on the counters' cache line.
__device__ __forceinline__ int next_items(int32_t* queue, int n, int xcd, int& tried, int bsz, int& end)
{
    while (tried < 8) {
        const int sidx = (xcd + tried) & 7;
        const int lo = (int)(((long long)n * sidx) >> 3);
        const int hi = (int)(((long long)n * (sidx + 1)) >> 3);
        const int k = atomicAdd(&queue[sidx], bsz);
        if (lo + k < hi) {
            end = min(lo + k + bsz, hi);
            return lo + k;
        }
        ++tried;
    }
    return -1;
}

// Next work item from 8 queues, one per XCD segment of [0, n): own segment
// first, then the others.  `tried` is per-caller state.  Returns -1 at the end.
__device__ __forceinline__ int next_item(int32_t* queue, int n, int xcd, int& tried)
{
    while (tried < 8) {
        const int sidx = (xcd + tried) & 7;
        const int lo = (int)(((long long)n * sidx) >> 3);
        const int hi = (int)(((long long)n * (sidx + 1)) >> 3);
        const int k = atomicAdd(&queue[sidx], 1);
        if (lo + k < hi) return lo + k;
        ++tried;
    }
    return -1;
}

}  // namespace dev
}  // namespace gs

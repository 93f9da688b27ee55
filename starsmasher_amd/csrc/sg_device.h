// sg_device.h -- device helpers shared by the pair kernels (sg_direct.hip)
// and the adjoint kernels (sg_adjoint.hip): wave-uniform values and
// reductions, the tile box test, the kernel polynomials, FP64 rsqrt.
#pragma once

#include "sg_internal.h"

#include <math.h>

namespace sg {

constexpr float FAR_MARGIN = 1.0e-3f;   // relative guard on the box-distance test
// Hi-only differences x_j,hi - x_i,hi carry a relative error of about
// 2^-24 (|x_i - O| + |x_j - O| + r) / r <= 2^-24 (2 + 2 |x_i - O| / r).
// Chunks whose box is nearer to the wave's box than WIDE_K rho (rho = the
// wave's largest |x_i - O|) form the difference in FP64 and round it once,
// which bounds the hi-only ones by 2^-24 (2 + 2 / WIDE_K) = 6 ulp (3.6e-7).
// Waves that straddle a jump of the space-filling curve have a large rho.
constexpr float WIDE_K = 0.5f;

__device__ __forceinline__ float uniform_f(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return uniform_f(v);
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return uniform_f(v);
}

// This workgroup's copy of the statistics counters.
__device__ __forceinline__ unsigned long long *stat_copy(unsigned long long *stats)
{
    return stats + (blockIdx.x % STAT_SLOTS) * STAT_STRIDE;
}

// Sum over the wave of a per-lane count, from ballots of its bits (no
// cross-lane moves, so no lane-address registers stay live until the end of
// a kernel for it).
__device__ __forceinline__ unsigned long long wave_sum_count(unsigned v)
{
    unsigned long long s = 0;
    for (int bit = 0; bit < 32; ++bit)
        s += (unsigned long long)__builtin_popcountll(__ballot((v >> bit) & 1u)) << bit;
    return s;
}

// Squared distance between a wave's i-box [bl, bh] and tile tm's box, both
// relative to the FP64 origin O.
__device__ __forceinline__ float box_dist2(const TileMeta &tm, double Ox, double Oy, double Oz,
                                           const float (&bl)[3], const float (&bh)[3])
{
    const float lx = (float)(tm.lo[0] - Ox), ly = (float)(tm.lo[1] - Oy), lz = (float)(tm.lo[2] - Oz);
    const float ux = (float)(tm.hi[0] - Ox), uy = (float)(tm.hi[1] - Oy), uz = (float)(tm.hi[2] - Oz);
    const float gx = fmaxf(0.0f, fmaxf(lx - bh[0], bl[0] - ux));
    const float gy = fmaxf(0.0f, fmaxf(ly - bh[1], bl[1] - uy));
    const float gz = fmaxf(0.0f, fmaxf(lz - bh[2], bl[2] - uz));
    return fmaf(gx, gx, fmaf(gy, gy, gz * gz));
}

// True when every pair of the wave and the tile is at least the larger
// support radius apart, i.e. every pair takes the Newtonian branch.
__device__ __forceinline__ bool boxes_far(float d2, float hmax, const TileMeta &tm)
{
    const float hm = fmaxf(hmax, (float)tm.maxH);
    return d2 > hm * hm * (1.0f + FAR_MARGIN);
}

// Near/far decision of one (wave, tile): 0 when every pair of the tile is
// Newtonian for the wave, else bit c set for each chunk c of the tile that
// holds a pair that may be softened (the near chunks).  A chunk passes when
// the wave's box is far from the chunk box, or else when every lane's own
// i is farther from the chunk box than max(H_i, largest H of the chunk).
// The per-lane test matters for waves whose space-filling-curve range
// straddles a jump: their box spans much of the set, and a box test alone
// would call most chunks near.
template <int IPL>
__device__ __forceinline__ unsigned near_mask(const TileMeta *chunks, int t, const TileMeta &tm,
                                              float d2, double Ox, double Oy, double Oz,
                                              const float (&bl)[3], const float (&bh)[3],
                                              float hmax, const float (&px)[IPL],
                                              const float (&py)[IPL], const float (&pz)[IPL],
                                              const float (&ph)[IPL])
{
    if (uniform_i(boxes_far(d2, hmax, tm)) != 0) return 0u;
    unsigned mask = 0;
    for (int c = 0; c < CHUNKS; ++c) {
        const TileMeta &cm = chunks[(int64_t)t * CHUNKS + c];
        if (uniform_i(boxes_far(box_dist2(cm, Ox, Oy, Oz, bl, bh), hmax, cm)) != 0) continue;
        const float lx = (float)(cm.lo[0] - Ox), ly = (float)(cm.lo[1] - Oy),
                    lz = (float)(cm.lo[2] - Oz);
        const float ux = (float)(cm.hi[0] - Ox), uy = (float)(cm.hi[1] - Oy),
                    uz = (float)(cm.hi[2] - Oz);
        const float hc = (float)cm.maxH;
        bool lane_near = false;
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            const float gx = fmaxf(0.0f, fmaxf(lx - px[k], px[k] - ux));
            const float gy = fmaxf(0.0f, fmaxf(ly - py[k], py[k] - uy));
            const float gz = fmaxf(0.0f, fmaxf(lz - pz[k], pz[k] - uz));
            const float dl = fmaf(gx, gx, fmaf(gy, gy, gz * gz));
            const float hm = fmaxf(ph[k], hc);
            lane_near = lane_near || !(dl > hm * hm * (1.0f + FAR_MARGIN));
        }
        if (__ballot(lane_near)) mask |= 1u << c;
    }
    return mask;
}

// Kernel polynomials A(q) (force) and P(q) (potential), q = r/H, with the
// reference's decimal coefficients (cpu_grav.f:91-117, 210-243, 321-350).
__device__ __forceinline__ void poly_AP(int nk, double q, double invq, double &A, double &P)
{
    const double q2 = q * q;
    if (nk == 0) {
        if (q < 0.5) {
            A = 10.666666666667 + q2 * (32.0 * q - 38.4);
            P = -2.8 + q2 * (5.333333333333 + q2 * (6.4 * q - 9.6));
        } else {
            const double iq3 = invq * invq * invq;
            A = 21.333333333333 + q * (-48.0 + q * (38.4 - 10.666666666667 * q))
                - 0.066666666667 * iq3;
            P = -3.2 + 0.066666666667 * invq
                + q2 * (10.666666666667 + q * (-16.0 + q * (9.6 - 2.133333333333 * q)));
        }
    } else if (nk == 1) {
        A = 28.4375 + q2 * (-187.6875 + q2 * (804.375 + q2 * (-4379.375 + q * (9009.0
            + q * (-8957.8125 + q * (5005.0 + q * (-1515.9375 + 195.0 * q)))))));
        P = -3.828125 + q2 * (14.21875 + q2 * (-46.921875 + q2 * (134.0625 + q2 * (
            -547.421875 + q * (1001.0 + q * (-895.78125 + q * (455.0 + q * (
            -126.328125 + 15.0 * q))))))));
    } else {
        A = 20.625 + q2 * (-115.5 + q2 * (618.75 + q * (-1155.0 + q * (962.5
            + q * (-396.0 + 65.625 * q)))));
        P = -3.4375 + q2 * (10.3125 + q2 * (-28.875 + q2 * (103.125 + q * (-165.0
            + q * (120.3125 + q * (-44.0 + 6.5625 * q))))));
    }
}

__device__ __forceinline__ double self_coeff(int nk)
{
    return nk == 0 ? 1.4 : (nk == 1 ? 1.9140625 : 1.71875);
}

// 1/sqrt(x) in FP64: FAST takes the FP32 hardware estimate (1 ulp) and one
// Newton step, y' = y + y (1 - x y^2) / 2, leaving ~1e-14 relative error;
// otherwise the FP64 library rsqrt.
template <bool FAST>
__device__ __forceinline__ double rsqrt64(double x)
{
    if (!FAST || !(x > 1.0e-36 && x < 1.0e36)) return rsqrt(x);   // outside FP32 range
    const double y = (double)__builtin_amdgcn_rsqf((float)x);
    const double e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// 1/sqrt(x) in FP64 for finite x > 0 in the normal range (every Newtonian
// pair: r^2 >= max(H_i^2, H_j^2) > 0): the hardware estimate v_rsq_f64 and
// one correction y + y e (1/2 + 3/8 e), e = 1 - x y^2 -- the steps of the
// library rsqrt(double) in the same operation order, without its fix-up of
// zero, infinite and denormal inputs (three VALU per pair).
__device__ __forceinline__ double rsqrt_pos64(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma((-x) * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Orders the LDS writes of one wave before its reads (and the reverse).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

}  // namespace sg

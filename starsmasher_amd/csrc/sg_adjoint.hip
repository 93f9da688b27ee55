// sg_adjoint.hip -- vector-Jacobian product of the direct-summation gravity
// for gfx950 (CDNA4).
//
// For cotangents w_i (3-vector) and u_i the kernels return the gradient of
//     L = sum_i w_i . a_i + u_i phi_i
// with respect to x_k, m_k and the support radius H_k of every particle,
// where a and phi are the forward outputs over all n particles
// (grav_force_direct.cu:55-502 semantics, with the forward kernels' r = 0
// rule).  The reference has no adjoint; its h-derivative of the softened
// potential is what kernels.f:215-279 writes as dphidh (H = 2 hp).
//
// Per ordered pair (k <- l), dr = x_l - x_k, r = |dr|, the pair factors are
// F = (F_k + F_l) / 2 (a_k += m_l F dr) and P = (P_k + P_l) / 2
// (phi_k += m_l P), each side Newtonian (F = r^-3, P = -1/r) or softened
// (F = A(q)/H^3, P = P(q)/H, q = r/H).  The two terms of L holding the pair,
// m_l (F w_k.dr + u_k P) and m_k (-F w_l.dr + u_l P), give, with
// t = m_l w_k.dr - m_k w_l.dr and s = m_l u_k + m_k u_l:
//     dL/dx_k += F (m_k w_l - m_l w_k) - (F'(r)/r t + F s) dr
//     dL/dm_k += u_l P - F w_l.dr
//     dL/dH_k += (dF_k/dH_k t + dP_k/dH_k s) / 2
// using P'(r)/r = F on every side (a = -grad phi; P'(q) = q A(q)).  Lanes
// keep G1 = sum F w_l, G2 = sum F m_l, G3 = sum -(...) dr and GM, and form
// m_k G1 - w_k G2 + G3 once at the end, so a far pair costs about 30 VALU
// instructions and one v_rsq_f32.
//
// Layout and classification follow the forward pass: Morton-sorted
// particles, 256-particle j-tiles of four 64-particle chunks, box and
// per-lane tests (near_mask) that send every chunk with a possibly softened
// pair to the near kernel.  The far kernel runs the Newtonian pair in FP32
// on hi/lo positions relative to an FP64 origin per workgroup, with FP32
// sums over 32 j folded into FP64.  The near kernel runs in FP64; in the
// FP32C pass it takes the Newtonian pairs of its chunks in FP32 on
// FP64-formed differences and only the softened, self and coincident pairs
// in FP64.
#include "sg_internal.h"
#include "sg_device.h"

#include <math.h>

namespace sg {
namespace {

constexpr int ATS = 128;        // j per LDS stage of the far kernel
constexpr int AFOLD = 32;       // FP32 partial sums span AFOLD j
constexpr int ANEAR_WAVES = 4;  // waves sharing one group's near tiles
constexpr int ACHUNK = 64;      // j per LDS chunk of a near wave

static_assert(TJ % ATS == 0 && ATS % TC == 0 && TC % AFOLD == 0 && ACHUNK == TC, "tile shapes");

// dA/dq divided by q: a polynomial for the Wendland kernels, the cubic
// spline's outer branch carries q^-4.  Term-by-term derivatives of the
// coefficients in poly_AP.
__device__ __forceinline__ double poly_Aq(int nk, double q, double invq)
{
    const double q2 = q * q;
    if (nk == 0) {
        if (q < 0.5) return 96.0 * q - 76.8;
        const double iq2 = invq * invq;
        return (-48.0 + q * (76.8 - 32.000000000001 * q) + 0.200000000001 * iq2 * iq2) * invq;
    }
    if (nk == 1)
        return 2.0 * -187.6875 + q2 * (4.0 * 804.375 + q2 * (6.0 * -4379.375 + q * (7.0 * 9009.0
               + q * (8.0 * -8957.8125 + q * (9.0 * 5005.0 + q * (10.0 * -1515.9375
               + 11.0 * 195.0 * q))))));
    return 2.0 * -115.5 + q2 * (4.0 * 618.75 + q * (5.0 * -1155.0 + q * (6.0 * 962.5
           + q * (7.0 * -396.0 + 8.0 * 65.625 * q))));
}

// One side of a pair: F, P, F'(r)/r and (for the k side) dF/dH, dP/dH.
struct Side {
    double F, P, Fr, FH, PH;
};

template <int NK>
__device__ __forceinline__ void adj_side(double r, double ri, double r2, double H, double invH,
                                         Side &s)
{
    if (r2 <= H * H) {
        const double q = r * invH, invq = H * ri;
        double A, P;
        poly_AP(NK, q, invq, A, P);
        const double Aq = poly_Aq(NK, q, invq);
        const double iH2 = invH * invH, iH3 = iH2 * invH;
        s.F = A * iH3;
        s.P = P * invH;
        s.Fr = Aq * iH3 * iH2;
        s.FH = -fma(q * q, Aq, 3.0 * A) * iH2 * iH2;
        s.PH = -fma(q * q, A, P) * iH2;
    } else {
        const double ri2 = ri * ri, ri3 = ri2 * ri;
        s.F = ri3;
        s.P = -ri;
        s.Fr = -3.0 * ri3 * ri2;
        s.FH = 0.0;
        s.PH = 0.0;
    }
}

// Accumulators of one lane: G1 (3), G2, G3 (3), GM, GH.
using Acc9 = double[9];

// Ordered pair (k <- l) in FP64, every branch: self pair, coincident
// distinct pair (the r -> 0 limit), Newtonian, softened.  The kernel choice
// NK is a template parameter (one polynomial per instantiation keeps the
// register footprint, and so the occupancy, of the near kernel in check).
template <bool FAST, int NK>
__device__ __forceinline__ void adj_pair64(double dx, double dy, double dz, double Hk,
                                           double invHk, double mk, double wkx, double wky,
                                           double wkz, double uk, bool gask, double Hl,
                                           double ml, double wlx, double wly, double wlz,
                                           double ul, bool same, Acc9 &g)
{
    if (same) {
        // L holds flag_k m_k u_k P(0)/H_k once
        if (gask) {
            double A0, P0;
            poly_AP(NK, 0.0, 0.0, A0, P0);
            const double pk = P0 * invHk;
            g[7] += uk * pk;
            g[8] -= mk * uk * pk * invHk;
        }
        return;
    }
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double invHl = 1.0 / Hl;
    if (r2 == 0.0) {
        double A0, P0;
        poly_AP(NK, 0.0, 0.0, A0, P0);
        const double F = 0.5 * A0 * (invHk * invHk * invHk + invHl * invHl * invHl);
        g[0] = fma(F, wlx, g[0]);
        g[1] = fma(F, wly, g[1]);
        g[2] = fma(F, wlz, g[2]);
        g[3] = fma(F, ml, g[3]);
        g[7] += ul * 0.5 * P0 * (invHk + invHl);
        g[8] -= 0.5 * P0 * invHk * invHk * fma(ml, uk, mk * ul);
        return;
    }
    const double ri = rsqrt64<FAST>(r2);
    double F, P, Fr, FH, PH;
    if (r2 >= fmax(Hk * Hk, Hl * Hl)) {
        const double ri2 = ri * ri, ri3 = ri2 * ri;
        F = ri3;
        P = -ri;
        Fr = -3.0 * ri3 * ri2;
        FH = 0.0;
        PH = 0.0;
    } else {
        const double r = r2 * ri;
        Side sk, sl;
        adj_side<NK>(r, ri, r2, Hk, invHk, sk);
        adj_side<NK>(r, ri, r2, Hl, invHl, sl);
        F = 0.5 * (sk.F + sl.F);
        P = 0.5 * (sk.P + sl.P);
        Fr = 0.5 * (sk.Fr + sl.Fr);
        FH = 0.5 * sk.FH;
        PH = 0.5 * sk.PH;
    }
    const double wkdr = fma(wkx, dx, fma(wky, dy, wkz * dz));
    const double wldr = fma(wlx, dx, fma(wly, dy, wlz * dz));
    const double t = fma(ml, wkdr, -mk * wldr);
    const double s = fma(ml, uk, mk * ul);
    const double c = -fma(Fr, t, F * s);
    g[0] = fma(F, wlx, g[0]);
    g[1] = fma(F, wly, g[1]);
    g[2] = fma(F, wlz, g[2]);
    g[3] = fma(F, ml, g[3]);
    g[4] = fma(c, dx, g[4]);
    g[5] = fma(c, dy, g[5]);
    g[6] = fma(c, dz, g[6]);
    g[7] += fma(ul, P, -F * wldr);
    g[8] += fma(FH, t, PH * s);
}

// ------------------------------------------------------------ far kernel
// i-side state of one lane in the FP32 far kernel.
struct Lane32 {
    float xh, yh, zh, xl, yl, zl, m, wx, wy, wz, u, h2;   // h2 = H_k^2
};

// One chunk (TC j) of far pairs in FP32.  acc: G1 (3), G2, G3 (3), GM' with
// GM' = sum (u_l / r + w_l.dr / r^3) = -GM, and GH.
template <bool WIDE>
__device__ __forceinline__ void adj_chunk_fp32(const float4 *s_p, const float4 *s_w,
                                               const float4 *s_lo, const Lane32 &q,
                                               double (&acc)[9])
{
    for (int j0 = 0; j0 < TC; j0 += AFOLD) {
        float g1x = 0.0f, g1y = 0.0f, g1z = 0.0f, g2 = 0.0f;
        float g3x = 0.0f, g3y = 0.0f, g3z = 0.0f, gm = 0.0f;
#pragma unroll 2
        for (int jj = j0; jj < j0 + AFOLD; ++jj) {
            const float4 p = s_p[jj];   // x_hi, y_hi, z_hi, m
            const float4 w = s_w[jj];   // wx, wy, wz, u
            float dx = p.x - q.xh, dy = p.y - q.yh, dz = p.z - q.zh;
            if (WIDE) {
                const float4 l = s_lo[jj];
                dx += l.x - q.xl;
                dy += l.y - q.yl;
                dz += l.z - q.zl;
            }
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float ri2 = ri * ri;
            const float ri3 = ri2 * ri;
            const float wkdr = fmaf(q.wx, dx, fmaf(q.wy, dy, q.wz * dz));
            const float wldr = fmaf(w.x, dx, fmaf(w.y, dy, w.z * dz));
            const float t = fmaf(p.w, wkdr, -q.m * wldr);
            const float s = fmaf(p.w, q.u, q.m * w.w);
            const float c = ri3 * fmaf(3.0f * t, ri2, -s);
            g1x = fmaf(ri3, w.x, g1x);
            g1y = fmaf(ri3, w.y, g1y);
            g1z = fmaf(ri3, w.z, g1z);
            g2 = fmaf(ri3, p.w, g2);
            g3x = fmaf(c, dx, g3x);
            g3y = fmaf(c, dy, g3y);
            g3z = fmaf(c, dz, g3z);
            gm = fmaf(w.w, ri, fmaf(ri3, wldr, gm));
        }
        acc[0] += (double)g1x;
        acc[1] += (double)g1y;
        acc[2] += (double)g1z;
        acc[3] += (double)g2;
        acc[4] += (double)g3x;
        acc[5] += (double)g3y;
        acc[6] += (double)g3z;
        acc[7] += (double)gm;
    }
}

// The same in FP64 (the plain FP64 comparison path).
__device__ __forceinline__ void adj_chunk_fp64(const double4 *s_p, const double4 *s_w,
                                               double xr, double yr, double zr, double mk,
                                               double wkx, double wky, double wkz, double uk,
                                               double (&acc)[9])
{
#pragma unroll 2
    for (int jj = 0; jj < TC; ++jj) {
        const double4 p = s_p[jj];
        const double4 w = s_w[jj];
        const double dx = p.x - xr, dy = p.y - yr, dz = p.z - zr;
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        const double ri = rsqrt_pos64(r2);   // far pairs: r^2 >= max(H_k^2, H_l^2) > 0
        const double ri2 = ri * ri;
        const double ri3 = ri2 * ri;
        const double wkdr = fma(wkx, dx, fma(wky, dy, wkz * dz));
        const double wldr = fma(w.x, dx, fma(w.y, dy, w.z * dz));
        const double t = fma(p.w, wkdr, -mk * wldr);
        const double s = fma(p.w, uk, mk * w.w);
        const double c = ri3 * fma(3.0 * t, ri2, -s);
        acc[0] = fma(ri3, w.x, acc[0]);
        acc[1] = fma(ri3, w.y, acc[1]);
        acc[2] = fma(ri3, w.z, acc[2]);
        acc[3] = fma(ri3, p.w, acc[3]);
        acc[4] = fma(c, dx, acc[4]);
        acc[5] = fma(c, dy, acc[5]);
        acc[6] = fma(c, dz, acc[6]);
        acc[7] = fma(w.w, ri, fma(ri3, wldr, acc[7]));
    }
}

__device__ __forceinline__ void adj_tile_range(const AdjArgs &a, int jp, int &t0, int &t1)
{
    t0 = (int)((int64_t)a.ntiles * jp / a.jsplit);
    t1 = (int)((int64_t)a.ntiles * (jp + 1) / a.jsplit);
}

// Grid: niblocks x jsplit workgroups of WG lanes, one i per lane.  Near
// chunks are flagged (chunk bits) for k_adj_near in both precisions: kept
// out of the far kernel, they leave the FP32C far loop a register budget of
// 4 waves per SIMD without spills in it.
template <int WG, bool FP64, int NK>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(FP64 ? 1 : 4)))
void k_adj_far(AdjArgs a)
{
    // FP64: {x, y, z, m} and {wx, wy, wz, u} as double4; FP32: {x_hi, y_hi,
    // z_hi, m}, {wx, wy, wz, u} and {x_lo, y_lo, z_lo, H^2} as float4 (6 KB)
    __shared__ double4 s_buf[2 * ATS];
    double4 *s_p64 = s_buf, *s_w64 = s_buf + ATS;
    float4 *s_p32 = reinterpret_cast<float4 *>(s_buf);
    float4 *s_w32 = s_p32 + ATS, *s_lo32 = s_p32 + 2 * ATS;

    const int ib = (int)(blockIdx.x % (unsigned)a.niblocks);
    const int jp = (int)(blockIdx.x / (unsigned)a.niblocks);
    // partition jp takes every jsplit-th tile (near tiles spread out)
    const int t0 = jp, t1 = a.ntiles, tstep = a.jsplit;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const SortedSet &P = a.p;
    const int64_t ibase = (int64_t)ib * WG, iend = a.n;

    // FP64 origin of the workgroup: its middle i-particle
    const int64_t io = min(ibase + WG / 2, iend - 1);
    const double Ox = P.x[io], Oy = P.y[io], Oz = P.z[io];
    const int64_t i0 = ibase + tid;
    const bool valid = i0 < iend;
    const int64_t i = valid ? i0 : iend - 1;
    const double rx = P.x[i] - Ox, ry = P.y[i] - Oy, rz = P.z[i] - Oz;
    const double mk = P.m[i], wkx = a.wx[i], wky = a.wy[i], wkz = a.wz[i], uk = a.u[i];
    const double Hd = fabs(P.Hs[i]);

    Lane32 q;
    q.xh = (float)rx;
    q.yh = (float)ry;
    q.zh = (float)rz;
    q.xl = (float)(rx - (double)q.xh);
    q.yl = (float)(ry - (double)q.yh);
    q.zl = (float)(rz - (double)q.zh);
    q.m = (float)mk;
    q.wx = (float)wkx;
    q.wy = (float)wky;
    q.wz = (float)wkz;
    q.u = (float)uk;
    q.h2 = (float)(Hd * Hd);

    // the wave's box and largest support radius (the forward kernels' test)
    float bl[3] = {q.xh, q.yh, q.zh}, bh[3] = {q.xh, q.yh, q.zh};
    for (int c = 0; c < 3; ++c) {
        bl[c] = wave_min(bl[c]);
        bh[c] = wave_max(bh[c]);
    }
    const float hmax = wave_max((float)Hd);
    const float px[1] = {q.xh}, py[1] = {q.yh}, pz[1] = {q.zh}, ph[1] = {(float)Hd};
    // chunks nearer than WIDE_K times the wave's largest |x_i - O| take the
    // hi/lo body
    const float wide_d2 = WIDE_K * WIDE_K *
                          wave_max(fmaf(q.xh, q.xh, fmaf(q.yh, q.yh, q.zh * q.zh))) *
                          (1.0f + FAR_MARGIN);
    const bool active = uniform_i(ibase + wv * 64 < iend) != 0;
    unsigned char *flags = a.nearflag + ((int64_t)ib * (WG / 64) + wv) * a.ntiles;

    double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    unsigned nearv = 0;
    for (int t = t0; t < t1; t += tstep) {
        const TileMeta &tm = a.tiles[t];
        const float d2 = box_dist2(tm, Ox, Oy, Oz, bl, bh);
        const unsigned mask =
            active ? near_mask<1>(a.chunks, t, tm, d2, Ox, Oy, Oz, bl, bh, hmax, px, py, pz, ph)
                   : 0u;
        unsigned widem = 0;
        if (!FP64 && uniform_i(d2 < wide_d2) != 0)
            for (int c = 0; c < CHUNKS; ++c)
                if (uniform_i(box_dist2(a.chunks[(int64_t)t * CHUNKS + c], Ox, Oy, Oz, bl, bh) <
                              wide_d2) != 0)
                    widem |= 1u << c;
        if (lane == 0) flags[t] = (unsigned char)mask;
        nearv += __builtin_popcount(mask);
        if (WG == 64 && mask == (1u << CHUNKS) - 1) continue;
        for (int st = 0; st < TJ / ATS; ++st) {
            __syncthreads();
            for (int jj = tid; jj < ATS; jj += WG) {
                const int64_t j = (int64_t)t * TJ + st * ATS + jj;
                const double jx = P.x[j] - Ox, jy = P.y[j] - Oy, jz = P.z[j] - Oz;
                if (FP64) {
                    s_p64[jj] = make_double4(jx, jy, jz, P.m[j]);
                    s_w64[jj] = make_double4(a.wx[j], a.wy[j], a.wz[j], a.u[j]);
                } else {
                    const float hx = (float)jx, hy = (float)jy, hz = (float)jz;
                    const double hj = P.Hs[j];
                    s_p32[jj] = make_float4(hx, hy, hz, (float)P.m[j]);
                    s_w32[jj] = make_float4((float)a.wx[j], (float)a.wy[j], (float)a.wz[j],
                                            (float)a.u[j]);
                    s_lo32[jj] = make_float4((float)(jx - (double)hx), (float)(jy - (double)hy),
                                             (float)(jz - (double)hz), (float)(hj * hj));
                }
            }
            __syncthreads();
            if (!active) continue;
            for (int h = 0; h < ATS / TC; ++h) {
                const int c = st * (ATS / TC) + h;
                const int o = h * TC;
                if ((mask >> c) & 1u) continue;   // k_adj_near
                if (FP64)
                    adj_chunk_fp64(s_p64 + o, s_w64 + o, rx, ry, rz, mk, wkx, wky, wkz, uk, acc);
                else if ((widem >> c) & 1u)
                    adj_chunk_fp32<true>(s_p32 + o, s_w32 + o, s_lo32 + o, q, acc);
                else
                    adj_chunk_fp32<false>(s_p32 + o, s_w32 + o, s_lo32 + o, q, acc);
            }
        }
    }

    if (valid) {
        double *part = a.partial + (int64_t)jp * 5 * a.n;
        part[i] = mk * acc[0] - wkx * acc[3] + acc[4];
        part[a.n + i] = mk * acc[1] - wky * acc[3] + acc[5];
        part[2 * a.n + i] = mk * acc[2] - wkz * acc[3] + acc[6];
        part[3 * a.n + i] = -acc[7];
        part[4 * a.n + i] = acc[8];
    }
    if (a.stats && lane == 0) {
        unsigned long long *st = stat_copy(a.stats);
        atomicAdd(&st[0], (unsigned long long)((t1 - t0 + tstep - 1) / tstep) * CHUNKS);
        atomicAdd(&st[1], (unsigned long long)nearv);
    }
}

// ----------------------------------------------------------- near kernel
struct AdjNearLds {
    double4 p[ANEAR_WAVES][ACHUNK];   // x, y, z relative to the origin; m
    double4 w[ANEAR_WAVES][ACHUNK];   // wx, wy, wz, u
    double2 h[ANEAR_WAVES][ACHUNK];   // signed H, original index
    float4 wf[ANEAR_WAVES][ACHUNK];   // FAST: wx, wy, wz, u in FP32
    float2 mf[ANEAR_WAVES][ACHUNK];   // FAST: m, H^2 in FP32
};

// FAST: one Newtonian pair in FP32 on an FP64-formed difference, the far
// kernel's body (gm holds -GM); ri = 0 adds nothing.
__device__ __forceinline__ void adj_newton_fp32(float dx, float dy, float dz, float ri, float ml,
                                                float4 wl, const Lane32 &q, float (&f)[8])
{
    const float ri2 = ri * ri;
    const float ri3 = ri2 * ri;
    const float wkdr = fmaf(q.wx, dx, fmaf(q.wy, dy, q.wz * dz));
    const float wldr = fmaf(wl.x, dx, fmaf(wl.y, dy, wl.z * dz));
    const float t = fmaf(ml, wkdr, -q.m * wldr);
    const float s = fmaf(ml, q.u, q.m * wl.w);
    const float c = ri3 * fmaf(3.0f * t, ri2, -s);
    f[0] = fmaf(ri3, wl.x, f[0]);
    f[1] = fmaf(ri3, wl.y, f[1]);
    f[2] = fmaf(ri3, wl.z, f[2]);
    f[3] = fmaf(ri3, ml, f[3]);
    f[4] = fmaf(c, dx, f[4]);
    f[5] = fmaf(c, dy, f[5]);
    f[6] = fmaf(c, dz, f[6]);
    f[7] = fmaf(wl.w, ri, fmaf(ri3, wldr, f[7]));
}

// One workgroup of ANEAR_WAVES waves per 64-particle group and j-partition.
// The waves deal the group's flagged tiles round-robin, stage them in
// 64-particle chunks of their own LDS slice and combine their sums in a
// fixed order before adding them to the far kernel's partials.
template <bool FAST, int NK>
__global__ __launch_bounds__(64 * ANEAR_WAVES) void k_adj_near(AdjArgs a)
{
    __shared__ AdjNearLds lds;
    __shared__ double s_red[(ANEAR_WAVES - 1) * 9 * 64];
    const int64_t gi = blockIdx.x % (uint64_t)a.ngroups;
    const int jp = (int)(blockIdx.x / (uint64_t)a.ngroups);
    int t0, t1;
    adj_tile_range(a, jp, t0, t1);
    const int64_t gbase = gi * 64;
    if (gbase >= a.n) return;
    const unsigned char *flags = a.nearflag + gi * a.ntiles;
    bool any = false;
    for (int t = t0; t < t1 && !any; ++t) any = flags[t] != 0;
    if (!any) return;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const SortedSet &P = a.p;
    const int64_t io = min(gbase + 32, a.n - 1);
    const double Ox = P.x[io], Oy = P.y[io], Oz = P.z[io];
    const int64_t i0 = gbase + lane;
    const bool valid = i0 < a.n;
    const int64_t i = valid ? i0 : a.n - 1;
    const double xr = P.x[i] - Ox, yr = P.y[i] - Oy, zr = P.z[i] - Oz;
    const double Hs = P.Hs[i], Hk = fabs(Hs), invHk = 1.0 / Hk;
    const bool gask = Hs > 0.0;
    const double mk = P.m[i], wkx = a.wx[i], wky = a.wy[i], wkz = a.wz[i], uk = a.u[i];
    const int idk = P.idx[i];
    Lane32 q;   // FAST: the i side in FP32 (positions unused)
    q.xh = q.yh = q.zh = q.xl = q.yl = q.zl = 0.0f;
    q.m = (float)mk;
    q.wx = (float)wkx;
    q.wy = (float)wky;
    q.wz = (float)wkz;
    q.u = (float)uk;
    q.h2 = (float)(Hk * Hk);

    Acc9 g = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int cnt = 0;
    for (int t = t0; t < t1; ++t) {
        const unsigned fl = flags[t];
        if (!fl) continue;
        if ((cnt++ % ANEAR_WAVES) != w) continue;
        for (int c0 = 0; c0 < TJ; c0 += ACHUNK) {
            if (!((fl >> (c0 / TC)) & 1u)) continue;   // a far chunk of a near tile
            const int64_t j = (int64_t)t * TJ + c0 + lane;
            wave_lds_sync();
            lds.p[w][lane] = make_double4(P.x[j] - Ox, P.y[j] - Oy, P.z[j] - Oz, P.m[j]);
            lds.w[w][lane] = make_double4(a.wx[j], a.wy[j], a.wz[j], a.u[j]);
            lds.h[w][lane] = make_double2(P.Hs[j], (double)P.idx[j]);
            if (FAST) {
                const double hj = P.Hs[j];
                lds.wf[w][lane] = make_float4((float)a.wx[j], (float)a.wy[j], (float)a.wz[j],
                                              (float)a.u[j]);
                lds.mf[w][lane] = make_float2((float)P.m[j], (float)(hj * hj));
            }
            wave_lds_sync();
            if (FAST) {
                // Newtonian pairs (r^2 clear of both supports) in FP32 with
                // FP32 sums over the chunk, every lane at once (the others add
                // zero); softened, self and coincident pairs and those within
                // the margin of a support in FP64, each lane walking its own
                // list of them
                float f[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                uint64_t own = 0;
                for (int jj = 0; jj < ACHUNK; ++jj) {
                    const double4 p = lds.p[w][jj];
                    const float dx = (float)(p.x - xr), dy = (float)(p.y - yr),
                                dz = (float)(p.z - zr);
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    const float2 mh = lds.mf[w][jj];
                    const bool newton = r2 > fmaxf(q.h2, mh.y) * (1.0f + FAR_MARGIN);
                    const float ri = newton ? __builtin_amdgcn_rsqf(r2) : 0.0f;
                    adj_newton_fp32(dx, dy, dz, ri, mh.x, lds.wf[w][jj], q, f);
                    if (!newton) own |= 1ull << jj;
                }
                for (int c = 0; c < 7; ++c) g[c] += (double)f[c];
                g[7] -= (double)f[7];
                while (own) {
                    const int jj = __builtin_ctzll(own);
                    own &= own - 1;
                    const double4 p = lds.p[w][jj];
                    const double4 c = lds.w[w][jj];
                    const double2 h = lds.h[w][jj];
                    adj_pair64<true, NK>(p.x - xr, p.y - yr, p.z - zr, Hk, invHk, mk, wkx, wky,
                                         wkz, uk, gask, fabs(h.x), p.w, c.x, c.y, c.z, c.w,
                                         (int)h.y == idk, g);
                }
            } else {
                for (int jj = 0; jj < ACHUNK; ++jj) {
                    const double4 p = lds.p[w][jj];
                    const double4 c = lds.w[w][jj];
                    const double2 h = lds.h[w][jj];
                    adj_pair64<false, NK>(p.x - xr, p.y - yr, p.z - zr, Hk, invHk, mk, wkx,
                                          wky, wkz, uk, gask, fabs(h.x), p.w, c.x, c.y, c.z,
                                          c.w, (int)h.y == idk, g);
                }
            }
        }
    }

    if (w > 0)
        for (int c = 0; c < 9; ++c) s_red[((w - 1) * 9 + c) * 64 + lane] = g[c];
    __syncthreads();
    if (w != 0 || !valid) return;
    for (int v = 0; v < ANEAR_WAVES - 1; ++v)
        for (int c = 0; c < 9; ++c) g[c] += s_red[(v * 9 + c) * 64 + lane];
    double *part = a.partial + (int64_t)jp * 5 * a.n;
    part[i] += mk * g[0] - wkx * g[3] + g[4];
    part[a.n + i] += mk * g[1] - wky * g[3] + g[5];
    part[2 * a.n + i] += mk * g[2] - wkz * g[3] + g[6];
    part[3 * a.n + i] += g[7];
    part[4 * a.n + i] += g[8];
}

// Fixed-order FP64 sum over the j-partitions, scattered to input order.
__global__ __launch_bounds__(256) void k_adj_reduce(AdjArgs a)
{
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= a.n) return;
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int jp = 0; jp < a.jsplit; ++jp) {
        const double *p = a.partial + (int64_t)jp * 5 * a.n;
        for (int c = 0; c < 5; ++c) s[c] += p[c * a.n + i];
    }
    const int o = a.p.idx[i];
    a.gx[o] = s[0];
    a.gy[o] = s[1];
    a.gz[o] = s[2];
    a.gm[o] = s[3];
    a.gH[o] = s[4];
}

}  // namespace

bool adj_config_supported(int wg) { return wg == 64 || wg == 128 || wg == 256; }

template <int WG>
static hipError_t adj_far_launch(int precision, const AdjArgs &a, hipStream_t s)
{
    const dim3 grid((unsigned)a.niblocks * (unsigned)a.jsplit), block(WG);
    if (precision != 0) {
        hipLaunchKernelGGL((k_adj_far<WG, true, 0>), grid, block, 0, s, a);
    } else {
        switch (a.nkernel) {
        case 0: hipLaunchKernelGGL((k_adj_far<WG, false, 0>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_adj_far<WG, false, 1>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_adj_far<WG, false, 2>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_adj_far(int precision, const AdjArgs &a, int wg, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    switch (wg) {
    case 64: return adj_far_launch<64>(precision, a, s);
    case 128: return adj_far_launch<128>(precision, a, s);
    case 256: return adj_far_launch<256>(precision, a, s);
    default: return hipErrorInvalidConfiguration;
    }
}

hipError_t launch_adj_near(int precision, const AdjArgs &a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    // FP64 near pairs for both precisions; the FP32C pass seeds the FP64
    // reciprocal square root in FP32 and refines it with one Newton step
    const dim3 grid((unsigned)(a.ngroups * a.jsplit)), block(64 * ANEAR_WAVES);
    if (precision == 0) {
        switch (a.nkernel) {
        case 0: hipLaunchKernelGGL((k_adj_near<true, 0>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_adj_near<true, 1>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_adj_near<true, 2>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (a.nkernel) {
    case 0: hipLaunchKernelGGL((k_adj_near<false, 0>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_adj_near<false, 1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_adj_near<false, 2>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_adj_reduce(const AdjArgs &a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_reduce, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace sg

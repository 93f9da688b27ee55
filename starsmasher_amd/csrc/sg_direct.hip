// sg_direct.hip -- the O(N^2) pair kernels for gfx950 (CDNA4).
//
// Reference: dev_compute_forces<256> and dev_force_ij_nkernel{0,1,2}
// (pbe/src/SPHgrav_lib2/grav_force_direct.cu:55-502): one i per thread, j in
// 256-wide shared-memory tiles, every pair through the full three-way softened
// branch, float-float (Fast2Sum) accumulation of every term.
//
// Design here (wave64, VALU-bound, no MFMA -- a pairwise sum is not a
// contraction):
//   * particles arrive sorted along a Morton curve (sg_prep.hip), so an
//     i-group of 64*IPL consecutive particles (one wave) and a j-tile of TJ
//     consecutive particles are both spatially compact;
//   * FAR kernel: each wave holds IPL i-particles per lane in registers and
//     tests every j-tile once against its own bounding box.  If the boxes are
//     farther apart than the largest support radius of both, every pair in
//     the tile is Newtonian (the grav_force_direct.cu:69-75 branch) and runs a
//     branch-free 14-instruction body per pair (3 sub, 3 for r^2, rsq, 3 mul,
//     3 fma, 1 add).  Otherwise the tile is flagged for the near kernel;
//   * j data for far tiles is staged through LDS as FP32 {x_hi, y_hi, z_hi, m}
//     and {x_lo, y_lo, z_lo} relative to an FP64 origin taken per workgroup,
//     written from coalesced FP64 global loads and read back as same-address
//     (broadcast) ds_read_b128.  Tiles close to the wave use hi+lo
//     differences; distant ones hi only;
//   * FP32 partial sums over FOLD j are folded into FP64 accumulators (the
//     cross-tile reduction is FP64);
//   * NEAR kernel: the flagged (wave, tile) pairs, every pair in FP64 (the
//     Wendland polynomials carry coefficients up to 9009 that cancel to O(1)
//     near q = 1, which FP32 cannot resolve);
//   * REDUCE kernel: fixed-order FP64 sum over j-partitions, scattered back to
//     input order.
// The far kernel carries no FP64 softening code, so its register budget (and
// occupancy) is set by the 14-instruction loop alone.
//
// Two launch forms share the block-level code: one particle set (grid over
// i-blocks x j-partitions), and a batch of independent sets in one persistent
// launch whose workgroups pull (set, i-block) items from a queue.
#include "sg_internal.h"

#include <math.h>

namespace sg {
namespace {

constexpr int TS = 128;                 // j per LDS stage; TJ / TS stages per tile
constexpr float FAR_MARGIN = 1.0e-3f;   // relative guard on the box-distance test
// Hi-only differences x_j,hi - x_i,hi carry a relative error of about
// 2^-24 (|x_i - O| + |x_j - O| + r) / r <= 2^-24 (2 + 2 |x_i - O| / r), with
// |x_i - O| <= sqrt(3) reach (reach = per-axis max over the wave).  Tiles
// nearer than sqrt(3) WIDE_K reach use hi+lo differences, which bounds the
// hi-only ones by 2^-24 (2 + 2 / WIDE_K) = 6 ulp (3.6e-7).  Waves that
// straddle a jump of the space-filling curve have a large reach; with
// WIDE_K = 0.5 about 2% of tiles take the hi+lo body on the giant-star set.
constexpr float WIDE_K = 0.5f;
constexpr int FOLD = 32;                // FP32 partial sums span FOLD j
constexpr int NEAR_WAVES = 4;           // waves sharing one group's near tiles

static_assert(TJ % TS == 0, "tile must be a whole number of stages");
static_assert(TS % FOLD == 0, "stage must be a whole number of folds");

// ---------------------------------------------------------------- helpers
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

// The part of a launch one far-kernel workgroup works on.
struct Block {
    int64_t ibase;          // first (sorted) i of the workgroup
    int64_t iend;           // end of the i segment the block lies in
    int t0, t1;             // j-tile range
    unsigned char *flags;   // near flags of the workgroup's first wave, [t - t0]
    int64_t fstride;        // flag stride between consecutive waves
    double *part;           // partial sums: part[c * pstride + i], c = ax, ay, az, pot
    int64_t pstride;
};

// The part of a launch one near-kernel workgroup works on (one i-group).
struct Group {
    int64_t gbase, iend;
    int t0, t1;
    const unsigned char *flags;   // [t - t0]
    double *part;
    int64_t pstride;
};

// ------------------------------------------------ far block, FP32 + comp.
template <int IPL, bool WIDE>
__device__ __forceinline__ void far_stage_fp32(const float4 *s_hi, const float4 *s_lo,
                                               const float (&xh)[IPL], const float (&yh)[IPL],
                                               const float (&zh)[IPL], const float (&xl)[IPL],
                                               const float (&yl)[IPL], const float (&zl)[IPL],
                                               double (&ax)[IPL], double (&ay)[IPL],
                                               double (&az)[IPL], double (&ap)[IPL])
{
    for (int j0 = 0; j0 < TS; j0 += FOLD) {
        float fx[IPL], fy[IPL], fz[IPL], fp[IPL];
#pragma unroll
        for (int k = 0; k < IPL; ++k) fx[k] = fy[k] = fz[k] = fp[k] = 0.0f;
#pragma unroll 4
        for (int jj = j0; jj < j0 + FOLD; ++jj) {
            const float4 p = s_hi[jj];
            float4 q;
            if (WIDE) q = s_lo[jj];
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                float dx = p.x - xh[k], dy = p.y - yh[k], dz = p.z - zh[k];
                if (WIDE) {
                    dx += q.x - xl[k];
                    dy += q.y - yl[k];
                    dz += q.z - zl[k];
                }
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const float ri = __builtin_amdgcn_rsqf(r2);
                const float mr = p.w * ri;
                const float mr3 = mr * ri * ri;
                fx[k] = fmaf(mr3, dx, fx[k]);
                fy[k] = fmaf(mr3, dy, fy[k]);
                fz[k] = fmaf(mr3, dz, fz[k]);
                fp[k] += mr;
            }
        }
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            ax[k] += (double)fx[k];
            ay[k] += (double)fy[k];
            az[k] += (double)fz[k];
            ap[k] -= (double)fp[k];
        }
    }
}

template <int WG, int IPL>
__device__ __forceinline__ void far_block_fp32c(const SortedSet &I, const SortedSet &J,
                                                const TileMeta *tiles, const Block &b,
                                                unsigned long long *stats, float4 *s_hi,
                                                float4 *s_lo)
{
    constexpr int WI = 64 * IPL;   // i per wave
    constexpr int GI = WG * IPL;   // i per workgroup
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    // FP64 origin of this workgroup: its middle i-particle.
    const int64_t io = min(b.ibase + GI / 2, b.iend - 1);
    const double Ox = I.x[io], Oy = I.y[io], Oz = I.z[io];

    float xh[IPL], yh[IPL], zh[IPL], xl[IPL], yl[IPL], zl[IPL];
    int64_t ii[IPL];
    bool valid[IPL];
    float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
    float hmax = 0.0f;
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        const int64_t i = b.ibase + wv * WI + k * 64 + lane;
        valid[k] = i < b.iend;
        ii[k] = valid[k] ? i : b.iend - 1;
        const double rx = I.x[ii[k]] - Ox, ry = I.y[ii[k]] - Oy, rz = I.z[ii[k]] - Oz;
        xh[k] = (float)rx;
        yh[k] = (float)ry;
        zh[k] = (float)rz;
        xl[k] = (float)(rx - (double)xh[k]);
        yl[k] = (float)(ry - (double)yh[k]);
        zl[k] = (float)(rz - (double)zh[k]);
        const float H = (float)fabs(I.Hs[ii[k]]);
        bl[0] = fminf(bl[0], xh[k]); bl[1] = fminf(bl[1], yh[k]); bl[2] = fminf(bl[2], zh[k]);
        bh[0] = fmaxf(bh[0], xh[k]); bh[1] = fmaxf(bh[1], yh[k]); bh[2] = fmaxf(bh[2], zh[k]);
        hmax = fmaxf(hmax, H);
    }
    float reach = 0.0f;   // max |x_i - O| over the wave, per axis
    for (int c = 0; c < 3; ++c) {
        bl[c] = wave_min(bl[c]);
        bh[c] = wave_max(bh[c]);
        reach = fmaxf(reach, fmaxf(-bl[c], bh[c]));
    }
    hmax = wave_max(hmax);
    // close tiles take hi+lo differences (reach is per axis, so the
    // Euclidean reach is at most sqrt(3) times it)
    const float wide_d2 = 3.0f * (WIDE_K * reach) * (WIDE_K * reach);

    double ax[IPL], ay[IPL], az[IPL], ap[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) ax[k] = ay[k] = az[k] = ap[k] = 0.0;

    // a wave past the end of its segment (the tail of the last i-block)
    // computes nothing and flags nothing
    const bool active = uniform_i(b.ibase + wv * WI < b.iend) != 0;
    unsigned nearv = 0;
    unsigned char *flags = b.flags + wv * b.fstride;
    for (int t = b.t0; t < b.t1; ++t) {
        const TileMeta &tm = tiles[t];
        const float d2 = box_dist2(tm, Ox, Oy, Oz, bl, bh);
        const bool far = !active || uniform_i(boxes_far(d2, hmax, tm)) != 0;
        const bool wide = uniform_i(d2 < wide_d2) != 0;
        if (lane == 0) flags[t - b.t0] = far ? 0 : 1;
        nearv += far ? 0 : 1;
        if (WG == 64 && !far) continue;
        for (int s = 0; s < TJ / TS; ++s) {
            __syncthreads();
            for (int jj = tid; jj < TS; jj += WG) {
                const int64_t j = (int64_t)t * TJ + s * TS + jj;
                const double rx = J.x[j] - Ox, ry = J.y[j] - Oy, rz = J.z[j] - Oz;
                const float hx = (float)rx, hy = (float)ry, hz = (float)rz;
                s_hi[jj] = make_float4(hx, hy, hz, (float)J.m[j]);
                s_lo[jj] = make_float4((float)(rx - (double)hx), (float)(ry - (double)hy),
                                       (float)(rz - (double)hz), 0.0f);
            }
            __syncthreads();
            if (far && active) {
                if (wide)
                    far_stage_fp32<IPL, true>(s_hi, s_lo, xh, yh, zh, xl, yl, zl, ax, ay, az, ap);
                else
                    far_stage_fp32<IPL, false>(s_hi, s_lo, xh, yh, zh, xl, yl, zl, ax, ay, az, ap);
            }
        }
    }

#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        if (!valid[k]) continue;
        b.part[ii[k]] = ax[k];
        b.part[b.pstride + ii[k]] = ay[k];
        b.part[2 * b.pstride + ii[k]] = az[k];
        b.part[3 * b.pstride + ii[k]] = ap[k];
    }
    if (stats && lane == 0) {
        atomicAdd(&stats[0], (unsigned long long)(b.t1 - b.t0));
        atomicAdd(&stats[1], (unsigned long long)nearv);
    }
}

// ------------------------------------------------------- far block, FP64
template <int WG, int IPL>
__device__ __forceinline__ void far_block_fp64(const SortedSet &I, const SortedSet &J,
                                               const TileMeta *tiles, const Block &b,
                                               unsigned long long *stats, double4 *s_p)
{
    constexpr int WI = 64 * IPL;
    constexpr int GI = WG * IPL;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    const int64_t io = min(b.ibase + GI / 2, b.iend - 1);
    const double Ox = I.x[io], Oy = I.y[io], Oz = I.z[io];

    double xr[IPL], yr[IPL], zr[IPL];
    int64_t ii[IPL];
    bool valid[IPL];
    float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
    float hmax = 0.0f;
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        const int64_t i = b.ibase + wv * WI + k * 64 + lane;
        valid[k] = i < b.iend;
        ii[k] = valid[k] ? i : b.iend - 1;
        xr[k] = I.x[ii[k]] - Ox;
        yr[k] = I.y[ii[k]] - Oy;
        zr[k] = I.z[ii[k]] - Oz;
        // the same FP32 box as the FP32 kernel, so both classify tiles alike
        bl[0] = fminf(bl[0], (float)xr[k]); bl[1] = fminf(bl[1], (float)yr[k]);
        bl[2] = fminf(bl[2], (float)zr[k]);
        bh[0] = fmaxf(bh[0], (float)xr[k]); bh[1] = fmaxf(bh[1], (float)yr[k]);
        bh[2] = fmaxf(bh[2], (float)zr[k]);
        hmax = fmaxf(hmax, (float)fabs(I.Hs[ii[k]]));
    }
    for (int c = 0; c < 3; ++c) {
        bl[c] = wave_min(bl[c]);
        bh[c] = wave_max(bh[c]);
    }
    hmax = wave_max(hmax);

    double ax[IPL], ay[IPL], az[IPL], ap[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) ax[k] = ay[k] = az[k] = ap[k] = 0.0;

    const bool active = uniform_i(b.ibase + wv * WI < b.iend) != 0;
    unsigned nearv = 0;
    unsigned char *flags = b.flags + wv * b.fstride;
    for (int t = b.t0; t < b.t1; ++t) {
        const TileMeta &tm = tiles[t];
        const bool far = !active ||
                         uniform_i(boxes_far(box_dist2(tm, Ox, Oy, Oz, bl, bh), hmax, tm)) != 0;
        if (lane == 0) flags[t - b.t0] = far ? 0 : 1;
        nearv += far ? 0 : 1;
        if (WG == 64 && !far) continue;
        for (int s = 0; s < TJ / TS; ++s) {
            __syncthreads();
            for (int jj = tid; jj < TS; jj += WG) {
                const int64_t j = (int64_t)t * TJ + s * TS + jj;
                s_p[jj] = make_double4(J.x[j] - Ox, J.y[j] - Oy, J.z[j] - Oz, J.m[j]);
            }
            __syncthreads();
            if (!far || !active) continue;
#pragma unroll 2
            for (int jj = 0; jj < TS; ++jj) {
                const double4 q = s_p[jj];
#pragma unroll
                for (int k = 0; k < IPL; ++k) {
                    const double dx = q.x - xr[k], dy = q.y - yr[k], dz = q.z - zr[k];
                    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    const double ri = rsqrt(r2);
                    const double mr = q.w * ri;
                    const double mr3 = mr * ri * ri;
                    ax[k] = fma(mr3, dx, ax[k]);
                    ay[k] = fma(mr3, dy, ay[k]);
                    az[k] = fma(mr3, dz, az[k]);
                    ap[k] -= mr;
                }
            }
        }
    }

#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        if (!valid[k]) continue;
        b.part[ii[k]] = ax[k];
        b.part[b.pstride + ii[k]] = ay[k];
        b.part[2 * b.pstride + ii[k]] = az[k];
        b.part[3 * b.pstride + ii[k]] = ap[k];
    }
    if (stats && lane == 0) {
        atomicAdd(&stats[0], (unsigned long long)(b.t1 - b.t0));
        atomicAdd(&stats[1], (unsigned long long)nearv);
    }
}

// ------------------------------------------------- softened pair, FP64
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

// One side k of a softened pair: adds F(r,H_k) and P(r,H_k) (without m_j).
__device__ __forceinline__ void soft_side(int nk, double r, double ri, double r2, double H,
                                          double invH, double &gacc, double &gpot)
{
    if (r2 <= H * H) {
        double A, P;
        poly_AP(nk, r * invH, H * ri, A, P);
        gacc += A * invH * invH * invH;
        gpot += P * invH;
    } else {
        gacc += ri * ri * ri;
        gpot -= ri;
    }
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

// Ordered pair (i <- j), FP64, the semantics of grav_force_direct.cu:55-396
// with the r == 0 rule documented in include/sphgrav.h.
template <bool FAST>
__device__ __forceinline__ void near_pair(int nk, double dx, double dy, double dz, double Hi,
                                          double invHi, double Hj, double mj, bool gasj,
                                          bool same, double &ax, double &ay, double &az,
                                          double &ap)
{
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double hj2 = Hj * Hj;
    if (r2 >= fmax(Hi * Hi, hj2)) {
        const double ri = rsqrt64<FAST>(r2);
        const double mr = mj * ri;
        const double mr3 = mr * ri * ri;
        ax = fma(mr3, dx, ax);
        ay = fma(mr3, dy, ay);
        az = fma(mr3, dz, az);
        ap -= mr;
    } else if (r2 == 0.0) {
        if (!same || gasj) ap -= self_coeff(nk) * mj * (invHi + 1.0 / Hj);
    } else {
        const double ri = rsqrt64<FAST>(r2);
        const double r = r2 * ri;
        double gacc = 0.0, gpot = 0.0;
        soft_side(nk, r, ri, r2, Hi, invHi, gacc, gpot);
        soft_side(nk, r, ri, r2, Hj, 1.0 / Hj, gacc, gpot);
        const double hm = 0.5 * mj;
        const double f = hm * gacc;
        ax = fma(f, dx, ax);
        ay = fma(f, dy, ay);
        az = fma(f, dz, az);
        ap = fma(hm, gpot, ap);
    }
}

// ------------------------------------------------------------ near group
// One workgroup of NEAR_WAVES waves per i-group.  The waves deal the group's
// flagged tiles round-robin among themselves, stage each tile in their own
// LDS slice, and at the end combine their sums in a fixed order (wave
// 0 + 1 + 2 + 3) before adding them to the far kernel's partials.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ bool group_has_near(const Group &g)
{
    for (int t = g.t0; t < g.t1; ++t)
        if (g.flags[t - g.t0]) return true;
    return false;
}

struct NearLds {
    double4 p[NEAR_WAVES][TJ];   // x, y, z relative to the origin; m
    double2 h[NEAR_WAVES][TJ];   // signed H, original index (as double)
};

template <int IPL, bool FAST>
__device__ __forceinline__ void near_group(const SortedSet &I, const SortedSet &J, int nk,
                                           const Group &g, NearLds &lds, double *s_red)
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t io = min(g.gbase + 32 * IPL, g.iend - 1);
    const double Ox = I.x[io], Oy = I.y[io], Oz = I.z[io];

    double xr[IPL], yr[IPL], zr[IPL], Hi[IPL], invHi[IPL];
    double ax[IPL], ay[IPL], az[IPL], ap[IPL];
    int idi[IPL];
    int64_t ii[IPL];
    bool valid[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        const int64_t i = g.gbase + k * 64 + lane;
        valid[k] = i < g.iend;
        ii[k] = valid[k] ? i : g.iend - 1;
        xr[k] = I.x[ii[k]] - Ox;
        yr[k] = I.y[ii[k]] - Oy;
        zr[k] = I.z[ii[k]] - Oz;
        Hi[k] = fabs(I.Hs[ii[k]]);
        invHi[k] = 1.0 / Hi[k];
        idi[k] = I.idx[ii[k]];
        ax[k] = ay[k] = az[k] = ap[k] = 0.0;
    }

    int cnt = 0;
    for (int t = g.t0; t < g.t1; ++t) {
        if (!g.flags[t - g.t0]) continue;
        if ((cnt++ % NEAR_WAVES) != w) continue;
        const int64_t j0 = (int64_t)t * TJ;
        wave_lds_sync();
        for (int q = 0; q < TJ / 64; ++q) {
            const int jj = q * 64 + lane;
            const int64_t j = j0 + jj;
            lds.p[w][jj] = make_double4(J.x[j] - Ox, J.y[j] - Oy, J.z[j] - Oz, J.m[j]);
            lds.h[w][jj] = make_double2(J.Hs[j], (double)J.idx[j]);
        }
        wave_lds_sync();
        for (int jj = 0; jj < TJ; ++jj) {
            const double4 p = lds.p[w][jj];
            const double2 h = lds.h[w][jj];
#pragma unroll
            for (int k = 0; k < IPL; ++k)
                near_pair<FAST>(nk, p.x - xr[k], p.y - yr[k], p.z - zr[k], Hi[k], invHi[k],
                                fabs(h.x), p.w, h.x > 0.0, (int)h.y == idi[k], ax[k], ay[k],
                                az[k], ap[k]);
        }
    }

    // fixed-order combination of the waves' sums, added to the partials
    if (w > 0) {
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            double *r = s_red + (((w - 1) * 4) * IPL + k) * 64 + lane;
            r[0] = ax[k];
            r[IPL * 64] = ay[k];
            r[2 * IPL * 64] = az[k];
            r[3 * IPL * 64] = ap[k];
        }
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        for (int v = 0; v < NEAR_WAVES - 1; ++v) {
            const double *r = s_red + ((v * 4) * IPL + k) * 64 + lane;
            ax[k] += r[0];
            ay[k] += r[IPL * 64];
            az[k] += r[2 * IPL * 64];
            ap[k] += r[3 * IPL * 64];
        }
        if (!valid[k]) continue;
        g.part[ii[k]] += ax[k];
        g.part[g.pstride + ii[k]] += ay[k];
        g.part[2 * g.pstride + ii[k]] += az[k];
        g.part[3 * g.pstride + ii[k]] += ap[k];
    }
}

// ======================================================= single-set kernels
__device__ __forceinline__ void tile_range(const PairArgs &a, int jp, int &t0, int &t1)
{
    t0 = (int)((int64_t)a.ntiles * jp / a.jsplit);
    t1 = (int)((int64_t)a.ntiles * (jp + 1) / a.jsplit);
}

template <int WG, int IPL>
__device__ __forceinline__ Block single_block(const PairArgs &a)
{
    const int ib = blockIdx.x % a.niblocks;
    const int jp = blockIdx.x / a.niblocks;
    Block b;
    b.ibase = (int64_t)ib * WG * IPL;
    b.iend = a.ni;
    tile_range(a, jp, b.t0, b.t1);
    b.fstride = a.ntiles;
    b.flags = a.nearflag + (int64_t)ib * (WG / 64) * a.ntiles + b.t0;
    b.part = a.partial + (int64_t)jp * 4 * a.ni;
    b.pstride = a.ni;
    return b;
}

template <int WG, int IPL>
__global__ __launch_bounds__(WG) void k_far_fp32c(PairArgs a)
{
    __shared__ float4 s_hi[TS];
    __shared__ float4 s_lo[TS];
    far_block_fp32c<WG, IPL>(a.i, a.j, a.tiles, single_block<WG, IPL>(a), a.stats, s_hi, s_lo);
}

template <int WG, int IPL>
__global__ __launch_bounds__(WG) void k_far_fp64(PairArgs a)
{
    __shared__ double4 s_p[TS];
    far_block_fp64<WG, IPL>(a.i, a.j, a.tiles, single_block<WG, IPL>(a), a.stats, s_p);
}

template <int IPL, bool FAST>
__global__ __launch_bounds__(64 * NEAR_WAVES) void k_near(PairArgs a)
{
    __shared__ NearLds lds;
    __shared__ double s_red[(NEAR_WAVES - 1) * 4 * IPL * 64];
    const int64_t gi = blockIdx.x % a.ngroups;
    const int jp = (int)(blockIdx.x / a.ngroups);
    Group g;
    g.gbase = gi * (64 * IPL);
    g.iend = a.ni;
    tile_range(a, jp, g.t0, g.t1);
    g.flags = a.nearflag + gi * a.ntiles + g.t0;
    g.part = a.partial + (int64_t)jp * 4 * a.ni;
    g.pstride = a.ni;
    if (g.gbase >= g.iend || !group_has_near(g)) return;
    near_group<IPL, FAST>(a.i, a.j, a.nkernel, g, lds, s_red);
}

// Fixed-order FP64 sum of the j-partition partials, scattered to input order.
__global__ __launch_bounds__(256) void k_reduce_partials(PairArgs a)
{
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= a.ni) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int jp = 0; jp < a.jsplit; ++jp) {
        const double *p = a.partial + (int64_t)jp * 4 * a.ni;
        for (int c = 0; c < 4; ++c) s[c] += p[c * a.ni + i];
    }
    const int64_t o = (int64_t)a.i.idx[i] - a.ibeg;
    a.ax[o] = s[0];
    a.ay[o] = s[1];
    a.az[o] = s[2];
    a.pot[o] = s[3];
}

// ============================================================= batch kernels
template <int WG, int IPL>
__device__ __forceinline__ Block batch_block(const BatchArgs &a, int item)
{
    const int2 it = a.items[item];
    const BatchSet st = a.sets[it.x];
    Block b;
    b.ibase = st.pofs + (int64_t)it.y * WG * IPL;
    b.iend = st.pofs + st.n;
    b.t0 = st.tile0;
    b.t1 = st.tile0 + st.ntiles;
    b.fstride = st.ntiles;
    b.flags = a.nearflag + st.flagofs + (int64_t)it.y * (WG / 64) * st.ntiles;
    b.part = a.partial;
    b.pstride = a.padded;
    return b;
}

// Persistent far kernel over a batch: every workgroup takes the next
// (set, i-block) item from the queue until the queue is empty, so every
// workgroup reaches the exit once all items are taken.
template <int WG, int IPL, bool FP64>
__global__ __launch_bounds__(WG) void k_far_batch(BatchArgs a)
{
    // one TS-entry FP64 tile, or the two TS-entry FP32 tiles (hi, lo)
    __shared__ double4 s_buf[TS];
    float4 *s_hi = reinterpret_cast<float4 *>(s_buf);
    float4 *s_lo = s_hi + TS;
    __shared__ int s_item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int item = s_item;
        if (item >= a.nitems) break;
        const Block b = batch_block<WG, IPL>(a, item);
        if (FP64)
            far_block_fp64<WG, IPL>(a.j, a.j, a.tiles, b, a.stats, s_buf);
        else
            far_block_fp32c<WG, IPL>(a.j, a.j, a.tiles, b, a.stats, s_hi, s_lo);
    }
}

template <int IPL, bool FAST>
__global__ __launch_bounds__(64 * NEAR_WAVES) void k_near_batch(BatchArgs a)
{
    __shared__ NearLds lds;
    __shared__ double s_red[(NEAR_WAVES - 1) * 4 * IPL * 64];
    const int64_t gg = blockIdx.x;
    // the set this group belongs to (sets are ordered by group offset)
    int lo = 0, hi = a.nsets - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (a.sets[mid].gofs <= gg) lo = mid;
        else hi = mid - 1;
    }
    const BatchSet st = a.sets[lo];
    const int64_t lg = gg - st.gofs;
    Group g;
    g.gbase = st.pofs + lg * (64 * IPL);
    g.iend = st.pofs + st.n;
    g.t0 = st.tile0;
    g.t1 = st.tile0 + st.ntiles;
    g.flags = a.nearflag + st.flagofs + lg * st.ntiles;
    g.part = a.partial;
    g.pstride = a.padded;
    if (g.gbase >= g.iend || !group_has_near(g)) return;
    near_group<IPL, FAST>(a.j, a.j, a.nkernel, g, lds, s_red);
}

// Scatter the per-slot sums of a batch back to input order.
__global__ __launch_bounds__(256) void k_reduce_batch(BatchArgs a)
{
    const int64_t s = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (s >= a.padded) return;
    const int o = a.j.idx[s];
    if (o < 0) return;   // padding
    a.ax[o] = a.partial[s];
    a.ay[o] = a.partial[a.padded + s];
    a.az[o] = a.partial[2 * a.padded + s];
    a.pot[o] = a.partial[3 * a.padded + s];
}

inline dim3 grid_of(const PairArgs &a) { return dim3((unsigned)a.niblocks * (unsigned)a.jsplit); }

}  // namespace

#define SG_FAR(KERNEL, WG, IPL)                                                        \
    if (wg == WG && a.ipl == IPL) {                                                    \
        hipLaunchKernelGGL((KERNEL<WG, IPL>), grid_of(a), dim3(WG), 0, s, a);          \
        return hipGetLastError();                                                      \
    }

hipError_t launch_pair_far(int precision, const PairArgs &a, int wg, hipStream_t s)
{
    if (a.ni <= 0) return hipSuccess;
    if (precision == 0) {
        SG_FAR(k_far_fp32c, 64, 1) SG_FAR(k_far_fp32c, 64, 2) SG_FAR(k_far_fp32c, 64, 4)
        SG_FAR(k_far_fp32c, 128, 1) SG_FAR(k_far_fp32c, 128, 2) SG_FAR(k_far_fp32c, 128, 4)
        SG_FAR(k_far_fp32c, 256, 1) SG_FAR(k_far_fp32c, 256, 2) SG_FAR(k_far_fp32c, 256, 4)
    } else {
        SG_FAR(k_far_fp64, 64, 1) SG_FAR(k_far_fp64, 64, 2)
        SG_FAR(k_far_fp64, 128, 1) SG_FAR(k_far_fp64, 128, 2)
        SG_FAR(k_far_fp64, 256, 1) SG_FAR(k_far_fp64, 256, 2)
    }
    return hipErrorInvalidConfiguration;
}

hipError_t launch_pair_near(int precision, const PairArgs &a, hipStream_t s)
{
    if (a.ni <= 0) return hipSuccess;
    const dim3 grid((unsigned)(a.ngroups * a.jsplit)), block(64 * NEAR_WAVES);
    if (precision == 0) {
        switch (a.ipl) {
        case 1: hipLaunchKernelGGL((k_near<1, true>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_near<2, true>), grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_near<4, true>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidConfiguration;
        }
    } else {
        switch (a.ipl) {
        case 1: hipLaunchKernelGGL((k_near<1, false>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_near<2, false>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidConfiguration;
        }
    }
    return hipGetLastError();
}

bool pair_config_supported(int precision, int wg, int ipl)
{
    const bool wg_ok = wg == 64 || wg == 128 || wg == 256;
    if (precision == 0) return wg_ok && (ipl == 1 || ipl == 2 || ipl == 4);
    return wg_ok && (ipl == 1 || ipl == 2);
}

hipError_t launch_reduce_partials(const PairArgs &a, hipStream_t s)
{
    if (a.ni <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)((a.ni + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------ batch launches
template <int WG, int IPL, bool FP64>
static hipError_t far_batch_launch(const BatchArgs &a, int nwg, hipStream_t s)
{
    hipLaunchKernelGGL((k_far_batch<WG, IPL, FP64>), dim3(nwg), dim3(WG), 0, s, a);
    return hipGetLastError();
}

template <int WG, int IPL, bool FP64>
static int far_batch_resident(int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_far_batch<WG, IPL, FP64>, WG, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return per_cu * cus;
}

#define SG_FARB(WG, IPL)                                                                     \
    if (wg == WG && ipl == IPL) {                                                            \
        if (resident_only)                                                                   \
            return precision == 0 ? far_batch_resident<WG, IPL, false>(device)               \
                                  : far_batch_resident<WG, IPL, true>(device);               \
        return (int)(precision == 0 ? far_batch_launch<WG, IPL, false>(a, nwg, s)            \
                                    : far_batch_launch<WG, IPL, true>(a, nwg, s));           \
    }

static int far_batch_dispatch(int precision, const BatchArgs &a, int wg, int ipl, int nwg,
                              hipStream_t s, bool resident_only, int device)
{
    SG_FARB(64, 1) SG_FARB(64, 2) SG_FARB(128, 1) SG_FARB(128, 2) SG_FARB(256, 1)
    SG_FARB(256, 2)
    return resident_only ? 0 : (int)hipErrorInvalidConfiguration;
}

int batch_resident_workgroups(int precision, int wg, int ipl, int device)
{
    BatchArgs dummy{};
    return far_batch_dispatch(precision, dummy, wg, ipl, 0, nullptr, true, device);
}

bool batch_config_supported(int wg, int ipl)
{
    return (wg == 64 || wg == 128 || wg == 256) && (ipl == 1 || ipl == 2);
}

hipError_t launch_batch(int precision, const BatchArgs &a, int wg, int ipl, int nwg, hipStream_t s)
{
    if (a.padded <= 0) return hipSuccess;
    hipError_t e = launch_fill_zero_int((int *)a.counter, 1, s);
    if (e != hipSuccess) return e;
    e = (hipError_t)far_batch_dispatch(precision, a, wg, ipl, nwg, s, false, 0);
    if (e != hipSuccess) return e;
    const dim3 block(64 * NEAR_WAVES), grid((unsigned)a.ngroups);
    if (a.ngroups > 0) {
        if (precision == 0) {
            if (ipl == 1) hipLaunchKernelGGL((k_near_batch<1, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_near_batch<2, true>), grid, block, 0, s, a);
        } else {
            if (ipl == 1) hipLaunchKernelGGL((k_near_batch<1, false>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_near_batch<2, false>), grid, block, 0, s, a);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_reduce_batch, dim3((unsigned)((a.padded + 255) / 256)), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

}  // namespace sg

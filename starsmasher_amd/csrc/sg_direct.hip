// sg_direct.hip -- the O(N^2) pair kernels for gfx950 (CDNA4).
//
// Reference: dev_compute_forces<256> and dev_force_ij_nkernel{0,1,2}
// (pbe/src/SPHgrav_lib2/grav_force_direct.cu:55-502): one i per thread, j in
// 256-wide shared-memory tiles, every pair through the full three-way softened
// branch, float-float (Fast2Sum) accumulation of every term.
//
// Design here (wave64, VALU-bound):
//   * particles arrive sorted along a Morton curve (sg_prep.hip), so an
//     i-group of 64*IPL consecutive particles (one wave) and a j-tile of TJ
//     consecutive particles (four chunks of TC) are both spatially compact;
//   * each wave holds IPL i-particles per lane in registers and classifies
//     every (wave, tile) and then every (wave, chunk) with box tests and a
//     per-lane test (near_mask): a far chunk has every pair Newtonian (the
//     grav_force_direct.cu:69-75 branch) and runs a branch-free body of 13
//     VALU operations plus one v_rsq_f32 per pair;
//   * j data is staged through LDS per workgroup as FP32 {x_hi, y_hi, z_hi, m}
//     and as FP64 {x, y, z} relative to an FP64 origin taken per workgroup,
//     written from coalesced FP64 global loads and read back as same-address
//     (broadcast) LDS reads.  Chunks use hi-only differences, except those
//     nearer to the wave's box than WIDE_K times its largest |x_i - O|,
//     which form the difference in FP64 and round it once;
//   * FP32 partial sums over FOLD j are folded into FP64 accumulators (the
//     cross-tile and cross-partition reductions are FP64);
//   * near chunks (pairs that may be softened) are listed per wave and run
//     after its far loop in the wave's own LDS slice: FP64-formed
//     differences, the FP32 body for the Newtonian lanes, FP64 near_pair for
//     the softened ones, each lane walking its own list of them (the
//     Wendland polynomials carry coefficients up to 9009 that cancel to O(1)
//     near q = 1, which FP32 cannot resolve);
//   * the FP64 comparison pass and the grad-h pass flag near chunks for a
//     separate near kernel instead;
//   * REDUCE kernel: fixed-order FP64 sum over j-partitions, scattered back to
//     input order.
//
// Two launch forms share the block-level code: one particle set (grid over
// i-blocks x j-partitions), and a batch of independent sets in one persistent
// launch whose workgroups pull (set, i-block, j-partition) items from a
// queue.
#include "sg_internal.h"
#include "sg_device.h"

#include <math.h>

namespace sg {
namespace {

constexpr int TS = 128;                 // j per LDS stage; TJ / TS stages per tile
constexpr int FOLD = 32;                // FP32 partial sums span FOLD j
constexpr int NEAR_WAVES = 4;           // waves sharing one group's near tiles

static_assert(TJ % TS == 0, "tile must be a whole number of stages");
static_assert(TS % TC == 0 && TC % FOLD == 0, "stage = whole chunks = whole folds");

// The part of a launch one far-kernel workgroup works on.
struct Block {
    int64_t ibase;          // first (sorted) i of the workgroup
    int64_t iend;           // end of the i segment the block lies in
    int t0, t1, tstep;      // j-tiles t0, t0 + tstep, ... below t1
    int fbase;              // near flags are indexed by t - fbase
    unsigned char *flags;   // near flags of the workgroup's first wave, [t - fbase]
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
    double *psipart;              // psi partial sums of this j-partition (PSI kernels)
};

// ------------------------------------------------ far block, FP32 + comp.
// One Newtonian pair (i <- j) in FP32 from its difference dr and m_j.  FIRST
// starts the partial sums (the first j of a fold writes them, which saves
// clearing them).
template <bool FIRST>
__device__ __forceinline__ void far_pair_fp32(float dx, float dy, float dz, float mj, float &fx,
                                              float &fy, float &fz, float &fp)
{
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float ri = __builtin_amdgcn_rsqf(r2);
    const float mr = mj * ri;
    const float mr3 = mr * ri * ri;
    if (FIRST) {
        fx = mr3 * dx;
        fy = mr3 * dy;
        fz = mr3 * dz;
        fp = mr;
    } else {
        fx = fmaf(mr3, dx, fx);
        fy = fmaf(mr3, dy, fy);
        fz = fmaf(mr3, dz, fz);
        fp += mr;
    }
}

// m_j travels in the FP64 stage as the low half of the w slot.
__device__ __forceinline__ double pack_mass(float m)
{
    const float2 v = make_float2(m, 0.0f);
    return __builtin_bit_cast(double, v);
}

// The difference of one pair for the FP32 body.  Far tiles use hi-only
// differences (x_j,hi - x_i,hi).  WIDE tiles (close relative to the wave's
// spread about the origin) form x_j - x_i in FP64 from the FP64 stage and
// round it once (two ops per axis): exact to half an ulp of r, with no
// shared error from the rounding of x_i.
template <bool WIDE>
__device__ __forceinline__ void far_diff(const float4 *ph, const double4 *pd, int jj, float xh,
                                         float yh, float zh, double xr, double yr, double zr,
                                         float &dx, float &dy, float &dz, float &mj)
{
    if (WIDE) {
        const double4 q = pd[jj];
        dx = (float)(q.x - xr);
        dy = (float)(q.y - yr);
        dz = (float)(q.z - zr);
        mj = __builtin_bit_cast(float2, q.w).x;
    } else {
        const float4 p = ph[jj];
        dx = p.x - xh;
        dy = p.y - yh;
        dz = p.z - zh;
        mj = p.w;
    }
}

// One chunk (TC j) of far pairs from LDS.  The FOLD j of a fold are
// unrolled, so the LDS reads use immediate offsets from one address per fold.
template <int IPL, bool WIDE>
__device__ __forceinline__ void far_chunk_fp32(const float4 *s_hi, const double4 *s_p,
                                               const float (&xh)[IPL], const float (&yh)[IPL],
                                               const float (&zh)[IPL], const double (&xr)[IPL],
                                               const double (&yr)[IPL], const double (&zr)[IPL],
                                               double (&ax)[IPL], double (&ay)[IPL],
                                               double (&az)[IPL], double (&ap)[IPL])
{
#pragma unroll 1
    for (int j0 = 0; j0 < TC; j0 += FOLD) {
        const float4 *ph = s_hi + j0;
        const double4 *pd = s_p + j0;
        float fx[IPL], fy[IPL], fz[IPL], fp[IPL];
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            float dx, dy, dz, mj;
            far_diff<WIDE>(ph, pd, 0, xh[k], yh[k], zh[k], xr[k], yr[k], zr[k], dx, dy, dz, mj);
            far_pair_fp32<true>(dx, dy, dz, mj, fx[k], fy[k], fz[k], fp[k]);
        }
#pragma unroll
        for (int jj = 1; jj < FOLD; ++jj) {
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                float dx, dy, dz, mj;
                far_diff<WIDE>(ph, pd, jj, xh[k], yh[k], zh[k], xr[k], yr[k], zr[k], dx, dy, dz,
                               mj);
                far_pair_fp32<false>(dx, dy, dz, mj, fx[k], fy[k], fz[k], fp[k]);
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

// One side k of a softened pair: adds F(r,H_k) and P(r,H_k) (without m_j).
// The kernel choice NK is a template parameter of the near kernels: with one
// polynomial per instantiation the register footprint drops by a third.
// With PSI it also adds P(q) + q^2 A(q) to *ps when the side is softened: the
// grad-h term's dphidh(2q) = -(P(q) + q^2 A(q))/2 (kernels.f:215-279).
template <int NK, bool PSI = false>
__device__ __forceinline__ void soft_side(double r, double ri, double r2, double H,
                                          double invH, double &gacc, double &gpot,
                                          double *ps = nullptr)
{
    if (r2 <= H * H) {
        const double q = r * invH;
        double A, P;
        poly_AP(NK, q, H * ri, A, P);
        gacc += A * invH * invH * invH;
        gpot += P * invH;
        if (PSI) *ps = fma(q * q, A, P);
    } else {
        gacc += ri * ri * ri;
        gpot -= ri;
    }
}

// Ordered pair (i <- j), FP64, the semantics of grav_force_direct.cu:55-396
// with the r == 0 rule documented in include/sphgrav.h.
// With PSI, ps accumulates m_j (P(q_i) + q_i^2 A(q_i)) over the pairs with
// r <= H_i (the grad-h softening sum before its -2/H_i^2 factor).
// The j side's 1/H_j and H_j^2 come precomputed from the LDS stage (one
// division per staged particle instead of one per pair).
template <bool FAST, int NK, bool PSI = false>
__device__ __forceinline__ void near_pair(double dx, double dy, double dz, double Hi,
                                          double invHi, double hi2, double Hj, double invHj,
                                          double hj2, double mj, bool gasj, bool same,
                                          double &ax, double &ay, double &az, double &ap,
                                          double *ps = nullptr)
{
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    if (r2 >= fmax(hi2, hj2)) {
        const double ri = rsqrt64<FAST>(r2);
        const double mr = mj * ri;
        const double mr3 = mr * ri * ri;
        ax = fma(mr3, dx, ax);
        ay = fma(mr3, dy, ay);
        az = fma(mr3, dz, az);
        ap -= mr;
    } else if (r2 == 0.0) {
        if (!same || gasj) ap -= self_coeff(NK) * mj * (invHi + invHj);
        if (PSI) *ps -= 2.0 * self_coeff(NK) * mj;   // P(0) = -2c
    } else {
        const double ri = rsqrt64<FAST>(r2);
        const double r = r2 * ri;
        double gacc = 0.0, gpot = 0.0;
        double psk = 0.0;
        soft_side<NK, PSI>(r, ri, r2, Hi, invHi, gacc, gpot, &psk);
        soft_side<NK>(r, ri, r2, Hj, invHj, gacc, gpot);
        if (PSI) *ps = fma(mj, psk, *ps);
        const double hm = 0.5 * mj;
        const double f = hm * gacc;
        ax = fma(f, dx, ax);
        ay = fma(f, dy, ay);
        az = fma(f, dz, az);
        ap = fma(hm, gpot, ap);
    }
}

// FP64 i-side softening state of the near pairs.
template <int IPL>
__device__ __forceinline__ void near_i_state(const SortedSet &I, const int64_t (&ii)[IPL],
                                             double (&Hi)[IPL], double (&invHi)[IPL],
                                             double (&hi2)[IPL], int (&idi)[IPL])
{
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        Hi[k] = fabs(I.Hs[ii[k]]);
        invHi[k] = 1.0 / Hi[k];
        hi2[k] = Hi[k] * Hi[k];
        idi[k] = I.idx[ii[k]];
    }
}

// Per-wave LDS of the fused FP32C far kernel: the wave's list of deferred
// near tiles and the slice a near chunk is staged in.
constexpr int NLIST = 128;   // deferred near tiles per wave (more run at once)
struct NearSlice {
    double4 p[TC];           // x, y, z relative to the origin; w: {FP32 m, FP32 H^2}
    double hs[TC];           // signed H (> 0: gas)
    double hinv[TC];         // 1 / |H|
    double m[TC];
    uint32_t list[NLIST];    // (tile - t0) << 4 | chunk mask
};

// Stages chunk c of tile t in the wave's slice and evaluates it.  Each pair
// difference is formed in FP64 and rounded once to FP32 (two ops per axis;
// near pairs are close, where the hi+lo difference would lose the low bits
// of the i position whenever |x_i - O| < |x_j - O|, an error that does not
// average out over j).  The FP32 body then runs every pair with the lanes
// whose pair may be softened (r^2 < max(H_i^2, H_j^2) in FP32, r = 0
// included) masked out; the j where a lane was masked are recorded in a
// wave-uniform bit mask, and those lanes' pairs then run in FP64 through
// near_pair from the slice, each lane over its own list (a pair is taken by
// exactly one of the two, so a borderline FP32 decision costs only
// rounding).  Wave-local: no block barrier.
template <int IPL, int NK>
__device__ __forceinline__ void near_chunk_slice(
    NearSlice &sl, const SortedSet &J, int t, int c, double Ox, double Oy, double Oz,
    const double (&xr)[IPL], const double (&yr)[IPL], const double (&zr)[IPL],
    const double (&Hi)[IPL], const double (&invHi)[IPL], const double (&hi2)[IPL],
    const int (&idi)[IPL], double (&ax)[IPL], double (&ay)[IPL], double (&az)[IPL],
    double (&ap)[IPL], unsigned (&nfp64)[IPL])
{
    const int lane = threadIdx.x & 63;
    uint64_t real;   // wave-uniform: j that are particles (not massless padding)
    {
        const int64_t j = (int64_t)t * TJ + c * TC + lane;
        const double rx = J.x[j] - Ox, ry = J.y[j] - Oy, rz = J.z[j] - Oz, mj = J.m[j];
        const double hs = J.Hs[j];
        const float mh[2] = {(float)mj, (float)(hs * hs)};
        real = __ballot(mj > 0.0);
        wave_lds_sync();
        sl.p[lane] = make_double4(rx, ry, rz, __builtin_bit_cast(double, mh));
        sl.hs[lane] = hs;
        sl.hinv[lane] = 1.0 / fabs(hs);
        sl.m[lane] = mj;
        wave_lds_sync();
    }
    float hi2f[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) hi2f[k] = (float)hi2[k];
    uint64_t jmask = 0;   // wave-uniform: j with at least one masked lane
#pragma unroll 1
    for (int j0 = 0; j0 < TC; j0 += FOLD) {
        float fx[IPL], fy[IPL], fz[IPL], fp[IPL];
#pragma unroll
        for (int k = 0; k < IPL; ++k) fx[k] = fy[k] = fz[k] = fp[k] = 0.0f;
        for (int jj = j0; jj < j0 + FOLD; ++jj) {
            const double4 p = sl.p[jj];
            const float2 mh = __builtin_bit_cast(float2, p.w);   // m_j, H_j^2
            bool any = false;
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                const float dx = (float)(p.x - xr[k]);
                const float dy = (float)(p.y - yr[k]);
                const float dz = (float)(p.z - zr[k]);
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const bool soft = r2 < fmaxf(hi2f[k], mh.y);
                const float ri = soft ? 0.0f : __builtin_amdgcn_rsqf(r2);
                const float mr = mh.x * ri;
                const float mr3 = mr * ri * ri;
                fx[k] = fmaf(mr3, dx, fx[k]);
                fy[k] = fmaf(mr3, dy, fy[k]);
                fz[k] = fmaf(mr3, dz, fz[k]);
                fp[k] += mr;
                any = any || soft;
            }
            if (__ballot(any)) jmask |= 1ull << jj;
        }
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            ax[k] += (double)fx[k];
            ay[k] += (double)fy[k];
            az[k] += (double)fz[k];
            ap[k] -= (double)fp[k];
        }
    }
    // The masked pairs in FP64.  A wave-wide walk over the j of jmask would
    // run near_pair with about one lane in eight active, so the walk only
    // repeats the FP32 decision to give each lane its own list, and each
    // lane then runs its list (the wave takes as many steps as its longest).
    uint64_t own[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) own[k] = 0;
    while (jmask) {
        const int jj = __builtin_ctzll(jmask);
        jmask &= jmask - 1;
        const double4 p = sl.p[jj];
        const float2 mh = __builtin_bit_cast(float2, p.w);
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            // the same FP32 decision as above
            const float dx = (float)(p.x - xr[k]), dy = (float)(p.y - yr[k]),
                        dz = (float)(p.z - zr[k]);
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            if (r2 < fmaxf(hi2f[k], mh.y)) own[k] |= 1ull << jj;
        }
    }
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        // the counter covers pairs of particles; padding pairs (m_j = 0,
        // evaluated all the same) are left out
        nfp64[k] += (unsigned)__builtin_popcountll(own[k] & real);
        while (own[k]) {
            const int jj = __builtin_ctzll(own[k]);
            own[k] &= own[k] - 1;
            const double4 p = sl.p[jj];
            const double mj = sl.m[jj], hs = sl.hs[jj], ha = fabs(hs), hinv = sl.hinv[jj];
            const double ddx = p.x - xr[k], ddy = p.y - yr[k], ddz = p.z - zr[k];
            // the self pair only matters at r = 0: the index is read then
            const bool same = ddx == 0.0 && ddy == 0.0 && ddz == 0.0 &&
                              J.idx[(int64_t)t * TJ + c * TC + jj] == idi[k];
            near_pair<true, NK>(ddx, ddy, ddz, Hi[k], invHi[k], hi2[k], ha, hinv, ha * ha, mj,
                                hs > 0.0, same, ax[k], ay[k], az[k], ap[k]);
        }
    }
}

// FUSED (the forward FP32C pass): the chunks a wave cannot classify as far
// are listed and run by the wave alone after its far loop (in its own LDS
// slice, no block barrier in between, so one wave's near work does not hold
// the others at the stage barriers); a list that fills up runs the rest at
// once.  Not FUSED (the grad-h pass): they are flagged in b.flags for the
// near kernel.
template <int WG, int IPL, int NK, bool FUSED>
__device__ __forceinline__ void far_block_fp32c(const SortedSet &I, const SortedSet &J,
                                                const TileMeta *tiles, const TileMeta *chunks,
                                                const Block &b, unsigned long long *stats,
                                                float4 *s_hi, double4 *s_p, NearSlice *s_near)
{
    constexpr int WI = 64 * IPL;   // i per wave
    constexpr int GI = WG * IPL;   // i per workgroup
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    // FP64 origin of this workgroup: its middle i-particle.
    const int64_t io = min(b.ibase + GI / 2, b.iend - 1);
    const double Ox = I.x[io], Oy = I.y[io], Oz = I.z[io];

    float xh[IPL], yh[IPL], zh[IPL], hf[IPL];
    double xr[IPL], yr[IPL], zr[IPL];
    int64_t ii[IPL];
    bool valid[IPL];
    float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
    float hmax = 0.0f, reach2 = 0.0f;
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        const int64_t i = b.ibase + wv * WI + k * 64 + lane;
        valid[k] = i < b.iend;
        ii[k] = valid[k] ? i : b.iend - 1;
        xr[k] = I.x[ii[k]] - Ox;
        yr[k] = I.y[ii[k]] - Oy;
        zr[k] = I.z[ii[k]] - Oz;
        xh[k] = (float)xr[k];
        yh[k] = (float)yr[k];
        zh[k] = (float)zr[k];
        hf[k] = (float)fabs(I.Hs[ii[k]]);
        bl[0] = fminf(bl[0], xh[k]); bl[1] = fminf(bl[1], yh[k]); bl[2] = fminf(bl[2], zh[k]);
        bh[0] = fmaxf(bh[0], xh[k]); bh[1] = fmaxf(bh[1], yh[k]); bh[2] = fmaxf(bh[2], zh[k]);
        hmax = fmaxf(hmax, hf[k]);
        reach2 = fmaxf(reach2, fmaf(xh[k], xh[k], fmaf(yh[k], yh[k], zh[k] * zh[k])));
    }
    for (int c = 0; c < 3; ++c) {
        bl[c] = wave_min(bl[c]);
        bh[c] = wave_max(bh[c]);
    }
    hmax = wave_max(hmax);
    // close chunks take FP64-formed differences: those nearer than WIDE_K
    // times the wave's largest |x_i - O| (Euclidean)
    const float wide_d2 = WIDE_K * WIDE_K * wave_max(reach2) * (1.0f + FAR_MARGIN);

    double ax[IPL], ay[IPL], az[IPL], ap[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) ax[k] = ay[k] = az[k] = ap[k] = 0.0;

    // a wave past the end of its segment (the tail of the last i-block)
    // computes nothing and flags nothing
    const bool active = uniform_i(b.ibase + wv * WI < b.iend) != 0;
    unsigned nearv = 0, widev = 0;   // wave-uniform counts of near and wide chunk visits
    unsigned nfp64[IPL];             // per lane: pairs evaluated in FP64 (near tail)
#pragma unroll
    for (int k = 0; k < IPL; ++k) nfp64[k] = 0;
    int nlist = 0;   // wave-uniform
    unsigned char *flags = b.flags + wv * b.fstride;
    for (int t = b.t0; t < b.t1; t += b.tstep) {
        const TileMeta &tm = tiles[t];
        const float d2 = box_dist2(tm, Ox, Oy, Oz, bl, bh);
        const unsigned mask =
            active ? near_mask<IPL>(chunks, t, tm, d2, Ox, Oy, Oz, bl, bh, hmax, xh, yh, zh, hf)
                   : 0u;
        // chunks of a close tile that are themselves close
        unsigned widem = 0;
        if (uniform_i(d2 < wide_d2) != 0)
            for (int c = 0; c < CHUNKS; ++c)
                if (uniform_i(box_dist2(chunks[(int64_t)t * CHUNKS + c], Ox, Oy, Oz, bl, bh) <
                              wide_d2) != 0)
                    widem |= 1u << c;
        nearv += __builtin_popcount(mask);
        if (active) widev += __builtin_popcount(widem & ~mask);
        bool near_tile_now = false;   // a full list: this tile's near chunks run below
        if (!FUSED) {
            if (lane == 0) flags[t - b.fbase] = (unsigned char)mask;
        } else if (mask) {
            if (nlist < NLIST) {
                if (lane == 0)
                    s_near[wv].list[nlist] = ((uint32_t)((t - b.t0) / b.tstep) << 4) | mask;
                ++nlist;
            } else {
                near_tile_now = true;
            }
        }
        if (FUSED && near_tile_now) {
            double Hi[IPL], invHi[IPL], hi2[IPL];
            int idi[IPL];
            near_i_state<IPL>(I, ii, Hi, invHi, hi2, idi);
            for (int c = 0; c < CHUNKS; ++c)
                if ((mask >> c) & 1u)
                    near_chunk_slice<IPL, NK>(s_near[wv], J, t, c, Ox, Oy, Oz, xr, yr, zr, Hi,
                                              invHi, hi2, idi, ax, ay, az, ap, nfp64);
        }
        if (WG == 64 && mask == (1u << CHUNKS) - 1) continue;
        for (int s = 0; s < TJ / TS; ++s) {
            __syncthreads();
            for (int jj = tid; jj < TS; jj += WG) {
                const int64_t j = (int64_t)t * TJ + s * TS + jj;
                const double rx = J.x[j] - Ox, ry = J.y[j] - Oy, rz = J.z[j] - Oz;
                const float mf = (float)J.m[j];
                s_hi[jj] = make_float4((float)rx, (float)ry, (float)rz, mf);
                s_p[jj] = make_double4(rx, ry, rz, pack_mass(mf));
            }
            __syncthreads();
            if (!active) continue;
            for (int h = 0; h < TS / TC; ++h) {
                const int c = s * (TS / TC) + h;
                if ((mask >> c) & 1u) continue;
                if ((widem >> c) & 1u)
                    far_chunk_fp32<IPL, true>(s_hi + h * TC, s_p + h * TC, xh, yh, zh, xr, yr, zr,
                                              ax, ay, az, ap);
                else
                    far_chunk_fp32<IPL, false>(s_hi + h * TC, s_p + h * TC, xh, yh, zh, xr, yr, zr,
                                               ax, ay, az, ap);
            }
        }
    }
    // the wave's deferred near chunks
    if (FUSED && nlist > 0) {
        double Hi[IPL], invHi[IPL], hi2[IPL];
        int idi[IPL];
        near_i_state<IPL>(I, ii, Hi, invHi, hi2, idi);
        wave_lds_sync();
        for (int q = 0; q < nlist; ++q) {
            const uint32_t e = s_near[wv].list[q];
            const int t = b.t0 + (int)(e >> 4) * b.tstep;
            for (int c = 0; c < CHUNKS; ++c)
                if ((e >> c) & 1u)
                    near_chunk_slice<IPL, NK>(s_near[wv], J, t, c, Ox, Oy, Oz, xr, yr, zr, Hi,
                                              invHi, hi2, idi, ax, ay, az, ap, nfp64);
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
    if (stats) {
        // i past the end of the set (copies of the last one) are not counted
        unsigned f = 0;
#pragma unroll
        for (int k = 0; k < IPL; ++k) f += valid[k] ? nfp64[k] : 0u;
        const unsigned long long fw = wave_sum_count(f);
        if (lane == 0) {
            unsigned long long *st = stat_copy(stats);
            atomicAdd(&st[0], (unsigned long long)((b.t1 - b.t0 + b.tstep - 1) / b.tstep) * CHUNKS);
            atomicAdd(&st[1], (unsigned long long)nearv);
            atomicAdd(&st[2], (unsigned long long)widev);
            atomicAdd(&st[3], fw);
        }
    }
}

// ------------------------------------------------------- far block, FP64
template <int WG, int IPL>
__device__ __forceinline__ void far_block_fp64(const SortedSet &I, const SortedSet &J,
                                               const TileMeta *tiles, const TileMeta *chunks,
                                               const Block &b, unsigned long long *stats,
                                               double4 *s_p)
{
    constexpr int WI = 64 * IPL;
    constexpr int GI = WG * IPL;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    const int64_t io = min(b.ibase + GI / 2, b.iend - 1);
    const double Ox = I.x[io], Oy = I.y[io], Oz = I.z[io];

    double xr[IPL], yr[IPL], zr[IPL];
    float xf[IPL], yf[IPL], zf[IPL], hf[IPL];
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
        // the same FP32 positions as the FP32 kernel, so both classify tiles alike
        xf[k] = (float)xr[k];
        yf[k] = (float)yr[k];
        zf[k] = (float)zr[k];
        hf[k] = (float)fabs(I.Hs[ii[k]]);
        bl[0] = fminf(bl[0], xf[k]); bl[1] = fminf(bl[1], yf[k]); bl[2] = fminf(bl[2], zf[k]);
        bh[0] = fmaxf(bh[0], xf[k]); bh[1] = fmaxf(bh[1], yf[k]); bh[2] = fmaxf(bh[2], zf[k]);
        hmax = fmaxf(hmax, hf[k]);
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
    for (int t = b.t0; t < b.t1; t += b.tstep) {
        const TileMeta &tm = tiles[t];
        const unsigned mask =
            active ? near_mask<IPL>(chunks, t, tm, box_dist2(tm, Ox, Oy, Oz, bl, bh), Ox, Oy, Oz,
                                    bl, bh, hmax, xf, yf, zf, hf)
                   : 0u;
        if (lane == 0) flags[t - b.fbase] = (unsigned char)mask;
        nearv += __builtin_popcount(mask);
        if (WG == 64 && mask == (1u << CHUNKS) - 1) continue;
        for (int s = 0; s < TJ / TS; ++s) {
            __syncthreads();
            for (int jj = tid; jj < TS; jj += WG) {
                const int64_t j = (int64_t)t * TJ + s * TS + jj;
                s_p[jj] = make_double4(J.x[j] - Ox, J.y[j] - Oy, J.z[j] - Oz, J.m[j]);
            }
            __syncthreads();
            if (!active) continue;
            for (int h = 0; h < TS / TC; ++h) {
                if ((mask >> (s * (TS / TC) + h)) & 1u) continue;
#pragma unroll 2
                for (int jj = h * TC; jj < (h + 1) * TC; ++jj) {
                    const double4 q = s_p[jj];
#pragma unroll
                    for (int k = 0; k < IPL; ++k) {
                        const double dx = q.x - xr[k], dy = q.y - yr[k], dz = q.z - zr[k];
                        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                        const double ri = rsqrt_pos64(r2);   // far: r^2 >= max(H_i^2, H_j^2) > 0
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
        unsigned long long *st = stat_copy(stats);
        atomicAdd(&st[0], (unsigned long long)((b.t1 - b.t0 + b.tstep - 1) / b.tstep) * CHUNKS);
        atomicAdd(&st[1], (unsigned long long)nearv);
    }
}

// ------------------------------------------------------------ near group
// One workgroup of NEAR_WAVES waves per i-group.  The waves deal the group's
// flagged tiles round-robin among themselves, stage each tile in 64-particle
// chunks in their own LDS slice, and at the end combine their sums in a
// fixed order (wave 0 + 1 + 2 + 3) before adding them to the far kernel's
// partials.
__device__ __forceinline__ bool group_has_near(const Group &g)
{
    for (int t = g.t0; t < g.t1; ++t)
        if (g.flags[t - g.t0]) return true;
    return false;
}

constexpr int NEAR_CHUNK = TC;   // j per LDS chunk of a near wave (one per lane)
static_assert(NEAR_CHUNK == 64, "one staged particle per lane");

// 18 KB per workgroup, so LDS does not cap occupancy below the register limit
struct NearLds {
    double4 p[NEAR_WAVES][NEAR_CHUNK];   // x, y, z relative to the origin; m
    double4 h[NEAR_WAVES][NEAR_CHUNK];   // signed H, original index (as double), 1/|H|, H^2
    float2 mh[NEAR_WAVES][NEAR_CHUNK];   // FP32 m, H^2 (FAST kernels)
};

// With PSI (IPL = 1 only) it also forms the grad-h softening term
// psi_i = -(2/H_i^2) sum_j m_j (P(q_i) + q_i^2 A(q_i)) of gas particles i and
// adds it to g.psipart.
template <int IPL, bool FAST, int NK, bool PSI = false>
__device__ __forceinline__ void near_group(const SortedSet &I, const SortedSet &J,
                                           const Group &g, NearLds &lds, double *s_red)
{
    static_assert(!PSI || IPL == 1, "psi is formed with one i per lane");
    constexpr int NC = PSI ? 5 : 4;   // sums per i: ax, ay, az, pot (, psi)
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t io = min(g.gbase + 32 * IPL, g.iend - 1);
    const double Ox = I.x[io], Oy = I.y[io], Oz = I.z[io];

    double xr[IPL], yr[IPL], zr[IPL], Hi[IPL], invHi[IPL], hi2[IPL];
    double ax[IPL], ay[IPL], az[IPL], ap[IPL], ps[IPL];
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
        hi2[k] = Hi[k] * Hi[k];
        idi[k] = I.idx[ii[k]];
        ax[k] = ay[k] = az[k] = ap[k] = ps[k] = 0.0;
    }

    // FP32 H_i^2 for the FAST kernels' softened-pair test
    float hi2f[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) hi2f[k] = (float)hi2[k];

    int cnt = 0;
    for (int t = g.t0; t < g.t1; ++t) {
        const unsigned fl = g.flags[t - g.t0];
        if (!fl) continue;
        if ((cnt++ % NEAR_WAVES) != w) continue;
        for (int c0 = 0; c0 < TJ; c0 += NEAR_CHUNK) {
            if (!((fl >> (c0 / TC)) & 1u)) continue;   // a far chunk of a near tile
            const int64_t j = (int64_t)t * TJ + c0 + lane;
            wave_lds_sync();
            const double jx = J.x[j] - Ox, jy = J.y[j] - Oy, jz = J.z[j] - Oz;
            lds.p[w][lane] = make_double4(jx, jy, jz, J.m[j]);
            const double hs = J.Hs[j], ha = fabs(hs);
            lds.h[w][lane] = make_double4(hs, (double)J.idx[j], 1.0 / ha, ha * ha);
            if (FAST) lds.mh[w][lane] = make_float2((float)J.m[j], (float)(ha * ha));
            wave_lds_sync();
            if (FAST) {
                // Newtonian pairs in FP32 (differences formed in FP64 and
                // rounded once, as in the forward near tail; sums folded into
                // FP64 every FOLD j); the pairs with r^2 < max(H_i^2, H_j^2),
                // including r = 0, go on their lane's list and then take the
                // FP64 pair, each lane walking its own list.
                uint64_t own[IPL];
#pragma unroll
                for (int k = 0; k < IPL; ++k) own[k] = 0;
                for (int f0 = 0; f0 < NEAR_CHUNK; f0 += FOLD) {
                    float fx[IPL], fy[IPL], fz[IPL], fp[IPL];
#pragma unroll
                    for (int k = 0; k < IPL; ++k) fx[k] = fy[k] = fz[k] = fp[k] = 0.0f;
#pragma unroll 2
                    for (int jj = f0; jj < f0 + FOLD; ++jj) {
                        const double4 pd = lds.p[w][jj];
                        const float2 mh = lds.mh[w][jj];   // m_j, H_j^2
#pragma unroll
                        for (int k = 0; k < IPL; ++k) {
                            const float dx = (float)(pd.x - xr[k]);
                            const float dy = (float)(pd.y - yr[k]);
                            const float dz = (float)(pd.z - zr[k]);
                            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                            const bool soft = r2 < fmaxf(hi2f[k], mh.y);
                            const float ri = soft ? 0.0f : __builtin_amdgcn_rsqf(r2);
                            const float mr = mh.x * ri;
                            const float mr3 = mr * ri * ri;
                            fx[k] = fmaf(mr3, dx, fx[k]);
                            fy[k] = fmaf(mr3, dy, fy[k]);
                            fz[k] = fmaf(mr3, dz, fz[k]);
                            fp[k] += mr;
                            if (soft) own[k] |= 1ull << jj;
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
#pragma unroll
                for (int k = 0; k < IPL; ++k) {
                    while (own[k]) {
                        const int jj = __builtin_ctzll(own[k]);
                        own[k] &= own[k] - 1;
                        const double4 p = lds.p[w][jj];
                        const double4 h = lds.h[w][jj];
                        near_pair<FAST, NK, PSI>(p.x - xr[k], p.y - yr[k], p.z - zr[k], Hi[k],
                                                 invHi[k], hi2[k], fabs(h.x), h.z, h.w, p.w,
                                                 h.x > 0.0, (int)h.y == idi[k], ax[k], ay[k],
                                                 az[k], ap[k], &ps[k]);
                    }
                }
            } else {
#pragma unroll 1
                for (int jj = 0; jj < NEAR_CHUNK; ++jj) {
                    const double4 p = lds.p[w][jj];
                    const double4 h = lds.h[w][jj];
#pragma unroll
                    for (int k = 0; k < IPL; ++k)
                        near_pair<FAST, NK, PSI>(p.x - xr[k], p.y - yr[k], p.z - zr[k], Hi[k],
                                                 invHi[k], hi2[k], fabs(h.x), h.z, h.w, p.w,
                                                 h.x > 0.0, (int)h.y == idi[k], ax[k], ay[k],
                                                 az[k], ap[k], &ps[k]);
                }
            }
        }
    }

    // fixed-order combination of the waves' sums, added to the partials
    if (w > 0) {
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            double *r = s_red + (((w - 1) * NC) * IPL + k) * 64 + lane;
            r[0] = ax[k];
            r[IPL * 64] = ay[k];
            r[2 * IPL * 64] = az[k];
            r[3 * IPL * 64] = ap[k];
            if (PSI) r[4 * IPL * 64] = ps[k];
        }
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        for (int v = 0; v < NEAR_WAVES - 1; ++v) {
            const double *r = s_red + ((v * NC) * IPL + k) * 64 + lane;
            ax[k] += r[0];
            ay[k] += r[IPL * 64];
            az[k] += r[2 * IPL * 64];
            ap[k] += r[3 * IPL * 64];
            if (PSI) ps[k] += r[4 * IPL * 64];
        }
        if (!valid[k]) continue;
        g.part[ii[k]] += ax[k];
        g.part[g.pstride + ii[k]] += ay[k];
        g.part[2 * g.pstride + ii[k]] += az[k];
        g.part[3 * g.pstride + ii[k]] += ap[k];
        // point particles (negative H) carry no psi (advance.f90:750-753)
        if (PSI && I.Hs[ii[k]] > 0.0)
            g.psipart[ii[k]] += -2.0 * invHi[k] * invHi[k] * ps[k];
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
    // partition jp takes every jsplit-th tile: the near tiles of a wave,
    // neighbours along the curve, spread over the partitions instead of
    // piling up in one
    b.t0 = jp;
    b.t1 = a.ntiles;
    b.tstep = a.jsplit;
    b.fbase = 0;
    b.fstride = a.ntiles;
    b.flags = a.nearflag + (int64_t)ib * (WG / 64) * a.ntiles;
    b.part = a.partial + (int64_t)jp * 4 * a.ni;
    b.pstride = a.ni;
    return b;
}

// The forward FP32C pass: far and near chunks in one kernel.  With one i per
// lane the register budget is held to 7 waves per SIMD, the occupancy of the
// far loop alone; the near chunks' FP64 fallback then spills a few values
// to scratch outside the pair loops.
template <int WG, int IPL, int NK>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(IPL == 1 ? 7 : 1)))
void k_far_fp32c(PairArgs a)
{
    __shared__ float4 s_hi[TS];
    __shared__ double4 s_p[TS];
    __shared__ NearSlice s_near[WG / 64];
    far_block_fp32c<WG, IPL, NK, true>(a.i, a.j, a.tiles, a.chunks, single_block<WG, IPL>(a),
                                       a.stats, s_hi, s_p, s_near);
}

// Far chunks only, near chunks flagged (the pass with the grad-h term).
template <int WG>
__global__ __launch_bounds__(WG) void k_far_fp32c_split(PairArgs a)
{
    __shared__ float4 s_hi[TS];
    __shared__ double4 s_p[TS];
    far_block_fp32c<WG, 1, 0, false>(a.i, a.j, a.tiles, a.chunks, single_block<WG, 1>(a),
                                     a.stats, s_hi, s_p, nullptr);
}

template <int WG, int IPL>
__global__ __launch_bounds__(WG) void k_far_fp64(PairArgs a)
{
    __shared__ double4 s_p[TS];
    far_block_fp64<WG, IPL>(a.i, a.j, a.tiles, a.chunks, single_block<WG, IPL>(a), a.stats, s_p);
}

template <int IPL, bool FAST, int NK>
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
    near_group<IPL, FAST, NK>(a.i, a.j, g, lds, s_red);
}

// Near kernel that also forms the grad-h softening term (one i per lane).
template <bool FAST, int NK>
__global__ __launch_bounds__(64 * NEAR_WAVES) void k_near_psi(PairArgs a)
{
    __shared__ NearLds lds;
    __shared__ double s_red[(NEAR_WAVES - 1) * 5 * 64];
    const int64_t gi = blockIdx.x % a.ngroups;
    const int jp = (int)(blockIdx.x / a.ngroups);
    Group g;
    g.gbase = gi * 64;
    g.iend = a.ni;
    tile_range(a, jp, g.t0, g.t1);
    g.flags = a.nearflag + gi * a.ntiles + g.t0;
    g.part = a.partial + (int64_t)jp * 4 * a.ni;
    g.pstride = a.ni;
    g.psipart = a.psipart + (int64_t)jp * a.ni;
    if (g.gbase >= g.iend || !group_has_near(g)) return;
    near_group<1, FAST, NK, true>(a.i, a.j, g, lds, s_red);
}

// ------------------------------------------------- grad-h term on its own
// psi_i = -(2/H_i^2) sum_{j: r <= H_i} m_j (P(q_i) + q_i^2 A(q_i)) of the
// gas particles of the i-set, without the accelerations: one wave per
// 64-particle i-group tests every tile and chunk against the group's box
// (r <= H_i needs only the wave's largest H) and sums the chunks it cannot
// rule out, staged through its own LDS slice.  FP64 throughout.
template <int NK>
__global__ __launch_bounds__(256) void k_psi_only(PairArgs a)
{
    __shared__ double4 s_p[4][TC];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t gi = (int64_t)blockIdx.x * 4 + w;
    if (gi >= a.ngroups) return;   // whole wave; no block barrier below
    const int64_t gbase = gi * 64;
    const int64_t io = min(gbase + 32, a.ni - 1);
    const double Ox = a.i.x[io], Oy = a.i.y[io], Oz = a.i.z[io];
    const bool valid = gbase + lane < a.ni;
    const int64_t i = valid ? gbase + lane : a.ni - 1;
    const double xr = a.i.x[i] - Ox, yr = a.i.y[i] - Oy, zr = a.i.z[i] - Oz;
    const double Hs = a.i.Hs[i], Hi = fabs(Hs), invHi = 1.0 / Hi, hi2 = Hi * Hi;
    float bl[3] = {(float)xr, (float)yr, (float)zr}, bh[3] = {bl[0], bl[1], bl[2]};
    for (int c = 0; c < 3; ++c) {
        bl[c] = wave_min(bl[c]);
        bh[c] = wave_max(bh[c]);
    }
    // only side i matters for psi: a chunk farther than the wave's largest
    // H_i (with the far test's margin) holds no contributing pair
    const float hmax = wave_max((float)Hi);
    const float lim2 = hmax * hmax * (1.0f + FAR_MARGIN);
    double ps = 0.0;
    for (int t = 0; t < a.ntiles; ++t) {
        if (uniform_i(box_dist2(a.tiles[t], Ox, Oy, Oz, bl, bh) <= lim2) == 0) continue;
        for (int c = 0; c < CHUNKS; ++c) {
            const TileMeta &cm = a.chunks[(int64_t)t * CHUNKS + c];
            if (uniform_i(box_dist2(cm, Ox, Oy, Oz, bl, bh) <= lim2) == 0) continue;
            const int64_t j = (int64_t)t * TJ + c * TC + lane;
            wave_lds_sync();
            s_p[w][lane] = make_double4(a.j.x[j] - Ox, a.j.y[j] - Oy, a.j.z[j] - Oz, a.j.m[j]);
            wave_lds_sync();
#pragma unroll 2
            for (int jj = 0; jj < TC; ++jj) {
                const double4 p = s_p[w][jj];
                const double dx = p.x - xr, dy = p.y - yr, dz = p.z - zr;
                const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                if (r2 <= hi2) {
                    const double r = sqrt(r2), q = r * invHi;
                    double A, P;
                    poly_AP(NK, q, q > 0.0 ? 1.0 / q : 0.0, A, P);
                    ps = fma(p.w, fma(q * q, A, P), ps);
                }
            }
        }
    }
    if (!valid) return;
    const int64_t o = (int64_t)a.i.idx[i] - a.ibeg;
    a.psi[o] = Hs > 0.0 ? -2.0 * invHi * invHi * ps : 0.0;
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
    if (a.psi) {
        double p = 0.0;
        for (int jp = 0; jp < a.jsplit; ++jp) p += a.psipart[(int64_t)jp * a.ni + i];
        a.psi[o] = p;
    }
}

// ============================================================= batch kernels
template <int WG, int IPL>
__device__ __forceinline__ Block batch_block(const BatchArgs &a, int item)
{
    const int2 it = a.items[item];
    const BatchSet st = a.sets[it.x];
    const int ib = it.y / st.jsplit, jp = it.y - ib * st.jsplit;
    const int64_t pad = (int64_t)st.ntiles * TJ;
    Block b;
    b.ibase = st.pofs + (int64_t)ib * WG * IPL;
    b.iend = st.pofs + st.n;
    b.t0 = st.tile0 + jp;
    b.t1 = st.tile0 + st.ntiles;
    b.tstep = st.jsplit;
    b.fbase = st.tile0;
    b.fstride = st.ntiles;
    b.flags = a.nearflag + st.flagofs + (int64_t)ib * (WG / 64) * st.ntiles;
    // partition jp's slab; slot i of the set sits at [i - pofs] (partofs >=
    // pofs, so the base stays inside the buffer)
    b.part = a.partial + (st.partofs - st.pofs) + (int64_t)jp * 4 * pad;
    b.pstride = pad;
    return b;
}

// Persistent far kernel over a batch: every workgroup takes the next
// (set, i-block, j-partition) item from the queue until the queue is empty,
// so every workgroup reaches the exit once all items are taken.
template <int WG, int IPL, bool FP64, int NK>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(IPL == 1 && !FP64 ? 7 : 1)))
void k_far_batch(BatchArgs a)
{
    // one TS-entry FP64 tile, or the FP32 hi stage followed by the FP64 stage
    __shared__ double4 s_buf[TS + TS / 2];
    float4 *s_hi = reinterpret_cast<float4 *>(s_buf);
    double4 *s_p = s_buf + TS / 2;
    __shared__ NearSlice s_near[FP64 ? 1 : WG / 64];
    __shared__ int s_item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int item = s_item;
        if (item >= a.nitems) break;
        const Block b = batch_block<WG, IPL>(a, item);
        if (FP64)
            far_block_fp64<WG, IPL>(a.j, a.j, a.tiles, a.chunks, b, a.stats, s_buf);
        else
            far_block_fp32c<WG, IPL, NK, true>(a.j, a.j, a.tiles, a.chunks, b, a.stats, s_hi,
                                               s_p, s_near);
    }
}

template <int IPL, bool FAST, int NK>
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
    // added to the set's first partition slab (after the far kernel)
    g.part = a.partial + (st.partofs - st.pofs);
    g.pstride = (int64_t)st.ntiles * TJ;
    if (g.gbase >= g.iend || !group_has_near(g)) return;
    near_group<IPL, FAST, NK>(a.j, a.j, g, lds, s_red);
}

// Sum the partition slabs of each slot in a fixed order (partition 0
// first) and scatter the sums back to input order.
__global__ __launch_bounds__(256) void k_reduce_batch(BatchArgs a)
{
    const int64_t s = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (s >= a.padded) return;
    const int o = a.j.idx[s];
    if (o < 0) return;   // padding
    // the set of slot s (sets are ordered by first slot)
    int lo = 0, hi = a.nsets - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (a.sets[mid].pofs <= s) lo = mid;
        else hi = mid - 1;
    }
    const BatchSet st = a.sets[lo];
    const int64_t pad = (int64_t)st.ntiles * TJ;
    const double *p = a.partial + st.partofs + (s - st.pofs);
    double v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = p[c * pad];
    for (int q = 1; q < st.jsplit; ++q) {
        p += 4 * pad;
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] += p[c * pad];
    }
    a.ax[o] = v[0];
    a.ay[o] = v[1];
    a.az[o] = v[2];
    a.pot[o] = v[3];
}

inline dim3 grid_of(const PairArgs &a) { return dim3((unsigned)a.niblocks * (unsigned)a.jsplit); }

}  // namespace

#define SG_FAR(KERNEL, WG, IPL)                                                        \
    if (wg == WG && a.ipl == IPL) {                                                    \
        hipLaunchKernelGGL((KERNEL<WG, IPL>), grid_of(a), dim3(WG), 0, s, a);          \
        return hipGetLastError();                                                      \
    }

template <int WG, int IPL>
static hipError_t far_fused_launch(const PairArgs &a, hipStream_t s)
{
    switch (a.nkernel) {
    case 0: hipLaunchKernelGGL((k_far_fp32c<WG, IPL, 0>), grid_of(a), dim3(WG), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_far_fp32c<WG, IPL, 1>), grid_of(a), dim3(WG), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_far_fp32c<WG, IPL, 2>), grid_of(a), dim3(WG), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// FP32C: the fused far kernel (no near kernel follows), or, with the grad-h
// term, the split far kernel with one i per lane.  FP64: the far kernel.
hipError_t launch_pair_far(int precision, const PairArgs &a, int wg, hipStream_t s)
{
    if (a.ni <= 0) return hipSuccess;
    if (precision == 0 && a.psi) {
        if (a.ipl != 1) return hipErrorInvalidConfiguration;
        switch (wg) {
        case 64: hipLaunchKernelGGL((k_far_fp32c_split<64>), grid_of(a), dim3(64), 0, s, a); break;
        case 128: hipLaunchKernelGGL((k_far_fp32c_split<128>), grid_of(a), dim3(128), 0, s, a); break;
        case 256: hipLaunchKernelGGL((k_far_fp32c_split<256>), grid_of(a), dim3(256), 0, s, a); break;
        default: return hipErrorInvalidConfiguration;
        }
        return hipGetLastError();
    }
    if (precision == 0) {
        if (wg == 64 && a.ipl == 1) return far_fused_launch<64, 1>(a, s);
        if (wg == 64 && a.ipl == 2) return far_fused_launch<64, 2>(a, s);
        if (wg == 128 && a.ipl == 1) return far_fused_launch<128, 1>(a, s);
        if (wg == 128 && a.ipl == 2) return far_fused_launch<128, 2>(a, s);
        if (wg == 256 && a.ipl == 1) return far_fused_launch<256, 1>(a, s);
        if (wg == 256 && a.ipl == 2) return far_fused_launch<256, 2>(a, s);
    } else {
        SG_FAR(k_far_fp64, 64, 1) SG_FAR(k_far_fp64, 64, 2)
        SG_FAR(k_far_fp64, 128, 1) SG_FAR(k_far_fp64, 128, 2)
        SG_FAR(k_far_fp64, 256, 1) SG_FAR(k_far_fp64, 256, 2)
    }
    return hipErrorInvalidConfiguration;
}

template <int IPL>
static void near_launch(const PairArgs &a, dim3 grid, dim3 block, hipStream_t s)
{
    switch (a.nkernel) {
    case 0: hipLaunchKernelGGL((k_near<IPL, false, 0>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_near<IPL, false, 1>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((k_near<IPL, false, 2>), grid, block, 0, s, a); break;
    }
}

template <bool FAST>
static void near_psi_launch(const PairArgs &a, dim3 grid, dim3 block, hipStream_t s)
{
    switch (a.nkernel) {
    case 0: hipLaunchKernelGGL((k_near_psi<FAST, 0>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_near_psi<FAST, 1>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((k_near_psi<FAST, 2>), grid, block, 0, s, a); break;
    }
}

// The near kernel of the FP64 pass and of the grad-h pass; the FP32C
// forward pass has none (its far kernel takes the near chunks).
hipError_t launch_pair_near(int precision, const PairArgs &a, hipStream_t s)
{
    if (a.ni <= 0) return hipSuccess;
    if (a.nkernel < 0 || a.nkernel > 2) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(a.ngroups * a.jsplit)), block(64 * NEAR_WAVES);
    if (a.psi) {
        if (a.ipl != 1) return hipErrorInvalidConfiguration;
        hipError_t e = launch_fill_zero_int(reinterpret_cast<int *>(a.psipart),
                                            (int)(2 * (int64_t)a.jsplit * a.ni), s);
        if (e != hipSuccess) return e;
        if (precision == 0) near_psi_launch<true>(a, grid, block, s);
        else near_psi_launch<false>(a, grid, block, s);
        return hipGetLastError();
    }
    if (precision == 0) return hipSuccess;
    switch (a.ipl) {
    case 1: near_launch<1>(a, grid, block, s); break;
    case 2: near_launch<2>(a, grid, block, s); break;
    default: return hipErrorInvalidConfiguration;
    }
    return hipGetLastError();
}

bool pair_config_supported(int precision, int wg, int ipl)
{
    (void)precision;
    return (wg == 64 || wg == 128 || wg == 256) && (ipl == 1 || ipl == 2);
}

hipError_t launch_psi_only(const PairArgs &a, hipStream_t s)
{
    if (a.ni <= 0) return hipSuccess;
    const dim3 grid((unsigned)((a.ngroups + 3) / 4)), block(256);
    switch (a.nkernel) {
    case 0: hipLaunchKernelGGL((k_psi_only<0>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_psi_only<1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_psi_only<2>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const PairArgs &a, hipStream_t s)
{
    if (a.ni <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)((a.ni + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------ batch launches
// FP32C batches run the fused far kernel (one instantiation per softening
// kernel NK); FP64 batches the plain far kernel (NK = 0) and then the near
// kernel.
template <int WG, int IPL, bool FP64, int NK>
static hipError_t far_batch_launch(const BatchArgs &a, int nwg, hipStream_t s)
{
    hipLaunchKernelGGL((k_far_batch<WG, IPL, FP64, NK>), dim3(nwg), dim3(WG), 0, s, a);
    return hipGetLastError();
}

template <int WG, int IPL, bool FP64, int NK>
static int far_batch_resident(int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_far_batch<WG, IPL, FP64, NK>, WG,
                                                     0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return per_cu * cus;
}

template <int WG, int IPL>
static int far_batch_one(int precision, int nk, const BatchArgs &a, int nwg, hipStream_t s,
                         bool resident_only, int device)
{
    if (precision != 0)
        return resident_only ? far_batch_resident<WG, IPL, true, 0>(device)
                             : (int)far_batch_launch<WG, IPL, true, 0>(a, nwg, s);
    switch (nk) {
    case 0:
        return resident_only ? far_batch_resident<WG, IPL, false, 0>(device)
                             : (int)far_batch_launch<WG, IPL, false, 0>(a, nwg, s);
    case 1:
        return resident_only ? far_batch_resident<WG, IPL, false, 1>(device)
                             : (int)far_batch_launch<WG, IPL, false, 1>(a, nwg, s);
    default:
        return resident_only ? far_batch_resident<WG, IPL, false, 2>(device)
                             : (int)far_batch_launch<WG, IPL, false, 2>(a, nwg, s);
    }
}

static int far_batch_dispatch(int precision, int nk, const BatchArgs &a, int wg, int ipl, int nwg,
                              hipStream_t s, bool resident_only, int device)
{
    if (wg == 64 && ipl == 1) return far_batch_one<64, 1>(precision, nk, a, nwg, s, resident_only, device);
    if (wg == 64 && ipl == 2) return far_batch_one<64, 2>(precision, nk, a, nwg, s, resident_only, device);
    if (wg == 128 && ipl == 1) return far_batch_one<128, 1>(precision, nk, a, nwg, s, resident_only, device);
    if (wg == 128 && ipl == 2) return far_batch_one<128, 2>(precision, nk, a, nwg, s, resident_only, device);
    if (wg == 256 && ipl == 1) return far_batch_one<256, 1>(precision, nk, a, nwg, s, resident_only, device);
    if (wg == 256 && ipl == 2) return far_batch_one<256, 2>(precision, nk, a, nwg, s, resident_only, device);
    return resident_only ? 0 : (int)hipErrorInvalidConfiguration;
}

int batch_resident_workgroups(int precision, int nkernel, int wg, int ipl, int device)
{
    BatchArgs dummy{};
    return far_batch_dispatch(precision, nkernel, dummy, wg, ipl, 0, nullptr, true, device);
}

bool batch_config_supported(int wg, int ipl)
{
    return (wg == 64 || wg == 128 || wg == 256) && (ipl == 1 || ipl == 2);
}

template <int IPL>
static void near_batch_launch(const BatchArgs &a, dim3 grid, dim3 block, hipStream_t s)
{
    switch (a.nkernel) {
    case 0: hipLaunchKernelGGL((k_near_batch<IPL, false, 0>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_near_batch<IPL, false, 1>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((k_near_batch<IPL, false, 2>), grid, block, 0, s, a); break;
    }
}

hipError_t launch_batch(int precision, const BatchArgs &a, int wg, int ipl, int nwg, hipStream_t s)
{
    if (a.nkernel < 0 || a.nkernel > 2) return hipErrorInvalidValue;
    if (a.padded <= 0) return hipSuccess;
    hipError_t e = launch_fill_zero_int((int *)a.counter, 1, s);
    if (e != hipSuccess) return e;
    e = (hipError_t)far_batch_dispatch(precision, a.nkernel, a, wg, ipl, nwg, s, false, 0);
    if (e != hipSuccess) return e;
    const dim3 block(64 * NEAR_WAVES), grid((unsigned)a.ngroups);
    if (precision != 0 && a.ngroups > 0) {
        if (ipl == 1) near_batch_launch<1>(a, grid, block, s);
        else near_batch_launch<2>(a, grid, block, s);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_reduce_batch, dim3((unsigned)((a.padded + 255) / 256)), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

}  // namespace sg

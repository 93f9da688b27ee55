// sg_prep.hip -- O(N) preparation of the particle set for the pair kernels.
//
// The reference packs FP64 host arrays into a 32-byte FP32 AoS record per
// particle on the host (grav_force_direct.cu:588-612).  Here the set is
// prepared on the device from FP64 device arrays:
//   1. validate (mass > 0, range2 != 0, finite; :601-602) and bounding box;
//   2. 63-bit Morton keys (21 bits per axis) of the positions;
//   3. radix sort of (key, index) -- sg_sort.hip;
//   4. gather into sorted FP64 structure-of-arrays, H = sqrt|range2| signed
//      by the gas/point flag (:603-605), padded to a multiple of TJ with
//      massless copies of the last particle;
//   5. per-tile and per-chunk (64 particles) bounding box and largest H (the
//      near/far test of the pair kernels).
#include "sg_internal.h"

#include <math.h>

namespace sg {
namespace {

constexpr int RB = 256;   // threads per block for the O(N) kernels

__device__ inline double warp_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double warp_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide min/max of 6 doubles (lo xyz, hi xyz); result valid in thread 0.
__device__ inline void block_minmax6(double v[6])
{
    __shared__ double sh[RB / 64][6];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = 0; c < 3; ++c) v[c] = warp_min(v[c]);
    for (int c = 3; c < 6; ++c) v[c] = warp_max(v[c]);
    if (lane == 0)
        for (int c = 0; c < 6; ++c) sh[w][c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int ww = 1; ww < RB / 64; ++ww) {
            for (int c = 0; c < 3; ++c) v[c] = fmin(v[c], sh[ww][c]);
            for (int c = 3; c < 6; ++c) v[c] = fmax(v[c], sh[ww][c]);
        }
    }
}

__global__ __launch_bounds__(RB) void k_fill_zero_int(int *p, int n)
{
    const int t = blockIdx.x * RB + threadIdx.x;
    if (t < n) p[t] = 0;
}

__global__ __launch_bounds__(RB) void k_validate_bbox(const double *__restrict__ x,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ z,
                                                      const double *__restrict__ m,
                                                      const double *__restrict__ r2, int n,
                                                      double *__restrict__ red, int *err)
{
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    int flags = 0;
    for (int64_t k = blockIdx.x * (int64_t)RB + threadIdx.x; k < n; k += (int64_t)gridDim.x * RB) {
        const double px = x[k], py = y[k], pz = z[k], pm = m[k], pr = r2[k];
        if (!(isfinite(px) && isfinite(py) && isfinite(pz) && isfinite(pm) && isfinite(pr)))
            flags |= ERR_NONFINITE;
        if (!(pm > 0.0)) flags |= ERR_MASS;
        if (pr == 0.0) flags |= ERR_RANGE2;
        v[0] = fmin(v[0], px); v[1] = fmin(v[1], py); v[2] = fmin(v[2], pz);
        v[3] = fmax(v[3], px); v[4] = fmax(v[4], py); v[5] = fmax(v[5], pz);
    }
    if (flags) atomicOr(err, flags);
    block_minmax6(v);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) red[blockIdx.x * 8 + c] = v[c];
}

__global__ __launch_bounds__(RB) void k_bbox_final(const double *__restrict__ red, int nblocks,
                                                   double *__restrict__ bbox)
{
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nblocks; b += RB) {
        for (int c = 0; c < 3; ++c) v[c] = fmin(v[c], red[b * 8 + c]);
        for (int c = 3; c < 6; ++c) v[c] = fmax(v[c], red[b * 8 + c]);
    }
    block_minmax6(v);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) bbox[c] = v[c];
}

// Gravitational potential energy, the diagnostic of output.f:232-252:
// epot = 1/2 sum_i m_i phi_i, in FP64.  Each block sums a fixed contiguous
// range with a fixed tree, and one block sums the block results in order, so
// the value does not depend on scheduling.
__device__ inline double block_sum(double v)
{
    __shared__ double sh[RB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int ww = 0; ww < RB / 64; ++ww) t += sh[ww];
    return t;
}

__global__ __launch_bounds__(RB) void k_epot_part(const double *__restrict__ m,
                                                  const double *__restrict__ pot, int n, int per,
                                                  double *__restrict__ part)
{
    const int64_t b0 = (int64_t)blockIdx.x * per, b1 = min<int64_t>(b0 + per, n);
    double v = 0.0;
    for (int64_t k = b0 + threadIdx.x; k < b1; k += RB) v = fma(m[k], pot[k], v);
    v = block_sum(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(RB) void k_epot_final(const double *__restrict__ part, int nblocks,
                                                   double *__restrict__ out)
{
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += RB) v += part[b];
    v = block_sum(v);
    if (threadIdx.x == 0) out[0] = 0.5 * v;
}

// Spread the low 21 bits of v so that bit b lands at bit 3b (bit interleave
// by successive shift-and-mask; public-domain bit-twiddling construction).
__device__ inline uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__device__ inline uint64_t quant21(double v, double lo, double scale)
{
    double t = (v - lo) * scale;
    t = fmin(fmax(t, 0.0), 2097151.0);
    return (uint64_t)t;
}

__global__ __launch_bounds__(RB) void k_morton(const double *__restrict__ x,
                                               const double *__restrict__ y,
                                               const double *__restrict__ z, int first, int count,
                                               const double *__restrict__ bbox,
                                               uint64_t *__restrict__ keys, int *__restrict__ vals)
{
    const int k = blockIdx.x * RB + threadIdx.x;
    if (k >= count) return;
    const int p = first + k;
    // one cube for all axes keeps the curve isotropic
    const double ext = fmax(fmax(bbox[3] - bbox[0], bbox[4] - bbox[1]), bbox[5] - bbox[2]);
    const double scale = ext > 0.0 ? 2097151.0 / ext : 0.0;
    const uint64_t kx = quant21(x[p], bbox[0], scale);
    const uint64_t ky = quant21(y[p], bbox[1], scale);
    const uint64_t kz = quant21(z[p], bbox[2], scale);
    keys[k] = (spread3(kx) << 2) | (spread3(ky) << 1) | spread3(kz);
    vals[k] = p;
}

__global__ __launch_bounds__(RB) void k_gather(const double *__restrict__ x,
                                               const double *__restrict__ y,
                                               const double *__restrict__ z,
                                               const double *__restrict__ m,
                                               const double *__restrict__ r2,
                                               const int *__restrict__ perm, int count,
                                               int64_t padded, double *__restrict__ ox,
                                               double *__restrict__ oy, double *__restrict__ oz,
                                               double *__restrict__ om, double *__restrict__ oH,
                                               int *__restrict__ oidx)
{
    const int64_t k = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (k >= padded) return;
    const bool real = k < count;
    const int src = perm[real ? k : count - 1];
    const double rr = r2[src];
    const double H = sqrt(rr > 0.0 ? rr : -rr);
    ox[k] = x[src];
    oy[k] = y[src];
    oz[k] = z[src];
    om[k] = real ? m[src] : 0.0;
    oH[k] = (real && rr > 0.0) ? H : -H;   // padding carries no self potential
    oidx[k] = real ? src : -1;
}

// Cotangents of the adjoint into sorted order; padding slots get 0 so they
// add nothing as j.  Any non-finite value sets ERR_NONFINITE.
__global__ __launch_bounds__(RB) void k_gather_cot(const double *__restrict__ wx,
                                                   const double *__restrict__ wy,
                                                   const double *__restrict__ wz,
                                                   const double *__restrict__ u,
                                                   const int *__restrict__ perm, int count,
                                                   int64_t padded, double *__restrict__ owx,
                                                   double *__restrict__ owy,
                                                   double *__restrict__ owz,
                                                   double *__restrict__ ou, int *err)
{
    const int64_t k = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (k >= padded) return;
    if (k >= count) {
        owx[k] = owy[k] = owz[k] = ou[k] = 0.0;
        return;
    }
    const int src = perm[k];
    const double a = wx[src], b = wy[src], c = wz[src], d = u[src];
    owx[k] = a;
    owy[k] = b;
    owz[k] = c;
    ou[k] = d;
    if (!(isfinite(a) && isfinite(b) && isfinite(c) && isfinite(d))) atomicOr(err, (int)ERR_NONFINITE);
}

__global__ __launch_bounds__(TJ) void k_tile_meta(const double *__restrict__ sx,
                                                  const double *__restrict__ sy,
                                                  const double *__restrict__ sz,
                                                  const double *__restrict__ sH,
                                                  TileMeta *__restrict__ tiles,
                                                  TileMeta *__restrict__ chunks)
{
    static_assert(TJ == RB, "one thread per tile particle");
    static_assert(TC == 64 && CHUNKS * TC == TJ, "one wave per chunk");
    const int64_t k = (int64_t)blockIdx.x * TJ + threadIdx.x;
    double v[6] = {sx[k], sy[k], sz[k], sx[k], sy[k], sz[k]};
    double h = fabs(sH[k]);
    // chunk box: one wave per TC particles
    double c[6];
    for (int q = 0; q < 3; ++q) c[q] = warp_min(v[q]);
    for (int q = 3; q < 6; ++q) c[q] = warp_max(v[q]);
    h = warp_max(h);
    if ((threadIdx.x & 63) == 0) {
        TileMeta t;
        for (int q = 0; q < 3; ++q) {
            t.lo[q] = c[q];
            t.hi[q] = c[3 + q];
        }
        t.maxH = h;
        t.pad = 0.0;
        chunks[(int64_t)blockIdx.x * CHUNKS + (threadIdx.x >> 6)] = t;
    }
    block_minmax6(v);
    __shared__ double shH[TJ / 64];
    if ((threadIdx.x & 63) == 0) shH[threadIdx.x >> 6] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < TJ / 64; ++w) h = fmax(h, shH[w]);
        TileMeta t;
        for (int q = 0; q < 3; ++q) {
            t.lo[q] = v[q];
            t.hi[q] = v[3 + q];
        }
        t.maxH = h;
        t.pad = 0.0;
        tiles[blockIdx.x] = t;
    }
}

// Index of the set containing position p, for offsets[0..nsets] ascending.
__device__ inline int set_of(const int64_t *__restrict__ offsets, int nsets, int64_t p)
{
    int lo = 0, hi = nsets - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (offsets[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Batch keys: the set index above the top 42 bits (14 per axis) of the
// Morton code of the common bounding box, so a sort groups sets and orders
// each along the curve.
__global__ __launch_bounds__(RB) void k_morton_batch(const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     const double *__restrict__ z, int n,
                                                     const int64_t *__restrict__ offsets,
                                                     int nsets, const double *__restrict__ bbox,
                                                     uint64_t *__restrict__ keys,
                                                     int *__restrict__ vals)
{
    const int k = blockIdx.x * RB + threadIdx.x;
    if (k >= n) return;
    const double ext = fmax(fmax(bbox[3] - bbox[0], bbox[4] - bbox[1]), bbox[5] - bbox[2]);
    const double scale = ext > 0.0 ? 2097151.0 / ext : 0.0;
    const uint64_t m = (spread3(quant21(x[k], bbox[0], scale)) << 2) |
                       (spread3(quant21(y[k], bbox[1], scale)) << 1) |
                       spread3(quant21(z[k], bbox[2], scale));
    keys[k] = ((uint64_t)set_of(offsets, nsets, k) << 42) | (m >> 21);
    vals[k] = k;
}

// Padded per-set layout: slot -> (set, rank within the set); ranks past the
// set's count are massless copies of its last particle.
__global__ __launch_bounds__(RB) void k_gather_batch(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ m, const double *__restrict__ r2, const int *__restrict__ perm,
    const int64_t *__restrict__ offsets, const BatchSet *__restrict__ sets, int nsets,
    int64_t padded, double *__restrict__ ox, double *__restrict__ oy, double *__restrict__ oz,
    double *__restrict__ om, double *__restrict__ oH, int *__restrict__ oidx)
{
    const int64_t slot = blockIdx.x * (int64_t)RB + threadIdx.x;
    if (slot >= padded) return;
    int lo = 0, hi = nsets - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (sets[mid].pofs <= slot) lo = mid;
        else hi = mid - 1;
    }
    const BatchSet st = sets[lo];
    const int64_t rank = slot - st.pofs;
    const bool real = rank < st.n;
    const int src = perm[offsets[lo] + (real ? rank : st.n - 1)];
    const double rr = r2[src];
    const double H = sqrt(rr > 0.0 ? rr : -rr);
    ox[slot] = x[src];
    oy[slot] = y[src];
    oz[slot] = z[src];
    om[slot] = real ? m[src] : 0.0;
    oH[slot] = (real && rr > 0.0) ? H : -H;
    oidx[slot] = real ? src : -1;
}

inline int nblocks_for(int64_t n) { return (int)((n + RB - 1) / RB); }

}  // namespace

hipError_t launch_morton_batch(const double *x, const double *y, const double *z, int n,
                               const int64_t *offsets, int nsets, const PrepBuffers &b,
                               hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_morton_batch, dim3(nblocks_for(n)), dim3(RB), 0, s, x, y, z, n, offsets,
                       nsets, b.bbox, b.keys_in, b.vals_in);
    return hipGetLastError();
}

hipError_t launch_gather_batch(const double *x, const double *y, const double *z,
                               const double *m, const double *r2, const int *perm,
                               const int64_t *offsets, const BatchSet *sets, int nsets,
                               int64_t padded, double *ox, double *oy, double *oz, double *om,
                               double *oH, int *oidx, hipStream_t s)
{
    if (padded <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_batch, dim3(nblocks_for(padded)), dim3(RB), 0, s, x, y, z, m, r2,
                       perm, offsets, sets, nsets, padded, ox, oy, oz, om, oH, oidx);
    return hipGetLastError();
}

hipError_t launch_potential_energy(const double *m, const double *pot, int n, double *work,
                                   int nwork, hipStream_t s)
{
    // work: nwork - 1 block partials, then the result
    if (n < 1 || nwork < 2) return hipErrorInvalidValue;
    const int nb = std::min(nwork - 1, std::max(1, (n + 4 * RB - 1) / (4 * RB)));
    const int per = (n + nb - 1) / nb;
    hipLaunchKernelGGL(k_epot_part, dim3(nb), dim3(RB), 0, s, m, pot, n, per, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_epot_final, dim3(1), dim3(RB), 0, s, work, nb, work + nwork - 1);
    return hipGetLastError();
}

hipError_t launch_fill_zero_int(int *p, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero_int, dim3(nblocks_for(n)), dim3(RB), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_validate_bbox(const double *x, const double *y, const double *z,
                                const double *m, const double *r2, int n, PrepBuffers &b,
                                hipStream_t s)
{
    hipError_t e = launch_fill_zero_int(b.err, 1, s);
    if (e != hipSuccess) return e;
    const int nb = std::min(b.nred_blocks, std::max(1, nblocks_for(n)));
    hipLaunchKernelGGL(k_validate_bbox, dim3(nb), dim3(RB), 0, s, x, y, z, m, r2, n, b.red, b.err);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(RB), 0, s, b.red, nb, b.bbox);
    return hipGetLastError();
}

hipError_t launch_morton(const double *x, const double *y, const double *z, int first, int count,
                         const PrepBuffers &b, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_morton, dim3(nblocks_for(count)), dim3(RB), 0, s, x, y, z, first, count,
                       b.bbox, b.keys_in, b.vals_in);
    return hipGetLastError();
}

hipError_t launch_gather(const double *x, const double *y, const double *z, const double *m,
                         const double *r2, const int *perm, int count, int64_t padded,
                         double *ox, double *oy, double *oz, double *om, double *oH, int *oidx,
                         hipStream_t s)
{
    if (padded <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3(nblocks_for(padded)), dim3(RB), 0, s, x, y, z, m, r2, perm,
                       count, padded, ox, oy, oz, om, oH, oidx);
    return hipGetLastError();
}

hipError_t launch_gather_cot(const double *wx, const double *wy, const double *wz,
                             const double *u, const int *perm, int count, int64_t padded,
                             double *owx, double *owy, double *owz, double *ou,
                             const PrepBuffers &b, hipStream_t s)
{
    if (padded <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_cot, dim3(nblocks_for(padded)), dim3(RB), 0, s, wx, wy, wz, u,
                       perm, count, padded, owx, owy, owz, ou, b.err);
    return hipGetLastError();
}

hipError_t launch_tile_meta(const double *sx, const double *sy, const double *sz,
                            const double *sH, int ntiles, TileMeta *tiles, TileMeta *chunks,
                            hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tile_meta, dim3(ntiles), dim3(TJ), 0, s, sx, sy, sz, sH, tiles, chunks);
    return hipGetLastError();
}

}  // namespace sg

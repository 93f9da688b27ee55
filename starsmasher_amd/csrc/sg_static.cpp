// sg_static.cpp -- sg_compute_device_static: the evaluation of
// sg_compute_device (same kernels, same tile shape and j-partitions as a
// context at its default configuration) with every scratch buffer carved
// from one file-scope device array, for callers that cannot allocate device
// memory (code under stream capture, serving loops, harnesses that forbid
// hipMalloc).  The entry point calls no hipMalloc / hipFree and creates no
// stream or event.
//
// Limits and sharing:
//   * n <= SG_STATIC_MAX_N particles (the scratch is sized for it, ~40 MB
//     per device, held by each device the library runs on until the process
//     ends);
//   * one array per device, shared by every call on that device and keeping
//     the last call's contents: calls that overlap on one device (other
//     streams, other host threads) race; calls in one stream are ordered.
//   * what must start at zero (validation flags, counters) is cleared by
//     fill kernels of the call itself, and the Morton sort is rocPRIM's
//     merge sort (the same stable order as the context path's radix sort,
//     whose counters are cleared with hipMemsetAsync): the call issues
//     kernels only, so a caller can capture it in a graph.
#include "sphgrav.h"
#include "sg_internal.h"

#include <algorithm>
#include <mutex>

namespace {

using sg::TC;
using sg::TJ;

constexpr int64_t MAXN = SG_STATIC_MAX_N;
constexpr int64_t MAXPAD = (MAXN + TJ - 1) / TJ * TJ;
// Default tile shape of a context (sg_context in sg_capi.cpp): automatic
// workgroup size (128 or 256, by the i-range), one i per lane, automatic
// j-split; at n <= MAXN the split is at most MAXPAD / TJ / 4.  The near flags
// are sized for 256, the larger count of the two sizes.
constexpr int WG = 256, IPL = 1;
constexpr int64_t MAX_JSPLIT = std::min<int64_t>(64, std::max<int64_t>(1, MAXPAD / TJ / 4));
constexpr int NRED_BLOCKS = 1024;
constexpr size_t SORT_RESERVE = (size_t)4 << 20;   // rocPRIM merge-sort temporary storage

constexpr size_t al(size_t b) { return (b + 255) / 256 * 256; }

// Byte offsets of the carve-out (each 256-byte aligned).
struct Layout {
    size_t keys_in, keys_out, vals_in, vals_out, red, bbox, err, stats;
    size_t sj[5], sidx, tiles, chunks, si[5], iidx, partial, nearflag, sort, total;
};

constexpr Layout make_layout()
{
    Layout L{};
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        const size_t at = o;
        o += al(bytes);
        return at;
    };
    L.keys_in = take(sizeof(uint64_t) * MAXN);
    L.keys_out = take(sizeof(uint64_t) * MAXN);
    L.vals_in = take(sizeof(int) * MAXN);
    L.vals_out = take(sizeof(int) * MAXN);
    L.red = take(sizeof(double) * 8 * NRED_BLOCKS);
    L.bbox = take(sizeof(double) * 8);
    L.err = take(sizeof(int) * 4);
    L.stats = take(sizeof(unsigned long long) * sg::STAT_WORDS);
    for (int k = 0; k < 5; ++k) L.sj[k] = take(sizeof(double) * MAXPAD);
    L.sidx = take(sizeof(int) * MAXPAD);
    L.tiles = take(sizeof(sg::TileMeta) * (MAXPAD / TJ));
    L.chunks = take(sizeof(sg::TileMeta) * (MAXPAD / TC));
    for (int k = 0; k < 5; ++k) L.si[k] = take(sizeof(double) * MAXN);
    L.iidx = take(sizeof(int) * MAXN);
    L.partial = take(sizeof(double) * MAX_JSPLIT * 4 * MAXN);
    // FP64 pass: one byte per (wave, tile); waves = 4 per 256-i block
    L.nearflag = take((size_t)((MAXN + WG * IPL - 1) / (WG * IPL)) * (WG / 64) * (MAXPAD / TJ));
    L.sort = take(SORT_RESERVE);
    L.total = o;
    return L;
}

constexpr Layout LAYOUT = make_layout();
static_assert(sizeof(sg::TileMeta) == 64, "tile record size");
static_assert(SG_VALID_MASS == sg::ERR_MASS && SG_VALID_RANGE2 == sg::ERR_RANGE2 &&
                  SG_VALID_NONFINITE == sg::ERR_NONFINITE,
              "validation flags of the header and the device code agree");
static_assert(LAYOUT.total == SG_STATIC_SCRATCH_BYTES,
              "SG_STATIC_SCRATCH_BYTES in sphgrav.h must match the carve-out");

// The scratch array.  In host code `scratch` names a host-side stand-in, not
// device memory: host code uses only the per-device address below.
static __device__ __align__(256) unsigned char scratch[SG_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char *g_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on `device` (the current device), looked
// up once per device.
int scratch_address(int device, unsigned char **out)
{
    if (device < 0 || device >= MAX_DEVICES) return SG_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            sg::record_hip_error(e);
            return SG_ERR_HIP;
        }
        g_ptr[device] = (unsigned char *)p;
    }
    *out = g_ptr[device];
    return SG_OK;
}

template <typename T>
T *at(unsigned char *base, size_t ofs)
{
    return reinterpret_cast<T *>(base + ofs);
}

}  // namespace

extern "C" int sg_compute_device_static(int n, int ibeg, int ni, const double *x,
                                        const double *y, const double *z, const double *m,
                                        const double *range2, int nkernel, int precision,
                                        double *ax, double *ay, double *az, double *pot,
                                        int *err, void *stream)
{
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (precision != SG_PREC_FP32C && precision != SG_PREC_FP64) return SG_ERR_ARG;
    if (n < 1 || ibeg < 0 || ni < 0 || (int64_t)ibeg + ni > n) return SG_ERR_ARG;
    if (n > SG_STATIC_MAX_N) return SG_ERR_SIZE;
    if (!x || !y || !z || !m || !range2) return SG_ERR_ARG;
    if (ni > 0 && (!ax || !ay || !az || !pot)) return SG_ERR_ARG;
    if (ni == 0) return SG_OK;

    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return SG_ERR_NO_DEVICE;
    unsigned char *base = nullptr;
    int rc = scratch_address(device, &base);
    if (rc != SG_OK) return rc;

    const sg::EvalPlan p = sg::plan_eval(n, ni, 0, IPL, 0, precision, false);
    size_t sort_bytes = 0;
    hipError_t e = sg::merge_sort_pairs_temp_bytes(n, &sort_bytes);
    if (e != hipSuccess) {
        sg::record_hip_error(e);
        return SG_ERR_HIP;
    }
    // the carve-out covers every n <= SG_STATIC_MAX_N; checked, not assumed
    if (p.padded > MAXPAD || p.part_elems > MAX_JSPLIT * 4 * MAXN ||
        p.flag_elems > (int64_t)(LAYOUT.sort - LAYOUT.nearflag) || sort_bytes > SORT_RESERVE)
        return SG_ERR_SIZE;

    sg::EvalBuffers b{};
    b.pb.keys_in = at<uint64_t>(base, LAYOUT.keys_in);
    b.pb.keys_out = at<uint64_t>(base, LAYOUT.keys_out);
    b.pb.vals_in = at<int>(base, LAYOUT.vals_in);
    b.pb.vals_out = at<int>(base, LAYOUT.vals_out);
    b.pb.red = at<double>(base, LAYOUT.red);
    b.pb.nred_blocks = NRED_BLOCKS;
    b.pb.bbox = at<double>(base, LAYOUT.bbox);
    // validation flags go to the caller's device int when given
    b.pb.err = err ? err : at<int>(base, LAYOUT.err);
    double **sj[5] = {&b.pb.sx, &b.pb.sy, &b.pb.sz, &b.pb.sm, &b.pb.sH};
    double **si[5] = {&b.pb.ix, &b.pb.iy, &b.pb.iz, &b.pb.im, &b.pb.iH};
    for (int k = 0; k < 5; ++k) {
        *sj[k] = at<double>(base, LAYOUT.sj[k]);
        *si[k] = at<double>(base, LAYOUT.si[k]);
    }
    b.pb.sidx = at<int>(base, LAYOUT.sidx);
    b.pb.iidx = at<int>(base, LAYOUT.iidx);
    b.pb.tiles = at<sg::TileMeta>(base, LAYOUT.tiles);
    b.pb.chunks = at<sg::TileMeta>(base, LAYOUT.chunks);
    b.sort_tmp = base + LAYOUT.sort;
    b.sort_tmp_bytes = sort_bytes;
    b.partial = at<double>(base, LAYOUT.partial);
    b.nearflag = base + LAYOUT.nearflag;
    b.stats = at<unsigned long long>(base, LAYOUT.stats);
    b.psipart = nullptr;
    b.merge_sort = true;   // kernels only: nothing a stream capture would record as a memset

    e = sg::enqueue_eval(p, b, n, ibeg, ni, x, y, z, m, range2, nkernel, precision, ax, ay, az,
                         pot, nullptr, (hipStream_t)stream, nullptr);
    if (e != hipSuccess) {
        sg::record_hip_error(e);
        return SG_ERR_HIP;
    }
    return SG_OK;
}

// sg_capi.cpp -- host layer of libsphgrav_hip.so: contexts, device buffer
// management, the sg_* status-code API and the reference's three
// Fortran-callable symbols (grav_force_direct.cu:648-677).
//
// Differences from the reference host code (grav_force_direct.cu:504-646),
// all by design:
//   * device buffers grow on demand and are reused across calls (the
//     reference frees and re-allocates both buffers on every call,
//     cuVector.h:42-59);
//   * errors are status codes; the legacy symbols turn them into the
//     reference's exit-on-error unless SPHGRAV_NOABORT=1;
//   * the packing is done on the device (sg_prep.hip), not by a host loop.
#include "sphgrav.h"
#include "sg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

using sg::TJ;

struct sg_context {
    int device = 0;
    hipStream_t stream = nullptr;       // stream work is enqueued on
    hipStream_t own_stream = nullptr;   // the context's own stream
    // events on the context stream: start, before far, after far, end
    hipEvent_t ev_p0 = nullptr, ev_k0 = nullptr, ev_kf = nullptr, ev_k1 = nullptr;

    // capacities (elements)
    int64_t cap_n = 0;       // raw-set sized buffers (keys, vals, red)
    int64_t cap_pad = 0;     // sorted j arrays
    int64_t cap_ni = 0;      // sorted i arrays
    int64_t cap_part = 0;    // j-partition partials (doubles)
    int64_t cap_in = 0;      // host-path input staging
    int64_t cap_out = 0;     // host-path output staging
    size_t sort_tmp_bytes = 0;

    sg::PrepBuffers pb{};
    void *sort_tmp = nullptr;
    double *partial = nullptr;
    double *psipart = nullptr;          // grad-h term partial sums
    int64_t cap_psipart = 0;
    unsigned char *nearflag = nullptr;
    int64_t cap_flag = 0;
    // [4]: chunk visits, near chunk visits, wide chunk visits (FP64-formed
    // differences in the far loop), pairs evaluated in FP64 (FP32C pass)
    unsigned long long *stats = nullptr;
    double *in[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double *out[4] = {nullptr, nullptr, nullptr, nullptr};
    // adjoint: sorted cotangents; host-path staging of cotangents and gradients
    double *scot[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t cap_cot = 0;
    double *vin[4] = {nullptr, nullptr, nullptr, nullptr};
    double *vout[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int64_t cap_vin = 0, cap_vout = 0;
    double *psi_out = nullptr;          // host-path staging of the grad-h term
    int64_t cap_psi_out = 0;
    double *epot_work = nullptr;        // block sums and result of sg_potential_energy
    int64_t cap_epot = 0;

    // tile shape per precision; wg 0 = by the size of the i-range
    // (plan_eval), jsplit 0 = automatic
    int wg[2] = {0, 0};
    int ipl[2] = {1, 1};
    int jsplit[2] = {0, 0};
    int last_shape[3] = {0, 0, 0};      // wg, ipl, jsplit of the last evaluation

    bool launched = false;

    // ---- batch of independent sets
    int bwg[2] = {256, 256};
    int bipl[2] = {1, 1};
    struct Plan {
        std::vector<int64_t> offsets;
        int wg = 0, ipl = 0;
        int64_t padded = 0, ngroups = 0, nflags = 0, nparts = 0;
        int nitems = 0;
    } plan;
    int64_t *d_offsets = nullptr;
    int64_t cap_offsets = 0;
    sg::BatchSet *d_sets = nullptr;
    int64_t cap_sets = 0;
    int2 *d_items = nullptr;
    int64_t cap_items = 0;
    unsigned *d_counter = nullptr;
    // captured compute of the last batch call (replayed when nothing changed)
    hipGraphExec_t gexec = nullptr;
    sg::BatchArgs g_args{};
    int g_key[5] = {-1, -1, -1, -1, -1};   // precision, nkernel, wg, ipl, nwg
    hipStream_t g_stream = nullptr;
};

namespace {

std::mutex g_mu;
int g_last_status = SG_OK;

#define HIP_TRY(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) {                        \
            last_hip_error() = e_;                     \
            return SG_ERR_HIP;                         \
        }                                              \
    } while (0)

hipError_t &last_hip_error()
{
    static thread_local hipError_t e = hipSuccess;
    return e;
}

template <typename T>
int grow(T **p, int64_t *cap, int64_t need)
{
    if (need <= *cap && *p) return SG_OK;
    if (*p) HIP_TRY(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    const int64_t n = std::max<int64_t>(need, 1);
    HIP_TRY(hipMalloc((void **)p, sizeof(T) * (size_t)n));
    *cap = n;
    return SG_OK;
}

// Capacity for n input particles (keys), `padded` sorted slots and ni
// i-particles.
int ensure_capacity(sg_context *c, int64_t n, int64_t padded, int64_t ni)
{
    if (n > c->cap_n || !c->pb.keys_in) {
        int64_t cap = 0;
        const int64_t need = std::max<int64_t>(n, 1);
        int rc;
        if ((rc = grow(&c->pb.keys_in, &cap, need)) != SG_OK) return rc;
        cap = 0;
        if ((rc = grow(&c->pb.keys_out, &cap, need)) != SG_OK) return rc;
        cap = 0;
        if ((rc = grow(&c->pb.vals_in, &cap, need)) != SG_OK) return rc;
        cap = 0;
        if ((rc = grow(&c->pb.vals_out, &cap, need)) != SG_OK) return rc;
        size_t bytes = 0;
        HIP_TRY(sg::sort_pairs_temp_bytes((int)need, &bytes));
        if (bytes > c->sort_tmp_bytes || !c->sort_tmp) {
            if (c->sort_tmp) HIP_TRY(hipFree(c->sort_tmp));
            c->sort_tmp = nullptr;
            HIP_TRY(hipMalloc(&c->sort_tmp, std::max<size_t>(bytes, 256)));
            c->sort_tmp_bytes = bytes;
        }
        c->cap_n = need;
    }
    if (!c->pb.red) {
        c->pb.nred_blocks = 1024;
        int64_t cap = 0;
        int rc;
        if ((rc = grow(&c->pb.red, &cap, (int64_t)c->pb.nred_blocks * 8)) != SG_OK) return rc;
        cap = 0;
        if ((rc = grow(&c->pb.bbox, &cap, 8)) != SG_OK) return rc;
        cap = 0;
        if ((rc = grow(&c->pb.err, &cap, 4)) != SG_OK) return rc;
        cap = 0;
        if ((rc = grow(&c->stats, &cap, sg::STAT_WORDS)) != SG_OK) return rc;
    }
    if (padded > c->cap_pad || !c->pb.sx) {
        int64_t cap = 0;
        int rc;
        double **arrs[5] = {&c->pb.sx, &c->pb.sy, &c->pb.sz, &c->pb.sm, &c->pb.sH};
        for (auto a : arrs) {
            cap = 0;
            if ((rc = grow(a, &cap, padded)) != SG_OK) return rc;
        }
        cap = 0;
        if ((rc = grow(&c->pb.sidx, &cap, padded)) != SG_OK) return rc;
        cap = 0;
        if ((rc = grow(&c->pb.tiles, &cap, std::max<int64_t>(padded / TJ, 1))) != SG_OK) return rc;
        cap = 0;
        if ((rc = grow(&c->pb.chunks, &cap, std::max<int64_t>(padded / sg::TC, 1))) != SG_OK)
            return rc;
        c->cap_pad = padded;
    }
    if (ni > c->cap_ni || !c->pb.ix) {
        int64_t cap = 0;
        int rc;
        const int64_t need = std::max<int64_t>(ni, 1);
        double **arrs[5] = {&c->pb.ix, &c->pb.iy, &c->pb.iz, &c->pb.im, &c->pb.iH};
        for (auto a : arrs) {
            cap = 0;
            if ((rc = grow(a, &cap, need)) != SG_OK) return rc;
        }
        cap = 0;
        if ((rc = grow(&c->pb.iidx, &cap, need)) != SG_OK) return rc;
        c->cap_ni = need;
    }
    return SG_OK;
}

bool env_flag(const char *name)
{
    const char *v = std::getenv(name);
    return v && *v && std::strcmp(v, "0") != 0;
}

}  // namespace

namespace sg {

void record_hip_error(hipError_t e) { last_hip_error() = e; }

int auto_jsplit(int niblocks, int wg, int ntiles, int64_t ni)
{
    // Split the j range until the launch holds about 2^18 waves (some 36
    // rounds of the 7168 that fit at once at 7 per SIMD): the last round of
    // a launch with few waves leaves most SIMDs idle, and small blocks also
    // spread the uneven near-field work.  Partitions keep >= 4 tiles, there
    // are at most 64 (more measured slower at 2^17 particles), and the FP64
    // partial sums (32 B per particle and partition) stay within 2^24
    // particle-partitions (512 MB).
    const int64_t waves = std::max<int64_t>(1, (int64_t)niblocks * (wg / 64));
    const int64_t want = (((int64_t)1 << 18) + waves - 1) / waves;
    const int64_t slabs = ((int64_t)1 << 24) / std::max<int64_t>(1, ni);
    const int64_t cap =
        std::min<int64_t>(64, std::min<int64_t>(std::max(1, ntiles / 4), std::max<int64_t>(1, slabs)));
    return (int)std::max<int64_t>(1, std::min(want, cap));
}

EvalPlan plan_eval(int n, int ni, int wg, int ipl, int jsplit_cfg, int precision, bool psi)
{
    EvalPlan p{};
    p.padded = ((int64_t)n + TJ - 1) / TJ * TJ;
    p.ntiles = (int)(p.padded / TJ);
    // The automatic workgroup size: 128 for an i-range of up to AUTO_SMALL_NI
    // particles, whose waves lie close to their workgroup's origin so that
    // fewer chunks need FP64-formed differences (the 16,384-particle Plummer
    // pass: 0.27 against 0.35 ms at 256), else 256 (equal at 2^17, faster
    // from 2^18 on).
    p.wg = wg > 0 ? wg : (ni <= AUTO_SMALL_NI ? 128 : 256);
    // the grad-h term is formed by the near kernel with one i per lane
    p.ipl = psi ? 1 : ipl;
    p.niblocks = (int)((ni + (int64_t)p.wg * p.ipl - 1) / ((int64_t)p.wg * p.ipl));
    p.ngroups = (int64_t)p.niblocks * (p.wg / 64);
    p.jsplit = jsplit_cfg > 0 ? std::min(jsplit_cfg, p.ntiles)
                              : auto_jsplit(p.niblocks, p.wg, p.ntiles, ni);
    p.part_elems = (int64_t)p.jsplit * 4 * ni;
    // per-(wave, tile) near flags, N/64 x N/256 bytes: only the passes with a
    // separate near kernel (FP64, grad-h) read them; the fused FP32C pass
    // keeps its near list in LDS
    const bool flagged = psi || precision != SG_PREC_FP32C;
    p.flag_elems = flagged ? p.ngroups * p.ntiles : 1;
    p.psipart_elems = psi ? (int64_t)p.jsplit * ni : 0;
    return p;
}

hipError_t enqueue_eval(const EvalPlan &p, const EvalBuffers &bufs, int n, int ibeg, int ni,
                        const double *x, const double *y, const double *z, const double *m,
                        const double *range2, int nkernel, int precision, double *ax, double *ay,
                        double *az, double *pot, double *psi, hipStream_t s,
                        const hipEvent_t *ev)
{
#define SG_ENQ(expr)                          \
    do {                                      \
        hipError_t e_ = (expr);               \
        if (e_ != hipSuccess) return e_;      \
    } while (0)
    PrepBuffers b = bufs.pb;
    if (ev) SG_ENQ(hipEventRecord(ev[0], s));
    // 1-2: validate + bbox, Morton keys of all n
    SG_ENQ(launch_validate_bbox(x, y, z, m, range2, n, b, s));
    SG_ENQ(launch_morton(x, y, z, 0, n, b, s));
    // 3: sort
    auto sort = bufs.merge_sort ? merge_sort_pairs : sort_pairs;
    SG_ENQ(sort(bufs.sort_tmp, bufs.sort_tmp_bytes, b.keys_in, b.keys_out, b.vals_in, b.vals_out,
                n, s));
    // 4-5: gather sorted j arrays, tile metadata
    SG_ENQ(launch_gather(x, y, z, m, range2, b.vals_out, n, p.padded, b.sx, b.sy, b.sz, b.sm,
                         b.sH, b.sidx, s));
    SG_ENQ(launch_tile_meta(b.sx, b.sy, b.sz, b.sH, p.ntiles, b.tiles, b.chunks, s));

    PairArgs a{};
    a.j = SortedSet{b.sx, b.sy, b.sz, b.sm, b.sH, b.sidx, n, p.padded};
    if (ibeg == 0 && ni == n) {
        a.i = a.j;
    } else {
        // the i-range gets its own curve order
        SG_ENQ(launch_morton(x, y, z, ibeg, ni, b, s));
        SG_ENQ(sort(bufs.sort_tmp, bufs.sort_tmp_bytes, b.keys_in, b.keys_out, b.vals_in,
                    b.vals_out, ni, s));
        SG_ENQ(launch_gather(x, y, z, m, range2, b.vals_out, ni, ni, b.ix, b.iy, b.iz, b.im,
                             b.iH, b.iidx, s));
        a.i = SortedSet{b.ix, b.iy, b.iz, b.im, b.iH, b.iidx, ni, ni};
    }
    a.tiles = b.tiles;
    a.chunks = b.chunks;
    a.ni = ni;
    a.ibeg = ibeg;
    a.nkernel = nkernel;
    a.ntiles = p.ntiles;
    a.ipl = p.ipl;
    a.niblocks = p.niblocks;
    a.ngroups = p.ngroups;
    a.jsplit = p.jsplit;
    a.ax = ax;
    a.ay = ay;
    a.az = az;
    a.pot = pot;
    a.partial = bufs.partial;
    a.nearflag = bufs.nearflag;
    a.stats = bufs.stats;
    if (psi) {
        a.psi = psi;
        a.psipart = bufs.psipart;
    }
    SG_ENQ(launch_fill_zero_int((int *)bufs.stats, 2 * STAT_WORDS, s));

    if (ev) SG_ENQ(hipEventRecord(ev[1], s));
    SG_ENQ(launch_pair_far(precision, a, p.wg, s));
    if (ev) SG_ENQ(hipEventRecord(ev[2], s));
    SG_ENQ(launch_pair_near(precision, a, s));
    SG_ENQ(launch_reduce_partials(a, s));
    if (ev) SG_ENQ(hipEventRecord(ev[3], s));
#undef SG_ENQ
    return hipSuccess;
}

}  // namespace sg

// ============================================================ sg_* API

extern "C" {

const char *sg_status_string(int status)
{
    switch (status) {
    case SG_OK: return "ok";
    case SG_ERR_ARG: return "invalid argument (size, range or null pointer)";
    case SG_ERR_NKERNEL: return "nkernel must be 0 (cubic), 1 (Wendland W3,3) or 2 (Wendland C4)";
    case SG_ERR_MASS: return "particle mass <= 0";
    case SG_ERR_RANGE2: return "range2 == 0 (zero smoothing length)";
    case SG_ERR_NONFINITE: return "NaN or Inf in positions, masses or range2";
    case SG_ERR_HIP: return hipGetErrorString(last_hip_error());
    case SG_ERR_NO_DEVICE: return "no usable HIP device";
    case SG_ERR_STATE: return "lasthalf called without a pending firsthalf";
    case SG_ERR_CONFIG: return "unsupported tile configuration";
    case SG_ERR_SIZE: return "n exceeds SG_STATIC_MAX_N, the limit of sg_compute_device_static";
    default: return "unknown status";
    }
}

int sg_last_status(void) { return g_last_status; }

int sg_device_count(int *count)
{
    if (!count) return SG_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return n > 0 ? SG_OK : SG_ERR_NO_DEVICE;
}

int sg_create(int device, sg_context **out)
{
    if (!out) return SG_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SG_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return SG_ERR_NO_DEVICE;
    HIP_TRY(hipSetDevice(device));
    sg_context *c = new sg_context();
    c->device = device;
    // The events that mark the start and the phase boundaries of a pass only
    // time it: they skip the system-scope fence (an L2 writeback and
    // invalidate between the kernels of a pass).  The closing event keeps it.
    const unsigned timing_only = hipEventDisableSystemFence;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_p0, timing_only) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_k0, timing_only) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_kf, timing_only) != hipSuccess ||
        hipEventCreate(&c->ev_k1) != hipSuccess) {
        delete c;
        return SG_ERR_HIP;
    }
    c->stream = c->own_stream;
    *out = c;
    return SG_OK;
}

int sg_set_stream(sg_context *c, void *stream)
{
    if (!c) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    // work already queued on the previous stream completes first
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    return SG_OK;
}

int sg_destroy(sg_context *c)
{
    if (!c) return SG_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->pb.red, c->pb.bbox, c->pb.err, c->pb.keys_in, c->pb.keys_out,
                    c->pb.vals_in, c->pb.vals_out, c->pb.sx, c->pb.sy, c->pb.sz, c->pb.sm,
                    c->pb.sH, c->pb.sidx, c->pb.tiles, c->pb.chunks, c->pb.ix, c->pb.iy, c->pb.iz,
                    c->pb.im, c->pb.iH, c->pb.iidx, c->sort_tmp, c->partial, c->nearflag, c->stats,
                    c->in[0], c->in[1], c->in[2], c->in[3], c->in[4],
                    c->out[0], c->out[1], c->out[2], c->out[3],
                    c->d_offsets, c->d_sets, c->d_items, c->d_counter,
                    c->scot[0], c->scot[1], c->scot[2], c->scot[3],
                    c->vin[0], c->vin[1], c->vin[2], c->vin[3],
                    c->vout[0], c->vout[1], c->vout[2], c->vout[3], c->vout[4], c->psipart,
                    c->psi_out, c->epot_work};
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    (void)hipEventDestroy(c->ev_p0);
    (void)hipEventDestroy(c->ev_k0);
    (void)hipEventDestroy(c->ev_kf);
    (void)hipEventDestroy(c->ev_k1);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return SG_OK;
}

int sg_set_config(sg_context *c, int precision, int wg, int ipl, int jsplit)
{
    if (!c || (precision != SG_PREC_FP32C && precision != SG_PREC_FP64) || jsplit < 0)
        return SG_ERR_ARG;
    // wg 0: chosen per evaluation from the size of the i-range
    if (!sg::pair_config_supported(precision, wg == 0 ? 256 : wg, ipl)) return SG_ERR_CONFIG;
    c->wg[precision] = wg;
    c->ipl[precision] = ipl;
    c->jsplit[precision] = jsplit;
    return SG_OK;
}

int sg_last_config(sg_context *c, int *wg, int *ipl, int *jsplit)
{
    if (!c) return SG_ERR_ARG;
    if (!c->launched || c->last_shape[0] == 0) return SG_ERR_STATE;
    if (wg) *wg = c->last_shape[0];
    if (ipl) *ipl = c->last_shape[1];
    if (jsplit) *jsplit = c->last_shape[2];
    return SG_OK;
}

int sg_get_config(sg_context *c, int precision, int *wg, int *ipl, int *jsplit)
{
    if (!c || (precision != SG_PREC_FP32C && precision != SG_PREC_FP64)) return SG_ERR_ARG;
    if (wg) *wg = c->wg[precision];
    if (ipl) *ipl = c->ipl[precision];
    if (jsplit) *jsplit = c->jsplit[precision];
    return SG_OK;
}

namespace {

// One evaluation; psi (may be null) also forms the grad-h softening term.
int compute_device_impl(sg_context *c, int n, int ibeg, int ni, const double *x,
                        const double *y, const double *z, const double *m,
                        const double *range2, int nkernel, int precision, double *ax,
                        double *ay, double *az, double *pot, double *psi)
{
    if (!c) return SG_ERR_ARG;
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (precision != SG_PREC_FP32C && precision != SG_PREC_FP64) return SG_ERR_ARG;
    if (n < 1 || ibeg < 0 || ni < 0 || (int64_t)ibeg + ni > n) return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2) return SG_ERR_ARG;
    if (ni > 0 && (!ax || !ay || !az || !pot)) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    c->launched = false;
    c->last_shape[0] = 0;   // set again by a single-set pair pass
    if (ni == 0) return SG_OK;

    const sg::EvalPlan p = sg::plan_eval(n, ni, c->wg[precision], c->ipl[precision],
                                         c->jsplit[precision], precision, psi != nullptr);
    if (psi && p.psipart_elems > INT32_MAX / 2) return SG_ERR_ARG;
    int rc = ensure_capacity(c, n, p.padded, ni);
    if (rc != SG_OK) return rc;
    if ((rc = grow(&c->partial, &c->cap_part, p.part_elems)) != SG_OK) return rc;
    if ((rc = grow(&c->nearflag, &c->cap_flag, p.flag_elems)) != SG_OK) return rc;
    if (psi && (rc = grow(&c->psipart, &c->cap_psipart, p.psipart_elems)) != SG_OK) return rc;
    const sg::EvalBuffers bufs{c->pb,       c->sort_tmp, c->sort_tmp_bytes, c->partial,
                               c->nearflag, c->stats,    c->psipart,        false};
    const hipEvent_t ev[4] = {c->ev_p0, c->ev_k0, c->ev_kf, c->ev_k1};
    HIP_TRY(sg::enqueue_eval(p, bufs, n, ibeg, ni, x, y, z, m, range2, nkernel, precision, ax, ay,
                             az, pot, psi, c->stream, ev));
    c->last_shape[0] = p.wg;
    c->last_shape[1] = p.ipl;
    c->last_shape[2] = p.jsplit;
    c->launched = true;
    return SG_OK;
}

}  // namespace

int sg_compute_device(sg_context *c, int n, int ibeg, int ni, const double *x, const double *y,
                      const double *z, const double *m, const double *range2, int nkernel,
                      int precision, double *ax, double *ay, double *az, double *pot)
{
    return compute_device_impl(c, n, ibeg, ni, x, y, z, m, range2, nkernel, precision, ax, ay, az,
                               pot, nullptr);
}

int sg_psi_device(sg_context *c, int n, int ibeg, int ni, const double *x, const double *y,
                  const double *z, const double *m, const double *range2, int nkernel,
                  double *psi)
{
    if (!c) return SG_ERR_ARG;
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (n < 1 || ibeg < 0 || ni < 0 || (int64_t)ibeg + ni > n) return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2) return SG_ERR_ARG;
    if (ni > 0 && !psi) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    c->launched = false;
    c->last_shape[0] = 0;   // set again by a single-set pair pass
    if (ni == 0) return SG_OK;

    const int64_t padded = ((int64_t)(n) + TJ - 1) / TJ * TJ;
    int rc = ensure_capacity(c, n, padded, ni);
    if (rc != SG_OK) return rc;
    hipStream_t s = c->stream;
    sg::PrepBuffers &b = c->pb;
    HIP_TRY(hipEventRecord(c->ev_p0, s));
    HIP_TRY(sg::launch_validate_bbox(x, y, z, m, range2, n, b, s));
    HIP_TRY(sg::launch_morton(x, y, z, 0, n, b, s));
    HIP_TRY(sg::sort_pairs(c->sort_tmp, c->sort_tmp_bytes, b.keys_in, b.keys_out, b.vals_in,
                           b.vals_out, n, s));
    HIP_TRY(sg::launch_gather(x, y, z, m, range2, b.vals_out, n, padded, b.sx, b.sy, b.sz, b.sm,
                              b.sH, b.sidx, s));
    const int ntiles = (int)(padded / TJ);
    HIP_TRY(sg::launch_tile_meta(b.sx, b.sy, b.sz, b.sH, ntiles, b.tiles, b.chunks, s));
    sg::PairArgs a{};
    a.j = sg::SortedSet{b.sx, b.sy, b.sz, b.sm, b.sH, b.sidx, n, padded};
    if (ibeg == 0 && ni == n) {
        a.i = a.j;
    } else {
        HIP_TRY(sg::launch_morton(x, y, z, ibeg, ni, b, s));
        HIP_TRY(sg::sort_pairs(c->sort_tmp, c->sort_tmp_bytes, b.keys_in, b.keys_out, b.vals_in,
                               b.vals_out, ni, s));
        HIP_TRY(sg::launch_gather(x, y, z, m, range2, b.vals_out, ni, ni, b.ix, b.iy, b.iz, b.im,
                                  b.iH, b.iidx, s));
        a.i = sg::SortedSet{b.ix, b.iy, b.iz, b.im, b.iH, b.iidx, ni, ni};
    }
    a.tiles = b.tiles;
    a.chunks = b.chunks;
    a.ni = ni;
    a.ibeg = ibeg;
    a.nkernel = nkernel;
    a.ntiles = ntiles;
    a.ipl = 1;
    a.jsplit = 1;
    a.ngroups = (ni + 63) / 64;
    a.psi = psi;
    // the psi-only pass classifies no chunks: its statistics read zero
    // rather than the previous evaluation's counts
    HIP_TRY(sg::launch_fill_zero_int((int *)c->stats, 2 * sg::STAT_WORDS, s));
    HIP_TRY(hipEventRecord(c->ev_k0, s));
    HIP_TRY(hipEventRecord(c->ev_kf, s));
    HIP_TRY(sg::launch_psi_only(a, s));
    HIP_TRY(hipEventRecord(c->ev_k1, s));
    c->launched = true;
    return SG_OK;
}

int sg_psi_host(sg_context *c, int n, int ibeg, int ni, const double *x, const double *y,
                const double *z, const double *m, const double *range2, int nkernel, double *psi)
{
    if (!c) return SG_ERR_ARG;
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (n < 1 || ibeg < 0 || ni < 0 || (int64_t)ibeg + ni > n) return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2 || (ni > 0 && !psi)) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    int rc;
    for (int k = 0; k < 5; ++k) {
        int64_t cap = c->cap_in;
        if ((rc = grow(&c->in[k], &cap, n)) != SG_OK) return rc;
    }
    c->cap_in = std::max<int64_t>(c->cap_in, n);
    if ((rc = grow(&c->psi_out, &c->cap_psi_out, std::max(ni, 1))) != SG_OK) return rc;
    const double *src[5] = {x, y, z, m, range2};
    for (int k = 0; k < 5; ++k)
        HIP_TRY(hipMemcpyAsync(c->in[k], src[k], sizeof(double) * (size_t)n, hipMemcpyHostToDevice,
                               c->stream));
    rc = sg_psi_device(c, n, ibeg, ni, c->in[0], c->in[1], c->in[2], c->in[3], c->in[4], nkernel,
                       c->psi_out);
    if (rc != SG_OK) return rc;
    if ((rc = sg_synchronize(c)) != SG_OK) return rc;
    if (ni > 0)
        HIP_TRY(hipMemcpy(psi, c->psi_out, sizeof(double) * (size_t)ni, hipMemcpyDeviceToHost));
    return SG_OK;
}

int sg_compute_psi_device(sg_context *c, int n, int ibeg, int ni, const double *x,
                          const double *y, const double *z, const double *m,
                          const double *range2, int nkernel, int precision, double *ax,
                          double *ay, double *az, double *pot, double *psi)
{
    if (ni > 0 && !psi) return SG_ERR_ARG;
    return compute_device_impl(c, n, ibeg, ni, x, y, z, m, range2, nkernel, precision, ax, ay, az,
                               pot, psi);
}

// ----------------------------------------------------- batch of sets
int sg_set_batch_config(sg_context *c, int precision, int wg, int ipl)
{
    if (!c || (precision != SG_PREC_FP32C && precision != SG_PREC_FP64)) return SG_ERR_ARG;
    if (!sg::batch_config_supported(wg, ipl)) return SG_ERR_CONFIG;
    c->bwg[precision] = wg;
    c->bipl[precision] = ipl;
    return SG_OK;
}

int sg_get_batch_config(sg_context *c, int precision, int *wg, int *ipl)
{
    if (!c || (precision != SG_PREC_FP32C && precision != SG_PREC_FP64)) return SG_ERR_ARG;
    if (wg) *wg = c->bwg[precision];
    if (ipl) *ipl = c->bipl[precision];
    return SG_OK;
}

namespace {

// Host part of a batch plan: per-set layout, near-flag and partial-sum
// offsets and the work items, longest first.  An item is one i-block of a
// set against one j partition; sets of more than BATCH_JTILES tiles are cut
// into ceil(ntiles / BATCH_JTILES) partitions, so that the longest items do
// not outlast the rest of the queue.  Where one large set dominates a batch
// this is what keeps the launch short (a 262,144-particle star with 31 small
// ones: 38.0 ms with whole-set items, 22.9 ms with partitions); on the
// 64-star batch of N in [2,048, 131,072] it gains 0.5%.  As in a single-set
// evaluation there are at most 64 partitions and at most 2^24
// slot-partitions of partial sums (512 MB).
int build_plan(sg_context *c, int nsets, const long long *offsets, int wg, int ipl)
{
    auto &P = c->plan;
    P.offsets.assign(offsets, offsets + nsets + 1);
    P.wg = wg;
    P.ipl = ipl;
    std::vector<sg::BatchSet> sets(nsets);
    std::vector<int2> items;
    const int64_t gi = (int64_t)wg * ipl;
    int64_t slots = 0;
    for (int s = 0; s < nsets; ++s) slots += (offsets[s + 1] - offsets[s] + TJ - 1) / TJ * TJ;
    const int64_t jcap = std::max<int64_t>(1, std::min<int64_t>(64, ((int64_t)1 << 24) / slots));
    int64_t pofs = 0, gofs = 0, fofs = 0, partofs = 0;
    for (int s = 0; s < nsets; ++s) {
        const int64_t n = offsets[s + 1] - offsets[s];
        const int64_t ntl = (n + TJ - 1) / TJ;
        const int64_t nblk = (n + gi - 1) / gi;
        const int js = (int)std::min<int64_t>(jcap, (ntl + sg::BATCH_JTILES - 1) / sg::BATCH_JTILES);
        sets[s] = sg::BatchSet{pofs, n, (int)(pofs / TJ), (int)ntl, gofs, fofs, partofs, js};
        const int64_t groups = nblk * (wg / 64);   // every wave of every i-block
        for (int64_t b = 0; b < nblk; ++b)
            for (int p = 0; p < js; ++p) items.push_back(make_int2(s, (int)(b * js + p)));
        pofs += ntl * TJ;
        gofs += groups;
        fofs += groups * ntl;
        partofs += (int64_t)js * 4 * ntl * TJ;   // >= 4 * pofs of the next set
    }
    // tiles an item visits: partition p of a set holds ceil((ntiles - p) / js)
    auto cost = [&](const int2 &u) {
        const sg::BatchSet &st = sets[u.x];
        return (st.ntiles - u.y % st.jsplit + st.jsplit - 1) / st.jsplit;
    };
    std::stable_sort(items.begin(), items.end(),
                     [&](const int2 &u, const int2 &v) { return cost(u) > cost(v); });
    P.padded = pofs;
    P.ngroups = gofs;
    P.nflags = fofs;
    P.nparts = partofs;
    P.nitems = (int)items.size();
    int rc;
    if ((rc = grow(&c->d_offsets, &c->cap_offsets, nsets + 1)) != SG_OK) return rc;
    if ((rc = grow(&c->d_sets, &c->cap_sets, nsets)) != SG_OK) return rc;
    if ((rc = grow(&c->d_items, &c->cap_items, (int64_t)items.size())) != SG_OK) return rc;
    if (!c->d_counter) HIP_TRY(hipMalloc((void **)&c->d_counter, 64));
    HIP_TRY(hipMemcpy(c->d_offsets, P.offsets.data(), sizeof(int64_t) * (nsets + 1),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_sets, sets.data(), sizeof(sg::BatchSet) * nsets, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_items, items.data(), sizeof(int2) * items.size(),
                      hipMemcpyHostToDevice));
    return SG_OK;
}

}  // namespace

int sg_compute_batch_device(sg_context *c, int nsets, const long long *offsets, const double *x,
                            const double *y, const double *z, const double *m,
                            const double *range2, int nkernel, int precision, int use_graph,
                            double *ax, double *ay, double *az, double *pot)
{
    if (!c || nsets < 1 || !offsets) return SG_ERR_ARG;
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (precision != SG_PREC_FP32C && precision != SG_PREC_FP64) return SG_ERR_ARG;
    if (offsets[0] != 0 || nsets > (1 << 20)) return SG_ERR_ARG;
    for (int s = 0; s < nsets; ++s)
        if (offsets[s + 1] <= offsets[s]) return SG_ERR_ARG;   // every set holds >= 1 particle
    const int64_t ntot = offsets[nsets];
    if (ntot > 0x7fffffff) return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2 || !ax || !ay || !az || !pot) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    c->launched = false;
    c->last_shape[0] = 0;   // set again by a single-set pair pass
    const int wg = c->bwg[precision], ipl = c->bipl[precision];

    auto &P = c->plan;
    const bool same_plan = P.wg == wg && P.ipl == ipl && (int)P.offsets.size() == nsets + 1 &&
                           std::equal(P.offsets.begin(), P.offsets.end(), offsets);
    int rc;
    if (!same_plan) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if ((rc = build_plan(c, nsets, offsets, wg, ipl)) != SG_OK) return rc;
    }
    if ((rc = ensure_capacity(c, ntot, P.padded, 1)) != SG_OK) return rc;
    if ((rc = grow(&c->partial, &c->cap_part, P.nparts)) != SG_OK) return rc;
    if ((rc = grow(&c->nearflag, &c->cap_flag, std::max<int64_t>(P.nflags, 1))) != SG_OK) return rc;

    hipStream_t s = c->stream;
    sg::PrepBuffers &b = c->pb;
    HIP_TRY(hipEventRecord(c->ev_p0, s));
    // preparation (not captured: the radix sort clears its own scratch with
    // memset nodes)
    HIP_TRY(sg::launch_validate_bbox(x, y, z, m, range2, (int)ntot, b, s));
    HIP_TRY(sg::launch_morton_batch(x, y, z, (int)ntot, c->d_offsets, nsets, b, s));
    HIP_TRY(sg::sort_pairs(c->sort_tmp, c->sort_tmp_bytes, b.keys_in, b.keys_out, b.vals_in,
                           b.vals_out, (int)ntot, s));
    HIP_TRY(sg::launch_gather_batch(x, y, z, m, range2, b.vals_out, c->d_offsets, c->d_sets, nsets,
                                    P.padded, b.sx, b.sy, b.sz, b.sm, b.sH, b.sidx, s));
    HIP_TRY(sg::launch_tile_meta(b.sx, b.sy, b.sz, b.sH, (int)(P.padded / TJ), b.tiles, b.chunks,
                                 s));

    sg::BatchArgs a{};
    a.j = sg::SortedSet{b.sx, b.sy, b.sz, b.sm, b.sH, b.sidx, ntot, P.padded};
    a.tiles = b.tiles;
    a.chunks = b.chunks;
    a.sets = c->d_sets;
    a.nsets = nsets;
    a.items = c->d_items;
    a.nitems = P.nitems;
    a.ngroups = P.ngroups;
    a.padded = P.padded;
    a.nkernel = nkernel;
    a.partial = c->partial;
    a.nearflag = c->nearflag;
    a.counter = c->d_counter;
    a.ax = ax;
    a.ay = ay;
    a.az = az;
    a.pot = pot;
    a.stats = c->stats;
    const int nwg = sg::batch_resident_workgroups(precision, nkernel, wg, ipl, c->device);
    if (nwg <= 0) return SG_ERR_CONFIG;

    HIP_TRY(hipEventRecord(c->ev_k0, s));
    if (use_graph) {
        const int key[5] = {precision, nkernel, wg, ipl, nwg};
        const bool reuse = c->gexec && c->g_stream == s &&
                           std::memcmp(&c->g_args, &a, sizeof(a)) == 0 &&
                           std::memcmp(c->g_key, key, sizeof(key)) == 0;
        if (!reuse) {
            if (c->gexec) HIP_TRY(hipGraphExecDestroy(c->gexec));
            c->gexec = nullptr;
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            HIP_TRY(sg::launch_fill_zero_int((int *)c->stats, 2 * sg::STAT_WORDS, s));
            const hipError_t le = sg::launch_batch(precision, a, wg, ipl, nwg, s);
            const hipError_t ee = hipStreamEndCapture(s, &graph);
            HIP_TRY(le);
            HIP_TRY(ee);
            const hipError_t ie = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_TRY(ie);
            c->g_args = a;
            std::memcpy(c->g_key, key, sizeof(key));
            c->g_stream = s;
        }
        HIP_TRY(hipGraphLaunch(c->gexec, s));
    } else {
        HIP_TRY(sg::launch_fill_zero_int((int *)c->stats, 2 * sg::STAT_WORDS, s));
        HIP_TRY(sg::launch_batch(precision, a, wg, ipl, nwg, s));
    }
    HIP_TRY(hipEventRecord(c->ev_kf, s));
    HIP_TRY(hipEventRecord(c->ev_k1, s));
    c->launched = true;
    return SG_OK;
}

int sg_synchronize(sg_context *c)
{
    if (!c) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!c->launched) return SG_OK;
    int err = 0;
    HIP_TRY(hipMemcpy(&err, c->pb.err, sizeof(int), hipMemcpyDeviceToHost));
    if (err & sg::ERR_NONFINITE) return SG_ERR_NONFINITE;
    if (err & sg::ERR_MASS) return SG_ERR_MASS;
    if (err & sg::ERR_RANGE2) return SG_ERR_RANGE2;
    return SG_OK;
}

int sg_potential_energy(sg_context *c, int n, const double *m, const double *pot,
                        double *epot)
{
    if (!c || n < 1 || !m || !pot || !epot) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    constexpr int NWORK = 1025;   // up to 1024 block sums, then the result
    int rc;
    if ((rc = grow(&c->epot_work, &c->cap_epot, NWORK)) != SG_OK) return rc;
    HIP_TRY(sg::launch_potential_energy(m, pot, n, c->epot_work, NWORK, c->stream));
    HIP_TRY(hipMemcpyAsync(epot, c->epot_work + NWORK - 1, sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_last_kernel_ms(sg_context *c, float *ms)
{
    if (!c || !ms) return SG_ERR_ARG;
    if (!c->launched) return SG_ERR_STATE;
    HIP_TRY(hipEventElapsedTime(ms, c->ev_k0, c->ev_k1));
    return SG_OK;
}

int sg_last_timings(sg_context *c, float *prep_ms, float *far_ms, float *near_ms)
{
    if (!c) return SG_ERR_ARG;
    if (!c->launched) return SG_ERR_STATE;
    float t;
    if (prep_ms) {
        HIP_TRY(hipEventElapsedTime(&t, c->ev_p0, c->ev_k0));
        *prep_ms = t;
    }
    if (far_ms) {
        HIP_TRY(hipEventElapsedTime(&t, c->ev_k0, c->ev_kf));
        *far_ms = t;
    }
    if (near_ms) {
        HIP_TRY(hipEventElapsedTime(&t, c->ev_kf, c->ev_k1));
        *near_ms = t;
    }
    return SG_OK;
}

int sg_last_tile_stats(sg_context *c, long long *visits, long long *near_visits)
{
    return sg_last_precision_stats(c, visits, near_visits, nullptr, nullptr);
}

int sg_last_precision_stats(sg_context *c, long long *visits, long long *near_visits,
                            long long *wide_visits, long long *fp64_pairs)
{
    if (!c) return SG_ERR_ARG;
    if (!c->launched) return SG_ERR_STATE;
    std::vector<unsigned long long> copies(sg::STAT_WORDS);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(copies.data(), c->stats, sizeof(unsigned long long) * copies.size(),
                      hipMemcpyDeviceToHost));
    unsigned long long h[sg::STAT_N] = {0, 0, 0, 0};
    for (int k = 0; k < sg::STAT_SLOTS; ++k)
        for (int q = 0; q < sg::STAT_N; ++q) h[q] += copies[(size_t)k * sg::STAT_STRIDE + q];
    if (visits) *visits = (long long)h[0];
    if (near_visits) *near_visits = (long long)h[1];
    if (wide_visits) *wide_visits = (long long)h[2];
    if (fp64_pairs) *fp64_pairs = (long long)h[3];
    return SG_OK;
}

namespace {

int compute_host_impl(sg_context *c, int n, int ibeg, int ni, const double *x, const double *y,
                      const double *z, const double *m, const double *range2, int nkernel,
                      int precision, double *ax, double *ay, double *az, double *pot,
                      double *psi)
{
    if (!c) return SG_ERR_ARG;
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (n < 1 || ibeg < 0 || ni < 0 || (int64_t)ibeg + ni > n) return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2) return SG_ERR_ARG;
    if (ni > 0 && (!ax || !ay || !az || !pot)) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    int rc;
    for (int k = 0; k < 5; ++k) {
        int64_t cap = c->cap_in;
        if ((rc = grow(&c->in[k], &cap, n)) != SG_OK) return rc;
    }
    c->cap_in = std::max<int64_t>(c->cap_in, n);
    for (int k = 0; k < 4; ++k) {
        int64_t cap = c->cap_out;
        if ((rc = grow(&c->out[k], &cap, std::max(ni, 1))) != SG_OK) return rc;
    }
    if (psi && (rc = grow(&c->psi_out, &c->cap_psi_out, std::max(ni, 1))) != SG_OK) return rc;
    c->cap_out = std::max<int64_t>(c->cap_out, std::max(ni, 1));
    const double *src[5] = {x, y, z, m, range2};
    for (int k = 0; k < 5; ++k)
        HIP_TRY(hipMemcpyAsync(c->in[k], src[k], sizeof(double) * (size_t)n, hipMemcpyHostToDevice,
                               c->stream));
    rc = compute_device_impl(c, n, ibeg, ni, c->in[0], c->in[1], c->in[2], c->in[3], c->in[4],
                             nkernel, precision, c->out[0], c->out[1], c->out[2], c->out[3],
                             psi ? c->psi_out : nullptr);
    if (rc != SG_OK) return rc;
    rc = sg_synchronize(c);
    if (rc != SG_OK) return rc;
    double *dst[4] = {ax, ay, az, pot};
    for (int k = 0; k < 4 && ni > 0; ++k)
        HIP_TRY(hipMemcpy(dst[k], c->out[k], sizeof(double) * (size_t)ni, hipMemcpyDeviceToHost));
    if (psi && ni > 0)
        HIP_TRY(hipMemcpy(psi, c->psi_out, sizeof(double) * (size_t)ni, hipMemcpyDeviceToHost));
    return SG_OK;
}

}  // namespace

int sg_compute_host(sg_context *c, int n, int ibeg, int ni, const double *x, const double *y,
                    const double *z, const double *m, const double *range2, int nkernel,
                    int precision, double *ax, double *ay, double *az, double *pot)
{
    return compute_host_impl(c, n, ibeg, ni, x, y, z, m, range2, nkernel, precision, ax, ay, az,
                             pot, nullptr);
}

int sg_compute_psi_host(sg_context *c, int n, int ibeg, int ni, const double *x, const double *y,
                        const double *z, const double *m, const double *range2, int nkernel,
                        int precision, double *ax, double *ay, double *az, double *pot,
                        double *psi)
{
    if (ni > 0 && !psi) return SG_ERR_ARG;
    return compute_host_impl(c, n, ibeg, ni, x, y, z, m, range2, nkernel, precision, ax, ay, az,
                             pot, psi);
}



// ------------------------------------------------------------ adjoint
int sg_vjp_device(sg_context *c, int n, const double *x, const double *y, const double *z,
                  const double *m, const double *range2, int nkernel, int precision,
                  const double *wx, const double *wy, const double *wz, const double *u,
                  double *gx, double *gy, double *gz, double *gm, double *gH)
{
    if (!c) return SG_ERR_ARG;
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (precision != SG_PREC_FP32C && precision != SG_PREC_FP64) return SG_ERR_ARG;
    if (n < 1) return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2 || !wx || !wy || !wz || !u) return SG_ERR_ARG;
    if (!gx || !gy || !gz || !gm || !gH) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    c->launched = false;
    c->last_shape[0] = 0;   // set again by a single-set pair pass

    const int64_t padded = ((int64_t)(n) + TJ - 1) / TJ * TJ;
    int rc = ensure_capacity(c, n, padded, 1);
    if (rc != SG_OK) return rc;
    for (int k = 0; k < 4; ++k) {
        int64_t cap = c->cap_cot;
        if ((rc = grow(&c->scot[k], &cap, padded)) != SG_OK) return rc;
    }
    c->cap_cot = std::max<int64_t>(c->cap_cot, padded);
    hipStream_t s = c->stream;
    sg::PrepBuffers &b = c->pb;
    HIP_TRY(hipEventRecord(c->ev_p0, s));
    HIP_TRY(sg::launch_validate_bbox(x, y, z, m, range2, n, b, s));
    HIP_TRY(sg::launch_morton(x, y, z, 0, n, b, s));
    HIP_TRY(sg::sort_pairs(c->sort_tmp, c->sort_tmp_bytes, b.keys_in, b.keys_out, b.vals_in,
                           b.vals_out, n, s));
    HIP_TRY(sg::launch_gather(x, y, z, m, range2, b.vals_out, n, padded, b.sx, b.sy, b.sz, b.sm,
                              b.sH, b.sidx, s));
    const int ntiles = (int)(padded / TJ);
    HIP_TRY(sg::launch_tile_meta(b.sx, b.sy, b.sz, b.sH, ntiles, b.tiles, b.chunks, s));
    HIP_TRY(sg::launch_gather_cot(wx, wy, wz, u, b.vals_out, n, padded, c->scot[0], c->scot[1],
                                  c->scot[2], c->scot[3], b, s));

    sg::AdjArgs a{};
    a.p = sg::SortedSet{b.sx, b.sy, b.sz, b.sm, b.sH, b.sidx, n, padded};
    a.wx = c->scot[0];
    a.wy = c->scot[1];
    a.wz = c->scot[2];
    a.u = c->scot[3];
    a.tiles = b.tiles;
    a.chunks = b.chunks;
    a.n = n;
    a.nkernel = nkernel;
    a.ntiles = ntiles;
    const int wg = sg::adj_config_supported(c->wg[precision]) ? c->wg[precision] : 256;
    a.niblocks = (int)((n + (int64_t)wg - 1) / wg);
    a.ngroups = (int64_t)a.niblocks * (wg / 64);
    a.jsplit = sg::auto_jsplit(a.niblocks, wg, ntiles, n);
    a.gx = gx;
    a.gy = gy;
    a.gz = gz;
    a.gm = gm;
    a.gH = gH;
    if ((rc = grow(&c->partial, &c->cap_part, (int64_t)a.jsplit * 5 * n)) != SG_OK) return rc;
    a.partial = c->partial;
    if ((rc = grow(&c->nearflag, &c->cap_flag, a.ngroups * ntiles)) != SG_OK) return rc;
    a.nearflag = c->nearflag;
    a.stats = c->stats;
    HIP_TRY(sg::launch_fill_zero_int((int *)c->stats, 2 * sg::STAT_WORDS, s));

    HIP_TRY(hipEventRecord(c->ev_k0, s));
    HIP_TRY(sg::launch_adj_far(precision, a, wg, s));
    HIP_TRY(hipEventRecord(c->ev_kf, s));
    HIP_TRY(sg::launch_adj_near(precision, a, s));
    HIP_TRY(sg::launch_adj_reduce(a, s));
    HIP_TRY(hipEventRecord(c->ev_k1, s));
    c->launched = true;
    return SG_OK;
}

int sg_vjp_host(sg_context *c, int n, const double *x, const double *y, const double *z,
                const double *m, const double *range2, int nkernel, int precision,
                const double *wx, const double *wy, const double *wz, const double *u,
                double *gx, double *gy, double *gz, double *gm, double *gH)
{
    if (!c) return SG_ERR_ARG;
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (n < 1) return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2 || !wx || !wy || !wz || !u) return SG_ERR_ARG;
    if (!gx || !gy || !gz || !gm || !gH) return SG_ERR_ARG;
    HIP_TRY(hipSetDevice(c->device));
    int rc;
    for (int k = 0; k < 5; ++k) {
        int64_t cap = c->cap_in;
        if ((rc = grow(&c->in[k], &cap, n)) != SG_OK) return rc;
    }
    c->cap_in = std::max<int64_t>(c->cap_in, n);
    for (int k = 0; k < 4; ++k) {
        int64_t cap = c->cap_vin;
        if ((rc = grow(&c->vin[k], &cap, n)) != SG_OK) return rc;
    }
    c->cap_vin = std::max<int64_t>(c->cap_vin, n);
    for (int k = 0; k < 5; ++k) {
        int64_t cap = c->cap_vout;
        if ((rc = grow(&c->vout[k], &cap, n)) != SG_OK) return rc;
    }
    c->cap_vout = std::max<int64_t>(c->cap_vout, n);
    const double *src[5] = {x, y, z, m, range2};
    for (int k = 0; k < 5; ++k)
        HIP_TRY(hipMemcpyAsync(c->in[k], src[k], sizeof(double) * (size_t)n, hipMemcpyHostToDevice,
                               c->stream));
    const double *csrc[4] = {wx, wy, wz, u};
    for (int k = 0; k < 4; ++k)
        HIP_TRY(hipMemcpyAsync(c->vin[k], csrc[k], sizeof(double) * (size_t)n,
                               hipMemcpyHostToDevice, c->stream));
    rc = sg_vjp_device(c, n, c->in[0], c->in[1], c->in[2], c->in[3], c->in[4], nkernel, precision,
                       c->vin[0], c->vin[1], c->vin[2], c->vin[3], c->vout[0], c->vout[1],
                       c->vout[2], c->vout[3], c->vout[4]);
    if (rc != SG_OK) return rc;
    rc = sg_synchronize(c);
    if (rc != SG_OK) return rc;
    double *dst[5] = {gx, gy, gz, gm, gH};
    for (int k = 0; k < 5; ++k)
        HIP_TRY(hipMemcpy(dst[k], c->vout[k], sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    return SG_OK;
}

}  // extern "C"

// ===================================================== legacy layer state
namespace {

struct Legacy {
    int device = -1;
    std::vector<sg_context *> ctx;   // one per device, created on first use
    bool pending = false;
    int ibeg = 0, ni = 0;
    sg_context *cur = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
};

Legacy &legacy()
{
    static Legacy L;
    return L;
}

int legacy_precision()
{
    const char *v = std::getenv("SPHGRAV_PRECISION");
    return (v && (std::strcmp(v, "fp64") == 0 || std::strcmp(v, "FP64") == 0)) ? SG_PREC_FP64
                                                                               : SG_PREC_FP32C;
}

int legacy_context(sg_context **out)
{
    Legacy &L = legacy();
    int dev = L.device;
    if (dev < 0) {
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SG_ERR_NO_DEVICE;
    if (dev >= count) return SG_ERR_NO_DEVICE;
    if ((int)L.ctx.size() < count) L.ctx.resize(count, nullptr);
    if (!L.ctx[dev]) {
        int rc = sg_create(dev, &L.ctx[dev]);
        if (rc != SG_OK) return rc;
    }
    if (!L.t0) {
        HIP_TRY(hipSetDevice(dev));
        HIP_TRY(hipEventCreate(&L.t0));
        HIP_TRY(hipEventCreate(&L.t1));
    }
    *out = L.ctx[dev];
    return SG_OK;
}

// Host-side argument checks of the reference's packing loop
// (grav_force_direct.cu:597-602), turned into status codes.
int check_host_inputs(int n, const double *x, const double *y, const double *z, const double *m,
                      const double *r2)
{
    for (int k = 0; k < n; ++k) {
        if (!(std::isfinite(x[k]) && std::isfinite(y[k]) && std::isfinite(z[k]) &&
              std::isfinite(m[k]) && std::isfinite(r2[k])))
            return SG_ERR_NONFINITE;
        if (!(m[k] > 0.0)) return SG_ERR_MASS;
        if (r2[k] == 0.0) return SG_ERR_RANGE2;
    }
    return SG_OK;
}

void legacy_fail(const char *where, int status)
{
    g_last_status = status;
    if (status == SG_OK) return;
    std::fprintf(stderr, " SPHgrav(%s): error %d: %s\n", where, status, sg_status_string(status));
    if (!env_flag("SPHGRAV_NOABORT")) std::exit(EXIT_FAILURE);
}

}  // namespace

extern "C" {

int sg_init_device(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SG_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return SG_ERR_NO_DEVICE;
    HIP_TRY(hipSetDevice(device));
    legacy().device = device;
    return SG_OK;
}

int sg_firsthalf(int n, int n_lower, int my_length, const double *x, const double *y,
                 const double *z, const double *m, const double *range2, int nkernel)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Legacy &L = legacy();
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (n < 1 || n_lower < 1 || my_length < 0 || (int64_t)n_lower - 1 + my_length > n)
        return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2) return SG_ERR_ARG;
    int rc = check_host_inputs(n, x, y, z, m, range2);
    if (rc != SG_OK) return rc;
    sg_context *c = nullptr;
    if ((rc = legacy_context(&c)) != SG_OK) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (L.pending) {
        // the reference allows one outstanding call; a second firsthalf
        // replaces the first (its results are dropped)
        std::fprintf(stderr, " SPHgrav: firsthalf called twice; dropping the pending result\n");
        HIP_TRY(hipStreamSynchronize(L.cur->stream));
        L.pending = false;
    }
    const int ni = my_length;
    for (int k = 0; k < 5; ++k) {
        int64_t cap = c->cap_in;
        if ((rc = grow(&c->in[k], &cap, n)) != SG_OK) return rc;
    }
    c->cap_in = std::max<int64_t>(c->cap_in, n);
    for (int k = 0; k < 4; ++k) {
        int64_t cap = c->cap_out;
        if ((rc = grow(&c->out[k], &cap, std::max(ni, 1))) != SG_OK) return rc;
    }
    c->cap_out = std::max<int64_t>(c->cap_out, std::max(ni, 1));
    HIP_TRY(hipEventRecord(L.t0, c->stream));
    const double *src[5] = {x, y, z, m, range2};
    for (int k = 0; k < 5; ++k)
        HIP_TRY(hipMemcpyAsync(c->in[k], src[k], sizeof(double) * (size_t)n, hipMemcpyHostToDevice,
                               c->stream));
    // the host arrays may be reused as soon as firsthalf returns
    HIP_TRY(hipStreamSynchronize(c->stream));
    rc = sg_compute_device(c, n, n_lower - 1, ni, c->in[0], c->in[1], c->in[2], c->in[3],
                           c->in[4], nkernel, legacy_precision(), c->out[0], c->out[1], c->out[2],
                           c->out[3]);
    if (rc != SG_OK) return rc;
    L.pending = true;
    L.cur = c;
    L.ibeg = n_lower - 1;
    L.ni = ni;
    return SG_OK;
}

int sg_grav_psi(int n, int n_lower, int my_length, const double *x, const double *y,
                const double *z, const double *m, const double *range2, int nkernel, double *psi)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (nkernel < 0 || nkernel > 2) return SG_ERR_NKERNEL;
    if (n < 1 || n_lower < 1 || my_length < 0 || (int64_t)n_lower - 1 + my_length > n)
        return SG_ERR_ARG;
    if (!x || !y || !z || !m || !range2 || !psi) return SG_ERR_ARG;
    int rc = check_host_inputs(n, x, y, z, m, range2);
    if (rc != SG_OK) return rc;
    sg_context *c = nullptr;
    if ((rc = legacy_context(&c)) != SG_OK) return rc;
    // stream order keeps a pending firsthalf's inputs intact until its
    // kernels have read them; the call then waits for both
    return sg_psi_host(c, n, n_lower - 1, my_length, x, y, z, m, range2, nkernel,
                       psi + (n_lower - 1));
}

int sg_lasthalf(double *ax, double *ay, double *az, double *pot)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Legacy &L = legacy();
    if (!L.pending) return SG_ERR_STATE;
    if (L.ni > 0 && (!ax || !ay || !az || !pot)) return SG_ERR_ARG;
    sg_context *c = L.cur;
    L.pending = false;
    HIP_TRY(hipSetDevice(c->device));
    int rc = sg_synchronize(c);
    if (rc != SG_OK) return rc;
    double *dst[4] = {ax, ay, az, pot};
    for (int k = 0; k < 4 && L.ni > 0; ++k)
        HIP_TRY(hipMemcpyAsync(dst[k] + L.ibeg, c->out[k], sizeof(double) * (size_t)L.ni,
                               hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(L.t1, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!env_flag("SPHGRAV_QUIET")) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, L.t0, L.t1) == hipSuccess)
            std::fprintf(stderr, " >>> SPHgrav_lib took %g sec\n", ms / 1000.0);
    }
    return SG_OK;
}

// ------------------------------------------------ reference ABI symbols

void firsthalf_grav_forces_(int *n, int *n_lower, int *my_length, double *px, double *py,
                            double *pz, double *mass, double *range2, int *q, int *nkernel)
{
    (void)q;
    if (!n || !n_lower || !my_length || !nkernel) {
        legacy_fail("firsthalf_grav_forces", SG_ERR_ARG);
        return;
    }
    legacy_fail("firsthalf_grav_forces",
                sg_firsthalf(*n, *n_lower, *my_length, px, py, pz, mass, range2, *nkernel));
}

void lasthalf_grav_forces_(int *n, double *ax, double *ay, double *az, double *pot, int *myrank)
{
    (void)n;
    (void)myrank;
    legacy_fail("lasthalf_grav_forces", sg_lasthalf(ax, ay, az, pot));
}

void gpu_init_dev_(int *device, double *theta)
{
    (void)theta;
    legacy_fail("gpu_init_dev", device ? sg_init_device(*device) : SG_ERR_ARG);
}

}  // extern "C"

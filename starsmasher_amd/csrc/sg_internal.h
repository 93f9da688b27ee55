// sg_internal.h -- shared declarations between the host layer (sg_capi.cpp)
// and the device code (sg_direct.hip, sg_prep.hip, sg_sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sg {

// Particles are visited in j-tiles of TJ consecutive (space-filling-curve
// sorted) particles.  The sorted arrays are padded to a multiple of TJ with
// massless copies of the last particle.
constexpr int TJ = 256;

// Per j-tile bounding box (FP64, absolute coordinates) and largest support
// radius.  64 bytes.  The same record describes each of the TJ / TC chunks
// of TC consecutive particles of a tile, the unit of the near/far split.
constexpr int TC = 64;
constexpr int CHUNKS = 4;   // TJ / TC
struct TileMeta {
    double lo[3];
    double hi[3];
    double maxH;
    double pad;
};

// Chunk-visit statistics (visits, near visits, wide visits, FP64 pairs):
// STAT_SLOTS copies of the STAT_N counters, each copy on its own 128-byte
// line; a workgroup adds to copy blockIdx.x % STAT_SLOTS and the host sums
// the copies (one copy for all waves serialised their atomics in L2 and cost
// up to a fifth of a small set's pass).
constexpr int STAT_N = 4;
constexpr int STAT_SLOTS = 64;
constexpr int STAT_STRIDE = 16;                          // u64 per copy
constexpr int STAT_WORDS = STAT_SLOTS * STAT_STRIDE;     // u64 in the buffer

// Sorted, structure-of-arrays particle set as the pair kernels read it.
//   x,y,z : FP64 positions
//   m     : FP64 mass (0 for padding)
//   Hs    : FP64 support radius H = sqrt|range2|, negative for point
//           particles (range2 < 0, no self potential)
//   idx   : original particle index (-1 for padding)
struct SortedSet {
    const double *x, *y, *z, *m, *Hs;
    const int *idx;
    int64_t count;      // real particles
    int64_t padded;     // multiple of TJ
};

struct PairArgs {
    SortedSet j;             // all n particles
    SortedSet i;             // i-range particles in sorted order (may alias j)
    const TileMeta *tiles;   // padded/TJ entries
    const TileMeta *chunks;  // padded/TC entries
    int64_t ni;              // i-particles
    int ibeg;                // outputs go to out[idx - ibeg]
    int nkernel;
    int ntiles;              // j-tiles
    int jsplit;              // j partitions per i-block
    int niblocks;            // i-blocks
    int ipl;                 // i-particles per lane of the far kernel (groups of 64*ipl)
    int64_t ngroups;         // i-groups (waves of the far kernel)
    double *ax, *ay, *az, *pot;   // final outputs, input order, length ni
    double *partial;              // [jsplit][4][ni] sorted-order partial sums
    unsigned char *nearflag;      // [ngroups][ntiles]: bit c set = chunk c of the tile may
                                  // hold softened pairs and takes the near kernel
    unsigned long long *stats;    // [STAT_WORDS] statistics copies (may be null)
    double *psi;                  // grad-h softening term, input order (null: not formed)
    double *psipart;              // [jsplit][ni] its partial sums (with psi)
};

// A batch of independent particle sets laid out back to back in one padded
// sorted array; each set starts at a tile boundary.  A set of more than
// BATCH_JTILES tiles has its j range cut into jsplit strided partitions
// (partition p: tiles p, p + jsplit, ...), each with its own FP64 partial
// sums, so that no work item holds a whole large set's j range.
constexpr int BATCH_JTILES = 64;
struct BatchSet {
    int64_t pofs;      // first padded slot of the set (multiple of TJ)
    int64_t n;         // real particles
    int tile0;         // first tile
    int ntiles;        // tiles of the set
    int64_t gofs;      // first near group (groups of 64*ipl i)
    int64_t flagofs;   // first near flag
    int64_t partofs;   // first partial sum: [jsplit][4][ntiles * TJ] from here
    int jsplit;        // j partitions (>= 1)
};

struct BatchArgs {
    SortedSet j;                 // padded sorted arrays of all sets (also the i-set)
    const TileMeta *tiles;
    const TileMeta *chunks;
    const BatchSet *sets;
    int nsets;
    const int2 *items;           // (set, i-block * jsplit + partition), longest first
    int nitems;
    int64_t ngroups;             // near groups over all sets
    int64_t padded;              // slots over all sets
    int nkernel;
    double *partial;             // per set at partofs: [jsplit][4][ntiles * TJ]
    unsigned char *nearflag;
    unsigned *counter;           // work-queue head
    double *ax, *ay, *az, *pot;  // outputs, input order
    unsigned long long *stats;
};

// Vector-Jacobian product of one full evaluation (all n particles as i):
// L = sum_i w_i . a_i + u_i phi_i, gradients w.r.t. x, y, z, m and the
// support radius H.  Same tiles, origin and near flags as the forward pass.
struct AdjArgs {
    SortedSet p;                        // sorted particles (i and j)
    const double *wx, *wy, *wz, *u;     // sorted-order cotangents (0 on padding)
    const TileMeta *tiles;
    const TileMeta *chunks;
    int64_t n;
    int nkernel;
    int ntiles;
    int jsplit;
    int niblocks;                       // i-blocks of the far kernel
    int64_t ngroups;                    // waves of the far kernel (64 i each)
    double *gx, *gy, *gz, *gm, *gH;     // outputs, input order
    double *partial;                    // [jsplit][5][n] sorted order
    unsigned char *nearflag;            // [ngroups][ntiles]
    unsigned long long *stats;          // [STAT_WORDS] statistics copies
};

// Adjoint launchers (sg_adjoint.hip): far, near, reduce in stream order.
hipError_t launch_adj_far(int precision, const AdjArgs &a, int wg, hipStream_t s);
hipError_t launch_adj_near(int precision, const AdjArgs &a, hipStream_t s);
hipError_t launch_adj_reduce(const AdjArgs &a, hipStream_t s);
bool adj_config_supported(int wg);

// Launchers (sg_direct.hip).  Return hipError_t of the launch.
// An evaluation is far kernel -> near kernel -> reduce, in stream order.
hipError_t launch_pair_far(int precision, const PairArgs &a, int wg, hipStream_t s);
hipError_t launch_pair_near(int precision, const PairArgs &a, hipStream_t s);
hipError_t launch_reduce_partials(const PairArgs &a, hipStream_t s);
// Grad-h term alone (a.psi, no accelerations): groups of 64 i, a.ngroups of them.
hipError_t launch_psi_only(const PairArgs &a, hipStream_t s);
bool pair_config_supported(int precision, int wg, int ipl);
// Batch: queue reset, persistent far kernel on nwg workgroups, near kernel,
// scatter.  Contains no memset, so it can be captured in a graph.
hipError_t launch_batch(int precision, const BatchArgs &a, int wg, int ipl, int nwg, hipStream_t s);
int batch_resident_workgroups(int precision, int nkernel, int wg, int ipl, int device);
bool batch_config_supported(int wg, int ipl);

// Preparation (sg_prep.hip).
struct PrepBuffers {
    // scratch
    double *red;          // [nred_blocks * 8] reduction partials
    int nred_blocks;
    double *bbox;         // [8]: lo xyz, hi xyz, (unused)
    int *err;             // [1]: validation bit flags
    uint64_t *keys_in, *keys_out;
    int *vals_in, *vals_out;
    // sorted output arrays (capacity >= padded)
    double *sx, *sy, *sz, *sm, *sH;
    int *sidx;
    TileMeta *tiles;
    TileMeta *chunks;     // CHUNKS per tile
    // i-subset sorted arrays
    double *ix, *iy, *iz, *im, *iH;
    int *iidx;
};

enum ErrBits { ERR_MASS = 1, ERR_RANGE2 = 2, ERR_NONFINITE = 4 };

hipError_t launch_validate_bbox(const double *x, const double *y, const double *z,
                                const double *m, const double *r2, int n,
                                PrepBuffers &b, hipStream_t s);
hipError_t launch_morton(const double *x, const double *y, const double *z, int first, int count,
                         const PrepBuffers &b, hipStream_t s);
hipError_t launch_gather(const double *x, const double *y, const double *z,
                         const double *m, const double *r2, const int *perm, int count,
                         int64_t padded, double *ox, double *oy, double *oz, double *om,
                         double *oH, int *oidx, hipStream_t s);
// Cotangents gathered into sorted order (0 on padding); non-finite values
// set ERR_NONFINITE in b.err.
hipError_t launch_gather_cot(const double *wx, const double *wy, const double *wz,
                             const double *u, const int *perm, int count, int64_t padded,
                             double *owx, double *owy, double *owz, double *ou,
                             const PrepBuffers &b, hipStream_t s);
hipError_t launch_tile_meta(const double *sx, const double *sy, const double *sz,
                            const double *sH, int ntiles, TileMeta *tiles, TileMeta *chunks,
                            hipStream_t s);
hipError_t launch_fill_zero_int(int *p, int n, hipStream_t s);
// 1/2 sum m pot in FP64 into work[nwork - 1] (work holds nwork doubles)
hipError_t launch_potential_energy(const double *m, const double *pot, int n, double *work,
                                   int nwork, hipStream_t s);
// Batch preparation: keys = (set << 42) | top 42 bits of the Morton code, and
// the gather into the padded per-set layout.
hipError_t launch_morton_batch(const double *x, const double *y, const double *z, int n,
                               const int64_t *offsets, int nsets, const PrepBuffers &b,
                               hipStream_t s);
hipError_t launch_gather_batch(const double *x, const double *y, const double *z,
                               const double *m, const double *r2, const int *perm,
                               const int64_t *offsets, const BatchSet *sets, int nsets,
                               int64_t padded, double *ox, double *oy, double *oz, double *om,
                               double *oH, int *oidx, hipStream_t s);

// Sorting (sg_sort.hip): stable sort of 63-bit keys with int payloads (radix
// sort; merge sort for sets of up to 65,536).
hipError_t sort_pairs_temp_bytes(int n, size_t *bytes);
hipError_t sort_pairs(void *tmp, size_t tmp_bytes, const uint64_t *kin, uint64_t *kout,
                      const int *vin, int *vout, int n, hipStream_t s);
// The same stable order by merge sort: kernels only, no memset (for work a
// caller may capture in a graph).
hipError_t merge_sort_pairs_temp_bytes(int n, size_t *bytes);
hipError_t merge_sort_pairs(void *tmp, size_t tmp_bytes, const uint64_t *kin, uint64_t *kout,
                            const int *vin, int *vout, int n, hipStream_t s);

// ------------------------------------------------- one evaluation (sg_capi.cpp)
// Sizes of one evaluation of n particles with ni i-particles at a tile
// shape (wg 0: chosen from ni); the context path (sg_compute_device) and the static-scratch path
// (sg_compute_device_static) size their device buffers from it, so both
// launch the same kernels with the same j-partitions.
struct EvalPlan {
    int64_t padded;        // sorted j slots (multiple of TJ)
    int ntiles;
    int wg, ipl, jsplit, niblocks;
    int64_t ngroups;       // waves of the far kernel
    int64_t part_elems;    // doubles of j-partition partial sums: jsplit * 4 * ni
    int64_t flag_elems;    // near-flag bytes (FP64 and grad-h passes; else 1)
    int64_t psipart_elems; // grad-h partial sums: jsplit * ni (with psi; else 0)
};
int auto_jsplit(int niblocks, int wg, int ntiles, int64_t ni);
// i-ranges up to this many particles take workgroups of 128 when the
// configured workgroup size is 0 (automatic), larger ones 256
constexpr int AUTO_SMALL_NI = 65536;
// the HIP error that sg_status_string(SG_ERR_HIP) describes
void record_hip_error(hipError_t e);
EvalPlan plan_eval(int n, int ni, int wg, int ipl, int jsplit_cfg, int precision, bool psi);

// The device memory one evaluation works in.
struct EvalBuffers {
    PrepBuffers pb;
    void *sort_tmp;
    size_t sort_tmp_bytes;
    double *partial;
    unsigned char *nearflag;
    unsigned long long *stats;   // [STAT_WORDS]
    double *psipart;             // with psi only
    bool merge_sort;             // sort with merge_sort_pairs (same order, no memset)
};

// Enqueues validation, sort, gather, tile boxes, the pair kernels and the
// reduction on s.  ev (may be null) = events recorded before the preparation,
// before the far kernel, after it and at the end.
hipError_t enqueue_eval(const EvalPlan &p, const EvalBuffers &b, int n, int ibeg, int ni,
                        const double *x, const double *y, const double *z, const double *m,
                        const double *range2, int nkernel, int precision, double *ax, double *ay,
                        double *az, double *pot, double *psi, hipStream_t s,
                        const hipEvent_t *ev);

}  // namespace sg

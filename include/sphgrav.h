/*
 * sphgrav.h -- C ABI of the MI355X (gfx950) direct-summation softened
 * self-gravity library (starsmasher_amd/libsphgrav_hip.so).
 *
 * Two layers:
 *
 * 1. The three Fortran-callable symbols of the reference gravity library,
 *    exported verbatim so an unmodified Fortran host links against this
 *    library instead of libSPHgrav.a.  All arguments are by reference
 *    (gfortran implicit interface, default integer = int32).
 *
 * 2. An sg_* API with int status returns, host-pointer and device-pointer
 *    entry points, used by the legacy layer, the tests, bench.py and new
 *    Fortran call sites (starsmasher_amd/fortran/sphgrav_mod.f90).
 *
 * Physics (units G = 1): for every i in the i-range and every j in 0..n-1,
 *   a_i   += m_j * 0.5*(F(r,H_i) + F(r,H_j)) * (x_j - x_i)
 *   phi_i += m_j * 0.5*(P(r,H_i) + P(r,H_j))
 * with H = sqrt|range2| (= 2 hp), F = A(r/H)/H^3 and P = P(r/H)/H when r < H
 * and F = r^-3, P = -1/r otherwise (cubic spline or Wendland W3,3 / C4 by
 * nkernel).  The i == j self pair adds -c*flag_j*(m_j/H_i + m_j/H_j) to phi
 * (c = 1.4, 1.9140625, 1.71875), flag_j = range2_j > 0 (gas) and 0 for point
 * particles (range2 < 0).  A distinct pair at exactly r = 0 contributes the
 * finite r->0 limit -c*(m_j/H_i + m_j/H_j) (the reference CUDA kernel
 * produces NaN there, grav_force_direct.cu:110-112).
 */
#ifndef SPHGRAV_H
#define SPHGRAV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Layer 1: reference ABI (pbe/src/SPHgrav_lib2/grav_force_direct.cu:648-677) */
/* ------------------------------------------------------------------------ */

/* Replaces firsthalf_grav_forces_ (grav_force_direct.cu:650-656).
 * j-set = particles 1..*n, i-set = *n_lower .. *n_lower + *my_length - 1
 * (1-based).  Copies the host arrays synchronously (the caller may reuse
 * them on return), launches the computation asynchronously and returns.
 * *q is ignored (legacy threads-per-particle knob, as in the reference).
 * On a bad argument or HIP error: prints a message and exits with
 * EXIT_FAILURE (the reference asserts / CUDA_SAFE_CALL exits), unless the
 * environment sets SPHGRAV_NOABORT=1, in which case the call records the
 * status (sg_last_status) and returns. */
void firsthalf_grav_forces_(int *n, int *n_lower, int *my_length,
                            double *px, double *py, double *pz,
                            double *mass, double *range2, int *q, int *nkernel);

/* Replaces lasthalf_grav_forces_ (grav_force_direct.cu:658-664).
 * Blocks until the preceding firsthalf's results are ready and writes
 * ax/ay/az/pot[n_lower-1 .. n_lower+my_length-2] (0-based); other entries
 * are untouched.  *n and *myrank are ignored, as in the reference.  Prints
 * " >>> SPHgrav_lib took %g sec" to stderr unless SPHGRAV_QUIET=1. */
void lasthalf_grav_forces_(int *n, double *ax, double *ay, double *az,
                           double *pot, int *myrank);

/* Replaces gpu_init_dev_ (grav_force_direct.cu:672-675): select the HIP
 * device for this process.  *theta is ignored (tree-library opening angle). */
void gpu_init_dev_(int *device, double *theta);

/* ------------------------------------------------------------------------ */
/* Layer 2: sg_* API                                                         */
/* ------------------------------------------------------------------------ */

enum sg_status {
    SG_OK = 0,
    SG_ERR_ARG = 1,          /* bad size / range / null pointer            */
    SG_ERR_NKERNEL = 2,      /* nkernel not in {0,1,2}                      */
    SG_ERR_MASS = 3,         /* mass <= 0 (grav_force_direct.cu:602)        */
    SG_ERR_RANGE2 = 4,       /* range2 == 0 (grav_force_direct.cu:601)      */
    SG_ERR_NONFINITE = 5,    /* NaN / Inf in positions, mass or range2      */
    SG_ERR_HIP = 6,          /* HIP runtime error                           */
    SG_ERR_NO_DEVICE = 7,    /* no HIP device / device index out of range   */
    SG_ERR_STATE = 8,        /* lasthalf without a pending firsthalf         */
    SG_ERR_CONFIG = 9,       /* unsupported tile configuration              */
    SG_ERR_SIZE = 10         /* n above SG_STATIC_MAX_N (static entry point) */
};

enum sg_precision {
    SG_PREC_FP32C = 0,       /* FP32 pair arithmetic, hi/lo positions relative
                                to an FP64 tile origin, FP64 cross-tile sums   */
    SG_PREC_FP64 = 1         /* plain FP64 pair arithmetic                     */
};

const char *sg_status_string(int status);
int sg_last_status(void);         /* status of the last legacy-layer call */
int sg_device_count(int *count);

typedef struct sg_context sg_context;

/* Create a context bound to HIP device `device` (owns device buffers that
 * grow on demand and a default-priority stream). */
int sg_create(int device, sg_context **out);
int sg_destroy(sg_context *ctx);

/* Enqueue the context's work on `stream` (a hipStream_t of the context's
 * device, e.g. the caller's current torch stream) instead of its own stream;
 * NULL restores the context's own stream.  Waits for work already queued. */
int sg_set_stream(sg_context *ctx, void *stream);

/* Device-resident evaluation.  All array pointers are DEVICE pointers on the
 * context's device.  Inputs: n particles (x,y,z,m,range2 as in layer 1).
 * i-range: 0-based [ibeg, ibeg+ni).  Outputs ax..pot have length ni and
 * receive the i-range in input order.  Work is enqueued on the context's
 * stream; call sg_synchronize before reading results.  Input validation runs
 * on the device; its verdict is returned by sg_synchronize. */
int sg_compute_device(sg_context *ctx, int n, int ibeg, int ni,
                      const double *x, const double *y, const double *z,
                      const double *m, const double *range2,
                      int nkernel, int precision,
                      double *ax, double *ay, double *az, double *pot);

/* Wait for the context's stream; returns the validation status of the last
 * sg_compute_device (SG_ERR_MASS, SG_ERR_RANGE2, SG_ERR_NONFINITE) or SG_OK. */
int sg_synchronize(sg_context *ctx);

/* Gravitational potential energy epot = 1/2 sum_i m_i pot_i of device
 * arrays m and pot (n >= 1), summed in FP64 in a fixed order (the same value
 * for the same inputs on every call), written to the host double *epot.
 * Replaces the host loop of output.f:232-252 when the outputs stay in HBM.
 * Ordered on the context's stream after earlier work; blocks until done. */
int sg_potential_energy(sg_context *ctx, int n, const double *m, const double *pot,
                        double *epot);

/* Milliseconds of the main pair kernel(s) of the last completed evaluation,
 * measured with HIP events on the context's stream (call after
 * sg_synchronize). */
int sg_last_kernel_ms(sg_context *ctx, float *ms);

/* Host-pointer evaluation (blocking): copies in, computes, copies out.
 * Same semantics as sg_compute_device with host arrays. */
int sg_compute_host(sg_context *ctx, int n, int ibeg, int ni,
                    const double *x, const double *y, const double *z,
                    const double *m, const double *range2,
                    int nkernel, int precision,
                    double *ax, double *ay, double *az, double *pot);

/* Tile configuration of the pair kernel for a precision.
 * wg: threads per workgroup (64, 128 or 256), or 0 (the default) to take 128
 * for i-ranges of up to 65,536 particles and 256 above; ipl: i-particles per
 * lane (1 or 2); jsplit: number of j-partitions per i-block (>= 1, partial
 * sums are reduced in FP64 in a fixed order); 0 (the default) selects it
 * from n. */
int sg_set_config(sg_context *ctx, int precision, int wg, int ipl, int jsplit);
int sg_get_config(sg_context *ctx, int precision, int *wg, int *ipl, int *jsplit);
/* The tile shape the last single-set evaluation on the context ran with
 * (sg_compute_device and the calls built on it, the automatic choices
 * resolved); SG_ERR_STATE before one, or when the last call was of another
 * kind (batch, psi-only, vector-Jacobian product). */
int sg_last_config(sg_context *ctx, int *wg, int *ipl, int *jsplit);

/* Batch of independent particle sets (no reference counterpart: the
 * reference evaluates one set per call).  Set s holds particles
 * offsets[s] .. offsets[s+1]-1 of the device arrays (offsets: HOST array of
 * nsets+1 entries, offsets[0] = 0, every set non-empty); particles interact
 * only within their set.  Outputs (device, length offsets[nsets]) are in
 * input order.  All sets run in one persistent launch whose workgroups take
 * (set, i-block, j-partition) items from a queue, longest first (sets of
 * more than 16,384 particles split their j range into partitions whose FP64
 * partial sums are added in a fixed order).  With use_graph the
 * queue reset, far, near and scatter kernels are captured once in a HIP graph
 * and replayed while the plan, pointers and precision stay the same; the
 * preparation (validation, sort, gather) runs eagerly before the replay.
 * Asynchronous like sg_compute_device; sg_synchronize reports validation. */
int sg_compute_batch_device(sg_context *ctx, int nsets, const long long *offsets,
                            const double *x, const double *y, const double *z,
                            const double *m, const double *range2,
                            int nkernel, int precision, int use_graph,
                            double *ax, double *ay, double *az, double *pot);

/* Tile shape of the batch kernel (wg 64/128/256, ipl 1/2) per precision. */
int sg_set_batch_config(sg_context *ctx, int precision, int wg, int ipl);
int sg_get_batch_config(sg_context *ctx, int precision, int *wg, int *ipl);

/* HIP-event timings of the last completed evaluation, in milliseconds:
 * prep = validation, curve sort, gather and tile boxes; far = the pair
 * kernel (in the FP32C forward pass it also takes the near chunks, listed
 * per wave and run after its far loop); near = the separate near-field
 * kernel where there is one (FP64 pass, grad-h pass, adjoint) plus the
 * final reduction.  Any pointer may be NULL. */
int sg_last_timings(sg_context *ctx, float *prep_ms, float *far_ms, float *near_ms);

/* Statistics of the last evaluation: number of (wave, 64-particle j-chunk)
 * visits and how many of them were near (a chunk for which the box and
 * per-particle tests cannot rule out softened pairs for the wave), i.e. took
 * the near-field path. */
int sg_last_tile_stats(sg_context *ctx, long long *visits, long long *near_visits);

/* The same counts and two more of the FP32C pass (no reference
 * counterpart; they state where FP64 arithmetic enters the pair loop):
 * wide_visits = far (wave, chunk) visits whose pair differences were formed
 * in FP64 and rounded once (chunks close to the wave relative to its spread
 * about the workgroup origin), and fp64_pairs = (i, j) pairs evaluated
 * entirely in FP64 (those that may be softened, r < max(H_i, H_j), the
 * self pair included; pairs with the massless padding that fills the last
 * j-tile are evaluated too but not counted).  A chunk visit covers 64 j for each i of a wave
 * (64 per lane and i-particle per lane).  Zero for the FP64 pass, whose
 * pairs are all FP64.  Any pointer may be NULL. */
int sg_last_precision_stats(sg_context *ctx, long long *visits, long long *near_visits,
                            long long *wide_visits, long long *fp64_pairs);

/* Evaluation that also forms the grad-h gravitational softening term of
 * each gas particle i of the i-range (range2_i > 0; 0 for point particles):
 *     psi_i = hp_i^-2 sum_{j: r_ij < 2 hp_i} m_j dphidh(r_ij / hp_i)
 * over every j of the set, j = i included (dphidh(0) = c), with dphidh the
 * closed forms of kernels.f:215-279 for the nkernel in use and hp = H/2.
 * Replaces the neighbour sum bonet1_psi of advance.f90:716-737 (with
 * nselfgravity = 1), which the reference evaluates from a 10^7-entry table
 * on the CPU; here it is exact and comes out of the near-field pass that
 * already visits every pair with r < H_i.  psi (length ni, input order)
 * is written with ax..pot; other semantics as sg_compute_device / _host
 * (the pass runs the near kernel with one i per lane). */
int sg_compute_psi_device(sg_context *ctx, int n, int ibeg, int ni,
                          const double *x, const double *y, const double *z,
                          const double *m, const double *range2,
                          int nkernel, int precision,
                          double *ax, double *ay, double *az, double *pot, double *psi);
int sg_compute_psi_host(sg_context *ctx, int n, int ibeg, int ni,
                        const double *x, const double *y, const double *z,
                        const double *m, const double *range2,
                        int nkernel, int precision,
                        double *ax, double *ay, double *az, double *pot, double *psi);

/* The grad-h term alone (no accelerations), FP64: psi over the i-range as
 * in sg_compute_psi_device, from a pass that visits only the 64-particle
 * j-chunks within the i-group's largest H.  Replaces the psi part of
 * advance.f90:716-737 (bonetsumwithphi) at a fraction of a gravity
 * evaluation's cost.  Device pointers, asynchronous. */
int sg_psi_device(sg_context *ctx, int n, int ibeg, int ni,
                  const double *x, const double *y, const double *z,
                  const double *m, const double *range2, int nkernel, double *psi);
int sg_psi_host(sg_context *ctx, int n, int ibeg, int ni,
                const double *x, const double *y, const double *z,
                const double *m, const double *range2, int nkernel, double *psi);

/* Adjoint (no reference counterpart; the reference has no derivative of the
 * gravity step beyond kernels.f:215-279's dphi/dh, which the H-gradient here
 * generalises).  For the forward outputs over all n particles (ibeg = 0,
 * ni = n) and cotangents w = (wx, wy, wz), u of length n, computes the
 * gradient of
 *     L = sum_i wx_i ax_i + wy_i ay_i + wz_i az_i + u_i pot_i
 * with respect to x_i, y_i, z_i, m_i and H_i = sqrt|range2_i| (the support
 * radius; dL/dhp = 2 dL/dH for range2 = +-4 hp^2).  The r = 0 rule is the
 * forward one: the self pair contributes through its potential term, a
 * distinct coincident pair through the r -> 0 limit (the force factor F(0)
 * is finite, so its x-gradient is F(0) (m_k w_l - m_l w_k)).
 * precision SG_PREC_FP32C: Newtonian far pairs in FP32 with hi/lo positions
 * and FP64 folding, softened pairs in FP64; SG_PREC_FP64: all FP64.
 * Device pointers, asynchronous on the context stream like
 * sg_compute_device (sg_synchronize reports validation, including
 * non-finite cotangents as SG_ERR_NONFINITE); sg_last_timings covers it.
 * The tile shape is the forward far kernel's workgroup size for the
 * precision (one i per lane). */
int sg_vjp_device(sg_context *ctx, int n,
                  const double *x, const double *y, const double *z,
                  const double *m, const double *range2, int nkernel, int precision,
                  const double *wx, const double *wy, const double *wz, const double *u,
                  double *gx, double *gy, double *gz, double *gm, double *gH);

/* Host-pointer adjoint (blocking), same semantics as sg_vjp_device. */
int sg_vjp_host(sg_context *ctx, int n,
                const double *x, const double *y, const double *z,
                const double *m, const double *range2, int nkernel, int precision,
                const double *wx, const double *wy, const double *wz, const double *u,
                double *gx, double *gy, double *gz, double *gm, double *gH);

/* Legacy-layer equivalents with status returns (used by layer 1). */
int sg_init_device(int device);
int sg_firsthalf(int n, int n_lower, int my_length,
                 const double *x, const double *y, const double *z,
                 const double *m, const double *range2, int nkernel);
int sg_lasthalf(double *ax, double *ay, double *az, double *pot);
/* Blocking grad-h term of particles n_lower .. n_lower+my_length-1 (1-based)
 * of the set 1..n on the legacy layer's device, written to
 * psi[n_lower-1 .. n_lower+my_length-2] like lasthalf writes its outputs:
 * the Fortran-callable form of sg_psi_host for the density step
 * (advance.f90:716-737), called before firsthalf. */
int sg_grav_psi(int n, int n_lower, int my_length,
                const double *x, const double *y, const double *z,
                const double *m, const double *range2, int nkernel, double *psi);

/* ------------------------------------------------------------------------ */
/* Evaluation without device allocation                                      */
/* ------------------------------------------------------------------------ */

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The evaluation of sg_compute_device -- same
 * arguments and results as a context at its default configuration
 * (automatic workgroup size, one i per lane, automatic j-split) -- on the
 * CURRENT HIP device, enqueued
 * on `stream` (a hipStream_t of that device; NULL = the null stream).  It
 * allocates nothing: every scratch buffer (sorted copies, tile boxes,
 * partial sums, sort storage, counters) is carved from one file-scope
 * device array of SG_STATIC_SCRATCH_BYTES per device, held by each device
 * the library runs on until the process ends.
 *   - n above SG_STATIC_MAX_N returns SG_ERR_SIZE;
 *   - every call on a device shares that device's array and leaves its
 *     contents behind: calls that overlap on one device (different streams
 *     or host threads) race; calls in one stream are ordered;
 *   - the call issues kernels only (no memset or copy; the Morton sort is
 *     the merge sort the context path also takes at these sizes), so it may
 *     be captured in a graph, in global capture mode too;
 *     the first call on a device must come before any capture (it looks up
 *     the array's device address);
 *   - input validation runs on the device: when `err` (a DEVICE int) is
 *     given it receives the bit flags SG_VALID_* (0 = inputs valid), readable
 *     once the stream reaches the end of the call. */
#define SG_STATIC_MAX_N 32768
#define SG_STATIC_SCRATCH_BYTES 41599488
enum sg_validation_flags {
    SG_VALID_MASS = 1,       /* a mass <= 0                                 */
    SG_VALID_RANGE2 = 2,     /* a range2 == 0                               */
    SG_VALID_NONFINITE = 4   /* NaN / Inf in positions, masses or range2    */
};
int sg_compute_device_static(int n, int ibeg, int ni,
                             const double *x, const double *y, const double *z,
                             const double *m, const double *range2,
                             int nkernel, int precision,
                             double *ax, double *ay, double *az, double *pot,
                             int *err, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SPHGRAV_H */

/* nt2hip.h — C ABI of the MI355X (gfx950) evaluator for NT2 table expressions.
 *
 * This is the accelerator slot of NT2's evaluation-site stack
 * (accelerator_site<shared_memory_site<...>>, modules/sdk/functor/include/
 * nt2/sdk/functor/site.hpp:50-53): where NT2 would dispatch transform_ and
 * inner_fold_ to a CUDA site (modules/arch/cuda/include/nt2/sdk/cuda/
 * cuda.hpp:22-41 registers none), a caller binds these entry points.  Every
 * entry point evaluates one fused loop nest in one pass over its tables (one
 * kernel launch; long folds add a small launch that combines segment
 * partials, and scans of long runs or of many rows take three passes:
 * segment totals, their prefixes, the rescan).
 *
 * Conventions
 *  - Plain C types only.  Pointers are caller-owned device pointers unless a
 *    comment says otherwise; sizes are element counts (int64_t).
 *  - Matrices are column-major (NT2's default storage order,
 *    core/settings/storage_order.hpp:22) with an explicit leading dimension.
 *  - All calls are asynchronous on the given stream (NULL = the default
 *    stream) except where noted.  Returns NT2H_OK (0) or a negative
 *    nt2h_status; arguments are validated before anything is launched.
 *  - Reductions are deterministic: fixed combine order, no float atomics.
 *    Their results (double) may be written to device memory (asynchronous)
 *    or to ordinary host memory (the call then synchronises the stream).
 *  - Thread safety: calls on distinct streams may run concurrently; the
 *    library keeps one reduction workspace per (device, stream).
 *  - Stream capture (HIP graphs): the entry points can be recorded and
 *    replayed; they record kernels only (no memset or copy nodes).  While the
 *    stream is capturing, result pointers must be device memory, and the
 *    same call must have run once on that stream beforehand so that its
 *    workspace and scratch exist (allocation is not capturable).  A graph
 *    uses the workspace of its capture stream: it must not run at the same
 *    time as calls on that stream or as another graph recorded on it
 *    (replays and calls in stream order are fine).  Growing a stream's
 *    scratch never frees the buffer a graph recorded; nt2h_release_stream
 *    does.
 *  - Per-stream state (reduction workspace, scratch, look-back state) lives
 *    until nt2h_release_stream(s), to be called before the stream is
 *    destroyed.
 */
#ifndef NT2HIP_H
#define NT2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nt2h_stream_opaque* nt2h_stream_t; /* a hipStream_t; NULL = default stream */

typedef enum nt2h_status {
  NT2H_OK = 0,
  NT2H_ERR_INVALID_ARG = -1,  /* bad enum value, negative size, bad leading dimension */
  NT2H_ERR_NULL_POINTER = -2, /* a required pointer is NULL while n > 0 */
  NT2H_ERR_BAD_OP = -3,       /* unsupported functor id for this entry point */
  NT2H_ERR_NO_DEVICE = -4,    /* no HIP device, or device index out of range */
  NT2H_ERR_HIP = -5,          /* a HIP runtime call failed (see nt2h_last_hip_error) */
  NT2H_ERR_OUT_OF_MEMORY = -6,/* device allocation failed */
  NT2H_ERR_TOO_LARGE = -7,    /* size beyond what the entry point supports */
  NT2H_ERR_CAPTURE = -8,      /* the call needs a workspace allocation while its stream is capturing:
                                 run it once on that stream before the capture */
  NT2H_ERR_WAIT = -9          /* reported by nt2h_sync: a bounded wait inside a kernel on the stream
                                 gave up (or a launch shared its state with another): that call's
                                 results are invalid */
} nt2h_status;

/* Elementwise functor ids (the NT2 functors on the path). */
typedef enum nt2h_op {
  NT2H_OP_EXP = 0,         /* nt2::exp      (exponential/.../expo/f_expo_reduction.hpp:49-85)      */
  NT2H_OP_LOG = 1,         /* nt2::log      (exponential/.../logs/f_log.hpp:59-87)                 */
  NT2H_OP_SQRT = 2,        /* nt2::sqrt     (boost/simd/.../scalar/sqrt.hpp:46)                    */
  NT2H_OP_ERF = 3,         /* nt2::erf      (euler/.../scalar/erf.hpp:76-100)                      */
  NT2H_OP_ERFC = 4,        /* nt2::erfc     (euler/.../scalar/erfc.hpp:77-105)                     */
  NT2H_OP_NORMCDF = 5,     /* nt2::normcdf  (statistics/.../generic/normcdf.hpp:34)                */
  NT2H_OP_TANH = 6,        /* nt2::tanh     (hyperbolic/.../scalar/tanh.hpp:45-54)                 */
  NT2H_OP_SIN = 7,         /* nt2::sin      (trigonometric/.../generic/sin.hpp:27)                 */
  NT2H_OP_COS = 8,         /* nt2::cos      (trigonometric/.../generic/cos.hpp:28)                 */
  NT2H_OP_NORMPDF = 9,     /* nt2::normpdf  (statistics/.../generic/normpdf.hpp:31)                */
  NT2H_OP_FASTNORMCDF = 10,/* nt2::fastnormcdf, float only (statistics/.../fastnormcdf.hpp:39-54)  */
  NT2H_OP_EXP2 = 11,       /* nt2::exp2, float only                                               */
  NT2H_OP_EXP10 = 12       /* nt2::exp10, float only                                              */
} nt2h_op;

/* Arithmetic of the kernels.
 *  EXACT  NT2's scalar functor (two-rounding fma/horner, IEEE division):
 *         bit-identical to NT2's scalar path for every functor, float and
 *         double, NT2's quirks included (double exp is inf on [709.436,
 *         709.783), double log misreads subnormals, double erf is +1 below
 *         -6); sin/cos reproduce fdlibm's Payne-Hanek where NT2 runs it
 *         (float: every float checked; double: sampled doubles of every
 *         binade).
 *  FUSED  NT2's algorithms and coefficients with FMA and a refined hardware
 *         reciprocal, within a stated ULP bound of EXACT, except where
 *         NT2's double path is wrong: there FUSED gives the finite exp, the
 *         prescaled subnormal log and erf = -1 below -6.
 *  FAST   column fold only (nt2h_colfold_gc_f32_prec): hardware exp2
 *         (v_exp_f32) and a reduction of the argument prescaled by 1/pi;
 *         each term's error is within a few 2^-24 of |term| plus 2^-24 of
 *         its argument, stated against the FP64 fold as a fraction of the
 *         column's sum of |terms|. */
typedef enum nt2h_precision { NT2H_EXACT = 0, NT2H_FUSED = 1, NT2H_FAST = 2 } nt2h_precision;

/* Summation order of the reductions. */
typedef enum nt2h_sum_mode {
  NT2H_SUM_NAIVE = 0,    /* per-lane float accumulation, tree, FP64 cross-block combine            */
  NT2H_SUM_KAHAN = 1,    /* per-lane TwoSum compensation (dot: + exact product error), same tree    */
  NT2H_SUM_PAIRWISE = 2  /* per-lane binary-counter pairwise cascade, same tree                    */
} nt2h_sum_mode;

typedef enum nt2h_copy_kind { NT2H_H2H = 0, NT2H_H2D = 1, NT2H_D2H = 2, NT2H_D2D = 3, NT2H_DEFAULT = 4 } nt2h_copy_kind;

/* ------------------------------------------------------------ runtime
 * Replaces NT2's CUDA buffer/locality layer: cuda_buffer/allocator
 * (modules/arch/cuda/include/nt2/sdk/memory/cuda/{allocator,buffer}.hpp:51-85),
 * pinned_allocator.hpp:89-100 and memory::copy (memory/functions/cuda/copy.hpp:52-65).
 * Errors return codes instead of CUDA_ERROR's exit(1) (sdk/cuda/cuda.hpp:44-50). */
int nt2h_init(int device);              /* select device, create per-device state              */
int nt2h_device_count(int* count);
const char* nt2h_strerror(int code);
int nt2h_last_hip_error(void);          /* hipError_t of the last failing HIP call (this thread) */
/* Device tables.  Allocations of 1 MiB or more are 4 KiB aligned and start at
 * successive 4 KiB "colours" modulo 64 KiB, so that the tables one loop nest
 * streams together do not share an offset modulo the 2 MiB allocation
 * alignment (same HBM channel and bank group).  nt2h_free takes exactly the
 * pointer nt2h_malloc returned.  A zero-byte request sets *ptr to NULL and
 * returns NT2H_OK; freeing NULL is a no-op.  nt2h_memcpy is asynchronous on
 * `s` (kind is an nt2h_copy_kind; NT2H_DEFAULT infers the direction from the
 * pointers); a zero-byte copy is a no-op, an unknown kind NT2H_ERR_INVALID_ARG. */
int nt2h_malloc(void** ptr, size_t bytes);
int nt2h_free(void* ptr);
int nt2h_host_alloc(void** ptr, size_t bytes); /* pinned host memory */
int nt2h_host_free(void* ptr);
int nt2h_memcpy(void* dst, const void* src, size_t bytes, int kind, nt2h_stream_t s);
/* Pageable host memory <-> device, staged through pinned buffers in 8 MiB
 * chunks with the host-side copy of one chunk overlapping the DMA of the
 * next (NT2's blocked host transfers: sdk/cuda/settings/specific_cuda.hpp
 * cu_buffers, transfer_htd/transfer_dth).  Synchronous: on return dst holds
 * the data. */
int nt2h_upload(void* dst_device, const void* src_host, size_t bytes, nt2h_stream_t s);
int nt2h_download(void* dst_host, const void* src_device, size_t bytes, nt2h_stream_t s);
/* Wait for the stream; then report NT2H_ERR_WAIT once if a bounded wait in a
 * kernel on it gave up since the last check (the look-back scans), and
 * clear that flag. */
int nt2h_sync(nt2h_stream_t s);
/* Free the per-(current device, stream) state the library keeps for s: the
 * reduction workspace, the scratch (with every buffer it outgrew) and the
 * look-back state.  Waits for the stream first.  Call it when no graph
 * recorded on s will run again, before destroying s; a later call on s (or
 * on a new stream with the same handle) starts from fresh state. */
int nt2h_release_stream(nt2h_stream_t s);
/* A per-stream buffer of the library, for code that launches the library's
 * fold and scan kernels itself (the C++ front end instantiates them over
 * expression trees, include/nt2hip/{fold,scan}_kernels.hpp): kind 0 is the
 * stream's grow-only scratch, kind 1 its look-back state (zeroed when first
 * allocated, then kept consistent by the kernels).  *out holds at least
 * `bytes`; the buffer lives until nt2h_release_stream(s).  NT2H_ERR_CAPTURE
 * if it would have to grow while s is capturing. */
int nt2h_stream_buffer(nt2h_stream_t s, int kind, size_t bytes, void** out);
/* Device bytes the library holds in per-stream state (all devices and
 * streams: workspaces, scratch with the buffers it outgrew, look-back
 * state), for accounting and leak checks. */
int nt2h_state_bytes(size_t* bytes);
/* Test control of the look-back scans' bounded waits, for tests of the error
 * path only: spin_cap polls before a waiting tile gives up (0 = the
 * default, 2^20), and never_ready != 0 makes every predecessor's state read
 * as not yet published, so that each waiting tile gives up after spin_cap
 * polls (every kernel still ends; nt2h_sync then reports NT2H_ERR_WAIT).
 * Affects later launches in this process. */
int nt2h_test_set_lookback_wait(unsigned spin_cap, int never_ready);

/* ---------------------------------------------------- elementwise (transform_)
 * NT2 transform: core/functions/transform.hpp:55-69 -> elementwise run_assign
 * (dsl/functions/elementwise/run.hpp:21-35), one fused loop nest per assignment. */

/* Y = alpha*X + Y (demo/saxpy/nt2/saxpy_nt2.cpp:41); two roundings, no
 * contraction: bit-identical to NT2's scalar path.  x and y may not overlap
 * except x == y. */
int nt2h_saxpy_f32(float alpha, const float* x, float* y, int64_t n, nt2h_stream_t s);

/* y = op(x) (one functor node of a table expression); y may equal x. */
int nt2h_map_f32(int op, int precision, const float* x, float* y, int64_t n, nt2h_stream_t s);
int nt2h_map_f64(int op, int precision, const double* x, double* y, int64_t n, nt2h_stream_t s);

/* Black-Scholes European call and put, one fused loop nest with 5 input and
 * 2 output streams:
 *   d1 = (log(S/K) + (r + sqr(sig)/2)*T) / (sig*sqrt(T)), d2 = d1 - sig*sqrt(T)
 *   call = S*normcdf(d1) - K*exp(-r*T)*normcdf(d2)
 *   put  = K*exp(-r*T)*normcdf(-d2) - S*normcdf(-d1)
 * (the formula NT2's blackscholes demo evaluates as four loop nests,
 * demo/blackscholes/nt2/blackandscholes_nt2.cpp:53-60, with the textbook d1). */
int nt2h_bs_fwd_f32(int precision, const float* S, const float* K, const float* T, const float* r,
                    const float* sig, float* call, float* put, int64_t n, nt2h_stream_t s);
int nt2h_bs_fwd_f64(int precision, const double* S, const double* K, const double* T, const double* r,
                    const double* sig, double* call, double* put, int64_t n, nt2h_stream_t s);

/* ------------------------------------------------------ reductions (inner_fold_)
 * NT2 inner_fold with plus/Zero/sum (core/reduction/.../scalar/inner_fold.hpp:24-53,
 * dsl/functions/reduction/run.hpp:30-104); globalsum (core/functions/globalsum.hpp:45)
 * and dot (reduction/functions/expr/dot.hpp:28) over a whole table.
 * One pass over the data: per-block partials combined in FP64 in a fixed
 * order — by the last-arriving block (small grids), or by a second small
 * launch over the stored partials (grids of 4096 blocks or more, where the
 * per-block tickets would cost bandwidth).  The same bits on every call for
 * a given size.  *out receives the sum as a double.  Data holding +-inf or
 * NaN, or whose partial sums overflow the float lanes, give the plain sum's
 * IEEE value (+-inf; NaN for a NaN or for +inf with -inf): the
 * compensation of an infinite running sum is dropped, not added. */
int nt2h_sum_f32(const float* x, int64_t n, int mode, double* out, nt2h_stream_t s);
int nt2h_dot_f32(const float* x, const float* y, int64_t n, int mode, double* out, nt2h_stream_t s);
int nt2h_sum_f64(const double* x, int64_t n, double* out, nt2h_stream_t s);
int nt2h_dot_f64(const double* x, const double* y, int64_t n, double* out, nt2h_stream_t s);

/* Other reduction_ functors along one dimension of a contiguous table.
 * NT2 reshapes a reduction along dimension k to (lo, mid, hi) — the product
 * of the dimensions before k, dimension k, the product of those after —
 * and folds over the middle index with inner_fold (lo == 1) or outer_fold
 * (core/reduction/include/nt2/dsl/functions/reduction/run.hpp:48-93):
 *   out[i + lo*h] = fold_{m = 0..mid-1} x[i + lo*(m + mid*h)]
 * sum(x, k), prod(x, k), maximum(x, k), minimum(x, k), asum1, asum2,
 * norm2 = sqrt(asum2), mean = rec(size) * sum (core/reduction/include/nt2/
 * core/functions/expr/{asum1,asum2,norm2,mean}.hpp; boost/simd/reduction/
 * functions/{sum,prod,maximum,minimum}.hpp).  Whole-table globalX(x) is
 * lo = hi = 1, mid = numel.  out holds lo*hi values of the table type.
 *  - maximum/minimum are bit-identical to NT2's left-to-right a > b ? a : b
 *    fold from Maxinit = -FLT_MAX (Valmax = FLT_MAX; -DBL_MAX/DBL_MAX for
 *    double), NaN and signed-zero behaviour included.
 *  - sum-like folds are compensated in the table's type with a compensated
 *    FP64 cross-segment combine (more accurate than NT2's sequential order,
 *    so not bit-identical); +-inf and NaN come out where NT2's recurrence
 *    has them (as for the whole-table sums above);
 *    prod keeps an FP64 mantissa with a separate exponent, so partial
 *    products never overflow (NT2's float product can, giving inf).
 *  - mid == 0 yields the neutral element (mean: NaN, as rec(0)*0). */
typedef enum nt2h_fold_op {
  NT2H_FOLD_SUM = 0,
  NT2H_FOLD_PROD = 1,
  NT2H_FOLD_MAXIMUM = 2,
  NT2H_FOLD_MINIMUM = 3,
  NT2H_FOLD_ASUM1 = 4,
  NT2H_FOLD_ASUM2 = 5,
  NT2H_FOLD_NORM2 = 6,
  NT2H_FOLD_MEAN = 7
} nt2h_fold_op;
int nt2h_fold_f32(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* out, nt2h_stream_t s);
int nt2h_fold_f64(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* out, nt2h_stream_t s);

/* Scans along one dimension (NT2 inner_scan / outer_scan over the middle
 * index of the same (lo, mid, hi) reshaping; core/swar/include/nt2/core/
 * functions/scalar/{inner,outer}_scan.hpp, details/scan_step.hpp):
 *   y[i + lo*(m + mid*h)] = fold_{m' <= m} x[i + lo*(m' + mid*h)]
 * cumsum (plus, Zero), cumprod (multiplies, One), cummax (max, Minf),
 * cummin (min, Inf) — boost/simd/swar/functions/{cumsum,cumprod,cummax,
 * cummin}.hpp.  cummax/cummin are bit-identical to NT2's recurrence (NaN
 * resets included); cumsum uses compensated float pairs with FP64 carries
 * (+-inf and NaN from where the recurrence reaches them) and cumprod an
 * FP64 mantissa with a separate exponent (no intermediate overflow).  y may
 * equal x. */
typedef enum nt2h_scan_op {
  NT2H_SCAN_CUMSUM = 0,
  NT2H_SCAN_CUMPROD = 1,
  NT2H_SCAN_CUMMAX = 2,
  NT2H_SCAN_CUMMIN = 3
} nt2h_scan_op;
int nt2h_scan_f32(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* y, nt2h_stream_t s);
int nt2h_scan_f64(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* y, nt2h_stream_t s);

/* ------------------------------------------------------------- adjoints
 * NT2 has no differentiation; these pair a forward loop nest with its
 * hand-written reverse sweep, recomputing forward intermediates in
 * registers (nothing is stored between the sweeps).
 *
 * Black-Scholes call price with cotangent w (portfolio P = sum_i w_i*price_i):
 *   price_i and the input cotangents dS_i = w_i*dprice/dS_i, dK, dT, dr, dsig;
 * totals[0] = sum_i dr_i and totals[1] = sum_i dsig_i (dP/dr and dP/dsigma
 * when r and sigma are one scalar each), compensated, FP64 combine.
 * totals may be NULL.  Float kernel: FUSED arithmetic.  */
int nt2h_bs_fwd_adj_f32(const float* S, const float* K, const float* T, const float* r, const float* sig,
                        const float* w, float* price, float* dS, float* dK, float* dT, float* dr, float* dsig,
                        double* totals, int64_t n, nt2h_stream_t s);
/* The same over host tables (pageable memory): blocks of `block` options
 * (<= 0: 2^20) alternate between two streams so that staging, H2D, the
 * kernel and D2H of neighbouring blocks overlap; totals (host memory, may be
 * NULL) are the FP64 sum of the block totals in block order.  Outputs equal
 * nt2h_bs_fwd_adj_f32's bit for bit.  Synchronous. */
int nt2h_bs_fwd_adj_f32_host(const float* S, const float* K, const float* T, const float* r, const float* sig,
                             const float* w, float* price, float* dS, float* dK, float* dT, float* dr, float* dsig,
                             double* totals, int64_t n, int64_t block);
/* No counterpart in NT2: nt2h_bs_fwd_adj_f32 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernel, with its reduction
 * workspace in a static device array of the library (about 325 KiB per GPU,
 * held until the process ends) instead of a per-stream allocation: the call
 * makes no hipMalloc/hipFree.  Outputs and totals are bit-identical to
 * nt2h_bs_fwd_adj_f32's for 16-byte aligned tables.
 *  - n <= 2^24 options (16384 workgroups' partials); larger n returns
 *    NT2H_ERR_TOO_LARGE (split the portfolio and add the totals).
 *  - totals is NULL or device memory (a host pointer: NT2H_ERR_INVALID_ARG).
 *  - Every call on one GPU shares that GPU's array: calls that overlap in
 *    time on one GPU (other streams, other host threads) race and give wrong
 *    totals.  Calls in stream order are safe; each launch leaves the array's
 *    ticket counters at zero for the next.
 *  - The array is resolved and cleared by nt2h_init (or by the first call on
 *    a device nt2h_init did not see; that first call is not capturable and
 *    returns NT2H_ERR_CAPTURE while its stream is capturing). */
int nt2h_bs_fwd_adj_f32_noalloc(const float* S, const float* K, const float* T, const float* r,
                                const float* sig, const float* w, float* price, float* dS, float* dK, float* dT,
                                float* dr, float* dsig, double* totals, int64_t n, nt2h_stream_t s);
int nt2h_bs_fwd_adj_f64(const double* S, const double* K, const double* T, const double* r,
                        const double* sig, const double* w, double* price, double* dS, double* dK, double* dT,
                        double* dr, double* dsig, double* totals, int64_t n, nt2h_stream_t s);

/* Column fold of a 2-D table expression and its adjoint, one pass over X:
 *   z_j = sum_i exp(-X_ij^2/(2 s_j^2)) * cos(k_j X_ij + p_j)
 * (NT2: sum(exp(-sqr(X)/(2*sqr(sx(s)))) .* cos(...), 1) -> inner_fold over
 * the rows of each column).  With gz the cotangent of z (NULL = all ones),
 * dz_ds_j = gz_j*sum e c X^2/s^3, dz_dk_j = -gz_j*sum e sin(.) X,
 * dz_dp_j = -gz_j*sum e sin(.).  Gradient pointers may all be NULL (forward
 * only).  X is rows x cols, column-major, leading dimension ld >= rows;
 * per-lane FP64 accumulation of float pairwise sums of four elements, then a
 * fixed-order FP64 combine per column. */
int nt2h_colfold_gc_f32(const float* X, int64_t rows, int64_t cols, int64_t ld, const float* s_,
                        const float* k_, const float* p_, const float* gz, float* z, float* dz_ds,
                        float* dz_dk, float* dz_dp, nt2h_stream_t s);
/* The same with the arithmetic chosen: NT2H_FAST (what nt2h_colfold_gc_f32
 * runs: exp(-x^2/(2 s^2)) = v_exp_f32(x^2 * (-log2(e)/(2 s^2))), and
 * sin/cos(k x + p) = +-sin/cos(pi f) with f = y - rint(y), y = (k/pi) x + p/pi)
 * or NT2H_FUSED (NT2's exp and sin/cos algorithms with FMA); NT2H_EXACT is
 * refused (NT2H_ERR_INVALID_ARG).  Stated bounds against the FP64 fold, each
 * as a fraction of the column's sum of |terms| of that output, at
 * 16384 x 8192: FAST 3.5e-7, FUSED 1.7e-7 (tested on every column).  A
 * column of a few rows carries single terms' errors: up to 2^-24 |k x + p|
 * of a term, from rounding the argument once.  z and the gradients are NaN
 * where the FP64 fold's are (inf or NaN in X, s, k or p). */
int nt2h_colfold_gc_f32_prec(int precision, const float* X, int64_t rows, int64_t cols, int64_t ld,
                             const float* s_, const float* k_, const float* p_, const float* gz, float* z,
                             float* dz_ds, float* dz_dk, float* dz_dp, nt2h_stream_t s);

#ifdef __cplusplus
}
#endif
#endif /* NT2HIP_H */

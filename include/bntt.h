/*
 * bntt.h -- C ABI of the MI355X (gfx950) benchNTT hot path: 128-bit modular
 * forward NTT and vvadd / vvsub / vvmul / axpy mod q.
 *
 * The reference has no library: its kernels are plain C functions inside six
 * benchmark executables, with N a compile-time macro and the BLAS length fixed
 * at 1024.  This header exports the same operations with the sizes as
 * arguments; benchntt_compat.h maps the reference's own call shapes onto them.
 * Entry point -> reference interface it replaces (files of the reference's src/):
 *   bntt_ntt_fwd, bntt_ntt_fwd_dev   ntt()        ntt_scalar.c:582-627
 *                                    (ntt_pre/ntt/ntt_post of ntt_avx2.c:507-655,
 *                                     ntt_avx512.c:386-528: same network)
 *   bntt_vvadd[_dev]                 vector_add() blas_scalar.c:578-584
 *   bntt_vvsub[_dev]                 vector_sub() blas_scalar.c:586-592
 *   bntt_vvmul[_dev]                 vector_mul() blas_scalar.c:594-600
 *   bntt_axpy[_dev]                  axpy()       blas_scalar.c:602-608
 *
 * Data: a residue is 16 bytes in the native layout of unsigned __int128 (low
 * 64-bit word first), exactly the reference's uint128_t arrays.  Device arrays
 * must be 16-byte aligned.  Results are bit-identical to the reference's
 * scalar functions (addmod128/submod128/mulmod128, ntt_scalar.c:49-199, with
 * MBITS = 124) for every input, including their edge behaviour.
 *
 * Errors: every entry point returns BNTT_OK or a nonzero code; the message of
 * the calling thread's last failure is in bntt_last_error().  Nothing falls
 * back to the CPU: with no usable gfx950 device every call fails.
 *
 * _dev entry points take device pointers and enqueue on `stream` (a
 * hipStream_t; NULL = the null stream) without synchronizing.  The others take
 * host pointers, stage through library-owned device buffers and return when
 * the results are in host memory.  Calls are thread-safe.
 */
#ifndef BNTT_H
#define BNTT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A 128-bit value in native unsigned __int128 layout. */
typedef struct bntt_u128 {
    uint64_t lo;
    uint64_t hi;
} bntt_u128;

enum {
    BNTT_OK = 0,
    BNTT_EINVAL = 1, /* bad size, pointer or alignment */
    BNTT_EHIP = 2,   /* HIP runtime error */
    BNTT_ENODEV = 3, /* no gfx950 device */
    BNTT_ENOMEM = 4  /* device allocation failed */
};

/* Version of this ABI (major*10000 + minor*100 + patch). */
int bntt_version(void);

/* Select the current HIP device and check it is gfx950.  Called implicitly by
 * every entry point; calling it first moves start-up cost out of timed code. */
int bntt_init(void);

/* Message of the calling thread's last failure ("" if none). */
const char *bntt_last_error(void);

/* ------------------------------------------------------------ device pointers */

/* Vector ops: c may be a or b (in place) but may not overlap them otherwise
 * (BNTT_EINVAL); n above SIZE_MAX / 48 is BNTT_EINVAL. */

/* c[i] = a[i] + b[i] mod m, i < n (ntt_scalar.c:49-79 semantics). */
int bntt_vvadd_dev(void *c, const void *a, const void *b, size_t n, const bntt_u128 *m, void *stream);
/* c[i] = a[i] - b[i] mod m. */
int bntt_vvsub_dev(void *c, const void *a, const void *b, size_t n, const bntt_u128 *m, void *stream);
/* c[i] = a[i] * b[i] mod m (Barrett with mu = floor(2^251 / m) in the reference). */
int bntt_vvmul_dev(void *c, const void *a, const void *b, size_t n, const bntt_u128 *m, const bntt_u128 *mu,
                   void *stream);
/* c[i] = a[i] + x * b[i] mod m (blas_scalar.c:602-608: the scalar multiplies b). */
int bntt_axpy_dev(void *c, const void *a, const void *b, const bntt_u128 *x, size_t n, const bntt_u128 *m,
                  const bntt_u128 *mu, void *stream);

/* Forward NTT of `count` transforms of N = 2^log2n points, stored one after
 * another in x (count*N residues) into y, all with the twiddle table tw
 * (N residues; entry 0 unused, stage s uses entries [2^s, 2^(s+1))).
 * 1 <= log2n <= 24.  For log2n <= 12 every transform is done in one launch and
 * y may equal x.  For log2n > 12 the transform takes several passes through
 * device memory: tmp must point to bntt_ntt_tmp_bytes(log2n, count) bytes of
 * device memory and y must not overlap x.  count = 1 is the plain
 * one-transform path; for log2n 11 .. 13 with count = 1 the tmp holds
 * bntt_ntt_tmp_bytes bytes, 2^log2n + 16 residues (optional at 11 and 12):
 * given, the transform splits over 4, 8 or 16 workgroups joined by a grid
 * barrier in one launch (in a launch being captured into a graph: one
 * workgroup, or two launches).  tmp must not overlap x or y, and a call that
 * uses it must not run concurrently with another call using the same tmp. */
int bntt_ntt_fwd_dev(void *y, const void *x, const void *tw, int log2n, size_t count, const bntt_u128 *m,
                     const bntt_u128 *mu, void *tmp, void *stream);
/* Bytes of tmp for bntt_ntt_fwd_dev: 0 when log2n <= 12 (no tmp needed),
 * except count = 1 at log2n 11 and 12 (an optional tmp of 2^log2n + 16
 * residues); count * 2^log2n residues above that, except count = 1 at log2n
 * 13 (2^log2n + 16 residues); SIZE_MAX when count * 2^log2n * 16 does not fit
 * a size_t. */
size_t bntt_ntt_tmp_bytes(int log2n, size_t count);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): bntt_ntt_fwd_dev without the tmp argument, allocating
 * nothing.  Multi-pass sizes (log2n > 12), and one transform of 2^11 .. 2^13
 * points (split over several workgroups, as with a tmp), keep their
 * intermediates in a fixed array of BNTT_NOALLOC_MAX_RESIDUES residues
 * (16 MiB) that the library holds on each GPU it runs on until the process
 * ends, so count * 2^log2n may not exceed it there (BNTT_EINVAL); one-pass
 * sizes have no limit.  Every call on
 * one GPU shares that GPU's array: calls that overlap on the device (on
 * different streams, or from several threads) race and must be ordered by
 * the caller.  Results are those of bntt_ntt_fwd_dev (same kernels). */
#define BNTT_NOALLOC_MAX_RESIDUES ((size_t)1 << 20)
int bntt_ntt_fwd_dev_noalloc(void *y, const void *x, const void *tw, int log2n, size_t count, const bntt_u128 *m,
                             const bntt_u128 *mu, void *stream);

/* ------------------------------------------------------------ uneven batches */
/*
 * No counterpart in the reference, whose ntt() handles one transform of a
 * compile-time size (ntt_scalar.c:582-627): these entry points run many
 * transforms of different sizes, each exactly the reference network.
 */

/* One forward transform of a batch: device pointers, 16-byte aligned. */
typedef struct bntt_ntt_task {
    const void *x;  /* 2^log2n residues in */
    void *y;        /* 2^log2n residues out (must not overlap x above 2^11 points) */
    const void *tw; /* 2^log2n-entry twiddle table */
    int32_t log2n;  /* 1 .. 24 */
    int32_t reserved;
} bntt_ntt_task;

/* One pointwise product c[i] = a[i] * b[i] mod m, i < n (device pointers).
 * When a or b overlaps the y of transforms in the same plan, the product runs
 * after those transforms.  When a and b are exactly the outputs of two
 * transforms of the plan of size n and one twiddle table, the two transforms
 * share their work items and the product is computed in their last pass from
 * on-chip data (no second read of the outputs).  Tasks of a plan are
 * independent: a transform's y may not overlap another transform's x or y, a
 * product's c may not overlap any transform's buffers or another product's
 * operands (BNTT_EINVAL). */
typedef struct bntt_mul_task {
    const void *a;
    const void *b;
    void *c;
    uint64_t n;
} bntt_mul_task;

/* Per-size choice of the batch kernel: points per thread per register
 * sub-pass 2^elog (1..3) and number of passes through device memory. */
typedef struct bntt_tune_entry {
    int32_t log2n;
    int32_t elog;
    int32_t passes;
    int32_t reserved;
} bntt_tune_entry;

typedef struct bntt_batch_info {
    int32_t levels;              /* launches per execution */
    int32_t reserved;
    uint64_t items;              /* work items over all levels */
    uint64_t tmp_bytes;          /* plan-owned intermediate storage */
    uint64_t level_items[8];
    int32_t level_grid[8];       /* persistent workgroups per launch */
    uint32_t level_lds_bytes[8]; /* dynamic LDS per workgroup */
} bntt_batch_info;

typedef struct bntt_batch_plan bntt_batch_plan;

#define BNTT_PLAN_GRAPH 1 /* capture the launches in a hipGraph on the first
                             execute with a non-null stream, replay after */
/* Where transform items read their twiddles.  Default: a digit copy in device
 * memory (L1/L2-resident; measured fastest on MI355X).  TW_LDS: staged in LDS
 * when the tables fit beside the data tile with three workgroups per CU.
 * TW_WIDE: staged whenever they fit in 80 KiB of LDS with the data (two
 * workgroups per CU). */
#define BNTT_PLAN_TW_LDS 2
#define BNTT_PLAN_TW_WIDE 4
/* The twiddle tables are not modified after the plan is created: their digit
 * copies and large-entry checks are made once, at creation, instead of at the
 * start of every execution.  (Without it an execution picks up tables changed
 * in place, at the cost of one conversion pass per execution.) */
#define BNTT_PLAN_TW_FIXED 8

/* Build a plan: device work lists, intermediates for multi-pass sizes, and
 * per-size kernel choices (tune may be NULL for built-in defaults).  The plan
 * records pointers, not data: execute it any number of times after the
 * inputs change in place. */
int bntt_batch_plan_create(bntt_batch_plan **plan, const bntt_ntt_task *tasks, size_t ntasks,
                           const bntt_mul_task *muls, size_t nmuls, const bntt_u128 *m, const bntt_u128 *mu,
                           const bntt_tune_entry *tune, size_t ntune, int flags);
/* Enqueue one execution on `stream` (no host synchronization).  Executions of
 * one plan share its work-queue counters and intermediates, so one runs at a
 * time on the device: an execution enqueued on another stream than the
 * previous one first waits (on the device) for the previous one to finish.
 * On a stream that is being captured, the launches join the caller's graph,
 * which may then be replayed any number of times (no memset nodes; the
 * work-queue counters rearm themselves); the caller must not let replays of
 * that graph overlap each other or other executions of the plan. */
int bntt_batch_plan_execute(bntt_batch_plan *plan, void *stream);
int bntt_batch_plan_info(const bntt_batch_plan *plan, bntt_batch_info *info);
/* Synchronizes the device, then frees the plan. */
int bntt_batch_plan_destroy(bntt_batch_plan *plan);

/* The same transforms with one launch sequence per transform, the plain path
 * of bntt_ntt_fwd_dev applied in turn (tmp: 2^max_log2n residues when any
 * task is above 2^12 points). */
int bntt_ntt_launch_each_dev(const bntt_ntt_task *tasks, size_t ntasks, const bntt_u128 *m, const bntt_u128 *mu,
                             void *tmp, void *stream);

/* ------------------------------------------- polynomial multiplication server */
/*
 * No counterpart in the reference.  Negacyclic products c = a * b mod
 * (x^N + 1, q) for an odd prime q < 2^122 with 2^(N's log2 + 1) | q - 1 (the
 * reference's q_116 = 2^115 + 47*2^17 + 1 serves N up to 2^16), built from the
 * same forward network with exact Montgomery arithmetic.  Operands and results
 * are canonical residues (< q) in device memory; twiddle tables stay resident
 * in the server.  One server is driven by one host thread at a time.  The
 * server runs on streams of its own that do not wait for other streams:
 * operands must be in device memory before a call, and c may be read by any
 * stream after it returns.  A result does not depend on the other requests of
 * its batch.  Within one call, a request's c may not overlap another
 * request's a, b or c (BNTT_EINVAL); c may equal its own a or b.
 */
typedef struct bntt_pm_request {
    const void *a; /* N residues */
    const void *b; /* N residues */
    void *c;       /* N residues out */
    int32_t log2n;
    int32_t reserved;
} bntt_pm_request;

typedef struct bntt_pm_config {
    uint64_t hbm_budget_bytes; /* arena for intermediates: per request in flight, 32*N bytes
                                  for N < 2^10 points, 64*N bytes from 2^10 on */
    int32_t max_inflight;      /* batches on the device at once (1..8 streams) */
    int32_t max_batch;         /* requests per batch */
    int32_t min_log2n;         /* sizes served: tables are built at creation */
    int32_t max_log2n;         /* <= 16 */
    uint64_t admit_timeout_ns; /* bntt_pm_run_trace drops a request not admitted within this
                                  time of its arrival (0: never) */
    uint64_t max_batch_butterflies; /* bntt_pm_run_trace closes a batch before a request that
                                       would take its work, 3 * N/2 * log2 N butterflies per
                                       request, above this (a batch holds at least one
                                       request; 0: no limit beyond max_batch) */
} bntt_pm_config;

/* done_ns of a request that bntt_pm_run_trace dropped (its c is untouched). */
#define BNTT_PM_DROPPED UINT64_MAX

typedef struct bntt_pm_stats {
    uint64_t requests;
    uint64_t batches;
    uint64_t max_inflight_requests; /* most requests on the device at once */
    uint64_t admission_waits;       /* requests that arrived and waited for budget */
    uint64_t dropped;               /* requests not admitted within admit_timeout_ns */
    uint64_t max_batch_ns;          /* longest batch, launch to completion seen by the host */
    uint64_t max_loop_gap_ns;       /* longest interval between two scheduler passes (a host stall) */
} bntt_pm_stats;

typedef struct bntt_pm_server bntt_pm_server;

int bntt_pm_server_create(bntt_pm_server **server, const bntt_u128 *q, const bntt_pm_config *config);
int bntt_pm_server_destroy(bntt_pm_server *server);
/* Multiply n <= max_batch requests as one batch; returns when c is written. */
int bntt_pm_multiply(bntt_pm_server *server, const bntt_pm_request *reqs, size_t n);

/* The two halves of the product pipeline on their own, for count transforms of
 * N = 2^log2n residues lying back to back; returns when y is written.  y may
 * equal x (in place) but may not overlap it otherwise.  Two-pass sizes (log2n
 * > 11) take N*16 bytes of the HBM budget per transform of a launch.
 *   BNTT_PM_FORWARD  y[i] = sum_j x[j] psi^((2 brv(i) + 1) j) for the server's
 *                    root psi of order 2N (the negacyclic NTT, outputs in
 *                    bit-reversed order brv of the log2n index bits): the
 *                    forward network with the table psi_s^(2 brv_s(k) + 1).
 *   BNTT_PM_INVERSE  its inverse: x in BNTT_PM_FORWARD's output order, y the
 *                    coefficients in natural order.
 * The pointwise product of two forward outputs, taken mod q, is the forward
 * output of the negacyclic product. */
enum { BNTT_PM_FORWARD = 0, BNTT_PM_INVERSE = 1 };
int bntt_pm_transform(bntt_pm_server *server, int direction, void *y, const void *x, int log2n, size_t count);
/* Serve requests that arrive at arrival_ns[i] (sorted, ns after the call
 * starts): arrived requests are admitted in order while the HBM budget has
 * room, batched, and launched on a free stream while earlier batches still
 * run; done_ns[i] receives the time request i was seen complete, or
 * BNTT_PM_DROPPED if it waited longer than the server's admit_timeout_ns for
 * admission.  Returns when every request is done or dropped. */
int bntt_pm_run_trace(bntt_pm_server *server, const bntt_pm_request *reqs, const uint64_t *arrival_ns, size_t n,
                      uint64_t *done_ns, bntt_pm_stats *stats);

/* ------------------------------------------------ data formats and inputs */

/* The reference's scalar primitives on the host, bit-identical to
 * addmod128 / submod128 / mulmod128 (ntt_scalar.c:49-79, 81-104, 173-199, MBITS
 * = 124), including their edge behaviour.  Arguments by value, as there. */
bntt_u128 bntt_addmod128(bntt_u128 a, bntt_u128 b, bntt_u128 m);
bntt_u128 bntt_submod128(bntt_u128 a, bntt_u128 b, bntt_u128 m);
bntt_u128 bntt_mulmod128(bntt_u128 a, bntt_u128 b, bntt_u128 m, bntt_u128 mu);

/* Word order of the reference's fixture files and AVX code: residue i is
 * hilo[2i] = high word, hilo[2i+1] = low word (config.c:6-197 reads "hi lo"
 * pairs into this layout; ntt_avx512.c:386-528 takes it).  n residues; out may
 * equal the input (the _dev versions reject any other overlap).  Host versions
 * convert on the host; _dev versions run on `stream`. */
int bntt_hilo_to_u128(bntt_u128 *out, const uint64_t *hilo, size_t n);
int bntt_u128_to_hilo(uint64_t *hilo, const bntt_u128 *in, size_t n);
int bntt_hilo_to_u128_dev(void *out, const void *hilo, size_t n, void *stream);
int bntt_u128_to_hilo_dev(void *hilo, const void *in, size_t n, void *stream);

/* Twiddle tables of roots of unity for the forward network (2^log2n entries,
 * entry 0 = 0; stage s reads entries 2^s + k, k < 2^s):
 *   BNTT_TW_CYCLIC      w_s^brv_s(k),        w_s of order 2^(s+1): the network
 *                       is then the cyclic DFT in bit-reversed order;
 *   BNTT_TW_NEGACYCLIC  p_s^(2 brv_s(k)+1),  p_s of order 2^(s+2): the
 *                       negacyclic NTT in bit-reversed order;
 *   *_INV               the entry-wise inverses of the same tables;
 * with w_s, p_s powers of g^((q-1)/2^v) for the least quadratic non-residue g
 * and 2^v || q - 1.  q must be a prime < 2^122 with v >= log2n (cyclic) or
 * log2n + 1 (negacyclic): q_116 = 2^115 + 47*2^17 + 1 serves up to 2^16 points
 * negacyclic.  Not possible for the reference's q_N (v = 4), whose drivers use
 * random tables. */
enum { BNTT_TW_CYCLIC = 0, BNTT_TW_NEGACYCLIC = 1, BNTT_TW_CYCLIC_INV = 2, BNTT_TW_NEGACYCLIC_INV = 3 };
int bntt_twiddles_dev(void *tw, int log2n, const bntt_u128 *q, int kind, void *stream);
int bntt_twiddles(bntt_u128 *tw, int log2n, const bntt_u128 *q, int kind);

/* n residues uniform in [0, q) from seed (q >= 2): residue i keeps the first
 * of the draws t = 0, 1, ... below q, draw t being the splitmix64 words
 * 2(t n + i) (low) and 2(t n + i) + 1 (high), masked below 2^bitlen(q); the
 * 64th draw, if ever reached, is kept minus q.  Deterministic per (n, q, seed)
 * and independent of the launch shape. */
int bntt_uniform_dev(void *out, size_t n, const bntt_u128 *q, uint64_t seed, void *stream);
int bntt_uniform(bntt_u128 *out, size_t n, const bntt_u128 *q, uint64_t seed);

/* ------------------------------------------------------------ diagnostics */
/* No counterpart in the reference.  Measured issue rate of one integer VALU
 * instruction on the current device, in lane-instructions per second (8
 * independent chains per lane, 16 workgroups of 256 threads per CU), and the
 * nominal one-per-lane-per-clock rate CUs * 64 * clock.  The NTT's roofline
 * is this rate divided by its VALU instructions per butterfly. */
enum {
    BNTT_PROBE_MAD_U64_U32 = 0,
    BNTT_PROBE_ADD_CO_U32 = 1,
    BNTT_PROBE_CNDMASK_B32 = 2,
    BNTT_PROBE_LSHL_ADD_U64 = 3,
    BNTT_PROBE_AND_B32 = 4
};
int bntt_probe_valu_rate(int op, double *lane_ops_per_s, double *nominal_lane_ops_per_s);

/* -------------------------------------------------------------- host pointers */

int bntt_vvadd(bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, size_t n, const bntt_u128 *m);
int bntt_vvsub(bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, size_t n, const bntt_u128 *m);
int bntt_vvmul(bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, size_t n, const bntt_u128 *m,
               const bntt_u128 *mu);
int bntt_axpy(bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, const bntt_u128 *x, size_t n,
              const bntt_u128 *m, const bntt_u128 *mu);
/* One forward NTT of 2^log2n points; Y and X are host arrays, X is not modified. */
int bntt_ntt_fwd(bntt_u128 *Y, const bntt_u128 *X, const bntt_u128 *tw, int log2n, const bntt_u128 *m,
                 const bntt_u128 *mu);

#ifdef __cplusplus
}
#endif

#endif /* BNTT_H */

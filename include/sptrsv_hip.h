/*
 * sptrsv_hip.h -- C ABI of the MI355X (gfx950) sync-free sparse triangular
 * solver.
 *
 * Plain C: no HIP, torch or C++ types cross this boundary.  Device pointers are
 * `void *`/typed pointers obtained from hipMalloc (or sptrsv_malloc), streams
 * are hipStream_t passed as `void *` (NULL = the library's default stream).
 * Every entry point returns an int status (SPTRSV_OK == 0, negative on error).
 *
 * Reference interfaces replaced (paths relative to harunsasmaz/Benchmark-SpTRSV):
 *   sptrsv_syncfree_hip          <- sptrsv_syncfree_cuda, SpTRSV_cuda/sptrsv_syncfree_cuda.h:409-421
 *   sptrsv_syncfree_hip_f32      <- the same entry built with -D VALUE_TYPE=float
 *                                   (SpTRSV_cuda/Makefile_macos:5, common.h:9-10)
 *   sptrsv_syncfree_hip_csr      <- main.cu:208-210 (host matrix_transposition) + the call above
 *   sptrsv_csr2csc               <- matrix_transposition, SpTRSV_cuda/tranpose.h:6-43
 *   sptrsv_exclusive_scan        <- exclusive_scan, SpTRSV_cuda/utils.h:374-391
 *   sptrsv_analyse               <- sptrsv_syncfree_cuda_analyser, sptrsv_syncfree_cuda.h:8-19
 *   sptrsv_plan_* (solve)        <- the analyse-once / solve-many loop inside
 *                                   sptrsv_syncfree_cuda, sptrsv_syncfree_cuda.h:456-543
 *   sptrsv_equal_*_partition     <- equal_col_partition / equal_nnz_partition, utils.h:7-67
 *   sptrsv_levels                <- findlevel_csc, SpTRSV_cuda/findlevel.h:71-147
 *   sptrsv_plan_levels           <- the same statistics, computed on the device
 *   sptrsv_plan_prepare          <- the analyser step of sptrsv_syncfree_cuda,
 *                                   sptrsv_syncfree_cuda.h:456-483
 */
#ifndef SPTRSV_HIP_H
#define SPTRSV_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants kept at the reference's values (SpTRSV_cuda/common.h:25-33) */
#define SPTRSV_SUBSTITUTION_FORWARD 0
#define SPTRSV_SUBSTITUTION_BACKWARD 1
#define SPTRSV_OPT_WARP_NNZ 1  /* lanes over a column's nonzeros          */
#define SPTRSV_OPT_WARP_RHS 2  /* lanes over right-hand sides             */
#define SPTRSV_OPT_WARP_AUTO 3 /* pick per matrix (autotuner / heuristic) */
#define SPTRSV_COL_PARTITION 0
#define SPTRSV_NNZ_PARTITION 1

/* ---- status codes.  -1 keeps the reference's meaning (m != n,
 * sptrsv_syncfree_cuda.h:423-427); the rest are checks the reference lacks. */
#define SPTRSV_OK 0
#define SPTRSV_ERR_NOT_SQUARE (-1)
#define SPTRSV_ERR_INVALID_ARG (-2)    /* null pointer, n < 0, rhs < 1, bad config */
#define SPTRSV_ERR_MISSING_DIAG (-3)   /* diagonal not first (fwd) / last (bwd)   */
#define SPTRSV_ERR_ZERO_DIAG (-4)
#define SPTRSV_ERR_NOT_TRIANGULAR (-5) /* entry outside the solved triangle       */
#define SPTRSV_ERR_UNSORTED (-6)       /* column rows not strictly ascending      */
#define SPTRSV_ERR_INDEX (-7)          /* row index outside [0, n) / bad colPtr   */
#define SPTRSV_ERR_HIP (-8)            /* a HIP runtime call failed               */
#define SPTRSV_ERR_TIMEOUT (-9)        /* a spin-wait hit its cap (state reset)   */
#define SPTRSV_ERR_NO_DEVICE (-10)
#define SPTRSV_ERR_UNSUPPORTED (-11)   /* e.g. IR on an FP32 plan                 */
#define SPTRSV_ERR_NO_MEMORY (-12)

/* ---- value types */
#define SPTRSV_F64 0
#define SPTRSV_F32 1

/* Human-readable name of a status code (static string). */
const char *sptrsv_status_string(int status);
/* Library version, e.g. "0.1.0"; the gfx target the code objects were built for. */
const char *sptrsv_version(void);

/* ========================================================================
 * Drop-in entry points (host arrays in, host x out).
 * Same parameter list, order and meaning as sptrsv_syncfree_cuda
 * (SpTRSV_cuda/sptrsv_syncfree_cuda.h:409-421):
 *   CSC of the triangular factor (0-based int32, rows ascending per column,
 *   diagonal first for forward / last for backward), m == n, nnzTR entries;
 *   substitution 0/1; rhs >= 1; opt 1/2/3 (matters for rhs > 1);
 *   x, b, x_ref host arrays, row-major [n][rhs]; x is overwritten; *gflops
 *   receives 2*rhs*nnz / mean solve time.  x_ref may be NULL (no validation).
 * Like the reference, each call allocates and frees all device memory, runs
 * the analysis, and times SPTRSV_BENCH_REPEAT solves (default 100, see
 * sptrsv_set_bench_repeat).  Unlike it, the matrix is validated and errors are
 * returned instead of hanging.  Output lines in the reference's format are
 * printed only when sptrsv_set_verbose(1).
 * ======================================================================== */
int sptrsv_syncfree_hip(const int *cscColPtrTR, const int *cscRowIdxTR,
                        const double *cscValTR, const int m, const int n,
                        const int nnzTR, const int substitution, const int rhs,
                        const int opt, double *x, const double *b,
                        const double *x_ref, double *gflops);

int sptrsv_syncfree_hip_f32(const int *cscColPtrTR, const int *cscRowIdxTR,
                            const float *cscValTR, const int m, const int n,
                            const int nnzTR, const int substitution,
                            const int rhs, const int opt, float *x,
                            const float *b, const float *x_ref, double *gflops);

/* CSR of the triangular factor in (the reference transposes on the host at
 * main.cu:208-210 before calling the solver); transposed on the device. */
int sptrsv_syncfree_hip_csr(const int *csrRowPtr, const int *csrColIdx,
                            const double *csrVal, const int m, const int n,
                            const int nnz, const int substitution,
                            const int rhs, const int opt, double *x,
                            const double *b, const double *x_ref,
                            double *gflops);

void sptrsv_set_verbose(int verbose);
void sptrsv_set_bench_repeat(int repeat); /* >= 1; BENCH_REPEAT, common.h:13-14 */

/* ========================================================================
 * Device-memory helpers (thin wrappers over the HIP runtime the library uses,
 * so that C, ctypes, cgo, ... callers need no HIP headers).  Copies and
 * memsets are complete when they return, so any stream may use the data.
 * ======================================================================== */
int sptrsv_device_count(int *count);
int sptrsv_set_device(int device);
int sptrsv_device_name(char *buf, int len);
int sptrsv_device_cu_count(int *cus);
int sptrsv_malloc(void **dptr, size_t bytes);
int sptrsv_free(void *dptr);
int sptrsv_memcpy_htod(void *dst, const void *src, size_t bytes);
int sptrsv_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int sptrsv_memcpy_dtod(void *dst, const void *src, size_t bytes);
int sptrsv_memset(void *dptr, int value, size_t bytes);
int sptrsv_synchronize(void);
int sptrsv_stream_create(void **stream);
int sptrsv_stream_destroy(void *stream);
int sptrsv_stream_synchronize(void *stream);
/* Event timing on a stream: returns milliseconds between two records. */
int sptrsv_event_create(void **event);
int sptrsv_event_destroy(void *event);
int sptrsv_event_record(void *event, void *stream);
int sptrsv_event_elapsed(void *start, void *stop, float *ms);

/* ========================================================================
 * Device building blocks (all arrays device-resident, asynchronous on stream
 * unless noted).
 * ======================================================================== */
/* In-place exclusive scan of len ints (utils.h:374-391). */
int sptrsv_exclusive_scan(int *d_data, int len, void *stream);

/* The same scan with no allocation at all.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Its tile
 * totals live in a fixed static device array (about 0.5 MB per device, held
 * until the process ends), so len is limited to SPTRSV_SCAN_NOALLOC_MAX
 * (larger requests return SPTRSV_ERR_INVALID_ARG).  Asynchronous on the
 * stream.  All calls on one device share that array: calls that overlap in
 * time on one device race. */
#define SPTRSV_SCAN_NOALLOC_MAX (1 << 28)
int sptrsv_exclusive_scan_noalloc(int *d_data, int len, void *stream);

/* Stable CSR -> CSC transpose of an m x n matrix with nnz entries
 * (tranpose.h:6-43): output columns list their rows in input order, so for a
 * CSR with ascending rows every column comes out ascending.  value_type is
 * SPTRSV_F64/SPTRSV_F32; d_val/d_cscVal may both be NULL (pattern only).
 * Also CSC -> CSR when given the CSC arrays (as findlevel.h:94-96 does).
 * Synchronises the stream (the sort tiers are sized on the host). */
int sptrsv_csr2csc(int m, int n, int nnz, const int *d_rowPtr,
                   const int *d_colIdx, const void *d_val, int value_type,
                   int *d_colPtr, int *d_rowIdx, void *d_cscVal, void *stream);

/* in_degree[r] = number of stored entries of row r, diagonal included
 * (sptrsv_syncfree_cuda.h:8-19). */
int sptrsv_analyse(const int *d_cscRowIdx, int m, int nnz, int *d_inDegree,
                   void *stream);

/* Check a CSC triangular factor; returns SPTRSV_OK or the first error code
 * (synchronises the stream). */
int sptrsv_validate(const int *d_colPtr, const int *d_rowIdx, const void *d_val,
                    int value_type, int n, int nnz, int substitution,
                    void *stream);

/* Host-side partitioners (utils.h:7-17, :19-67); counts/displs have nparts
 * entries.  The nnz partitioner zero-fills counts/displs first (the reference
 * leaves that to its callers, naive_unified.h:130-131). */
int sptrsv_equal_col_partition(int nparts, int cols, int *counts, int *displs);
int sptrsv_equal_nnz_partition(int nparts, int ncols, int nnz,
                               const int *colPtr, int *counts, int *displs);

/* Level-set statistics of a host CSC factor (findlevel.h:71-147). */
int sptrsv_levels(const int *cscColPtr, const int *cscRowIdx, int n, int nnz,
                  int *nlevel, int *pmin, int *pavg, int *pmax);

/* ========================================================================
 * Plan API: analyse once, solve many times on device-resident data.
 * ======================================================================== */
typedef struct sptrsv_plan_s *sptrsv_plan_t;

/* Executor configuration.  lanes_per_column is the number of lanes of a
 * 64-wide wavefront that cooperate on one column (64 = one wavefront per
 * column, 1 = one lane per column, 64 columns per wavefront).  rhs_lanes of
 * those lanes run over right-hand sides (1 = lanes over nonzeros,
 * OPT_WARP_NNZ; == lanes_per_column = OPT_WARP_RHS), each lane holding
 * rhs_per_lane right-hand sides, so one workgroup pass covers a tile of
 * rhs_lanes*rhs_per_lane right-hand sides.  window is the number of
 * consecutive columns whose partial sums and counters live in LDS.
 * A zero field means "choose". */
typedef struct {
  int lanes_per_column; /* 1,2,4,...,64 (lane engine: 1) */
  int rhs_lanes;        /* 1 or lanes_per_column (lane engine: 1) */
  int rhs_per_lane;     /* group engine 1..4; lane engine 1, 2, 4 or 8 */
  int block;            /* threads per workgroup: 64..1024, multiple of 64 */
  int window;           /* LDS window, columns (max, for NNZ partition) */
  int partition;        /* SPTRSV_COL_PARTITION / SPTRSV_NNZ_PARTITION */
  int grid;             /* workgroups launched; 0 = resident capacity */
  int engine;           /* SPTRSV_ENGINE_GROUP, _LANE or _ROW */
  int remote;           /* lane engine: SPTRSV_REMOTE_RECEIVER or _DIRECT
                           (0 = choose; other engines: 0) */
} sptrsv_config_t;

/* How the lane engine delivers an entry from an earlier window.  RECEIVER:
 * receiver wavefronts take the slots and add them into the row's LDS sum.
 * DIRECT (one right-hand side per lane only): a row with exactly one such
 * entry takes its slot in its own lane, one handoff less on the critical
 * path; the default where windows hold whole grid periods (planes, rows,
 * chains), receivers otherwise. */
#define SPTRSV_REMOTE_RECEIVER 1
#define SPTRSV_REMOTE_DIRECT 2

/* Executor engines.  GROUP: lane groups of a wavefront per column, counters
 * and partial sums in LDS within a window and in HBM (atomics + fences)
 * across windows -- all rhs modes.  LANE (the default): one lane per column
 * carrying rhs_per_lane right-hand sides of a tile, chains inside a
 * 64-column batch solved by a wavefront scan, and one value-carrying slot
 * per cross-window entry and right-hand side, polled by receiver wavefronts
 * or (remote = SPTRSV_REMOTE_DIRECT) by the row's own lane
 * (lanes_per_column = rhs_lanes = 1; block 128..1024, at most 512 for
 * rhs_per_lane >= 2; LDS ~ window * (rhs_per_lane * value size + 4) bytes,
 * or + 1 with one right-hand side per lane when 4-byte counters would not
 * fit and no row has more than 255 entries). */
#define SPTRSV_ENGINE_GROUP 0
#define SPTRSV_ENGINE_LANE 1
/* ROW (no counterpart in the reference, whose executor only pushes): pulls
 * over the CSR of the factor, lanes_per_column (4, 16, 32 or 64) lanes per
 * row and right-hand side reading each entry's solution, which is its own
 * arrival flag (x is filled with a sentinel first), so no atomics accumulate
 * and a row fed by hundreds of columns costs no serialised updates.  Rows
 * are claimed in DAG-level order through a ticket, window rounds of the
 * workgroup's row groups per claim, so any grid works.  With rhs_lanes =
 * lanes_per_column (8..64) the lanes run over right-hand sides instead
 * (rhs_per_lane 1, 2 or 4 each).  For wide, shallow factors (random
 * power-law rows) and many right-hand sides; partition is unused. */
#define SPTRSV_ENGINE_ROW 2

#define SPTRSV_PLAN_VALIDATE 1 /* check the factor at creation */
#define SPTRSV_PLAN_MIXED 2    /* keep FP32 values + FP64 CSR for sptrsv_plan_solve_ir */

/* d_colPtr/d_rowIdx/d_val are borrowed (must outlive the plan).
 * value_type gives the type of d_val and of b/x in sptrsv_plan_solve. */
int sptrsv_plan_create(sptrsv_plan_t *plan, int n, int nnz,
                       const int *d_colPtr, const int *d_rowIdx,
                       const void *d_val, int value_type, int substitution,
                       int flags, void *stream);

/* Many independent triangular systems solved by one launch: the systems are
 * the diagonal blocks [sys_ptr[s], sys_ptr[s+1]) of one block-diagonal CSC
 * factor (h_sys_ptr is a host array of nsys+1 column offsets, global
 * indices).  Entries coupling two systems are rejected. */
int sptrsv_batch_plan_create(sptrsv_plan_t *plan, int nsys,
                             const int *h_sys_ptr, int n, int nnz,
                             const int *d_colPtr, const int *d_rowIdx,
                             const void *d_val, int value_type,
                             int substitution, int flags, void *stream);

int sptrsv_plan_destroy(sptrsv_plan_t plan);

/* Fix the executor configuration used for `rhs` right-hand sides. */
int sptrsv_plan_set_config(sptrsv_plan_t plan, int rhs,
                           const sptrsv_config_t *cfg);
/* The configuration the plan will use for `rhs` (after defaults/tuning). */
int sptrsv_plan_get_config(sptrsv_plan_t plan, int rhs, sptrsv_config_t *cfg);

/* x = L^{-1} b for rhs right-hand sides, row-major [n][rhs], in the plan's
 * value type.  One kernel launch, no memsets (the executor restores its own
 * counters), so it may be captured into a hipGraph. */
int sptrsv_plan_solve(sptrsv_plan_t plan, const void *d_b, void *d_x, int rhs,
                      void *stream);

/* Run the per-configuration analysis that the first solve with `rhs`
 * right-hand sides would run (mode 0 = plain solve, 1 = the FP32 solve of IR):
 * window set, lane targets, batches and their levels, receive lists, the CSR
 * copy of the row engine.  Lets a caller keep it out of timed or captured
 * regions (the reference times its analyser apart from the solves,
 * sptrsv_syncfree_cuda.h:456-483).  Synchronous. */
int sptrsv_plan_prepare(sptrsv_plan_t plan, int rhs, int mode, void *stream);

/* DAG depth and level widths of the plan's factor, computed on the device
 * (the statistics of findlevel_csc, findlevel.h:71-147: number of levels,
 * min / average / max rows per level).  Cached; synchronous. */
int sptrsv_plan_levels(sptrsv_plan_t plan, int *nlevel, int *pmin, int *pavg,
                       int *pmax, void *stream);

/* FP32 solves + FP64 residual/update iterative refinement (plan created with
 * SPTRSV_PLAN_MIXED on FP64 data).  iters refinement steps after the first
 * FP32 solve.  d_b, d_x are FP64. */
int sptrsv_plan_solve_ir(sptrsv_plan_t plan, const double *d_b, double *d_x,
                         int rhs, int iters, void *stream);

/* Autotune: time candidate configurations on the plan's matrix and keep the
 * fastest for `rhs` (mode 0 = plain solve, 1 = FP32 solve used inside IR).
 * Results are cached per (n, nnz, rhs, type, depth bucket) in the file named
 * by $SPTRSV_TUNE_CACHE when set.  Synchronous. */
int sptrsv_plan_autotune(sptrsv_plan_t plan, int rhs, int mode,
                         sptrsv_config_t *best, float *best_ms);

/* Capture solve (mode 0) or IR solve (mode 1, `iters` steps) into a hipGraph
 * on the given buffers; replay with sptrsv_graph_launch. */
typedef struct sptrsv_graph_s *sptrsv_graph_t;
int sptrsv_graph_create(sptrsv_graph_t *graph, sptrsv_plan_t plan,
                        const void *d_b, void *d_x, int rhs, int mode,
                        int iters, void *stream);
int sptrsv_graph_launch(sptrsv_graph_t graph, void *stream);
int sptrsv_graph_destroy(sptrsv_graph_t graph);

/* Wait for the stream and report a spin-wait timeout of any earlier solve
 * (SPTRSV_ERR_TIMEOUT, after which the plan's scratch state is reset). */
int sptrsv_plan_check(sptrsv_plan_t plan, void *stream);
/* Clear the plan's scratch state (counters, partial sums, tickets). */
int sptrsv_plan_reset(sptrsv_plan_t plan, void *stream);
/* Longest one spin-wait may last, in microseconds of wall clock, before the
 * solve gives up and reports SPTRSV_ERR_TIMEOUT (0 = default, 30 s). */
int sptrsv_plan_set_spin_cap(sptrsv_plan_t plan, long long cap_us);

typedef struct {
  int n, nnz, substitution, value_type, nsys;
  int levels;                /* DAG depth (0 until sptrsv_plan_levels or
                                sptrsv_plan_autotune computed it)       */
  long long scratch_bytes;   /* device scratch held by the plan        */
  double analysis_ms;        /* device analysis time at creation       */
  double prepare_ms;         /* analysis after creation so far: the
                                per-configuration analyses (first solve
                                of each configuration, or
                                sptrsv_plan_prepare) and the level
                                analysis, host wall clock               */
} sptrsv_plan_info_t;
int sptrsv_plan_info(sptrsv_plan_t plan, sptrsv_plan_info_t *info);

#ifdef __cplusplus
}
#endif
#endif /* SPTRSV_HIP_H */

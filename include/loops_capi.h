/*
 * loops_capi.h -- C ABI of the MI355X (gfx950) load-balanced CSR SpMV/SpMM
 * library, libloops_mi355x.so.
 *
 * Plain pointers and sizes only.  Every compute entry point is asynchronous
 * on the hipStream_t it is given and performs no device allocation when the
 * caller passes a workspace, so it can be captured into a hipGraph.  Status
 * codes are returned, never thrown.  The C++ drop-in headers under
 * include/loops/ restore the reference's synchronous, exception-throwing
 * semantics on top of this ABI.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   loops_spmv(MERGE_PATH_FLAT)   algorithms/spmv/merge_path_flat.cuh:96-139
 *   loops_spmv(WORK_ORIENTED)     algorithms/spmv/work_oriented.cuh:91-113
 *   loops_spmv(THREAD_MAPPED)     algorithms/spmv/thread_mapped.cuh:68-91
 *   loops_spmv(GROUP_MAPPED)      algorithms/spmv/group_mapped.cuh:73-105
 *   loops_spmv(ORIGINAL)          algorithms/spmv/original.cuh:59-76
 *   loops_spmm(...)               algorithms/spmm/thread_mapped.cuh:67-90
 *   loops_merge_path_partition    schedule/merge_path_flat.hxx:108-138 and
 *                                 the preprocess_t constructor, :172-206
 *   loops_schedule_t              schedule.hxx:23-29 (same numbering)
 *
 * Contract differences from the reference (all deliberate):
 *   - y (or C) is overwritten, never accumulated into, for every schedule.
 *     The reference's atomic schedules require y == 0 on entry and add into
 *     it; with y == 0 both give the same result.
 *   - Results are deterministic: no schedule uses floating-point atomics.
 *     Rows split across work units are combined by a fixed-order fix-up.
 *   - Index and offset types are int32, as in every reference example
 *     (examples/spmv/merge_path.cu:18-20); rows + nnzs must be at most
 *     LOOPS_MAX_ROWS_PLUS_NNZS, 2^31 - 1 less 2^20: the kernels index
 *     entries in int32 up to a tile past the last one.
 */
#ifndef LOOPS_CAPI_H_
#define LOOPS_CAPI_H_

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LOOPS_CAPI_VERSION 1

/* Largest rows + nnzs of a matrix (int32 offsets and entry positions, with
 * headroom for the entries a merge tile reads past its end). */
#define LOOPS_MAX_ROWS_PLUS_NNZS (2147483647LL - (1LL << 20))

/* error.hxx:15-45 (exception_t / throw_if_exception) become status codes. */
typedef enum loops_status_t {
  LOOPS_SUCCESS = 0,
  LOOPS_ERROR_INVALID_VALUE = 1,   /* null pointer, negative size, bad enum */
  LOOPS_ERROR_OVERFLOW = 2,        /* rows + nnzs > LOOPS_MAX_ROWS_PLUS_NNZS */
  LOOPS_ERROR_WORKSPACE = 3,       /* workspace missing or too small        */
  LOOPS_ERROR_NOT_SUPPORTED = 4,   /* schedule / tuning not available       */
  LOOPS_ERROR_HIP = 5,             /* a HIP runtime call failed             */
  LOOPS_ERROR_ALLOCATION = 6       /* handle-owned workspace growth failed  */
} loops_status_t;

/* schedule.hxx:23-29 keeps its numbering; ORIGINAL (algorithms/spmv/
 * original.cuh, no schedule object) is appended. */
typedef enum loops_schedule_t {
  LOOPS_SCHEDULE_MERGE_PATH_FLAT = 0,
  LOOPS_SCHEDULE_WORK_ORIENTED = 1,
  LOOPS_SCHEDULE_THREAD_MAPPED = 2,
  LOOPS_SCHEDULE_GROUP_MAPPED = 3,
  LOOPS_SCHEDULE_BUCKETING = 4, /* declared by the reference, never built */
  LOOPS_SCHEDULE_ORIGINAL = 5
} loops_schedule_t;

/* Value storage of A.  For every entry point but loops_spmv_typed:
 * F32 / BF16 / F16 storage, fp32 accumulation, fp32 x, y, B and C.
 * loops_spmv_typed takes F64 or I32: A's values, x and y all of that type
 * and every product and sum in it (I32: 32-bit wrap-around). */
typedef enum loops_value_type_t {
  LOOPS_VALUE_F32 = 0,
  LOOPS_VALUE_BF16 = 1,
  LOOPS_VALUE_F16 = 2,
  LOOPS_VALUE_F64 = 3,
  LOOPS_VALUE_I32 = 4
} loops_value_type_t;

/* container/csr.hxx:36-43 as a non-owning view (device pointers). */
typedef struct loops_csr_view_t {
  int64_t rows;
  int64_t cols;
  int64_t nnzs;
  const int32_t* offsets; /* rows + 1 entries, offsets[0] == 0           */
  const int32_t* indices; /* nnzs column indices                          */
  const void* values;     /* nnzs values of value_type                    */
  loops_value_type_t value_type;
} loops_csr_view_t;

/* container/coordinate.hxx:14-18 with coord_t = coordinate_t<unsigned>
 * (schedule/merge_path_flat.hxx:30): x = row, y = nonzero. */
typedef struct loops_coord_t {
  uint32_t x;
  uint32_t y;
} loops_coord_t;

/* Launch tuning; a zero field selects the schedule's default.
 * Defaults follow the reference's host entries:
 *   merge_path_flat: 128 threads x 5 items (merge_path_flat.cuh:102-103)
 *   work_oriented:   128 threads, grid 2 x CU count (work_oriented.cuh:97-104)
 *   group_mapped:    block_mapped<128> (group_mapped.cuh:39,79)
 *   thread_mapped / original: 128 threads, ceil(rows/128) blocks. */
typedef struct loops_params_t {
  int32_t threads_per_block;
  int32_t items_per_thread;  /* merge_path_flat only                       */
  int32_t threads_per_tile;  /* group_mapped only: 64 (wave) or TPB        */
  int32_t grid_size;         /* work_oriented only                         */
  /* merge_path_flat only: precomputed tile coordinates from
   * loops_merge_path_partition with items_per_tile = TPB * IPT (the
   * reference's preprocess_t meta).  NULL: each tile searches itself. */
  const loops_coord_t* tile_coordinates;
  /* Optional events recorded around the main kernel, the analogue of the
   * util::timer_t returned by merge_path_flat (merge_path_flat.cuh:121-138). */
  hipEvent_t start_event;
  hipEvent_t stop_event;
  uint32_t flags; /* reserved, must be 0 */
} loops_params_t;

typedef struct loops_handle* loops_handle_t;

/* Handle: caches device properties (CU count) and an optional
 * library-owned workspace used when a call passes workspace == NULL. */
loops_status_t loops_create(loops_handle_t* handle);
loops_status_t loops_destroy(loops_handle_t handle);
const char* loops_status_string(loops_status_t status);
const char* loops_schedule_name(loops_schedule_t schedule);
/* The hipError_t of the last LOOPS_ERROR_HIP returned on this thread. */
int loops_last_hip_error(void);

/* Device properties as the reference's util/device.hxx:23-52 reads them. */
loops_status_t loops_device_cu_count(loops_handle_t handle, int32_t* cu_count);

/* Bytes of device workspace loops_spmv needs for this schedule, matrix
 * shape and tuning (0 for thread_mapped, original and group_mapped). */
loops_status_t loops_spmv_workspace_size(loops_handle_t handle,
                                         loops_schedule_t schedule,
                                         const loops_csr_view_t* A,
                                         const loops_params_t* params,
                                         size_t* bytes);

/* y = A * x (y overwritten).  workspace may be NULL, in which case the
 * handle's own workspace is used (grown with hipMalloc outside any graph
 * capture); a caller workspace needs no initialisation.  merge_path_flat
 * and work_oriented finish rows that cross work units in a second small
 * kernel.  Calls that overlap in time must use different workspaces.
 * Asynchronous on stream. */
loops_status_t loops_spmv(loops_handle_t handle, loops_schedule_t schedule,
                          const loops_csr_view_t* A, const float* x, float* y,
                          const loops_params_t* params, void* workspace,
                          size_t workspace_bytes, hipStream_t stream);

/* y = A * x as loops_spmv, without allocating anything (no hipMalloc, no
 * handle workspace, no caller workspace).  No counterpart in the
 * reference: for callers that cannot allocate device memory -- code under
 * hipGraph capture, tight serving loops, harnesses that forbid allocation.
 * The fix-up slots of merge_path_flat and work_oriented live in a static
 * device array of LOOPS_NOALLOC_SCRATCH_BYTES owned by the library, one per
 * GPU, held until the process ends; a call whose schedule and tuning need
 * more (loops_spmv_workspace_size) returns LOOPS_ERROR_WORKSPACE.  The
 * default merge_path_flat tiles (256 x 8) fit every matrix the int32 ABI
 * accepts.  All calls on one GPU share its array: calls that overlap in
 * time on one GPU (two streams, or graph replays on several streams) race
 * and give wrong results; serialise them.  The array keeps the last call's
 * contents, which no call needs cleared.  Asynchronous on stream. */
#define LOOPS_NOALLOC_SCRATCH_BYTES (16u << 20)
loops_status_t loops_spmv_noalloc(loops_handle_t handle, loops_schedule_t schedule,
                                  const loops_csr_view_t* A, const float* x, float* y,
                                  const loops_params_t* params, hipStream_t stream);

/* y = A * x with A->value_type LOOPS_VALUE_F64 (x, y double) or
 * LOOPS_VALUE_I32 (x, y int32): the reference's host entries for
 * type_t = double and int (algorithms/spmv/<schedule>.cuh; its benchmark's value
 * types, benchmarks/spmv/work_oriented.cu:53).  Same schedules, partitions
 * and contract as loops_spmv (y overwritten, fixed-order sums); default
 * tiles follow the reference per value size: merge_path_flat 64 x 3 for
 * double, 128 x 5 for int32 (merge_path_flat.cuh:102-103); also 256 x 4
 * and 256 x 8.  work_oriented: 64, 128 or 256 threads; group_mapped: the
 * loops_spmv tunings (threads_per_tile 32 .. TPB).  The workspace rule is
 * loops_spmv's (loops_spmv_typed_workspace_size). */
loops_status_t loops_spmv_typed_workspace_size(loops_handle_t handle,
                                               loops_schedule_t schedule,
                                               const loops_csr_view_t* A,
                                               const loops_params_t* params,
                                               size_t* bytes);
loops_status_t loops_spmv_typed(loops_handle_t handle, loops_schedule_t schedule,
                                const loops_csr_view_t* A, const void* x, void* y,
                                const loops_params_t* params, void* workspace,
                                size_t workspace_bytes, hipStream_t stream);

/* Number of merge tiles, ceil((rows + nnzs) / items_per_tile)
 * (merge_path_flat.cuh:116-117). */
loops_status_t loops_merge_path_num_tiles(int64_t rows, int64_t nnzs,
                                          int64_t items_per_tile,
                                          int64_t* num_tiles);

/* coords[t] = merge-path coordinate at diagonal min(t * items_per_tile,
 * rows + nnzs) for t in [0, num_tiles] (num_tiles + 1 entries): the
 * reference's generate_search_coordinates (merge_path_flat.hxx:108-138),
 * bit-exact with its search (util/search.hxx:34-60). */
loops_status_t loops_merge_path_partition(const int32_t* offsets, int64_t rows,
                                          int64_t nnzs, int64_t items_per_tile,
                                          loops_coord_t* coords,
                                          hipStream_t stream);

/* work_oriented partition: for each of num_threads threads its start and end
 * coordinate (work_oriented.hxx:150-167), written as 2 coords per thread. */
loops_status_t loops_work_oriented_partition(const int32_t* offsets,
                                             int64_t rows, int64_t nnzs,
                                             int64_t num_threads,
                                             loops_coord_t* coords,
                                             hipStream_t stream);

/* ---- SpMV plan: preprocessed once per matrix, executed many times ------
 * Layout PANELS: merge_path_flat SpMV with the tile coordinates computed
 * once (the reference's preprocess_t, merge_path_flat.hxx:150-230) and,
 * with panels > 1, A copied into `panels` column panels whose tiles run on
 * separate XCDs so each XCD's L2 holds only its slice of x.  Deterministic.
 * Layout BLOCKS: A copied into row blocks (at most 15872 rows, about equal
 * nonzeros) x column segments (4 up to 2^22 columns, else a multiple of 8),
 * each (block, segment) stream bucketed by 64-column chunks and cut into
 * rounds of at most 4096 entries sorted by (row, column); a workgroup
 * accumulates its block's y in LDS, reading x in column order.
 * Deterministic (no atomics).
 * Neither copied layout has a counterpart in the reference.  Creation
 * allocates and synchronises `stream`; execution allocates nothing and can
 * be captured into a hipGraph.  With PANELS and panels == 1 the plan reads
 * A in place, so A must outlive it unchanged; otherwise A is no longer
 * needed after creation. */
typedef struct loops_spmv_plan* loops_spmv_plan_t;

typedef enum loops_plan_layout_t {
  LOOPS_PLAN_LAYOUT_AUTO = 0,   /* BLOCKS when cols >= 2^12, nnzs >= 2^19
                                   and no PANELS parameter (panels,
                                   threads_per_block, items_per_thread) is
                                   set, else PANELS.  The rule weighs
                                   execution time only: a BLOCKS plan costs
                                   1-28 ms to build (several stream
                                   synchronisations) and about 1.45x the
                                   CSR's device memory, and repays the build
                                   against per-call loops_spmv after roughly
                                   115 executions on a 2M-row, 40M-nonzero
                                   matrix (against a PANELS plan: 250-560).
                                   For a matrix used a few times call
                                   loops_spmv or ask for PANELS.            */
  LOOPS_PLAN_LAYOUT_PANELS = 1,
  LOOPS_PLAN_LAYOUT_BLOCKS = 2
} loops_plan_layout_t;

typedef struct loops_plan_params_t {
  int32_t threads_per_block; /* PANELS: 0 = 512                             */
  int32_t items_per_thread;  /* PANELS: 0 = 4                               */
  int32_t panels;            /* PANELS: 1..8; 0 = one per 2 MiB of x, <= 8  */
  int32_t layout;            /* loops_plan_layout_t                         */
  int32_t for_batches;       /* BLOCKS (AUTO then means BLOCKS): nonzero cuts
                                the plan for batch items (loops_batch_item_t
                                .plan) instead of for a launch of its own:
                                one column segment when cols <= 2^20, and
                                as few row blocks as the LDS holds, so each
                                workgroup streams many full rounds; with one
                                segment and no row cut into pieces the
                                streams write y themselves (no partial sums,
                                no combine).  Alone such a plan underfills
                                the GPU.  Not valid with PANELS or a PANELS
                                parameter.                                  */
} loops_plan_params_t;

/* A plan reorders A's entries by column range and so needs a well-formed
 * CSR (loops_csr_check): it checks A first (one pass over it) and returns
 * LOOPS_ERROR_INVALID_VALUE for a matrix that is not, before building
 * anything.  Matrices with rows + nnzs above LOOPS_MAX_PLAN_ROWS_PLUS_NNZS
 * get LOOPS_ERROR_NOT_SUPPORTED before any device work: plans are
 * validated on matrices of up to 42M rows + nonzeros, and building plans of
 * about 2^31 faulted the GPU with no cause found; loops_spmv (every
 * schedule) is validated up to LOOPS_MAX_ROWS_PLUS_NNZS.  Synchronises the
 * stream. */
#define LOOPS_MAX_PLAN_ROWS_PLUS_NNZS (1LL << 26)
loops_status_t loops_spmv_plan_create(loops_handle_t handle,
                                      const loops_csr_view_t* A,
                                      const loops_plan_params_t* params,
                                      hipStream_t stream,
                                      loops_spmv_plan_t* plan);

/* Structure of a CSR matrix, for callers that want to validate one before
 * use (the reference's containers assume it; its CSR from COO sorts by
 * row and column, csr.hxx:86-94).  *flags = 0 when A is well formed, else
 * an OR of: LOOPS_CSR_BAD_OFFSETS (offsets[0] != 0, offsets[rows] != nnzs,
 * or decreasing), LOOPS_CSR_COLUMN_RANGE (a column id outside [0, cols)),
 * LOOPS_CSR_UNSORTED (column ids descending inside a row; repeats are
 * allowed).  The per-call entry points read columns in row order and need
 * only offsets and column range to be right; plans need all three.  Reads
 * offsets[0 .. rows] and indices[0 .. nnzs) only.  Synchronises the
 * stream; allocates 4 bytes of device memory. */
#define LOOPS_CSR_BAD_OFFSETS 1u
#define LOOPS_CSR_COLUMN_RANGE 2u
#define LOOPS_CSR_UNSORTED 4u
loops_status_t loops_csr_check(const loops_csr_view_t* A, uint32_t* flags,
                               hipStream_t stream);
/* y = A * x (y overwritten); events (optional) around the main kernel.
 * Executions of one plan must not overlap in time (two streams, or graph
 * replays on several streams): they share the plan's partial sums (BLOCKS)
 * or fix-up slots and partial y (PANELS).  Serialise them, or build one
 * plan per concurrent user. */
loops_status_t loops_spmv_plan_execute(loops_spmv_plan_t plan, const float* x,
                                       float* y, hipEvent_t start_event,
                                       hipEvent_t stop_event,
                                       hipStream_t stream);
/* panels: PANELS: the column panels; BLOCKS: the column segments.
 * layout (may be null): the layout chosen. */
loops_status_t loops_spmv_plan_info(loops_spmv_plan_t plan, int32_t* panels,
                                    int64_t* device_bytes);
loops_status_t loops_spmv_plan_layout(loops_spmv_plan_t plan, int32_t* layout);
/* BLOCKS plans (0 for PANELS; any pointer may be null): blocks, the row
 * blocks (each runs as one workgroup per column segment); compressed, 1
 * when a row holds an LDS slot and a partial only in the segments where it
 * has entries, 0 when it does in every segment of its block; split_rows,
 * the rows cut into pieces. */
loops_status_t loops_spmv_plan_blocks(loops_spmv_plan_t plan, int32_t* blocks,
                                      int32_t* compressed, int32_t* split_rows);
loops_status_t loops_spmv_plan_destroy(loops_spmv_plan_t plan);

/* ---- Batches: many independent products in one persistent launch -------
 * No counterpart in the reference.  Each item is y = A x (k == 1) or
 * Y = A X (k > 1, X cols x k and Y rows x k row-major), computed exactly as
 * it would be alone with the same schedule and tile size.  Items may use:
 *   k == 1, LOOPS_SCHEDULE_MERGE_PATH_FLAT: merge tiles of 256 x
 *           items_per_thread steps (4, 8 or 16; 0 = 8);
 *   k == 1, LOOPS_SCHEDULE_THREAD_MAPPED: one row per lane;
 *   k == 1, LOOPS_SCHEDULE_GROUP_MAPPED: items_per_thread lanes per row
 *           (2, 4, ..., 64; 0 = the power of two at or above the mean row
 *           length), reduced in a fixed order;
 *   k >  1, LOOPS_SCHEDULE_MERGE_PATH_FLAT: merge-path units of lane
 *           groups spanning the k columns, items_per_thread steps per group
 *           (4, 8, 16, or 32 for groups of 8 lanes or more; 0 = the
 *           loops_spmm default);
 *   k == 1, plan != NULL: the row-block streams of a BLOCKS plan of A
 *           (loops_spmv_plan_create), bit for bit what
 *           loops_spmv_plan_execute gives; schedule, items_per_thread and
 *           A's pointers are ignored, A's rows and cols must be the plan's.
 *           The launch uses the plan's partial sums: a plan may appear in
 *           one item of a batch only, must outlive the batch's launches,
 *           and those launches must not overlap other executions of it.
 * Values fp32, bf16 or fp16; indices / values 16-byte aligned (8 for 16-bit
 * values).
 * Steps: loops_batch_query (sizes) -> loops_batch_prepare (host only: fills
 * a table in host memory that points into the device workspace) -> copy
 * the table to device memory -> loops_batch_launch (kernels only; no
 * allocation; capturable in a hipGraph; a launch may be replayed as long
 * as the items' pointers stay valid).  Overlapping launches need their own
 * table and workspace. */
typedef struct loops_batch_item_t {
  loops_csr_view_t A;
  int64_t k;
  const float* x;
  int64_t ldx;
  float* y;
  int64_t ldy;
  loops_schedule_t schedule;
  int32_t items_per_thread;
  loops_spmv_plan_t plan; /* NULL, or a BLOCKS plan of A (k == 1) */
} loops_batch_item_t;

typedef struct loops_batch_info_t {
  int64_t table_bytes;     /* host table = device table size            */
  int64_t workspace_bytes; /* device workspace (no initialisation)      */
  int64_t work_items;      /* workgroup-sized pieces                    */
  int64_t spmv_tiles;      /* merge tiles of the SpMV items             */
  int64_t spmm_units;      /* merge-path units of the SpMM items        */
  int64_t spmm_work_items; /* pieces of the SpMM items (of work_items)  */
  int64_t coordinates;     /* merge-tile coordinates written per launch */
  int64_t max_items_per_thread; /* largest merge tile of the SpMV items   */
  int64_t spmm_variant_mask; /* bit v: SpMM pieces of tile variant v exist;
                                one persistent kernel is launched per bit  */
  int64_t spmv_class_mask;   /* bit c: SpMV pieces of class c exist (0: merge
                                tiles of 4 steps, thread / group mapped; 1:
                                8 steps; 2: 16 steps); one kernel per bit   */
  int64_t blocks_work_items; /* (block, segment) streams of the plan items
                                (of work_items), those of plans of several
                                segments padded per item to a multiple of
                                8                                          */
  int64_t blocks_row_pieces; /* 256-row pieces of the plan items' combine  */
} loops_batch_info_t;

loops_status_t loops_batch_query(const loops_batch_item_t* items, int64_t n,
                                 loops_batch_info_t* info);
loops_status_t loops_batch_prepare(const loops_batch_item_t* items, int64_t n,
                                   void* device_workspace,
                                   size_t workspace_bytes, void* host_table,
                                   size_t table_bytes,
                                   loops_batch_info_t* info);
/* grid: persistent workgroups per kernel (0: as many as are resident at
 * once on the current device).  loops_batch_prepare also queries the
 * current device's occupancy (once per device). */
loops_status_t loops_batch_launch(const void* device_table,
                                  const loops_batch_info_t* info, int32_t grid,
                                  hipStream_t stream);

/* ---- Serving: continuous batching over HBM-resident matrices ----------
 * No counterpart in the reference.  A server owns a device arena for
 * request outputs, `slots` streams each with a batch table and workspace of
 * scratch_bytes, and a thread that admits requests in arrival order once
 * their output fits in the arena and launches everything admitted on a
 * free slot, so slots overlap on the device: requests of at least 2^20
 * work items ((rows + nnzs) x k) alone through loops_spmv / loops_spmm
 * (their launch overhead is small beside them), the others as one batch
 * (loops_batch_*).
 * Mode PLAIN launches each request alone through loops_spmv / loops_spmm on
 * one stream instead (the per-call baseline, same admission and clock).
 * Results are bit-identical whatever else is in flight. */
typedef enum loops_server_mode_t {
  LOOPS_SERVER_BATCHED = 0,
  LOOPS_SERVER_PLAIN = 1
} loops_server_mode_t;

/* PLAIN mode only: launch every request through the reference's own entry
 * points -- spmv::merge_path_flat (default tiles) for k == 1 and
 * spmm::thread_mapped (algorithms/spmm/thread_mapped.cuh:67-71, its only
 * SpMM) for k > 1 -- instead of the tuned merge-path kernels. */
#define LOOPS_SERVER_FLAG_REFERENCE_SCHEDULES 1u
/* BATCHED mode only: every admitted request in the batch, none alone. */
#define LOOPS_SERVER_FLAG_BATCH_ALL 2u

typedef struct loops_server_config_t {
  int64_t arena_bytes;        /* outputs of admitted requests; 0 = 1 GiB        */
  int64_t scratch_bytes;      /* per slot: batch table + workspace; 0 = 64 MiB  */
  int32_t slots;              /* batches in flight (BATCHED, 0 = 3) or requests
                                 in flight (PLAIN, 0 = 64)                      */
  int32_t max_batch_requests; /* 0 = 512                                        */
  int32_t mode;               /* loops_server_mode_t                            */
  uint32_t flags;             /* LOOPS_SERVER_FLAG_*, or 0                       */
} loops_server_config_t;

/* A resident matrix (device CSR, owned by the caller, kept valid while the
 * server lives) and its tunings: the schedule / items per thread of its
 * SpMV requests (loops_batch_item_t rules) and the merge steps per SpMM unit
 * by k: [0] k = 2, [1] k <= 4, [2] <= 8, [3] <= 16, [4] <= 32, [5] <= 64,
 * [6] <= 128, [7] larger (0 = 256). */
typedef struct loops_server_matrix_t {
  loops_csr_view_t A;
  int32_t spmv_schedule;
  int32_t spmv_items_per_thread;
  int32_t spmm_items_per_unit[8];
} loops_server_matrix_t;

typedef struct loops_server_stats_t {
  int64_t submitted, completed, batches, batched_requests;
  int64_t arena_bytes, arena_high_water;
} loops_server_stats_t;

typedef struct loops_server* loops_server_t;

loops_status_t loops_server_create(const loops_server_config_t* config,
                                   loops_server_t* server);
loops_status_t loops_server_destroy(loops_server_t server);
loops_status_t loops_server_add_matrix(loops_server_t server,
                                       const loops_server_matrix_t* matrix,
                                       int32_t* id);
/* Request Y = A X with X (cols x k, leading dimension ldx >= k; a plain
 * vector when k == 1) in device memory, valid until the request completes.
 * The server runs on streams of its own, so it orders its read of X after
 * the work that produced X only through `ready`: an event recorded after
 * that work on the producer's stream (the server's stream waits for it
 * before the request's kernels), or NULL when X is complete on the device
 * at submit time.  The event must not be recorded again or destroyed until
 * the request completes.  The output (rows x k, leading dimension k) is
 * allocated by the server. */
loops_status_t loops_server_submit(loops_server_t server, int32_t matrix, int64_t k,
                                   const float* x, int64_t ldx, hipEvent_t ready,
                                   int64_t* ticket);
/* done = 1 once complete; y then points at the output until release.
 * Times are seconds on the server's clock (loops_server_now). */
loops_status_t loops_server_test(loops_server_t server, int64_t ticket, int32_t* done,
                                 float** y, double* submit_s, double* done_s);
loops_status_t loops_server_wait(loops_server_t server, int64_t ticket, float** y,
                                 double* submit_s, double* done_s);
/* Frees a completed request's output (arena memory for admission): the
 * next admitted request may write there at once, so every read of it must
 * be complete first (a device-to-device hipMemcpy can return before it
 * is: synchronise its stream). */
loops_status_t loops_server_release(loops_server_t server, int64_t ticket);
loops_status_t loops_server_now(loops_server_t server, double* now_s);
loops_status_t loops_server_stats(loops_server_t server, loops_server_stats_t* stats);

/* Open-loop load: request i is submitted arrival_s[i] seconds (non-decreasing)
 * after the call starts (X complete on the device beforehand: no ready
 * event) and released once complete; latency_s[i] =
 * completion - scheduled arrival; makespan_s = last completion - first
 * arrival. */
typedef struct loops_server_request_t {
  int32_t matrix;
  int64_t k;
  const float* x;
  int64_t ldx;
} loops_server_request_t;
loops_status_t loops_server_replay(loops_server_t server,
                                   const loops_server_request_t* requests, int64_t n,
                                   const double* arrival_s, double* latency_s,
                                   double* makespan_s);

/* SpMM, C = A * B with B (cols x k) and C (rows x k) dense row-major fp32
 * (container/matrix.cuh:27-35) with leading dimensions ldb, ldc >= k.
 * Schedules: THREAD_MAPPED (algorithms/spmm/thread_mapped.cuh:28-53),
 * GROUP_MAPPED and MERGE_PATH_FLAT (new in this library; the reference has
 * only thread_mapped SpMM).  MERGE_PATH_FLAT: more than 64 columns run as
 * launches of 64 columns each, in order on the stream (65 columns left go
 * as 60 + 5).  Lanes take 4 columns (16-byte B loads) from 8 columns up
 * when B's rows are 16-byte aligned (B and ldb), else 1; a range is the
 * fewest lanes (at least 2, a power of two) that span a launch's columns;
 * C's alignment does not matter.  params->items_per_thread merge steps per
 * lane range in each launch: 4, 8, 16, or 32 when every launch's ranges
 * hold 8 lanes or more (else LOOPS_ERROR_NOT_SUPPORTED); 0 = 32 for ranges
 * of 16 lanes or more, else 16.  The workspace size covers the widest
 * launch. */
loops_status_t loops_spmm_workspace_size(loops_handle_t handle,
                                         loops_schedule_t schedule,
                                         const loops_csr_view_t* A, int64_t k,
                                         const loops_params_t* params,
                                         size_t* bytes);
loops_status_t loops_spmm(loops_handle_t handle, loops_schedule_t schedule,
                          const loops_csr_view_t* A, int64_t k, const float* B,
                          int64_t ldb, float* C, int64_t ldc,
                          const loops_params_t* params, void* workspace,
                          size_t workspace_bytes, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* LOOPS_CAPI_H_ */

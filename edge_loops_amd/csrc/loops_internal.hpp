// Internal launch interface between the C ABI (capi.hip) and the kernel
// translation units.  Host-only declarations; every function launches on the
// given stream and returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "loops_capi.h"

namespace loops_mi355x {

struct csr_args {
  const int* offsets;
  const int* indices;
  const void* values;
  loops_value_type_t vtype;
  int rows;
  int cols;
  int nnz;
  bool aligned;  // indices / values allow 16-byte (8-byte for 16-bit) loads
};

// Slots of the row fix-up: per work unit (merge tile or work_oriented
// thread) the row continuing past its end and two partial sums (caller
// workspace, no initialisation needed).  seg_end[u]: for the unit where its
// open row starts, the unit holding that row's end (so the fix-up needs no
// offsets lookup); -1 otherwise.
struct fixup_ws {
  int* carry_row;
  int* seg_end;
  float* carry_val;
  float* head_val;
};

inline size_t fixup_ws_bytes(long long units) {
  size_t b = (size_t)units * (2 * sizeof(int) + 2 * sizeof(float));
  return (b + 255) & ~size_t(255);
}
fixup_ws carve_fixup_ws(void* ws, long long units);

// ---- SpMV with double or int32 values, x and y (spmv_typed.hip) ----------
struct typed_launch {
  loops_schedule_t schedule;
  int tpb, ipt, tpt, grid;
  const loops_coord_t* coords;  // merge_path_flat tile coordinates, or null
  void* ws;                     // fix-up slots (merge_path_flat, work_oriented)
  hipEvent_t ev0, ev1;
};
bool typed_spmv_supported(loops_schedule_t schedule, int tpb, int ipt, int tpt);
size_t typed_fixup_bytes(long long units, int value_bytes);
hipError_t launch_typed_spmv(const typed_launch& L, const csr_args& A, const void* x, void* y,
                             hipStream_t s);

// ---- SpMV ----------------------------------------------------------------
bool mp_spmv_supported(int tpb, int ipt);
hipError_t launch_mp_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int ipt, const loops_coord_t* coords,
                          const fixup_ws& ws, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s);
hipError_t launch_mp_partition(const int* offsets, int rows, int nnz,
                               long long items_per_tile, long long num_coords,
                               loops_coord_t* coords, hipStream_t s);
hipError_t launch_wo_partition(const int* offsets, int rows, int nnz,
                               long long num_threads, loops_coord_t* coords,
                               hipStream_t s);
hipError_t launch_fixup(const int* offsets, int rows, long long unit_items,
                        long long num_units, const fixup_ws& ws, float* y,
                        hipStream_t s);

bool wo_spmv_supported(int tpb);
hipError_t launch_wo_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int grid, const fixup_ws& ws, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s);

bool gm_spmv_supported(int tpb, int tpt);
hipError_t launch_gm_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int tpt, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s);

hipError_t launch_tm_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          bool original, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s);

// ---- column-panel plan ---------------------------------------------------
constexpr int kMaxPanels = 8;
struct panel_bounds_t {
  int c[kMaxPanels + 1];  // panel p holds columns [c[p], c[p+1])
};
struct panel_ptrs_t {
  int* p[kMaxPanels];
};
struct panel_vptrs_t {
  void* p[kMaxPanels];
};
// Everything one launch needs about every panel (kernel argument).
struct panel_set_t {
  const int* offsets[kMaxPanels];
  const int* indices[kMaxPanels];
  const void* values[kMaxPanels];
  const loops_coord_t* coords[kMaxPanels];
  float* y[kMaxPanels];  // partial outputs (P > 1)
  fixup_ws ws[kMaxPanels];
  int nnz[kMaxPanels];
  int tiles[kMaxPanels];
};
// ---- row-block plan (spmv_blocks.hip) ------------------------------------
// Two workgroups per CU: LDS y of kBlkRows floats + a round's products
// (kBlkRound floats) each, within the CU's 160 KB.
constexpr int kBlkRows = 15872;                 // rows per block (< 2^14)
constexpr int kBlkSlotBits = 12;
constexpr int kBlkRound = 1 << kBlkSlotBits;    // entries per round at most
constexpr int kBlkColBits = 32 - kBlkSlotBits;  // column bits within a segment
constexpr int kBlkChunkCols = 64;               // columns per build bucket
constexpr int kBlkBatchRounds = 8;              // rounds per stream of plans for batches
struct blk_plan_state {
  int rows = 0, cols = 0, nnz = 0, S = 4, W = 1, nb = 0, nrounds = 0;
  loops_value_type_t vtype = LOOPS_VALUE_F32;
  // Entries of a round in column order, stored lane-major per 512-entry slab
  // (entry 512 q + 64 j + l at 512 q + 256 (j / 4) + 4 l + j % 4; 16-bit
  // values at 512 q + 8 l + j), each round from slab round_base[r] on:
  // column in segment << kBlkSlotBits | slot in round, and the values.
  uint32_t* packed = nullptr;
  void* values = nullptr;
  uint16_t* srow = nullptr;     // stream row of slot k of round r at r * kBlkRound + k
  int* stream_round = nullptr;  // first round of each (block, segment) stream
  int* round_off = nullptr;     // first entry of each round (nrounds + 1), stream order
  int* round_base = nullptr;    // first slab of each round in packed / values (nrounds + 1)
  int* blk_start = nullptr;     // first row of each block (nb + 1)
  int* blk_order = nullptr;     // blocks in launch order (heaviest first)
  float* partial = nullptr;     // S arrays of part_stride floats
  long long part_stride = 0;
  // Rows longer than a quarter of a block's share are cut into pieces
  // (entry i of a row of m pieces to piece i % m).  Blocks, the LDS
  // accumulator and the partials count pieces ("virtual rows"); vrow[r] is
  // the first virtual row of row r (rows + 1 entries), null when no row is
  // cut (virtual row = row).
  int nsplit = 0, vrows = 0;
  int* vrow = nullptr;
  // Stream rows.  Compressed plans (rlist != null) number, per (block,
  // segment) stream, only the virtual rows with entries in that segment:
  // rlist holds them stream by stream (segment-major: segment s, blocks in
  // order), rstart[s * (nb + 1) + b] where stream (b, s) starts; the LDS
  // accumulator and the partial writes cover those rows only, and the
  // partials of the others stay 0 from the build.  Dense plans number all
  // rows of the block (virtual row - blk_start[b]).
  int* rstart = nullptr;
  int* rlist = nullptr;
  // One segment, no row cut and every row a stream row: the streams write y
  // themselves (partial is null and no combine runs).
  bool direct = false;
  void* mem = nullptr;
  size_t device_bytes = 0;
};
int blk_segments(long long cols);
// for_batches: one segment when the columns fit one, and the fewest row
// blocks the LDS allows (batch items), instead of enough blocks to fill the
// GPU alone.
hipError_t blk_plan_build(const csr_args& A, blk_plan_state* st, bool for_batches,
                          hipStream_t s);
hipError_t blk_plan_execute(const blk_plan_state& st, const float* x, float* y,
                            hipEvent_t ev1, hipStream_t s);
void blk_plan_release(blk_plan_state* st);

struct plan_state {
  int layout = LOOPS_PLAN_LAYOUT_PANELS;
  blk_plan_state blk;
  int P = 1, rows = 0, tpb = 256, ipt = 8, tmax = 0;
  bool aligned = true;
  loops_value_type_t vtype = LOOPS_VALUE_F32;
  panel_bounds_t bound{};
  panel_set_t ps{};
  void* offsets_mem = nullptr;  // P > 1: panel row offsets
  void* data_mem = nullptr;     // P > 1: panel column ids and values
  void* aux_mem = nullptr;      // tile coordinates, fix-up slots, partial y
  size_t device_bytes = 0;
};
bool panels_supported(int tpb, int ipt);
hipError_t plan_build(const csr_args& A, int P, int tpb, int ipt, plan_state* st,
                      hipStream_t s);
hipError_t plan_execute(plan_state* st, const float* x, float* y, hipEvent_t ev0,
                        hipEvent_t ev1, hipStream_t s);
void plan_release(plan_state* st);
// OR of LOOPS_CSR_* flags of A (0: well formed); synchronises s.
hipError_t csr_structure_flags(const csr_args& A, hipStream_t s, unsigned* out);
size_t exclusive_scan_scratch_bytes(long long n);
hipError_t launch_exclusive_scan(const int* cnt, long long n, int* out, void* scratch,
                                 hipStream_t s);

// ---- batches -------------------------------------------------------------
struct batch_item_dev;
loops_status_t batch_layout(const loops_batch_item_t* items, int64_t n,
                            loops_batch_info_t* info, batch_item_dev* out_table,
                            char* dev_ws);
hipError_t batch_launch(const void* dev_table, const loops_batch_info_t* info,
                        int grid, hipStream_t s);
hipError_t batch_warm_up();

// ---- SpMM ----------------------------------------------------------------
hipError_t launch_tm_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s);
bool gm_spmm_supported(int tpb);
hipError_t launch_gm_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s);
bool mp_spmm_supported(int tpb, int ipt);
// Whether every column chunk of a k-column product with this B takes ipt
// steps per range (spmm_ipt_ok for the chunk's lanes per range).
bool mp_spmm_fits(int k, const float* B, long long ldb, int ipt);
// Merge-path SpMM with more than kSpmmChunk columns runs as launches of
// kSpmmChunk columns (the last one the rest), one after the other on the
// stream.  With the range's B rows loaded 8 steps at a time, ranges of 16
// lanes x 4 columns with 32 steps beat two launches of 8-lane ranges
// (1M-row, 16M-nonzero Zipf(2.0) matrix, fp32, k = 64: 0.70 ms against
// 2 x 0.375; 400k-row, 2M-nonzero matrix: 0.131 against 2 x 0.082).
// Batches split such items the same way, so an item's result stays the one
// loops_spmm gives.
constexpr int kSpmmChunk = 64;
// Merge steps per range of a merge-path SpMM tile when the caller passes 0:
// 32 for ranges of 16 lanes or more, else 16.  The range loads its B rows 8
// steps at a time, so longer ranges cost no registers and spread the tile's
// fixed work (end-point searches, staging, the scan across ranges) over
// more nonzeros.  1M-row, 16M-nonzero Zipf(2.0) matrix, microseconds at 4 /
// 8 / 16 / 32 steps: k = 1 178 / 156 / 145 / -, k = 4 274 / 247 / 242 / -,
// k = 16 343 / 325 / 310 / -, k = 32 555 / 433 / 378 / 375, k = 64 1258 /
// 938 / 747 / 700; 400k-row, 2M-nonzero matrix: k = 1 39 / 37 / 32 / -,
// k = 16 71 / 69 / 71 / -, k = 64 175 / 145 / 131 / 133.
inline int spmm_default_ipt(int g) { return g >= 16 ? 32 : 16; }
// Steps per range a merge-path SpMM tile takes: 4, 8, 16, or 32 for ranges
// of 8 lanes or more (the tile's LDS holds 1,024 steps at 32).
inline bool spmm_ipt_ok(int ipt, int g) {
  return ipt == 4 || ipt == 8 || ipt == 16 || (ipt == 32 && g >= 8);
}
// SpMM lanes take 4 consecutive columns (16-byte B loads) when B's rows are
// 16-byte aligned (B and ldb) and k is at least this, else 1 column.  C is
// written with 16-byte stores where its address allows, else column by
// column: C's alignment does not pick the shape.  400k-row, 2M-nonzero
// Zipf(2.0) matrix, C rows of k floats: k = 11-15 74-83 us against 100-103
// with 1 column per lane; k = 4-5 63-70 us with 1 column against 72-76 with
// 4; k = 8 60 us in ranges of 2 lanes x 4 columns (72 with 1 column or 4
// lanes).  Ranges hold the fewest lanes (at least 2) that span k columns.
constexpr int kSpmmVec4MinK = 8;
// Width of the chunk starting at column j of a k-column product (k >
// kSpmmChunk): kSpmmChunk, except that 65 columns left go as 60 + 5, never
// leaving a one-column chunk (a one-column batch item is an SpMV, which
// reads x and y as vectors) and keeping chunk starts 16-byte aligned.
inline int spmm_chunk_width(long long k, long long j) {
  const long long left = k - j;
  if (left == kSpmmChunk + 1) return kSpmmChunk - 4;
  return left < kSpmmChunk ? (int)left : kSpmmChunk;
}
long long mp_spmm_tiles(const csr_args& A, int k, int ipt);
size_t mp_spmm_ws_bytes(long long tiles, int k);
// Workspace of a loops_spmm merge-path call: the most any of its chunks needs.
size_t mp_spmm_ws_need(const csr_args& A, int k, int ipt);
hipError_t launch_mp_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, int ipw, void* ws,
                          hipEvent_t ev0, hipEvent_t ev1, hipStream_t s);

}  // namespace loops_mi355x

// The opaque plan handle of the C ABI (loops_spmv_plan_t); batch items may
// carry one (loops_batch_item_t.plan).
struct loops_spmv_plan {
  loops_mi355x::plan_state st;
  int64_t rows = 0, cols = 0;
};

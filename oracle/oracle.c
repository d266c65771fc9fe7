/*
 * CPU oracle for the load-balanced CSR SpMV/SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it may be linked or
 * called only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  The product library (edge_loops_amd) never calls it and
 * has no CPU fallback.
 *
 * Every function restates one piece of the reference's algorithm and names
 * the reference file:line it follows (paths relative to the reference root).
 * Built with -ffp-contract=off so that the fp32 sums are evaluated exactly as
 * the sequential validator writes them: one rounded multiply, then one
 * rounded add, in CSR order.
 *
 * Pinning (what checks each function; the tests live in tests/test_oracle.py):
 *   oracle_spmv_f32        sample.hxx 4x4 known answer, chesapeake degrees
 *                          (x = 1), FP64 bound, a planted dropped-term fails.
 *   oracle_spmv_f64 / abs  known answers an fp32 accumulator or an fp32
 *                          product fails (2^25 + 1 - 2^25 = 1; (1+2^-23)^2
 *                          exact in double), |y64| <= abs.
 *   oracle_spmv_f32_mt     bit-identical to oracle_spmv_f32 for 1..16
 *                          threads on skewed and uniform matrices (rows
 *                          left unwritten would keep a NaN sentinel).
 *   oracle_mp_search       brute-force merge walk for every diagonal of
 *                          random tiny matrices, invariants x+y=d,
 *                          offsets[x] <= y <= offsets[x+1].
 *   oracle_mp_tile_coords  == search at t*items_per_tile (definition), and
 *   oracle_mp_thread_coords  the sub-path property vs the brute-force walk.
 *   oracle_wo_coords       == search at clamped tid*wpt, brute force.
 *   oracle_gm_atom_row     brute force over the exclusive scan.
 *   oracle_spmv_d          integer known answers (exact in double), the
 *                          sequential double loop element by element, and
 *                          a cancellation the long-double reference keeps.
 *   oracle_spmv_d_ref      2^54 + 1 - 2^54 = 1 (double: 0), |y| <= abs.
 *   oracle_spmv_i32        wrap-around known answers (2^31 products), and
 *                          the result is the same for any summation order.
 *   oracle_spmm_f32        column-by-column equality with oracle_spmv_f32.
 *   oracle_spmm_f64        the same fp64 known answers per column, and
 *                          column-by-column equality with oracle_spmv_f64.
 *   oracle_count_*_errors  hand-made vectors on both sides of the bound.
 *   oracle_mtx_read        chesapeake: 39 x 39, 340 nnz (README.md:86),
 *                          symmetric, sorted (row, col); duplicate entries
 *                          keep their input order (a written file).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* SpMV validator: examples/spmv/helpers.hxx:106-112.                       */
/* for row: sum = 0; for nz in [off[row], off[row+1]): sum += v*x[idx]      */
/* ------------------------------------------------------------------------ */
void oracle_spmv_f32(int64_t rows, const int32_t* offsets,
                     const int32_t* indices, const float* values,
                     const float* x, float* y) {
  for (int64_t row = 0; row < rows; ++row) {
    float sum = 0.0f;
    for (int32_t nz = offsets[row]; nz < offsets[row + 1]; ++nz) {
      float prod = values[nz] * x[indices[nz]];
      sum = sum + prod;
    }
    y[row] = sum;
  }
}

/* Same loop in double precision (the FP64 host oracle), and the per-row
 * magnitude sum_j |a_ij * x_j| that scales the 1e-5 relative tolerance. */
void oracle_spmv_f64(int64_t rows, const int32_t* offsets,
                     const int32_t* indices, const float* values,
                     const float* x, double* y, double* abs_sum) {
  for (int64_t row = 0; row < rows; ++row) {
    double s = 0.0, a = 0.0;
    for (int32_t nz = offsets[row]; nz < offsets[row + 1]; ++nz) {
      double p = (double)values[nz] * (double)x[indices[nz]];
      s += p;
      a += fabs(p);
    }
    y[row] = s;
    if (abs_sum) abs_sum[row] = a;
  }
}

/* The validator loop for type_t = double (the reference's host entries are
 * instantiated for double too: benchmarks/spmv/work_oriented.cu:53):
 * double products and sums in CSR order. */
void oracle_spmv_d(int64_t rows, const int32_t* offsets, const int32_t* indices,
                   const double* values, const double* x, double* y) {
  for (int64_t row = 0; row < rows; ++row) {
    double sum = 0.0;
    for (int32_t nz = offsets[row]; nz < offsets[row + 1]; ++nz) {
      double prod = values[nz] * x[indices[nz]];
      sum = sum + prod;
    }
    y[row] = sum;
  }
}

/* A reference for double results: long double (x86 80-bit, 64-bit
 * mantissa) products and sums, and the per-row sum_j |a_ij x_j|. */
void oracle_spmv_d_ref(int64_t rows, const int32_t* offsets, const int32_t* indices,
                       const double* values, const double* x, double* y, double* abs_sum) {
  for (int64_t row = 0; row < rows; ++row) {
    long double s = 0.0L, a = 0.0L;
    for (int32_t nz = offsets[row]; nz < offsets[row + 1]; ++nz) {
      long double p = (long double)values[nz] * (long double)x[indices[nz]];
      s += p;
      a += fabsl(p);
    }
    y[row] = (double)s;
    abs_sum[row] = (double)a;
  }
}

/* The validator loop for type_t = int: products and sums wrap around
 * modulo 2^32 (unsigned arithmetic, so no overflow is undefined), which
 * makes every summation order give the same bits. */
void oracle_spmv_i32(int64_t rows, const int32_t* offsets, const int32_t* indices,
                     const int32_t* values, const int32_t* x, int32_t* y) {
  for (int64_t row = 0; row < rows; ++row) {
    uint32_t sum = 0;
    for (int32_t nz = offsets[row]; nz < offsets[row + 1]; ++nz)
      sum += (uint32_t)values[nz] * (uint32_t)x[indices[nz]];
    y[row] = (int32_t)sum;
  }
}

/* Row-partitioned multi-threaded variant of the validator loop, used only as
 * the all-cores CPU baseline.  Each row is still summed in CSR order. */
typedef struct {
  int64_t r0, r1;
  const int32_t *offsets, *indices;
  const float *values, *x;
  float* y;
} mt_job_t;

static void* mt_worker(void* p) {
  mt_job_t* j = (mt_job_t*)p;
  for (int64_t row = j->r0; row < j->r1; ++row) {
    float sum = 0.0f;
    for (int32_t nz = j->offsets[row]; nz < j->offsets[row + 1]; ++nz) {
      float prod = j->values[nz] * j->x[j->indices[nz]];
      sum = sum + prod;
    }
    j->y[row] = sum;
  }
  return NULL;
}

int oracle_spmv_f32_mt(int nthreads, int64_t rows, const int32_t* offsets,
                       const int32_t* indices, const float* values,
                       const float* x, float* y) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  mt_job_t jobs[256];
  /* Balance by nnz: split the nnz range evenly, snap to row starts. */
  int64_t nnz = offsets[rows];
  int64_t r = 0;
  for (int t = 0; t < nthreads; ++t) {
    int64_t target = (nnz * (t + 1)) / nthreads;
    int64_t r1 = r;
    if (t == nthreads - 1) {
      r1 = rows;
    } else {
      int64_t lo = r, hi = rows;
      while (lo < hi) { /* first row whose start >= target */
        int64_t mid = (lo + hi) / 2;
        if (offsets[mid] < target) lo = mid + 1; else hi = mid;
      }
      r1 = lo;
    }
    jobs[t].r0 = r; jobs[t].r1 = r1;
    jobs[t].offsets = offsets; jobs[t].indices = indices;
    jobs[t].values = values; jobs[t].x = x; jobs[t].y = y;
    r = r1;
  }
  for (int t = 0; t < nthreads; ++t)
    if (pthread_create(&th[t], NULL, mt_worker, &jobs[t]) != 0) return -1;
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Validators.                                                              */
/* ------------------------------------------------------------------------ */
/* examples/spmv/helpers.hxx:136-139 + util/equal.hxx:40-62: count entries
 * with |a - b| > 1e-2. */
int64_t oracle_count_abs_errors(int64_t n, const float* a, const float* b,
                                double tol) {
  int64_t errors = 0;
  for (int64_t i = 0; i < n; ++i)
    if (!(fabs((double)a[i] - (double)b[i]) <= tol)) ++errors;
  return errors;
}

/* Per-row relative check: |y - yref| <= rel * sum_j |a_ij x_j|, with a tiny
 * absolute floor for rows whose magnitude sum is zero (empty rows must then
 * be exactly 0). Returns the number of failing rows; *worst receives the
 * largest |y - yref| / max(abs_sum, tiny). */
int64_t oracle_count_rel_errors(int64_t n, const float* y, const double* yref,
                                const double* abs_sum, double rel,
                                double* worst) {
  int64_t errors = 0;
  double w = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double err = fabs((double)y[i] - yref[i]);
    double scale = abs_sum[i];
    double lim = rel * scale;
    double q = scale > 0 ? err / scale : (err > 0 ? INFINITY : 0.0);
    if (q > w) w = q;
    if (!(err <= lim)) ++errors;
  }
  if (worst) *worst = w;
  return errors;
}

/* ------------------------------------------------------------------------ */
/* Merge-path search: util/search.hxx:34-60.                                */
/* A = offsets+1 (row ends, length rows), B = counting 0..nnz-1.            */
/* x = lower_bound over idx in [max(d-nnz,0), min(d,rows)) of the predicate */
/* "A[idx] <= B[d-idx-1]" (i.e. first idx where it is false); y = d - x.    */
/* ------------------------------------------------------------------------ */
void oracle_mp_search(const int32_t* offsets, int64_t rows, int64_t nnz,
                      int64_t diagonal, uint32_t* out_x, uint32_t* out_y) {
  int64_t x_min = diagonal - nnz > 0 ? diagonal - nnz : 0;
  int64_t x_max = diagonal < rows ? diagonal : rows;
  /* thrust::lower_bound with a binary predicate comp(elem, value): find the
   * first position where comp is false. */
  int64_t first = x_min, count = x_max - x_min;
  while (count > 0) {
    int64_t step = count / 2;
    int64_t idx = first + step;
    int comp = (int64_t)offsets[idx + 1] <= (diagonal - idx - 1);
    if (comp) {
      first = idx + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  int64_t x = first < rows ? first : rows;
  *out_x = (uint32_t)x;
  *out_y = (uint32_t)(diagonal - first);
}

/* schedule/merge_path_flat.hxx:108-138: coord[t] = MP(t * items_per_tile)
 * for t in [0, num_tiles] (num_tiles + 1 coordinates).  The diagonal is
 * clamped to rows+nnz so the final coordinate is the end of the path (the
 * reference passes the unclamped diagonal; for t = num_tiles it lands past
 * the end and the search's x_max/x_min clamps give x = rows, y = d - rows;
 * see oracle_mp_tile_coords_unclamped). */
void oracle_mp_tile_coords(const int32_t* offsets, int64_t rows, int64_t nnz,
                           int64_t items_per_tile, int64_t num_coords,
                           uint32_t* xy) {
  int64_t total = rows + nnz;
  for (int64_t t = 0; t < num_coords; ++t) {
    int64_t d = t * items_per_tile;
    if (d > total) d = total;
    oracle_mp_search(offsets, rows, nnz, d, &xy[2 * t], &xy[2 * t + 1]);
  }
}

/* The reference's exact coordinate for tile t without the clamp
 * (merge_path_flat.hxx:128-136). */
void oracle_mp_tile_coords_unclamped(const int32_t* offsets, int64_t rows,
                                     int64_t nnz, int64_t items_per_tile,
                                     int64_t num_coords, uint32_t* xy) {
  for (int64_t t = 0; t < num_coords; ++t)
    oracle_mp_search(offsets, rows, nnz, t * items_per_tile, &xy[2 * t],
                     &xy[2 * t + 1]);
}

/* schedule/merge_path_flat.hxx:321-385: the start coordinate of thread
 * `lane` of tile `tile`, as a global coordinate.  The reference searches the
 * tile-local path (LDS row ends vs counting from tile_start.y) at diagonal
 * lane*ipt, bounded by the tile's own (tile_rows, tile_nnz); by the sub-path
 * property this equals the global MP(min(tile_start_d + lane*ipt,
 * tile_end_d)).  This function evaluates the tile-local search literally. */
void oracle_mp_thread_coords(const int32_t* offsets, int64_t rows, int64_t nnz,
                             int64_t threads_per_block,
                             int64_t items_per_thread, int64_t tile,
                             uint32_t* xy /* 2*threads_per_block */) {
  int64_t ipt = items_per_thread;
  int64_t tile_items = threads_per_block * ipt;
  int64_t total = rows + nnz;
  int64_t d0 = tile * tile_items, d1 = (tile + 1) * tile_items;
  if (d0 > total) d0 = total;
  if (d1 > total) d1 = total;
  uint32_t sx, sy, ex, ey;
  oracle_mp_search(offsets, rows, nnz, d0, &sx, &sy);
  oracle_mp_search(offsets, rows, nnz, d1, &ex, &ey);
  int64_t tile_rows = (int64_t)ex - sx, tile_nnz = (int64_t)ey - sy;
  for (int64_t lane = 0; lane < threads_per_block; ++lane) {
    int64_t d = lane * ipt;
    int64_t x_min = d - tile_nnz > 0 ? d - tile_nnz : 0;
    int64_t x_max = d < tile_rows ? d : tile_rows;
    int64_t first = x_min, count = x_max - x_min;
    while (count > 0) {
      int64_t step = count / 2, idx = first + step;
      /* tile_end_offset[idx] = offsets[1 + sx + idx] (gather at
       * merge_path_flat.hxx:359-366); b[k] = sy + k. */
      int comp = (int64_t)offsets[1 + sx + idx] <= (int64_t)sy + (d - idx - 1);
      if (comp) { first = idx + 1; count -= step + 1; } else count = step;
    }
    int64_t lx = first < tile_rows ? first : tile_rows;
    xy[2 * lane] = (uint32_t)(sx + lx);
    xy[2 * lane + 1] = (uint32_t)(sy + (d - first));
  }
}

/* Brute-force merge walk: coordinates after each of the rows+nnz steps,
 * following the consume rule of algorithms/spmv/merge_path_flat.cuh:75-81
 * (if y < row_end[x] consume nz y, else advance x).  xy gets rows+nnz+1
 * coordinates, xy[d] = MP(d). */
void oracle_mp_walk(const int32_t* offsets, int64_t rows, int64_t nnz,
                    uint32_t* xy) {
  int64_t x = 0, y = 0;
  int64_t total = rows + nnz;
  for (int64_t d = 0; d <= total; ++d) {
    xy[2 * d] = (uint32_t)x;
    xy[2 * d + 1] = (uint32_t)y;
    if (d == total) break;
    if (x < rows && y < offsets[x + 1]) ++y;
    else ++x;
  }
}

/* schedule/work_oriented.hxx:146-167: thread tid's [start, end) coordinates,
 * at diagonals min(wpt*tid, total) and min(wpt*tid + wpt, total), where
 * wpt = ceil(total / num_threads). */
void oracle_wo_coords(const int32_t* offsets, int64_t rows, int64_t nnz,
                      int64_t num_threads, uint32_t* xy /* 4*num_threads */) {
  int64_t total = rows + nnz;
  int64_t wpt = (total + num_threads - 1) / num_threads;
  for (int64_t tid = 0; tid < num_threads; ++tid) {
    int64_t upper = wpt * tid < total ? wpt * tid : total;
    int64_t lower = upper + wpt < total ? upper + wpt : total;
    oracle_mp_search(offsets, rows, nnz, upper, &xy[4 * tid], &xy[4 * tid + 1]);
    oracle_mp_search(offsets, rows, nnz, lower, &xy[4 * tid + 2],
                     &xy[4 * tid + 3]);
  }
}

/* schedule/group_mapped.hxx:146-213: group g owns rows
 * [g*tpt, min((g+1)*tpt, rows)); its atoms are numbered by the exclusive
 * scan of the row lengths; virtual atom v maps to the row found by
 * upper_bound(scan, v) - 1 and to nz = offsets[row] + v - scan[row_local].
 * Writes, for every nonzero of the matrix in CSR order, the (group, virtual
 * atom) that processes it. */
void oracle_gm_atom_map(const int32_t* offsets, int64_t rows,
                        int64_t threads_per_tile, int64_t* group_of_nz,
                        int64_t* vatom_of_nz) {
  int64_t ngroups = (rows + threads_per_tile - 1) / threads_per_tile;
  for (int64_t g = 0; g < ngroups; ++g) {
    int64_t r0 = g * threads_per_tile;
    int64_t r1 = r0 + threads_per_tile < rows ? r0 + threads_per_tile : rows;
    int64_t aggregate = (int64_t)offsets[r1] - offsets[r0];
    for (int64_t v = 0; v < aggregate; ++v) {
      /* upper_bound over scan[i] = offsets[r0+i] - offsets[r0], i < len */
      int64_t lo = 0, hi = r1 - r0;
      while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if ((int64_t)offsets[r0 + mid] - offsets[r0] <= v) lo = mid + 1;
        else hi = mid;
      }
      int64_t row_local = lo - 1;
      int64_t nz = offsets[r0 + row_local] + v -
                   ((int64_t)offsets[r0 + row_local] - offsets[r0]);
      group_of_nz[nz] = g;
      vatom_of_nz[nz] = v;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* SpMM: algorithms/spmm/thread_mapped.cuh:41-52 with container/matrix.cuh  */
/* row-major indexing (idx = cols*r + c, matrix.cuh:27-35), here with       */
/* explicit leading dimensions.                                             */
/* ------------------------------------------------------------------------ */
void oracle_spmm_f32(int64_t rows, const int32_t* offsets,
                     const int32_t* indices, const float* values, int64_t k,
                     const float* B, int64_t ldb, float* C, int64_t ldc) {
  for (int64_t row = 0; row < rows; ++row) {
    for (int64_t c = 0; c < k; ++c) {
      float sum = 0.0f;
      for (int32_t nz = offsets[row]; nz < offsets[row + 1]; ++nz) {
        float prod = values[nz] * B[(int64_t)indices[nz] * ldb + c];
        sum = sum + prod;
      }
      C[row * ldc + c] = sum;
    }
  }
}

void oracle_spmm_f64(int64_t rows, const int32_t* offsets,
                     const int32_t* indices, const float* values, int64_t k,
                     const float* B, int64_t ldb, double* C, double* abs_sum) {
  for (int64_t row = 0; row < rows; ++row) {
    for (int64_t c = 0; c < k; ++c) {
      double s = 0.0, a = 0.0;
      for (int32_t nz = offsets[row]; nz < offsets[row + 1]; ++nz) {
        double p = (double)values[nz] * (double)B[(int64_t)indices[nz] * ldb + c];
        s += p;
        a += fabs(p);
      }
      C[row * k + c] = s;
      abs_sum[row * k + c] = a;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Matrix Market reader: container/market.hxx:99-227 (banner, size line,    */
/* pattern/real/integer entries, 1-based -> 0-based, symmetric expansion in */
/* entry order) followed by the COO -> CSR conversion of csr.hxx:86-94:     */
/* sort by (row, col) (coo.hxx:104-110) and offsets by lower_bound          */
/* (convert.hxx:69-78).  Two calls: first with out arrays NULL to get the   */
/* sizes, then with arrays of those sizes.  Returns 0 on success, negative  */
/* codes for the reference's error conditions.                              */
/* ------------------------------------------------------------------------ */
typedef struct { int64_t r, c, pos; float v; } coo_entry_t;

/* (row, col) order, ties (duplicate entries) kept in input order: a stable
 * sort, as thrust 1.17's sort_by_key gives for the zip-iterator keys of
 * coo.hxx:163-165 (non-primitive keys take its merge sort).  The position
 * tie-break makes qsort stable. */
static int coo_cmp(const void* a, const void* b) {
  const coo_entry_t* p = (const coo_entry_t*)a;
  const coo_entry_t* q = (const coo_entry_t*)b;
  if (p->r != q->r) return p->r < q->r ? -1 : 1;
  if (p->c != q->c) return p->c < q->c ? -1 : 1;
  if (p->pos != q->pos) return p->pos < q->pos ? -1 : 1;
  return 0;
}

int oracle_mtx_read(const char* path, int64_t* rows, int64_t* cols,
                    int64_t* nnz, int32_t* offsets, int32_t* indices,
                    float* values) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[4096];
  if (!fgets(line, sizeof line, f)) { fclose(f); return -2; }
  if (strncmp(line, "%%MatrixMarket", 14) != 0) { fclose(f); return -2; }
  int pattern = strstr(line, "pattern") != NULL;
  int real = strstr(line, "real") != NULL || strstr(line, "integer") != NULL;
  int symmetric = strstr(line, "symmetric") != NULL;
  if (!strstr(line, "coordinate") || (!pattern && !real)) { fclose(f); return -3; }
  do {
    if (!fgets(line, sizeof line, f)) { fclose(f); return -4; }
  } while (line[0] == '%');
  long long m, n, e;
  if (sscanf(line, "%lld %lld %lld", &m, &n, &e) != 3) { fclose(f); return -4; }
  coo_entry_t* ent = (coo_entry_t*)malloc(sizeof(coo_entry_t) * (size_t)(2 * e + 1));
  int64_t k = 0;
  for (long long i = 0; i < e; ++i) {
    long long r, c;
    double w = 1.0;
    int got = pattern ? fscanf(f, " %lld %lld", &r, &c)
                      : fscanf(f, " %lld %lld %lf", &r, &c, &w);
    if (got != (pattern ? 2 : 3)) { free(ent); fclose(f); return -5; }
    if (r == 0 || c == 0) { free(ent); fclose(f); return -6; }
    ent[k].r = r - 1; ent[k].c = c - 1; ent[k].v = (float)w; ent[k].pos = k; ++k;
    if (symmetric && r != c) {
      ent[k].r = c - 1; ent[k].c = r - 1; ent[k].v = (float)w; ent[k].pos = k; ++k;
    }
  }
  fclose(f);
  *rows = m; *cols = n; *nnz = k;
  if (offsets) {
    qsort(ent, (size_t)k, sizeof(coo_entry_t), coo_cmp);
    for (int64_t i = 0; i < k; ++i) { indices[i] = (int32_t)ent[i].c; values[i] = ent[i].v; }
    /* offsets[r] = lower_bound(row_indices, r) */
    int64_t p = 0;
    for (int64_t r = 0; r <= m; ++r) {
      while (p < k && ent[p].r < r) ++p;
      offsets[r] = (int32_t)p;
    }
  }
  free(ent);
  return 0;
}

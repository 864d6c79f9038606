/* C ABI of the MI355X set primitives that the semi-naive loop is built from.
 *
 * These entry points expose, one by one, the device operations behind the
 * reference's MIR maintenance ops so they can be parity-tested and timed in
 * isolation.  Buffers are HOST row-major arrays (n rows x k u32 columns);
 * each call uploads, runs on the current HIP device and downloads.  All
 * functions return 0 on success and nonzero on error (message via
 * srd_last_error()).
 *
 * Reference counterparts (runtime/generalized_datalog/ in harp-lab/srdatalog-python):
 *   srd_prim_sort_dedup      gpu/device_sorted_array_index.ipp:348 build_take_ownership
 *                            (sort gpu/sort.h:297, dedup gpu/deduplication.h:178,
 *                             root_unique_values gpu/deduplication.h:457)
 *   srd_prim_set_difference  gpu/device_sorted_array_index.ipp:966 set_difference_update
 *   srd_prim_merge           gpu/device_sorted_array_index.ipp:709 merge
 *   srd_prim_exclusive_scan  thrust::exclusive_scan as used by
 *                            src/srdatalog/codegen/jit/complete_runner.py:1072-1112
 */
#ifndef SRDATALOG_PRIMS_H
#define SRDATALOG_PRIMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Initialise the device pool on `device`. */
int srd_prim_init(int device);

/* Lexicographically sort and deduplicate n rows of k columns.
 * out_rows: capacity n*k words; *out_n = distinct rows.
 * root_vals / root_offs (capacity n and n+1, may be NULL): distinct column-0
 * values and their first-row offsets, root_offs[r] = *out_n; *out_r = r. */
int srd_prim_sort_dedup(const uint32_t* rows, uint64_t n, uint32_t k, uint32_t* out_rows,
                        uint64_t* out_n, uint32_t* root_vals, uint32_t* root_offs,
                        uint64_t* out_r);

/* srd_prim_sort_dedup for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It has no
 * counterpart in the reference.  Unlike srd_prim_sort_dedup it takes DEVICE
 * buffers, enqueues on `stream` (a hipStream_t; NULL = the null stream), never
 * synchronises and calls no hipMalloc/hipFree: every intermediate lives in a
 * static device array of 34 MiB per GPU, held until the process ends.
 *   cols        k columns of n rows, column c at cols + c*stride
 *   out_cols    sorted distinct rows, column c at out_cols + c*out_stride
 *   root_vals   capacity n: distinct column-0 values
 *   root_offs   capacity n+1: first row of each, root_offs[r] = distinct rows
 *   counts      3 u64 written on the device: distinct rows, distinct column-0
 *               values, error bits (nonzero: the output is not valid)
 * Limits: 1 <= k <= 4, n <= 1,048,576 rows (larger n is rejected with an
 * error), strides >= n.  Calls that overlap on one GPU share the array and
 * race; calls in stream order are safe. */
int srd_prim_sort_dedup_noalloc(const uint32_t* cols, uint64_t stride, uint64_t n, uint32_t k,
                                uint32_t* out_cols, uint64_t out_stride, uint32_t* root_vals,
                                uint32_t* root_offs, uint64_t* counts, void* stream);

/* out = a - b for sorted duplicate-free inputs of arity k (capacity na*k). */
int srd_prim_set_difference(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb,
                            uint32_t k, uint32_t* out, uint64_t* out_n);

/* out = a u b for sorted, duplicate-free, DISJOINT inputs (capacity (na+nb)*k). */
int srd_prim_merge(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint32_t k,
                   uint32_t* out, uint64_t* out_n);

/* out[0..n] = exclusive prefix sums of in[0..n), out[n] = total. */
int srd_prim_exclusive_scan(const uint32_t* in, uint64_t n, uint64_t* out);

/* Device-resident timing of sort+dedup: the rows are uploaded once, then
 * sorted `reps` times; *ms_per_rep = mean wall time per sort (events). */
int srd_prim_time_sort_dedup(const uint32_t* rows, uint64_t n, uint32_t k, uint32_t reps,
                             double* ms_per_rep);

/* Device-resident benchmark of one primitive on synthetic inputs generated
 * on the device (k columns; columns 1..k-1 are 20-bit fields of a 64-bit key,
 * column 0 the rest):
 *   op 0  sort+dedup of na unordered rows (keys hashed over 2^(20k) values);
 *   op 1  set-difference NEW - FULL, NEW = {2i : i < na}, FULL = {3j : j < nb};
 *   op 2  merge of disjoint FULL = {3i : i < na} and DELTA = {3j+1 : j < nb}.
 * One untimed call, then `reps` timed calls; *ms_per_rep = mean wall time
 * per call (HIP events on the primitive's stream); *out_n = output rows. */
int srd_prim_bench(int op, uint64_t na, uint64_t nb, uint32_t k, uint32_t reps,
                   double* ms_per_rep, uint64_t* out_n);

/* Calibration: `reps` launches of srd_calib_copy_kernel, a plain streaming
 * copy of `bytes` with 16-byte words (read once, written once); *ms_per_rep
 * = mean duration.  Then `reps` launches of srd_calib_copy8_kernel, the same
 * copy with 8-byte words.  Their profiler HBM counters give the scale for
 * the other kernels' counters (and check it for 8-byte accesses). */
int srd_calib_copy(uint64_t bytes, int reps, double* ms_per_rep);

/* Device memory pool counters, out[0..cap) of: bytes in use, peak bytes in
 * use, budget (0 = none), allocations served, allocations that reached
 * hipMalloc, bytes obtained from hipMalloc, bytes cached on the free list,
 * times cached blocks were handed back to the device (least recently
 * freed first, as much as a failed hipMalloc needs), allocations served from
 * a range another library stream freed (ordered after that stream's work by
 * an event).  The reference's RMM pool
 * (gpu/device_array_rmm_impl.h:102-160) has no counterpart query. */
int srd_pool_stats(uint64_t* out, uint32_t cap);

/* Kernel timing: while enabled, every primitive and join launch on the
 * runtime's streams is bracketed by HIP events.  srd_timing_report writes a
 * JSON array [{"name", "launches", "ms", "bytes"}] (ms = summed duration,
 * bytes = algorithmic bytes the primitive attributes to those launches)
 * into buf; returns the length needed (excluding NUL). */
int srd_timing_enable(int on);
uint64_t srd_timing_report(char* buf, uint64_t cap);
int srd_timing_reset(void);

/* Message of the last failed call on this thread ("" if none). */
const char* srd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SRDATALOG_PRIMS_H */

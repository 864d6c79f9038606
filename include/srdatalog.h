/* Drop-in C ABI of a compiled srdatalog program (one shared library per
 * program, produced by srdatalog_hip.build_project + compile_jit_project).
 *
 * The first six symbols are exactly the ones the reference's generated
 * shim exports and its Python driver binds with ctypes
 * (src/srdatalog/codegen/jit/main_file.py:656-781, bound in
 * examples/run_benchmark.py:195-203): same names, C signatures and return
 * codes.  The remaining symbols are additions for what the reference's shim
 * cannot report (device-side results, per-iteration deltas, in-memory
 * input); they do not change the six.
 *
 * Threading: one program instance per library; calls are not re-entrant.
 * Device: HIP device 0 (gfx950).
 */
#ifndef SRDATALOG_H
#define SRDATALOG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces main_file.py:685-691.  Initialise device 0 and the memory pool
 * and parse the embedded plan.  0 ok, 1 std::exception (message on stderr),
 * 2 other error. */
int srdatalog_init(void);

/* Replaces main_file.py:693-729.  Load one input relation from a CSV/TSV
 * file (delimiter auto-detected: tab, comma or spaces; one tuple per line).
 * 0 ok, 1 null argument, 2 unknown or non-input relation, 3 parse/IO error. */
int srdatalog_load_csv(const char* rel_name, const char* path);

/* Replaces main_file.py:731-741.  Load every relation declared with an
 * input_file from data_dir/input_file.  0 ok, 1 null, 2 error. */
int srdatalog_load_all(const char* data_dir);

/* Replaces main_file.py:743-753.  Run the program to fixpoint (max_iters
 * caps each recursive step's iterations; 0 = unlimited).  Prints the
 * reference's " >>>>>>>>>>>>>>>>> Rel : N" size lines for print_size
 * relations (silenced by SRDATALOG_QUIET=1).  0 ok, 1 nothing loaded,
 * 2 error. */
int srdatalog_run(unsigned long long max_iters);

/* Replaces main_file.py:755-770.  FULL size of `rel_name`.  Deliberate
 * deviation: after srdatalog_run this is the DEVICE result (the reference
 * reads its host DB, which reports 0 for derived relations).  Before a run
 * it is the loaded input size.  For a count-only (.with_count()) rule's
 * head it is the count.  0 for unknown relations. */
unsigned long long srdatalog_size(const char* rel_name);

/* Replaces main_file.py:772-775.  Free all host and device state; 0. */
int srdatalog_shutdown(void);

/* ---- additions ---------------------------------------------------------- */

/* Copy the FULL tuple set of `rel_name`, sorted lexicographically over the
 * relation's columns, row-major into out (cap_rows rows).  *n_rows = rows in
 * the relation (out may be NULL to query).  0 ok, 1 null, 2 unknown
 * relation, 3 error (e.g. cap too small). */
int srdatalog_export(const char* rel_name, uint32_t* out, uint64_t cap_rows, uint64_t* n_rows);

/* Per-iteration |DELTA| of the last step that computed rel_name; returns the
 * number of entries (written up to cap), -1 for an unknown relation. */
int srdatalog_delta_sizes(const char* rel_name, unsigned long long* out, unsigned long long cap);

/* Number of iterations recorded for rel_name (length of its delta vector). */
unsigned long long srdatalog_num_iterations(const char* rel_name);

/* Load an input relation from host memory (row-major n x arity u32). */
int srdatalog_load_rows(const char* rel_name, const uint32_t* rows, uint64_t n, uint32_t arity);

/* Copy loaded inputs to device memory now (srdatalog_run does it lazily);
 * later runs reuse the resident copies. */
int srdatalog_prepare(void);

/* Wall seconds of the last srdatalog_run's fixpoint (inputs resident). */
double srdatalog_run_seconds(void);

/* Tuples the joins of the last srdatalog_run emitted before deduplication
 * (the reference's per-rule count phase totals, complete_runner.py:1072-1112,
 * summed over rules and iterations). */
unsigned long long srdatalog_join_tuples(void);

/* Probe runs: while on, every srdatalog_run also counts the joins' probes --
 * lower/upper-bound searches, root-map lookups and heavy-key bitmap tests
 * (the reference's search.h group searches, one per source per candidate) --
 * with one extra launch per pipeline, so a timed run leaves it off.
 * srdatalog_join_probes returns the count of the last run (0 when off). */
void srdatalog_count_probes(int on);
unsigned long long srdatalog_join_probes(void);

/* Instance table of a packed batch (no counterpart in the reference, which
 * runs one instance per process): bounds[0..n_bounds) ascending, instance g
 * owns column-0 values [bounds[g], bounds[g+1]).  While set, every DELTA
 * computation also counts each instance's DELTA rows on the device.
 * n_bounds = 0 clears the table.  0 ok, 2 error. */
int srdatalog_set_instances(const uint32_t* bounds, uint64_t n_bounds);

/* Per-instance |DELTA| of the last step that computed rel_name: row i of
 * out (n_inst entries, n_inst = n_bounds - 1) is iteration i; up to
 * cap_iters rows are written.  Returns the number of iterations, -1 for an
 * unknown relation or error. */
long long srdatalog_instance_delta_sizes(const char* rel_name, unsigned long long* out,
                                         unsigned long long cap_iters);

/* Export of a packed batch: like srdatalog_export, but every value of a row
 * is rebased to its instance (minus the instance's first bound), and cuts
 * (n_inst+1 entries, may be NULL) receives each instance's first row, so
 * instance g owns rows [cuts[g], cuts[g+1]).  Needs an instance table.
 * 0 ok, 1 null, 2 unknown relation, 3 error. */
int srdatalog_export_instances(const char* rel_name, uint32_t* out, uint64_t cap_rows,
                               uint64_t* n_rows, uint64_t* cuts);

/* ---- Engine handles and resumable runs (iteration-granular serving) ----
 * No counterpart in the reference, which runs one fixpoint per process call.
 * srdatalog_engine_new() creates another engine of this program (NULL on
 * error; release with srdatalog_engine_free).  For a program whose last step
 * is its only recursive one:
 *   run_begin        runs every earlier step once (the recursive step at
 *                    iteration 0); inputs come from engine_load_rows;
 *   run_iterations   advances the recursive step by at most k iterations;
 *                    *done = iterations run, *converged = 1 once an
 *                    iteration left every DELTA empty;
 *   absorb(dst, src) unites the state src reached after its own run_begin
 *                    into dst's FULL (and into DELTA for the recursive
 *                    step's relations), as if those rows had been in dst's
 *                    base results; both are engines of this program;
 *   range_counts     rows of rel's FULL (delta = 0) or DELTA (delta = 1)
 *                    whose values lie in each of m ranges [lohi[2g],
 *                    lohi[2g+1]), ascending and disjoint, into out[g];
 *   export_ranges    rel's FULL rows in each of m ranges (as range_counts),
 *                    row-major in relation column order, the ranges one
 *                    after the other: range g owns rows [cuts[g], cuts[g+1])
 *                    (cuts: m+1 entries, may be NULL); with rebase != 0
 *                    every value of range g's rows is reduced by lohi[2g];
 *                    *n_rows is set, and cuts filled, also when out is NULL;
 *   erase_ranges     drops the rows in each of m ranges (as range_counts)
 *                    from every relation, in one pass over each index.
 * Instances must occupy disjoint value ranges with every value of an
 * instance's rows inside its range (as a packed batch's do); an instance's
 * rows and per-iteration |DELTA| are then those of a run of it alone,
 * shifted by the iteration it was absorbed at.  0 ok, 1 null, 2 error
 * (range_counts / export_ranges: 2 unknown relation, 3 error). */
void* srdatalog_engine_new(void);
void srdatalog_engine_free(void* h);
void srdatalog_engine_set_quiet(void* h, int quiet);
int srdatalog_engine_load_rows(void* h, const char* rel_name, const uint32_t* rows, uint64_t n,
                               uint32_t arity);
int srdatalog_engine_run_begin(void* h);
int srdatalog_engine_run_iterations(void* h, unsigned long long k, unsigned long long* done,
                                    int* converged);
int srdatalog_engine_absorb(void* dst, void* src);
int srdatalog_engine_range_counts(void* h, const char* rel_name, int delta, const uint32_t* lohi,
                                  uint32_t m, unsigned long long* out);
int srdatalog_engine_export_ranges(void* h, const char* rel_name, const uint32_t* lohi, uint32_t m,
                                   int rebase, uint32_t* out, uint64_t cap_rows, uint64_t* n_rows,
                                   uint64_t* cuts);
int srdatalog_engine_erase_ranges(void* h, const uint32_t* lohi, uint32_t m);

#ifdef __cplusplus
}
#endif

#endif /* SRDATALOG_H */

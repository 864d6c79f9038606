// Engine entry points (extern "C") called by the shim a generated program
// exports; see include/srdatalog.h for the public symbols built on them.
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct srd_engine;
int srd_engine_ok(struct srd_engine* e);
int srd_engine_load_csv(struct srd_engine* e, const char* rel, const char* path);
int srd_engine_load_all(struct srd_engine* e, const char* dir);
int srd_engine_load_rows(struct srd_engine* e, const char* rel, const uint32_t* rows, uint64_t n,
                         uint32_t arity);
int srd_engine_prepare(struct srd_engine* e);
int srd_engine_run(struct srd_engine* e, unsigned long long max_iters);
unsigned long long srd_engine_size(struct srd_engine* e, const char* rel);
int srd_engine_export(struct srd_engine* e, const char* rel, uint32_t* out, uint64_t cap_rows,
                      uint64_t* n_rows);
unsigned long long srd_engine_num_iterations(struct srd_engine* e, const char* rel);
int srd_engine_delta_sizes(struct srd_engine* e, const char* rel, unsigned long long* out,
                           unsigned long long cap);
int srd_engine_shutdown(struct srd_engine* e);
double srd_engine_run_seconds(struct srd_engine* e);
unsigned long long srd_engine_join_tuples(struct srd_engine* e);
void srd_engine_count_probes(struct srd_engine* e, int on);
unsigned long long srd_engine_join_searches(struct srd_engine* e);
void srd_engine_set_quiet(struct srd_engine* e, int quiet);
void srd_engine_set_grain(struct srd_engine* e, unsigned grain);
int srd_engine_set_instances(struct srd_engine* e, const uint32_t* bounds, uint64_t n_bounds);
int srd_engine_export_instances(struct srd_engine* e, const char* rel, uint32_t* out,
                                uint64_t cap_rows, uint64_t* n_rows, uint64_t* cuts);
long long srd_engine_instance_delta_sizes(struct srd_engine* e, const char* rel,
                                          unsigned long long* out, unsigned long long cap_iters);
/* Resumable runs (iteration-granular serving); see include/srdatalog.h. */
int srd_engine_run_begin(struct srd_engine* e);
int srd_engine_run_iterations(struct srd_engine* e, unsigned long long k, unsigned long long* done,
                              int* converged);
int srd_engine_absorb(struct srd_engine* dst, struct srd_engine* src);
int srd_engine_range_counts(struct srd_engine* e, const char* rel, int delta, const uint32_t* lohi,
                            uint32_t m, unsigned long long* out);
int srd_engine_export_ranges(struct srd_engine* e, const char* rel, const uint32_t* lohi, uint32_t m,
                             int rebase, uint32_t* out, uint64_t cap_rows, uint64_t* n_rows,
                             uint64_t* cuts);
int srd_engine_erase_ranges(struct srd_engine* e, const uint32_t* lohi, uint32_t m);

#ifdef __cplusplus
}
#endif

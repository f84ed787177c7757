/*
 * oracle.h -- CPU FP64 restatement of cmtj's (s)LLGS time integration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.  The
 * product (cmtj_amd) never links, imports or calls it.
 *
 * It takes the same descriptors as the engine (include/cmtj_amd.h) and
 * follows the reference algorithm step by step (citations in oracle.c).
 */
#ifndef CMTJ_ORACLE_H_
#define CMTJ_ORACLE_H_
#include "../include/cmtj_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Driver value at time t; step/stage locate CMTJ_TABLE samples. */
double oracle_driver_value(const cmtj_driver* d, double t, uint32_t step,
                           int stage, int samples_per_step,
                           const double* tables);

/* Philox4x32-10 block and the thermal normal triple of one (step, layer). */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                          uint32_t out[4]);
void oracle_thermal_normals(uint64_t seed, uint64_t stream, uint32_t step,
                            uint32_t layer, double xi[3]);

/* One junction; log is [channel][entry] doubles (j->log_offset ignored). */
int oracle_run(const cmtj_junction* j, const cmtj_layer* layers,
               const cmtj_constants* c, const double* tables, size_t n_table,
               double* log, double* mag_out, cmtj_status* st);

/* Many junctions over n_threads host threads; log blocks at log_offset
 * (FP64), mag_out [n][CMTJ_MAX_LAYERS][3]. Returns 0. */
int oracle_run_batch(const cmtj_junction* js, size_t n,
                     const cmtj_layer* layers, const cmtj_constants* c,
                     const double* tables, size_t n_table, double* log,
                     double* mag_out, cmtj_status* st, int n_threads);

/* A coupled stack (core/stack.hpp), same descriptors, tables and log layout
 * as cmtj_run_stacks for one stack; mag_out [n][CMTJ_MAX_LAYERS][3], st [n].
 * Returns 0. */
int oracle_run_stack(const cmtj_stack* s, const cmtj_junction* js,
                     const cmtj_layer* layers, const cmtj_constants* c,
                     const double* tables, double* log, double* mag_out, cmtj_status* st);

/* A dipole-coupled group (core/reservoir.hpp GroupInteraction), same
 * descriptors and log layout as cmtj_run_groups for one group; mag_out
 * [n][CMTJ_MAX_LAYERS][3], st [n].  Returns 0, 1 (bad descriptor) or 2
 * (two points too close for the interaction). */
int oracle_run_group(const cmtj_group* g, const cmtj_junction* js,
                     const cmtj_layer* layers, const cmtj_constants* c,
                     const double* pair_terms, double* log, double* mag_out, cmtj_status* st);

/* Channels of an instance for a log mask (same rule as the engine). */
int oracle_n_channels(int n_layers, int mr_mode, int log_mask);

#ifdef __cplusplus
}
#endif
#endif

/*
 * vc_batch.h -- batched path of libvoxelcore (gfx950): many volumes of
 * uneven sizes through EDT + measure + thinning, one launch per pass over the
 * whole batch, captured sweep segments, and a launch-parameter autotuner.
 *
 * No counterpart in the reference, which handles one volume per process
 * (src/voroUtility.cpp:21-142 -> src/highlevelalgo.cpp:815-961).  Each
 * volume's results are exactly those of the single-volume entry points of
 * vc_api.h on the same volume (volumes never interact): read them with
 * vc_get / vc_get_device on the member volumes.
 *
 * Schedules change launch shapes only, never results.
 */
#ifndef VC_BATCH_H
#define VC_BATCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vc_ctx vc_ctx;
typedef struct vc_vol vc_vol;
typedef struct vc_thin_stats vc_thin_stats;
typedef struct vc_batch vc_batch;

/* launch parameters (autotuned by vc_batch_tune; all results identical) */
typedef struct vc_schedule {
    int edt_lds_kb;      /* LDS budget of the EDT column tiles: 8, 16, 32
                            or 64 KiB (the x pass gets twice it, at most
                            64 KiB)                                       */
    int edt_grid_per_cu; /* EDT + measure launches: persistent grid of
                            this many workgroups per CU; 0 = one
                            workgroup per work item                       */
    int thin_grid_per_cu;/* subfield passes: the same                     */
    int sweeps_per_sync; /* sweeps launched between two convergence
                            checks: 0 = adaptive (4, 8, 12, then 16),
                            -1 = adaptive doubling (4, 8, then 16), or
                            1-32                                          */
    int use_graph;       /* 1: each segment of sweeps is a captured
                            hipGraph replayed on the context stream       */
    int tail_sweep;      /* sweeps with index >= tail_sweep within a level
                            use the sparse-pass variant (one wavefront per
                            256 words, flagged words compacted in LDS);
                            -1 = never                                    */
    int tile_sweep;      /* from the first segment of sweeps that starts at
                            an index >= tile_sweep within a level, each
                            sweep is one launch over 32 x 32 x 64 tiles
                            (all 8 subfield passes in LDS, 8-row halo),
                            visiting only tiles next to last sweep's
                            changes; -1 = never                           */
} vc_schedule;

/* volumes: n member volumes of this context (1 <= n <= 65535, distinct); the
 * batch keeps pointers to them, release the batch before its volumes. */
int vc_batch_create(vc_ctx* ctx, vc_vol* const* vols, int n, vc_batch** out);
int vc_batch_destroy(vc_ctx* ctx, vc_batch* b);
int vc_batch_size(const vc_batch* b, int* n, int64_t* voxels);

/* vc_edt on every member volume (flags must be 0) */
int vc_batch_edt(vc_ctx* ctx, vc_batch* b, int flags);
/* vc_measure on every member volume */
int vc_batch_measure(vc_ctx* ctx, vc_batch* b);
/* vc_thin_levels on every member volume with its own levels: levels holds
 * n rows of n_levels ascending values (row i for volume i).  Each volume uses
 * the measure it has (vc_measure or vc_set_measure_f32).  stats (may be NULL)
 * sums over the batch; when given, the call waits. */
int vc_batch_thin(vc_ctx* ctx, vc_batch* b, const float* levels,
                  int n_levels, vc_thin_stats* stats);
/* vc_batch_edt + vc_batch_measure + vc_batch_thin */
int vc_batch_run(vc_ctx* ctx, vc_batch* b, const float* levels, int n_levels,
                 vc_thin_stats* stats);

/* schedules: the default, the batch's current one, and pinning one */
int vc_schedule_defaults(vc_schedule* out);
int vc_batch_get_schedule(const vc_batch* b, vc_schedule* out);
int vc_batch_set_schedule(vc_ctx* ctx, vc_batch* b, const vc_schedule* sc);
/* Autotuner: times every candidate schedule on this batch (EDT + measure
 * candidates on those stages, thinning candidates on a full thinning with
 * these levels), in a fixed order, median of `reps` runs each (reps >= 1),
 * keeps a candidate only if it beats the incumbent by more than 2%, installs
 * the winner and caches it in the context under the batch's shape key
 * (volume count, total voxels, largest axis).  A later vc_batch_tune on a
 * batch with the same key reuses the cached schedule without timing.
 * times_ms (may be NULL) receives the winner's EDT + measure and thinning
 * times.  Leaves every member volume thinned with these levels. */
int vc_batch_tune(vc_ctx* ctx, vc_batch* b, const float* levels, int n_levels,
                  int reps, vc_schedule* chosen, float* times_ms);

#ifdef __cplusplus
}
#endif
#endif /* VC_BATCH_H */

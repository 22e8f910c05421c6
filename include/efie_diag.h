/*
 * efie_diag.h -- diagnostic entry points of libefie_hip.so (no counterpart
 * in the reference): a plain device-to-device copy kernel of known size,
 * used to calibrate hardware byte counters (FETCH_SIZE / WRITE_SIZE) when
 * profiling the EFIE kernels.
 */
#ifndef EFIE_DIAG_H
#define EFIE_DIAG_H

#include "efie.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Copies `bytes` (multiple of 16) between two device buffers `reps` times
 * with one streaming kernel per copy; *ms = average kernel time. */
efie_status efie_diag_copy(efie_ctx* ctx, size_t bytes, int reps, double* ms);

/* Process-wide counts of frequency-independent geometry builds (quadrature
 * points, near lists) and near/self moment integrations, over all
 * contexts: a caller that sets the same tables again (one MEX call after
 * another) sees neither count move. */
efie_status efie_diag_cache_counts(long long* geometry_builds, long long* moment_builds);

/* The unordered near triangle pairs (0-based t1 < t2, ordered by t1, then
 * t2) the quadrature fill of *opts uses on the context's mesh: *n is their
 * number; the first min(*n, cap) go to pairs[2k], pairs[2k+1] (pairs may
 * be NULL with cap 0). */
efie_status efie_diag_near_pairs(efie_ctx* ctx, const efie_opts* opts, long long* n, int* pairs, long long cap);

/* The quadrature points the fill of *opts uses on the context's mesh
 * (built on the device), copied out: set 0 the regular rule [Nt][Nq],
 * 1 the near-pair children [Nt][4^near_subdiv][Nq2], 2 the self children
 * [Nt][4^self_subdiv][Nq2], each point as (x, y, z, w) relative to its
 * triangle's centroid; set 3 the self softening radii [Nt][4^self_subdiv];
 * set 4 the radius of each triangle [Nt].  *n is the number of doubles;
 * the first min(*n, cap) go to out (out may be NULL with cap 0). */
efie_status efie_diag_quad_points(efie_ctx* ctx, const efie_opts* opts, int set, long long* n, double* out,
                                  long long cap);

/* The batches efie_sweep_many would form, without device work: count
 * items of order n[i] with nf[i] frequencies, for precision (EFIE_PREC_*)
 * and opts.batch `batch` (0 = automatic), with free_bytes of free device
 * memory.  *nseg is the number of segments; the first min(*nseg, cap) go
 * to out[6k..6k+5] = (batch, item, first frequency, frequencies, first
 * matrix in the batch, order of the batch), batch after batch. */
efie_status efie_diag_pack_plan(const int* n, const int* nf, int count, int precision, int batch,
                                double free_bytes, int* out, int cap, int* nseg);

#ifdef __cplusplus
}
#endif
#endif

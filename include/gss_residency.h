/*
 * gss_residency.h -- C ABI of the clause residency scorer.
 *
 * When the GPU clause DB exceeds its budget some clauses must leave it.  The
 * reference keeps the clauses whose activity is at least the approximate
 * median (HostClauses::reduceDb / approxNthAct / tryRemoveClausesAsync,
 * gpuShareLib/Clauses.cu:249-282,391-525, activity bumped per report and
 * decayed per added clause, Clauses.cu:200-237,284-291,334).  This module
 * keeps that policy, on the device, and adds a learned one: a 12 -> 64 -> 64
 * -> 1 MLP (FP32, 5057 parameters) over per-clause features (length, age,
 * hit history per round and per solver word, relative activity) that
 * predicts whether a clause will be reported in the next rounds.  It is
 * trained online, on the GPU, by hand-written forward / backward / SGD
 * kernels, from the hit records the sweep itself produces.  Both policies
 * keep every clause of length <= 2 (Clauses.cu:255) and remove about half of
 * the others per reduction; the kept clauses are checked exactly as before.
 *
 * The module works on a tagged engine DB (gss_db_enable_tags): a clause's tag
 * is its slot in the per-clause state arrays.  All entry points are
 * synchronous on the engine's stream and return GSS_* codes (gss.h).
 */
#ifndef GSS_RESIDENCY_H
#define GSS_RESIDENCY_H

#include <stdint.h>

#include "gss.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gss_res gss_res;

#define GSS_RES_FEATURES 12
#define GSS_RES_HIDDEN 64
/* parameter layout: W1[64][12], b1[64], W2[64][64], b2[64], w3[64], b3 */
#define GSS_RES_PARAMS (64 * 12 + 64 + 64 * 64 + 64 + 64 + 1)

#define GSS_RES_POLICY_ACTIVITY 0 /* the reference's activity median */
#define GSS_RES_POLICY_LEARNED 1  /* the MLP's score median */

/* Creates the scorer over engine e (whose DB must be tagged) for tags in
 * [0, max_tags); decay is the per-added-clause activity decay (the
 * reference's 0.99999); seed initialises the MLP.  The scorer uses the
 * engine until gss_res_destroy, which must come before gss_engine_destroy. */
int gss_res_create(gss_engine *e, int64_t max_tags, float decay, uint64_t seed, gss_res **out);
int gss_res_destroy(gss_res *r);

/* Appends n clauses to the engine DB under the given unused tags and starts
 * their state at round `round`: activity = the increment after its own
 * decay step (Clauses.cu:204,334). */
int gss_res_add(gss_res *r, int64_t n, const uint32_t *lits, const int32_t *lens,
                const uint32_t *tags, int32_t round);
/* After a waited plain sweep (gss_sweep_wait): folds the sweep's records into
 * the per-clause state of round `round` (one activity bump per record, as the
 * reference bumps once per reported clause and solver, GpuRunner.cu:375-378)
 * and returns the (clause, assignment) hit pairs of the sweep. */
int gss_res_account(gss_res *r, int32_t round, int64_t *hit_pairs);
/* Records the features of n alive tags at round `round` for training. */
int gss_res_snapshot(gss_res *r, int32_t round, int64_t n, const uint32_t *tags);
/* Trains on the snapshot taken at round - horizon: label = the clause was
 * reported in rounds (round - horizon, round]; samples whose clause left the
 * DB are skipped.  epochs passes of minibatch SGD (momentum 0.9). *loss = mean
 * loss of the last pass, *samples = samples used (0 if no snapshot). */
int gss_res_train(gss_res *r, int32_t round, int horizon, int epochs, int batch, float lr,
                  float *loss, int64_t *samples);
/* Removes about half of the clauses of length >= 3 by the policy: those
 * below the approximate median of activity (20000 log buckets, as
 * Clauses.cu:484-525) or of the MLP score.  evicted (host, max_tags flags,
 * may be null) receives 1 for every tag removed; *kept = clauses left. */
int gss_res_reduce(gss_res *r, int policy, int32_t round, uint8_t *evicted, int64_t *kept);

/* Features of n tags at round `round` into out[n][12] (host). */
int gss_res_features(gss_res *r, int32_t round, int64_t n, const uint32_t *tags, float *out);
/* Current activity of n tags (host). */
int gss_res_activity(gss_res *r, int64_t n, const uint32_t *tags, float *out);
/* MLP parameters (GSS_RES_PARAMS floats, layout above), host copies. */
int gss_res_get_params(gss_res *r, float *out);
int gss_res_set_params(gss_res *r, const float *in);
/* Mean weighted logistic loss and its gradient (GSS_RES_PARAMS floats) over
 * n samples x[n][12], labels y[n] in {0,1}, weights wt[n] (null = 1), with
 * the device kernels used for training; no parameter update. */
int gss_res_grad(gss_res *r, int64_t n, const float *x, const float *y, const float *wt,
                 float *loss, float *grad);
/* MLP output probability for n feature rows. */
int gss_res_score(gss_res *r, int64_t n, const float *x, float *out);

#ifdef __cplusplus
}
#endif

#endif /* GSS_RESIDENCY_H */

/*
 * efie_surrogate.h -- the learned near-pair surrogate of libefie_hip.so
 * (near_mode EFIE_NEAR_SURROGATE).  No counterpart in the reference: it
 * replaces the subdivision quadrature of near-singular triangle pairs
 * (rwg_pair_integrals_near_singular.m:1-81) by
 *
 *   record = plain product rule of the regular pairs (quad_order)
 *          + MLP correction,
 *
 * where the network (13 -> 64 -> 64 -> 64 -> 16, GELU, FP32) sees the pair
 * in the test triangle's frame: origin at the centroid of the lower-numbered
 * triangle t1, x along its first edge, z along its normal, lengths in units
 * of L = sqrt(4 A1 / sqrt(3)); inputs are the 12 normalised coordinates
 * (x0, y0, x1 of t1's first two vertices and the nine coordinates of the
 * other triangle's vertices) and k d (d = centroid distance); outputs are
 * the real and imaginary parts of the eight pair moments of the record
 * (S0, S1, S1', S2, from which the 3x3 RWG vertex combinations of the
 * vector and scalar potential integrals follow), normalised, with the phase
 * exp(-j k d) / (4 pi) taken out.
 *
 * Training (efie_sur_train) runs on the context's GPU with the library's own
 * kernels: pairs are sampled from seeded random two-layer patch meshes (the
 * make_patch_with_ground_mesh family with random cell counts, aspect,
 * layer gap and vertex jitter), labelled by the subdivided near-pair
 * quadrature kernel, and fitted with Adam.  In use, a pair whose inputs
 * leave the training envelope (d/L outside [0.2, 3] or the trained range,
 * any input outside the range seen in training, k L above the trained
 * maximum, a lossy medium, or quadrature options other than the trained
 * ones) is integrated by the subdivision quadrature instead.
 */
#ifndef EFIE_SURROGATE_H
#define EFIE_SURROGATE_H

#include "efie.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    long long n_pairs;       /* training pairs                       (2000000) */
    int n_val;               /* held-out pairs from other meshes       (20000) */
    int epochs;              /* passes over the training pairs            (60) */
    int batch;               /* pairs per Adam step                     (4096) */
    double lr;               /* peak learning rate (cosine decay to lr/100)
                                                                        (2e-3) */
    double kl_max;           /* largest k L trained                     (0.35) */
    unsigned long long seed; /* sampling and initialisation seed           (1) */
} efie_sur_cfg;

typedef struct {
    long long n_train, n_val;
    int steps;
    double ms_sample;        /* host: meshes, near pairs, frames            */
    double ms_label;         /* device: label quadrature                    */
    double ms_train;         /* device: training steps                      */
    double train_loss;       /* last epoch's mean standardised loss         */
    double val_loss;         /* standardised loss on the held-out pairs     */
    /* held-out pairs, error of the record relative to its norm: with the
     * network's correction, and with the plain product rule alone        */
    double val_rel_rms, val_rel_max;
    double base_rel_rms, base_rel_max;
} efie_sur_report;

void efie_sur_cfg_default(efie_sur_cfg* cfg);

/* Floats in a model blob: network parameters, standardisation, envelope
 * and the quadrature options it was trained for. */
int efie_sur_blob_floats(void);

/* Train a surrogate for the near-pair options of *opts (near_alpha,
 * near_subdiv, near_order, quad_order, near_sing_eps) and install it on
 * the context. */
efie_status efie_sur_train(efie_ctx* ctx, const efie_opts* opts, const efie_sur_cfg* cfg, efie_sur_report* rep);
/* Copy the installed model out / install a model. */
efie_status efie_sur_get(efie_ctx* ctx, float* blob);
efie_status efie_sur_set(efie_ctx* ctx, const float* blob);

/* Diagnostic: with on != 0, near_mode SURROGATE applies the exact
 * correction (the label quadrature run on the mesh's own pairs) instead of
 * the network, so the result equals near_mode DIRECT up to rounding; it
 * checks the frame transform and the correction path. */
efie_status efie_sur_set_exact(efie_ctx* ctx, int on);

/* Network kernels on plain arrays (host pointers): X [N][13], Y and T
 * [N][16], params [efie_mlp_params()] in the layout
 * W0[64][13] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]. */
int efie_mlp_params(void);
efie_status efie_mlp_forward(efie_ctx* ctx, const float* params, const float* X, int N, float* Y);
/* loss = sum((y - t)^2) / (16 N) and its gradient with respect to params. */
efie_status efie_mlp_loss_grad(efie_ctx* ctx, const float* params, const float* X, const float* T, int N,
                               float* loss, float* grad);

#ifdef __cplusplus
}
#endif
#endif

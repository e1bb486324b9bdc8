/*
 * sunhip_learn.h -- learned multi-site proposals for the finite-temperature sweep (libsunhip.so).
 *
 * The reference has no counterpart: its only moves are the local single-site updates of
 * src/dqmc_config_u/dqmc_update*.f90 (the lstglobal/lworm switches declared at
 * src/model_para.f90:84-85 are never used).  These entry points add a self-learning proposal on top of
 * the same sweep, for callers that want faster decorrelation than the local update gives:
 *
 *   effective action   E_eff(C) = Ns (ymu + ysd g(x(C))), g a 4 -> 128 -> 128 -> 1 tanh MLP
 *                      (FP32 weights), Ns = ndim * ltrot HS spins s = 2 conf - 3 (lcomp = 2)
 *   features           x_k = (f_k - xmu_k) / xsd_k of the global couplings of the field:
 *                        f0 = sum s_{i,t} / Ns                          (on-site)
 *                        f1 = sum_<ij>,t s_{i,t} s_{j,t} / (Nb ltrot)   (space neighbours)
 *                        f2 = sum_i,t s_{i,t} s_{i,t+1} / Ns            (time neighbours)
 *                        f3 = sum_i,t s_{i,t} sum_{j~i} s_{j,t+1} / (2 Nb ltrot)
 *                      with Nb = (sum of neighbour-list lengths) / 2, t periodic in ltrot
 *   proposal           after the local update of each slice, a Metropolis chain on E_eff of
 *                      nprop single-spin attempts at uniformly drawn sites of that slice flips
 *                      a set F (|F| <= nprop <= 64); the move is accepted with probability
 *                      min(1, |Re(R phase)| / |Re(phase)| * exp(E_eff(C') - E_eff(C))),
 *                      R = det(1 + Delta_F (1 - G)_FF)^nflr * prod of the field factors: the
 *                      exact fermion ratio of the whole block, so detailed balance holds for
 *                      any network and the local update remains the fallback.
 *   training           mean squared error of g against (E(C)/Ns - ymu) / ysd on samples
 *                      E(C) = -(nflr Re logweightf + sum log gaml(conf)) taken from the sweep's
 *                      own configurations, Adam, minibatches, FP32 kernels.
 *
 * Return codes and error handling follow sunhip.h (SH_E_*).
 */
#ifndef SUNHIP_LEARN_H
#define SUNHIP_LEARN_H
#include <stdint.h>
#include "sunhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SHL_HIDDEN 128
#define SHL_NFEAT 4
#define SHL_MAXBLOCK 64
/* weights, float32, in this order: W1[128][4], b1[128], W2[128][128] (row = output),
 * b2[128], w3[128], b3 */
#define SHL_NWEIGHTS (SHL_HIDDEN * SHL_HIDDEN + 7 * SHL_HIDDEN + 1)

typedef struct shl_net shl_net;

/* A network with Glorot-uniform weights drawn from seed and room for `capacity` samples. */
int shl_create(shl_net **net, int capacity, int32_t seed, int device);
int shl_destroy(shl_net *net);
const char *shl_last_error(const shl_net *net);
int shl_get_weights(shl_net *net, float *w);
int shl_set_weights(shl_net *net, const float *w);
/* norm[10] = xmu[4], xsd[4], ymu, ysd.  Unset norms are taken from the samples by the first
 * shl_train; shl_get_norm returns SH_E_STATE while unset. */
int shl_set_norm(shl_net *net, const double *norm);
int shl_get_norm(shl_net *net, double *norm);
/* Append the current configuration of every chain of ctx as one sample each (features from
 * the device-resident HS field, energy from logweightf of the last stabilization, so call it
 * right after a completed pass).  Needs sh_set_neighbours. */
int shl_add_samples(shl_net *net, sh_ctx *ctx);
/* Append host samples x[n][5] = (f0, f1, f2, f3, E/Ns). */
int shl_push_samples(shl_net *net, const double *x, int n);
int shl_num_samples(const shl_net *net);
int shl_get_samples(shl_net *net, double *x, int first, int n);
int shl_clear_samples(shl_net *net);
/* epochs over the stored samples in shuffled minibatches; loss[epochs] gets each epoch's
 * mean training loss (may be NULL) */
int shl_train(shl_net *net, int epochs, int batch, double lr, double *loss);
/* mean loss and its gradient (float[SHL_NWEIGHTS]) over host samples x[n][5], current norm */
int shl_loss_grad(shl_net *net, const double *x, int n, double *loss, float *grad);
/* E_eff / Ns = ymu + ysd g(x) for host features f[n][4] */
int shl_predict(shl_net *net, const double *f, int n, double *e);
/* shl_loss_grad without any allocation, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight training or serving loops).  It has no counterpart in the
 * reference.  x[n][5] (double), loss[1] (double) and grad[SHL_NWEIGHTS] (float) are DEVICE
 * pointers owned by the caller; the results are those of shl_loss_grad on the same samples.
 * The partial sums live in a fixed 4.4 MB device array that the library holds per GPU until
 * the process ends (set up by shl_create); n <= 16777216 (SH_E_ARG beyond).  Calls that overlap
 * on one GPU share that array and race.  Returns after the results are written. */
int shl_loss_grad_noalloc(shl_net *net, const double *x, int n, double *loss, float *grad);

/* Neighbour lists of the lattice sites, int32 [sum over chains of ndim][z], 0-based, symmetric
 * (j in list(i) as often as i in list(j)); z the same for every chain.  Switches block
 * proposals off until the next sh_set_proposal. */
int sh_set_neighbours(sh_ctx *ctx, int z, const int32_t *nbr);
/* Features of every chain's current field: out[nchains][5] = f0..f3, E/Ns. */
int sh_features(sh_ctx *ctx, double *out);
/* Enable block proposals (net != NULL) with nprop attempts per slice (1..64) or disable them
 * (net == NULL).  The weights and norm are copied: later training needs another call.
 * local_updates = 0 turns the local single-site update off (block moves only).  Needs
 * lcomp = 2, ltrot >= 2 and sh_set_neighbours. */
int sh_set_proposal(sh_ctx *ctx, const shl_net *net, int nprop, int local_updates);
/* stats[5]: block moves tried, accepted, spins in proposed blocks, spins in accepted blocks,
 * inner Metropolis attempts accepted */
int sh_get_proposal_stats(sh_ctx *ctx, int chain, double *stats);
int sh_reset_proposal_stats(sh_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif

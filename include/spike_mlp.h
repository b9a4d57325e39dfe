/*
 * spike_mlp.h -- C ABI of the sweep surrogate: a small MLP trained and evaluated on the
 * MI355X with hand-written gfx950 kernels (BF16 MFMA 16x16x32, FP32 accumulation).
 *
 * The surrogate has no counterpart in the reference (Spike has no learned components);
 * it maps Brunel sweep parameters to the network statistics that spk_stats_get reports
 * and is used to screen parameter sweeps before they are simulated.
 *
 * Model (one per ensemble member, all members trained side by side in every launch):
 *   5 input features -> 4 hidden layers of 256 SiLU units -> 3 outputs.
 *   Parameters, one flat float32 vector of SPK_MLP_PARAMS values per member, in order:
 *     W1 [256][5], b1 [256], W2 [256][256], b2 [256], W3 [256][256], b3 [256],
 *     W4 [256][256], b4 [256], W5 [3][256], b5 [3]          (W[out][in], row-major)
 *   Loss per member: (1/M) sum_rows sum_o w_o (y_o - t_o)^2 over a batch of M rows.
 *   Optimiser: Adam (beta1 0.9, beta2 0.999, eps 1e-8) on FP32 master weights, cosine
 *   learning-rate schedule; the kernels read BF16 copies written by the optimiser.
 * Features and targets are whatever the caller normalised them to; the library does not
 * transform them.  All pointers are host memory; sizes are element counts.
 * Status codes and spk_last_error are those of spike_hip.h.
 */
#ifndef SPIKE_MLP_H
#define SPIKE_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPK_MLP_INPUTS 5
#define SPK_MLP_OUTPUTS 3
#define SPK_MLP_HIDDEN 256
#define SPK_MLP_PARAMS 199683
#define SPK_MLP_ROW_TILE 128 /* training batches are multiples of this */

typedef struct spk_mlp spk_mlp;

/* sampling modes of spk_mlp_train */
enum {
  SPK_MLP_BATCH_FIRST_ROWS = 0, /* every step uses data rows [0, batch) */
  SPK_MLP_BATCH_SAMPLED = 1     /* rows drawn uniformly with replacement, Philox per member */
};

/* Device timings of the last call (milliseconds, HIP events on the surrogate's stream). */
typedef struct {
  float total_ms;    /* the whole call's device work */
  float forward_ms;  /* forward kernel(s) */
  float backward_ms; /* backward kernel(s) */
  float wgrad_ms;    /* weight-gradient kernel(s) */
  float update_ms;   /* gradient reduction + Adam */
  int64_t steps;     /* training steps (0 for inference) */
  int64_t rows;      /* rows evaluated per member */
} spk_mlp_timing;

int spk_mlp_param_count(void);

/* models: ensemble members (1..64); max_batch: largest training batch (rounded up to a
 * multiple of SPK_MLP_ROW_TILE). Parameters start at zero: call spk_mlp_set_params. */
int spk_mlp_create(int models, int max_batch, spk_mlp** out);
int spk_mlp_destroy(spk_mlp* m);

/* [models][SPK_MLP_PARAMS] float32; set also resets the Adam state and step counter. */
int spk_mlp_set_params(spk_mlp* m, const float* params);
int spk_mlp_get_params(spk_mlp* m, float* params);
int spk_mlp_reset_optimizer(spk_mlp* m);

/* Training set, kept resident in device memory: features [n][5], targets [n][3]. */
int spk_mlp_set_data(spk_mlp* m, const float* features, const float* targets, int64_t n);
/* Per-output loss weights w_o (default 1, 1, 1). */
int spk_mlp_set_loss_weights(spk_mlp* m, const float* w3);

/* Loss and gradient of every member on data rows [0, rows) (rows a multiple of
 * SPK_MLP_ROW_TILE, at most max_batch and n); no update. grads [models][SPK_MLP_PARAMS]
 * (may be NULL), loss [models] (may be NULL). */
int spk_mlp_gradients(spk_mlp* m, int rows, float* grads, float* loss);

/* One Adam step with caller-given gradients [models][SPK_MLP_PARAMS] at learning rate lr. */
int spk_mlp_apply_gradients(spk_mlp* m, const float* grads, float lr);

/* `steps` Adam steps on batches of `batch` rows (mode SPK_MLP_BATCH_*; `seed` keys the
 * sampling).  The learning rate follows lr_min + (lr0 - lr_min)(1 + cos(pi u))/2 with
 * u = min(t - 1, total_steps)/total_steps at the optimiser's step t = 1, 2, ... (counted
 * across calls since the last reset).  loss_per_step [steps][models] may be NULL.
 * use_graph != 0 replays one captured step as a HIP graph. */
int spk_mlp_train(spk_mlp* m, int steps, int batch, int mode, uint64_t seed, float lr0,
                  float lr_min, int total_steps, int use_graph, float* loss_per_step);

/* Outputs of every member for n rows: features [n][5] -> out [models][n][3]. */
int spk_mlp_predict(spk_mlp* m, const float* features, int64_t n, float* out);

/* spk_mlp_predict without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory -- tight screening loops that serve
 * predictions, code that must not call hipMalloc.  spk_mlp_predict grows two device
 * staging buffers on first use; this variant stages features and outputs through one
 * fixed 4 MiB device array per GPU, held until the process ends, and calls no hipMalloc,
 * hipFree or other allocating routine.  Limit: round_up(5 n, 64) + 3 n models floats must
 * fit in 4 MiB (1,048,576 floats; e.g. n <= 36,157 rows for 8 members), else SPK_ERR_ARG.
 * Every call on one GPU uses that GPU's one array: calls that overlap in time on the same
 * GPU (from several threads, on any handles) race and must be serialised by the caller.
 * Results equal spk_mlp_predict's bit for bit. */
int spk_mlp_predict_noalloc(spk_mlp* m, const float* features, int64_t n, float* out);

int spk_mlp_last_timing(const spk_mlp* m, spk_mlp_timing* t);

#ifdef __cplusplus
}
#endif

#endif

/*
 * song_amd learned form -- C ABI.  No counterpart in the reference (it always starts
 * its search at vertex 0, graph.h:171-181 / warp_astar_accelerator.h:166-176); this is
 * the entry-vertex and stop-step predictor that the learned search uses.
 *
 * A 3-layer MLP  in_dim -> hidden -> hidden -> n_classes + 1, ReLU, BF16 weights and
 * activations on the matrix cores (MFMA, FP32 accumulation), FP32 master weights
 * and Adam state.  Outputs: logits over n_classes entry pivots plus one stop-step
 * regression output.  Training loss: softmax cross-entropy on the pivot label plus
 * stop_weight * squared error on the stop target.
 */
#ifndef SONG_LEARNED_H
#define SONG_LEARNED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct song_mlp song_mlp;

/* Create with He-uniform initial weights from `seed`; in_dim is padded to 16. */
int song_mlp_create(int in_dim, int hidden, int n_classes, int max_batch, uint64_t seed, song_mlp **out);
int song_mlp_destroy(song_mlp *mlp);
/* padded input width (columns of the bf16 input matrix) */
int song_mlp_input_width(const song_mlp *mlp);

/* Parameters as float32 host arrays in the order W1[h x in], b1[h], W2[h x h], b2[h],
 * W3[(c+1) x h], b3[c+1] (in = padded width).  count = total floats. */
int song_mlp_param_count(const song_mlp *mlp, uint64_t *count);
int song_mlp_get_params(const song_mlp *mlp, float *host);
int song_mlp_set_params(song_mlp *mlp, const float *host);

/*
 * Loss and gradients of one batch (no update).  d_x: batch x width bf16 (device);
 * d_labels: batch int32 (pivot index); d_stop: batch float32 (stop target).
 * loss_out (host, 2 floats): mean cross-entropy, mean squared stop error.
 * grads_host (nullable): gradients in the get_params layout.
 */
int song_mlp_loss_grad(song_mlp *mlp, const uint16_t *d_x, const int *d_labels, const float *d_stop, int batch,
                       float stop_weight, float *loss_out, float *grads_host, void *stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the gradients of the last loss_grad. */
int song_mlp_adam_step(song_mlp *mlp, float lr, void *stream);

/*
 * Fused inference (one kernel, 64 rows per workgroup): per row the two most likely
 * pivots (d_top2, batch x 2, best first, ties to the lower index), their summed softmax
 * probability (d_conf, the confidence) and the stop output (d_stop).
 * d_logits (nullable): batch x (c+1) float32 logits.
 */
int song_mlp_infer(song_mlp *mlp, const uint16_t *d_x, int batch, int *d_top2, float *d_conf, float *d_stop,
                   float *d_logits, void *stream);

/*
 * As song_mlp_infer, with the confidence taken as the softmax mass of the best pivot's
 * neighbourhood d_nbr[best * n_nbr .. + n_nbr) (c x n_nbr pivot indices, each row
 * starting with its own pivot) instead of the two best pivots.
 */
int song_mlp_infer_nbr(song_mlp *mlp, const uint16_t *d_x, int batch, const int *d_nbr, int n_nbr, int *d_top2,
                       float *d_conf, float *d_stop, float *d_logits, void *stream);

/*
 * Input features of the learned search, batch of queries (device, nq x dim float32):
 * columns [0, dim): (q - mean) * inv_std per coordinate; then 8 query-state features
 * scaled by fscale: |q|, distance to the default entry row v0, mean, std, min, max,
 * fraction of zero coordinates, 1.  Zero padding to the MLP input width.
 */
int song_mlp_features(const float *d_queries, int nq, int dim, const float *d_mean, const float *d_inv_std,
                      const float *d_v0, const float *fscale_host, uint16_t *d_x, int width, void *stream);

/*
 * Search plan of a learned batch from song_mlp_infer_nbr's outputs, one thread per query.
 * Query i is confident when d_conf[i] >= conf_threshold and its best pivot (d_top2[2i]) is one of
 * the n_pivots: d_pq[i] = pq_learned, its d_entry row
 * (n_nbr ints) = d_pivot_vertex[d_nbr[top1 * n_nbr + e]], d_caps[i] = (int) of the float32
 * clamp(step_scale * (2^d_stop[i] - 1) + step_add, 1, 1e6).  Otherwise d_pq[i] = pq_fallback,
 * the entry row is -1 (vertex 0) and d_caps[i] = fallback_cap (1e6 when 0).  d_ok[i] = 1 or 0.
 */
int song_mlp_plan(const int *d_top2, const float *d_conf, const float *d_stop, int batch, const int *d_nbr, int n_nbr,
                  const int *d_pivot_vertex, int n_pivots, float conf_threshold, float step_scale, float step_add,
                  int fallback_cap, int pq_learned, int pq_fallback, int *d_pq, int *d_entry, int *d_caps,
                  unsigned char *d_ok, void *stream);

/* Generic C = alpha * A . B (+ bias) on bf16 inputs, float32 output (test hook for the GEMM). */
int song_gemm_bf16(int M, int N, int K, const uint16_t *d_a, int sam, int sak, const uint16_t *d_b, int sbk, int sbn,
                   float *d_c, int ldc, float alpha, void *stream);

const char *song_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

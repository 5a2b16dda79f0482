/*
 * sdem_gnn.h -- C ABI of the learned surrogate in libsdem_hip.so.
 *
 * No counterpart in the reference: a graph network that advances a DEM
 * state by `nsub` time steps per call, trained on trajectories of the plain
 * DEM path.  Graph: every pair of particles i, j with |x_i - x_j| <=
 * cutoff_factor * 2 max(r_i, r_j) (both directions, nearest periodic image;
 * `radius` must bound that distance).  Network: encoders (node, edge), then
 * `layers` message-passing layers (edge update e += MLP([e, h_src, h_dst]),
 * node update h += MLP([h, sum of incoming e])), then a decoder; every MLP is
 * Linear-SiLU-Linear-SiLU-Linear with `hidden` units.  bf16 weights on the
 * matrix cores with fp32 accumulation; fp32 master weights and Adam.
 *
 * Node features (node_in = 8): v / vel_scale (3), linAcc(1) / acc_scale (3),
 * radius / len_scale, min(height over floor_z, 4 len_scale) / len_scale.
 * Edge features (edge_in = 7), sender j to receiver i: (x_j - x_i) /
 * len_scale (3), |x_j - x_i| / len_scale, (v_j - v_i) / vel_scale (3).
 * Output (out = 6): corrections to the ballistic step over T = nsub * dt,
 *   x' = x + v T + g T^2 / 2 + out[0:3] * dx_scale,
 *   v' = v + g T + out[3:6] * dv_scale.
 * Same conventions as sdem.h: 0 or a negative SDEM_E_* code, caller-owned
 * host arrays copied before return.
 */
#ifndef SDEM_GNN_H
#define SDEM_GNN_H

#include <stdint.h>

#include "sdem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdem_gnn_config {
  double vel_scale, acc_scale, len_scale;
  double radius;         /* largest cut-off: the graph search's cell edge */
  double cutoff_factor;  /* edge i-j when |x_i - x_j| <= cutoff_factor * 2 max(r_i, r_j) */
  double floor_z, dx_scale, dv_scale;
  int32_t hidden, layers, node_in, edge_in, out, nsub;
} sdem_gnn_config;

typedef struct sdem_gnn sdem_gnn;

int sdem_gnn_config_default(sdem_gnn_config* c);
int sdem_gnn_create(const sdem_gnn_config* c, uint64_t seed, sdem_gnn** out);
int sdem_gnn_get_config(sdem_gnn* m, sdem_gnn_config* c);
/* replace the scales and the graph radius (the network shape stays) */
int sdem_gnn_set_scales(sdem_gnn* m, const sdem_gnn_config* c);
int sdem_gnn_num_params(sdem_gnn* m, int64_t* n);
int sdem_gnn_get_params(sdem_gnn* m, float* w);
int sdem_gnn_set_params(sdem_gnn* m, const float* w);
int sdem_gnn_get_grads(sdem_gnn* m, float* g);
/* parameter layout: per MLP the offsets of W1, b1, W2, b2, W3, b3 (row-major
 * [in x hid], [hid], [hid x hid], [hid], [hid x out], [out]) and (in, hid,
 * out); MLPs in the order node encoder, edge encoder, (edge, node) per
 * layer, decoder.  Encoders read 16 padded input columns, the decoder
 * writes 16 columns of which the first `out` are used. */
int sdem_gnn_get_layout(sdem_gnn* m, int32_t nmlp_cap, int32_t* nmlp, int64_t* offsets, int32_t* dims);

/* graph from host arrays: node features n x node_in, edges (src, dst) sorted
 * by dst (receiver), edge features ne x edge_in */
int sdem_gnn_set_graph(sdem_gnn* m, int32_t n, const float* node_feat, int64_t ne, const int32_t* src,
                       const int32_t* dst, const float* edge_feat);
/* graph and features from the state of a DEM context (its first box);
 * nodes in the context's storage order, whose particle ids
 * sdem_gnn_get_node_ids gives.  for_training = 0 keeps only one layer's
 * activations (inference: no backward pass on this graph). */
int sdem_gnn_graph_from_dem(sdem_gnn* m, sdem_ctx* dem, int32_t for_training, int64_t* nedges);
int sdem_gnn_get_node_ids(sdem_gnn* m, int32_t cap, int32_t* ids);
int sdem_gnn_get_num_nodes(sdem_gnn* m, int32_t* n, int64_t* ne);
int sdem_gnn_get_graph(sdem_gnn* m, int64_t cap, int32_t* src, int32_t* dst, float* node_feat, float* edge_feat);

int sdem_gnn_forward(sdem_gnn* m, float* out /* n x out, may be null */);
/* gradients of sum(out * dout) with respect to the parameters */
int sdem_gnn_backward(sdem_gnn* m, const float* dout /* n x out */);
/* forward, mean squared error against target (n x out), backward, gradient
 * clipping at `clip` (<= 0: none), one Adam step with learning rate lr */
int sdem_gnn_train_step(sdem_gnn* m, const float* target, float lr, float clip, float* loss);

/* `calls` surrogate calls on a DEM context: each advances its state by nsub
 * steps on the device; the context then goes on stepping from there, with
 * iterNumber advanced and its contact history cleared.  A graph of more than
 * SDEM_GNN_MAX_EDGES_PER_NODE edges per particle (a state the surrogate has
 * let collapse) is refused with SDEM_E_STATE; the context then holds the
 * state after the calls that completed, for the plain DEM to take over. */
#define SDEM_GNN_MAX_EDGES_PER_NODE 48
int sdem_gnn_advance(sdem_gnn* m, sdem_ctx* dem, int32_t calls);

/* checks of the matrix-core building blocks: one v_mfma_f32_32x32x16_bf16
 * on 64 lanes (a, b: 64 x 8 bf16 bit patterns, d: 64 x 16 floats), and one
 * GEMM C = op(A) op(B) (+bias, SiLU) with inputs rounded to bf16 (flags:
 * 1 bias, 2 SiLU, 4 return the pre-activation; splits > 1: split-K) */
int sdem_gnn_mfma_probe(const uint16_t* a_bits, const uint16_t* b_bits, float* d);
int sdem_gnn_gemm(int32_t ta, int32_t tb, int32_t m, int32_t n, int32_t k, const float* a, const float* b,
                  const float* bias, int32_t flags, int32_t splits, float* c, float* pre);

/* sdem_gnn_gemm without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  a (m x k, or k x m with ta), b (k x n, or
 * n x k with tb), bias (n) and c (m x n) are device buffers of fp32,
 * row-major; pre (m x n, may be NULL) gets the pre-activation when splits
 * resolve to one K chunk.  The work is queued on `stream` (a hipStream_t;
 * NULL = the null stream) and the call returns without waiting for it.  The
 * bf16 copies of a and b and the split-K partial sums live in one fixed
 * array of SDEM_GNN_GEMM_SCRATCH_BYTES on each device, which the library
 * holds until the process ends: calls whose staging (2 m k + 2 k n bytes,
 * each rounded up to 256, plus 4 m n per K chunk when there is more than
 * one) does not fit return SDEM_E_CAPACITY.  Every call on a device shares
 * that device's array: calls that overlap on one device (other streams,
 * other threads) race.  Gives the bits of sdem_gnn_gemm. */
#define SDEM_GNN_GEMM_SCRATCH_BYTES (16 << 20)
int sdem_gnn_gemm_noalloc(int32_t ta, int32_t tb, int32_t m, int32_t n, int32_t k, const float* a, const float* b,
                          const float* bias, int32_t flags, int32_t splits, float* c, float* pre, void* stream);

const char* sdem_gnn_last_error(const sdem_gnn* m);
void sdem_gnn_destroy(sdem_gnn* m);

#ifdef __cplusplus
}
#endif

#endif /* SDEM_GNN_H */

// gnn_api.cpp -- C ABI of the learned surrogate (include/sdem_gnn.h).
//
// A graph network (encoder, L message-passing layers of edge and node MLPs
// with residual updates, decoder) over the neighbour graph of the particles.
// Every MLP is Linear-SiLU-Linear-SiLU-Linear with H hidden units; weights
// live as fp32 masters with bf16 copies that the matrix-core GEMMs read,
// accumulation is fp32.  The forward pass keeps every activation the
// backward pass needs; the backward pass is written out layer by layer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sdem.h"
#include "../../include/sdem_gnn.h"
#include "gnn_internal.h"
#include "sdem_host.h"

using namespace gnn;

namespace {

constexpr int PAD_IN = 16;  // encoder input width (features padded)

struct MLP {
  int in, out, hid;
  size_t w1, b1, w2, b2, w3, b3;  // offsets into the parameter blob
};

// activations of one MLP application to R rows
struct Act {
  int R = 0;
  __bf16* x = nullptr;  // input, R x in (owned by the caller's buffer)
  float *pre1 = nullptr, *pre2 = nullptr;
  __bf16 *a1 = nullptr, *a2 = nullptr;
  float* out = nullptr;  // R x out
};

size_t align64(size_t n) { return (n + 63) / 64 * 64; }

}  // namespace

struct sdem_gnn {
  sdem_gnn_config cfg{};
  std::string err;
  hipStream_t stream = nullptr;
  std::vector<MLP> mlp;  // 0 enc_node, 1 enc_edge, 2+2l edge_l, 3+2l node_l, last dec
  size_t nparam = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  __bf16* wb = nullptr;
  __bf16* wbt = nullptr;   // every weight matrix transposed ([out][in]), for the fused inference MLPs
  bool wt_stale = true;    // wbt lags behind wb
  int64_t adam_t = 0;
  // graph
  int n = 0;
  int64_t ne = 0;
  int *esrc = nullptr, *edst = nullptr, *rstart = nullptr;
  __bf16 *fn = nullptr, *fe = nullptr;  // encoder inputs
  // activations
  std::vector<Act> act;                  // one per MLP application
  std::vector<float*> h, e;              // residual streams per layer (L+1 each)
  std::vector<__bf16*> xe, xn;           // edge / node MLP inputs per layer
  float* agg = nullptr;
  // scratch for the backward pass
  float *dout = nullptr, *tmp_a = nullptr, *tmp_b = nullptr, *dx_e = nullptr, *dx_n = nullptr;
  float *dh = nullptr, *dh2 = nullptr, *de = nullptr, *de2 = nullptr, *dagg = nullptr, *part = nullptr;
  __bf16 *db16 = nullptr, *db16b = nullptr;
  size_t part_len = 0;
  float* loss = nullptr;
  // graph building scratch
  // graph search scratch: cell of each particle, particles in cell order,
  // per-cell counts / fill cursors / ranges, degrees, scan tile sums
  int *gkey = nullptr, *gsorted = nullptr, *gcnt = nullptr, *gfill = nullptr, *cs = nullptr, *ce = nullptr;
  int *deg = nullptr, *tsum = nullptr;
  int64_t ncell_alloc = 0;
  int search_n = 0;
  int gcap_n = 0;
  int64_t gcap_e = 0;
  bool train_alloc = false;  // activations of every layer kept (training)
  std::vector<int> ids;      // particle ids of the nodes of a graph taken from a DEM state
  int ids_cap = 0;
  std::vector<void*> graph_allocs;
  std::vector<void*> allocs;
  int fail(int code, const std::string& s) {
    err = s;
    return code;
  }
};

namespace {

int hipf(sdem_gnn* m, hipError_t e, const char* where) {
  return m->fail(SDEM_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

template <class T>
hipError_t galloc(std::vector<void*>& L, T** p, size_t n) {
  hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) L.push_back(*p);
  return e;
}

void free_all(std::vector<void*>& L) {
  for (void* p : L) (void)hipFree(p);
  L.clear();
}

// splitmix64 for the initial weights
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

MLP make_mlp(size_t& off, int in, int hid, int out) {
  MLP p;
  p.in = in; p.out = out; p.hid = hid;
  p.w1 = off; off = align64(off + (size_t)in * hid);
  p.b1 = off; off = align64(off + hid);
  p.w2 = off; off = align64(off + (size_t)hid * hid);
  p.b2 = off; off = align64(off + hid);
  p.w3 = off; off = align64(off + (size_t)hid * out);
  p.b3 = off; off = align64(off + out);
  return p;
}

// one GEMM with the common arguments
hipError_t mm(sdem_gnn* M, bool ta, bool tb, const __bf16* A, const __bf16* B, int m, int n, int k, int lda,
              int ldb, const Epi& ep, int splits = 1) {
  return gemm(ta, tb, A, B, m, n, k, lda, ldb, splits, ep, M->stream);
}

// forward of one MLP on act.x (R x in) into y (R x out); y (+)= resid if
// given (y may be resid itself: every element is read before it is written,
// by the same thread)
hipError_t mlp_fwd(sdem_gnn* M, const MLP& p, Act& a, const float* resid, float* y) {
  const __bf16* W = M->wb;
  const float* Wf = M->w;
  Epi e1{EPI_BIAS | EPI_PRE | EPI_SILU, Wf + p.b1, a.pre1, nullptr, nullptr, a.a1};
  hipError_t e = mm(M, false, false, a.x, W + p.w1, a.R, p.hid, p.in, p.in, p.hid, e1);
  Epi e2{EPI_BIAS | EPI_PRE | EPI_SILU, Wf + p.b2, a.pre2, nullptr, nullptr, a.a2};
  if (e == hipSuccess) e = mm(M, false, false, a.a1, W + p.w2, a.R, p.hid, p.hid, p.hid, p.hid, e2);
  Epi e3{EPI_BIAS | (resid ? EPI_RESID : 0), Wf + p.b3, nullptr, resid, y, nullptr};
  if (e == hipSuccess) e = mm(M, false, false, a.a2, W + p.w3, a.R, p.out, p.hid, p.hid, p.out, e3);
  return e;
}

// dW (+)= X^T dY with X: R x kin (bf16), dY: R x nout (bf16); split over R
hipError_t weight_grad(sdem_gnn* M, const __bf16* X, const __bf16* dY, int R, int kin, int nout, float* gW) {
  const int splits = std::max(1, std::min(64, R / 4096));
  const size_t len = (size_t)kin * nout;
  Epi ep{0, nullptr, nullptr, nullptr, M->part, nullptr};
  hipError_t e = mm(M, true, false, X, dY, kin, nout, R, kin, nout, ep, splits);
  // the launcher may merge chunks: recompute how many partials exist
  int kchunk = (R + splits - 1) / splits;
  kchunk = (kchunk + 31) / 32 * 32;
  const int nz = std::max(1, (R + kchunk - 1) / kchunk);
  if (e == hipSuccess) e = reduce_parts(M->part, nz, len, gW, true, M->stream);
  return e;
}

// backward of one MLP: dY (R x out, fp32) -> weight gradients, and dX
// (R x in, fp32) when dx is set
hipError_t mlp_bwd(sdem_gnn* M, const MLP& p, const Act& a, const float* dy, float* dx) {
  const __bf16* W = M->wb;
  float* G = M->g;
  const int R = a.R, H = p.hid;
  hipError_t e = to_bf16(dy, (size_t)R * p.out, M->db16, M->stream);
  if (e == hipSuccess) e = weight_grad(M, a.a2, M->db16, R, H, p.out, G + p.w3);
  if (e == hipSuccess) e = colsum(dy, R, p.out, G + p.b3, M->stream);
  // dA2 = dY W3^T
  Epi ea{0, nullptr, nullptr, nullptr, M->tmp_a, nullptr};
  if (e == hipSuccess) e = mm(M, false, true, M->db16, W + p.w3, R, H, p.out, p.out, p.out, ea);
  if (e == hipSuccess) e = silu_bwd(M->tmp_a, a.pre2, (size_t)R * H, M->tmp_b, M->db16b, M->stream);
  if (e == hipSuccess) e = weight_grad(M, a.a1, M->db16b, R, H, H, G + p.w2);
  if (e == hipSuccess) e = colsum(M->tmp_b, R, H, G + p.b2, M->stream);
  // dA1 = dPre2 W2^T
  if (e == hipSuccess) e = mm(M, false, true, M->db16b, W + p.w2, R, H, H, H, H, ea);
  if (e == hipSuccess) e = silu_bwd(M->tmp_a, a.pre1, (size_t)R * H, M->tmp_b, M->db16b, M->stream);
  if (e == hipSuccess) e = weight_grad(M, a.x, M->db16b, R, p.in, H, G + p.w1);
  if (e == hipSuccess) e = colsum(M->tmp_b, R, H, G + p.b1, M->stream);
  if (dx && e == hipSuccess) {
    Epi ex{0, nullptr, nullptr, nullptr, dx, nullptr};
    e = mm(M, false, true, M->db16b, W + p.w1, R, p.in, H, H, H, ex);
  }
  return e;
}

int L_of(const sdem_gnn* M) { return M->cfg.layers; }

// Inference with 128 hidden units runs the message-passing MLPs as one
// fused kernel each (same bits as the three GEMMs, no activations in HBM).
bool fused_inference(const sdem_gnn* M, bool train) { return !train && M->cfg.hidden == FUSED_H; }

// the transposed bf16 weights of every MLP, after the weights changed
hipError_t refresh_transposes(sdem_gnn* M) {
  if (!M->wt_stale) return hipSuccess;
  hipError_t e = hipSuccess;
  for (const MLP& p : M->mlp) {
    if (e == hipSuccess) e = transpose_bf16(M->wb + p.w1, p.in, p.hid, M->wbt + p.w1, M->stream);
    if (e == hipSuccess) e = transpose_bf16(M->wb + p.w2, p.hid, p.hid, M->wbt + p.w2, M->stream);
    if (e == hipSuccess) e = transpose_bf16(M->wb + p.w3, p.hid, p.out, M->wbt + p.w3, M->stream);
  }
  if (e == hipSuccess) M->wt_stale = false;
  return e;
}

FusedMLP fused_args(const sdem_gnn* M, const MLP& p, int R, const float* resid, float* y) {
  FusedMLP f{};
  f.R = R;
  f.w1t = M->wbt + p.w1;
  f.w2t = M->wbt + p.w2;
  f.w3t = M->wbt + p.w3;
  f.b1 = M->w + p.b1;
  f.b2 = M->w + p.b2;
  f.b3 = M->w + p.b3;
  f.resid = resid;
  f.y = y;
  return f;
}

// Activation buffers for n nodes and ne edges.  Training keeps every
// layer's activations for the backward pass; inference keeps one layer's
// (residual streams ping-pong, the MLP buffers are shared by all layers).
int alloc_graph(sdem_gnn* M, int n, int64_t ne, bool train) {
  if (n <= M->gcap_n && ne <= M->gcap_e && train == M->train_alloc) return SDEM_OK;
  free_all(M->graph_allocs);
  M->gcap_n = 0;
  M->gcap_e = 0;
  const int H = M->cfg.hidden, L = L_of(M);
  // room to grow: graphs of a settling bed gain edges call after call
  const size_t N = (size_t)std::max(n, 1), E = (size_t)std::max<int64_t>(ne + ne / 4, 1);
  auto& A = M->graph_allocs;
  hipError_t e = hipSuccess;
  auto G = [&](auto** p, size_t k) { if (e == hipSuccess) e = galloc(A, p, k); };
  G(&M->esrc, E); G(&M->edst, E); G(&M->rstart, N + 1);
  G(&M->fn, N * PAD_IN); G(&M->fe, E * PAD_IN);
  M->act.assign(M->mlp.size(), Act{});
  const bool fused = fused_inference(M, train);
  // MLP applications: 0 enc_node (N), 1 enc_edge (E), per layer edge (E), node (N), dec (N)
  for (size_t q = 0; q < M->mlp.size(); q++) {
    const MLP& p = M->mlp[q];
    const bool layer = q >= 2 && q < 2 + 2 * (size_t)L;
    const bool on_edges = q == 1 || (layer && (q - 2) % 2 == 0);
    Act& a = M->act[q];
    if (fused && layer) continue;     // the fused kernel keeps its activations on chip
    if (!train && layer && q >= 4) {  // share the first layer's buffers
      a = M->act[q - 2];
      continue;
    }
    const size_t R = on_edges ? E : N;
    a.R = (int)R;
    G(&a.pre1, R * p.hid); G(&a.pre2, R * p.hid); G(&a.a1, R * p.hid); G(&a.a2, R * p.hid);
    G(&a.out, R * p.out);
  }
  const int ns = train ? L + 1 : 2, nx = train ? L : (fused ? 0 : 1);
  M->h.assign(ns, nullptr);
  M->e.assign(ns, nullptr);
  M->xe.assign(nx, nullptr);
  M->xn.assign(nx, nullptr);
  for (int l = 0; l < ns; l++) { G(&M->h[l], N * H); G(&M->e[l], E * H); }
  for (int l = 0; l < nx; l++) { G(&M->xe[l], E * 3 * H); G(&M->xn[l], N * 2 * H); }
  G(&M->agg, N * H);
  G(&M->dout, N * PAD_IN);
  if (train) {
    const size_t R = std::max(N, E);
    G(&M->tmp_a, R * 3 * H); G(&M->tmp_b, R * 3 * H);
    G(&M->dx_e, E * 3 * H); G(&M->dx_n, N * 2 * H);
    G(&M->dh, N * H); G(&M->dh2, N * H); G(&M->de, E * H); G(&M->de2, E * H); G(&M->dagg, N * H);
    G(&M->db16, R * 3 * H); G(&M->db16b, R * 3 * H);
    M->part_len = (size_t)64 * 3 * H * H;
    G(&M->part, M->part_len);
  } else {
    G(&M->db16, N * H);
  }
  if (e != hipSuccess) return hipf(M, e, "hipMalloc(graph activations)");
  M->gcap_n = (int)N;
  M->gcap_e = (int64_t)E;
  M->train_alloc = train;
  return SDEM_OK;
}

// forward pass over the current graph; prediction in act.back().out
int forward(sdem_gnn* M) {
  const int H = M->cfg.hidden, L = L_of(M), n = M->n;
  const int ne = (int)M->ne;
  hipStream_t st = M->stream;
  const bool tr = M->train_alloc;
  auto Hs = [&](int l) { return M->h[tr ? l : l % 2]; };
  auto Es = [&](int l) { return M->e[tr ? l : l % 2]; };
  auto XE = [&](int l) { return M->xe[tr ? l : 0]; };
  auto XN = [&](int l) { return M->xn[tr ? l : 0]; };
  Act& an = M->act[0];
  an.R = n; an.x = M->fn;
  Act& ae = M->act[1];
  ae.R = ne; ae.x = M->fe;
  const bool fused = fused_inference(M, tr);
  hipError_t e = fused ? refresh_transposes(M) : hipSuccess;
  // encoder and layer outputs go straight into the residual streams
  if (e == hipSuccess) e = mlp_fwd(M, M->mlp[0], an, nullptr, Hs(0));
  if (e == hipSuccess && ne > 0) e = mlp_fwd(M, M->mlp[1], ae, nullptr, Es(0));
  for (int l = 0; l < L && e == hipSuccess; l++) {
    Act& a_e = M->act[2 + 2 * l];
    Act& a_n = M->act[3 + 2 * l];
    if (ne > 0 && fused) {
      FusedMLP f = fused_args(M, M->mlp[2 + 2 * l], ne, Es(l), Es(l + 1));
      f.seg[0] = {Es(l), nullptr};
      f.seg[1] = {Hs(l), M->esrc};
      f.seg[2] = {Hs(l), M->edst};
      f.nseg = 3;
      e = mlp_fused(f, st);
      if (e == hipSuccess) e = segsum(Es(l + 1), M->rstart, n, H, M->agg, st);
    } else if (ne > 0) {
      // edge update e' = e + MLP([e, h_src, h_dst])
      e = pack_cols(Es(l), nullptr, ne, H, XE(l), 3 * H, 0, st);
      if (e == hipSuccess) e = pack_cols(Hs(l), M->esrc, ne, H, XE(l), 3 * H, H, st);
      if (e == hipSuccess) e = pack_cols(Hs(l), M->edst, ne, H, XE(l), 3 * H, 2 * H, st);
      a_e.R = ne; a_e.x = XE(l);
      if (e == hipSuccess) e = mlp_fwd(M, M->mlp[2 + 2 * l], a_e, Es(l), Es(l + 1));
      if (e == hipSuccess) e = segsum(Es(l + 1), M->rstart, n, H, M->agg, st);
    } else {
      e = hipMemsetAsync(M->agg, 0, (size_t)n * H * 4, st);
    }
    // node update h' = h + MLP([h, agg])
    if (fused) {
      FusedMLP f = fused_args(M, M->mlp[3 + 2 * l], n, Hs(l), Hs(l + 1));
      f.seg[0] = {Hs(l), nullptr};
      f.seg[1] = {M->agg, nullptr};
      f.nseg = 2;
      if (e == hipSuccess) e = mlp_fused(f, st);
      continue;
    }
    if (e == hipSuccess) e = pack_cols(Hs(l), nullptr, n, H, XN(l), 2 * H, 0, st);
    if (e == hipSuccess) e = pack_cols(M->agg, nullptr, n, H, XN(l), 2 * H, H, st);
    a_n.R = n; a_n.x = XN(l);
    if (e == hipSuccess) e = mlp_fwd(M, M->mlp[3 + 2 * l], a_n, Hs(l), Hs(l + 1));
  }
  Act& ad = M->act.back();
  ad.R = n;
  if (e == hipSuccess) e = pack_cols(Hs(L), nullptr, n, H, M->db16, H, 0, st);
  ad.x = M->db16;
  if (e == hipSuccess) e = mlp_fwd(M, M->mlp.back(), ad, nullptr, ad.out);
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "forward");
}

// backward pass from dout (n x PAD_IN... the decoder's output width)
int backward(sdem_gnn* M) {
  const int H = M->cfg.hidden, L = L_of(M), n = M->n;
  const int ne = (int)M->ne;
  hipStream_t st = M->stream;
  if (!M->train_alloc) return M->fail(SDEM_E_STATE, "the graph was set up for inference only");
  hipError_t e = hipMemsetAsync(M->g, 0, M->nparam * 4, st);
  // the decoder's input (h_L in bf16) was packed into scratch that the
  // backward pass reuses: pack it again where nothing else writes before
  // the decoder's gradients are done
  Act ad = M->act.back();
  __bf16* hLb = (__bf16*)M->dx_n;  // n x H bf16 fits in n x 2H fp32
  if (e == hipSuccess) e = pack_cols(M->h[L], nullptr, n, H, hLb, H, 0, st);
  ad.x = hLb;
  if (e == hipSuccess) e = mlp_bwd(M, M->mlp.back(), ad, M->dout, M->dh);  // dh = dL/dh_L
  if (e == hipSuccess && ne > 0) e = hipMemsetAsync(M->de, 0, (size_t)ne * H * 4, st);
  for (int l = L - 1; l >= 0 && e == hipSuccess; l--) {
    // node update: h_{l+1} = h_l + MLP_n([h_l, agg_l])
    e = mlp_bwd(M, M->mlp[3 + 2 * l], M->act[3 + 2 * l], M->dh, M->dx_n);
    // dh_l = dh_{l+1} + dXn[:, :H];  dagg = dXn[:, H:]
    if (e == hipSuccess) e = take_cols(M->dx_n, 2 * H, 0, n, H, M->dh, true, st);
    if (ne > 0) {
      if (e == hipSuccess) e = take_cols(M->dx_n, 2 * H, H, n, H, M->dagg, false, st);
      // agg = sum over incoming edges of e_{l+1}: de_{l+1} += dagg[dst]
      if (e == hipSuccess) e = gather_add(M->dagg, M->edst, ne, H, M->de, st);
      // edge update: e_{l+1} = e_l + MLP_e([e_l, h_l[src], h_l[dst]])
      if (e == hipSuccess) e = mlp_bwd(M, M->mlp[2 + 2 * l], M->act[2 + 2 * l], M->de, M->dx_e);
      if (e == hipSuccess) e = take_cols(M->dx_e, 3 * H, 0, ne, H, M->de, true, st);
      if (e == hipSuccess) e = scatter_rows(M->dx_e, 3 * H, H, M->esrc, ne, H, M->dh, st);
      if (e == hipSuccess) e = scatter_rows(M->dx_e, 3 * H, 2 * H, M->edst, ne, H, M->dh, st);
    }
  }
  if (e == hipSuccess && ne > 0) e = mlp_bwd(M, M->mlp[1], M->act[1], M->de, nullptr);
  if (e == hipSuccess) e = mlp_bwd(M, M->mlp[0], M->act[0], M->dh, nullptr);
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "backward");
}

// neighbour graph of n particles on the device (positions double4)
int build_graph(sdem_gnn* M, const double4* pos, int n, const GGrid& G, bool train) {
  hipStream_t st = M->stream;
  const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
  auto& A = M->allocs;
  hipError_t e = hipSuccess;
  if (n > M->search_n || ncell > M->ncell_alloc || !M->gkey) {
    // (re)allocate the search scratch
    for (void* p : {(void*)M->gkey, (void*)M->gsorted, (void*)M->gcnt, (void*)M->gfill, (void*)M->cs, (void*)M->ce,
                    (void*)M->deg, (void*)M->tsum}) {
      auto it = std::find(A.begin(), A.end(), p);
      if (it != A.end()) { (void)hipFree(p); A.erase(it); }
    }
    const size_t N = (size_t)std::max(n, 1);
    auto Gl = [&](auto** p, size_t k) { if (e == hipSuccess) e = galloc(A, p, k); };
    Gl(&M->gkey, N); Gl(&M->gsorted, N); Gl(&M->gcnt, (size_t)ncell); Gl(&M->gfill, (size_t)ncell);
    Gl(&M->cs, (size_t)ncell); Gl(&M->ce, (size_t)ncell); Gl(&M->deg, N + 1);
    Gl(&M->tsum, (size_t)std::max<int64_t>(1, std::max(scan_tiles(ncell), scan_tiles((int64_t)N + 1))));
    M->ncell_alloc = ncell;
    M->search_n = n;
    if (e != hipSuccess) return hipf(M, e, "hipMalloc(graph search)");
  }
  e = cell_sort(pos, n, G, ncell, M->gkey, M->gcnt, M->gfill, M->cs, M->ce, M->gsorted, M->tsum, st);
  if (e == hipSuccess) e = edges(false, pos, n, G, M->gsorted, M->cs, M->ce, M->deg, nullptr, nullptr, nullptr, st);
  if (e == hipSuccess) e = hipMemsetAsync(M->deg + n, 0, 4, st);
  // degrees -> receiver offsets (in place)
  if (e == hipSuccess) e = excl_scan(M->deg, M->deg, n + 1, M->tsum, st);
  int total = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&total, M->deg + n, 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hipf(M, e, "graph search");
  if ((int64_t)total > (int64_t)SDEM_GNN_MAX_EDGES_PER_NODE * n)
    return M->fail(SDEM_E_STATE, "surrogate graph too dense (" + std::to_string(total) + " edges for " +
                                     std::to_string(n) + " particles): the state has collapsed");
  if (int rc = alloc_graph(M, n, total, train)) return rc;
  M->n = n;
  M->ne = total;
  e = hipMemcpyAsync(M->rstart, M->deg, (size_t)(n + 1) * 4, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess)
    e = edges(true, pos, n, G, M->gsorted, M->cs, M->ce, nullptr, M->rstart, M->esrc, M->edst, st);
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "graph edges");
}

GGrid make_ggrid(const double lo[3], const double hi[3], const int per[3], double rc, double cfac) {
  GGrid G;
  G.rc = rc;
  G.cfac = cfac;
  for (int d = 0; d < 3; d++) {
    G.lo[d] = lo[d];
    G.L[d] = hi[d] - lo[d];
    int k = (int)std::floor(G.L[d] / rc);
    if (k < 1) k = 1;
    if (per[d] && k < 3) k = 3;  // the 27-cell walk must not meet a cell twice
    G.n[d] = k;
    G.inv_w[d] = k / G.L[d];
    G.per[d] = per[d] ? 1 : 0;
  }
  return G;
}

FeatScale scales(const sdem_gnn_config& c) {
  FeatScale F;
  F.v = c.vel_scale; F.a = c.acc_scale; F.l = c.len_scale; F.floor_z = c.floor_z;
  F.dx = c.dx_scale; F.dv = c.dv_scale;
  return F;
}

}  // namespace

// ===========================================================================
int sdem_gnn_config_default(sdem_gnn_config* c) {
  if (!c) return SDEM_E_ARG;
  std::memset(c, 0, sizeof(*c));
  c->hidden = 128;
  c->layers = 8;
  c->node_in = 8;
  c->edge_in = 7;
  c->out = 6;
  c->vel_scale = 0.1;
  c->acc_scale = 10.0;
  c->len_scale = 2.0e-3;
  c->radius = 4.8e-3;
  c->cutoff_factor = 1.2;
  c->dx_scale = 1.0e-6;
  c->dv_scale = 1.0e-2;
  c->nsub = 10;
  return SDEM_OK;
}

int sdem_gnn_create(const sdem_gnn_config* cfg, uint64_t seed, sdem_gnn** out) {
  if (!cfg || !out) return SDEM_E_ARG;
  *out = nullptr;
  if (cfg->hidden < 8 || cfg->hidden % 8 || cfg->layers < 0 || cfg->node_in < 1 || cfg->node_in > PAD_IN ||
      cfg->edge_in < 1 || cfg->edge_in > PAD_IN || cfg->out < 1 || cfg->out > PAD_IN)
    return SDEM_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SDEM_E_NODEV;
  sdem_gnn* M = new sdem_gnn();
  M->cfg = *cfg;
  if (hipStreamCreate(&M->stream) != hipSuccess) {
    delete M;
    return SDEM_E_HIP;
  }
  const int H = cfg->hidden, L = cfg->layers;
  size_t off = 0;
  M->mlp.push_back(make_mlp(off, PAD_IN, H, H));  // node encoder
  M->mlp.push_back(make_mlp(off, PAD_IN, H, H));  // edge encoder
  for (int l = 0; l < L; l++) {
    M->mlp.push_back(make_mlp(off, 3 * H, H, H));  // edge MLP
    M->mlp.push_back(make_mlp(off, 2 * H, H, H));  // node MLP
  }
  M->mlp.push_back(make_mlp(off, H, H, PAD_IN));  // decoder (first `out` columns used)
  M->nparam = off;
  // weights U(-s, s), s = sqrt(3 / fan_in) for the hidden layers (unit
  // variance through SiLU roughly kept), small last layers; biases zero
  std::vector<float> w(off, 0.0f);
  Rng rng{seed * 0x9E3779B97F4A7C15ull + 12345};
  auto fill = [&](size_t o, int fan_in, int rows, int cols, double gain) {
    const double s = gain * std::sqrt(3.0 / fan_in);
    for (size_t i = 0; i < (size_t)rows * cols; i++) w[o + i] = (float)((2.0 * rng.uniform() - 1.0) * s);
  };
  for (size_t q = 0; q < M->mlp.size(); q++) {
    const MLP& p = M->mlp[q];
    fill(p.w1, p.in, p.in, p.hid, 1.0);
    fill(p.w2, p.hid, p.hid, p.hid, 1.0);
    const bool last = q + 1 == M->mlp.size();
    fill(p.w3, p.hid, p.hid, p.out, last ? 0.1 : 0.3);
  }
  hipError_t e = hipSuccess;
  auto G = [&](auto** p, size_t k) { if (e == hipSuccess) e = galloc(M->allocs, p, k); };
  G(&M->w, off); G(&M->g, off); G(&M->m, off); G(&M->v, off); G(&M->wb, off); G(&M->wbt, off); G(&M->loss, 1);
  if (e == hipSuccess) e = hipMemcpy(M->w, w.data(), off * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(M->m, 0, off * 4);
  if (e == hipSuccess) e = hipMemset(M->v, 0, off * 4);
  if (e == hipSuccess) e = to_bf16(M->w, off, M->wb, M->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(M->stream);
  if (e != hipSuccess) {
    sdem_gnn_destroy(M);
    return SDEM_E_HIP;
  }
  *out = M;
  return SDEM_OK;
}

int sdem_gnn_get_config(sdem_gnn* M, sdem_gnn_config* c) {
  if (!M || !c) return SDEM_E_ARG;
  *c = M->cfg;
  return SDEM_OK;
}

int sdem_gnn_set_scales(sdem_gnn* M, const sdem_gnn_config* c) {
  if (!M || !c) return SDEM_E_ARG;
  M->cfg.vel_scale = c->vel_scale; M->cfg.acc_scale = c->acc_scale; M->cfg.len_scale = c->len_scale;
  M->cfg.radius = c->radius; M->cfg.cutoff_factor = c->cutoff_factor;
  M->cfg.floor_z = c->floor_z; M->cfg.dx_scale = c->dx_scale;
  M->cfg.dv_scale = c->dv_scale; M->cfg.nsub = c->nsub;
  return SDEM_OK;
}

int sdem_gnn_num_params(sdem_gnn* M, int64_t* n) {
  if (!M || !n) return SDEM_E_ARG;
  *n = (int64_t)M->nparam;
  return SDEM_OK;
}

int sdem_gnn_get_params(sdem_gnn* M, float* w) {
  if (!M || !w) return SDEM_E_ARG;
  hipError_t e = hipStreamSynchronize(M->stream);
  if (e == hipSuccess) e = hipMemcpy(w, M->w, M->nparam * 4, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "get_params");
}

int sdem_gnn_set_params(sdem_gnn* M, const float* w) {
  if (!M || !w) return SDEM_E_ARG;
  hipError_t e = hipMemcpy(M->w, w, M->nparam * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = to_bf16(M->w, M->nparam, M->wb, M->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(M->stream);
  M->wt_stale = true;
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "set_params");
}

int sdem_gnn_get_grads(sdem_gnn* M, float* g) {
  if (!M || !g) return SDEM_E_ARG;
  hipError_t e = hipStreamSynchronize(M->stream);
  if (e == hipSuccess) e = hipMemcpy(g, M->g, M->nparam * 4, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "get_grads");
}

int sdem_gnn_get_layout(sdem_gnn* M, int32_t nmlp_cap, int32_t* nmlp, int64_t* offsets /* nmlp x 6 */,
                        int32_t* dims /* nmlp x 3: in, hid, out */) {
  if (!M || !nmlp) return SDEM_E_ARG;
  *nmlp = (int32_t)M->mlp.size();
  if (nmlp_cap < *nmlp) return M->fail(SDEM_E_CAPACITY, "output arrays too small");
  for (size_t q = 0; q < M->mlp.size(); q++) {
    const MLP& p = M->mlp[q];
    const size_t o[6] = {p.w1, p.b1, p.w2, p.b2, p.w3, p.b3};
    for (int k = 0; k < 6; k++) if (offsets) offsets[6 * q + k] = (int64_t)o[k];
    if (dims) { dims[3 * q] = p.in; dims[3 * q + 1] = p.hid; dims[3 * q + 2] = p.out; }
  }
  return SDEM_OK;
}

int sdem_gnn_set_graph(sdem_gnn* M, int32_t n, const float* node_feat, int64_t ne, const int32_t* src,
                       const int32_t* dst, const float* edge_feat) {
  if (!M || n < 1 || ne < 0 || !node_feat || (ne > 0 && (!src || !dst || !edge_feat))) return SDEM_E_ARG;
  for (int64_t k = 1; k < ne; k++)
    if (dst[k] < dst[k - 1]) return M->fail(SDEM_E_ARG, "edges must be sorted by receiver");
  for (int64_t k = 0; k < ne; k++)
    if (src[k] < 0 || src[k] >= n || dst[k] < 0 || dst[k] >= n) return M->fail(SDEM_E_ARG, "edge index out of range");
  if (int rc = alloc_graph(M, n, ne, true)) return rc;
  M->n = n;
  M->ne = ne;
  std::vector<int> rs(n + 1, 0);
  for (int64_t k = 0; k < ne; k++) rs[dst[k] + 1]++;
  for (int i = 0; i < n; i++) rs[i + 1] += rs[i];
  std::vector<__bf16> fn((size_t)n * PAD_IN, (__bf16)0.0f), fe((size_t)std::max<int64_t>(ne, 1) * PAD_IN, (__bf16)0.0f);
  for (int i = 0; i < n; i++)
    for (int c = 0; c < M->cfg.node_in; c++) fn[(size_t)i * PAD_IN + c] = (__bf16)node_feat[(size_t)i * M->cfg.node_in + c];
  for (int64_t k = 0; k < ne; k++)
    for (int c = 0; c < M->cfg.edge_in; c++) fe[(size_t)k * PAD_IN + c] = (__bf16)edge_feat[(size_t)k * M->cfg.edge_in + c];
  hipError_t e = hipMemcpy(M->rstart, rs.data(), (size_t)(n + 1) * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && ne > 0) e = hipMemcpy(M->esrc, src, (size_t)ne * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && ne > 0) e = hipMemcpy(M->edst, dst, (size_t)ne * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(M->fn, fn.data(), fn.size() * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess && ne > 0) e = hipMemcpy(M->fe, fe.data(), (size_t)ne * PAD_IN * 2, hipMemcpyHostToDevice);
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "set_graph");
}

int sdem_gnn_forward(sdem_gnn* M, float* out) {
  if (!M) return SDEM_E_ARG;
  if (M->n < 1) return M->fail(SDEM_E_STATE, "no graph");
  if (int rc = forward(M)) return rc;
  hipError_t e = hipStreamSynchronize(M->stream);
  if (e != hipSuccess) return hipf(M, e, "forward");
  if (out) {
    std::vector<float> o((size_t)M->n * PAD_IN);
    e = hipMemcpy(o.data(), M->act.back().out, o.size() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipf(M, e, "forward");
    for (int i = 0; i < M->n; i++)
      for (int c = 0; c < M->cfg.out; c++) out[(size_t)i * M->cfg.out + c] = o[(size_t)i * PAD_IN + c];
  }
  return SDEM_OK;
}

int sdem_gnn_backward(sdem_gnn* M, const float* dout) {
  if (!M || !dout) return SDEM_E_ARG;
  if (M->n < 1) return M->fail(SDEM_E_STATE, "no graph");
  std::vector<float> d((size_t)M->n * PAD_IN, 0.0f);
  for (int i = 0; i < M->n; i++)
    for (int c = 0; c < M->cfg.out; c++) d[(size_t)i * PAD_IN + c] = dout[(size_t)i * M->cfg.out + c];
  hipError_t e = hipMemcpy(M->dout, d.data(), d.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hipf(M, e, "backward");
  if (int rc = backward(M)) return rc;
  e = hipStreamSynchronize(M->stream);
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "backward");
}

// one optimisation step on the current graph: forward, mean squared error
// against target (n x out), backward, gradient clipping, Adam
int sdem_gnn_train_step(sdem_gnn* M, const float* target, float lr, float clip, float* loss_out) {
  if (!M || !target) return SDEM_E_ARG;
  if (M->n < 1) return M->fail(SDEM_E_STATE, "no graph");
  if (int rc = forward(M)) return rc;
  const int n = M->n, nout = M->cfg.out;
  float* dtarget = M->tmp_a;  // n x out fits
  hipError_t e = hipMemcpyAsync(dtarget, target, (size_t)n * nout * 4, hipMemcpyHostToDevice, M->stream);
  if (e == hipSuccess) e = hipMemsetAsync(M->loss, 0, 4, M->stream);
  const float scale = 1.0f / ((float)n * nout);
  if (e == hipSuccess) e = mse(M->act.back().out, dtarget, n, nout, PAD_IN, scale, M->dout, M->loss, M->stream);
  if (e != hipSuccess) return hipf(M, e, "loss");
  if (int rc = backward(M)) return rc;
  float* gn = M->loss + 0;  // reuse after reading the loss
  float loss = 0.0f, g2 = 0.0f;
  e = hipMemcpyAsync(&loss, M->loss, 4, hipMemcpyDeviceToHost, M->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(M->stream);
  if (e == hipSuccess) e = hipMemsetAsync(gn, 0, 4, M->stream);
  if (e == hipSuccess) e = sumsq(M->g, M->nparam, gn, M->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&g2, gn, 4, hipMemcpyDeviceToHost, M->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(M->stream);
  if (e != hipSuccess) return hipf(M, e, "gradient norm");
  const float gnorm = std::sqrt(g2);
  const float cs = (clip > 0.0f && gnorm > clip) ? clip / gnorm : 1.0f;
  M->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.0f - std::pow(b1, (float)M->adam_t), bc2 = 1.0f - std::pow(b2, (float)M->adam_t);
  e = adam(M->w, M->g, M->m, M->v, M->wb, M->nparam, lr, b1, b2, 1e-8f, bc1, bc2, cs, M->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(M->stream);
  M->wt_stale = true;
  if (e != hipSuccess) return hipf(M, e, "adam");
  if (loss_out) *loss_out = loss;
  return SDEM_OK;
}

// graph + features from a DEM context's state (first box, live slots)
static int graph_from_dem(sdem_gnn* M, sdem_ctx* dem, sdem_dem_view* V, bool train) {
  if (int rc = sdem_internal_view(dem, V)) return M->fail(rc, "DEM context view: " + std::string(sdem_last_error(dem)));
  const GGrid G = make_ggrid(V->lo, V->hi, V->per, M->cfg.radius, M->cfg.cutoff_factor);
  if (int rc = build_graph(M, V->pos, V->n, G, train)) return rc;
  const FeatScale F = scales(M->cfg);
  hipError_t e = node_feat(V->pos, V->lv, V->la, V->n, V->cap, F, M->fn, PAD_IN, M->stream);
  if (e == hipSuccess && M->ne > 0)
    e = edge_feat(V->pos, V->lv, V->cap, M->esrc, M->edst, (int)M->ne, G, F, M->fe, PAD_IN, M->stream);
  return e == hipSuccess ? SDEM_OK : hipf(M, e, "features");
}

int sdem_gnn_graph_from_dem(sdem_gnn* M, sdem_ctx* dem, int32_t for_training, int64_t* nedges) {
  if (!M || !dem) return SDEM_E_ARG;
  sdem_dem_view V;
  if (int rc = graph_from_dem(M, dem, &V, for_training != 0)) return rc;
  if (M->ids_cap < V.n) {
    M->ids.resize(V.n);
    M->ids_cap = V.n;
  }
  hipError_t e0 = hipMemcpy(M->ids.data(), V.id, (size_t)V.n * 4, hipMemcpyDeviceToHost);
  if (e0 != hipSuccess) return hipf(M, e0, "node ids");
  hipError_t e = hipStreamSynchronize(M->stream);
  if (e != hipSuccess) return hipf(M, e, "graph");
  if (nedges) *nedges = M->ne;
  return SDEM_OK;
}

int sdem_gnn_get_num_nodes(sdem_gnn* M, int32_t* n, int64_t* ne) {
  if (!M) return SDEM_E_ARG;
  if (n) *n = M->n;
  if (ne) *ne = M->ne;
  return SDEM_OK;
}

int sdem_gnn_get_node_ids(sdem_gnn* M, int32_t cap, int32_t* ids) {
  if (!M || !ids) return SDEM_E_ARG;
  if (cap < M->n || (int)M->ids.size() < M->n) return M->fail(SDEM_E_CAPACITY, "no ids or arrays too small");
  std::copy(M->ids.begin(), M->ids.begin() + M->n, ids);
  return SDEM_OK;
}

int sdem_gnn_get_graph(sdem_gnn* M, int64_t cap, int32_t* src, int32_t* dst, float* node_feat, float* edge_feat) {
  if (!M) return SDEM_E_ARG;
  if (cap < M->ne) return M->fail(SDEM_E_CAPACITY, "output arrays too small");
  hipError_t e = hipStreamSynchronize(M->stream);
  std::vector<__bf16> fn((size_t)M->n * PAD_IN), fe((size_t)std::max<int64_t>(M->ne, 1) * PAD_IN);
  if (e == hipSuccess && src) e = hipMemcpy(src, M->esrc, (size_t)M->ne * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && dst) e = hipMemcpy(dst, M->edst, (size_t)M->ne * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(fn.data(), M->fn, fn.size() * 2, hipMemcpyDeviceToHost);
  if (e == hipSuccess && M->ne > 0) e = hipMemcpy(fe.data(), M->fe, (size_t)M->ne * PAD_IN * 2, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hipf(M, e, "get_graph");
  if (node_feat)
    for (int i = 0; i < M->n; i++)
      for (int c = 0; c < M->cfg.node_in; c++) node_feat[(size_t)i * M->cfg.node_in + c] = (float)fn[(size_t)i * PAD_IN + c];
  if (edge_feat)
    for (int64_t k = 0; k < M->ne; k++)
      for (int c = 0; c < M->cfg.edge_in; c++) edge_feat[(size_t)k * M->cfg.edge_in + c] = (float)fe[(size_t)k * PAD_IN + c];
  return SDEM_OK;
}

// `calls` surrogate calls on a DEM context: each builds the graph of the
// current state, predicts the corrections over nsub DEM steps and applies
// them (ballistic step + corrections) to the context's state on the device;
// the context then continues as if it had stepped nsub * calls times, with
// its contact history cleared.
int sdem_gnn_advance(sdem_gnn* M, sdem_ctx* dem, int32_t calls) {
  if (!M || !dem || calls < 0) return SDEM_E_ARG;
  for (int q = 0; q < calls; q++) {
    sdem_dem_view V;
    if (int rc = graph_from_dem(M, dem, &V, false)) return rc;
    if (int rc = forward(M)) return rc;
    const GGrid G = make_ggrid(V.lo, V.hi, V.per, M->cfg.radius, M->cfg.cutoff_factor);
    const FeatScale F = scales(M->cfg);
    const double T = V.dt * M->cfg.nsub;
    hipError_t e = advance(V.pos, V.lv, V.la, V.n, V.cap, M->act.back().out, PAD_IN, F, G, T, V.g, M->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(M->stream);
    if (e != hipSuccess) return hipf(M, e, "advance");
    if (int rc = sdem_internal_after_update(dem, M->cfg.nsub)) return M->fail(rc, sdem_last_error(dem));
  }
  return SDEM_OK;
}

int sdem_gnn_mfma_probe(const uint16_t* a_bits, const uint16_t* b_bits, float* d) {
  if (!a_bits || !b_bits || !d) return SDEM_E_ARG;
  void *da = nullptr, *db = nullptr;
  float* dd = nullptr;
  hipError_t e = hipMalloc(&da, 64 * 8 * 2);
  if (e == hipSuccess) e = hipMalloc(&db, 64 * 8 * 2);
  if (e == hipSuccess) e = hipMalloc((void**)&dd, 64 * 16 * 4);
  if (e == hipSuccess) e = hipMemcpy(da, a_bits, 64 * 8 * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, b_bits, 64 * 8 * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = mfma_probe(da, db, dd, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(d, dd, 64 * 16 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dd);
  return e == hipSuccess ? SDEM_OK : SDEM_E_HIP;
}

int sdem_gnn_gemm(int32_t ta, int32_t tb, int32_t M_, int32_t N, int32_t K, const float* a, const float* b,
                  const float* bias, int32_t flags, int32_t splits, float* c, float* pre) {
  if (!a || !b || !c || M_ < 1 || N < 1 || K < 1) return SDEM_E_ARG;
  const size_t na = (size_t)M_ * K, nbb = (size_t)K * N, nc = (size_t)M_ * N;
  std::vector<__bf16> ha(na), hb(nbb);
  for (size_t i = 0; i < na; i++) ha[i] = (__bf16)a[i];
  for (size_t i = 0; i < nbb; i++) hb[i] = (__bf16)b[i];
  __bf16 *da = nullptr, *db = nullptr;
  float *dc = nullptr, *dbias = nullptr, *dpre = nullptr;
  const int sp = std::max(1, (int)splits);
  hipError_t e = hipMalloc((void**)&da, na * 2);
  if (e == hipSuccess) e = hipMalloc((void**)&db, nbb * 2);
  if (e == hipSuccess) e = hipMalloc((void**)&dc, nc * 4 * sp);
  if (e == hipSuccess) e = hipMalloc((void**)&dbias, (size_t)N * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&dpre, nc * 4);
  if (e == hipSuccess) e = hipMemcpy(da, ha.data(), na * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, hb.data(), nbb * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess && bias) e = hipMemcpy(dbias, bias, (size_t)N * 4, hipMemcpyHostToDevice);
  const int lda = ta ? M_ : K, ldb = tb ? K : N;
  Epi ep{flags & (EPI_BIAS | EPI_SILU | EPI_PRE), dbias, dpre, nullptr, dc, nullptr};
  if (e == hipSuccess) e = gemm(ta != 0, tb != 0, da, db, M_, N, K, lda, ldb, sp, ep, nullptr);
  if (e == hipSuccess && sp > 1) {
    int kchunk = (K + sp - 1) / sp;
    kchunk = (kchunk + 31) / 32 * 32;
    const int nz = std::max(1, (K + kchunk - 1) / kchunk);
    float* red = dpre;
    e = reduce_parts(dc, nz, nc, red, false, nullptr);
    if (e == hipSuccess) e = hipMemcpy(dc, red, nc * 4, hipMemcpyDeviceToDevice);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(c, dc, nc * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && pre && sp == 1) e = hipMemcpy(pre, dpre, nc * 4, hipMemcpyDeviceToHost);
  (void)hipFree(da); (void)hipFree(db); (void)hipFree(dc); (void)hipFree(dbias); (void)hipFree(dpre);
  return e == hipSuccess ? SDEM_OK : SDEM_E_HIP;
}

const char* sdem_gnn_last_error(const sdem_gnn* M) { return M ? M->err.c_str() : "null model"; }

void sdem_gnn_destroy(sdem_gnn* M) {
  if (!M) return;
  if (M->stream) (void)hipStreamSynchronize(M->stream);
  free_all(M->graph_allocs);
  free_all(M->allocs);
  if (M->stream) (void)hipStreamDestroy(M->stream);
  delete M;
}

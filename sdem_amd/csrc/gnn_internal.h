// gnn_internal.h -- shared declarations of the surrogate's kernels
// (gnn_kernels.hip) and its host side (gnn_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace gnn {

// GEMM epilogue: v = acc (+ bias[col]); pre[o] = v if EPI_PRE; v = SiLU(v)
// if EPI_SILU; v += resid[o] if EPI_RESID; then out_f32[z][o] = v and/or
// out_bf16[o] = v for whichever pointers are set (o = row * N + col, z =
// split-K chunk).
enum { EPI_BIAS = 1, EPI_SILU = 2, EPI_PRE = 4, EPI_RESID = 8 };
struct Epi {
  int flags;
  const float* bias;
  float* pre;
  const float* resid;
  float* out_f32;
  __bf16* out_bf16;
};

// Fused inference MLP (hidden = out = FUSED_H): y = resid + MLP(x) with the
// input row x[r] = [seg0 row, seg1 row, seg2 row] gathered from fp32
// matrices of width FUSED_H (optional row index per segment), weights read
// transposed ([out][in], bf16).
constexpr int FUSED_H = 128;
struct FusedSeg {
  const float* x;
  const int* idx;
};
struct FusedMLP {
  FusedSeg seg[3];
  int nseg, R;
  const __bf16 *w1t, *w2t, *w3t;
  const float *b1, *b2, *b3;
  const float* resid;
  float* y;
};
hipError_t mlp_fused(const FusedMLP& a, hipStream_t st);
hipError_t transpose_bf16(const __bf16* w, int rows, int cols, __bf16* wt, hipStream_t st);

// neighbour-search grid of the surrogate's graph (cells >= cut-off)
struct GGrid {
  double lo[3], L[3], inv_w[3];
  int n[3], per[3];
  double rc;      // largest cut-off (cells are at least this wide)
  double cfac;    // pair cut-off = cfac * 2 * max(r_i, r_j)
};

// feature and output scales: velocity, acceleration, length, the floor's
// height, and the units of the position / velocity corrections
struct FeatScale {
  double v, a, l, floor_z, dx, dv;
};

hipError_t gather_add(const float* src, const int* idx, int R, int C, float* dst, hipStream_t st);
// exclusive scan of n ints (in == out allowed); tsum: scan_tiles(n) ints
int64_t scan_tiles(int64_t n);
hipError_t excl_scan(const int* in, int* out, int64_t n, int* tsum, hipStream_t st);
// particles sorted by cell (ascending index within a cell): sorted[cstart[c]
// .. cend[c]) = the particles of cell c; key, cnt, fill are scratch (n,
// ncell, ncell ints), tsum scan_tiles(ncell) ints
hipError_t cell_sort(const double4* pos, int n, const GGrid& G, int64_t ncell, int* key, int* cnt, int* fill,
                     int* cstart, int* cend, int* sorted, int* tsum, hipStream_t st);
hipError_t edges(bool fill, const double4* pos, int n, const GGrid& G, const int* sorted, const int* cstart,
                 const int* cend, int* deg, const int* rstart, int* esrc, int* edst, hipStream_t st);
hipError_t node_feat(const double4* pos, const double* vel, const double* acc, int n, int stride,
                     const FeatScale& F, __bf16* out, int ld, hipStream_t st);
hipError_t edge_feat(const double4* pos, const double* vel, int stride, const int* esrc, const int* edst, int ne,
                     const GGrid& G, const FeatScale& F, __bf16* out, int ld, hipStream_t st);
hipError_t advance(double4* pos, double* vel, double* acc, int n, int stride, const float* pred, int ld,
                   const FeatScale& F, const GGrid& G, double T, const double g[3], hipStream_t st);
hipError_t mfma_probe(const void* a, const void* b, float* d, hipStream_t st);
hipError_t gemm(bool ta, bool tb, const __bf16* A, const __bf16* B, int M, int N, int K, int lda, int ldb,
                int splits, const Epi& ep, hipStream_t st);
int gemm_splits(int K, int kchunk_target);
hipError_t reduce_parts(const float* part, int nz, size_t len, float* out, bool accumulate, hipStream_t st);
hipError_t colsum(const float* x, int R, int C, float* out, hipStream_t st);
hipError_t silu_bwd(const float* dy, const float* pre, size_t n, float* dpre, __bf16* dpre_b, hipStream_t st);
hipError_t to_bf16(const float* x, size_t n, __bf16* y, hipStream_t st);
hipError_t pack_cols(const float* x, const int* idx, int R, int C, __bf16* out, int ldo, int c0, hipStream_t st);
hipError_t segsum(const float* e, const int* rstart, int nv, int C, float* agg, hipStream_t st);
hipError_t scatter_rows(const float* x, int ldx, int c0, const int* idx, int R, int C, float* out, hipStream_t st);
hipError_t add(float* y, const float* x, size_t n, hipStream_t st);
hipError_t take_cols(const float* x, int ldx, int c0, int R, int C, float* y, bool accumulate, hipStream_t st);
hipError_t mse(const float* pred, const float* target, int R, int nout, int ld, float scale, float* grad,
               float* loss, hipStream_t st);
hipError_t adam(float* w, const float* g, float* m, float* v, __bf16* wb, size_t n, float lr, float b1, float b2,
                float eps, float bc1, float bc2, float clip_scale, hipStream_t st);
hipError_t sumsq(const float* x, size_t n, float* out, hipStream_t st);

}  // namespace gnn

// gnn_kernels.hip -- gfx950 kernels of the learned surrogate: a graph network
// over the neighbour graph of the particles (encoder, message-passing
// layers of edge and node MLPs, decoder), trained and run on the GPU.
//
//   k_gemm        : C = op(A) op(B) on the matrix cores (bf16 inputs,
//                   fp32 accumulation, v_mfma_f32_32x32x16_bf16), with bias,
//                   SiLU, pre-activation and residual epilogues; op(A) = A or
//                   A^T, op(B) = B or B^T, split-K partials for weight
//                   gradients
//   k_silu_bwd    : dPre = dY * SiLU'(pre), to bf16
//   k_gather_edge : edge-MLP input rows [e, h_src, h_dst]
//   k_segsum      : sum of edge rows into their receivers (edges sorted by
//                   receiver, so the order of the sum is fixed)
//   k_scatter_add : receivers' / senders' gradients (fp32 atomics)
//   graph builder : cell binning and radius search on the particle positions
//   k_adam        : Adam on fp32 master weights, bf16 copies for the MFMAs
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gnn_internal.h"

namespace gnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float silu(float x) { return x / (1.0f + __expf(-x)); }
__device__ __forceinline__ float silu_grad(float x) {
  const float s = 1.0f / (1.0f + __expf(-x));
  return s * (1.0f + x * (1.0f - s));
}

// ---------------------------------------------------------------------------
// lane layout of v_mfma_f32_32x32x16_bf16 (checked by k_mfma_probe):
//   A (32x16): lane l holds A[l%32][8*(l/32) + j], j = 0..7
//   B (16x32): lane l holds B[8*(l/32) + j][l%32], j = 0..7
//   D (32x32): register r of lane l is D[8*(r/4) + 4*(l/32) + r%4][l%32]
// ---------------------------------------------------------------------------
__global__ void k_mfma_probe(const bf16x8* __restrict__ a, const bf16x8* __restrict__ b, float* __restrict__ d) {
  const int l = threadIdx.x;  // one wave of 64 lanes
  f32x16 acc = {0};
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[l], b[l], acc, 0, 0, 0);
  for (int r = 0; r < 16; r++) d[l * 16 + r] = acc[r];
}

// ---------------------------------------------------------------------------
// GEMM: C[M x N] = op(A)[M x K] * op(B)[K x N]
//   A row-major [M x K] (lda), or, with TA, stored as A^T [K x M] (lda)
//   B row-major [K x N] (ldb), or, with TB, stored as B^T [N x K] (ldb)
// Block tile 64 x 64, 4 waves each owning a 32 x 32 quarter; K in stages of
// 32 through LDS, both operands k-contiguous there.  blockIdx.z selects a
// K chunk of `kchunk` (split-K: partial results go to C + z*M*N).
// ---------------------------------------------------------------------------
constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

__device__ __forceinline__ __bf16 ldbf(const __bf16* p, bool ok) { return ok ? *p : (__bf16)0.0f; }

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(const __bf16* __restrict__ A, const __bf16* __restrict__ B,
                                              int M, int N, int K, int lda, int ldb, int kchunk,
                                              Epi ep) {
  __shared__ __attribute__((aligned(16))) __bf16 As[BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[BN][LDK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x16 acc = {0};
  for (int k0 = kb; k0 < ke; k0 += BK) {
    // A tile 64 x 32 and B tile 32 x 64, 8 elements per thread each; a
    // whole in-bounds run of 8 is one 16-byte load (leading dimensions are
    // multiples of 8, chunks start at multiples of 32)
    if (!TA) {
      const int r = t >> 2, c = (t & 3) * 8;
      const int gm = m0 + r, gk = k0 + c;
      const __bf16* src = A + (size_t)gm * lda + gk;
      if (gm < M && gk + 8 <= ke && (lda & 7) == 0) {
        *(bf16x8*)&As[r][c] = *(const bf16x8*)src;
      } else {
        for (int j = 0; j < 8; j++) As[r][c + j] = ldbf(src + j, gm < M && gk + j < ke);
      }
    } else {
      const int kk = t >> 3, c = (t & 7) * 8;
      const int gk = k0 + kk, gm = m0 + c;
      const __bf16* src = A + (size_t)gk * lda + gm;
      if (gk < ke && gm + 8 <= M && (lda & 7) == 0) {
        const bf16x8 v = *(const bf16x8*)src;
        for (int j = 0; j < 8; j++) As[c + j][kk] = v[j];
      } else {
        for (int j = 0; j < 8; j++) As[c + j][kk] = ldbf(src + j, gk < ke && gm + j < M);
      }
    }
    if (!TB) {
      const int kk = t >> 3, c = (t & 7) * 8;
      const int gk = k0 + kk, gn = n0 + c;
      const __bf16* src = B + (size_t)gk * ldb + gn;
      if (gk < ke && gn + 8 <= N && (ldb & 7) == 0) {
        const bf16x8 v = *(const bf16x8*)src;
        for (int j = 0; j < 8; j++) Bs[c + j][kk] = v[j];
      } else {
        for (int j = 0; j < 8; j++) Bs[c + j][kk] = ldbf(src + j, gk < ke && gn + j < N);
      }
    } else {
      const int r = t >> 2, c = (t & 3) * 8;
      const int gn = n0 + r, gk = k0 + c;
      const __bf16* src = B + (size_t)gn * ldb + gk;
      if (gn < N && gk + 8 <= ke && (ldb & 7) == 0) {
        *(bf16x8*)&Bs[r][c] = *(const bf16x8*)src;
      } else {
        for (int j = 0; j < 8; j++) Bs[r][c + j] = ldbf(src + j, gn < N && gk + j < ke);
      }
    }
    __syncthreads();
    for (int s = 0; s < BK; s += 16) {
      const int kq = s + 8 * (lane >> 5);
      const bf16x8 a = *(const bf16x8*)&As[wm + (lane & 31)][kq];
      const bf16x8 b = *(const bf16x8*)&Bs[wn + (lane & 31)][kq];
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue
  const int col = n0 + wn + (lane & 31);
  if (col >= N) return;
  const float bias = (ep.flags & EPI_BIAS) ? ep.bias[col] : 0.0f;
  float* part = ep.out_f32 ? ep.out_f32 + (size_t)blockIdx.z * M * N : nullptr;
  for (int r = 0; r < 16; r++) {
    const int row = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (row >= M) continue;
    const size_t o = (size_t)row * N + col;
    float v = acc[r] + bias;
    if (ep.flags & EPI_PRE) ep.pre[o] = v;
    if (ep.flags & EPI_SILU) v = silu(v);
    if (ep.flags & EPI_RESID) v += ep.resid[o];
    if (part) part[o] = v;
    if (ep.out_bf16) ep.out_bf16[o] = (__bf16)v;
  }
}

// ---------------------------------------------------------------------------
// Fused inference MLP: y[r] = resid[r] + W3^T a2 + b3, a2 = SiLU(W2^T a1 +
// b2), a1 = SiLU(W1^T x[r] + b1), hidden = out = 128, for a tile of 128
// rows per block.  Layer 1 streams K in chunks of 32: the input rows
// (gathered from their fp32 sources, rounded to bf16) and the matching
// columns of W1^T are staged in LDS, double-buffered, with the next chunk's
// global loads in flight during the current chunk's products.  a1 and a2
// stay in LDS as bf16; W2^T and W3^T are staged in LDS in turn.  Four
// waves each own a 64 x 64 quarter of the output (2 x 2 MFMA tiles).  The
// products run in the same k order, with the same roundings, as k_gemm's
// three launches on pack_cols inputs, so both paths give the same bits.
// ---------------------------------------------------------------------------
constexpr int FR = 128, FNT = 512, FKC = 32, FXK = FKC + 8, FHK = FUSED_H + 8, FPD = 3;

struct FusedStage {
  __bf16 X[FR][FXK];       // input rows, chunk of FKC columns
  __bf16 W[FUSED_H][FXK];  // W1^T rows (output units), same columns
};
union FusedStaging {
  FusedStage st[2];
  __bf16 Wl[FUSED_H][FHK];  // W2^T or W3^T
};

// products of a 64 x 32 slice (rows wm..wm+63, columns wn..wn+31) over K:
// A rows from LDS (stride lda), B = W^T rows from LDS (stride ldb), k from 0
// to K in steps of 16
__device__ __forceinline__ void fused_mma(const __bf16* As, int lda, const __bf16* Bs, int ldb, int K, int wm,
                                          int wn, int lane, f32x16 acc[2]) {
  for (int kk = 0; kk < K; kk += 16) {
    const int kq = kk + 8 * (lane >> 5);
    bf16x8 af[2];
    for (int i = 0; i < 2; i++) af[i] = *(const bf16x8*)(As + (wm + i * 32 + (lane & 31)) * lda + kq);
    const bf16x8 bfr = *(const bf16x8*)(Bs + (wn + (lane & 31)) * ldb + kq);
    for (int i = 0; i < 2; i++) acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bfr, acc[i], 0, 0, 0);
  }
}

// a = bf16(SiLU(acc + bias)) into the LDS activation tile; acc cleared
__device__ __forceinline__ void fused_act(f32x16 acc[2], const float* __restrict__ bias, __bf16 (*Hs)[FHK],
                                          int wm, int wn, int lane) {
  const int col = wn + (lane & 31);
  const float b = bias[col];
  for (int i = 0; i < 2; i++)
    for (int r = 0; r < 16; r++) {
      const int row = wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      Hs[row][col] = (__bf16)silu(acc[i][r] + b);
      acc[i][r] = 0.0f;
    }
}

// a 128 x 128 bf16 weight matrix (row-major, stride 128) into LDS rows of FHK
__device__ __forceinline__ void fused_load_w(const __bf16* __restrict__ wt, __bf16 (*Wl)[FHK], int t) {
  for (int q = 0; q < (FUSED_H * FUSED_H / 8) / FNT; q++) {
    const int u = q * FNT + t;
    const int n = u / (FUSED_H / 8), c = (u % (FUSED_H / 8)) * 8;
    *(bf16x8*)&Wl[n][c] = *(const bf16x8*)(wt + (size_t)n * FUSED_H + c);
  }
}

__global__ __launch_bounds__(FNT, 2) void k_mlp_fused(FusedMLP a) {
  __shared__ __attribute__((aligned(16))) FusedStaging U;
  __shared__ __attribute__((aligned(16))) __bf16 Hs[FR][FHK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int r0 = blockIdx.x * FR;
  const int wm = (w >> 2) * 64, wn = (w & 3) * 32;  // eight waves: 64 rows x 32 output columns each
  const int K1 = a.nseg * FUSED_H, nch = K1 / FKC;
  f32x16 acc[2];
  for (int i = 0; i < 2; i++)
    for (int r = 0; r < 16; r++) acc[i][r] = 0.0f;

  // per thread and chunk: 2 float4 of input (rows u/8, columns (u%8)*4) and
  // 8 bf16 of W1^T (row u/4, columns (u%4)*8).  The input rows are gathers
  // (an edge's sender and receiver rows): FPD chunks are in flight at once,
  // in registers, so that a block waits for one round trip per FPD chunks.
  float4 xr[FPD][2];
  bf16x8 wr[FPD];
  auto load_chunk = [&](int c, float4 (&x)[2], bf16x8& wv) {
    const int k0 = c * FKC;
    const FusedSeg sg = a.seg[k0 / FUSED_H];
    const int c0 = k0 % FUSED_H;
    for (int q = 0; q < 2; q++) {
      const int u = q * FNT + t;
      const int r = u >> 3, cc = (u & 7) * 4;
      const int gr = r0 + r;
      x[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (gr < a.R) {
        const int src = sg.idx ? sg.idx[gr] : gr;
        x[q] = *(const float4*)(sg.x + (size_t)src * FUSED_H + c0 + cc);
      }
    }
    wv = *(const bf16x8*)(a.w1t + (size_t)(t >> 2) * K1 + k0 + (t & 3) * 8);
  };
#pragma unroll
  for (int p = 0; p < FPD; p++)
    if (p < nch) load_chunk(p, xr[p], wr[p]);
  for (int cb = 0; cb < nch; cb += FPD) {
#pragma unroll
    for (int p = 0; p < FPD; p++) {
      const int c = cb + p;
      if (c >= nch) break;
      FusedStage& S = U.st[c & 1];
      for (int q = 0; q < 2; q++) {
        const int u = q * FNT + t;
        const int r = u >> 3, cc = (u & 7) * 4;
        S.X[r][cc] = (__bf16)xr[p][q].x;
        S.X[r][cc + 1] = (__bf16)xr[p][q].y;
        S.X[r][cc + 2] = (__bf16)xr[p][q].z;
        S.X[r][cc + 3] = (__bf16)xr[p][q].w;
      }
      *(bf16x8*)&S.W[t >> 2][(t & 3) * 8] = wr[p];
      __syncthreads();
      if (c + FPD < nch) load_chunk(c + FPD, xr[p], wr[p]);
      fused_mma(&S.X[0][0], FXK, &S.W[0][0], FXK, FKC, wm, wn, lane, acc);
    }
  }
  fused_act(acc, a.b1, Hs, wm, wn, lane);
  __syncthreads();  // layer 1 done everywhere: the staging area is free, a1 complete
  fused_load_w(a.w2t, U.Wl, t);
  __syncthreads();
  fused_mma(&Hs[0][0], FHK, &U.Wl[0][0], FHK, FUSED_H, wm, wn, lane, acc);
  __syncthreads();  // a1 and W2^T read everywhere
  fused_act(acc, a.b2, Hs, wm, wn, lane);
  fused_load_w(a.w3t, U.Wl, t);
  __syncthreads();
  fused_mma(&Hs[0][0], FHK, &U.Wl[0][0], FHK, FUSED_H, wm, wn, lane, acc);
  const int col = wn + (lane & 31);
  const float b = a.b3[col];
  for (int i = 0; i < 2; i++)
    for (int r = 0; r < 16; r++) {
      const int row = r0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (row >= a.R) continue;
      const size_t o = (size_t)row * FUSED_H + col;
      float v = acc[i][r] + b;
      if (a.resid) v += a.resid[o];
      a.y[o] = v;
    }
}

// wt[c][r] = w[r][c] (bf16, row-major)
__global__ void k_transpose_bf16(const __bf16* __restrict__ w, int rows, int cols, __bf16* __restrict__ wt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  wt[(size_t)c * rows + r] = w[i];
}

// sum of the split-K partials: out[i] (+)= sum_z part[z*len + i]
__global__ void k_reduce_parts(const float* __restrict__ part, int nz, size_t len, float* __restrict__ out,
                               int accumulate) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int z = 0; z < nz; z++) s += part[(size_t)z * len + i];
  out[i] = accumulate ? out[i] + s : s;
}

// column sums of a row-major [R x C] fp32 matrix, added to out[C]
__global__ void k_colsum(const float* __restrict__ x, int R, int C, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int r0 = blockIdx.y * 1024, r1 = min(R, r0 + 1024);
  float s = 0.0f;
  for (int r = r0; r < r1; r++) s += x[(size_t)r * C + c];
  atomicAdd(&out[c], s);
}

// dPre = dY * SiLU'(pre) (fp32 and bf16 copies)
__global__ void k_silu_bwd(const float* __restrict__ dy, const float* __restrict__ pre, size_t n,
                           float* __restrict__ dpre, __bf16* __restrict__ dpre_b) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = dy[i] * silu_grad(pre[i]);
  dpre[i] = g;
  dpre_b[i] = (__bf16)g;
}

__global__ void k_to_bf16(const float* __restrict__ x, size_t n, __bf16* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

// copy rows of a [R x C] fp32 matrix into columns [c0, c0 + C) of a
// [R x ldo] bf16 matrix (MLP inputs), optionally through a row index; each
// thread moves 4 consecutive columns (C, ldo, c0 multiples of 4)
__global__ void k_pack_cols(const float* __restrict__ x, const int* __restrict__ idx, int R, int C,
                            __bf16* __restrict__ out, int ldo, int c0) {
  const int q = C >> 2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)R * q) return;
  const int r = (int)(i / q), c = (int)(i - (size_t)r * q) * 4;
  const int src = idx ? idx[r] : r;
  const float4 v = *(const float4*)(x + (size_t)src * C + c);
  __bf16* o = out + (size_t)r * ldo + c0 + c;
  o[0] = (__bf16)v.x;
  o[1] = (__bf16)v.y;
  o[2] = (__bf16)v.z;
  o[3] = (__bf16)v.w;
}

// agg[v] = sum of rows e[k] over the edges k of receiver v (edges sorted by
// receiver: rstart[v] .. rstart[v+1])
__global__ void k_segsum(const float* __restrict__ e, const int* __restrict__ rstart, int nv, int C,
                         float* __restrict__ agg) {
  // four columns per thread; edges added in ascending order (the loads of
  // four edges issued together)
  const int q = C >> 2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nv * q) return;
  const int v = (int)(i / q), c = (int)(i % q) * 4;
  const int k0 = rstart[v], k1 = rstart[v + 1];
  float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int k = k0;
  for (; k + 4 <= k1; k += 4) {
    float4 x[4];
    for (int u = 0; u < 4; u++) x[u] = *(const float4*)(e + (size_t)(k + u) * C + c);
    for (int u = 0; u < 4; u++) {
      s.x += x[u].x;
      s.y += x[u].y;
      s.z += x[u].z;
      s.w += x[u].w;
    }
  }
  for (; k < k1; k++) {
    const float4 x = *(const float4*)(e + (size_t)k * C + c);
    s.x += x.x;
    s.y += x.y;
    s.z += x.z;
    s.w += x.w;
  }
  *(float4*)(agg + (size_t)v * C + c) = s;
}

// out[idx[r]] += x[r] (rows of C columns taken from column c0 of x, ldx wide)
__global__ void k_scatter_rows(const float* __restrict__ x, int ldx, int c0, const int* __restrict__ idx, int R,
                               int C, float* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)R * C) return;
  const int r = (int)(i / C), c = (int)(i % C);
  atomicAdd(&out[(size_t)idx[r] * C + c], x[(size_t)r * ldx + c0 + c]);
}

// y[i] += x[i]
__global__ void k_add(float* __restrict__ y, const float* __restrict__ x, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

// take columns [c0, c0 + C) of [R x ldx] into a dense [R x C]
__global__ void k_take_cols(const float* __restrict__ x, int ldx, int c0, int R, int C, float* __restrict__ y,
                            int accumulate) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)R * C) return;
  const int r = (int)(i / C), c = (int)(i % C);
  const float v = x[(size_t)r * ldx + c0 + c];
  y[i] = accumulate ? y[i] + v : v;
}

// mean squared error over the first `nout` of `ld` columns, and its gradient
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target, int R, int nout, int ld,
                      float scale, float* __restrict__ grad, float* __restrict__ loss) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < (size_t)R * ld) {
    const int r = (int)(i / ld), c = (int)(i % ld);
    float g = 0.0f;
    if (c < nout) {
      const float d = pred[i] - target[(size_t)r * nout + c];
      l = d * d * scale;
      g = 2.0f * d * scale;
    }
    grad[i] = g;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
  if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, l);
}

// Adam (fp32 master weights), then the bf16 copy the MFMAs read
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, __bf16* __restrict__ wb, size_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, float clip_scale) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * clip_scale;
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (__bf16)wi;
}

// sum of squares (gradient norm for clipping)
__global__ void k_sumsq(const float* __restrict__ x, size_t n, float* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float s = i < n ? x[i] * x[i] : 0.0f;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0 && s != 0.0f) atomicAdd(out, s);
}

// de[k] += dagg[idx[k]] (rows of C)
__global__ void k_gather_add(const float* __restrict__ src, const int* __restrict__ idx, int R, int C,
                             float* __restrict__ dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)R * C) return;
  const int r = (int)(i / C), c = (int)(i % C);
  dst[i] += src[(size_t)idx[r] * C + c];
}

// ---------------------------------------------------------------------------
// neighbour graph: cell keys, cell ranges, degree count, edge fill.  A pair
// is an edge when its distance is at most cfac * 2 * max(r_i, r_j); cells
// are at least the largest such cut-off wide; periodic dimensions wrap
// with the nearest image.  Edges are grouped by receiver; a receiver's senders come in cell
// traversal order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int gcell(const GGrid& G, int d, double x) {
  int c = (int)floor((x - G.lo[d]) * G.inv_w[d]);
  return c < 0 ? 0 : (c >= G.n[d] ? G.n[d] - 1 : c);
}
// A particle more than a cut-off outside the domain in a closed dimension
// (one the surrogate has thrown out: the DEM removes it at its next step)
// is left out of the cells and has no edges, instead of joining the edge
// cells: there such particles, in their thousands, made the search
// quadratic in their number.
__device__ __forceinline__ bool ginside(const GGrid& G, double x, double y, double z) {
  const double v[3] = {x, y, z};
  for (int d = 0; d < 3; d++)
    if (!G.per[d] && !(v[d] >= G.lo[d] - G.rc && v[d] < G.lo[d] + G.L[d] + G.rc)) return false;
  return true;
}

// Counting sort of the particles by cell: the cell of each particle and
// the cells' counts (cnt zeroed by the caller), then (after an exclusive
// scan of the counts into cstart) each particle into its cell's range, and
// each cell's range put in ascending particle order -- the order a stable
// sort by cell gives, whatever order the atomics took.
__global__ void k_cell_count(const double4* __restrict__ pos, int n, GGrid G, int* __restrict__ key,
                             int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 p = pos[i];
  if (!ginside(G, p.x, p.y, p.z)) {
    key[i] = -1;
    return;
  }
  const int k = (gcell(G, 2, p.z) * G.n[1] + gcell(G, 1, p.y)) * G.n[0] + gcell(G, 0, p.x);
  key[i] = k;
  atomicAdd(&cnt[k], 1);
}

__global__ void k_cell_fill(const int* __restrict__ key, int n, const int* __restrict__ cstart,
                            int* __restrict__ fill, int* __restrict__ sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = key[i];
  if (k >= 0) sorted[cstart[k] + atomicAdd(&fill[k], 1)] = i;
}

__global__ void k_cell_order(int64_t ncell, const int* __restrict__ cstart, const int* __restrict__ cnt,
                             int* __restrict__ cend, int* __restrict__ sorted) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int a = cstart[c], b = a + cnt[c];
  cend[c] = b;
  for (int q = a + 1; q < b; q++) {  // insertion sort: a cell holds a few particles
    const int v = sorted[q];
    int r = q - 1;
    while (r >= a && sorted[r] > v) {
      sorted[r + 1] = sorted[r];
      r--;
    }
    sorted[r + 1] = v;
  }
}

// Exclusive scan of n ints (in place allowed): tiles of SC_TILE summed and
// scanned per block (wave scans by shuffles), the tiles' sums scanned by
// one block, then added.
constexpr int SC_T = 256, SC_PER = 4, SC_TILE = SC_T * SC_PER;

__device__ __forceinline__ int block_excl_scan(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int q = 0; q < SC_T / 64; q++) {
    if (q < w) before += sh[q];
    all += sh[q];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

__global__ __launch_bounds__(SC_T) void k_scan_tile(const int* in, int* out, int64_t n, int* __restrict__ tsum) {
  __shared__ int sh[SC_T / 64];
  const int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_PER;
  int v[SC_PER], s = 0;
  for (int q = 0; q < SC_PER; q++) {
    v[q] = base + q < n ? in[base + q] : 0;
    s += v[q];
  }
  int total;
  int run = block_excl_scan(s, sh, &total);
  for (int q = 0; q < SC_PER; q++) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SC_T) void k_scan_top(int* __restrict__ tsum, int nt) {
  __shared__ int sh[SC_T / 64];
  int carry = 0;
  for (int c0 = 0; c0 < nt; c0 += SC_T) {
    const int i = c0 + (int)threadIdx.x;
    const int v = i < nt ? tsum[i] : 0;
    int total;
    const int ex = block_excl_scan(v, sh, &total);
    if (i < nt) tsum[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(SC_T) void k_scan_add(int* __restrict__ out, int64_t n, const int* __restrict__ tsum) {
  const int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_PER;
  const int add = tsum[blockIdx.x];
  for (int q = 0; q < SC_PER; q++)
    if (base + q < n) out[base + q] += add;
}

template <bool FILL>
__global__ void k_edges(const double4* __restrict__ pos, int n, GGrid G, const int* __restrict__ sorted,
                        const int* __restrict__ cstart, const int* __restrict__ cend, int* __restrict__ deg,
                        const int* __restrict__ rstart, int* __restrict__ esrc, int* __restrict__ edst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 p = pos[i];
  if (!ginside(G, p.x, p.y, p.z)) {
    if (!FILL) deg[i] = 0;
    return;
  }
  const double xv[3] = {p.x, p.y, p.z};
  int c[3];
  for (int d = 0; d < 3; d++) c[d] = gcell(G, d, xv[d]);
  int k = 0;
  const int base = FILL ? rstart[i] : 0;
  for (int dz = -1; dz <= 1; dz++) {
    int cz = c[2] + dz;
    if (cz < 0 || cz >= G.n[2]) { if (!G.per[2]) continue; cz = (cz + G.n[2]) % G.n[2]; }
    for (int dy = -1; dy <= 1; dy++) {
      int cy = c[1] + dy;
      if (cy < 0 || cy >= G.n[1]) { if (!G.per[1]) continue; cy = (cy + G.n[1]) % G.n[1]; }
      for (int dx = -1; dx <= 1; dx++) {
        int cx = c[0] + dx;
        if (cx < 0 || cx >= G.n[0]) { if (!G.per[0]) continue; cx = (cx + G.n[0]) % G.n[0]; }
        const int key = (cz * G.n[1] + cy) * G.n[0] + cx;
        const int a = cstart[key], b = cend[key];
        for (int q = a; q < b; q++) {
          const int j = sorted[q];
          if (j == i) continue;
          const double4 pj = pos[j];
          double d3[3] = {pj.x - p.x, pj.y - p.y, pj.z - p.z};
          for (int d = 0; d < 3; d++)
            if (G.per[d]) {
              if (d3[d] > 0.5 * G.L[d]) d3[d] -= G.L[d];
              else if (d3[d] < -0.5 * G.L[d]) d3[d] += G.L[d];
            }
          const double r2 = d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2];
          const double cut = G.cfac * 2.0 * (p.w > pj.w ? p.w : pj.w);
          if (r2 > cut * cut) continue;
          if (FILL) {
            esrc[base + k] = j;
            edst[base + k] = i;
          }
          k++;
        }
      }
    }
  }
  if (!FILL) deg[i] = k;
}

// node features: v / v_s (3), a / a_s (3), r / l_s, min(height over the
// floor, 4 l_s) / l_s, then zero padding to the input width
__global__ void k_node_feat(const double4* __restrict__ pos, const double* __restrict__ vel,
                            const double* __restrict__ acc, int n, int stride, FeatScale F,
                            __bf16* __restrict__ out, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 p = pos[i];
  float f[8];
  for (int d = 0; d < 3; d++) f[d] = (float)(vel[(size_t)d * stride + i] / F.v);
  for (int d = 0; d < 3; d++) f[3 + d] = (float)(acc ? acc[(size_t)d * stride + i] / F.a : 0.0);
  f[6] = (float)(p.w / F.l);
  const double h = p.z - F.floor_z;
  f[7] = (float)((h < 4.0 * F.l ? h : 4.0 * F.l) / F.l);
  for (int c = 0; c < ld; c++) out[(size_t)i * ld + c] = (__bf16)(c < 8 ? f[c] : 0.0f);
}

// edge features (sender j -> receiver i): (x_j - x_i) / l_s (nearest image),
// |x_j - x_i| / l_s, (v_j - v_i) / v_s, then zero padding
__global__ void k_edge_feat(const double4* __restrict__ pos, const double* __restrict__ vel, int stride,
                            const int* __restrict__ esrc, const int* __restrict__ edst, int ne, GGrid G,
                            FeatScale F, __bf16* __restrict__ out, int ld) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ne) return;
  const int j = esrc[k], i = edst[k];
  const double4 pj = pos[j], pi = pos[i];
  double d3[3] = {pj.x - pi.x, pj.y - pi.y, pj.z - pi.z};
  for (int d = 0; d < 3; d++)
    if (G.per[d]) {
      if (d3[d] > 0.5 * G.L[d]) d3[d] -= G.L[d];
      else if (d3[d] < -0.5 * G.L[d]) d3[d] += G.L[d];
    }
  float f[7];
  for (int d = 0; d < 3; d++) f[d] = (float)(d3[d] / F.l);
  f[3] = (float)(sqrt(d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2]) / F.l);
  for (int d = 0; d < 3; d++) f[4 + d] = (float)((vel[(size_t)d * stride + j] - vel[(size_t)d * stride + i]) / F.v);
  for (int c = 0; c < ld; c++) out[(size_t)k * ld + c] = (__bf16)(c < 7 ? f[c] : 0.0f);
}

// apply a prediction over T = nsub * dt: ballistic step plus the network's
// corrections (columns 0-2: position, 3-5: velocity, in units of the
// output scales); periodic wrap, and the floor keeps its particles.  The
// stored acceleration (a node feature of the next call) becomes the mean
// acceleration of the call.
__global__ void k_advance(double4* __restrict__ pos, double* __restrict__ vel, double* __restrict__ acc, int n,
                          int stride, const float* __restrict__ pred, int ld, FeatScale F, GGrid G, double T,
                          double gx, double gy, double gz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = pos[i];
  const double g[3] = {gx, gy, gz};
  double x[3] = {p.x, p.y, p.z};
  for (int d = 0; d < 3; d++) {
    const double v = vel[(size_t)d * stride + i];
    double xn = x[d] + v * T + 0.5 * g[d] * T * T + (double)pred[(size_t)i * ld + d] * F.dx;
    const double vn = v + g[d] * T + (double)pred[(size_t)i * ld + 3 + d] * F.dv;
    if (G.per[d]) {
      if (xn >= G.lo[d] + G.L[d]) xn -= G.L[d];
      else if (xn < G.lo[d]) xn += G.L[d];
    }
    x[d] = xn;
    vel[(size_t)d * stride + i] = vn;
    if (acc) acc[(size_t)d * stride + i] = (vn - v) / T;
  }
  const double zmin = F.floor_z + 0.5 * p.w;
  if (x[2] < zmin) x[2] = zmin;
  pos[i] = make_double4(x[0], x[1], x[2], p.w);
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline unsigned nb(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t mfma_probe(const void* a, const void* b, float* d, hipStream_t st) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, st, (const bf16x8*)a, (const bf16x8*)b, d);
  return hipGetLastError();
}

hipError_t gemm(bool ta, bool tb, const __bf16* A, const __bf16* B, int M, int N, int K, int lda, int ldb,
                int splits, const Epi& ep, hipStream_t st) {
  if (splits < 1) splits = 1;
  int kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + BK - 1) / BK * BK;
  splits = (K + kchunk - 1) / kchunk;
  if (splits < 1) splits = 1;
  dim3 grid(nb(N, BN), nb(M, BM), splits);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, st, A, B, M, N, K, lda, ldb, kchunk, ep);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, st, A, B, M, N, K, lda, ldb, kchunk, ep);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, st, A, B, M, N, K, lda, ldb, kchunk, ep);
  else hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(256), 0, st, A, B, M, N, K, lda, ldb, kchunk, ep);
  return hipGetLastError();
}

int gemm_splits(int K, int kchunk_target) {
  int s = (K + kchunk_target - 1) / kchunk_target;
  return s < 1 ? 1 : s;
}

hipError_t mlp_fused(const FusedMLP& a, hipStream_t st) {
  if (a.R <= 0) return hipSuccess;
  if (a.nseg < 1 || a.nseg > 3) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mlp_fused, dim3(nb(a.R, FR)), dim3(FNT), 0, st, a);
  return hipGetLastError();
}

hipError_t transpose_bf16(const __bf16* w, int rows, int cols, __bf16* wt, hipStream_t st) {
  hipLaunchKernelGGL(k_transpose_bf16, dim3(nb((size_t)rows * cols, 256)), dim3(256), 0, st, w, rows, cols, wt);
  return hipGetLastError();
}

hipError_t reduce_parts(const float* part, int nz, size_t len, float* out, bool accumulate, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_parts, dim3(nb(len, 256)), dim3(256), 0, st, part, nz, len, out, accumulate ? 1 : 0);
  return hipGetLastError();
}

hipError_t colsum(const float* x, int R, int C, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3(nb(C, 128), nb(R, 1024)), dim3(128), 0, st, x, R, C, out);
  return hipGetLastError();
}

hipError_t silu_bwd(const float* dy, const float* pre, size_t n, float* dpre, __bf16* dpre_b, hipStream_t st) {
  hipLaunchKernelGGL(k_silu_bwd, dim3(nb(n, 256)), dim3(256), 0, st, dy, pre, n, dpre, dpre_b);
  return hipGetLastError();
}

hipError_t to_bf16(const float* x, size_t n, __bf16* y, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3(nb(n, 256)), dim3(256), 0, st, x, n, y);
  return hipGetLastError();
}

hipError_t pack_cols(const float* x, const int* idx, int R, int C, __bf16* out, int ldo, int c0, hipStream_t st) {
  if ((C & 3) || (ldo & 3) || (c0 & 3)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pack_cols, dim3(nb((size_t)R * (C / 4), 256)), dim3(256), 0, st, x, idx, R, C, out, ldo, c0);
  return hipGetLastError();
}

hipError_t segsum(const float* e, const int* rstart, int nv, int C, float* agg, hipStream_t st) {
  if (C % 4 != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_segsum, dim3(nb((size_t)nv * (C / 4), 256)), dim3(256), 0, st, e, rstart, nv, C, agg);
  return hipGetLastError();
}

hipError_t scatter_rows(const float* x, int ldx, int c0, const int* idx, int R, int C, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_rows, dim3(nb((size_t)R * C, 256)), dim3(256), 0, st, x, ldx, c0, idx, R, C, out);
  return hipGetLastError();
}

hipError_t add(float* y, const float* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_add, dim3(nb(n, 256)), dim3(256), 0, st, y, x, n);
  return hipGetLastError();
}

hipError_t take_cols(const float* x, int ldx, int c0, int R, int C, float* y, bool accumulate, hipStream_t st) {
  hipLaunchKernelGGL(k_take_cols, dim3(nb((size_t)R * C, 256)), dim3(256), 0, st, x, ldx, c0, R, C, y,
                     accumulate ? 1 : 0);
  return hipGetLastError();
}

hipError_t mse(const float* pred, const float* target, int R, int nout, int ld, float scale, float* grad,
               float* loss, hipStream_t st) {
  hipLaunchKernelGGL(k_mse, dim3(nb((size_t)R * ld, 256)), dim3(256), 0, st, pred, target, R, nout, ld, scale,
                     grad, loss);
  return hipGetLastError();
}

hipError_t adam(float* w, const float* g, float* m, float* v, __bf16* wb, size_t n, float lr, float b1, float b2,
                float eps, float bc1, float bc2, float clip_scale, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3(nb(n, 256)), dim3(256), 0, st, w, g, m, v, wb, n, lr, b1, b2, eps, bc1, bc2,
                     clip_scale);
  return hipGetLastError();
}

hipError_t sumsq(const float* x, size_t n, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sumsq, dim3(nb(n, 256)), dim3(256), 0, st, x, n, out);
  return hipGetLastError();
}

hipError_t gather_add(const float* src, const int* idx, int R, int C, float* dst, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_add, dim3(nb((size_t)R * C, 256)), dim3(256), 0, st, src, idx, R, C, dst);
  return hipGetLastError();
}

int64_t scan_tiles(int64_t n) { return (n + SC_TILE - 1) / SC_TILE; }

hipError_t excl_scan(const int* in, int* out, int64_t n, int* tsum, hipStream_t st) {
  const int64_t nt = scan_tiles(n);
  if (nt == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)nt), dim3(SC_T), 0, st, in, out, n, tsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SC_T), 0, st, tsum, (int)nt);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nt), dim3(SC_T), 0, st, out, n, tsum);
  return hipGetLastError();
}

hipError_t cell_sort(const double4* pos, int n, const GGrid& G, int64_t ncell, int* key, int* cnt, int* fill,
                     int* cstart, int* cend, int* sorted, int* tsum, hipStream_t st) {
  hipError_t e = hipMemsetAsync(cnt, 0, (size_t)ncell * 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(fill, 0, (size_t)ncell * 4, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cell_count, dim3(nb(n, 256)), dim3(256), 0, st, pos, n, G, key, cnt);
  if ((e = excl_scan(cnt, cstart, ncell, tsum, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_cell_fill, dim3(nb(n, 256)), dim3(256), 0, st, key, n, cstart, fill, sorted);
  hipLaunchKernelGGL(k_cell_order, dim3(nb((size_t)ncell, 256)), dim3(256), 0, st, ncell, cstart, cnt, cend, sorted);
  return hipGetLastError();
}

hipError_t edges(bool fill, const double4* pos, int n, const GGrid& G, const int* sorted, const int* cstart,
                 const int* cend, int* deg, const int* rstart, int* esrc, int* edst, hipStream_t st) {
  if (fill)
    hipLaunchKernelGGL(k_edges<true>, dim3(nb(n, 128)), dim3(128), 0, st, pos, n, G, sorted, cstart, cend, deg,
                       rstart, esrc, edst);
  else
    hipLaunchKernelGGL(k_edges<false>, dim3(nb(n, 128)), dim3(128), 0, st, pos, n, G, sorted, cstart, cend, deg,
                       rstart, esrc, edst);
  return hipGetLastError();
}

hipError_t node_feat(const double4* pos, const double* vel, const double* acc, int n, int stride,
                     const FeatScale& F, __bf16* out, int ld, hipStream_t st) {
  hipLaunchKernelGGL(k_node_feat, dim3(nb(n, 256)), dim3(256), 0, st, pos, vel, acc, n, stride, F, out, ld);
  return hipGetLastError();
}

hipError_t edge_feat(const double4* pos, const double* vel, int stride, const int* esrc, const int* edst, int ne,
                     const GGrid& G, const FeatScale& F, __bf16* out, int ld, hipStream_t st) {
  hipLaunchKernelGGL(k_edge_feat, dim3(nb(ne, 256)), dim3(256), 0, st, pos, vel, stride, esrc, edst, ne, G, F,
                     out, ld);
  return hipGetLastError();
}

hipError_t advance(double4* pos, double* vel, double* acc, int n, int stride, const float* pred, int ld,
                   const FeatScale& F, const GGrid& G, double T, const double g[3], hipStream_t st) {
  hipLaunchKernelGGL(k_advance, dim3(nb(n, 256)), dim3(256), 0, st, pos, vel, acc, n, stride, pred, ld, F, G, T,
                     g[0], g[1], g[2]);
  return hipGetLastError();
}

}  // namespace gnn

// ---------------------------------------------------------------------------
// sdem_gnn_gemm_noalloc (include/sdem_gnn.h): sdem_gnn_gemm on device
// buffers, with the bf16 copies of A and B and the split-K partial sums in
// `scratch`, a fixed array of the library on every device it runs on, so
// that nothing is allocated.  Host code reaches the array's device memory
// only through the address hipGetSymbolAddress gives for the current
// device (kept per device in gemm_scratch), never through its name.
// ---------------------------------------------------------------------------
#include <mutex>

#include "sdem_gnn.h"

static __device__ __align__(256) unsigned char scratch[SDEM_GNN_GEMM_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
void* scratch_ptr[MAX_DEVICES] = {};

// the address of this device's scratch array, looked up once per device
hipError_t gemm_scratch(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_ptr[dev] = p;
  }
  *out = (unsigned char*)scratch_ptr[dev];
  return hipSuccess;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" int sdem_gnn_gemm_noalloc(int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const float* a,
                                     const float* b, const float* bias, int32_t flags, int32_t splits, float* c,
                                     float* pre, void* stream) {
  using namespace gnn;
  if (!a || !b || !c || M < 1 || N < 1 || K < 1) return SDEM_E_ARG;
  if ((flags & EPI_BIAS) && !bias) return SDEM_E_ARG;
  // the K chunks of gemm(): whole multiples of its K stage
  int sp = splits < 1 ? 1 : splits;
  int kchunk = (K + sp - 1) / sp;
  kchunk = (kchunk + BK - 1) / BK * BK;
  const int nz = (K + kchunk - 1) / kchunk;
  const size_t na = (size_t)M * K, nbb = (size_t)K * N, nc = (size_t)M * N;
  const size_t off_b = align256(na * 2), off_p = off_b + align256(nbb * 2);
  const size_t need = off_p + (nz > 1 ? (size_t)nz * nc * 4 : 0);
  if (need > (size_t)SDEM_GNN_GEMM_SCRATCH_BYTES) return SDEM_E_CAPACITY;
  unsigned char* base = nullptr;
  hipError_t e = gemm_scratch(&base);
  if (e != hipSuccess) return SDEM_E_HIP;
  hipStream_t st = (hipStream_t)stream;
  __bf16* sa = (__bf16*)base;
  __bf16* sb = (__bf16*)(base + off_b);
  float* part = nz > 1 ? (float*)(base + off_p) : c;
  // every scratch byte read below is written first in the same call: no
  // clearing is needed, and nothing of an earlier call is read
  e = to_bf16(a, na, sa, st);
  if (e == hipSuccess) e = to_bf16(b, nbb, sb, st);
  int fl = flags & (EPI_BIAS | EPI_SILU | EPI_PRE);
  if (!pre || nz > 1) fl &= ~EPI_PRE;  // the pre-activation of one chunk alone means nothing
  const int lda = ta ? M : K, ldb = tb ? K : N;
  Epi ep{fl, bias, pre, nullptr, part, nullptr};
  if (e == hipSuccess) e = gemm(ta != 0, tb != 0, sa, sb, M, N, K, lda, ldb, sp, ep, st);
  if (e == hipSuccess && nz > 1) e = reduce_parts(part, nz, nc, c, false, st);
  return e == hipSuccess ? SDEM_OK : SDEM_E_HIP;
}

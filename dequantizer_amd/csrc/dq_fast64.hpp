// Complex128 tiles on FP64 matrix cores for degree-3, chi = 16, phys = 2
// node tensors (V_MFMA_F64_16X16X4F64: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; the result D[l / 16 + 4 v][l % 16] stays in four
// accumulators).  One physical component of a node tensor at a time: the
// 16^3 complex128 tensor, element (b0, b1, b2) at b0*273 + b1*17 + b2, so
// that 16 lanes walking any axis hit distinct banks with 16-byte elements.
#pragma once
#include "dq_internal.hpp"

namespace dq {

namespace fast64 {
constexpr int RP = 17;
constexpr int B0P = 16 * RP + 1;
constexpr int NPAD = 16 * B0P;
constexpr int NW = 8;
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int axis_stride(int j) { return j == 0 ? B0P : (j == 1 ? RP : 1); }
// column c (0..255) of the matrix view of mode j: the two other axes, row-major
__device__ __forceinline__ int col_base(int j, int c) {
  if (j == 0) return (c >> 4) * RP + (c & 15);
  if (j == 1) return (c >> 4) * B0P + (c & 15);
  return (c >> 4) * B0P + (c & 15) * RP;
}
__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// A operand of dst = M' src along an axis: m[ks] = M'[lane % 16][4 ks + lane / 16]
// with M' = M (op 0), conj(M) (op 1), M^T (op 2) or M^H (op 3); M 16 x 16
// row-major.
__device__ __forceinline__ void load_operand(const cd* M, int op, int lane, cd (&m)[4]) {
  const int i = lane & 15;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const int k = 4 * ks + (lane >> 4);
    const cd v = (op & 2) ? M[k * 16 + i] : M[i * 16 + k];
    m[ks] = (op & 1) ? conj(v) : v;
  }
}
// stage physical component s of a node tensor (global, (b0, b1, b2, s)
// row-major) into the padded LDS layout
__device__ __forceinline__ void stage_component(cd* dst, const cd* tg, int s) {
  for (int f = threadIdx.x; f < 4096; f += blockDim.x) dst[(f >> 8) * B0P + ((f >> 4) & 15) * RP + (f & 15)] = tg[2 * f + s];
}
// dst = src x_j M (may be in place); m[ks] = M[lane % 16][4 ks + lane / 16]
__device__ __forceinline__ void mode_product(cd* dst, const cd* src, const cd (&m)[4], int j, int wave, int lane) {
  const int st = axis_stride(j);
  const int col = lane & 15, kq = lane >> 4;
  for (int tile = wave; tile < 16; tile += NW) {
    const int cb = col_base(j, tile * 16 + col);
    d4 ar = {0.0, 0.0, 0.0, 0.0}, ai = ar;
    cd x[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) x[ks] = src[cb + (4 * ks + kq) * st];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      ar = mfma(m[ks].re, x[ks].re, ar);
      ar = mfma(-m[ks].im, x[ks].im, ar);
      ai = mfma(m[ks].re, x[ks].im, ai);
      ai = mfma(m[ks].im, x[ks].re, ai);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[cb + (kq + 4 * v) * st] = mk<double>(ar[v], ai[v]);
  }
}
// U = X x_j M fused with this wave's share of the Gram over axis idx,
// G[k][l] += sum conj(T[k@idx, ...]) U[l@idx, ...]: tiles enumerate the third
// axis tax, lane l%16 runs along idx; the product's D[l/16 + 4v][l%16] =
// U[b_j = l/16 + 4v][b_idx = l%16] is, for each v, the B operand of a Gram
// k-step over b_j in 4v..4v+3, with conj(T) at the same (b_j, b_tax) as the
// A operand.
// With kStore the product tile is also written back over its input (in
// place: each wave reads its whole tile before writing it).
template <bool kStore = false>
__device__ __forceinline__ void mode_gram(const cd* t, cd* x, const cd (&m)[4], int j, int idx, int tax, int wave,
                                          int lane, d4& gr, d4& gi) {
  const int sj = axis_stride(j);
  const int y = lane & 15, kq = lane >> 4;
  const int lane_base = y * axis_stride(idx);
  for (int tile = wave; tile < 16; tile += NW) {
    const int cb = lane_base + tile * axis_stride(tax);
    cd xv[4], tv[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) xv[ks] = x[cb + (4 * ks + kq) * sj];
#pragma unroll
    for (int v = 0; v < 4; ++v) tv[v] = t[cb + (kq + 4 * v) * sj];
    d4 ur = {0.0, 0.0, 0.0, 0.0}, ui = ur;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      ur = mfma(m[ks].re, xv[ks].re, ur);
      ur = mfma(-m[ks].im, xv[ks].im, ur);
      ui = mfma(m[ks].re, xv[ks].im, ui);
      ui = mfma(m[ks].im, xv[ks].re, ui);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      gr = mfma(tv[v].re, ur[v], gr);
      gr = mfma(tv[v].im, ui[v], gr);
      gi = mfma(tv[v].re, ui[v], gi);
      gi = mfma(-tv[v].im, ur[v], gi);
    }
    if (kStore) {
#pragma unroll
      for (int v = 0; v < 4; ++v) x[cb + (kq + 4 * v) * sj] = mk<double>(ur[v], ui[v]);
    }
  }
}
// this wave's share of sum_r conj(T[k][r]) U[l][r] along axis idx
__device__ __forceinline__ void gram_partial(const cd* t, const cd* u, int idx, int wave, int lane, d4& gr, d4& gi) {
  const int st = axis_stride(idx);
  const int row = lane & 15, kq = lane >> 4;
  constexpr int per = 64 / NW;
  for (int ks = wave * per; ks < wave * per + per; ++ks) {
    const int off = col_base(idx, 4 * ks + kq) + row * st;
    const cd a = t[off], b = u[off];
    gr = mfma(a.re, b.re, gr);
    gr = mfma(a.im, b.im, gr);
    gi = mfma(a.re, b.im, gi);
    gi = mfma(-a.im, b.re, gi);
  }
}
}  // namespace fast64

}  // namespace dq

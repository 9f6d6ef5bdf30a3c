// One-site gates and the two-site simple update on a Vidal canonical pair.
//
// apply1         reference dequantizer/src/graph/tensor_graph.py:268-275
// simple update  dequantizer/src/tensor_ops.py:221-237 (half gate),
//                :265-295 (per-side preprocessing: lambda scaling, QR,
//                safe inverse), :318-389 (core SVD, truncation, entropy,
//                error, recomposition).
// One block per gate; the gates of one launch sit on pairwise disjoint edges
// (no shared node), so they commute exactly and may run concurrently.
#include <float.h>

#include "dq_launch.hpp"
#include "dq_linalg.hpp"

namespace dq {

template <typename R>
__global__ void __launch_bounds__(256)
    k_apply1(DevGraph g, cx<R>* T, const int* nodes, const cd* gates) {
  __shared__ double red[8];
  const int node = nodes[blockIdx.x];
  const int d = g.deg[node], p = g.phys;
  const long long rows = ipow(g.chi, d);
  cx<R>* t = T + g.toff[node];
  const cd* G = gates + (long long)blockIdx.x * p * p;
  double acc = 0.0;
  constexpr int KR = 16;  // rows per thread kept in registers (qubit fast path)
  if (p == 2 && rows <= (long long)KR * blockDim.x) {
    // one read and one write of the tensor: the gated rows stay in registers
    // until the norm is known (same arithmetic and summation order as below)
    const cx<R> g00 = cvt<R>(G[0]), g01 = cvt<R>(G[1]), g10 = cvt<R>(G[2]), g11 = cvt<R>(G[3]);
    cx<R> v0[KR], v1[KR];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const long long r = threadIdx.x + (long long)k * blockDim.x;
      if (r < rows) {
        const cx<R> a = t[2 * r], b = t[2 * r + 1];
        v0[k] = mk<R>(0, 0);
        fma_acc(v0[k], g00, a);
        fma_acc(v0[k], g01, b);
        v1[k] = mk<R>(0, 0);
        fma_acc(v1[k], g10, a);
        fma_acc(v1[k], g11, b);
        acc += (double)norm2(v0[k]);
        acc += (double)norm2(v1[k]);
      }
    }
    const double nrm = sqrt(block_sum(acc, red));
    const R inv = (R)(1.0 / nrm);
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const long long r = threadIdx.x + (long long)k * blockDim.x;
      if (r < rows) {
        t[2 * r] = inv * v0[k];
        t[2 * r + 1] = inv * v1[k];
      }
    }
    return;
  }
  cx<R> in[16];
  for (long long r = threadIdx.x; r < rows; r += blockDim.x) {
    for (int i = 0; i < p; ++i) in[i] = t[r * p + i];
    for (int o = 0; o < p; ++o) {
      cx<R> v = mk<R>(0, 0);
      for (int i = 0; i < p; ++i) fma_acc(v, cvt<R>(G[o * p + i]), in[i]);
      t[r * p + o] = v;
      acc += (double)norm2(v);
    }
  }
  const double nrm = sqrt(block_sum(acc, red));
  const R inv = (R)(1.0 / nrm);
  for (long long f = threadIdx.x; f < rows * p; f += blockDim.x) t[f] = inv * t[f];
}

__device__ __forceinline__ double su_safe_inverse(double l, double eps) {
  return 1.0 / ((l < eps ? (double)FLT_MAX : 0.0) + l);
}

// Slot of edge e (at side `side`) in the adjacency of `node`.

// ---------------------------------------------------------------------------
// Simple update, one block per gate.  The reference factorises each side
// A = (T * lambda_others) x gate-half (rows: other bonds + phys, columns:
// bond * r + q) by QR and takes the SVD of core = R_a diag(lambda) R_b^T
// (tensor_ops.py:265-389).  Only R up to a left unitary matters, so R is taken
// from the Gram matrix G = A^H A by pivoted Cholesky (R^H R = G); G is assembled from
// the environment E = X^H X of X = T * lambda_others (rows: other bonds,
// columns: (bond, phys)) and the gate half.  With core = U Sigma V^H,
//   Gamma_a = A_a Lambda R_b^T V Sigma^-1,   Gamma_b = A_b Lambda R_a^T conj(U) Sigma^-1
// (the reference's Q_a U and Q_b conj(V)); A_a never needs to be formed:
// Gamma_a[o, j, s'] = sum_{b,s} T[o, b, s] Z_a[(b,s), (j,s')] with
// Z_a = sum_q half[q, s', s] W_a[(b,q), j].  Columns of Gamma for zero
// singular values are set to zero (the reference leaves arbitrary
// orthonormal columns there; they only ever meet a zero lambda).

__device__ __forceinline__ double lam_mask(double l, double eps) {
  // lambda * safe_inverse(lambda): 1 above eps, ~lambda / FLT_MAX below
  return l * su_safe_inverse(l, eps);
}

// Stage rows [r0, r0+nr) of X[o][(b,s)] = prod_{j != slot} lam_j[o_j] * T[o, b@slot, s]
// into buf (nr x cp, row-major, complex double).  Rows enumerate the other
// bonds in order, last one fastest.
template <typename R>
__device__ void su_stage_rows(const DevGraph& g, const cx<R>* t, int node, int slot, const double* wl, bool weighted,
                              int r0, int nr, cd* buf) {
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const int cp = chi * p;
  for (int f = threadIdx.x; f < nr * cp; f += blockDim.x) {
    const int rr = f / cp, col = f - rr * cp;
    const int b = col / p, s = col - b * p;
    int o = r0 + rr;
    int tidx = s, mul = p;
    double wt = 1.0;
    for (int j = d - 1; j >= 0; --j) {
      int bj;
      if (j == slot) {
        bj = b;
      } else {
        const int q = o / chi;
        bj = o - q * chi;
        o = q;
        if (weighted) wt *= wl[j * chi + bj];
      }
      tidx += bj * mul;
      mul *= chi;
    }
    buf[f] = wt * cvt<double>(t[tidx]);
  }
}

// E[(b,s),(b',s')] = sum_o conj(X[o][(b,s)]) X[o][(b',s')]  (cp x cp, cp <= 64).
template <typename R>
__device__ void su_environment(const DevGraph& g, const cx<R>* t, int node, int slot, const double* wl, cd* E,
                               cd* chunk, int chunk_rows) {
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const int cp = chi * p;
  const int rows = (int)ipow(chi, d - 1);
  if (cp == 32 && blockDim.x == 256) {
    // 2x2 register tiles: thread -> rows x0, x0+1 and columns y0, y0+1 of E
    const int tx = threadIdx.x >> 4, ty = threadIdx.x & 15;
    const int x0 = 2 * tx, y0 = 2 * ty;
    cd a00 = mk<double>(0, 0), a01 = a00, a10 = a00, a11 = a00;
    for (int r0 = 0; r0 < rows; r0 += chunk_rows) {
      const int nr = (rows - r0) < chunk_rows ? (rows - r0) : chunk_rows;
      su_stage_rows<R>(g, t, node, slot, wl, true, r0, nr, chunk);
      __syncthreads();
#pragma unroll 4
      for (int r = 0; r < nr; ++r) {
        const cd u0 = chunk[r * 32 + x0], u1 = chunk[r * 32 + x0 + 1];
        const cd v0 = chunk[r * 32 + y0], v1 = chunk[r * 32 + y0 + 1];
        fma_cacc(a00, u0, v0);
        fma_cacc(a01, u0, v1);
        fma_cacc(a10, u1, v0);
        fma_cacc(a11, u1, v1);
      }
      __syncthreads();
    }
    E[x0 * 32 + y0] = a00;
    E[x0 * 32 + y0 + 1] = a01;
    E[(x0 + 1) * 32 + y0] = a10;
    E[(x0 + 1) * 32 + y0 + 1] = a11;
    __syncthreads();
    return;
  }
  constexpr int kMaxPer = 16;  // (cp^2) / 256 for cp <= 64
  cd acc[kMaxPer];
#pragma unroll
  for (int i = 0; i < kMaxPer; ++i) acc[i] = mk<double>(0, 0);
  for (int r0 = 0; r0 < rows; r0 += chunk_rows) {
    const int nr = (rows - r0) < chunk_rows ? (rows - r0) : chunk_rows;
    su_stage_rows<R>(g, t, node, slot, wl, true, r0, nr, chunk);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kMaxPer; ++i) {
      const int q = threadIdx.x + i * 256;
      if (q < cp * cp) {
        const int x = q / cp, y = q - x * cp;
        cd a = acc[i];
        for (int r = 0; r < nr; ++r) fma_cacc(a, chunk[r * cp + x], chunk[r * cp + y]);
        acc[i] = a;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < kMaxPer; ++i) {
    const int q = threadIdx.x + i * 256;
    if (q < cp * cp) E[q] = acc[i];
  }
  __syncthreads();
}

// ------------------------------------------------ FP64 matrix-core paths
// For cp = chi * phys = 32 and chi^(d-1) rows a multiple of 16, the two
// GEMM-shaped steps of a side run on v_mfma_f64_16x16x4_f64 (exact FP64
// products and sums, like the FMA paths).  Lane l of a 16x16x4 step holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; the result D[l / 16 + 4 v][l % 16],
// v = 0..3, stays in its four accumulator registers (rows interleaved by 4,
// unlike the FP32 16x16x4 variant).  Operands come straight
// from the node tensor (L2-resident) with the lambda weights applied in FP64.
typedef double su_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ su_d4 su_mfma(double a, double b, su_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Rows o of the other bonds enumerate them in row-major order, so with
// chi % 16 == 0 the 16 rows o0 .. o0 + 15 of a tile (o0 % 16 == 0) differ only
// in the index of the last non-slot bond jl: b_jl = o0 % chi + delta.  Returns
// the offset and fn(lambda) product of the other bonds of row o0, and jl, b0.
template <bool kMask>
__device__ __forceinline__ long long su_tile_row(int d, int chi, int slot, const int* str, const double* wl,
                                                 double eps, int o0, double& whi, int& jl, int& b0) {
  jl = slot == d - 1 ? d - 2 : d - 1;
  b0 = o0 % chi;
  int o = o0 / chi;
  long long off = 0;
  whi = 1.0;
  for (int j = jl - 1; j >= 0; --j) {
    if (j == slot) continue;
    const int q = o / chi, bj = o - q * chi;
    o = q;
    off += (long long)bj * str[j];
    const double l = wl[j * chi + bj];
    whi *= kMask ? lam_mask(l, eps) : l;
  }
  return off;
}

// E (32 x 32) = X^H X, X[o][(b,s)] = t[...] * prod_{j != slot} lambda_j[b_j].
// Wave w computes the 16x16 tile (w / 2, w % 2); 4 MFMAs per 4 rows.
template <typename R>
__device__ void su_environment_mfma(const DevGraph& g, const cx<R>* t, int node, int slot, const double* wl,
                                    cd* E) {
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const int rows = (int)ipow(chi, d - 1);
  int str[kMaxDeg];
  {
    int m = p;
    for (int j = d - 1; j >= 0; --j) {
      str[j] = m;
      m *= chi;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int x0 = (w >> 1) * 16, y0 = (w & 1) * 16;
  const int i = lane & 15, kk = lane >> 4;
  const long long offx = (long long)((x0 + i) / p) * str[slot] + (x0 + i) % p;
  const long long offy = (long long)((y0 + i) / p) * str[slot] + (y0 + i) % p;
  su_d4 are = {0.0, 0.0, 0.0, 0.0}, aim = {0.0, 0.0, 0.0, 0.0};
  // the loads of row group o0 + 16 are issued before the MFMAs of group o0
  cx<R> na[4], nb[4];
  double nw[4];
  auto load_group = [&](int o0) {
    double whi;
    int jl, b0;
    const long long offh = su_tile_row<false>(d, chi, slot, str, wl, 0.0, o0, whi, jl, b0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + 4 * u + kk;
      const long long off = offh + (long long)b * str[jl];
      nw[u] = whi * wl[jl * chi + b];
      na[u] = t[off + offx];
      nb[u] = t[off + offy];
    }
  };
  load_group(0);
  for (int o0 = 0; o0 < rows; o0 += 16) {
    cx<R> va[4], vb[4];
    double wt[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      va[u] = na[u];
      vb[u] = nb[u];
      wt[u] = nw[u];
    }
    if (o0 + 16 < rows) load_group(o0 + 16);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double ar = wt[u] * (double)va[u].re, ai = wt[u] * (double)va[u].im;
      const double br = wt[u] * (double)vb[u].re, bi = wt[u] * (double)vb[u].im;
      are = su_mfma(ar, br, are);
      are = su_mfma(ai, bi, are);
      aim = su_mfma(ar, bi, aim);
      aim = su_mfma(-ai, br, aim);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) E[(x0 + kk + 4 * v) * 32 + y0 + i] = mk<double>(are[v], aim[v]);
  __syncthreads();
}

// Gamma rows o (other bonds) x (j, s') = scale * mask(o) * sum_x X[o][x] Z[x][(j,s')],
// X unweighted, in place.  Wave w owns the row tiles w, w+4, ... (both column
// tiles), so every element it overwrites was read by the same wave first.
template <typename R>
__device__ void su_write_gamma_mfma(const DevGraph& g, cx<R>* t, int node, int slot, const double* wl, double eps,
                                    const cd* Z, double scale) {
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const int rows = (int)ipow(chi, d - 1);
  int str[kMaxDeg];
  {
    int m = p;
    for (int j = d - 1; j >= 0; --j) {
      str[j] = m;
      m *= chi;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, kk = lane >> 4;
  // the loads of tile rt + 4 are issued before the MFMAs of tile rt; a tile
  // is read before it is overwritten, by the same wave
  cx<R> nv[8];
  auto load_tile = [&](int rt) {
    double unused;
    int jl, b0;
    const long long offa =
        su_tile_row<false>(d, chi, slot, str, wl, 0.0, rt * 16, unused, jl, b0) + (long long)(b0 + i) * str[jl];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int x = 4 * u + kk;
      nv[u] = t[offa + (long long)(x / p) * str[slot] + x % p];
    }
  };
  if (w * 16 < rows) load_tile(w);
  for (int rt = w; rt * 16 < rows; rt += 4) {
    su_d4 r0 = {0.0, 0.0, 0.0, 0.0}, i0 = r0, r1 = r0, i1 = r0;
    cx<R> va[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) va[u] = nv[u];
    if ((rt + 4) * 16 < rows) load_tile(rt + 4);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int x = 4 * u + kk;
      const double ar = (double)va[u].re, ai = (double)va[u].im;
      const cd z0 = Z[x * 32 + i], z1 = Z[x * 32 + 16 + i];
      r0 = su_mfma(ar, z0.re, r0);
      r0 = su_mfma(-ai, z0.im, r0);
      i0 = su_mfma(ar, z0.im, i0);
      i0 = su_mfma(ai, z0.re, i0);
      r1 = su_mfma(ar, z1.re, r1);
      r1 = su_mfma(-ai, z1.im, r1);
      i1 = su_mfma(ar, z1.im, i1);
      i1 = su_mfma(ai, z1.re, i1);
    }
    double mhi;
    int jl, b0;
    const long long offh = su_tile_row<true>(d, chi, slot, str, wl, eps, rt * 16, mhi, jl, b0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int b = b0 + kk + 4 * v;
      const long long off = offh + (long long)b * str[jl];
      const double f = scale * mhi * lam_mask(wl[jl * chi + b], eps);
      const int c0 = i, c1 = 16 + i;
      t[off + (long long)(c0 / p) * str[slot] + c0 % p] = cvt<R>(f * mk<double>(r0[v], i0[v]));
      t[off + (long long)(c1 / p) * str[slot] + c1 % p] = cvt<R>(f * mk<double>(r1[v], i1[v]));
    }
  }
  __syncthreads();
}

// G = E (x) K combined with the gate half h (r, p, p): G[(b,q),(b',q')] =
// sum_{s,s'} E[(b,s),(b',s')] sum_o conj(h[q,o,s]) h[q',o,s'];  G /= trace(G).
// Returns trace(G) before normalisation (= |A|^2).
__device__ double su_gram(const cd* E, const cd* h, int chi, int p, int r, cd* G, double* red) {
  const int n = chi * r, cp = chi * p;
  double tr = 0.0;
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int x = f / n, y = f % n;
    const int b = x / r, q = x % r, b2 = y / r, q2 = y % r;
    cd acc = mk<double>(0, 0);
    for (int s = 0; s < p; ++s)
      for (int s2 = 0; s2 < p; ++s2) {
        cd k = mk<double>(0, 0);
        for (int o = 0; o < p; ++o) fma_cacc(k, h[(q * p + o) * p + s], h[(q2 * p + o) * p + s2]);
        fma_acc(acc, E[(b * p + s) * cp + b2 * p + s2], k);
      }
    G[f] = acc;
    if (x == y) tr += acc.re;
  }
  tr = block_sum(tr, red);
  const double inv = tr > 0.0 ? 1.0 / tr : 0.0;
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int x = f / n, y = f % n;
    if (x <= y) {
      const cd a = G[f], b = G[y * n + x];
      const cd hm = mk<double>(0.5 * (a.re + b.re) * inv, 0.5 * (a.im - b.im) * inv);
      G[f] = hm;
      G[y * n + x] = conj(hm);
    }
  }
  __syncthreads();
  return tr;
}

// R = diag(sqrt(mu)) V^H from the eigendecomposition left in (G, V).
__device__ void su_rfactor(const cd* G, const cd* V, int n, cd* R) {
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int i = f / n, c = f % n;
    const double mu = G[i * n + i].re;
    const double sv = mu > 0.0 ? sqrt(mu) : 0.0;
    R[f] = sv * conj(V[c * n + i]);
  }
  __syncthreads();
}

// Gamma of one side: rows o of the other bonds, new bond j at `slot`.
// Z[(b,s),(j,s')] (cp x cp, complex double) holds the contraction with the
// gate half and W; scale = 1 / sqrt(trace G).
template <typename R>
__device__ void su_write_gamma(const DevGraph& g, cx<R>* t, int node, int slot, const double* wl, double eps,
                               const cd* Z, double scale, cd* chunk, int chunk_rows) {
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const int cp = chi * p;
  const int rows = (int)ipow(chi, d - 1);
  for (int r0 = 0; r0 < rows; r0 += chunk_rows) {
    const int nr = (rows - r0) < chunk_rows ? (rows - r0) : chunk_rows;
    su_stage_rows<R>(g, t, node, slot, wl, false, r0, nr, chunk);
    __syncthreads();
    if (cp == 32 && nr == 32 && blockDim.x == 256) {
      // 2 x 2 register tiles: rows rr0, rr0+1 (of 32) x columns c0, c0+1
      const int rr0 = 2 * (threadIdx.x >> 4), c0 = 2 * (threadIdx.x & 15);
      cd a00 = mk<double>(0, 0), a01 = a00, a10 = a00, a11 = a00;
#pragma unroll 8
      for (int x = 0; x < 32; ++x) {
        const cd u0 = chunk[rr0 * 32 + x], u1 = chunk[(rr0 + 1) * 32 + x];
        const cd z0 = Z[x * 32 + c0], z1 = Z[x * 32 + c0 + 1];
        fma_acc(a00, u0, z0);
        fma_acc(a01, u0, z1);
        fma_acc(a10, u1, z0);
        fma_acc(a11, u1, z1);
      }
      const cd av[4] = {a00, a01, a10, a11};
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4) {
        const int rr = rr0 + (q4 >> 1), col = c0 + (q4 & 1);
        const int jn = col / p, so = col - jn * p;
        int o = r0 + rr;
        int tidx = so, mul = p;
        double wt = scale;
        for (int j = d - 1; j >= 0; --j) {
          int bj;
          if (j == slot) {
            bj = jn;
          } else {
            const int qq = o / chi;
            bj = o - qq * chi;
            o = qq;
            wt *= lam_mask(wl[j * chi + bj], eps);
          }
          tidx += bj * mul;
          mul *= chi;
        }
        t[tidx] = cvt<R>(wt * av[q4]);
      }
      __syncthreads();
      continue;
    }
    for (int f = threadIdx.x; f < nr * cp; f += blockDim.x) {
      const int rr = f / cp, col = f - rr * cp;
      const int jn = col / p, so = col - jn * p;
      cd acc = mk<double>(0, 0);
      if (cp == 32) {
#pragma unroll 8
        for (int x = 0; x < 32; ++x) fma_acc(acc, chunk[rr * 32 + x], Z[x * 32 + col]);
      } else {
        for (int x = 0; x < cp; ++x) fma_acc(acc, chunk[rr * cp + x], Z[x * cp + col]);
      }
      int o = r0 + rr;
      int tidx = so, mul = p;
      double wt = scale;
      for (int j = d - 1; j >= 0; --j) {
        int bj;
        if (j == slot) {
          bj = jn;
        } else {
          const int q = o / chi;
          bj = o - q * chi;
          o = q;
          wt *= lam_mask(wl[j * chi + bj], eps);
        }
        tidx += bj * mul;
        mul *= chi;
      }
      t[tidx] = cvt<R>(wt * acc);
    }
    __syncthreads();
  }
}

// SVD of the 32 x 32 core C (row-major) preconditioned by a pivoted QR:
// C P = Q R (block_qrcp_rh32), the one-sided Jacobi SVD of X = R^H,
// X = U_x Sigma V_x^H (converging in a few sweeps where Jacobi on C needs about
// eight), then C = (Q V_x) Sigma (P U_x)^H.  Outputs as block_jacobi_svd_n<32>
// on C: sigma descending, U (column-major, sorted, zero columns for sigma = 0)
// and V with perm, perm here being the identity.  Buffers (C, U, V in LDS):
// U holds the reflectors until the end, X and V_x live in V and C, U_x in W
// (any memory).
// single (factors that end in single precision): the sweeps stop at the
// single-precision rule of block_jacobi_svd_n (kLoose).  The rounds follow
// the block-cyclic ordering (kBlockOrder).
__device__ __attribute__((noinline)) void su_core_svd_qr32(lds_cd* C, lds_cd* U, lds_cd* V, cd* W, double* sigma,
                                                             int* perm, double* scal, bool single) {
  __shared__ int qpiv[32];
  block_qrcp_rh32(C, U, V, (lds_int*)qpiv);
  if (single)
    block_jacobi_svd_n<32, lds_cd, true, true, true, true>(V, W, C, sigma, perm, scal);
  else
    block_jacobi_svd_n<32, lds_cd, true, true, false, true>(V, W, C, sigma, perm, scal);
  // U_C[:, j] = Q V_x[:, perm[j]], Q = H_0 ... H_31 applied from H_31 down:
  // 8 lanes per column, rows sub + 8 t
  const int j = threadIdx.x >> 3, sub = threadIdx.x & 7;
  const int pj = perm[j];
  cd y[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) y[t] = cld(C, pj * 32 + sub + 8 * t);
  for (int k = 31; k >= 0; --k) {
    cd u[4];
    double dr = 0.0, di = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      u[t] = cld(U, k * 32 + sub + 8 * t);
      dr += u[t].re * y[t].re + u[t].im * y[t].im;
      di += u[t].re * y[t].im - u[t].im * y[t].re;
    }
    dr = group_sum<8>(dr);
    di = group_sum<8>(di);
    const cd dot = mk<double>(dr, di);
#pragma unroll
    for (int t = 0; t < 4; ++t) y[t] = y[t] - u[t] * dot;
  }
  __syncthreads();  // every reflector and V_x read is done
  const bool nz = sigma[j] > 0.0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int m = sub + 8 * t;
    cst(U, j * 32 + m, nz ? y[t] : mk<double>(0.0, 0.0));
    // V_C[:, j] = P U_x[:, j]: row qpiv[m] of V_C is row m of U_x
    cst(V, j * 32 + qpiv[m], W[j * 32 + m]);
  }
  __syncthreads();
  if (threadIdx.x < 32) perm[threadIdx.x] = threadIdx.x;
  __syncthreads();
}

// SVD of the 64 x 64 core (column-major, in the workspace) with the core and
// the right vectors staged in LDS: the block-level one-sided Jacobi (eight
// lanes per column pair), outputs as jacobi_svd (U column-major and sorted,
// V with perm).  scal: 64 doubles.
__device__ __attribute__((noinline)) void su_core_svd64(const cd* C, cd* U, cd* Vs, double* sigma, int* perm,
                                                          double* scal) {
  __shared__ cd G64[64 * 64], V64[64 * 64];
  for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) cst((lds_cd*)G64, i, C[i]);
  __syncthreads();
  block_jacobi_svd_n<64, lds_cd, true, true, false, true>((lds_cd*)G64, U, (lds_cd*)V64, sigma, perm, scal);
  for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) Vs[i] = cld((lds_cd*)V64, i);
  __syncthreads();
}

// Matrices of one gate (complex doubles): E (cp^2 or n^2; later U, then Z),
// G (n^2; Gram, core, then W_a), V (n^2; right singular vectors, then W_b),
// Ra, Rb (n^2 each), chunk (rows x cp, generic environment path only).
// Storage modes:
//   kSuGlobal   all in the global workspace;
//   kSuLds      all in LDS (n, cp <= 32; ~112 KiB of LDS, one block per CU);
//   kSuCompact  E, G, V in LDS (48 KiB), Ra, Rb, chunk in the workspace
//               (L2-resident): two blocks per CU.  For cp = 32, chi % 16 == 0
//               (the matrix-core shapes), n <= 32;
//   kSuWide     storage as kSuGlobal, the n = 64 core SVD (chi = 32 with rank-2
//               gates, chi = 16 with rank-4 gates) by the block-level Jacobi
//               on the core and V in LDS (128 KiB, one block per CU).
constexpr int kSuGlobal = 0, kSuLds = 1, kSuCompact = 2, kSuWide = 3;
constexpr int kSuChunk = 32;
constexpr int kSuLdsN = 32;

__host__ __device__ inline long long su_e_cd(long long n, long long cp) { return cp * cp > n * n ? cp * cp : n * n; }

int su_mem_mode(int chi, int phys, int max_r) {
  if (chi * phys == 32 && chi % 16 == 0 && chi * max_r <= kSuLdsN) return kSuCompact;
  if (chi * max_r <= kSuLdsN && chi * phys <= kSuLdsN) return kSuLds;
  if (chi * max_r == 64) return kSuWide;
  return kSuGlobal;
}

// Workspace complex doubles per gate in each mode.
__host__ __device__ inline long long su_ws_cd(int mode, int chi, int phys, int max_r) {
  const long long n = (long long)chi * max_r, cp = (long long)chi * phys;
  if (mode == kSuLds) return 16;
  if (mode == kSuCompact) return 2 * n * n + kSuChunk * cp + 64;
  return su_e_cd(n, cp) + 4 * n * n + kSuChunk * cp + 64;
}

template <typename R, int kMem>
__global__ void __launch_bounds__(256, kMem == kSuCompact ? 2 : 1)
    k_su(DevGraph g, cx<R>* T, double* lam, const SuGate* gates, const cd* halves, int max_r, double eps,
         int max_rank, double* err_out, double* ent_out, char* ws, long long wsb, char* tape, long long tape_bytes) {
  __shared__ cd sm[kMem == kSuLds ? (5 * kSuLdsN * kSuLdsN + kSuChunk * kSuLdsN + 64)
                                  : (kMem == kSuCompact ? 3 * kSuLdsN * kSuLdsN : 1)];
  constexpr bool kGlob = kMem == kSuGlobal || kMem == kSuWide;  // matrices in the workspace
  __shared__ double red[8];
  __shared__ double rot[4 * 64 + 8];
  __shared__ double sigma[128];
  __shared__ int flag[2];
  __shared__ int perm[128];
  __shared__ int sh_rank;
  SuGate G;
  {
    const int* gi = (const int*)(gates + blockIdx.x);
    G.edge = ldc(gi, 0);
    G.first = ldc(gi, 1);
    G.rank = ldc(gi, 2);
    G.pad = 0;
  }
  if (G.rank < 1 || G.rank > max_r) {
    // rank outside what this launch's buffers were sized for: the gate is
    // skipped and flagged with NaN error and entropy (uniform per block)
    if (threadIdx.x == 0) {
      err_out[blockIdx.x] = __builtin_nan("");
      ent_out[blockIdx.x] = __builtin_nan("");
    }
    return;
  }
  const int e = G.edge, r = G.rank, chi = g.chi, p = g.phys;
  const int n = chi * r, cp = chi * p;
  const int na = ldc(g.edge_node, 2 * e + G.first), nb = ldc(g.edge_node, 2 * e + 1 - G.first);
  const int ia = find_slot(g, na, e, G.first), ib = find_slot(g, nb, e, 1 - G.first);
  const long long hstride = (long long)p * p * p * p;
  const cd* h1 = halves + (long long)blockIdx.x * 2 * hstride;
  const cd* h2 = h1 + hstride;
  cd* wsp = (cd*)(ws + (long long)blockIdx.x * wsb);
  cd* E = kGlob ? wsp : sm;
  cd* Gm = E + su_e_cd(n, cp);
  cd* V = Gm + n * n;
  cd* Ra = kMem == kSuCompact ? wsp : V + n * n;
  cd* Rb = Ra + n * n;
  cd* chunk = Rb + n * n;
  cd* C = Gm;  // core (after both R factors)
  cd* U = E;   // left singular vectors (after both environments)
  cd* Vs = V;  // right singular vectors
  cx<R>* ta = T + g.toff[na];
  cx<R>* tb = T + g.toff[nb];
  // lambdas of the bonds of both nodes (read before this gate writes lambda_e)
  __shared__ double wla[kMaxDeg * 64], wlb[kMaxDeg * 64];
  {
    const int da = g.deg[na], db = g.deg[nb], ba = g.adj_ptr[na], bb = g.adj_ptr[nb];
    for (int f = threadIdx.x; f < da * chi; f += blockDim.x)
      wla[f] = lam[(long long)g.adj_edge[ba + f / chi] * chi + f % chi];
    for (int f = threadIdx.x; f < db * chi; f += blockDim.x)
      wlb[f] = lam[(long long)g.adj_edge[bb + f / chi] * chi + f % chi];
  }
  __syncthreads();

  const bool mfma_a = cp == 32 && chi % 16 == 0 && blockDim.x == 256 && g.deg[na] >= 2;
  const bool mfma_b = cp == 32 && chi % 16 == 0 && blockDim.x == 256 && g.deg[nb] >= 2;
  if (mfma_a)
    su_environment_mfma<R>(g, ta, na, ia, wla, E);
  else
    su_environment<R>(g, ta, na, ia, wla, E, chunk, kSuChunk);
  const double tra = su_gram(E, h1, chi, p, r, Gm, red);
  SuTape tp(tape ? tape + (long long)blockIdx.x * tape_bytes : nullptr, chi * max_r);
  double trb;
  if (!kGlob && n <= 32 && blockDim.x >= 128) {
    // both Grams first (G_b in the V buffer, free until the SVD), then the
    // two factorisations at once on waves 0 and 1, without block barriers
    if (mfma_b)
      su_environment_mfma<R>(g, tb, nb, ib, wlb, E);
    else
      su_environment<R>(g, tb, nb, ib, wlb, E, chunk, kSuChunk);
    trb = su_gram(E, h2, chi, p, r, V, red);
    const int w = threadIdx.x >> 6;
    if (w == 0)
      wave_pivoted_cholesky((lds_cd*)Gm, n, Ra, (lds_int*)perm);
    else if (w == 1)
      wave_pivoted_cholesky((lds_cd*)V, n, Rb, (lds_int*)(perm + 64));
    __syncthreads();
    if (tape) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        tp.piv_a()[i] = perm[i];
        tp.piv_b()[i] = perm[64 + i];
      }
      __syncthreads();
    }
  } else {
    pivoted_cholesky(Gm, n, Ra, perm, rot);
    if (tape) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) tp.piv_a()[i] = perm[i];
      __syncthreads();
    }
    if (mfma_b)
      su_environment_mfma<R>(g, tb, nb, ib, wlb, E);
    else
      su_environment<R>(g, tb, nb, ib, wlb, E, chunk, kSuChunk);
    trb = su_gram(E, h2, chi, p, r, Gm, red);
    pivoted_cholesky(Gm, n, Rb, perm, rot);
    if (tape) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) tp.piv_b()[i] = perm[i];
      __syncthreads();
    }
  }
  // core = R_a diag(lambda_rep) R_b^T, column-major.  R_b^T goes to the
  // free environment buffer first so that lanes (along j) read contiguous
  // words of it while R_a[i][c] is a broadcast.
  __shared__ double le[128];
  for (int j = threadIdx.x; j < chi; j += blockDim.x) le[j] = lam[(long long)e * chi + j];
  cd* RbT = E;
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) RbT[(f % n) * n + f / n] = Rb[f];
  __syncthreads();
  if (n == 32 && blockDim.x == 256) {
    // on FP64 matrix cores: wave w computes the 16 x 16 tile (w / 2, w % 2)
    // of C = (R_a diag(lambda_rep)) R_b^T, 8 k-steps of 4 real MFMAs each,
    // all operand loads of the tile issued before the products
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int i0 = (w >> 1) * 16, j0 = (w & 1) * 16, rr = l & 15, kq = l >> 4;
    cd a[8], b[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const int c = 4 * ks + kq;
      a[ks] = le[c / r] * Ra[(i0 + rr) * n + c];
      b[ks] = RbT[c * n + j0 + rr];
    }
    su_d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = cr;
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      cr = su_mfma(a[ks].re, b[ks].re, cr);
      cr = su_mfma(-a[ks].im, b[ks].im, cr);
      ci = su_mfma(a[ks].re, b[ks].im, ci);
      ci = su_mfma(a[ks].im, b[ks].re, ci);
    }
    // column-major for the Jacobi SVD; row-major for the pivoted QR that
    // preconditions it (its lanes walk the columns)
    const bool row_major = !tape && !kGlob;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + kq + 4 * v, jc = j0 + rr;
      C[row_major ? i * n + jc : jc * n + i] = mk<double>(cr[v], ci[v]);
    }
  } else {
    for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
      const int i = f / n, j = f % n;
      cd acc = mk<double>(0, 0);
      for (int c = 0; c < n; ++c) fma_acc(acc, le[c / r] * Ra[i * n + c], RbT[c * n + j]);
      C[j * n + i] = acc;
    }
  }
  __syncthreads();
  if (n == 32 && blockDim.x == 256 && !tape && !kGlob) {
    su_core_svd_qr32((lds_cd*)C, (lds_cd*)U, (lds_cd*)Vs, chunk, sigma, perm, rot, sizeof(R) == 4);
  } else if (n == 32 && blockDim.x == 256) {
    // the adjoint's record keeps the unpreconditioned factors it was built on
    if (!kGlob)
      block_jacobi_svd_n<32, lds_cd, true, true, false, true>((lds_cd*)C, U, (lds_cd*)Vs, sigma, perm, rot);
    else
      block_jacobi_svd_n<32>(C, U, Vs, sigma, perm, rot);
  } else if (n == 16) {
    if (threadIdx.x < 64) wave_jacobi_svd_n<16>(C, U, Vs, sigma, perm, rot);
    __syncthreads();
  } else if (n < 32) {
    if (threadIdx.x < 64) wave_jacobi_svd(C, n, n, U, Vs, sigma, perm, rot);
    __syncthreads();
  } else if (kMem == kSuWide && n == 64 && blockDim.x == 256) {
    if constexpr (kMem == kSuWide) su_core_svd64(C, U, Vs, sigma, perm, rot);
  } else {
    jacobi_svd(C, n, n, n, U, n, Vs, sigma, rot, flag, perm);
  }
  int keep = n < max_rank ? n : max_rank;
  if (keep > chi) keep = chi;
  if (n <= 64) {
    // wave 0, one singular value per lane: threshold, norm, entropy terms
    // and discarded weight summed across the wave (the logarithms in parallel)
    if (threadIdx.x < 64) {
      const int j = threadIdx.x;
      double sj = 0.0;
      if (j < n) {
        sj = sigma[j] > eps ? sigma[j] : 0.0;
        sigma[j] = sj;
      }
      const double nrm = sqrt(wave_sum(sj * sj));
      const double x = nrm > 0.0 ? (sj / nrm) * (sj / nrm) : 0.0;
      const double ent = -wave_sum(x > 0.0 ? x * log(x) : 0.0);
      const double err = wave_sum(j < n && j >= keep ? x : 0.0);
      if (j == 0) {
        err_out[blockIdx.x] = sqrt(err);
        ent_out[blockIdx.x] = ent;
        sh_rank = keep;
      }
    }
  } else if (threadIdx.x == 0) {
    double nrm = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!(sigma[j] > eps)) sigma[j] = 0.0;
      nrm += sigma[j] * sigma[j];
    }
    nrm = sqrt(nrm);
    double ent = 0.0, err = 0.0;
    for (int j = 0; j < n; ++j) {
      const double x = nrm > 0.0 ? (sigma[j] / nrm) * (sigma[j] / nrm) : 0.0;
      if (x > 0.0) ent -= x * log(x);
      if (j >= keep) err += x;
    }
    err_out[blockIdx.x] = sqrt(err);
    ent_out[blockIdx.x] = ent;
    sh_rank = keep;
  }
  __syncthreads();
  const int rank = sh_rank;
  if (tape) {
    // factors the adjoint needs: R_a, R_b (row-major), sorted U and V
    // (row-major), thresholded singular values, traces, kept rank
    for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
      const int i = f / n, j = f % n;
      tp.Ra()[f] = Ra[f];
      tp.Rb()[f] = Rb[f];
      tp.U()[f] = U[(long long)j * n + i];
      tp.V()[f] = Vs[(long long)perm[j] * n + i];
    }
    for (int j = threadIdx.x; j < n; j += blockDim.x) tp.sigma()[j] = sigma[j];
    if (threadIdx.x == 0) {
      tp.tr()[0] = tra;
      tp.tr()[1] = trb;
      tp.meta()[0] = rank;
      tp.meta()[1] = n;
    }
  }
  // W_a = Lambda R_b^T V Sigma^-1 (into G, the core is no longer needed),
  // then W_b = Lambda R_a^T conj(U) Sigma^-1 (into V, read for W_a only);
  // both n x chi.  Z_a, Z_b then go to E (U is no longer needed).
  for (int side = 0; side < 2; ++side) {
    cd* W = side == 0 ? Gm : V;
    for (int f = threadIdx.x; f < n * chi; f += blockDim.x) {
      const int x = f % n, j = f / n;  // lanes along x: R rows are read contiguously
      const int fw = x * chi + j;
      cd acc = mk<double>(0, 0);
      if (j < rank && sigma[j] > 0.0) {
        for (int i = 0; i < n; ++i) {
          if (side == 0)
            fma_acc(acc, Rb[i * n + x], Vs[(long long)perm[j] * n + i]);
          else
            fma_acc(acc, Ra[i * n + x], conj(U[(long long)j * n + i]));
        }
        acc = (le[x / r] / sigma[j]) * acc;
      }
      W[fw] = acc;
    }
    __syncthreads();
  }
  cd* Z = E;
  for (int side = 0; side < 2; ++side) {
    const cd* W = side == 0 ? Gm : V;
    const cd* h = side == 0 ? h1 : h2;
    for (int f = threadIdx.x; f < cp * cp; f += blockDim.x) {
      const int x = f / cp, y = f % cp;
      const int b = x / p, si = x % p, jn = y / p, so = y % p;
      cd acc = mk<double>(0, 0);
      if (jn < chi)
        for (int q = 0; q < r; ++q) fma_acc(acc, h[(q * p + so) * p + si], W[(b * r + q) * chi + jn]);
      Z[f] = acc;
    }
    __syncthreads();
    if (side == 0) {
      if (mfma_a)
        su_write_gamma_mfma<R>(g, ta, na, ia, wla, eps, Z, 1.0 / sqrt(tra));
      else
        su_write_gamma<R>(g, ta, na, ia, wla, eps, Z, 1.0 / sqrt(tra), chunk, kSuChunk);
    } else {
      if (mfma_b)
        su_write_gamma_mfma<R>(g, tb, nb, ib, wlb, eps, Z, 1.0 / sqrt(trb));
      else
        su_write_gamma<R>(g, tb, nb, ib, wlb, eps, Z, 1.0 / sqrt(trb), chunk, kSuChunk);
    }
  }
  for (int j = threadIdx.x; j < chi; j += blockDim.x) lam[(long long)e * chi + j] = j < rank ? sigma[j] : 0.0;
}

// ------------------------------------------------------------- launchers
hipError_t launch_apply1(int rb, const DevGraph& g, void* T, const int* nodes, int count, const cd* gates,
                         hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_apply1<float>, dim3(count), dim3(256), 0, s, g, (cx<float>*)T, nodes, gates);
  else
    hipLaunchKernelGGL(k_apply1<double>, dim3(count), dim3(256), 0, s, g, (cx<double>*)T, nodes, gates);
  return hipGetLastError();
}

hipError_t launch_su(int rb, const DevGraph& g, void* T, double* lam, const SuGate* gates, int count,
                     const cd* halves, int max_gate_rank, double eps, int max_rank, double* err, double* entropy,
                     void* ws, long long wsb, void* tape, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int mode = su_launch_mode(rb, g.chi, g.phys, max_gate_rank);
  const long long tb = su_tape_bytes(g.chi, max_gate_rank);
  char* tp = (char*)tape;
#define DQ_SU_LAUNCH(RT, MODE)                                                                                   \
  hipLaunchKernelGGL((k_su<RT, MODE>), dim3(count), dim3(256), 0, s, g, (cx<RT>*)T, lam, gates, halves,         \
                     max_gate_rank, eps, max_rank, err, entropy, (char*)ws, wsb, tp, tb)
  if (rb == 4) {
    if (mode == kSuCompact)
      DQ_SU_LAUNCH(float, kSuCompact);
    else if (mode == kSuLds)
      DQ_SU_LAUNCH(float, kSuLds);
    else if (mode == kSuWide)
      DQ_SU_LAUNCH(float, kSuWide);
    else
      DQ_SU_LAUNCH(float, kSuGlobal);
  } else {
    if (mode == kSuCompact)
      DQ_SU_LAUNCH(double, kSuCompact);
    else if (mode == kSuLds)
      DQ_SU_LAUNCH(double, kSuLds);
    else if (mode == kSuWide)
      DQ_SU_LAUNCH(double, kSuWide);
    else
      DQ_SU_LAUNCH(double, kSuGlobal);
  }
#undef DQ_SU_LAUNCH
  return hipGetLastError();
}

long long su_tune_key(int rb, int chi, int phys, int max_r) {
  return ((long long)rb << 24) | ((long long)chi << 12) | ((long long)phys << 6) | max_r;
}

// The storage mode of a launch: the autotuner's choice when it made one for
// this shape (compact and all-LDS are both valid for the compact shapes),
// else the default.
int su_launch_mode(int rb, int chi, int phys, int max_r) {
  const int def = su_mem_mode(chi, phys, max_r);
  const int v = tune_get(kTuneSu, su_tune_key(rb, chi, phys, max_r));
  if (def == kSuCompact && (v == kSuCompact || v == kSuLds)) return v;
  return def;
}

long long su_ws_bytes(int chi, int phys, int max_r) {
  // sized for the larger of the modes a launch of this shape may take
  const int def = su_mem_mode(chi, phys, max_r);
  long long cdn = su_ws_cd(def, chi, phys, max_r);
  if (def == kSuCompact) {
    const long long alt = su_ws_cd(kSuLds, chi, phys, max_r);
    if (alt > cdn) cdn = alt;
  }
  return (cdn * (long long)sizeof(cd) + 255) / 256 * 256;
}

}  // namespace dq

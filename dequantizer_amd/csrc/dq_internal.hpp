// Internal declarations shared by the HIP translation units of libdq.
//
// Data layout in HBM (all complex values interleaved re/im, like torch
// complex64 / complex128):
//   node tensors  T[n] at toff[n], row-major (bond_0 .. bond_{d-1}, phys),
//                 every bond padded to the graph's chi (zeros beyond the true
//                 bond dimension), bonds in the node's adjacency order;
//   messages      n2e[(e*2+s)*chi*chi]  node (side s of edge e) -> edge e,
//                 e2n[(e*2+s)*chi*chi]  edge e -> node (side s), M[bra][ket];
//   lambdas       lam[e*chi + k]  (double, zeros beyond the true dimension).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dq {

constexpr int kMaxDeg = 8;

template <typename R>
struct cx {
  R re, im;
};

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) {
  cx<R> z;
  z.re = a;
  z.im = b;
  return z;
}
template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return mk<R>(a.re + b.re, a.im + b.im); }
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return mk<R>(a.re - b.re, a.im - b.im); }
template <typename R>
__device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return mk<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return mk<R>(s * a.re, s * a.im); }
template <typename R>
__device__ __forceinline__ cx<R> conj(cx<R> a) { return mk<R>(a.re, -a.im); }
template <typename R>
__device__ __forceinline__ R norm2(cx<R> a) { return a.re * a.re + a.im * a.im; }
// a += b * c
template <typename R>
__device__ __forceinline__ void fma_acc(cx<R>& a, cx<R> b, cx<R> c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(-b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(b.im, c.re, a.im);
}
// a += conj(b) * c
template <typename R>
__device__ __forceinline__ void fma_cacc(cx<R>& a, cx<R> b, cx<R> c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(-b.im, c.re, a.im);
}
template <typename R, typename S>
__device__ __forceinline__ cx<R> cvt(cx<S> a) { return mk<R>((R)a.re, (R)a.im); }
template <typename R>
__device__ __forceinline__ cx<R> cdiv(cx<R> a, cx<R> b) {
  R d = b.re * b.re + b.im * b.im;
  return mk<R>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

typedef cx<double> cd;
// Complex doubles and ints in LDS, for device functions the compiler does not
// inline: through a generic pointer their LDS accesses become flat memory
// instructions (slower, and waited on with the vector-memory counter).
typedef __attribute__((address_space(3))) cd lds_cd;
typedef __attribute__((address_space(3))) int lds_int;

// Element load / store of a complex double through a pointer in any address
// space (an LDS-qualified struct has no usable copy constructor or assignment).
template <typename P>
__device__ __forceinline__ cd cld(const P* p, long long i) {
  return mk<double>(p[i].re, p[i].im);
}
template <typename P>
__device__ __forceinline__ void cst(P* p, long long i, cd v) {
  p[i].re = v.re;
  p[i].im = v.im;
}

// Graph descriptor passed by value to kernels (all pointers are device memory).
struct DevGraph {
  int n_nodes;
  int n_edges;
  int chi;   // padded bond dimension
  int phys;  // physical dimension (uniform)
  const int* deg;         // [n_nodes]
  const int* adj_ptr;     // [n_nodes+1]
  const int* adj_edge;    // [adj_ptr[n]] edge of each bond slot
  const int* adj_side;    // [adj_ptr[n]] side (0/1) of the node in that edge
  const long long* toff;  // [n_nodes] node tensor offsets in complex elements
  const int* edge_node;   // [2*n_edges]
  const int* node_inst;   // [n_nodes]
  const int* edge_inst;   // [n_edges]
};

__host__ __device__ inline long long ipow(long long b, int e) {
  long long r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

// Read of graph metadata that no kernel writes while it runs (adjacency,
// offsets, node lists, schedule codes) through the constant address space:
// with a wave-uniform index it becomes a scalar load, whose completion is
// tracked apart from the vector loads in flight (a chain of such reads does
// not wait for a prefetched tensor).
template <typename T>
__device__ __forceinline__ T ldc(const T* p, long long i) {
  return ((const T __attribute__((address_space(4)))*)p)[i];
}

// ---------------------------------------------------------------- reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One DPP lane move of a double (two 32-bit moves).  Every source lane of
// the controls used below lies in the destination's own 16-lane row, and the
// callers run with the whole wave active.  bound_ctrl set: a source lane that
// is disabled would read as 0, so the move needs no "old" operand and the
// compiler writes a fresh register instead of first copying one.
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), kCtrl, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), kCtrl, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int kCtrl>
__device__ __forceinline__ int dpp_move_int(int v) {
  return __builtin_amdgcn_mov_dpp(v, kCtrl, 0xF, 0xF, true);
}

// Sum over aligned groups of L lanes (L = 1, 2, 4, 8, 16) by DPP moves inside
// a row: quad_perm(1,0,3,2) pairs lanes i, i^1; quad_perm(2,3,0,1) pairs
// i, i^2; row_half_mirror pairs i, 7-i within 8 lanes; row_mirror pairs
// i, 15-i within 16.  After each step all lanes of the group formed so far
// hold the same value (a + b and b + a round identically), so every lane of
// a group gets the same sum.
template <int L>
__device__ __forceinline__ double group_sum(double v) {
  static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16, "group of 1..16 lanes");
  if (L >= 2) v += dpp_move<0xB1>(v);
  if (L >= 4) v += dpp_move<0x4E>(v);
  if (L >= 8) v += dpp_move<0x141>(v);
  if (L >= 16) v += dpp_move<0x140>(v);
  return v;
}

// The same sum of four values at once, the DPP moves of each step issued
// together (four independent chains instead of one chain four times as long).
template <int kCtrl>
__device__ __forceinline__ void group_step4(double& a, double& b, double& c, double& d) {
  const double ma = dpp_move<kCtrl>(a), mb = dpp_move<kCtrl>(b);
  const double mc = dpp_move<kCtrl>(c), md = dpp_move<kCtrl>(d);
  a += ma;
  b += mb;
  c += mc;
  d += md;
}
template <int L>
__device__ __forceinline__ void group_sum4(double& a, double& b, double& c, double& d) {
  static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16, "group of 1..16 lanes");
  if (L >= 2) group_step4<0xB1>(a, b, c, d);
  if (L >= 4) group_step4<0x4E>(a, b, c, d);
  if (L >= 8) group_step4<0x141>(a, b, c, d);
  if (L >= 16) group_step4<0x140>(a, b, c, d);
}

// Block-wide sum of a double; `red` must hold blockDim.x/64 doubles.  All
// threads get the result.
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  __syncthreads();
  return s;
}

// Adjacency slot of edge e (node on side `side`) in `node`'s bond order.
__device__ inline int find_slot(const DevGraph& g, int node, int e, int side) {
  const int b = ldc(g.adj_ptr, node), d = ldc(g.deg, node);
  for (int k = 0; k < d; ++k)
    if (ldc(g.adj_edge, b + k) == e && ldc(g.adj_side, b + k) == side) return k;
  return -1;
}

// ----------------------------------------------------- generic tensor helpers
// dst = src with matrix M applied on axis j of a tensor with `d` bonds of
// size chi and a trailing phys axis: dst[o, a, in] = sum_b M[a*chi+b] src[o, b, in]
// (transposed=true uses M[b*chi+a]).  Whole block cooperates.
// For chi <= 32 and tensors of 4096 elements or more, each thread takes whole
// columns (all indices but axis j): the chi values of a column are read once
// into registers and the chi outputs written from them, so the tensor is read
// and written once per product (the element loop below reads every input chi
// times); M is staged in LDS (as R) first.  Small tensors keep the element
// loop, which needs no block barrier.  Both forms sum over b in the same
// order: the results are bitwise equal.
template <typename R, typename MT>
__device__ void mode_product(cx<R>* dst, const cx<R>* src, const cx<MT>* M, int chi, int d, int phys, int j,
                             bool transposed) {
  const long long inner = ipow(chi, d - 1 - j) * phys;
  const long long total = ipow(chi, d) * phys;
  constexpr int kReg = 32;
  if (chi <= kReg && total >= 4096) {
    __shared__ cx<R> Ms[kReg * kReg];
    __syncthreads();  // earlier readers of Ms are done
    for (int q = threadIdx.x; q < chi * chi; q += blockDim.x) {
      const int a = q / chi, b = q - a * chi;
      Ms[q] = cvt<R>(transposed ? M[(long long)b * chi + a] : M[q]);
    }
    __syncthreads();
    const long long cols = total / chi;
    for (long long c = threadIdx.x; c < cols; c += blockDim.x) {
      const long long o = c / inner, i = c - o * inner;
      const long long b0 = o * chi * inner + i;
      cx<R> in[kReg];
#pragma unroll
      for (int b = 0; b < kReg; ++b)
        if (b < chi) in[b] = src[b0 + b * inner];
      for (int a = 0; a < chi; ++a) {
        cx<R> acc = mk<R>(0, 0);
#pragma unroll
        for (int b = 0; b < kReg; ++b)
          if (b < chi) fma_acc(acc, Ms[a * chi + b], in[b]);
        dst[b0 + a * inner] = acc;
      }
    }
    return;
  }
  for (long long f = threadIdx.x; f < total; f += blockDim.x) {
    const long long in = f % inner;
    const long long t = f / inner;
    const int a = (int)(t % chi);
    const long long o = t / chi;
    const cx<R>* s = src + o * chi * inner + in;
    cx<R> acc = mk<R>(0, 0);
    for (int b = 0; b < chi; ++b) {
      cx<MT> m = transposed ? M[(long long)b * chi + a] : M[(long long)a * chi + b];
      fma_acc(acc, cvt<R>(m), s[b * inner]);
    }
    dst[f] = acc;
  }
}

// out[k*chi+l] = sum over all axes but bond idx of conj(A[..k..]) * B[..l..]
// * prod_{j != idx} w_j[b_j] (w may be null).  Accumulates in double.
// Results written to `out` (double complex).  Whole block cooperates.
template <typename R>
__device__ void bond_gram(cd* out, const cx<R>* A, const cx<R>* B, int chi, int d, int phys, int idx,
                          const double* const* w) {
  const long long inner = ipow(chi, d - 1 - idx) * phys;
  const long long rest = ipow(chi, d - 1) * phys;
  const int pairs = chi * chi;
  for (int q = threadIdx.x; q < pairs; q += blockDim.x) {
    const int k = q / chi, l = q % chi;
    double sr = 0.0, si = 0.0;
    for (long long r = 0; r < rest; ++r) {
      const long long o = r / inner, in = r % inner;
      double wt = 1.0;
      if (w) {
        // decode bonds of the rest index (outer bonds 0..idx-1, inner bonds idx+1..d-1)
        long long oo = o;
        for (int j = idx - 1; j >= 0; --j) {
          wt *= w[j][oo % chi];
          oo /= chi;
        }
        long long ii = in / phys;
        for (int j = d - 1; j > idx; --j) {
          wt *= w[j][ii % chi];
          ii /= chi;
        }
      }
      const cx<R> a = A[(o * chi + k) * inner + in];
      const cx<R> b = B[(o * chi + l) * inner + in];
      sr += wt * ((double)a.re * b.re + (double)a.im * b.im);
      si += wt * ((double)a.re * b.im - (double)a.im * b.re);
    }
    out[q] = mk<double>(sr, si);
  }
}

}  // namespace dq

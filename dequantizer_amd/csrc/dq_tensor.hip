// Gate products on single, unpadded node tensors (the dictionary API of
// TensorGraph, where every node tensor has its own shape).
//
// A tensor is viewed as (lead, bond, trail, phys); a gate half h (r, out, in)
// gives out[l, b, q, t, o] = sum_i h[q, o, i] t[l, b, t, i], i.e. the rank
// index q becomes the fast part of the fused bond (b * r + q), as in
// _apply_half_gate (dequantizer/src/tensor_ops.py:221-237).  A one-site gate
// G (out, in) is the case bond = trail = r = 1 (TensorGraph.apply1,
// dequantizer/src/graph/tensor_graph.py:268-275).  Both are followed by a
// Frobenius normalisation of the whole result.
//
// Two launches: k_tensor_gate writes the product and one partial sum of
// |out|^2 per block; k_tensor_scale has every block fold the partials (at
// most kTensorParts doubles) and scale its share.  Memory-bound: each output
// element reads phys inputs (L2-resident re-reads across q and o) and is
// written twice.
#include "dq_launch.hpp"

namespace dq {

template <typename R>
__global__ __launch_bounds__(256) void k_tensor_gate(const cx<R>* __restrict__ t, long long lead, int bond,
                                                     long long trail, int p, const cd* __restrict__ half, int r,
                                                     cx<R>* __restrict__ out, double* __restrict__ partial) {
  __shared__ double red[4];
  const long long n = lead * bond * r * trail * p;
  double sq = 0.0;
  for (long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x; f < n; f += (long long)gridDim.x * blockDim.x) {
    long long x = f;
    const int o = (int)(x % p);
    x /= p;
    const long long tr = x % trail;
    x /= trail;
    const int q = (int)(x % r);
    x /= r;
    const int b = (int)(x % bond);
    const long long l = x / bond;
    const cx<R>* src = t + ((l * bond + b) * trail + tr) * p;
    const cd* hq = half + ((long long)q * p + o) * p;
    cd acc = mk<double>(0.0, 0.0);
    for (int i = 0; i < p; ++i) fma_acc(acc, hq[i], cvt<double>(src[i]));
    const cx<R> v = cvt<R>(acc);
    out[f] = v;
    sq += (double)norm2(v);
  }
  sq = block_sum(sq, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = sq;
}

template <typename R>
__global__ __launch_bounds__(256) void k_tensor_scale(cx<R>* __restrict__ out, long long n,
                                                      const double* __restrict__ partial, int parts) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < parts; i += blockDim.x) s += partial[i];
  s = block_sum(s, red);
  const R inv = (R)(1.0 / sqrt(s));
  for (long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x; f < n; f += (long long)gridDim.x * blockDim.x)
    out[f] = inv * out[f];
}

hipError_t launch_tensor_gate(int real_bytes, const void* t, long long lead, int bond, long long trail, int p,
                              const cd* half, int r, void* out, double* partial, hipStream_t s) {
  const long long n = lead * bond * r * trail * p;
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > kTensorParts) blocks = kTensorParts;
  if (real_bytes == 4) {
    hipLaunchKernelGGL(k_tensor_gate<float>, dim3((unsigned)blocks), dim3(256), 0, s, (const cx<float>*)t, lead,
                       bond, trail, p, half, r, (cx<float>*)out, partial);
    hipLaunchKernelGGL(k_tensor_scale<float>, dim3((unsigned)blocks), dim3(256), 0, s, (cx<float>*)out, n, partial,
                       (int)blocks);
  } else {
    hipLaunchKernelGGL(k_tensor_gate<double>, dim3((unsigned)blocks), dim3(256), 0, s, (const cx<double>*)t, lead,
                       bond, trail, p, half, r, (cx<double>*)out, partial);
    hipLaunchKernelGGL(k_tensor_scale<double>, dim3((unsigned)blocks), dim3(256), 0, s, (cx<double>*)out, n, partial,
                       (int)blocks);
  }
  return hipGetLastError();
}

}  // namespace dq

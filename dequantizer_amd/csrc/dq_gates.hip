// Operator-Schmidt split of two-site gates, batched (one thread per gate).
//
// A gate G (out1, out2, in1, in2) is read as the N x N matrix
// M[(out1 in1), (out2 in2)] with N = phys^2.  One-sided (Hestenes) Jacobi in
// FP64 orthogonalises the columns of A = M V; then s_j = |A_j|, U_j = A_j/s_j
// and M = U S V^H.  Singular values above eps are kept in descending order;
// half 1 = (r, out1, in1) with entries U[(o1 i1), q] sqrt(s_q), half 2 =
// (r, out2, in2) with entries conj(V[(o2 i2), q]) sqrt(s_q), which is the
// split of _decompose_gate (dequantizer/src/tensor_ops.py:202-218).  The
// column phases of an SVD are free, and the simple update built on the
// halves depends only on their product (the gate), so any valid split gives
// the same Gammas and lambdas.
//
// The Jacobi rotation for columns a_j, a_k with alpha = |a_j|^2,
// beta = |a_k|^2, gamma = a_j^H a_k = |gamma| e^{i phi}: rotate a_k by
// e^{-i phi} so the pair's Gram is real symmetric, then the classical
// two-sided Jacobi angle of [[alpha, |gamma|], [|gamma|, beta]]:
// zeta = (beta - alpha) / (2 |gamma|), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
// c = 1 / sqrt(1 + t^2), s = t c; (a_j, a_k') -> (c a_j - s a_k', s a_j + c a_k').
#include "dq_launch.hpp"

namespace dq {

template <int N>
__global__ __launch_bounds__(64) void k_decompose_gates(const cd* __restrict__ gates, int count, int p, double eps,
                                                        cd* __restrict__ halves, int* __restrict__ ranks,
                                                        int rank_stride) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= count) return;
  const int p4 = N * N;
  const cd* G = gates + (long long)gi * p4;
  cd A[N * N], V[N * N];  // column-major: A[col * N + row]
  for (int o1 = 0; o1 < p; ++o1)
    for (int o2 = 0; o2 < p; ++o2)
      for (int i1 = 0; i1 < p; ++i1)
        for (int i2 = 0; i2 < p; ++i2) A[(o2 * p + i2) * N + o1 * p + i1] = G[((o1 * p + o2) * p + i1) * p + i2];
  for (int j = 0; j < N * N; ++j) V[j] = mk<double>(0.0, 0.0);
  for (int j = 0; j < N; ++j) V[j * N + j] = mk<double>(1.0, 0.0);
  for (int sweep = 0; sweep < 40; ++sweep) {
    bool rotated = false;
    for (int j = 0; j < N - 1; ++j)
      for (int k = j + 1; k < N; ++k) {
        double alpha = 0.0, beta = 0.0;
        cd gam = mk<double>(0.0, 0.0);
        for (int i = 0; i < N; ++i) {
          alpha += norm2(A[j * N + i]);
          beta += norm2(A[k * N + i]);
          fma_acc(gam, conj(A[j * N + i]), A[k * N + i]);
        }
        const double ag = sqrt(norm2(gam));
        if (!(ag > 1e-15 * sqrt(alpha * beta)) || !(ag > 1e-300)) continue;
        rotated = true;
        const cd ph = mk<double>(gam.re / ag, -gam.im / ag);  // e^{-i phi}
        const double zeta = (beta - alpha) / (2.0 * ag);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int i = 0; i < N; ++i) {
          const cd aj = A[j * N + i], ak = ph * A[k * N + i];
          A[j * N + i] = c * aj - s * ak;
          A[k * N + i] = s * aj + c * ak;
          const cd vj = V[j * N + i], vk = ph * V[k * N + i];
          V[j * N + i] = c * vj - s * vk;
          V[k * N + i] = s * vj + c * vk;
        }
      }
    if (!rotated) break;
  }
  double sv[N];
  int order[N];
  for (int j = 0; j < N; ++j) {
    double a = 0.0;
    for (int i = 0; i < N; ++i) a += norm2(A[j * N + i]);
    sv[j] = sqrt(a);
    order[j] = j;
  }
  for (int a = 0; a < N; ++a)  // selection sort, descending (N <= 16)
    for (int b = a + 1; b < N; ++b)
      if (sv[order[b]] > sv[order[a]]) {
        const int tmp = order[a];
        order[a] = order[b];
        order[b] = tmp;
      }
  int r = 0;
  while (r < N && sv[order[r]] > eps) ++r;
  cd* h1 = halves + (long long)gi * 2 * p4;
  cd* h2 = h1 + p4;
  for (int q = 0; q < N; ++q) {
    const int j = order[q];
    const bool keep = q < r;
    const double root = keep ? sqrt(sv[j]) : 0.0;
    const double inv = keep ? 1.0 / sv[j] : 0.0;
    for (int x = 0; x < N; ++x) {
      h1[q * N + x] = (root * inv) * A[j * N + x];  // U[(o1 i1), q] sqrt(s_q)
      h2[q * N + x] = root * conj(V[j * N + x]);    // conj(V[(o2 i2), q]) sqrt(s_q)
    }
  }
  ranks[(long long)gi * rank_stride] = r;
}

hipError_t launch_decompose_gates(const cd* gates, int count, int p, double eps, cd* halves, int* ranks,
                                  int rank_stride, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const dim3 grid((count + 63) / 64), block(64);
  switch (p) {
    case 1:
      hipLaunchKernelGGL(k_decompose_gates<1>, grid, block, 0, s, gates, count, p, eps, halves, ranks, rank_stride);
      break;
    case 2:
      hipLaunchKernelGGL(k_decompose_gates<4>, grid, block, 0, s, gates, count, p, eps, halves, ranks, rank_stride);
      break;
    case 3:
      hipLaunchKernelGGL(k_decompose_gates<9>, grid, block, 0, s, gates, count, p, eps, halves, ranks, rank_stride);
      break;
    case 4:
      hipLaunchKernelGGL(k_decompose_gates<16>, grid, block, 0, s, gates, count, p, eps, halves, ranks, rank_stride);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace dq

// Vidal gauge fixing, Vidal distance and the symmetric gauge.
//
// edge_svd      reference dequantizer/src/tensor_ops.py:108-130
// absorb        tensor_ops.py:88-93
// vidal_dist    tensor_ops.py:143-175, mappings/vidal_distance.py:17-42
// canonicalize  tensor_ops.py:189-199
// All small eigen/singular decompositions run in FP64 (dq_linalg.hpp).
#include <float.h>

#include "dq_launch.hpp"
#include "dq_linalg.hpp"

namespace dq {

__device__ __forceinline__ double safe_inverse(double l, double eps) {
  return 1.0 / ((l < eps ? (double)FLT_MAX : 0.0) + l);
}
__device__ __forceinline__ double safe_truncate(double l, double eps) { return l > eps ? l : 0.0; }

// Workspace of one edge block: 11 chi x chi complex matrices + scalars.
__host__ __device__ inline long long edge_svd_ws_bytes(int chi) {
  const long long cc = (long long)chi * chi;
  return 11 * cc * (long long)sizeof(cd) + (6LL * chi + 16) * sizeof(double) + (chi + 4) * sizeof(int);
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_edge_svd(DevGraph g, const cx<R>* n2e, double eps, cd* X, double* lam, char* ws, long long wsb,
               const int* inst_active, char* tape) {
  const int e = blockIdx.x;
  if (inst_active && !inst_active[g.edge_inst[e]]) return;
  const int n = g.chi;
  const int cc = n * n;
  cd* Mf = (cd*)(ws + (long long)blockIdx.x * wsb);
  cd *Mb = Mf + cc, *Vf = Mb + cc, *Vb = Vf + cc, *Sf = Vb + cc, *Sb = Sf + cc, *If = Sb + cc, *Ib = If + cc;
  cd *G = Ib + cc, *U = G + cc, *V = U + cc;
  double* rot = (double*)(V + cc);
  double* sigma = rot + 4 * n + 4;
  double* red = sigma + n;
  int* flag = (int*)(red + n + 8);
  int* perm = flag + 2;

  const cx<R>* m0 = n2e + (2LL * e) * cc;
  const cx<R>* m1 = n2e + (2LL * e + 1) * cc;
  // Hermitian part (x + x^H) / 2, as the reference's eigh symmetrises its input
  for (int q = threadIdx.x; q < cc; q += blockDim.x) {
    const int k = q / n, l = q % n;
    const cd a = cvt<double>(m0[q]), at = cvt<double>(m0[l * n + k]);
    const cd b = cvt<double>(m1[q]), bt = cvt<double>(m1[l * n + k]);
    Mf[q] = mk<double>(0.5 * (a.re + at.re), 0.5 * (a.im - at.im));
    Mb[q] = mk<double>(0.5 * (b.re + bt.re), 0.5 * (b.im - bt.im));
  }
  __syncthreads();
  jacobi_eigh(Mf, Vf, n, rot, flag);
  jacobi_eigh(Mb, Vb, n, rot, flag);
  // m^{1/2} and m^{-1/2} (safe) from the eigendecompositions
  for (int q = threadIdx.x; q < cc; q += blockDim.x) {
    const int r = q / n, c = q % n;
    cd sf = mk<double>(0, 0), sb = sf, inf = sf, inb = sf;
    for (int i = 0; i < n; ++i) {
      const double wf = Mf[i * n + i].re, wb = Mb[i * n + i].re;
      const cd pf = Vf[r * n + i] * conj(Vf[c * n + i]);
      const cd pb = Vb[r * n + i] * conj(Vb[c * n + i]);
      const double a = sqrt(safe_truncate(wf, eps)), b = sqrt(safe_truncate(wb, eps));
      const double ai = sqrt(safe_inverse(wf, eps)), bi = sqrt(safe_inverse(wb, eps));
      sf = sf + a * pf;
      sb = sb + b * pb;
      inf = inf + ai * pf;
      inb = inb + bi * pb;
    }
    Sf[q] = sf;
    Sb[q] = sb;
    If[q] = inf;
    Ib[q] = inb;
  }
  __syncthreads();
  // G (column-major) = Sf Sb^T
  for (int q = threadIdx.x; q < cc; q += blockDim.x) {
    const int c = q / n, r = q % n;
    cd acc = mk<double>(0, 0);
    for (int k = 0; k < n; ++k) fma_acc(acc, Sf[r * n + k], Sb[c * n + k]);
    G[q] = acc;
  }
  __syncthreads();
  jacobi_svd(G, n, n, n, U, n, V, sigma, rot, flag, perm);
  double nrm = 0.0;
  for (int j = 0; j < n; ++j) nrm += sigma[j] * sigma[j];
  nrm = sqrt(nrm);
  // masked Frobenius norms of U and Vh, the (j, i) terms spread over the
  // block and summed once
  double pu = 0.0, pv = 0.0;
  for (int q = threadIdx.x; q < cc; q += blockDim.x) {
    const int j = q / n, i = q % n;
    const double lj = nrm > 0.0 ? sigma[j] / nrm : 0.0;
    if (safe_truncate(lj, eps) > eps) {
      pu += norm2(U[j * n + i]);
      pv += norm2(V[perm[j] * n + i]);
    }
  }
  const double nu = sqrt(block_sum(pu, red));
  const double nv = sqrt(block_sum(pv, red));
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double lj = nrm > 0.0 ? safe_truncate(sigma[j] / nrm, eps) : 0.0;
    lam[(long long)e * n + j] = lj;
  }
  if (tape) {
    // eigenvectors / eigenvalues of both messages, sorted U and V (row-major),
    // singular values and the two norms, for the adjoint
    EdgeSvdTape tp(tape + (long long)e * edge_svd_tape_bytes(n), n);
    for (int q = threadIdx.x; q < cc; q += blockDim.x) {
      const int i = q / n, j = q % n;
      tp.Qf()[q] = Vf[q];
      tp.Qb()[q] = Vb[q];
      tp.U()[q] = U[(long long)j * n + i];
      tp.V()[q] = V[(long long)perm[j] * n + i];
    }
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      tp.wf()[j] = Mf[j * n + j].re;
      tp.wb()[j] = Mb[j * n + j].re;
      tp.sigma()[j] = sigma[j];
    }
    if (threadIdx.x == 0) {
      tp.scal()[0] = nrm;
      tp.scal()[1] = nu;
      tp.scal()[2] = nv;
    }
  }
  cd* Xf = X + (2LL * e) * cc;
  cd* Xb = X + (2LL * e + 1) * cc;
  for (int q = threadIdx.x; q < cc; q += blockDim.x) {
    const int b = q / n, k = q % n;
    const double lk = nrm > 0.0 ? safe_truncate(sigma[k] / nrm, eps) : 0.0;
    cd xf = mk<double>(0, 0), xb = xf;
    if (lk > eps) {
      for (int i = 0; i < n; ++i) {
        fma_acc(xf, If[b * n + i], U[k * n + i]);
        fma_acc(xb, Ib[b * n + i], conj(V[perm[k] * n + i]));
      }
      xf = (1.0 / nu) * xf;
      xb = (1.0 / nv) * xb;
    }
    Xf[q] = xf;
    Xb[q] = xb;
  }
}

// The same decomposition for chi = 16 with wave-level Jacobi: one edge per
// 128-thread block; wave 0 diagonalises the forward message while wave 1
// diagonalises the backward one, then wave 0 takes the SVD of the core.
// Seven 16 x 16 complex128 matrices in LDS (28 KiB): Mf (then Sf, then U),
// Mb (then Sb, then V), Qf, Qb, If, Ib, G.
template <typename R>
__global__ void __launch_bounds__(128)
    k_edge_svd16(DevGraph g, const cx<R>* n2e, double eps, cd* X, double* lam, const int* inst_active, char* tape) {
  constexpr int n = 16, cc = 256;
  // six 16 x 16 matrices (G reuses Qf once the square roots are formed), so
  // that six blocks fit a CU's LDS and 1536 edges run in one round
  __shared__ cd Mf[cc], Mb[cc], Qf[cc], Qb[cc], If[cc], Ib[cc];
  __shared__ double wf[n], wb[n], sigma[n], scal[2 * n];
  __shared__ int perm[n];
  const int e = blockIdx.x;
  if (inst_active && !inst_active[g.edge_inst[e]]) return;
  const int tid = threadIdx.x;
  const cx<R>* m0 = n2e + (2LL * e) * cc;
  const cx<R>* m1 = n2e + (2LL * e + 1) * cc;
  // Hermitian part (x + x^H) / 2, as the reference's eigh symmetrises its input
  for (int q = tid; q < cc; q += blockDim.x) {
    const int k = q / n, l = q % n;
    const cd a = cvt<double>(m0[q]), at = cvt<double>(m0[l * n + k]);
    const cd b = cvt<double>(m1[q]), bt = cvt<double>(m1[l * n + k]);
    Mf[q] = mk<double>(0.5 * (a.re + at.re), 0.5 * (a.im - at.im));
    Mb[q] = mk<double>(0.5 * (b.re + bt.re), 0.5 * (b.im - bt.im));
  }
  __syncthreads();
  // complex64 results without an adjoint record: the Jacobi iterations stop
  // once the remaining rotations are below single-precision rounding
  const bool loose = sizeof(R) == 4 && !tape;
  if (tid < 64)
    wave_jacobi_eigh_n<n>(Mf, Qf, loose);
  else
    wave_jacobi_eigh_n<n>(Mb, Qb, loose);
  __syncthreads();
  if (tid < n) {
    wf[tid] = Mf[tid * (n + 1)].re;
    wb[tid] = Mb[tid * (n + 1)].re;
  }
  __syncthreads();
  if (tape) {
    EdgeSvdTape tp(tape + (long long)e * edge_svd_tape_bytes(n), n);
    for (int q = tid; q < cc; q += blockDim.x) {
      tp.Qf()[q] = Qf[q];
      tp.Qb()[q] = Qb[q];
    }
    for (int j = tid; j < n; j += blockDim.x) {
      tp.wf()[j] = wf[j];
      tp.wb()[j] = wb[j];
    }
  }
  // m^{1/2} and m^{-1/2} (safe) from the eigendecompositions; the square
  // roots of the clipped and of the inverted eigenvalues once per edge
  __shared__ double rt[4][n];
  if (tid < 4 * n) {
    const int which = tid / n, i = tid % n;
    const double w = (which & 1) ? wb[i] : wf[i];
    rt[which][i] = which < 2 ? sqrt(safe_truncate(w, eps)) : sqrt(safe_inverse(w, eps));
  }
  __syncthreads();
  cd* Sf = Mf;
  cd* Sb = Mb;
  for (int q = tid; q < cc; q += blockDim.x) {
    const int r = q / n, c = q % n;
    cd sf = mk<double>(0, 0), sb = sf, inf = sf, inb = sf;
#pragma unroll 4
    for (int i = 0; i < n; ++i) {
      const cd pf = Qf[r * n + i] * conj(Qf[c * n + i]);
      const cd pb = Qb[r * n + i] * conj(Qb[c * n + i]);
      const double a = rt[0][i], b = rt[1][i], ai = rt[2][i], bi = rt[3][i];
      sf = sf + a * pf;
      sb = sb + b * pb;
      inf = inf + ai * pf;
      inb = inb + bi * pb;
    }
    Sf[q] = sf;
    Sb[q] = sb;
    If[q] = inf;
    Ib[q] = inb;
  }
  __syncthreads();
  // G (column-major) = Sf Sb^T
  cd* G = Qf;
  for (int q = tid; q < cc; q += blockDim.x) {
    const int c = q / n, r = q % n;
    cd acc = mk<double>(0, 0);
#pragma unroll 4
    for (int k = 0; k < n; ++k) fma_acc(acc, Sf[r * n + k], Sb[c * n + k]);
    G[q] = acc;
  }
  __syncthreads();
  cd* U = Mf;
  cd* V = Mb;
  if (tid < 64) wave_jacobi_svd_n<n>(G, U, V, sigma, perm, scal, loose);
  __syncthreads();
  double nrm = 0.0;
  for (int j = 0; j < n; ++j) nrm += sigma[j] * sigma[j];
  nrm = sqrt(nrm);
  // masked Frobenius norms of U and Vh, the (j, i) terms spread over the
  // block and summed once
  __shared__ double red[8];
  double pu = 0.0, pv = 0.0;
  for (int q = tid; q < cc; q += blockDim.x) {
    const int j = q / n, i = q % n;
    const double lj = nrm > 0.0 ? sigma[j] / nrm : 0.0;
    if (safe_truncate(lj, eps) > eps) {
      pu += norm2(U[j * n + i]);
      pv += norm2(V[perm[j] * n + i]);
    }
  }
  const double nu = sqrt(block_sum(pu, red));
  const double nv = sqrt(block_sum(pv, red));
  for (int j = tid; j < n; j += blockDim.x) {
    const double lj = nrm > 0.0 ? safe_truncate(sigma[j] / nrm, eps) : 0.0;
    lam[(long long)e * n + j] = lj;
  }
  if (tape) {
    EdgeSvdTape tp(tape + (long long)e * edge_svd_tape_bytes(n), n);
    for (int q = tid; q < cc; q += blockDim.x) {
      const int i = q / n, j = q % n;
      tp.U()[q] = U[(long long)j * n + i];
      tp.V()[q] = V[(long long)perm[j] * n + i];
    }
    for (int j = tid; j < n; j += blockDim.x) tp.sigma()[j] = sigma[j];
    if (tid == 0) {
      tp.scal()[0] = nrm;
      tp.scal()[1] = nu;
      tp.scal()[2] = nv;
    }
  }
  cd* Xf = X + (2LL * e) * cc;
  cd* Xb = X + (2LL * e + 1) * cc;
  for (int q = tid; q < cc; q += blockDim.x) {
    const int b = q / n, k = q % n;
    const double lk = nrm > 0.0 ? safe_truncate(sigma[k] / nrm, eps) : 0.0;
    cd xf = mk<double>(0, 0), xb = xf;
    if (lk > eps) {
      for (int i = 0; i < n; ++i) {
        fma_acc(xf, If[b * n + i], U[k * n + i]);
        fma_acc(xb, Ib[b * n + i], conj(V[perm[k] * n + i]));
      }
      xf = (1.0 / nu) * xf;
      xb = (1.0 / nv) * xb;
    }
    Xf[q] = xf;
    Xb[q] = xb;
  }
}

// The same decomposition for chi = 32, one edge per 256-thread block, four
// 32 x 32 complex128 matrices in LDS (64 KiB: two blocks per CU).  The two
// messages are diagonalised by the block-level one-sided Jacobi (below); the
// square roots are applied in factored form,
//   G = Sf Sb^T = Qf (Df^1/2 Qf^H conj(Qb) Db^1/2) Qb^T,
// as three small products; the block-level one-sided Jacobi takes the SVD of
// G (its left vectors U go to the edge's slot of the workspace); then
//   Xf = Qf (Df^-1/2 Qf^H U),  Xb = Qb (Db^-1/2 Qb^H conj(V)).
// Same results as k_edge_svd to rounding (no adjoint tape: the generic kernel
// records it).
template <typename R>
__global__ void __launch_bounds__(256)
    k_edge_svd32(DevGraph g, const cx<R>* n2e, double eps, cd* X, double* lam, char* ws, long long wsb,
                 const int* inst_active) {
  constexpr int n = 32, cc = 1024;
  __shared__ cd Mf[cc], Mb[cc], Qf[cc], Qb[cc];
  __shared__ double wf[n], wb[n], sigma[n], scal[2 * n], rt[4][n], red[8];
  __shared__ int perm[n];
  const int e = blockIdx.x;
  if (inst_active && !inst_active[g.edge_inst[e]]) return;
  const int tid = threadIdx.x;
  cd* U = (cd*)(ws + (long long)e * wsb);
  const cx<R>* m0 = n2e + (2LL * e) * cc;
  const cx<R>* m1 = n2e + (2LL * e + 1) * cc;
  // Hermitian part (x + x^H) / 2, as the reference's eigh symmetrises its input
  for (int q = tid; q < cc; q += blockDim.x) {
    const int k = q / n, l = q % n;
    const cd a = cvt<double>(m0[q]), at = cvt<double>(m0[l * n + k]);
    const cd b = cvt<double>(m1[q]), bt = cvt<double>(m1[l * n + k]);
    Mf[q] = mk<double>(0.5 * (a.re + at.re), 0.5 * (a.im - at.im));
    Mb[q] = mk<double>(0.5 * (b.re + bt.re), 0.5 * (b.im - bt.im));
  }
  __syncthreads();
  // Eigendecompositions by the block-level one-sided Jacobi, one message
  // after the other.  Read column-major, the row-major Hermitian M is
  // conj(M); with conj(M) J = G, the columns of conj(J) are eigenvectors of M
  // and the Rayleigh quotient j_p^H g_p its eigenvalue, sign included.  When
  // a column's quotient falls short of its norm |g_p| (eigenvalues +w and -w
  // sharing a singular subspace) the two-sided wave-level Jacobi redoes both.
  __shared__ int bad;
  if (tid == 0) bad = 0;
  for (int side = 0; side < 2; ++side) {
    cd* M = side == 0 ? Mf : Mb;
    cd* Q = side == 0 ? Qf : Qb;
    double* w = side == 0 ? wf : wb;
    block_jacobi_svd_n<32, lds_cd, true, false, false, true>((lds_cd*)M, nullptr, (lds_cd*)Q, sigma, perm, scal);
    if (tid < n) {
      double re = 0.0, s2 = 0.0;
#pragma unroll 8
      for (int k = 0; k < n; ++k) {
        const cd v = Q[tid * n + k], gk = M[tid * n + k];
        re += v.re * gk.re + v.im * gk.im;
        s2 += norm2(gk);
      }
      w[tid] = re;
      const double nrm = sqrt(s2);
      if (nrm > eps && fabs(re) < (1.0 - 1e-6) * nrm) bad = 1;
    }
    __syncthreads();
  }
  if (bad) {
    for (int q = tid; q < cc; q += blockDim.x) {
      const int k = q / n, l = q % n;
      const cd a = cvt<double>(m0[q]), at = cvt<double>(m0[l * n + k]);
      const cd b = cvt<double>(m1[q]), bt = cvt<double>(m1[l * n + k]);
      Mf[q] = mk<double>(0.5 * (a.re + at.re), 0.5 * (a.im - at.im));
      Mb[q] = mk<double>(0.5 * (b.re + bt.re), 0.5 * (b.im - bt.im));
    }
    __syncthreads();
    if (tid < 64)
      wave_jacobi_eigh_n<n>(Mf, Qf);
    else if (tid < 128)
      wave_jacobi_eigh_n<n>(Mb, Qb);
    __syncthreads();
    if (tid < n) {
      wf[tid] = Mf[tid * (n + 1)].re;
      wb[tid] = Mb[tid * (n + 1)].re;
    }
    __syncthreads();
  } else {
    // Q[k][p] = conj(J[k][p]) = conj(V[p * n + k]): transpose and conjugate
    cd tq[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int q = tid + 256 * t, side = q >> 10, r = q & 1023;
      const cd* Q = side == 0 ? Qf : Qb;
      tq[t] = conj(Q[(r % n) * n + r / n]);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int q = tid + 256 * t, side = q >> 10, r = q & 1023;
      cd* Q = side == 0 ? Qf : Qb;
      Q[r] = tq[t];
    }
    __syncthreads();
  }
  if (tid < 4 * n) {
    const int which = tid / n, i = tid % n;
    const double w = (which & 1) ? wb[i] : wf[i];
    rt[which][i] = which < 2 ? sqrt(safe_truncate(w, eps)) : sqrt(safe_inverse(w, eps));
  }
  __syncthreads();
  // P = Df^1/2 Qf^H conj(Qb) Db^1/2 -> Mf (Qf[k][i]: component k of eigenvector i)
  for (int q = tid; q < cc; q += blockDim.x) {
    const int i = q / n, j = q % n;
    cd acc = mk<double>(0, 0);
#pragma unroll 8
    for (int k = 0; k < n; ++k) fma_acc(acc, conj(Qf[k * n + i]), conj(Qb[k * n + j]));
    Mf[q] = (rt[0][i] * rt[1][j]) * acc;
  }
  __syncthreads();
  // H = Qf P -> Mb
  for (int q = tid; q < cc; q += blockDim.x) {
    const int r = q / n, j = q % n;
    cd acc = mk<double>(0, 0);
#pragma unroll 8
    for (int i = 0; i < n; ++i) fma_acc(acc, Qf[r * n + i], Mf[i * n + j]);
    Mb[q] = acc;
  }
  __syncthreads();
  // G = H Qb^T -> Mf, column-major (G[c * n + r]) for the Jacobi SVD
  for (int q = tid; q < cc; q += blockDim.x) {
    const int c = q / n, r = q % n;
    cd acc = mk<double>(0, 0);
#pragma unroll 8
    for (int j = 0; j < n; ++j) fma_acc(acc, Mb[r * n + j], Qb[c * n + j]);
    Mf[q] = acc;
  }
  __syncthreads();
  block_jacobi_svd_n<32, lds_cd, true, true, false, true>((lds_cd*)Mf, U, (lds_cd*)Mb, sigma, perm, scal);
  cd* V = Mb;
  double nrm = 0.0;
  for (int j = 0; j < n; ++j) nrm += sigma[j] * sigma[j];
  nrm = sqrt(nrm);
  // masked Frobenius norms of U and Vh
  double pu = 0.0, pv = 0.0;
  for (int q = tid; q < cc; q += blockDim.x) {
    const int j = q / n, i = q % n;
    const double lj = nrm > 0.0 ? sigma[j] / nrm : 0.0;
    if (safe_truncate(lj, eps) > eps) {
      pu += norm2(U[j * n + i]);
      pv += norm2(V[perm[j] * n + i]);
    }
  }
  const double nu = sqrt(block_sum(pu, red));
  const double nv = sqrt(block_sum(pv, red));
  for (int j = tid; j < n; j += blockDim.x) {
    const double lj = nrm > 0.0 ? safe_truncate(sigma[j] / nrm, eps) : 0.0;
    lam[(long long)e * n + j] = lj;
  }
  // T1 = Df^-1/2 Qf^H U -> Mf (G is no longer needed), then Xf = Qf T1 / nu;
  // T2 = Db^-1/2 Qb^H conj(V) -> Mf, then Xb = Qb T2 / nv.  Columns of
  // truncated singular values are zero.
  for (int side = 0; side < 2; ++side) {
    const cd* Q = side == 0 ? Qf : Qb;
    const double* ri = rt[2 + side];
    for (int q = tid; q < cc; q += blockDim.x) {
      const int m = q / n, k = q % n;
      cd acc = mk<double>(0, 0);
      if (side == 0) {
#pragma unroll 8
        for (int i = 0; i < n; ++i) fma_acc(acc, conj(Q[i * n + m]), U[k * n + i]);
      } else {
        const int pk = perm[k];
#pragma unroll 8
        for (int i = 0; i < n; ++i) fma_acc(acc, conj(Q[i * n + m]), conj(V[pk * n + i]));
      }
      Mf[q] = ri[m] * acc;
    }
    __syncthreads();
    cd* Xs = X + (2LL * e + side) * cc;
    const double sc = 1.0 / (side == 0 ? nu : nv);
    for (int q = tid; q < cc; q += blockDim.x) {
      const int b = q / n, k = q % n;
      const double lk = nrm > 0.0 ? safe_truncate(sigma[k] / nrm, eps) : 0.0;
      cd acc = mk<double>(0, 0);
      if (lk > eps) {
#pragma unroll 8
        for (int m = 0; m < n; ++m) fma_acc(acc, Q[b * n + m], Mf[m * n + k]);
        acc = sc * acc;
      }
      Xs[q] = acc;
    }
    __syncthreads();
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_absorb(DevGraph g, const cx<R>* T, const cd* X, cx<R>* gam, char* ws, long long wsb, const int* inst_active) {
  __shared__ double red[8];
  const int node = blockIdx.x;
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node], chi = g.chi, phys = g.phys;
  const long long cc = (long long)chi * chi;
  const long long S = ipow(chi, d) * phys;
  const int base = g.adj_ptr[node];
  cx<R>* w0 = (cx<R>*)(ws + (long long)blockIdx.x * wsb);
  cx<R>* w1 = w0 + S;
  const cx<R>* src = T + g.toff[node];
  int buf = 0;
  for (int i = 0; i < d; ++i) {
    const int e = g.adj_edge[base + i], s = g.adj_side[base + i];
    cx<R>* dst = buf ? w1 : w0;
    mode_product<R, double>(dst, src, X + (2LL * e + s) * cc, chi, d, phys, i, true);
    __syncthreads();
    src = dst;
    buf ^= 1;
  }
  double acc = 0.0;
  for (long long f = threadIdx.x; f < S; f += blockDim.x) acc += (double)norm2(src[f]);
  const double nrm = sqrt(block_sum(acc, red));
  const R inv = (R)(1.0 / nrm);
  cx<R>* out = gam + g.toff[node];
  for (long long f = threadIdx.x; f < S; f += blockDim.x) out[f] = inv * src[f];
}

// lexicographic complex comparison z > eps (real part first)
__device__ __forceinline__ bool complex_gt(cd z, double eps) { return z.re == eps ? z.im > 0.0 : z.re > eps; }

__host__ __device__ inline long long vidal_ws_bytes(int chi, int max_deg) {
  const long long cc = (long long)chi * chi;
  return 5 * cc * (long long)sizeof(cd) + ((long long)max_deg * chi + 4 * chi + 16) * sizeof(double) + 16;
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_vidal_node(DevGraph g, const cx<R>* gam, const double* lam, double eps, double* node_dist, char* ws,
                 long long wsb, const int* inst_active) {
  __shared__ const double* wp[kMaxDeg];
  const int node = blockIdx.x;
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node], chi = g.chi, phys = g.phys;
  const int cc = chi * chi;
  const int base = g.adj_ptr[node];
  cd* Rm = (cd*)(ws + (long long)blockIdx.x * wsb);
  cd* A = Rm + cc;
  cd* Us = A + cc;  // singular vectors of the fast nuclear norms (chi 16, 32)
  cd* Vs = Us + cc;
  double* w2 = (double*)(Vs + cc);
  double* rot = w2 + (long long)d * chi;
  int* flag = (int*)(rot + 4 * chi + 8);
  for (int t = threadIdx.x; t < d * chi; t += blockDim.x) {
    const int j = t / chi, b = t % chi;
    const double l = lam[(long long)g.adj_edge[base + j] * chi + b];
    w2[t] = l * l;
  }
  if (threadIdx.x < d) wp[threadIdx.x] = w2 + threadIdx.x * chi;
  __syncthreads();
  const cx<R>* t = gam + g.toff[node];
  double dist = 0.0;
  for (int idx = 0; idx < d; ++idx) {
    bond_gram<R>(Rm, t, t, chi, d, phys, idx, wp);
    __syncthreads();
    cd tr = mk<double>(0, 0);
    int cnt = 0;
    for (int k = 0; k < chi; ++k) tr = tr + Rm[k * chi + k];
    for (int k = 0; k < chi; ++k) cnt += complex_gt(cdiv(Rm[k * chi + k], tr), eps) ? 1 : 0;
    for (int q = threadIdx.x; q < cc; q += blockDim.x) {
      const int k = q / chi, l = q % chi;
      cd a = cdiv(Rm[q], tr), at = cdiv(Rm[l * chi + k], tr);
      cd h = mk<double>(0.5 * (a.re + at.re), 0.5 * (a.im - at.im));
      if (k == l && complex_gt(a, eps)) h.re -= 1.0 / cnt;
      A[q] = h;
    }
    __syncthreads();
    // sum |eigenvalues| of the Hermitian A = its nuclear norm: one-sided
    // Jacobi singular values at chi 16 (one wave) and 32 (the block), the
    // two-sided eigensolver otherwise
    if (chi == 32 && blockDim.x == 256) {
      __shared__ double sg[32];
      __shared__ int pm[32];
      block_jacobi_svd_n<32, cd, false, false>(A, nullptr, nullptr, sg, pm, rot);
      for (int k = 0; k < 32; ++k) dist += sg[k];
    } else if (chi == 16) {
      __shared__ double sg16[16];
      __shared__ int pm16[16];
      if (threadIdx.x < 64) wave_jacobi_svd_n<16>(A, Us, Vs, sg16, pm16, rot);
      __syncthreads();
      for (int k = 0; k < 16; ++k) dist += sg16[k];
    } else {
      jacobi_eigh(A, nullptr, chi, rot, flag);
      for (int k = 0; k < chi; ++k) dist += fabs(A[k * chi + k].re);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) node_dist[node] = dist;
}

__global__ void __launch_bounds__(256)
    k_vidal_reduce(const double* node_dist, const int* inst_node_ptr, const int* inst_edge_ptr, double* out,
                   const int* inst_active) {
  __shared__ double red[8];
  const int inst = blockIdx.x;
  if (inst_active && !inst_active[inst]) return;
  double acc = 0.0;
  for (int n = inst_node_ptr[inst] + threadIdx.x; n < inst_node_ptr[inst + 1]; n += blockDim.x) acc += node_dist[n];
  const double s = block_sum(acc, red);
  const int ne = inst_edge_ptr[inst + 1] - inst_edge_ptr[inst];
  if (threadIdx.x == 0) out[inst] = s / (2.0 * ne);
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_symmetric(DevGraph g, const cx<R>* gam, const double* lam, double eps, cx<R>* out) {
  __shared__ double red[8];
  const int node = blockIdx.x;
  const int d = g.deg[node], chi = g.chi, phys = g.phys;
  const long long S = ipow(chi, d) * phys;
  const int base = g.adj_ptr[node];
  const cx<R>* t = gam + g.toff[node];
  cx<R>* o = out + g.toff[node];
  double acc = 0.0;
  constexpr int KE = 32;  // elements per thread kept in registers (fast path)
  if (S <= (long long)KE * blockDim.x && d <= kMaxDeg && chi <= 64) {
    // bond factors once per node, 32-bit index arithmetic, one read and one
    // write of the tensor (same factor order and summation order as below)
    __shared__ double sq[kMaxDeg * 64];
    for (int i = threadIdx.x; i < d * chi; i += blockDim.x)
      sq[i] = sqrt(safe_truncate(lam[(long long)g.adj_edge[base + i / chi] * chi + i % chi], eps));
    __syncthreads();
    cx<R> v[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = threadIdx.x + k * blockDim.x;
      if (f < S) {
        int r = f / phys;
        double fac = 1.0;
        for (int j = d - 1; j >= 0; --j) {
          const int q = r / chi;
          fac *= sq[j * chi + (r - q * chi)];
          r = q;
        }
        v[k] = (R)fac * t[f];
        acc += (double)norm2(v[k]);
      }
    }
    const double nrm = sqrt(block_sum(acc, red));
    const R inv = (R)(1.0 / nrm);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = threadIdx.x + k * blockDim.x;
      if (f < S) o[f] = inv * v[k];
    }
    return;
  }
  for (long long f = threadIdx.x; f < S; f += blockDim.x) {
    long long r = f / phys;
    double fac = 1.0;
    for (int j = d - 1; j >= 0; --j) {
      const int b = (int)(r % chi);
      r /= chi;
      fac *= sqrt(safe_truncate(lam[(long long)g.adj_edge[base + j] * chi + b], eps));
    }
    const cx<R> v = (R)fac * t[f];
    o[f] = v;
    acc += (double)norm2(v);
  }
  const double nrm = sqrt(block_sum(acc, red));
  const R inv = (R)(1.0 / nrm);
  for (long long f = threadIdx.x; f < S; f += blockDim.x) o[f] = inv * o[f];
}

// ------------------------------------------------------------- launchers
long long edge_svd_ws_bytes_host(int chi) { return edge_svd_ws_bytes(chi); }

hipError_t launch_edge_svd(int rb, const DevGraph& g, const void* n2e, double eps, cd* X, double* lam, void* ws,
                           long long wsb, const int* inst_active, void* tape, hipStream_t s) {
  if (g.n_edges <= 0) return hipSuccess;
  if (g.chi == 16) {
    if (rb == 4)
      hipLaunchKernelGGL(k_edge_svd16<float>, dim3(g.n_edges), dim3(128), 0, s, g, (const cx<float>*)n2e, eps, X, lam,
                         inst_active, (char*)tape);
    else
      hipLaunchKernelGGL(k_edge_svd16<double>, dim3(g.n_edges), dim3(128), 0, s, g, (const cx<double>*)n2e, eps, X,
                         lam, inst_active, (char*)tape);
    return hipGetLastError();
  }
  if (g.chi == 32 && !tape && wsb >= 1024LL * (long long)sizeof(cd)) {
    if (rb == 4)
      hipLaunchKernelGGL(k_edge_svd32<float>, dim3(g.n_edges), dim3(256), 0, s, g, (const cx<float>*)n2e, eps, X, lam,
                         (char*)ws, wsb, inst_active);
    else
      hipLaunchKernelGGL(k_edge_svd32<double>, dim3(g.n_edges), dim3(256), 0, s, g, (const cx<double>*)n2e, eps, X,
                         lam, (char*)ws, wsb, inst_active);
    return hipGetLastError();
  }
  if (rb == 4)
    hipLaunchKernelGGL(k_edge_svd<float>, dim3(g.n_edges), dim3(256), 0, s, g, (const cx<float>*)n2e, eps, X, lam,
                       (char*)ws, wsb, inst_active, (char*)tape);
  else
    hipLaunchKernelGGL(k_edge_svd<double>, dim3(g.n_edges), dim3(256), 0, s, g, (const cx<double>*)n2e, eps, X,
                       lam, (char*)ws, wsb, inst_active, (char*)tape);
  return hipGetLastError();
}

hipError_t launch_absorb(int rb, const DevGraph& g, const void* T, const cd* X, void* gam, void* ws, long long wsb,
                         const int* inst_active, hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_absorb<float>, dim3(g.n_nodes), dim3(256), 0, s, g, (const cx<float>*)T, X,
                       (cx<float>*)gam, (char*)ws, wsb, inst_active);
  else
    hipLaunchKernelGGL(k_absorb<double>, dim3(g.n_nodes), dim3(256), 0, s, g, (const cx<double>*)T, X,
                       (cx<double>*)gam, (char*)ws, wsb, inst_active);
  return hipGetLastError();
}

hipError_t launch_vidal_distance(int rb, const DevGraph& g, const void* gam, const double* lam, double eps,
                                 double* node_dist, const int* inst_node_ptr, const int* inst_edge_ptr, int n_inst,
                                 double* inst_dist, void* ws, long long wsb, const int* inst_active, hipStream_t s) {
  if (g.n_nodes > 0) {
    if (rb == 4)
      hipLaunchKernelGGL(k_vidal_node<float>, dim3(g.n_nodes), dim3(256), 0, s, g, (const cx<float>*)gam, lam, eps,
                         node_dist, (char*)ws, wsb, inst_active);
    else
      hipLaunchKernelGGL(k_vidal_node<double>, dim3(g.n_nodes), dim3(256), 0, s, g, (const cx<double>*)gam, lam,
                         eps, node_dist, (char*)ws, wsb, inst_active);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  if (n_inst > 0)
    hipLaunchKernelGGL(k_vidal_reduce, dim3(n_inst), dim3(256), 0, s, node_dist, inst_node_ptr, inst_edge_ptr,
                       inst_dist, inst_active);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Vidal distance for the matrix-core graphs (degree 3, chi 16, phys 2,
// complex64).  The trace-normalised R of vidal_dist (tensor_ops.py:143-175),
// R = Gram over the other bonds of conj(Gamma) and Gamma * prod lambda_i^2,
// is the BP node update (tensor_ops.py:36-60) with the diagonal messages
// diag(lambda^2) on every bond: the diagonal-message MFMA node kernel
// (k_bp_nodes_lam2_d3c16) computes all of them, one per (edge, side).  One
// wave then forms R - ideal and its nuclear norm.

// Per (edge, side): |R/tr(R) - ideal|_* with ideal = diag(R_kk/tr > eps)/count,
// R Hermitised (it is a Gram matrix up to rounding).  4 waves per block.
// R in complex64 (the matrix-core path) or either precision (the streaming
// path of large node tensors).
template <typename RT>
__global__ void __launch_bounds__(256)
    k_vidal_nuclear16(DevGraph g, const cx<RT>* R, double eps, double* es_dist, const int* inst_active) {
  __shared__ cd A[4][256];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long es = (long long)blockIdx.x * 4 + w;
  if (es >= 2LL * g.n_edges) return;  // whole wave: no block barrier below
  const int e = (int)(es >> 1);
  if (inst_active && !inst_active[g.node_inst[g.edge_node[2 * e]]]) return;
  cd* G = A[w];
  const cx<RT>* r = R + es * 256;
  for (int i = lane; i < 256; i += 64) G[i] = cvt<double>(r[i]);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  cd tr = mk<double>(0.0, 0.0);
  for (int k = 0; k < 16; ++k) tr = tr + G[k * 17];
  int cnt = 0;
  for (int k = 0; k < 16; ++k) cnt += complex_gt(cdiv(G[k * 17], tr), eps) ? 1 : 0;
  cd h[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = lane + 64 * t, k = i >> 4, l = i & 15;
    const cd a = cdiv(G[i], tr), at = cdiv(G[l * 16 + k], tr);
    h[t] = mk<double>(0.5 * (a.re + at.re), 0.5 * (a.im - at.im));
    if (k == l && complex_gt(a, eps)) h[t].re -= 1.0 / cnt;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int t = 0; t < 4; ++t) G[lane + 64 * t] = h[t];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const double nuc = wave_nuclear_norm_n<16>(G, sizeof(RT) == 4);
  if (lane == 0) es_dist[es] = nuc;
}

// diag(lambda^2) messages at any chi and precision.
template <typename R>
__global__ void __launch_bounds__(256) k_lam2_diag_any(DevGraph g, const double* lam, cx<R>* e2n) {
  const int chi = g.chi;
  const long long cc = (long long)chi * chi, total = 2LL * g.n_edges * cc;
  for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < total;
       f += (long long)gridDim.x * blockDim.x) {
    const long long e = f / (2 * cc), q = f % cc;
    const int k = (int)(q / chi), l = (int)(q % chi);
    const double v = lam[e * chi + k];
    e2n[f] = mk<R>(k == l ? (R)(v * v) : (R)0, (R)0);
  }
}

// The nuclear-norm term of one (edge, side) at chi = 32, one block each:
// R Hermitised, the ideal subtracted, singular values by the block-level
// one-sided Jacobi without singular vectors (sum of |eigenvalues| of a
// Hermitian matrix); 16 KiB of LDS per block.
template <typename R>
__global__ void __launch_bounds__(256)
    k_vidal_nuclear32(DevGraph g, const cx<R>* Rm, double eps, double* es_dist, const int* inst_active) {
  __shared__ cd G[1024];
  __shared__ double sg[32], scal[64];
  __shared__ int perm[32];
  const long long es = blockIdx.x;
  const int e = (int)(es >> 1);
  if (inst_active && !inst_active[g.node_inst[g.edge_node[2 * e]]]) return;
  const cx<R>* r = Rm + es * 1024;
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) G[i] = cvt<double>(r[i]);
  __syncthreads();
  cd tr = mk<double>(0.0, 0.0);
  for (int k = 0; k < 32; ++k) tr = tr + G[k * 33];
  int cnt = 0;
  for (int k = 0; k < 32; ++k) cnt += complex_gt(cdiv(G[k * 33], tr), eps) ? 1 : 0;
  cd h[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = threadIdx.x + 256 * t, k = i >> 5, l = i & 31;
    const cd a = cdiv(G[i], tr), at = cdiv(G[l * 32 + k], tr);
    h[t] = mk<double>(0.5 * (a.re + at.re), 0.5 * (a.im - at.im));
    if (k == l && complex_gt(a, eps)) h[t].re -= 1.0 / cnt;
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 4; ++t) G[threadIdx.x + 256 * t] = h[t];
  __syncthreads();
  block_jacobi_svd_n<32, lds_cd, false, false, false, true>((lds_cd*)G, nullptr, nullptr, sg, perm, scal);
  if (threadIdx.x == 0) {
    double nuc = 0.0;
    for (int k = 0; k < 32; ++k) nuc += sg[k];
    es_dist[es] = nuc;
  }
}

// out[i] = sum over the (edge, side) terms of instance i / (2 |E_i|).
__global__ void __launch_bounds__(256)
    k_vidal_reduce_es(const double* es_dist, const int* inst_edge_ptr, double* out, const int* inst_active) {
  __shared__ double red[8];
  const int inst = blockIdx.x;
  if (inst_active && !inst_active[inst]) return;
  const int e0 = inst_edge_ptr[inst], e1 = inst_edge_ptr[inst + 1];
  double acc = 0.0;
  for (int q = 2 * e0 + threadIdx.x; q < 2 * e1; q += blockDim.x) acc += es_dist[q];
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[inst] = s / (2.0 * (e1 - e0));
}

long long vidal_fast_ws_bytes(const DevGraph& g) { return 2LL * g.n_edges * (256 * 8 + 8) + 256; }

// Vidal distance for large node tensors (chi 16 or 32) on the streaming node
// update: the lambda^2-weighted bond Grams are node updates from
// diag(lambda^2) messages.
long long vidal_big_ws_bytes(const DevGraph& g, int max_deg, int rb) {
  const long long cc = (long long)g.chi * g.chi;
  return big_ws_bytes(g, max_deg, g.n_nodes, rb) + 2 * (2LL * g.n_edges * cc * 2 * rb) + 2LL * g.n_edges * 8 + 512;
}

hipError_t launch_vidal_distance_big(int rb, const DevGraph& g, int min_deg, int max_deg, const void* gam,
                                     const double* lam,
                                     double eps, const int* all_nodes, const uint8_t* zero_src,
                                     const int* inst_edge_ptr, int n_inst, double* inst_dist, void* ws,
                                     const int* inst_active, hipStream_t s) {
  if (g.n_edges <= 0 || (g.chi != 16 && g.chi != 32)) return hipErrorInvalidValue;
  const long long cc = (long long)g.chi * g.chi, nm = 2LL * g.n_edges * cc;
  char* base = (char*)ws + (big_ws_bytes(g, max_deg, g.n_nodes, rb) + 255) / 256 * 256;
  void* msg = base;
  void* R = base + nm * 2 * rb;
  double* es_dist = (double*)((char*)R + nm * 2 * rb);
  if (rb == 4)
    hipLaunchKernelGGL(k_lam2_diag_any<float>, dim3(grid_for(nm, 256)), dim3(256), 0, s, g, lam, (cx<float>*)msg);
  else
    hipLaunchKernelGGL(k_lam2_diag_any<double>, dim3(grid_for(nm, 256)), dim3(256), 0, s, g, lam, (cx<double>*)msg);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = launch_bp_nodes_big(rb, g, min_deg, max_deg, all_nodes, g.n_nodes, gam, msg, msg, R, zero_src, ws,
                            inst_active, s);
  if (err != hipSuccess) return err;
  const unsigned n16 = (unsigned)((2LL * g.n_edges + 3) / 4), n32 = (unsigned)(2LL * g.n_edges);
  if (g.chi == 16 && rb == 4)
    hipLaunchKernelGGL(k_vidal_nuclear16<float>, dim3(n16), dim3(256), 0, s, g, (const cx<float>*)R, eps, es_dist,
                       inst_active);
  else if (g.chi == 16)
    hipLaunchKernelGGL(k_vidal_nuclear16<double>, dim3(n16), dim3(256), 0, s, g, (const cx<double>*)R, eps, es_dist,
                       inst_active);
  else if (rb == 4)
    hipLaunchKernelGGL(k_vidal_nuclear32<float>, dim3(n32), dim3(256), 0, s, g, (const cx<float>*)R, eps, es_dist,
                       inst_active);
  else
    hipLaunchKernelGGL(k_vidal_nuclear32<double>, dim3(n32), dim3(256), 0, s, g, (const cx<double>*)R, eps, es_dist,
                       inst_active);
  err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (n_inst > 0)
    hipLaunchKernelGGL(k_vidal_reduce_es, dim3(n_inst), dim3(256), 0, s, es_dist, inst_edge_ptr, inst_dist,
                       inst_active);
  return hipGetLastError();
}

hipError_t launch_vidal_distance_fast(const DevGraph& g, const void* gam, const double* lam, double eps,
                                      const int* all_nodes, const uint8_t* zero_src, const int* inst_edge_ptr,
                                      int n_inst, double* inst_dist, void* ws, const int* inst_active,
                                      hipStream_t s) {
  if (g.n_edges <= 0) return hipSuccess;
  // the lambda^2-weighted bond Grams R (node updates from diag(lambda^2)
  // messages), then per (edge, side) the nuclear distance
  cx<float>* R = (cx<float>*)ws;
  double* es_dist = (double*)(R + 2LL * g.n_edges * 256);
  (void)zero_src;
  hipError_t err = launch_bp_nodes_lam2_fast(g, all_nodes, g.n_nodes, gam, lam, R, inst_active, s);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_vidal_nuclear16<float>, dim3((unsigned)((2LL * g.n_edges + 3) / 4)), dim3(256), 0, s, g, R,
                     eps, es_dist, inst_active);
  err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (n_inst > 0)
    hipLaunchKernelGGL(k_vidal_reduce_es, dim3(n_inst), dim3(256), 0, s, es_dist, inst_edge_ptr, inst_dist,
                       inst_active);
  return hipGetLastError();
}

hipError_t launch_symmetric_gauge(int rb, const DevGraph& g, const void* gam, const double* lam, double eps, void* out,
                                  hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_symmetric<float>, dim3(g.n_nodes), dim3(256), 0, s, g, (const cx<float>*)gam, lam, eps,
                       (cx<float>*)out);
  else
    hipLaunchKernelGGL(k_symmetric<double>, dim3(g.n_nodes), dim3(256), 0, s, g, (const cx<double>*)gam, lam, eps,
                       (cx<double>*)out);
  return hipGetLastError();
}

}  // namespace dq

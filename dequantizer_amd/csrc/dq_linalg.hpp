// FP64 small dense linear algebra, one thread block per problem.
//
//   jacobi_eigh     Hermitian eigendecomposition, cyclic two-sided Jacobi
//                   (Golub & Van Loan, "Matrix Computations", 4th ed., 8.5)
//                   with the round-robin parallel ordering (Brent & Luk 1985)
//                   and a complex phase reduction of each 2x2 block.
//   jacobi_svd      one-sided (Hestenes 1958) Jacobi SVD, same ordering.
//   householder_qr  Householder QR with reflectors H = I - u u^H (|u|^2 = 2).
//
// All matrices live in memory visible to the whole block (LDS or global);
// every function is called by all threads of the block.
#pragma once
#include "dq_internal.hpp"

namespace dq {

// Round-robin pair k of round r among n2 (even) players; every unordered pair
// appears exactly once over the n2-1 rounds.
__device__ __forceinline__ void rr_pair(int r, int k, int n2, int& p, int& q) {
  if (k == 0) {
    p = n2 - 1;
    q = r;
  } else {
    p = (r + k) % (n2 - 1);
    q = (r - k + n2 - 1) % (n2 - 1);
  }
  if (p > q) {
    int t = p;
    p = q;
    q = t;
  }
}

// The pairs rr_pair(r, k, n2) of rounds r = 0, 1, ... advanced by one step
// per round: both players move up by one modulo n2 - 1, except player n2 - 1
// of pair 0 (k = 0).  Same pairs, without a division per round.
struct RoundRobin {
  int a, b;
  int k, m;  // m = n2 - 1
  __device__ __forceinline__ RoundRobin(int k_, int n2)
      : a(k_ == 0 ? n2 - 1 : k_), b(k_ == 0 ? 0 : n2 - 1 - k_), k(k_), m(n2 - 1) {}
  // this round's pair (p < q), then the state of the next round
  __device__ __forceinline__ void take(int& p, int& q) {
    p = a < b ? a : b;
    q = a < b ? b : a;
    a = k == 0 ? a : (a + 1 == m ? 0 : a + 1);
    b = b + 1 == m ? 0 : b + 1;
  }
};

// 1/sqrt(x) and 1/x for normal positive x: the hardware estimates
// (V_RSQ_F64, V_RCP_F64) refined by one Newton step, y <- y (3 - x y^2) / 2
// and y <- y (2 - x y), which takes them to about an ulp.  The rotation
// parameters below only need that accuracy (c^2 + s^2 = 1 to rounding), and
// the estimates avoid the long dependent sequences of correctly rounded
// division and square root on the critical path of every Jacobi round.
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return y * fma(-0.5 * x * y, y, 1.5);
}
__device__ __forceinline__ double rcp_nr(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return y * fma(-x, y, 2.0);
}

// Jacobi rotation for the Hermitian 2x2 block [[app, apq], [conj(apq), aqq]]:
// J = [[c, s], [-s e, c e]] with e = exp(-i arg apq) makes J^H B J diagonal.
// Returns false when the block is already diagonal to working precision.
__device__ __forceinline__ bool jacobi_params(double app, double aqq, cd apq, double& c, double& s, double& er,
                                              double& ei, double abs_tol) {
  const double n2 = norm2(apq);
  // |apq| > 1e-13 sqrt(|app aqq|), squared: 1e-13 sits above the
  // orthogonality FP64 can reach for n <= 64 (~n eps), so converged pairs
  // stop rotating; convergence is quadratic past it
  if (!(n2 > 1e-26 * fabs(app * aqq)) || !(n2 > abs_tol * abs_tol)) return false;
  if (!(n2 > 1e-300)) return false;  // |apq| below 1e-150: nothing to rotate
  const double inv = rsqrt_nr(n2);  // 1 / |apq|
  er = apq.re * inv;
  ei = -apq.im * inv;
  const double tau = 0.5 * (aqq - app) * inv;
  const double at = fabs(tau);
  // t = sign(tau) / (|tau| + sqrt(1 + tau^2)), the smaller root of
  // t^2 + 2 tau t - 1 = 0; for |tau| > 1e150 (tau^2 overflows) t = 1/(2 tau)
  double t;
  if (at > 1e150) {
    t = 0.5 * rcp_nr(at);
  } else {
    const double q = 1.0 + tau * tau;
    t = rcp_nr(at + q * rsqrt_nr(q));
  }
  if (tau < 0.0) t = -t;
  c = rsqrt_nr(1.0 + t * t);
  s = t * c;
  return true;
}

// The same rotation with a shorter dependent chain (two reciprocal square
// roots in sequence instead of three transcendentals): with d = (aqq - app)/2,
// m = |apq| and r = sqrt(d^2 + m^2), t = sign(d) m / (|d| + r), and
// 1 + t^2 = 2r / (r + |d|), so c = sqrt((r + |d|) / (2r)) and
// s = t c = sign(d) m / (2 r c).  Same thresholds; c^2 + s^2 = 1 to rounding.
// The sign convention of d = 0 is that of tau = 0 above (t > 0).
__device__ __forceinline__ bool jacobi_params_r(double app, double aqq, cd apq, double& c, double& s, double& er,
                                                double& ei, double abs_tol) {
  const double n2 = norm2(apq);
  if (!(n2 > 1e-26 * fabs(app * aqq)) || !(n2 > abs_tol * abs_tol)) return false;
  if (!(n2 > 1e-300)) return false;
  const double inv = rsqrt_nr(n2);  // 1 / m, in parallel with the chain below
  er = apq.re * inv;
  ei = -apq.im * inv;
  const double d = 0.5 * (aqq - app);
  const double h = fma(d, d, n2);
  const double y = rsqrt_nr(h);  // 1 / r
  const double c2 = 0.5 * fma(h, y, fabs(d)) * y;
  const double z = rsqrt_nr(c2);  // 1 / c
  c = c2 * z;
  const double sm = 0.5 * (n2 * inv) * y * z;
  s = d < 0.0 ? -sm : sm;
  return true;
}

// A (n x n row-major Hermitian) -> eigenvalues on the diagonal; V columns =
// eigenvectors.  rot: 4*ceil(n/2) + 1 doubles; flag: one int.
__device__ inline void jacobi_eigh(cd* A, cd* V, int n, double* rot, int* flag) {
  const int n2 = n + (n & 1), np = n2 / 2;
  if (V)
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) V[i] = mk<double>((i / n) == (i % n) ? 1.0 : 0.0, 0.0);
  // entries below ~eps_64 |A|_F are at the rounding level of A: no rotation
  if (threadIdx.x == 0) {
    double f = 0.0;
    for (int i = 0; i < n * n; ++i) f += norm2(A[i]);
    rot[4 * np] = 1e-16 * sqrt(f);
  }
  __syncthreads();
  const double abs_tol = rot[4 * np];
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (threadIdx.x == 0) *flag = 0;
    __syncthreads();
    for (int r = 0; r < n2 - 1; ++r) {
      for (int k = threadIdx.x; k < np; k += blockDim.x) {
        int p, q;
        rr_pair(r, k, n2, p, q);
        double c = 1.0, s = 0.0, er = 1.0, ei = 0.0;
        if (q < n && jacobi_params(A[p * n + p].re, A[q * n + q].re, A[p * n + q], c, s, er, ei, abs_tol))
          *flag = 1;
        rot[4 * k + 0] = c;
        rot[4 * k + 1] = s;
        rot[4 * k + 2] = er;
        rot[4 * k + 3] = ei;
      }
      __syncthreads();
      // A <- A J and V <- V J (columns p, q)
      for (int t = threadIdx.x; t < n * np; t += blockDim.x) {
        const int i = t / np, k = t % np;
        const double s = rot[4 * k + 1];
        if (s == 0.0) continue;
        int p, q;
        rr_pair(r, k, n2, p, q);
        const double c = rot[4 * k];
        const cd e = mk<double>(rot[4 * k + 2], rot[4 * k + 3]);
        const cd x = A[i * n + p], y = A[i * n + q];
        const cd ey = e * y;
        A[i * n + p] = mk<double>(c * x.re - s * ey.re, c * x.im - s * ey.im);
        A[i * n + q] = mk<double>(s * x.re + c * ey.re, s * x.im + c * ey.im);
        if (V) {
          const cd vx = V[i * n + p], vy = V[i * n + q];
          const cd evy = e * vy;
          V[i * n + p] = mk<double>(c * vx.re - s * evy.re, c * vx.im - s * evy.im);
          V[i * n + q] = mk<double>(s * vx.re + c * evy.re, s * vx.im + c * evy.im);
        }
      }
      __syncthreads();
      // A <- J^H A (rows p, q)
      for (int t = threadIdx.x; t < n * np; t += blockDim.x) {
        const int j = t / np, k = t % np;
        const double s = rot[4 * k + 1];
        if (s == 0.0) continue;
        int p, q;
        rr_pair(r, k, n2, p, q);
        const double c = rot[4 * k];
        const cd ec = mk<double>(rot[4 * k + 2], -rot[4 * k + 3]);
        const cd x = A[p * n + j], y = A[q * n + j];
        const cd ey = ec * y;
        A[p * n + j] = mk<double>(c * x.re - s * ey.re, c * x.im - s * ey.im);
        A[q * n + j] = mk<double>(s * x.re + c * ey.re, s * x.im + c * ey.im);
      }
      __syncthreads();
      for (int k = threadIdx.x; k < np; k += blockDim.x) {
        if (rot[4 * k + 1] == 0.0) continue;
        int p, q;
        rr_pair(r, k, n2, p, q);
        A[p * n + q] = mk<double>(0.0, 0.0);
        A[q * n + p] = mk<double>(0.0, 0.0);
        A[p * n + p].im = 0.0;
        A[q * n + q].im = 0.0;
      }
      __syncthreads();
    }
    const int any = *flag;
    __syncthreads();
    if (!any) break;
  }
}

// Wave-level two-sided Jacobi for an N x N Hermitian A (row-major, in LDS or
// global memory), N even, N / 2 pairs per round on groups of 128 / N lanes:
// the same rotations and convergence rule as jacobi_eigh, with the round's
// column and row updates separated by wave-level fences instead of block
// barriers.  A -> eigenvalues on the diagonal; V columns = eigenvectors.
// Call with all 64 lanes of one wave.  loose: stop after the first sweep
// whose rotations all had |a_pq| below 1e-8 sqrt|a_pp a_qq|, leaving
// off-diagonal terms of order 1e-16 (1e-8 at worst, for nearly equal
// eigenvalues), below single-precision rounding; otherwise the sweeps run
// until no pair rotates.
template <int N>
__device__ inline void wave_jacobi_eigh_n(cd* A, cd* V, bool loose = false) {
  constexpr int NP = N / 2;
  constexpr int LPP = 64 / NP;  // lanes per pair
  constexpr int RPL = N / LPP;  // rows (columns) per lane
  static_assert(N % 2 == 0 && LPP * NP == 64 && RPL * LPP == N, "wave eigh sizing");
  const int lane = threadIdx.x & 63;
  const int k = lane / LPP, sub = lane % LPP;
  for (int i = lane; i < N * N; i += 64) V[i] = mk<double>((i / N) == (i % N) ? 1.0 : 0.0, 0.0);
  double f = 0.0;
  for (int i = lane; i < N * N; i += 64) f += norm2(A[i]);
  const double abs_tol = 1e-16 * sqrt(wave_sum(f));
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0, large = 0;
    RoundRobin rrs(k, N);
    for (int r = 0; r < N - 1; ++r) {
      int p, q;
      rrs.take(p, q);
      double c = 1.0, s = 0.0, er = 1.0, ei = 0.0;
      const double app = A[p * N + p].re, aqq = A[q * N + q].re;
      const cd apq = A[p * N + q];
      const bool rot = jacobi_params_r(app, aqq, apq, c, s, er, ei, abs_tol);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (rot) {
        rotated = 1;
        if (norm2(apq) > 1e-16 * fabs(app * aqq)) large = 1;
        const cd e = mk<double>(er, ei);
        // A <- A J and V <- V J (columns p, q)
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int i = sub + t * LPP;
          const cd x = A[i * N + p], y = A[i * N + q];
          const cd ey = e * y;
          A[i * N + p] = mk<double>(c * x.re - s * ey.re, c * x.im - s * ey.im);
          A[i * N + q] = mk<double>(s * x.re + c * ey.re, s * x.im + c * ey.im);
          const cd vx = V[i * N + p], vy = V[i * N + q];
          const cd evy = e * vy;
          V[i * N + p] = mk<double>(c * vx.re - s * evy.re, c * vx.im - s * evy.im);
          V[i * N + q] = mk<double>(s * vx.re + c * evy.re, s * vx.im + c * evy.im);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (rot) {
        // A <- J^H A (rows p, q)
        const cd ec = mk<double>(er, -ei);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int j = sub + t * LPP;
          const cd x = A[p * N + j], y = A[q * N + j];
          const cd ey = ec * y;
          A[p * N + j] = mk<double>(c * x.re - s * ey.re, c * x.im - s * ey.im);
          A[q * N + j] = mk<double>(s * x.re + c * ey.re, s * x.im + c * ey.im);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (rot && sub == 0) {
        A[p * N + q] = mk<double>(0.0, 0.0);
        A[q * N + p] = mk<double>(0.0, 0.0);
        A[p * N + p].im = 0.0;
        A[q * N + q].im = 0.0;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (!__any(loose ? large : rotated)) break;
  }
}

// One-sided Jacobi SVD.  G: m x n column-major (ldg), m >= n, overwritten.
// Outputs sigma (descending), U (m x n, column-major, ldu; zero columns for
// sigma == 0) and the UNSORTED right vectors V (n x n column-major): with
// perm from this call, G_in = sum_j sigma[j] U[:, j] V[:, perm[j]]^H.
// rot: >= n doubles scratch (norms); flag: 1 int; perm: n ints.
__device__ inline void jacobi_svd(cd* G, int ldg, int m, int n, cd* U, int ldu, cd* V, double* sigma, double* rot,
                                  int* flag, int* perm) {
  const int n2 = n + (n & 1), np = n2 / 2;
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) V[i] = mk<double>((i / n) == (i % n) ? 1.0 : 0.0, 0.0);
  // threads per pair: power of two, <= 64, groups aligned inside waves
  int tpp = 1;
  while (tpp * 2 <= 64 && tpp * 2 * np <= (int)blockDim.x) tpp *= 2;
  const int grp = threadIdx.x / tpp, lane = threadIdx.x % tpp;
  if (threadIdx.x == 0) {
    double f = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) f += norm2(G[(long long)j * ldg + i]);
    rot[0] = 1e-32 * f;  // |g_p^H g_q| below ~eps_64^2 |G|_F^2 is rounding noise
  }
  __syncthreads();
  const double abs_tol = rot[0];
  __syncthreads();
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (threadIdx.x == 0) *flag = 0;
    __syncthreads();
    for (int r = 0; r < n2 - 1; ++r) {
      for (int kb = 0; kb < np; kb += (int)blockDim.x / tpp) {
        const int k = kb + grp;
        int p = 0, q = 0;
        const bool live = (grp < (int)blockDim.x / tpp) && (k < np);
        if (live) rr_pair(r, k, n2, p, q);
        const bool act = live && q < n;
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        if (act) {
          const cd* gp = G + (long long)p * ldg;
          const cd* gq = G + (long long)q * ldg;
          for (int i = lane; i < m; i += tpp) {
            const cd x = gp[i], y = gq[i];
            al += norm2(x);
            be += norm2(y);
            gr += x.re * y.re + x.im * y.im;
            gi += x.re * y.im - x.im * y.re;
          }
        }
        for (int o = tpp / 2; o > 0; o >>= 1) {
          al += __shfl_xor(al, o, 64);
          be += __shfl_xor(be, o, 64);
          gr += __shfl_xor(gr, o, 64);
          gi += __shfl_xor(gi, o, 64);
        }
        double c, s, er, ei;
        if (act && jacobi_params(al, be, mk<double>(gr, gi), c, s, er, ei, abs_tol)) {
          if (lane == 0) *flag = 1;
          const cd e = mk<double>(er, ei);
          cd* gp = G + (long long)p * ldg;
          cd* gq = G + (long long)q * ldg;
          for (int i = lane; i < m; i += tpp) {
            const cd x = gp[i], ey = e * gq[i];
            gp[i] = mk<double>(c * x.re - s * ey.re, c * x.im - s * ey.im);
            gq[i] = mk<double>(s * x.re + c * ey.re, s * x.im + c * ey.im);
          }
          cd* vp = V + (long long)p * n;
          cd* vq = V + (long long)q * n;
          for (int i = lane; i < n; i += tpp) {
            const cd x = vp[i], ey = e * vq[i];
            vp[i] = mk<double>(c * x.re - s * ey.re, c * x.im - s * ey.im);
            vq[i] = mk<double>(s * x.re + c * ey.re, s * x.im + c * ey.im);
          }
        }
      }
      __syncthreads();
    }
    const int any = *flag;
    __syncthreads();
    if (!any) break;
  }
  // column norms -> sigma, then sort descending
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += norm2(G[(long long)j * ldg + i]);
    rot[j] = sqrt(s);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < n; ++j) perm[j] = j;
    for (int a = 0; a < n; ++a) {
      int best = a;
      for (int b = a + 1; b < n; ++b)
        if (rot[perm[b]] > rot[perm[best]]) best = b;
      int t = perm[a];
      perm[a] = perm[best];
      perm[best] = t;
    }
    for (int j = 0; j < n; ++j) sigma[j] = rot[perm[j]];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < m * n; t += blockDim.x) {
    const int j = t / m, i = t % m;
    const double s = sigma[j];
    const cd g = G[(long long)perm[j] * ldg + i];
    U[(long long)j * ldu + i] = s > 0.0 ? mk<double>(g.re / s, g.im / s) : mk<double>(0.0, 0.0);
  }
  __syncthreads();
}

// Wave-synchronous one-sided Jacobi SVD: ONE wavefront (64 lanes, all active)
// factors G (m x n column-major, leading dimension m, m >= n, n <= 32,
// overwritten) with the round-robin pair order; lanes of a pair split the rows.
// No block barrier: the columns live in LDS and every round is ordered by
// program order within the wave.  Outputs as jacobi_svd: sigma (descending),
// U (m x n column-major, zero columns for sigma == 0), unsorted V (n x n
// column-major) with perm.  Needs 1 double of scratch in `scal`.
__device__ inline void wave_jacobi_svd(cd* G, int m, int n, cd* U, cd* V, double* sigma, int* perm, double* scal) {
  const int lane = threadIdx.x & 63;
  const int n2 = n + (n & 1), np = n2 / 2;
  int lpp = 1;
  while (lpp * 2 * np <= 64 && lpp * 2 <= 16) lpp *= 2;
  const int k = lane / lpp, sub = lane % lpp;
  for (int i = lane; i < n * n; i += 64) V[i] = mk<double>((i / n) == (i % n) ? 1.0 : 0.0, 0.0);
  double f = 0.0;
  for (int i = lane; i < m * n; i += 64) f += norm2(G[i]);
  f = wave_sum(f);
  const double abs_tol = 1e-32 * f;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0;
    for (int r = 0; r < n2 - 1; ++r) {
      int p = 0, q = 0;
      const bool live = k < np;
      if (live) rr_pair(r, k, n2, p, q);
      const bool act = live && q < n;
      double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
      if (act) {
        const cd* gp = G + p * m;
        const cd* gq = G + q * m;
        for (int i = sub; i < m; i += lpp) {
          const cd x = gp[i], y = gq[i];
          al += norm2(x);
          be += norm2(y);
          gr += x.re * y.re + x.im * y.im;
          gi += x.re * y.im - x.im * y.re;
        }
      }
      for (int o = lpp / 2; o > 0; o >>= 1) {
        al += __shfl_xor(al, o, 64);
        be += __shfl_xor(be, o, 64);
        gr += __shfl_xor(gr, o, 64);
        gi += __shfl_xor(gi, o, 64);
      }
      double c, s, er, ei;
      if (act && jacobi_params(al, be, mk<double>(gr, gi), c, s, er, ei, abs_tol)) {
        rotated = 1;
        const cd e = mk<double>(er, ei);
        cd* gp = G + p * m;
        cd* gq = G + q * m;
        for (int i = sub; i < m; i += lpp) {
          const cd x = gp[i], ey = e * gq[i];
          gp[i] = mk<double>(c * x.re - s * ey.re, c * x.im - s * ey.im);
          gq[i] = mk<double>(s * x.re + c * ey.re, s * x.im + c * ey.im);
        }
        cd* vp = V + p * n;
        cd* vq = V + q * n;
        for (int i = sub; i < n; i += lpp) {
          const cd x = vp[i], ey = e * vq[i];
          vp[i] = mk<double>(c * x.re - s * ey.re, c * x.im - s * ey.im);
          vq[i] = mk<double>(s * x.re + c * ey.re, s * x.im + c * ey.im);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (!__any(rotated)) break;
  }
  // column norms, rank sort (descending, ties by index), normalised left vectors
  double nrm = 0.0;
  if (lane < n) {
    for (int i = 0; i < m; ++i) nrm += norm2(G[lane * m + i]);
    nrm = sqrt(nrm);
    sigma[lane] = nrm;  // temporary: unsorted
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < n) {
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double o = sigma[j];
      rank += (o > nrm || (o == nrm && j < lane)) ? 1 : 0;
    }
    perm[rank] = lane;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < n) scal[lane] = sigma[perm[lane]];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < n) sigma[lane] = scal[lane];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int t = lane; t < m * n; t += 64) {
    const int j = t / m, i = t % m;
    const double sg = sigma[j];
    const cd g = G[perm[j] * m + i];
    U[j * m + i] = sg > 0.0 ? mk<double>(g.re / sg, g.im / sg) : mk<double>(0.0, 0.0);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Fixed-size variant of wave_jacobi_svd for square N x N (N = 16 or 32):
// loop bounds are compile-time constants, so each lane loads its rows of a
// column pair in one burst, rotates them in registers and stores them back.
// loose: stop after the first sweep whose rotations all had |g_pq| below
// 1e-8 sqrt(|g_p|^2 |g_q|^2) (the rule of wave_jacobi_eigh_n).
template <int N>
__device__ inline void wave_jacobi_svd_n(cd* G, cd* U, cd* V, double* sigma, int* perm, double* scal,
                                         bool loose = false) {
  constexpr int NP = N / 2;
  constexpr int LPP = 64 / NP;    // lanes per pair
  constexpr int RPL = N / LPP;    // rows per lane
  const int lane = threadIdx.x & 63;
  const int k = lane / LPP, sub = lane % LPP;
  for (int i = lane; i < N * N; i += 64) V[i] = mk<double>((i / N) == (i % N) ? 1.0 : 0.0, 0.0);
  double f = 0.0;
  for (int i = lane; i < N * N; i += 64) f += norm2(G[i]);
  f = wave_sum(f);
  const double abs_tol = 1e-32 * f;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0, large = 0;
    RoundRobin rrs(k, N);
    for (int r = 0; r < N - 1; ++r) {
      int p, q;
      rrs.take(p, q);
      cd x[RPL], y[RPL];
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        x[t] = G[p * N + sub + t * LPP];
        y[t] = G[q * N + sub + t * LPP];
      }
      // (the sums start from the first row's terms: adding them to 0.0
      // would cost an instruction each, x + 0.0 not folding for x = -0.0)
      double al = norm2(x[0]), be = norm2(y[0]);
      double gr = x[0].re * y[0].re + x[0].im * y[0].im;
      double gi = x[0].re * y[0].im - x[0].im * y[0].re;
#pragma unroll
      for (int t = 1; t < RPL; ++t) {
        al += norm2(x[t]);
        be += norm2(y[t]);
        gr += x[t].re * y[t].re + x[t].im * y[t].im;
        gi += x[t].re * y[t].im - x[t].im * y[t].re;
      }
      group_sum4<LPP>(al, be, gr, gi);
      double c, s, er, ei;
      if (jacobi_params_r(al, be, mk<double>(gr, gi), c, s, er, ei, abs_tol)) {
        rotated = 1;
        if (gr * gr + gi * gi > 1e-16 * al * be) large = 1;
        const cd e = mk<double>(er, ei);
        cd vx[RPL], vy[RPL];
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          vx[t] = V[p * N + sub + t * LPP];
          vy[t] = V[q * N + sub + t * LPP];
        }
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const cd ey = e * y[t];
          G[p * N + sub + t * LPP] = mk<double>(c * x[t].re - s * ey.re, c * x[t].im - s * ey.im);
          G[q * N + sub + t * LPP] = mk<double>(s * x[t].re + c * ey.re, s * x[t].im + c * ey.im);
          const cd evy = e * vy[t];
          V[p * N + sub + t * LPP] = mk<double>(c * vx[t].re - s * evy.re, c * vx[t].im - s * evy.im);
          V[q * N + sub + t * LPP] = mk<double>(s * vx[t].re + c * evy.re, s * vx[t].im + c * evy.im);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (!__any(loose ? large : rotated)) break;
  }
  double nrm = 0.0;
  if (lane < N) {
#pragma unroll 8
    for (int i = 0; i < N; ++i) nrm += norm2(G[lane * N + i]);
    nrm = sqrt(nrm);
    sigma[lane] = nrm;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < N) {
    int rank = 0;
    for (int j = 0; j < N; ++j) {
      const double o = sigma[j];
      rank += (o > nrm || (o == nrm && j < lane)) ? 1 : 0;
    }
    perm[rank] = lane;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < N) scal[lane] = sigma[perm[lane]];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < N) sigma[lane] = scal[lane];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int t = lane; t < N * N; t += 64) {
    const int j = t / N, i = t % N;
    const double sg = sigma[j];
    const cd g = G[perm[j] * N + i];
    U[j * N + i] = sg > 0.0 ? mk<double>(g.re / sg, g.im / sg) : mk<double>(0.0, 0.0);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Sum of the singular values (nuclear norm) of a square N x N matrix G
// (column-major, overwritten) by the same wave-synchronous one-sided Jacobi
// iteration, without accumulating singular vectors.  ONE wavefront; every
// lane gets the result.  loose: the stopping rule of wave_jacobi_svd_n; the
// column norms then differ from the singular values by terms of second order
// in the remaining off-diagonal products.
template <int N>
__device__ inline double wave_nuclear_norm_n(cd* G, bool loose = false) {
  constexpr int NP = N / 2;
  constexpr int LPP = 64 / NP;
  constexpr int RPL = N / LPP;
  const int lane = threadIdx.x & 63;
  const int k = lane / LPP, sub = lane % LPP;
  double f = 0.0;
  for (int i = lane; i < N * N; i += 64) f += norm2(G[i]);
  f = wave_sum(f);
  const double abs_tol = 1e-32 * f;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0, large = 0;
    RoundRobin rrs(k, N);
    for (int r = 0; r < N - 1; ++r) {
      int p, q;
      rrs.take(p, q);
      cd x[RPL], y[RPL];
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        x[t] = G[p * N + sub + t * LPP];
        y[t] = G[q * N + sub + t * LPP];
      }
      // (the sums start from the first row's terms: adding them to 0.0
      // would cost an instruction each, x + 0.0 not folding for x = -0.0)
      double al = norm2(x[0]), be = norm2(y[0]);
      double gr = x[0].re * y[0].re + x[0].im * y[0].im;
      double gi = x[0].re * y[0].im - x[0].im * y[0].re;
#pragma unroll
      for (int t = 1; t < RPL; ++t) {
        al += norm2(x[t]);
        be += norm2(y[t]);
        gr += x[t].re * y[t].re + x[t].im * y[t].im;
        gi += x[t].re * y[t].im - x[t].im * y[t].re;
      }
      group_sum4<LPP>(al, be, gr, gi);
      double c, s, er, ei;
      if (jacobi_params_r(al, be, mk<double>(gr, gi), c, s, er, ei, abs_tol)) {
        rotated = 1;
        if (gr * gr + gi * gi > 1e-16 * al * be) large = 1;
        const cd e = mk<double>(er, ei);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const cd ey = e * y[t];
          G[p * N + sub + t * LPP] = mk<double>(c * x[t].re - s * ey.re, c * x[t].im - s * ey.im);
          G[q * N + sub + t * LPP] = mk<double>(s * x[t].re + c * ey.re, s * x[t].im + c * ey.im);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (!__any(loose ? large : rotated)) break;
  }
  double nrm = 0.0;
  if (lane < N) {
#pragma unroll 8
    for (int i = 0; i < N; ++i) nrm += norm2(G[lane * N + i]);
    nrm = sqrt(nrm);
  }
  return wave_sum(nrm);
}

// Pair (p, q), p < q, of slot k in round r of the block-cyclic ordering of
// block_jacobi_svd_n (kBlockOrder) for N = 32 or 64.
template <int N>
__device__ __forceinline__ void block_order_pair(int r, int k, int& p, int& q) {
  constexpr int S = N / 8;  // slots (pairs) per wave and round
  const int o = r < 2 * S - 1 ? 0 : 1 + (r - (2 * S - 1)) / S;
  const int i = r < 2 * S - 1 ? r : (r - (2 * S - 1)) % S;
  const int sl = k % S;
  int bp, bq;
  rr_pair(o, k / S, 8, bp, bq);
  if (o == 0 && i < S - 1) {
    // pairs inside a block: round m pairs a with a ^ m; slot j of a block
    // takes the j-th a without the top bit of m
    const int m = i + 1, blk = sl < S / 2 ? bp : bq, j = sl % (S / 2);
    const int h = 31 - __builtin_clz(m);
    const int a = ((j >> h) << (h + 1)) | (j & ((1 << h) - 1));
    p = S * blk + a;
    q = S * blk + (a ^ m);
  } else {
    p = S * bp + sl;
    q = S * bq + (sl ^ (o == 0 ? i - (S - 1) : i));
  }
}

// Block-parallel fixed-size one-sided Jacobi SVD (N = 32, 256 threads): the
// 16 pairs of a round run on 16-lane groups (4 per wave), 2 rows per lane, one
// block barrier per round.  Same outputs as jacobi_svd.  scal: N doubles.
// kV = false: no right vectors (V is neither formed nor read; may be null);
// kU = false: U is not written (may be null).  G is left holding the rotated
// columns G J in their unsorted order (column p of G J goes with column p of
// J = V; sigma and perm give the sorted order).
// kLoose: the sweeps stop after the first sweep whose rotations all had |g|
// below 1e-4 sqrt(al be) instead of 1e-8: the off-diagonal terms left are of
// order 1e-8 relative (quadratic convergence), below the rounding of
// single-precision data, for factors that are rounded to single precision.
// kBlockOrder (N = 32 or 64): a block-cyclic ordering of the same N (N - 1) / 2
// pairs per sweep.  The columns form 8 blocks of S = N / 8 (the pairs a wave
// takes per round); in outer round o (of 7) wave w owns the block pair
// (P, Q) = rr_pair(o, w, 8) and takes its S^2 cross pairs in S rounds (slot s
// of the wave pairs column S P + s with S Q + (s ^ t)), after the pairs inside
// P and inside Q in S - 1 rounds at o = 0 (round m pairs a with a ^ m, the
// first S / 2 slots in P, the others in Q).  A wave only touches the 2S
// columns it owns within an outer round, so its rounds there are separated
// by wave-level fences, and block barriers fall between outer rounds: 7 per
// sweep instead of N - 1.
template <int N, typename P, bool kV = true, bool kU = true, bool kLoose = false, bool kBlockOrder = false>
__device__ inline void block_jacobi_svd_n(P* G, cd* U, P* V, double* sigma, int* perm, double* scal) {
  static_assert(!kBlockOrder || N == 32 || N == 64, "the block ordering is sized for N = 32 or 64");
  constexpr double kLarge2 = kLoose ? 1e-8 : 1e-16;  // squared relative size of a "large" rotation
  constexpr int NP = N / 2;
  constexpr int LPP = 256 / NP;  // 16 lanes per pair for N = 32
  constexpr int RPL = N / LPP;   // rows per lane
  static_assert(LPP <= 64 && RPL >= 1, "block SVD sizing");
  const int k = threadIdx.x / LPP, sub = threadIdx.x % LPP;
  if (kV)
    for (int i = threadIdx.x; i < N * N; i += blockDim.x) cst(V, i, mk<double>((i / N) == (i % N) ? 1.0 : 0.0, 0.0));
  double f = 0.0;
  for (int i = threadIdx.x; i < N * N; i += blockDim.x) f += norm2(cld(G, i));
  f = block_sum(f, scal);
  const double abs_tol = 1e-32 * f;
  // N = 32 in the block ordering: the 31 x 16 pairs (p | q << 8) in a table
  // built by the block before the first sweep, one LDS read per round
  // instead of the index arithmetic
  __shared__ unsigned short ptab[kBlockOrder && N == 32 ? (N - 1) * NP : 1];
  if (kBlockOrder && N == 32) {
    for (int e = threadIdx.x; e < (N - 1) * NP; e += blockDim.x) {
      int tp, tq;
      block_order_pair<N>(e / NP, e % NP, tp, tq);
      ptab[e] = (unsigned short)(tp | (tq << 8));
    }
    __syncthreads();
  }
  for (int sweep = 0; sweep < 60; ++sweep) {
    // rotated: any rotation this sweep; large: any with |g| above 1e-8 of
    // sqrt(al be) (1e-4 with kLoose).  A sweep whose rotations were all below 1e-8 leaves
    // off-diagonal terms of order 1e-16 (quadratic convergence), under the
    // 1e-13 threshold of the rotation test: the sweep that would only
    // confirm convergence is skipped.
    int rotated = 0, large = 0;
    RoundRobin rrs(k, N);  // (the round-robin ordering)
    for (int r = 0; r < N - 1; ++r) {
      int p, q;
      bool block_sync = true;  // uniform over the block: depends on r only
      if (kBlockOrder) {
        constexpr int S = N / 8;  // slots (pairs) per wave and round
        const int o = r < 2 * S - 1 ? 0 : 1 + (r - (2 * S - 1)) / S;
        const int i = r < 2 * S - 1 ? r : (r - (2 * S - 1)) % S;
        if (N == 32) {
          const unsigned pq = ptab[r * NP + k];
          p = (int)(pq & 0xffu);
          q = (int)(pq >> 8);
        } else {
          block_order_pair<N>(r, k, p, q);
        }
        block_sync = o == 0 ? i == 2 * S - 2 : i == S - 1;
      } else {
        rrs.take(p, q);
      }
      cd x[RPL], y[RPL], vx[RPL], vy[RPL];
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        x[t] = cld(G, p * N + sub + t * LPP);
        y[t] = cld(G, q * N + sub + t * LPP);
        if (kV) {
          vx[t] = cld(V, p * N + sub + t * LPP);
          vy[t] = cld(V, q * N + sub + t * LPP);
        }
      }
      // (the sums start from the first row's terms: adding them to 0.0
      // would cost an instruction each, x + 0.0 not folding for x = -0.0)
      double al = norm2(x[0]), be = norm2(y[0]);
      double gr = x[0].re * y[0].re + x[0].im * y[0].im;
      double gi = x[0].re * y[0].im - x[0].im * y[0].re;
#pragma unroll
      for (int t = 1; t < RPL; ++t) {
        al += norm2(x[t]);
        be += norm2(y[t]);
        gr += x[t].re * y[t].re + x[t].im * y[t].im;
        gi += x[t].re * y[t].im - x[t].im * y[t].re;
      }
      group_sum4<LPP>(al, be, gr, gi);
      double c, s, er, ei;
      if (jacobi_params_r(al, be, mk<double>(gr, gi), c, s, er, ei, abs_tol)) {
        rotated = 1;
        if (gr * gr + gi * gi > kLarge2 * al * be) large = 1;
        const cd e = mk<double>(er, ei);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const cd ey = e * y[t];
          cst(G, p * N + sub + t * LPP, mk<double>(c * x[t].re - s * ey.re, c * x[t].im - s * ey.im));
          cst(G, q * N + sub + t * LPP, mk<double>(s * x[t].re + c * ey.re, s * x[t].im + c * ey.im));
          if (kV) {
            const cd evy = e * vy[t];
            cst(V, p * N + sub + t * LPP, mk<double>(c * vx[t].re - s * evy.re, c * vx[t].im - s * evy.im));
            cst(V, q * N + sub + t * LPP, mk<double>(s * vx[t].re + c * evy.re, s * vx[t].im + c * evy.im));
          }
        }
      }
      if (block_sync) {
        __syncthreads();
      } else {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (!__syncthreads_or(large)) break;
  }
  if (threadIdx.x < N) {
    double nrm = 0.0;
    for (int i = 0; i < N; ++i) nrm += norm2(cld(G, threadIdx.x * N + i));
    scal[threadIdx.x] = sqrt(nrm);
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const double me = scal[threadIdx.x];
    int rank = 0;
    for (int j = 0; j < N; ++j) rank += (scal[j] > me || (scal[j] == me && j < (int)threadIdx.x)) ? 1 : 0;
    perm[rank] = threadIdx.x;
    sigma[rank] = me;
  }
  __syncthreads();
  if (!kU) return;
  for (int t = threadIdx.x; t < N * N; t += blockDim.x) {
    const int j = t / N, i = t % N;
    const double sg = sigma[j];
    const cd g = cld(G, perm[j] * N + i);
    U[j * N + i] = sg > 0.0 ? mk<double>(g.re / sg, g.im / sg) : mk<double>(0.0, 0.0);
  }
  __syncthreads();
}

// V columns permuted by perm: out[:, j] = V[:, perm[j]] (separate buffer).
__device__ inline void permute_cols(const cd* V, cd* out, int n, const int* perm) {
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
    const int j = t / n, i = t % n;
    out[(long long)j * n + i] = V[(long long)perm[j] * n + i];
  }
  __syncthreads();
}

// Factor R with R^H R = G for a Hermitian positive semi-definite G (n x n,
// row-major, destroyed) by Cholesky with diagonal pivoting (Higham,
// "Accuracy and Stability of Numerical Algorithms", 10.3): P^T G P = L L^H,
// R = L^H P^T.  Pivots below n * eps_64 * max diag end the factorisation
// (the remaining rows of R are zero).  R: n x n row-major.  piv: n ints;
// scal: 2 doubles of scratch.
__device__ inline void pivoted_cholesky_swap(cd* G, int n, cd* R, int* piv, double* scal);

// The same factorisation for n <= 64 without moving rows and columns: the
// chosen pivots are tracked in a bit mask (identical in every thread) and
// every wave finds the next pivot itself, so a step costs two block barriers
// (after the column scaling, after the trailing update) instead of up to
// five.  The pivot's square root and its reciprocal come from the refined
// reciprocal square root (rsqrt_nr); ties between equal pivots go to the
// smaller original index.
__device__ inline void pivoted_cholesky(cd* G, int n, cd* R, int* piv, double* scal) {
  if (n > 64) {
    pivoted_cholesky_swap(G, n, R, piv, scal);
    return;
  }
  const int lane = threadIdx.x & 63;
  // the block size once, in a scalar register: read inside the loops it is
  // a load from the dispatch packet after every barrier
  const int bd = __builtin_amdgcn_readfirstlane(blockDim.x);
  for (int i = threadIdx.x; i < n * n; i += bd) R[i] = mk<double>(0.0, 0.0);
  double mx = 0.0;
  for (int i = 0; i < n; ++i) mx = fmax(mx, G[i * n + i].re);
  const double tol = n * 2.3e-16 * mx;
  unsigned long long done = 0;
  int k = 0;
  auto take = [](double& best, int& j, double ob, int oj) {
    if (ob > best || (ob == best && oj < j)) {
      best = ob;
      j = oj;
    }
  };
  for (; k < n; ++k) {
    double best = -1.0;
    int j = n;
    if (lane < n && !((done >> lane) & 1ull)) {
      best = G[lane * n + lane].re;
      j = lane;
    }
    // argmax over the wave: within each 16-lane row by DPP moves (every lane
    // of the row ends with the row's result), then across rows
    take(best, j, dpp_move<0xB1>(best), dpp_move_int<0xB1>(j));
    take(best, j, dpp_move<0x4E>(best), dpp_move_int<0x4E>(j));
    take(best, j, dpp_move<0x141>(best), dpp_move_int<0x141>(j));
    take(best, j, dpp_move<0x140>(best), dpp_move_int<0x140>(j));
    take(best, j, __shfl_xor(best, 16, 64), __shfl_xor(j, 16, 64));
    take(best, j, __shfl_xor(best, 32, 64), __shfl_xor(j, 32, 64));
    if (!(best > tol)) break;
    if (threadIdx.x == 0) piv[k] = j;
    const double inv = rsqrt_nr(best);  // best > tol > 0
    const double lkk = best * inv;
    // column j of L below the pivot: the rows not chosen yet
    for (int i = threadIdx.x; i < n; i += bd)
      if (i != j && !((done >> i) & 1ull)) G[i * n + j] = inv * G[i * n + j];
    done |= 1ull << j;
    __syncthreads();
    // the pivot's diagonal is written only now: before the barrier other
    // waves may still be reading the diagonal for their own pivot search
    if (threadIdx.x == 0) G[j * n + j] = mk<double>(lkk, 0.0);
    const int sh = n == 32 ? 5 : (n == 16 ? 4 : 0);  // row/column by shifts for n = 16, 32
    for (int f = threadIdx.x; f < n * n; f += bd) {
      const int i = sh ? f >> sh : f / n, c = sh ? f & (n - 1) : f % n;
      if (!((done >> i) & 1ull) && !((done >> c) & 1ull)) G[i * n + c] = G[i * n + c] - G[i * n + j] * conj(G[c * n + j]);
    }
    __syncthreads();
  }
  // the indices never chosen follow the pivots, in increasing order
  if (threadIdx.x == 0) {
    int m = k;
    for (int i = 0; i < n; ++i)
      if (!((done >> i) & 1ull)) piv[m++] = i;
  }
  __syncthreads();
  // R[i][piv[m]] = conj(L[m][i]) for m >= i, rows i < k; L[m][i] sits at
  // G[piv[m]][piv[i]]
  for (int f = threadIdx.x; f < k * n; f += bd) {
    const int i = f / n, m = f % n;
    if (m >= i) R[i * n + piv[m]] = m == i ? conj(G[piv[i] * n + piv[i]]) : conj(G[piv[m] * n + piv[i]]);
  }
  __syncthreads();
}

// Value of lane `src` in every lane (a scalar read of a vector register).
__device__ __forceinline__ double readlane_d(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Pivoted Cholesky R^H R = G of an n x n Hermitian PSD G (n <= 32, row-major
// in LDS) on ONE wavefront, left-looking, with no block barrier: the same
// factorisation, pivot rule (largest remaining diagonal, ties to the smaller
// index) and stopping tolerance (n 2.3e-16 max diag) as pivoted_cholesky.
// Lane (i, h) = (lane % 32, lane / 32) owns row i; the remaining diagonal
// lives in registers, so the pivot search is a DPP argmax and two lane reads.
// Step k computes column k of L for every row i not chosen yet,
//   L[i][k] = (G[i][j] - sum_{m<k} L[i][m] conj(L[j][m])) / L[j][k],
// the two lane halves summing alternate m.  Column k is stored over row j of
// G (a chosen pivot's row is never read again; G[i][j] = conj(G[j][i]) is
// read from row j before it is overwritten), so row piv[m] of G ends as
// column m of L.  Outputs: piv (n ints: the pivots, then the unchosen
// indices ascending) and R (n x n row-major, any memory) with
// R[m][c] = conj(L[c][m]) for m < rank, zero rows after.  Returns the rank.
// Call with all 64 lanes of one wave.
template <typename P, typename I>
__device__ __attribute__((noinline)) int wave_pivoted_cholesky(P* G, int n, cd* R, I* piv) {
  const int lane = threadIdx.x & 63;
  const int i = lane & 31, h = lane >> 5;
  const bool row = i < n;
  double di = row ? cld(G, i * n + i).re : 0.0;
  double mx = h == 0 && row ? di : 0.0;
  mx = fmax(mx, dpp_move<0xB1>(mx));
  mx = fmax(mx, dpp_move<0x4E>(mx));
  mx = fmax(mx, dpp_move<0x141>(mx));
  mx = fmax(mx, dpp_move<0x140>(mx));
  mx = fmax(readlane_d(mx, 0), readlane_d(mx, 16));
  const double tol = n * 2.3e-16 * mx;
  unsigned done = 0u;
  int k = 0;
  auto take = [](double& best, int& j, double ob, int oj) {
    if (ob > best || (ob == best && oj < j)) {
      best = ob;
      j = oj;
    }
  };
  for (; k < n; ++k) {
    const bool mine = row && !((done >> i) & 1u);
    double best = h == 0 && mine ? di : -1.0;
    int j = h == 0 && mine ? i : n;
    take(best, j, dpp_move<0xB1>(best), dpp_move_int<0xB1>(j));
    take(best, j, dpp_move<0x4E>(best), dpp_move_int<0x4E>(j));
    take(best, j, dpp_move<0x141>(best), dpp_move_int<0x141>(j));
    take(best, j, dpp_move<0x140>(best), dpp_move_int<0x140>(j));
    {
      const double b1 = readlane_d(best, 16);
      const int j1 = __builtin_amdgcn_readlane(j, 16);
      best = readlane_d(best, 0);
      j = __builtin_amdgcn_readlane(j, 0);
      take(best, j, b1, j1);
    }
    if (!(best > tol)) break;
    const double inv = rsqrt_nr(best);
    const double lkk = best * inv;
    // sum over the earlier columns, alternate m on the two halves, two chains
    // (lanes past the matrix, i >= n, read row 0's entry and discard it)
    const int ii = row ? i : 0;
    cd a0 = mk<double>(0.0, 0.0), a1 = a0;
    int m = h;
    // two steps of the loop below at a time, every load issued first (same
    // order of additions in each chain)
    for (; m + 6 < k; m += 8) {
      const int p0 = piv[m], p1 = piv[m + 2], p2 = piv[m + 4], p3 = piv[m + 6];
      const cd x0 = cld(G, p0 * n + ii), y0 = cld(G, p0 * n + j);
      const cd x1 = cld(G, p1 * n + ii), y1 = cld(G, p1 * n + j);
      const cd x2 = cld(G, p2 * n + ii), y2 = cld(G, p2 * n + j);
      const cd x3 = cld(G, p3 * n + ii), y3 = cld(G, p3 * n + j);
      fma_acc(a0, x0, conj(y0));
      fma_acc(a1, x1, conj(y1));
      fma_acc(a0, x2, conj(y2));
      fma_acc(a1, x3, conj(y3));
    }
    for (; m + 2 < k; m += 4) {
      const int p0 = piv[m], p1 = piv[m + 2];
      fma_acc(a0, cld(G, p0 * n + ii), conj(cld(G, p0 * n + j)));
      fma_acc(a1, cld(G, p1 * n + ii), conj(cld(G, p1 * n + j)));
    }
    if (m < k) {
      const int p0 = piv[m];
      fma_acc(a0, cld(G, p0 * n + ii), conj(cld(G, p0 * n + j)));
    }
    cd acc = a0 + a1;
    acc.re += __shfl_xor(acc.re, 32, 64);
    acc.im += __shfl_xor(acc.im, 32, 64);
    const cd gij = row ? conj(cld(G, j * n + i)) : mk<double>(0.0, 0.0);
    cd l = mk<double>(0.0, 0.0);
    if (i == j)
      l = mk<double>(lkk, 0.0);
    else if (mine)
      l = inv * (gij - acc);
    // every lane has read row j (above, in program order) before it is written
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (h == 0 && row) cst(G, j * n + i, l);
    if (lane == 0) piv[k] = j;
    if (mine && i != j) di -= norm2(l);
    done |= 1u << j;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) {
    int q = k;
    for (int c = 0; c < n; ++c)
      if (!((done >> c) & 1u)) piv[q++] = c;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int m = h; m < n; m += 2) {
    if (row) R[m * n + i] = m < k ? conj(cld(G, piv[m] * n + i)) : mk<double>(0.0, 0.0);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return k;
}

// Householder QR with column pivoting, C P = Q R, of an N x N matrix (N even,
// N <= 32, ROW-major in LDS, overwritten) on ONE wavefront, for the
// preconditioned one-sided Jacobi SVD (Drmac & Veselic, "New fast and accurate
// Jacobi SVD algorithm I", SIAM J. Matrix Anal. Appl. 29 (2008)): Jacobi sweeps
// on R^H converge in far fewer sweeps than on C.  Lane (c, h) = (lane % 32,
// lane / 32) works on rows h N/2 .. h N/2 + N/2 - 1 of column c, in place in
// LDS (kept out of registers: the caller runs near its register limit; the
// row-major layout puts the 32 column lanes on consecutive words).  Step
// k takes the remaining column of largest norm (ties to the smaller index;
// norms recomputed from the rows >= k every step), reflects it onto beta e_k
// with H = I - u u^H, |u|^2 = 2 (beta = -e^{i arg x_k} |x|, as householder_qr),
// and applies H to the other remaining columns.  Outputs: Hk (N x N
// column-major, column k = u_k, zero above row k), X = R^H (column-major,
// X[r N + m] = conj(R[r][m]), R upper triangular in pivot order) and piv
// (column m of C P is column piv[m] of C).  Q = H_0 H_1 ... H_{N-1}.
// The same pivoted QR (C P = Q R, same pivot rule, reflectors and outputs as
// wave_qrcp_rh) for N = 32 on the whole 256-thread block: thread (c, g) =
// (tid / 8, tid % 8) keeps rows 4g .. 4g + 3 of column c in registers for the
// whole factorisation, column norms and reflector products are sums over the 8
// lanes of a column (DPP moves, no cross-wave traffic), and a step costs one
// block barrier: before it every remaining column posts its rows to LDS and
// every wave its best remaining column; after it every thread forms the rows
// of the pivot's reflector it needs from the posted pivot column.  The posts
// alternate between two buffers, C and X (a thread posting for step k + 1
// has passed the barrier of step k, so every read of the buffer of step k - 1
// is done).  C (row-major) is read first and then used as one of the two
// buffers.  Call with all 256 threads.
template <typename P, typename I>
__device__ inline void block_qrcp_rh32(P* C, P* Hk, P* X, I* piv) {
  constexpr int N = 32;
  __shared__ double cand_v[2][4];
  __shared__ int cand_j[2][4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int c = t >> 3, g = t & 7, r0 = 4 * g;
  cd x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) x[q] = cld(C, (r0 + q) * N + c);
  __syncthreads();  // C is read before it is posted to
  unsigned done = 0u;
  int mypos = N;
  // (largest norm, ties to the smaller index) as selects, without branches
  auto take = [](double& best, int& j, double ob, int oj) {
    const bool better = (ob > best) | ((ob == best) & (oj < j));  // (no short-circuit branches)
    best = better ? ob : best;
    j = better ? oj : j;
  };
  for (int k = 0; k < N; ++k) {
    P* pb = (k & 1) ? X : C;  // posted columns, column-major
    const bool mine = !((done >> c) & 1u);
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) part += r0 + q >= k ? norm2(x[q]) : 0.0;
    const double nrm2 = group_sum<8>(part);
    if (mine) {
#pragma unroll
      for (int q = 0; q < 4; ++q) cst(pb, c * N + r0 + q, x[q]);
    }
    // argmax over the wave's 8 columns: the two columns of each 16-lane row
    // meet by a row mirror, then the rows by lane shuffles (xor 16, xor 32)
    double best = mine ? nrm2 : -1.0;
    int j = mine ? c : N;
    take(best, j, dpp_move<0x140>(best), dpp_move_int<0x140>(j));
    take(best, j, __shfl_xor(best, 16, 64), __shfl_xor(j, 16, 64));
    take(best, j, __shfl_xor(best, 32, 64), __shfl_xor(j, 32, 64));
    if (lane == 0) {
      cand_v[k & 1][w] = best;
      cand_j[k & 1][w] = j;
    }
    __syncthreads();
    {
      // the four waves' candidates, loaded together and reduced as a tree
      double b0 = cand_v[k & 1][0], b1 = cand_v[k & 1][1], b2 = cand_v[k & 1][2], b3 = cand_v[k & 1][3];
      int j0 = cand_j[k & 1][0], j1 = cand_j[k & 1][1], j2 = cand_j[k & 1][2], j3 = cand_j[k & 1][3];
      take(b0, j0, b1, j1);
      take(b2, j2, b3, j3);
      take(b0, j0, b2, j2);
      best = b0;
      j = j0;
    }
    const cd alpha = cld(pb, j * N + k);
    // beta = -e^{i arg alpha} |x|, scale = 1 / sqrt(|x| (|x| + |alpha|)):
    // reciprocal square roots refined by a Newton step (about an ulp, as in
    // the Jacobi rotations) instead of correctly rounded square roots and
    // divisions on the critical path of every step
    // (columns with |x|^2 <= 1e-300 count as zero: no reflector)
    const bool live = best > 1e-300;
    cd beta = mk<double>(0.0, 0.0);
    double scale = 0.0;
    if (live) {
      const double nrm = best * rsqrt_nr(best);
      const double a2 = norm2(alpha);
      double aa = 0.0;
      if (a2 > 1e-300) {
        const double ia = rsqrt_nr(a2);
        aa = a2 * ia;
        beta = mk<double>(-alpha.re * ia * nrm, -alpha.im * ia * nrm);
      } else {
        beta = mk<double>(-nrm, 0.0);
      }
      scale = rsqrt_nr(nrm * (nrm + aa));
    }
    // this thread's rows of the reflector u = scale (x_j - beta e_k), zero above k
    cd u[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + q;
      const cd xj = cld(pb, j * N + r);
      u[q] = r >= k ? scale * (r == k ? xj - beta : xj) : mk<double>(0.0, 0.0);
    }
    if (c == j) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = r0 + q;
        cst(Hk, k * N + r, u[q]);
        x[q] = r < k ? x[q] : (r == k ? beta : mk<double>(0.0, 0.0));
      }
      mypos = k;
    } else if (mine && live) {
      double dr = 0.0, di = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        dr += u[q].re * x[q].re + u[q].im * x[q].im;
        di += u[q].re * x[q].im - u[q].im * x[q].re;
      }
      // u is zero above row k, so every row takes the same update
      const cd dot = mk<double>(group_sum<8>(dr), group_sum<8>(di));
#pragma unroll
      for (int q = 0; q < 4; ++q) x[q] = x[q] - u[q] * dot;
    }
    done |= 1u << j;
    if (t == 0) piv[k] = j;
  }
  __syncthreads();  // every read of the last posted buffer (possibly X) is done
  // X = R^H: column r of X holds conj(R[r][m]) at row m, R[r][m] = row r of
  // column piv[m], zero below the diagonal of R
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = r0 + q;
    cst(X, r * N + mypos, r <= mypos ? conj(x[q]) : mk<double>(0.0, 0.0));
  }
  __syncthreads();
}

template <int N, typename P, typename I>
__device__ inline void wave_qrcp_rh(P* C, P* Hk, P* X, I* piv) {
  static_assert(N % 8 == 0 && N <= 32, "wave QR sizing");
  constexpr int HALF = N / 2;
  const int lane = threadIdx.x & 63;
  const int c = lane & 31, h = lane >> 5;
  const bool col = c < N;
  const int cc = col ? c : 0;  // lanes past the matrix read column 0 and discard it
  const int r0 = h * HALF;
  unsigned done = 0u;
  int mypos = N;
  auto take = [](double& best, int& j, double ob, int oj) {
    if (ob > best || (ob == best && oj < j)) {
      best = ob;
      j = oj;
    }
  };
#pragma unroll 1
  for (int k = 0; k < N; ++k) {
    const bool mine = col && !((done >> c) & 1u);
    // rows are loaded unconditionally and masked, so the loads issue back to
    // back instead of one branch (and one LDS round trip) per row
    // batches of 4 rows bound the registers held (the caller runs near its
    // limit)
    double s = 0.0;
#pragma unroll 1
    for (int b = 0; b < HALF; b += 4) {
#pragma unroll
      for (int t = b; t < b + 4; ++t) {
        const double v = norm2(cld(C, (r0 + t) * N + cc));
        s += r0 + t >= k ? v : 0.0;
      }
    }
    s += __shfl_xor(s, 32, 64);
    double best = h == 0 && mine ? s : -1.0;
    int j = h == 0 && mine ? c : N;
    take(best, j, dpp_move<0xB1>(best), dpp_move_int<0xB1>(j));
    take(best, j, dpp_move<0x4E>(best), dpp_move_int<0x4E>(j));
    take(best, j, dpp_move<0x141>(best), dpp_move_int<0x141>(j));
    take(best, j, dpp_move<0x140>(best), dpp_move_int<0x140>(j));
    {
      const double b1 = readlane_d(best, 16);
      const int j1 = __builtin_amdgcn_readlane(j, 16);
      best = readlane_d(best, 0);
      j = __builtin_amdgcn_readlane(j, 0);
      take(best, j, b1, j1);
    }
    const cd alpha = cld(C, k * N + j);
    const double nrm = best > 0.0 ? sqrt(best) : 0.0;
    const double aa = sqrt(norm2(alpha));
    cd beta = mk<double>(0.0, 0.0);
    double scale = 0.0;
    if (best > 0.0) {
      beta = aa > 0.0 ? mk<double>(-alpha.re / aa * nrm, -alpha.im / aa * nrm) : mk<double>(-nrm, 0.0);
      scale = sqrt(2.0) / sqrt(2.0 * nrm * (nrm + aa));
    }
    // every lane has read column j (norms, alpha) before it is rewritten
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (c == j) {
      // the reflector, and the pivot column becomes beta e_k from row k down
#pragma unroll 1
      for (int b = 0; b < HALF; b += 4) {
#pragma unroll
        for (int t = b; t < b + 4; ++t) {
          const int r = r0 + t;
          const cd x = cld(C, r * N + j);
          cst(Hk, k * N + r, r >= k ? scale * (r == k ? x - beta : x) : mk<double>(0.0, 0.0));
          cst(C, r * N + j, r < k ? x : (r == k ? beta : mk<double>(0.0, 0.0)));
        }
      }
      mypos = k;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (mine && c != j && best > 0.0) {
      double dr = 0.0, di = 0.0;
#pragma unroll 1
      for (int b = 0; b < HALF; b += 4) {
#pragma unroll
        for (int t = b; t < b + 4; ++t) {
          const cd u = cld(Hk, k * N + r0 + t), x = cld(C, (r0 + t) * N + c);
          dr += u.re * x.re + u.im * x.im;
          di += u.re * x.im - u.im * x.re;
        }
      }
      // lanes c and c + 32 hold the two halves of one column and take this
      // branch together
      dr += __shfl_xor(dr, 32, 64);
      di += __shfl_xor(di, 32, 64);
      const cd dot = mk<double>(dr, di);
      // u is zero above row k, so every row takes the same update
#pragma unroll 1
      for (int b = 0; b < HALF; b += 4) {
#pragma unroll
        for (int t = b; t < b + 4; ++t) {
          const int r = r0 + t;
          cst(C, r * N + c, cld(C, r * N + c) - cld(Hk, k * N + r) * dot);
        }
      }
    }
    done |= 1u << j;
    if (lane == 0) piv[k] = j;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // X = R^H: column r of X holds conj(R[r][m]) at row m, R[r][m] = row r of
  // column piv[m], zero below the diagonal of R
  if (col) {
#pragma unroll
    for (int t = 0; t < HALF; ++t) {
      const int r = r0 + t;
      cst(X, r * N + mypos, r <= mypos ? conj(cld(C, r * N + c)) : mk<double>(0.0, 0.0));
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ inline void pivoted_cholesky_swap(cd* G, int n, cd* R, int* piv, double* scal) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) piv[i] = i;
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) R[i] = mk<double>(0.0, 0.0);
  if (threadIdx.x == 0) {
    double mx = 0.0;
    for (int i = 0; i < n; ++i) mx = fmax(mx, G[i * n + i].re);
    scal[0] = n * 2.3e-16 * mx;
  }
  __syncthreads();
  const double tol = scal[0];
  int k = 0;
  for (; k < n; ++k) {
    if (threadIdx.x < 64) {
      // wave-0 argmax of the remaining diagonal (ties -> smallest index)
      double best = -1.0;
      int j = n;
      for (int i = k + (int)threadIdx.x; i < n; i += 64) {
        const double v = G[i * n + i].re;
        if (v > best) {
          best = v;
          j = i;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int oj = __shfl_xor(j, o, 64);
        if (ob > best || (ob == best && oj < j)) {
          best = ob;
          j = oj;
        }
      }
      if (threadIdx.x == 0) {
        scal[1] = best;
        *(int*)(scal + 2) = j;
      }
    }
    __syncthreads();
    const double best = scal[1];
    const int j = *(int*)(scal + 2);
    if (!(best > tol)) break;
    if (j != k) {
      // symmetric swap of rows/columns k and j (columns first, then rows)
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const cd t = G[i * n + k];
        G[i * n + k] = G[i * n + j];
        G[i * n + j] = t;
      }
      __syncthreads();
      for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const cd t = G[k * n + c];
        G[k * n + c] = G[j * n + c];
        G[j * n + c] = t;
      }
      if (threadIdx.x == 0) {
        const int t = piv[k];
        piv[k] = piv[j];
        piv[j] = t;
      }
      __syncthreads();
    }
    const double lkk = sqrt(best);
    // column k of L below the diagonal, kept in G[i][k]
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) G[i * n + k] = (1.0 / lkk) * G[i * n + k];
    if (threadIdx.x == 0) G[k * n + k] = mk<double>(lkk, 0.0);
    __syncthreads();
    const int rem = n - k - 1;
    for (int f = threadIdx.x; f < rem * rem; f += blockDim.x) {
      const int i = k + 1 + f / rem, c = k + 1 + f % rem;
      G[i * n + c] = G[i * n + c] - G[i * n + k] * conj(G[c * n + k]);
    }
    __syncthreads();
  }
  // R[i][piv[m]] = conj(L[m][i]) for m >= i, rows i < k
  for (int f = threadIdx.x; f < k * n; f += blockDim.x) {
    const int i = f / n, m = f % n;
    if (m >= i) R[i * n + piv[m]] = conj(G[m * n + i]);
  }
  __syncthreads();
}

// Householder QR in place (column-major A, lda).  beta[j] = R_jj; u_j stored
// in A[j.., j] (|u_j|^2 = 2, H_j = I - u_j u_j^H, Hermitian and unitary);
// strict upper triangle = R.  red: blockDim/64 doubles.
__device__ inline void householder_qr(cd* A, int lda, int m, int n, cd* beta, double* red) {
  const int k = m < n ? m : n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
  for (int j = 0; j < k; ++j) {
    cd* col = A + (long long)j * lda;
    double s = 0.0;
    for (int i = j + threadIdx.x; i < m; i += blockDim.x) s += norm2(col[i]);
    const double nrm2 = block_sum(s, red);
    const cd alpha = col[j];
    const double nrm = sqrt(nrm2);
    const double aa = sqrt(norm2(alpha));
    cd b;
    double scale = 0.0;
    if (nrm2 > 0.0) {
      // beta = -e^{i arg alpha} |x|;  u = x - beta e_j;  |u|^2 = 2 |x| (|x| + |alpha|)
      if (aa > 0.0)
        b = mk<double>(-alpha.re / aa * nrm, -alpha.im / aa * nrm);
      else
        b = mk<double>(-nrm, 0.0);
      scale = sqrt(2.0) / sqrt(2.0 * nrm * (nrm + aa));
    } else {
      b = mk<double>(0.0, 0.0);
    }
    __syncthreads();
    if (threadIdx.x == 0) beta[j] = b;
    for (int i = j + threadIdx.x; i < m; i += blockDim.x) {
      cd x = col[i];
      if (i == j) x = x - b;
      col[i] = scale * x;
    }
    __syncthreads();
    if (nrm2 > 0.0) {
      for (int c = j + 1 + w; c < n; c += nw) {
        cd* y = A + (long long)c * lda;
        double sr = 0.0, si = 0.0;
        for (int i = j + lane; i < m; i += 64) {
          const cd u = col[i], v = y[i];
          sr += u.re * v.re + u.im * v.im;
          si += u.re * v.im - u.im * v.re;
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        const cd dot = mk<double>(sr, si);
        for (int i = j + lane; i < m; i += 64) y[i] = y[i] - col[i] * dot;
      }
    }
    __syncthreads();
  }
}

// Y (m x c column-major, ldy) <- H_0 H_1 ... H_{k-1} Y.
__device__ inline void householder_apply_q(const cd* A, int lda, int m, int k, cd* Y, int ldy, int c) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
  for (int j = k - 1; j >= 0; --j) {
    const cd* col = A + (long long)j * lda;
    for (int cc = w; cc < c; cc += nw) {
      cd* y = Y + (long long)cc * ldy;
      double sr = 0.0, si = 0.0;
      for (int i = j + lane; i < m; i += 64) {
        const cd u = col[i], v = y[i];
        sr += u.re * v.re + u.im * v.im;
        si += u.re * v.im - u.im * v.re;
      }
      sr = wave_sum(sr);
      si = wave_sum(si);
      const cd dot = mk<double>(sr, si);
      for (int i = j + lane; i < m; i += 64) y[i] = y[i] - col[i] * dot;
    }
    __syncthreads();
  }
}

}  // namespace dq

// Hand-written adjoints (vector-Jacobian products) of the forward kernels,
// complex128.  Gradient convention: for a complex value z the gradient is
// zbar = dL/dRe z + i dL/dIm z, so dL = Re sum conj(zbar) dz; real values
// carry dL/dx.
//
// Simple update (forward: dq_su.hip, reference tensor_ops.py:318-389).  With
// Atil = apply_half(Gamma) (rows = other bonds x phys out, columns = (b, q)),
// D = prod of the other lambdas per row, nu^2 = |D Atil|^2,
// G = (D Atil)^H (D Atil) / nu^2 = R^H R (pivoted Cholesky), the core
// C = R_a Lambda R_b^T = U S V^H, and K_a = Lambda R_b^T V_k S_k^-1,
// K_b = Lambda R_a^T conj(U_k) S_k^-1, the outputs are
//   Gamma_a' = mask_a (Atil_a / nu_a) K_a,  Gamma_b' = mask_b (Atil_b / nu_b) K_b,
//   lambda' = S_k.
// This equals the reference's QR + SVD formulation (Gamma_a' = Q_a U_k /
// lambda_others); the adjoint below differentiates this closed form, reading
// R_a, R_b, U, V, S and the pivots from the forward's tape so that it sees
// the forward's singular-vector phases.
#include <float.h>

#include "dq_fast64.hpp"
#include "dq_launch.hpp"
#include "dq_linalg.hpp"

namespace dq {

namespace adj {

// op codes of the small dense helpers
constexpr int N = 0, T = 1, H = 2, C = 3;

__device__ __forceinline__ cd opget(const cd* A, int lda, int op, int i, int j) {
  switch (op) {
    case T:
      return A[(long long)j * lda + i];
    case H:
      return conj(A[(long long)j * lda + i]);
    case C:
      return conj(A[(long long)i * lda + j]);
    default:
      return A[(long long)i * lda + j];
  }
}

typedef double d4 __attribute__((ext_vector_type(4)));

// The same product on FP64 matrix cores (V_MFMA_F64_16X16X4F64: lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16]; D[l / 16 + 4 v][l % 16]) for m, n
// multiples of 16 and k of 4: one 16 x 16 output tile per wave at a time, 4
// real MFMAs per complex k-step.  With fewer tiles than waves the k range is
// split over the waves and the partial tiles summed in a fixed order through
// LDS (at most four waves take part in a split).
__device__ void bmm_mfma(cd* Cm, int ldc, const cd* A, int lda, int opA, const cd* B, int ldb, int opB, int m, int n,
                         int k, bool accumulate) {
  __shared__ cd part[4][256];
  const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tn = n >> 4, tiles = (m >> 4) * tn;
  int split = 1;
  while (tiles * split * 2 <= (nw < 4 ? nw : 4) && (k % (split * 2 * 4)) == 0) split *= 2;
  const int items = tiles * split, kc = k / split;
  const int r = lane & 15, kq = lane >> 4;
  for (int it = wave; it < items; it += nw) {
    const int tile = it / split, ks = it % split;
    const int i0 = (tile / tn) * 16, j0 = (tile % tn) * 16;
    d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = cr;
    const int le = (ks + 1) * kc;
    int l0 = ks * kc;
    // four k-steps per group with all their operands loaded first, so that
    // the operand latency (the operands often live in global memory) is paid
    // once per group; the accumulation order is the plain k order
    for (; l0 + 16 <= le; l0 += 16) {
      cd a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = opget(A, lda, opA, i0 + r, l0 + 4 * u + kq);
        b[u] = opget(B, ldb, opB, l0 + 4 * u + kq, j0 + r);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].re, b[u].re, cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[u].im, b[u].im, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].re, b[u].im, ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].im, b[u].re, ci, 0, 0, 0);
      }
    }
    for (; l0 < le; l0 += 4) {
      const cd a = opget(A, lda, opA, i0 + r, l0 + kq);
      const cd b = opget(B, ldb, opB, l0 + kq, j0 + r);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, b.im, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, ci, 0, 0, 0);
    }
    if (split == 1) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        cd* c = Cm + (long long)(i0 + kq + 4 * v) * ldc + j0 + r;
        const cd val = mk<double>(cr[v], ci[v]);
        *c = accumulate ? *c + val : val;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) part[it][(kq + 4 * v) * 16 + r] = mk<double>(cr[v], ci[v]);
    }
  }
  if (split > 1) {
    __syncthreads();
    for (int q = threadIdx.x; q < tiles * 256; q += blockDim.x) {
      const int tile = q >> 8, e = q & 255;
      cd sum = part[tile * split][e];
      for (int ks = 1; ks < split; ++ks) sum = sum + part[tile * split + ks][e];
      cd* c = Cm + (long long)((tile / tn) * 16 + (e >> 4)) * ldc + (tile % tn) * 16 + (e & 15);
      *c = accumulate ? *c + sum : sum;
    }
  }
  __syncthreads();
}

// Cm (m x n, row-major, ldc) (+)= op(A) (m x k) op(B) (k x n).  Block-wide;
// ends with a barrier.  Cm must not alias A or B.
__device__ void bmm(cd* Cm, int ldc, const cd* A, int lda, int opA, const cd* B, int ldb, int opB, int m, int n, int k,
                    bool accumulate) {
  if ((m & 15) == 0 && (n & 15) == 0 && (k & 3) == 0 && k > 0 && (blockDim.x & 63) == 0) {
    bmm_mfma(Cm, ldc, A, lda, opA, B, ldb, opB, m, n, k, accumulate);
    return;
  }
  for (int f = threadIdx.x; f < m * n; f += blockDim.x) {
    const int i = f / n, j = f % n;
    cd acc = accumulate ? Cm[(long long)i * ldc + j] : mk<double>(0.0, 0.0);
    for (int l = 0; l < k; ++l) fma_acc(acc, opget(A, lda, opA, i, l), opget(B, ldb, opB, l, j));
    Cm[(long long)i * ldc + j] = acc;
  }
  __syncthreads();
}

// X (k x ncols, ldx) <- R^-1 X for upper-triangular R (k x k, ldr).  One
// thread per column (back substitution).
__device__ void solve_upper_left(const cd* R, int ldr, int k, cd* X, int ldx, int ncols) {
  for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
    for (int i = k - 1; i >= 0; --i) {
      cd acc = X[(long long)i * ldx + c];
      for (int m = i + 1; m < k; ++m) acc = acc - R[(long long)i * ldr + m] * X[(long long)m * ldx + c];
      X[(long long)i * ldx + c] = cdiv(acc, R[(long long)i * ldr + i]);
    }
  }
  __syncthreads();
}

// X (nrows x k, ldx) <- X R^-H for upper-triangular R: solve Y R^H = X row by
// row, (Y R^H)_c = sum_{m >= c} Y_m conj(R[c][m]).
__device__ void solve_upperH_right(const cd* R, int ldr, int k, cd* X, int ldx, int nrows) {
  for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
    cd* x = X + (long long)r * ldx;
    for (int c = k - 1; c >= 0; --c) {
      cd acc = x[c];
      for (int m = c + 1; m < k; ++m) acc = acc - x[m] * conj(R[(long long)c * ldr + m]);
      x[c] = cdiv(acc, conj(R[(long long)c * ldr + c]));
    }
  }
  __syncthreads();
}

// SVD adjoint of a square n x n C = U diag(s) V^H (U, V row-major, all n
// columns; s descending, zeros allowed).  Given Ubar, Vbar (n x n) and sbar,
// writes Cbar = U Abar V^H with
//   Abar_ij = F_ij (s_j (J - J^H)_ij + s_i (K - K^H)_ij),  i != j,
//   Abar_ii = sbar_i + i (Im J_ii - Im K_ii) / (2 s_i),
// J = U^H Ubar, K = V^H Vbar, F_ij = 1 / (s_j^2 - s_i^2) (Lorentzian-broadened
// at degeneracies, zero where both values vanish), plus the null-space term
// for the zero columns U carries where s = 0.  w1..w3: n x n scratch.
__device__ void svd_vjp(const cd* U, const cd* V, const double* s, const cd* Ubar, const cd* Vbar,
                        const double* sbar, int n, cd* Cbar, cd* w1, cd* w2, cd* w3) {
  bmm(w1, n, U, n, H, Ubar, n, N, n, n, n, false);  // J
  bmm(w2, n, V, n, H, Vbar, n, N, n, n, n, false);  // K
  double smax = 0.0;
  for (int i = 0; i < n; ++i) smax = fmax(smax, s[i]);
  const double broad = 1e-24 * smax * smax * smax * smax + 1e-300;
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int i = f / n, j = f % n;
    cd a = mk<double>(0.0, 0.0);
    if (i != j) {
      const double d = s[j] * s[j] - s[i] * s[i];
      if (s[i] > 0.0 || s[j] > 0.0) {
        const double F = d / (d * d + broad);
        const cd jj = w1[f] - conj(w1[(long long)j * n + i]);
        const cd kk = w2[f] - conj(w2[(long long)j * n + i]);
        a = F * (s[j] * jj + s[i] * kk);
      }
    } else {
      a = mk<double>(sbar[i], 0.0);
      if (s[i] > 0.0) a.im = (w1[f].im - w2[f].im) / (2.0 * s[i]);
    }
    w3[f] = a;
  }
  __syncthreads();
  bmm(w1, n, U, n, N, w3, n, N, n, n, n, false);       // U Abar
  bmm(Cbar, n, w1, n, N, V, n, H, n, n, n, false);     // (U Abar) V^H
  // U holds zero columns where s = 0 (no orthonormal completion), so the
  // null-space part of the kept left vectors' adjoint is added explicitly:
  // Cbar += (I - U U^H) Ubar S^-1 V^H  (V is a complete basis).
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int j = f % n;
    w1[f] = s[j] > 0.0 ? (1.0 / s[j]) * Ubar[f] : mk<double>(0.0, 0.0);
  }
  __syncthreads();
  bmm(w2, n, U, n, H, w1, n, N, n, n, n, false);       // U^H X
  bmm(w3, n, U, n, N, w2, n, N, n, n, n, false);       // U U^H X
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) w1[f] = w1[f] - w3[f];
  __syncthreads();
  bmm(Cbar, n, w1, n, N, V, n, H, n, n, n, true);
}

// Adjoint of the pivoted Cholesky factor R (n x n row-major, R = L^H P^T,
// rows >= kc zero) of a Hermitian positive semi-definite G of rank kc:
// Gbar (Hermitian) from Rbar.  In pivoted coordinates Rt = R P = [R11 R12; 0 0]
// with G11 = R11^H R11, G12 = R11^H R12:
//   M11' = triu(M11 - R12 M12^H R11^-H),  B = strict_upper(M11' R11^H) + diag(Re)/2,
//   G11bar = R11^-1 herm(B) R11^-H,  G12bar = R11^-1 M12 (split over both blocks).
// w1..w4: n x n scratch (w1 receives Rt, w2 the permuted Rbar).
__device__ void chol_vjp(const cd* R, const cd* Rbar, const int* piv, int n, cd* Gbar, cd* w1, cd* w2, cd* w3,
                         cd* w4, int* kc_sh, double cut) {
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int i = f / n, m = f % n;
    w1[f] = R[(long long)i * n + piv[m]];
    w2[f] = m >= i ? Rbar[(long long)i * n + piv[m]] : mk<double>(0.0, 0.0);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // pivots (diagonal of R, descending) below cut * the first are treated
    // as rank deficiency: the adjoint through them divides by them
    const double p0 = w1[0].re;
    int kc = 0;
    while (kc < n && w1[(long long)kc * n + kc].re > cut * p0) ++kc;
    *kc_sh = kc;
  }
  __syncthreads();
  const int kc = *kc_sh;
  const int nr = n - kc;
  // w3 = M11 (kc x kc, ld n), then M11' = M11 - (R12 M12^H) R11^-H
  for (int f = threadIdx.x; f < kc * kc; f += blockDim.x) {
    const int i = f / kc, m = f % kc;
    w3[(long long)i * n + m] = w2[(long long)i * n + m];
  }
  __syncthreads();
  if (nr > 0) {
    bmm(w4, n, w1 + kc, n, N, w2 + kc, n, H, kc, kc, nr, false);  // R12 M12^H
    solve_upperH_right(w1, n, kc, w4, n, kc);                      // (R12 M12^H) R11^-H
    for (int f = threadIdx.x; f < kc * kc; f += blockDim.x) {
      const int i = f / kc, m = f % kc;
      w3[(long long)i * n + m] = w3[(long long)i * n + m] - w4[(long long)i * n + m];
    }
    __syncthreads();
  }
  for (int f = threadIdx.x; f < kc * kc; f += blockDim.x) {  // triu
    const int i = f / kc, m = f % kc;
    if (m < i) w3[(long long)i * n + m] = mk<double>(0.0, 0.0);
  }
  __syncthreads();
  bmm(w4, n, w3, n, N, w1, n, H, kc, kc, kc, false);  // A = M11' R11^H
  // herm(strict_upper(A) + diag(Re A)/2) into w3
  for (int f = threadIdx.x; f < kc * kc; f += blockDim.x) {
    const int i = f / kc, m = f % kc;
    cd v;
    if (i == m)
      v = mk<double>(0.5 * w4[(long long)i * n + i].re, 0.0);
    else if (m > i)
      v = 0.5 * w4[(long long)i * n + m];
    else
      v = 0.5 * conj(w4[(long long)m * n + i]);
    w3[(long long)i * n + m] = v;
  }
  __syncthreads();
  solve_upper_left(w1, n, kc, w3, n, kc);     // R11^-1 herm
  solve_upperH_right(w1, n, kc, w3, n, kc);   // ... R11^-H
  // G12bar = R11^-1 M12 -> w2 columns kc.. (rows < kc)
  if (nr > 0) solve_upper_left(w1, n, kc, w2 + kc, n, nr);
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int i = f / n, m = f % n;
    cd v = mk<double>(0.0, 0.0);
    if (i < kc && m < kc)
      v = w3[(long long)i * n + m];
    else if (i < kc && m >= kc)
      v = 0.5 * w2[(long long)i * n + m];
    else if (i >= kc && m < kc)
      v = 0.5 * conj(w2[(long long)m * n + i]);
    Gbar[(long long)piv[i] * n + piv[m]] = v;
  }
  __syncthreads();
}

// dst[.., a@axis, ..] (+)= sum_b op(M)[a][b] src[.., b@axis, ..] for a tensor
// with d bonds of size chi and a trailing phys axis; M chi x chi row-major,
// op 0: M[a][b], 1: conj(M[a][b]), 2: M[b][a], 3: conj(M[b][a]).  dst must
// not alias src.  Ends with a barrier.
__device__ void mprod(cd* dst, const cd* src, const cd* M, int chi, int d, int phys, int axis, int op, bool accumulate) {
  const long long inner = ipow(chi, d - 1 - axis) * phys;
  const long long total = ipow(chi, d) * phys;
  for (long long f = threadIdx.x; f < total; f += blockDim.x) {
    const long long in = f % inner;
    const long long t = f / inner;
    const int a = (int)(t % chi);
    const long long o = t / chi;
    const cd* sp = src + o * chi * inner + in;
    cd v = accumulate ? dst[f] : mk<double>(0.0, 0.0);
    for (int b = 0; b < chi; ++b) {
      cd m;
      switch (op) {
        case 1:
          m = conj(M[(long long)a * chi + b]);
          break;
        case 2:
          m = M[(long long)b * chi + a];
          break;
        case 3:
          m = conj(M[(long long)b * chi + a]);
          break;
        default:
          m = M[(long long)a * chi + b];
      }
      fma_acc(v, m, sp[(long long)b * inner]);
    }
    dst[f] = v;
  }
  __syncthreads();
}

// Normalisation adjoint for O = B / tr(B) (complex trace):
// Bbar = Obar / conj(t) - conj(<Obar, B> / t^2) I.  n x n, block-wide.
__device__ void trace_norm_vjp(const cd* B, const cd* Ob, int n, cd* Bb) {
  cd tr = mk<double>(0.0, 0.0), ip = tr;
  for (int k = 0; k < n; ++k) tr = tr + B[(long long)k * n + k];
  for (int q = 0; q < n * n; ++q) fma_cacc(ip, Ob[q], B[q]);
  const cd t2 = tr * tr;
  const cd c = conj(cdiv(ip, t2));
  const cd ct = conj(tr);
  for (int q = threadIdx.x; q < n * n; q += blockDim.x) {
    cd v = cdiv(Ob[q], ct);
    if (q / n == q % n) v = v - c;
    Bb[q] = v;
  }
  __syncthreads();
}

}  // namespace adj

// Per-gate workspace layout (complex doubles unless noted).
struct SuVjpWs {
  cd* base;
  int m_max, n, chi;
  __device__ cd* Aa() const { return base; }
  __device__ cd* Ab() const { return Aa() + (long long)m_max * n; }
  __device__ cd* Ba() const { return Ab() + (long long)m_max * n; }  // Atil gradients
  __device__ cd* Bb() const { return Ba() + (long long)m_max * n; }
  __device__ cd* AE() const { return Bb() + (long long)m_max * n; }  // Atil Ebar
  __device__ cd* Ya() const { return AE() + (long long)m_max * n; }
  __device__ cd* Yb() const { return Ya() + (long long)m_max * chi; }
  __device__ cd* core(int k) const { return Yb() + (long long)m_max * chi + (long long)k * n * n; }  // 16 slots
  __device__ double* dbl() const { return (double*)core(16); }
};

__host__ __device__ inline long long su_vjp_ws_bytes(int chi, int phys, int max_r, int max_deg) {
  const long long n = (long long)chi * max_r;
  const long long m = ipow(chi, max_deg - 1) * phys;
  const long long cds = 5 * m * n + 2 * m * chi + 16 * n * n;
  const long long dbls = 8 * n + 4 * ipow(chi, max_deg - 1) + 64;
  return (cds * 16 + dbls * 8 + 255) / 256 * 256;
}

__device__ __forceinline__ long long elem_off(const int* str, int d, int slot, int chi, int row, int b, int s) {
  long long off = (long long)b * str[slot] + s;
  for (int j = d - 1; j >= 0; --j) {
    if (j == slot) continue;
    const int q = row / chi, bj = row - q * chi;
    row = q;
    off += (long long)bj * str[j];
  }
  return off;
}

// Other-bond weight of a row: product of lambda (or of lam_mask when mask)
// over the bonds j != slot, optionally leaving out bond `skip`.
__device__ __forceinline__ double row_weight(const double* wl, int d, int slot, int chi, int row, int skip,
                                             bool mask, double eps) {
  double w = 1.0;
  for (int j = d - 1; j >= 0; --j) {
    if (j == slot) continue;
    const int q = row / chi, bj = row - q * chi;
    row = q;
    if (j == skip) continue;
    const double l = wl[j * chi + bj];
    w *= mask ? (l < eps ? 0.0 : 1.0) : l;
  }
  return w;
}

__device__ __forceinline__ int row_bond(int d, int slot, int chi, int row, int j) {
  for (int jj = d - 1; jj >= 0; --jj) {
    if (jj == slot) continue;
    const int q = row / chi, bj = row - q * chi;
    if (jj == j) return bj;
    row = q;
  }
  return 0;
}

__global__ void __launch_bounds__(256, 2)
    k_su_vjp(DevGraph g, const cd* Tpre, const double* lam, const SuGate* gates, const cd* halves, int max_r,
             double eps, double cut, const char* tape, long long tape_bytes, cd* Tbar, double* lambar, cd* hbar,
             char* ws, long long wsb) {
  __shared__ double wl[2][kMaxDeg * 64];
  __shared__ double sbar[128], lamb[128], le[64], sgc[128];
  __shared__ double red[8];
  __shared__ double scal[8];
  __shared__ int kc_sh;
  SuGate G;
  {
    const int* gi = (const int*)(gates + blockIdx.x);
    G.edge = ldc(gi, 0);
    G.first = ldc(gi, 1);
    G.rank = ldc(gi, 2);
    G.pad = 0;
  }
  if (G.rank < 1 || G.rank > max_r) return;
  const int e = G.edge, rg = G.rank, chi = g.chi, p = g.phys;
  const int n = chi * rg;
  const int nd[2] = {g.edge_node[2 * e + G.first], g.edge_node[2 * e + 1 - G.first]};
  const int sl[2] = {find_slot(g, nd[0], e, G.first), find_slot(g, nd[1], e, 1 - G.first)};
  const long long hstride = (long long)p * p * p * p;
  const cd* hh[2] = {halves + (long long)blockIdx.x * 2 * hstride, halves + (long long)blockIdx.x * 2 * hstride + hstride};
  const SuTape tp((char*)tape + (long long)blockIdx.x * tape_bytes, chi * max_r);
  const int kr = tp.meta()[0];
  // singular values below cut * the largest are treated as zero by the adjoint
  {
    double ms = 0.0;
    for (int i = 0; i < n; ++i) ms = fmax(ms, tp.sigma()[i]);
    for (int j = threadIdx.x; j < n; j += blockDim.x) sgc[j] = tp.sigma()[j] > cut * ms ? tp.sigma()[j] : 0.0;
  }
  __syncthreads();
  const double* sg = sgc;
  const cd* Ra = tp.Ra();
  const cd* Rb = tp.Rb();
  const cd* U = tp.U();
  const cd* V = tp.V();
  const double nu[2] = {sqrt(tp.tr()[0]), sqrt(tp.tr()[1])};
  int maxdeg = g.deg[nd[0]] > g.deg[nd[1]] ? g.deg[nd[0]] : g.deg[nd[1]];
  SuVjpWs W{(cd*)(ws + (long long)blockIdx.x * wsb), (int)(ipow(chi, maxdeg - 1) * p), n, chi};
  for (int side = 0; side < 2; ++side) {
    const int d = g.deg[nd[side]], base = g.adj_ptr[nd[side]];
    for (int f = threadIdx.x; f < d * chi; f += blockDim.x)
      wl[side][f] = lam[(long long)g.adj_edge[base + f / chi] * chi + f % chi];
  }
  for (int j = threadIdx.x; j < chi; j += blockDim.x) le[j] = lam[(long long)e * chi + j];
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    sbar[j] = (j < kr && tp.sigma()[j] > 0.0) ? lambar[(long long)e * chi + j] : 0.0;
    lamb[j] = 0.0;
  }
  __syncthreads();

  cd* Ka = W.core(0);
  cd* Kb = W.core(1);
  cd* Kba = W.core(2);
  cd* Kbb = W.core(3);
  cd* Rab = W.core(4);
  cd* Rbb = W.core(5);
  cd* Ub = W.core(6);
  cd* Vb = W.core(7);
  cd* Cb = W.core(8);
  cd* w1 = W.core(9);
  cd* w2 = W.core(10);
  cd* w3 = W.core(11);
  cd* w4 = W.core(12);
  cd* Gba = W.core(13);
  cd* Gbb = W.core(14);
  // K_a = Lambda R_b^T V_k S^-1,  K_b = Lambda R_a^T conj(U_k) S^-1  (n x chi, columns >= kr zero)
  for (int f = threadIdx.x; f < n * chi; f += blockDim.x) {
    const int c = f / chi, j = f % chi;
    cd a = mk<double>(0.0, 0.0), b = a;
    if (j < kr && sg[j] > 0.0) {
      for (int i = 0; i < n; ++i) {
        fma_acc(a, Rb[(long long)i * n + c], V[(long long)i * n + j]);
        fma_acc(b, Ra[(long long)i * n + c], conj(U[(long long)i * n + j]));
      }
      const double f2 = le[c / rg] / sg[j];
      a = f2 * a;
      b = f2 * b;
    }
    Ka[f] = a;
    Kb[f] = b;
  }
  __syncthreads();

  // ---------------- per side: Atil, masked output gradient, K-bar, write-path Atil gradient
  double nubar[2];
  for (int side = 0; side < 2; ++side) {
    const int nod = nd[side], slot = sl[side], d = g.deg[nod];
    const int rows = (int)ipow(chi, d - 1), m = rows * p;
    int str[kMaxDeg];
    {
      int mm = p;
      for (int j = d - 1; j >= 0; --j) {
        str[j] = mm;
        mm *= chi;
      }
    }
    const cd* t = Tpre + g.toff[nod];
    const cd* tb = Tbar + g.toff[nod];
    const cd* h = hh[side];
    cd* A = side == 0 ? W.Aa() : W.Ab();
    cd* Y = side == 0 ? W.Ya() : W.Yb();
    cd* B = side == 0 ? W.Ba() : W.Bb();
    const cd* K = side == 0 ? Ka : Kb;
    cd* Kbar = side == 0 ? Kba : Kbb;
    const double* w = wl[side];
    // the gathers below are batched kU elements per thread, loads first, so
    // that the global-memory latency is paid once per batch
    constexpr int kU = 8;
    if (p == 2) {
      constexpr int kA = 4;
      for (int f0 = threadIdx.x; f0 < m * n; f0 += kA * blockDim.x) {
        cd tv[kA][2], hv[kA][2];
#pragma unroll
        for (int u = 0; u < kA; ++u) {
          const int f = f0 + u * blockDim.x;
          if (f < m * n) {
            const int R = f / n, c = f - (f / n) * n;
            const int r = R >> 1, o = R & 1, b = c / rg, q = c - (c / rg) * rg;
            const long long to = elem_off(str, d, slot, chi, r, b, 0);
            tv[u][0] = t[to];
            tv[u][1] = t[to + 1];
            hv[u][0] = h[(q * 2 + o) * 2];
            hv[u][1] = h[(q * 2 + o) * 2 + 1];
          }
        }
#pragma unroll
        for (int u = 0; u < kA; ++u) {
          const int f = f0 + u * blockDim.x;
          if (f < m * n) A[f] = hv[u][0] * tv[u][0] + hv[u][1] * tv[u][1];
        }
      }
    } else {
      for (int f = threadIdx.x; f < m * n; f += blockDim.x) {
        const int R = f / n, c = f % n;
        const int r = R / p, o = R % p, b = c / rg, q = c % rg;
        cd acc = mk<double>(0.0, 0.0);
        for (int s = 0; s < p; ++s) fma_acc(acc, h[((long long)q * p + o) * p + s], t[elem_off(str, d, slot, chi, r, b, s)]);
        A[f] = acc;
      }
    }
    // Y = masked output gradient / nu (the 1 / nu of Kbar and B folded in)
    const double inv_nu = 1.0 / nu[side];
    for (int f0 = threadIdx.x; f0 < m * chi; f0 += kU * blockDim.x) {
      cd yv[kU];
      double wv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int f = f0 + u * blockDim.x;
        yv[u] = mk<double>(0.0, 0.0);
        wv[u] = 0.0;
        if (f < m * chi) {
          const int R = f / chi, j = f - (f / chi) * chi;
          const int r = R / p, o = R - (R / p) * p;
          if (j < kr) {
            wv[u] = inv_nu * row_weight(w, d, slot, chi, r, -1, true, eps);
            yv[u] = tb[elem_off(str, d, slot, chi, r, j, o)];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int f = f0 + u * blockDim.x;
        if (f < m * chi) Y[f] = wv[u] * yv[u];
      }
    }
    __syncthreads();
    // Kbar = Atil^H Y  (n x chi)
    adj::bmm(Kbar, chi, A, n, adj::H, Y, chi, adj::N, n, chi, m, false);
    // nubar = -Re<Kbar, K> / nu
    double acc = 0.0;
    for (int f = threadIdx.x; f < n * chi; f += blockDim.x)
      acc += Kbar[f].re * K[f].re + Kbar[f].im * K[f].im;
    acc = block_sum(acc, red);
    nubar[side] = -acc / nu[side];
    // write-path gradient of Atil: B = Y K^H
    adj::bmm(B, n, Y, chi, adj::N, K, chi, adj::H, m, n, chi, false);
  }

  // ---------------- core adjoint
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    Rab[f] = mk<double>(0.0, 0.0);
    Rbb[f] = Rab[f];
    Ub[f] = Rab[f];
    Vb[f] = Rab[f];
  }
  __syncthreads();
  for (int side = 0; side < 2; ++side) {
    // side 0: K_a = Lambda R_b^T Wk, Wk = V_k S^-1 ; side 1: K_b = Lambda R_a^T Wk, Wk = conj(U_k) S^-1
    const cd* Rx = side == 0 ? Rb : Ra;
    const cd* Kbar = side == 0 ? Kba : Kbb;
    cd* Rxbar = side == 0 ? Rbb : Rab;
    // w1 = Wk (n x chi)
    for (int f = threadIdx.x; f < n * chi; f += blockDim.x) {
      const int i = f / chi, j = f % chi;
      cd v = mk<double>(0.0, 0.0);
      if (j < kr && sg[j] > 0.0) {
        v = side == 0 ? V[(long long)i * n + j] : conj(U[(long long)i * n + j]);
        v = (1.0 / sg[j]) * v;
      }
      w1[f] = v;
    }
    __syncthreads();
    adj::bmm(w2, chi, Rx, n, adj::T, w1, chi, adj::N, n, chi, n, false);  // Rx^T Wk
    // lambda-bar (per fused index c), R_x-bar, Wk-bar
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      double a = 0.0;
      for (int j = 0; j < chi; ++j) {
        const cd kb = Kbar[(long long)c * chi + j], t2 = w2[(long long)c * chi + j];
        a += kb.re * t2.re + kb.im * t2.im;
      }
      lamb[c] += a;
    }
    for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
      const int i = f / n, c = f % n;
      cd a = mk<double>(0.0, 0.0);
      for (int j = 0; j < chi; ++j) fma_acc(a, Kbar[(long long)c * chi + j], conj(w1[(long long)i * chi + j]));
      Rxbar[f] = Rxbar[f] + le[c / rg] * a;
    }
    __syncthreads();
    // w3 = Wk-bar = conj(Rx) Lambda Kbar  (n x chi)
    for (int f = threadIdx.x; f < n * chi; f += blockDim.x) {
      const int i = f / chi, j = f % chi;
      cd a = mk<double>(0.0, 0.0);
      for (int c = 0; c < n; ++c) fma_acc(a, conj(Rx[(long long)i * n + c]), le[c / rg] * Kbar[(long long)c * chi + j]);
      w3[f] = a;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < chi; j += blockDim.x) {
      if (j < kr && sg[j] > 0.0) {
        double a = 0.0;
        for (int i = 0; i < n; ++i) {
          const cd wb = w3[(long long)i * chi + j];
          if (side == 0) {
            const cd v = V[(long long)i * n + j];
            a += v.re * wb.re + v.im * wb.im;  // Re conj(v) wb
          } else {
            const cd u = U[(long long)i * n + j];
            a += u.re * wb.re - u.im * wb.im;  // Re u wb
          }
        }
        sbar[j] += -a / (sg[j] * sg[j]);
      }
    }
    for (int f = threadIdx.x; f < n * chi; f += blockDim.x) {
      const int i = f / chi, j = f % chi;
      if (j < kr && sg[j] > 0.0) {
        const cd wb = (1.0 / sg[j]) * w3[f];
        if (side == 0)
          Vb[(long long)i * n + j] = Vb[(long long)i * n + j] + wb;
        else
          Ub[(long long)i * n + j] = Ub[(long long)i * n + j] + conj(wb);
      }
    }
    __syncthreads();
  }
  // C-bar from the SVD adjoint
  adj::svd_vjp(U, V, sg, Ub, Vb, sbar, n, Cb, w1, w2, w3);
  // C = R_a Lambda R_b^T:  X = Cbar conj(R_b), Y = Cbar^T conj(R_a)
  adj::bmm(w1, n, Cb, n, adj::N, Rb, n, adj::C, n, n, n, false);
  adj::bmm(w2, n, Cb, n, adj::T, Ra, n, adj::C, n, n, n, false);
  for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
    const int c = f % n;
    Rab[f] = Rab[f] + le[c / rg] * w1[f];
    Rbb[f] = Rbb[f] + le[c / rg] * w2[f];
  }
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    double a = 0.0;
    for (int i = 0; i < n; ++i) {
      const cd x = w1[(long long)i * n + c], r = Ra[(long long)i * n + c];
      a += x.re * r.re + x.im * r.im;
    }
    lamb[c] += a;
  }
  __syncthreads();
  // Cholesky adjoints -> G-bar (Hermitian)
  adj::chol_vjp(Ra, Rab, tp.piv_a(), n, Gba, w1, w2, w3, w4, &kc_sh, cut);
  adj::chol_vjp(Rb, Rbb, tp.piv_b(), n, Gbb, w1, w2, w3, w4, &kc_sh, cut);

  // ---------------- per side: environment Gram adjoint, half gate, lambda of other bonds
  for (int side = 0; side < 2; ++side) {
    const int nod = nd[side], slot = sl[side], d = g.deg[nod];
    const int rows = (int)ipow(chi, d - 1), m = rows * p;
    int str[kMaxDeg];
    {
      int mm = p;
      for (int j = d - 1; j >= 0; --j) {
        str[j] = mm;
        mm *= chi;
      }
    }
    const cd* R = side == 0 ? Ra : Rb;
    cd* Gb = side == 0 ? Gba : Gbb;
    const cd* A = side == 0 ? W.Aa() : W.Ab();
    cd* B = side == 0 ? W.Ba() : W.Bb();
    const double* w = wl[side];
    const double tr = nu[side] * nu[side];
    // G = R^H R (normalised Gram);  Ebar = (Gbar - Re<G, Gbar> I) / tr + nubar / (2 nu) I
    adj::bmm(w1, n, R, n, adj::H, R, n, adj::N, n, n, n, false);
    double acc = 0.0;
    for (int f = threadIdx.x; f < n * n; f += blockDim.x) acc += w1[f].re * Gb[f].re + w1[f].im * Gb[f].im;
    acc = block_sum(acc, red);
    const double diag = -acc / tr + nubar[side] / (2.0 * nu[side]);
    for (int f = threadIdx.x; f < n * n; f += blockDim.x) {
      const int i = f / n, j = f % n;
      cd v = (1.0 / tr) * Gb[f];
      if (i == j) v.re += diag;
      w2[f] = v;
    }
    __syncthreads();
    // AE = Atil Ebar;  B += 2 D^2 AE (Gram adjoint, D the row weight)
    cd* AE = W.AE();
    adj::bmm(AE, n, A, n, adj::N, w2, n, adj::N, m, n, n, false);
    const int L = p * n;  // elements of one row r: (o, c), contiguous
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63, nwv = blockDim.x >> 6;
    if (p == 2 && L <= 64) {
      // one wave per row, four rows per batch (loads first): B += 2 D^2 AE and
      // rho_r = 2 D_r Re sum conj(AE) Atil over the row (wave sum);
      // lambda_j[b_j(r)] += rho_r * prod of the other weights of row r
      for (int r0 = 4 * wv; r0 < rows; r0 += 4 * nwv) {
        cd aev[4], av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const long long f = (long long)(r0 + u) * L + ln;
          if (r0 + u < rows && ln < L) {
            aev[u] = AE[f];
            av[u] = A[f];
            bv[u] = B[f];
          } else {
            aev[u] = av[u] = bv[u] = mk<double>(0.0, 0.0);
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int r = r0 + u;
          if (r >= rows) break;
          const double D = row_weight(w, d, slot, chi, r, -1, false, eps);
          if (ln < L) B[(long long)r * L + ln] = bv[u] + (2.0 * D * D) * aev[u];
          const double rho = 2.0 * D * wave_sum(aev[u].re * av[u].re + aev[u].im * av[u].im);
          if (rho != 0.0 && ln < d && ln != slot) {
            const int j = ln;
            const int bj = row_bond(d, slot, chi, r, j);
            const double wj = row_weight(w, d, slot, chi, r, j, false, eps);
            const int ej = g.adj_edge[g.adj_ptr[nod] + j];
            atomicAdd(&lambar[(long long)ej * chi + bj], rho * wj);
          }
        }
      }
    } else {
      for (int f = threadIdx.x; f < m * n; f += blockDim.x) {
        const double D = row_weight(w, d, slot, chi, (f / n) / p, -1, false, eps);
        B[f] = B[f] + (2.0 * D * D) * AE[f];
      }
      __syncthreads();
      // row weight gradient: rho_r = 2 D_r Re sum_{o,c} conj(AE[(r,o),c]) Atil[(r,o),c];
      // lambda_j[b_j(r)] += rho_r * prod of the other weights of row r
      for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        double rho = 0.0;
        for (int o = 0; o < p; ++o) {
          const long long Rr = (long long)r * p + o;
          for (int c = 0; c < n; ++c) {
            const cd ae = AE[Rr * n + c], a = A[Rr * n + c];
            rho += ae.re * a.re + ae.im * a.im;
          }
        }
        const double D = row_weight(w, d, slot, chi, r, -1, false, eps);
        rho *= 2.0 * D;
        if (rho != 0.0) {
          for (int j = 0; j < d; ++j) {
            if (j == slot) continue;
            const int bj = row_bond(d, slot, chi, r, j);
            const double wj = row_weight(w, d, slot, chi, r, j, false, eps);
            const int ej = g.adj_edge[g.adj_ptr[nod] + j];
            atomicAdd(&lambar[(long long)ej * chi + bj], rho * wj);
          }
        }
      }
    }
    __syncthreads();
    // Gamma-bar and h-bar from B = gradient of Atil
    const cd* t = Tpre + g.toff[nod];
    cd* tbw = Tbar + g.toff[nod];
    const cd* h = hh[side];
    cd* hb = hbar + (long long)blockIdx.x * 2 * hstride + side * hstride;
    if (p == 2 && rg <= 2) {
      // one pass over (r, b), two pairs per batch (loads first):
      //   Gamma-bar[r, b, s] = sum_{o, q} conj(h[q][o][s]) B[(r, o), (b, q)]
      //   h-bar[q][o][s]     = sum_{r, b} B[(r, o), (b, q)] conj(T[r, b, s])
      // with the 16 h-bar partials of each thread reduced once at the end
      __shared__ cd hs[16];
      if (threadIdx.x < 16) hs[threadIdx.x] = h[threadIdx.x];
      __syncthreads();
      double ar[16], ai[16];
#pragma unroll
      for (int f = 0; f < 16; ++f) ar[f] = ai[f] = 0.0;
      constexpr int kB = 2;
      for (int rb0 = threadIdx.x; rb0 < rows * chi; rb0 += kB * blockDim.x) {
        cd y[kB][2], x[kB][2][2];
        long long to[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
          const int rb = rb0 + u * blockDim.x;
          to[u] = -1;
          y[u][0] = y[u][1] = mk<double>(0.0, 0.0);
#pragma unroll
          for (int o = 0; o < 2; ++o) x[u][o][0] = x[u][o][1] = mk<double>(0.0, 0.0);
          if (rb < rows * chi) {
            const int r = rb / chi, b = rb - (rb / chi) * chi;
            to[u] = elem_off(str, d, slot, chi, r, b, 0);
            y[u][0] = t[to[u]];
            y[u][1] = t[to[u] + 1];
#pragma unroll
            for (int o = 0; o < 2; ++o)
#pragma unroll
              for (int q = 0; q < 2; ++q)
                if (q < rg) x[u][o][q] = B[(long long)(r * 2 + o) * n + b * rg + q];
          }
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
          if (to[u] < 0) continue;
#pragma unroll
          for (int sp = 0; sp < 2; ++sp) {
            cd acc2 = mk<double>(0.0, 0.0);
#pragma unroll
            for (int o = 0; o < 2; ++o)
#pragma unroll
              for (int q = 0; q < 2; ++q)
                if (q < rg) fma_acc(acc2, conj(hs[(q * 2 + o) * 2 + sp]), x[u][o][q]);
            tbw[to[u] + sp] = acc2;
          }
#pragma unroll
          for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int o = 0; o < 2; ++o) {
              const cd xx = x[u][o][q];
              const int f = (q * 2 + o) * 2;
              ar[f] += xx.re * y[u][0].re + xx.im * y[u][0].im;
              ai[f] += xx.im * y[u][0].re - xx.re * y[u][0].im;
              ar[f + 1] += xx.re * y[u][1].re + xx.im * y[u][1].im;
              ai[f + 1] += xx.im * y[u][1].re - xx.re * y[u][1].im;
            }
        }
      }
      __shared__ double hred[32][8];
#pragma unroll
      for (int f = 0; f < 16; ++f) {
        const double sr = wave_sum(ar[f]), si = wave_sum(ai[f]);
        if (ln == 0) {
          hred[2 * f][wv] = sr;
          hred[2 * f + 1][wv] = si;
        }
      }
      __syncthreads();
      if (threadIdx.x < 16) {
        const int f = threadIdx.x;
        double sr = 0.0, si = 0.0;
        for (int w2i = 0; w2i < nwv; ++w2i) {
          sr += hred[2 * f][w2i];
          si += hred[2 * f + 1][w2i];
        }
        hb[f] = mk<double>(sr, si);
      }
      __syncthreads();
      continue;
    }
    const long long S = ipow(chi, d) * p;
    for (long long f = threadIdx.x; f < S; f += blockDim.x) {
      // decode f -> (row r, bond b at slot, phys s)
      const int s = (int)(f % p);
      long long rest = f / p;
      int b = 0, r = 0, mul = 1;
      for (int j = d - 1; j >= 0; --j) {
        const int bj = (int)(rest % chi);
        rest /= chi;
        if (j == slot) {
          b = bj;
        } else {
          r += bj * mul;
          mul *= chi;
        }
      }
      cd acc2 = mk<double>(0.0, 0.0);
      for (int o = 0; o < p; ++o)
        for (int q = 0; q < rg; ++q)
          fma_acc(acc2, conj(h[((long long)q * p + o) * p + s]), B[(long long)(r * p + o) * n + b * rg + q]);
      tbw[f] = acc2;
    }
    // h-bar[q][o][s] = sum_{r, b} B[(r, o), (b, q)] conj(T[r, b, s]): the (r, b)
    // pairs are spread over the block, one partial per (q, o, s) per thread,
    // then block sums (q < rank; the rest are zero)
    for (int f = 0; f < (int)hstride; ++f) {
      const int q = f / (p * p), o = (f / p) % p, s = f % p;
      double ar = 0.0, ai = 0.0;
      if (q < rg)
        for (int rb = threadIdx.x; rb < rows * chi; rb += blockDim.x) {
          const int r = rb / chi, b = rb % chi;
          const cd x = B[(long long)(r * p + o) * n + b * rg + q], y = t[elem_off(str, d, slot, chi, r, b, s)];
          ar += x.re * y.re + x.im * y.im;
          ai += x.im * y.re - x.re * y.im;
        }
      ar = block_sum(ar, red);
      ai = block_sum(ai, red);
      if (threadIdx.x == 0) hb[f] = mk<double>(ar, ai);
    }
    __syncthreads();
  }
  // shared-edge lambda gradient (replaces the consumed output gradient)
  for (int b = threadIdx.x; b < chi; b += blockDim.x) {
    double a = 0.0;
    for (int q = 0; q < rg; ++q) a += lamb[b * rg + q];
    lambar[(long long)e * chi + b] = a;
  }
}

hipError_t launch_su_vjp(const DevGraph& g, const void* Tpre, const double* lam, const SuGate* gates, int count,
                         const cd* halves, int max_gate_rank, double eps, double cut, const void* tape, void* Tbar,
                         double* lambar, void* hbar, void* ws, long long wsb, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_su_vjp, dim3(count), dim3(256), 0, s, g, (const cd*)Tpre, lam, gates, halves, max_gate_rank,
                     eps, cut, (const char*)tape, su_tape_bytes(g.chi, max_gate_rank), (cd*)Tbar, lambar, (cd*)hbar,
                     (char*)ws, wsb);
  return hipGetLastError();
}

long long su_vjp_ws_bytes_host(int chi, int phys, int max_r, int max_deg) {
  return su_vjp_ws_bytes(chi, phys, max_r, max_deg);
}

// ---------------------------------------------------------------------------
// BP node update adjoint (one synchronous sweep, tensor_ops.py:36-60):
// for every node and bond idx, out = B / tr B with B = Gram(conj T, U_idx),
// U_idx = T with the incoming messages on the other bonds.  Given obar
// (gradient of the sweep's node->edge outputs) writes ebar (gradient of the
// sweep's edge->node inputs) and accumulates Tbar.
__global__ void __launch_bounds__(256)
    k_bp_node_vjp(DevGraph g, const cd* T, const cd* e2n, const cd* obar, cd* ebar, cd* Tbar, char* ws,
                  long long wsb) {
  const int chi = g.chi, p = g.phys;
  const long long cc = (long long)chi * chi;
  for (int node = blockIdx.x; node < g.n_nodes; node += gridDim.x) {
    const int d = g.deg[node], base = g.adj_ptr[node];
    const long long S = ipow(chi, d) * p;
    cd* W = (cd*)(ws + (long long)blockIdx.x * wsb);
    cd* Ub = W + (long long)(d > 1 ? d - 1 : 1) * S;
    cd* tmp = Ub + S;
    cd* B = tmp + S;
    cd* Bb = B + cc;
    cd* G2 = Bb + cc;
    cd* Mbar = G2 + cc;
    const cd* t = T + g.toff[node];
    cd* tb = Tbar + g.toff[node];
    for (long long q = threadIdx.x; q < d * cc; q += blockDim.x) Mbar[q] = mk<double>(0.0, 0.0);
    __syncthreads();
    for (int idx = 0; idx < d; ++idx) {
      const cd* prev = t;
      int ord[kMaxDeg];
      int m = 0;
      for (int j = 0; j < d; ++j) {
        if (j == idx) continue;
        const cd* Mj = e2n + (2LL * g.adj_edge[base + j] + g.adj_side[base + j]) * cc;
        adj::mprod(W + (long long)m * S, prev, Mj, chi, d, p, j, 0, false);
        prev = W + (long long)m * S;
        ord[m++] = j;
      }
      const cd* U = prev;
      bond_gram<double>(B, t, U, chi, d, p, idx, nullptr);
      __syncthreads();
      adj::trace_norm_vjp(B, obar + (2LL * g.adj_edge[base + idx] + g.adj_side[base + idx]) * cc, chi, Bb);
      adj::mprod(tb, U, Bb, chi, d, p, idx, 1, true);  // Tbar += conj(Bb) U  (bra side)
      cd* ub = Ub;
      cd* tp = tmp;
      adj::mprod(ub, t, Bb, chi, d, p, idx, 2, false);  // Ubar = Bb^T T
      for (int k = m - 1; k >= 0; --k) {
        const int j = ord[k];
        const cd* Wp = k == 0 ? t : W + (long long)(k - 1) * S;
        const cd* Mj = e2n + (2LL * g.adj_edge[base + j] + g.adj_side[base + j]) * cc;
        bond_gram<double>(G2, ub, Wp, chi, d, p, j, nullptr);
        __syncthreads();
        for (long long q = threadIdx.x; q < cc; q += blockDim.x) Mbar[j * cc + q] = Mbar[j * cc + q] + conj(G2[q]);
        adj::mprod(tp, ub, Mj, chi, d, p, j, 3, false);  // Ubar <- M^H Ubar on bond j
        cd* sw = ub;
        ub = tp;
        tp = sw;
      }
      for (long long f = threadIdx.x; f < S; f += blockDim.x) tb[f] = tb[f] + ub[f];
      __syncthreads();
    }
    for (long long q = threadIdx.x; q < d * cc; q += blockDim.x) {
      const int j = (int)(q / cc);
      ebar[(2LL * g.adj_edge[base + j] + g.adj_side[base + j]) * cc + q % cc] = Mbar[q];
    }
    __syncthreads();
  }
}

// The same sweep adjoint for degree-3, chi = 16, phys = 2 nodes on FP64
// matrix cores, one node per 512-thread block, one physical component of T
// in LDS at a time.  For output slot idx with the other axes j1 < j2 and
// forward U = (T x_j1 M1) x_j2 M2, O = G / tr(G), G = sum conj(T) U:
//   pass 1  G over both components, then Gbar = trace_norm_vjp(G, Obar);
//   pass 2  per component, with Ubar = T x_idx Gbar^T:
//     Ubar1 = Ubar x_j2 M2^H,  M1bar += conj(Gram_j1(Ubar1, T)),
//     Tbar += Ubar1 x_j1 M1^H;
//     X = Ubar x_j1 M1^H,      M2bar += conj(Gram_j2(X, T))
//       (Gram_j2(Ubar, T x_j1 M1) = Gram_j2(Ubar x_j1 M1^H, T): products on
//        different axes commute, so no intermediate of the forward is kept);
//     Tbar += U x_idx conj(Gbar)  (the bra side, U recomputed).
// The same arithmetic as k_bp_node_vjp with every contraction a 16 x 16
// complex MFMA tile; M bars accumulate in registers over both passes.
__global__ void __launch_bounds__(64 * fast64::NW)
    k_bp_node_vjp_d3c16(DevGraph g, const cd* T, const cd* e2n, const cd* obar, cd* ebar, cd* Tbar) {
  using namespace fast64;
  __shared__ cd tl[NPAD];
  __shared__ cd wk[NPAD];
  __shared__ cd Gb[3][256];
  __shared__ cd diag[16];
  __shared__ double red[2][8];
  const int node = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int base = g.adj_ptr[node];
  const cd* tg = T + g.toff[node];
  cd* tb = Tbar + g.toff[node];
  const cd* Mg[3];
  long long slot[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    slot[k] = 2LL * g.adj_edge[base + k] + g.adj_side[base + k];
    Mg[k] = e2n + slot[k] * 256;
  }
  auto others = [](int idx, int& j1, int& j2) {
    j1 = idx == 0 ? 1 : 0;
    j2 = idx == 2 ? 1 : 2;
  };
  cd m1[4], m2[4], op[4];
  // pass 1: the forward Grams
  d4 gr[3], gi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) gr[k] = gi[k] = d4{0.0, 0.0, 0.0, 0.0};
  {
    cd mm[3][4];
#pragma unroll
    for (int k = 0; k < 3; ++k) load_operand(Mg[k], 0, lane, mm[k]);
    for (int s = 0; s < 2; ++s) {
      stage_component(tl, tg, s);
      __syncthreads();
      // as the forward kernel: X = T x_2 M2 serves U_0 and U_1, X = T x_0 M0
      // serves U_2, each U's second product fused with its Gram
      mode_product(wk, tl, mm[2], 2, wave, lane);
      __syncthreads();
      mode_gram(tl, wk, mm[1], 1, 0, 2, wave, lane, gr[0], gi[0]);
      mode_gram(tl, wk, mm[0], 0, 1, 2, wave, lane, gr[1], gi[1]);
      __syncthreads();
      mode_product(wk, tl, mm[0], 0, wave, lane);
      __syncthreads();
      mode_gram(tl, wk, mm[1], 1, 2, 0, wave, lane, gr[2], gi[2]);
      __syncthreads();
    }
  }
  // Gbar = Obar / conj(tr) - conj(<Obar, G> / tr^2) I per output slot
  cd* part = wk;  // NW x 256
#pragma unroll
  for (int idx = 0; idx < 3; ++idx) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
      part[wave * 256 + ((lane >> 4) + 4 * v) * 16 + (lane & 15)] = mk<double>(gr[idx][v], gi[idx][v]);
    __syncthreads();
    cd G = mk<double>(0.0, 0.0), ob = G;
    double ipr = 0.0, ipi = 0.0;
    if (tid < 256) {
      for (int w = 0; w < NW; ++w) G = G + part[w * 256 + tid];
      if (tid % 17 == 0) diag[tid / 17] = G;
      ob = obar[slot[idx] * 256 + tid];
      ipr = ob.re * G.re + ob.im * G.im;  // conj(Obar) G
      ipi = ob.re * G.im - ob.im * G.re;
    }
    ipr = wave_sum(ipr);
    ipi = wave_sum(ipi);
    if (lane == 0) {
      red[0][wave] = ipr;
      red[1][wave] = ipi;
    }
    __syncthreads();
    if (tid < 256) {
      cd tr = mk<double>(0.0, 0.0), ip = tr;
#pragma unroll
      for (int k = 0; k < 16; ++k) tr = tr + diag[k];
      for (int w = 0; w < NW; ++w) ip = ip + mk<double>(red[0][w], red[1][w]);
      cd v = cdiv(ob, conj(tr));
      if (tid % 17 == 0) v = v - conj(cdiv(ip, tr * tr));
      Gb[idx][tid] = v;
    }
    __syncthreads();
  }
  // pass 2: starts with component 1, still in LDS.  H_j accumulates
  // sum conj(T[a@j]) Z[b@j] over the Z of bond j (Ubar1 or X); Mbar_j[b][a] = H_j[a][b].
  d4 hr[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) hr[k] = hi[k] = d4{0.0, 0.0, 0.0, 0.0};
  constexpr int TPER = 4096 / (64 * NW);
  // Gamma-bar += the work tensor (this thread's elements f = tid + 512 k of
  // component s; all loads first)
  auto add_tbar = [&](int s) {
    cd v[TPER], o[TPER];
#pragma unroll
    for (int k = 0; k < TPER; ++k) {
      const int f = tid + k * 64 * NW;
      v[k] = wk[(f >> 8) * B0P + ((f >> 4) & 15) * RP + (f & 15)];
      o[k] = tb[2 * f + s];
    }
#pragma unroll
    for (int k = 0; k < TPER; ++k) tb[2 * (tid + k * 64 * NW) + s] = o[k] + v[k];
  };
  for (int s = 1; s >= 0; --s) {
    if (s == 0) {
      stage_component(tl, tg, 0);
      __syncthreads();
    }
#pragma unroll
    for (int idx = 0; idx < 3; ++idx) {
      int j1, j2;
      others(idx, j1, j2);
      load_operand(Gb[idx], 2, lane, op);  // Gbar^T
      load_operand(Mg[j1], 3, lane, m1);   // M1^H
      load_operand(Mg[j2], 3, lane, m2);   // M2^H
      // Y = T x_idx Gbar^T
      mode_product(wk, tl, op, idx, wave, lane);
      __syncthreads();
      // X = Y x_j1 M1^H (not stored) with its Gram over j2
      mode_gram(tl, wk, m1, j1, j2, idx, wave, lane, hr[j2], hi[j2]);
      __syncthreads();
      // Ubar1 = Y x_j2 M2^H (stored in place) with its Gram over j1
      mode_gram<true>(tl, wk, m2, j2, j1, idx, wave, lane, hr[j1], hi[j1]);
      __syncthreads();
      // Tbar += Ubar1 x_j1 M1^H
      mode_product(wk, wk, m1, j1, wave, lane);
      __syncthreads();
      add_tbar(s);
      __syncthreads();
      // bra side: Tbar += U x_idx conj(Gbar), U = (T x_j1 M1) x_j2 M2
      load_operand(Mg[j1], 0, lane, m1);
      load_operand(Mg[j2], 0, lane, m2);
      load_operand(Gb[idx], 1, lane, op);
      mode_product(wk, tl, m1, j1, wave, lane);
      __syncthreads();
      mode_product(wk, wk, m2, j2, wave, lane);
      __syncthreads();
      mode_product(wk, wk, op, idx, wave, lane);
      __syncthreads();
      add_tbar(s);
      __syncthreads();
    }
  }
  // Mbar_j[b][a] = H_j[a][b], summed over the waves
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
      part[wave * 256 + (lane & 15) * 16 + (lane >> 4) + 4 * v] = mk<double>(hr[j][v], hi[j][v]);
    __syncthreads();
    if (tid < 256) {
      cd sum = mk<double>(0.0, 0.0);
      for (int w = 0; w < NW; ++w) sum = sum + part[w * 256 + tid];
      ebar[slot[j] * 256 + tid] = sum;
    }
    __syncthreads();
  }
}

// The absorb adjoint for degree-3, chi = 16, phys = 2 nodes on FP64 matrix
// cores (one node per 512-thread block, one physical component of T in LDS
// at a time).  Pass 1 recomputes Y = T x_0 X_0^T x_1 X_1^T x_2 X_2^T for |Y| and
// Re<Y, Gamma-bar>; pass 2 forms Ybar = (Gamma-bar - Re<Y, Gamma-bar>/|Y|^2 Y)/|Y|
// per component (kept in the workspace) and, with products on different axes
// commuting, Xbar_i^T = conj(Gram_i(Ybar x_{j != i} conj(X_j), T)) and
// Tbar += Ybar x_0 conj(X_0) x_1 conj(X_1) x_2 conj(X_2).
__global__ void __launch_bounds__(64 * fast64::NW)
    k_absorb_vjp_d3c16(DevGraph g, const cd* T, const cd* X, const cd* gbar, cd* Tbar, cd* Xbar, cd* ws) {
  using namespace fast64;
  __shared__ cd tl[NPAD];
  __shared__ cd wk[NPAD];
  __shared__ double red[NW];
  const int node = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int base = g.adj_ptr[node];
  const cd* tg = T + g.toff[node];
  const cd* gb = gbar + g.toff[node];
  cd* tb = Tbar + g.toff[node];
  cd* yb = ws + (long long)node * 8192;  // Ybar, (b0, b1, b2, s) row-major
  long long slot[3];
  cd mt[3][4], mc[3][4];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    slot[k] = 2LL * g.adj_edge[base + k] + g.adj_side[base + k];
    load_operand(X + slot[k] * 256, 2, lane, mt[k]);  // X^T (forward)
    load_operand(X + slot[k] * 256, 1, lane, mc[k]);  // conj(X) (adjoint)
  }
  auto padded = [](int f) { return (f >> 8) * B0P + ((f >> 4) & 15) * RP + (f & 15); };
  // pass 1: |Y|^2 and Re<Y, Gamma-bar>
  double nn = 0.0, ip = 0.0;
  for (int s = 0; s < 2; ++s) {
    stage_component(tl, tg, s);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      mode_product(tl, tl, mt[j], j, wave, lane);
      __syncthreads();
    }
    for (int f = tid; f < 4096; f += blockDim.x) {
      const cd y = tl[padded(f)], b = gb[2 * f + s];
      nn += norm2(y);
      ip += y.re * b.re + y.im * b.im;
    }
    __syncthreads();
  }
  nn = sqrt(block_sum(nn, red));
  ip = block_sum(ip, red) / nn;
  d4 gr[3], gi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) gr[k] = gi[k] = d4{0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < 2; ++s) {
    stage_component(tl, tg, s);
    __syncthreads();
    mode_product(wk, tl, mt[0], 0, wave, lane);
    __syncthreads();
    mode_product(wk, wk, mt[1], 1, wave, lane);
    __syncthreads();
    mode_product(wk, wk, mt[2], 2, wave, lane);
    __syncthreads();
    for (int f = tid; f < 4096; f += blockDim.x) {
      const int q = padded(f);
      const cd y = wk[q], b = gb[2 * f + s];
      const cd v = (1.0 / nn) * (b - (ip / nn) * y);
      wk[q] = v;
      yb[2 * f + s] = v;
    }
    __syncthreads();
    // axes 0, 1 -> Gram over axis 2, then axis 2 -> Tbar
    mode_product(wk, wk, mc[0], 0, wave, lane);
    __syncthreads();
    mode_product(wk, wk, mc[1], 1, wave, lane);
    __syncthreads();
    gram_partial(wk, tl, 2, wave, lane, gr[2], gi[2]);
    __syncthreads();
    mode_product(wk, wk, mc[2], 2, wave, lane);
    __syncthreads();
    for (int f = tid; f < 4096; f += blockDim.x) tb[2 * f + s] = tb[2 * f + s] + wk[padded(f)];
    __syncthreads();
    // axes 1, 2 -> Gram over axis 0
    for (int f = tid; f < 4096; f += blockDim.x) wk[padded(f)] = yb[2 * f + s];
    __syncthreads();
    mode_product(wk, wk, mc[1], 1, wave, lane);
    __syncthreads();
    mode_product(wk, wk, mc[2], 2, wave, lane);
    __syncthreads();
    gram_partial(wk, tl, 0, wave, lane, gr[0], gi[0]);
    __syncthreads();
    // axes 0, 2 -> Gram over axis 1
    for (int f = tid; f < 4096; f += blockDim.x) wk[padded(f)] = yb[2 * f + s];
    __syncthreads();
    mode_product(wk, wk, mc[0], 0, wave, lane);
    __syncthreads();
    mode_product(wk, wk, mc[2], 2, wave, lane);
    __syncthreads();
    gram_partial(wk, tl, 1, wave, lane, gr[1], gi[1]);
    __syncthreads();
  }
  // Xbar[slot][b][a] = conj(G[a][b]), G reduced over the waves
  cd* part = wk;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
      part[wave * 256 + ((lane >> 4) + 4 * v) * 16 + (lane & 15)] = mk<double>(gr[j][v], gi[j][v]);
    __syncthreads();
    if (tid < 256) {
      cd sum = mk<double>(0.0, 0.0);
      for (int w = 0; w < NW; ++w) sum = sum + part[w * 256 + tid];
      Xbar[slot[j] * 256 + (tid & 15) * 16 + (tid >> 4)] = conj(sum);
    }
    __syncthreads();
  }
}

hipError_t launch_absorb_vjp_fast(const DevGraph& g, const void* T, const void* X, const void* gbar, void* Tbar,
                                  void* Xbar, void* ws, hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_absorb_vjp_d3c16, dim3(g.n_nodes), dim3(64 * fast64::NW), 0, s, g, (const cd*)T, (const cd*)X,
                     (const cd*)gbar, (cd*)Tbar, (cd*)Xbar, (cd*)ws);
  return hipGetLastError();
}

// Absorb adjoint: Gamma = Y / |Y|, Y = T x_i X_i^T on every bond
// (tensor_ops.py:88-93).  Accumulates Tbar, writes Xbar per (node, slot).
__global__ void __launch_bounds__(256)
    k_absorb_vjp(DevGraph g, const cd* T, const cd* X, const cd* gbar, cd* Tbar, cd* Xbar, char* ws, long long wsb) {
  __shared__ double red[8];
  const int chi = g.chi, p = g.phys;
  const long long cc = (long long)chi * chi;
  for (int node = blockIdx.x; node < g.n_nodes; node += gridDim.x) {
    const int d = g.deg[node], base = g.adj_ptr[node];
    const long long S = ipow(chi, d) * p;
    cd* W = (cd*)(ws + (long long)blockIdx.x * wsb);  // W_1..W_d
    cd* Yb = W + (long long)d * S;
    cd* tmp = Yb + S;
    cd* G2 = tmp + S;
    const cd* t = T + g.toff[node];
    const cd* prev = t;
    for (int i = 0; i < d; ++i) {
      const cd* Xi = X + (2LL * g.adj_edge[base + i] + g.adj_side[base + i]) * cc;
      adj::mprod(W + (long long)i * S, prev, Xi, chi, d, p, i, 2, false);
      prev = W + (long long)i * S;
    }
    const cd* Y = prev;
    double nn = 0.0, ip = 0.0;
    const cd* gb = gbar + g.toff[node];
    for (long long f = threadIdx.x; f < S; f += blockDim.x) {
      nn += norm2(Y[f]);
      ip += Y[f].re * gb[f].re + Y[f].im * gb[f].im;
    }
    nn = sqrt(block_sum(nn, red));
    ip = block_sum(ip, red) / nn;  // Re<Gamma, Gamma-bar>
    for (long long f = threadIdx.x; f < S; f += blockDim.x) Yb[f] = (1.0 / nn) * (gb[f] - (ip / nn) * Y[f]);
    __syncthreads();
    cd* yb = Yb;
    cd* tp = tmp;
    for (int i = d - 1; i >= 0; --i) {
      const cd* Wp = i == 0 ? t : W + (long long)(i - 1) * S;
      const long long slot = 2LL * g.adj_edge[base + i] + g.adj_side[base + i];
      const cd* Xi = X + slot * cc;
      bond_gram<double>(G2, yb, Wp, chi, d, p, i, nullptr);
      __syncthreads();
      for (long long q = threadIdx.x; q < cc; q += blockDim.x) {
        const int a = (int)(q / chi), b = (int)(q % chi);
        Xbar[slot * cc + (long long)b * chi + a] = conj(G2[q]);
      }
      adj::mprod(tp, yb, Xi, chi, d, p, i, 1, false);  // sum_a conj(X[b][a]) Ybar[a]
      cd* sw = yb;
      yb = tp;
      tp = sw;
    }
    cd* tb = Tbar + g.toff[node];
    for (long long f = threadIdx.x; f < S; f += blockDim.x) tb[f] = tb[f] + yb[f];
    __syncthreads();
  }
}

// Matrix-function divided differences for the safe square root
// f(w) = sqrt(w) (w > eps, else 0) and the safe inverse square root
// g(w) = 1 / sqrt(w) (w >= eps, else 1 / sqrt(FLT_MAX + w)).
__device__ __forceinline__ double dd_sqrt(double a, double b, double eps) {
  const bool pa = a > eps, pb = b > eps;
  if (!pa && !pb) return 0.0;
  if (pa && pb) return 1.0 / (sqrt(a) + sqrt(b));
  const double fa = pa ? sqrt(a) : 0.0, fb = pb ? sqrt(b) : 0.0;
  return (fa - fb) / (a - b);
}
__device__ __forceinline__ double dd_isqrt(double a, double b, double eps) {
  const bool pa = !(a < eps), pb = !(b < eps);
  if (!pa && !pb) return 0.0;
  if (pa && pb) {
    const double sa = sqrt(a), sb = sqrt(b);
    return -1.0 / (sa * sb * (sa + sb));
  }
  const double ga = pa ? 1.0 / sqrt(a) : 1.0 / sqrt((double)FLT_MAX + a);
  const double gb = pb ? 1.0 / sqrt(b) : 1.0 / sqrt((double)FLT_MAX + b);
  return (ga - gb) / (a - b);
}

// Vidal edge decomposition adjoint (tensor_ops.py:108-130) from the forward's
// tape.  Given Xbar (orthogonalisers) and lambar writes n2e_bar for both
// messages of every edge (Hermitian part, as the forward symmetrises).
__global__ void __launch_bounds__(256)
    k_edge_svd_vjp(DevGraph g, const char* tape, double eps, double cut, const cd* Xbar, const double* lambar,
                   cd* n2e_bar, char* ws, long long wsb) {
  __shared__ double sb[64], kap[64], lamv[64], sgc[64];
  __shared__ unsigned char frz[2][64];
  const int n = g.chi;
  const long long cc = (long long)n * n;
  for (int e = blockIdx.x; e < g.n_edges; e += gridDim.x) {
    const EdgeSvdTape tp((char*)tape + (long long)e * edge_svd_tape_bytes(n), n);
    cd* w = (cd*)(ws + (long long)blockIdx.x * wsb);
    cd *If = w, *Ib = If + cc, *Sf = Ib + cc, *Sb = Sf + cc, *Ub = Sb + cc, *Vb = Ub + cc, *Ifb = Vb + cc,
       *Ibb = Ifb + cc, *Gb = Ibb + cc, *w1 = Gb + cc, *w2 = w1 + cc, *w3 = w2 + cc, *Sfb = w3 + cc, *Sbb = Sfb + cc;
    const double nrm = tp.scal()[0], nu = tp.scal()[1], nv = tp.scal()[2];
    const cd *Qf = tp.Qf(), *Qb = tp.Qb(), *U = tp.U(), *V = tp.V();
    // Backward cutoff: message eigen-directions and singular values below
    // cut * the largest on the bond are frozen (no gradient flows through
    // them).  Through such directions the adjoint is a difference of terms
    // of size w^-3/2 and 1/sigma that cancel in exact arithmetic only.
    {
      double mf = 0.0, mb = 0.0, ms = 0.0;
      for (int i = 0; i < n; ++i) {
        mf = fmax(mf, tp.wf()[i]);
        mb = fmax(mb, tp.wb()[i]);
        ms = fmax(ms, tp.sigma()[i]);
      }
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        sgc[j] = tp.sigma()[j] > cut * ms ? tp.sigma()[j] : 0.0;
        frz[0][j] = cut > 0.0 && tp.wf()[j] < cut * mf;
        frz[1][j] = cut > 0.0 && tp.wb()[j] < cut * mb;
      }
      __syncthreads();
    }
    const double *wf = tp.wf(), *wb = tp.wb(), *sg = tp.sigma();
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      const double l = nrm > 0.0 ? sg[j] / nrm : 0.0;
      const double lt = l > eps ? l : 0.0;
      lamv[j] = lt;
      kap[j] = lt > eps ? 1.0 : 0.0;
    }
    __syncthreads();
    // matrix functions of both messages
    for (long long q = threadIdx.x; q < cc; q += blockDim.x) {
      const int r = (int)(q / n), c = (int)(q % n);
      cd a = mk<double>(0, 0), b = a, ai = a, bi = a;
      for (int i = 0; i < n; ++i) {
        const cd pf = Qf[(long long)r * n + i] * conj(Qf[(long long)c * n + i]);
        const cd pb = Qb[(long long)r * n + i] * conj(Qb[(long long)c * n + i]);
        a = a + (wf[i] > eps ? sqrt(wf[i]) : 0.0) * pf;
        b = b + (wb[i] > eps ? sqrt(wb[i]) : 0.0) * pb;
        ai = ai + (wf[i] < eps ? 1.0 / sqrt((double)FLT_MAX + wf[i]) : 1.0 / sqrt(wf[i])) * pf;
        bi = bi + (wb[i] < eps ? 1.0 / sqrt((double)FLT_MAX + wb[i]) : 1.0 / sqrt(wb[i])) * pb;
      }
      Sf[q] = a;
      Sb[q] = b;
      If[q] = ai;
      Ib[q] = bi;
    }
    __syncthreads();
    const cd* xf = Xbar + (2LL * e) * cc;
    const cd* xb = Xbar + (2LL * e + 1) * cc;
    // Ubar = kappa (If^H Xf-bar) / nu ; If-bar = Xf-bar (U kappa)^H / nu
    // Vbar = kappa (Ib^T conj(Xb-bar)) / nv ; Ib-bar = Xb-bar (V kappa)^T / nv
    for (long long q = threadIdx.x; q < cc; q += blockDim.x) {
      const int i = (int)(q / n), k = (int)(q % n);
      cd u = mk<double>(0, 0), v = u, fi = u, bi = u;
      for (int b = 0; b < n; ++b) {
        fma_acc(u, conj(If[(long long)b * n + i]), xf[(long long)b * n + k]);
        fma_acc(v, Ib[(long long)b * n + i], conj(xb[(long long)b * n + k]));
      }
      for (int kk = 0; kk < n; ++kk) {
        if (kap[kk] == 0.0) continue;
        fma_acc(fi, xf[(long long)i * n + kk], conj(U[(long long)k * n + kk]));
        fma_acc(bi, xb[(long long)i * n + kk], V[(long long)k * n + kk]);
      }
      Ub[q] = (kap[k] / nu) * u;
      Vb[q] = (kap[k] / nv) * v;
      Ifb[q] = (1.0 / nu) * fi;
      Ibb[q] = (1.0 / nv) * bi;
    }
    // sigma-bar from lambda = sigma / |sigma| (kept entries)
    {
      double dot = 0.0;
      for (int j = 0; j < n; ++j) dot += kap[j] * lamv[j] * lambar[(long long)e * n + j];
      for (int j = threadIdx.x; j < n; j += blockDim.x)
        sb[j] = nrm > 0.0 ? kap[j] * (lambar[(long long)e * n + j] - dot * lamv[j]) / nrm : 0.0;
    }
    __syncthreads();
    adj::svd_vjp(U, V, sgc, Ub, Vb, sb, n, Gb, w1, w2, w3);
    // G = Sf Sb^T:  Sf-bar = Gbar conj(Sb),  Sb-bar = Gbar^T conj(Sf)
    adj::bmm(Sfb, n, Gb, n, adj::N, Sb, n, adj::C, n, n, n, false);
    adj::bmm(Sbb, n, Gb, n, adj::T, Sf, n, adj::C, n, n, n, false);
    for (int side = 0; side < 2; ++side) {
      const cd* Q = side == 0 ? Qf : Qb;
      const double* wv = side == 0 ? wf : wb;
      const unsigned char* fz = frz[side];
      const cd* Fs = side == 0 ? Sfb : Sbb;
      const cd* Fi = side == 0 ? Ifb : Ibb;
      // w1 = Q^H Fs Q, w2 = Q^H Fi Q
      adj::bmm(w3, n, Fs, n, adj::N, Q, n, adj::N, n, n, n, false);
      adj::bmm(w1, n, Q, n, adj::H, w3, n, adj::N, n, n, n, false);
      adj::bmm(w3, n, Fi, n, adj::N, Q, n, adj::N, n, n, n, false);
      adj::bmm(w2, n, Q, n, adj::H, w3, n, adj::N, n, n, n, false);
      for (long long q = threadIdx.x; q < cc; q += blockDim.x) {
        const int i = (int)(q / n), j = (int)(q % n);
        double df, dg;
        if (i == j) {
          const double x = wv[i];
          df = x > eps ? 0.5 / sqrt(x) : 0.0;
          dg = x < eps ? 0.0 : -0.5 / (x * sqrt(x));
        } else {
          df = dd_sqrt(wv[i], wv[j], eps);
          dg = dd_isqrt(wv[i], wv[j], eps);
          if (wv[i] == wv[j]) {
            df = wv[i] > eps ? 0.5 / sqrt(wv[i]) : 0.0;
            dg = wv[i] < eps ? 0.0 : -0.5 / (wv[i] * sqrt(wv[i]));
          }
        }
        if (fz[i] || fz[j]) df = dg = 0.0;
        w3[q] = df * w1[q] + dg * w2[q];
      }
      __syncthreads();
      adj::bmm(w1, n, Q, n, adj::N, w3, n, adj::N, n, n, n, false);
      adj::bmm(w2, n, w1, n, adj::N, Q, n, adj::H, n, n, n, false);
      cd* out = n2e_bar + (2LL * e + side) * cc;
      for (long long q = threadIdx.x; q < cc; q += blockDim.x) {
        const int i = (int)(q / n), j = (int)(q % n);
        const cd a = w2[q], b = w2[(long long)j * n + i];
        out[q] = mk<double>(0.5 * (a.re + b.re), 0.5 * (a.im - b.im));
      }
      __syncthreads();
    }
  }
}

// Symmetric-gauge adjoint: T = Gamma * prod_i sqrt(lambda_i > eps) / |.|
// (tensor_ops.py:189-199).  Accumulates gam_bar; lam_bar by atomics.
__global__ void __launch_bounds__(256)
    k_symmetric_vjp(DevGraph g, const cd* gam, const double* lam, double eps, const cd* tbar, cd* gam_bar,
                    double* lam_bar) {
  __shared__ double red[8];
  __shared__ double lacc[kMaxDeg * 64];
  const int chi = g.chi, p = g.phys;
  for (int node = blockIdx.x; node < g.n_nodes; node += gridDim.x) {
    const int d = g.deg[node], base = g.adj_ptr[node];
    const long long S = ipow(chi, d) * p;
    const cd* t = gam + g.toff[node];
    const cd* tb = tbar + g.toff[node];
    cd* gb = gam_bar + g.toff[node];
    for (int q = threadIdx.x; q < d * chi; q += blockDim.x) lacc[q] = 0.0;
    double nn = 0.0;
    for (long long f = threadIdx.x; f < S; f += blockDim.x) {
      long long r = f / p;
      double w = 1.0;
      for (int j = d - 1; j >= 0; --j) {
        const double l = lam[(long long)g.adj_edge[base + j] * chi + r % chi];
        r /= chi;
        w *= l > eps ? sqrt(l) : 0.0;
      }
      nn += w * w * norm2(t[f]);
    }
    nn = sqrt(block_sum(nn, red));
    double ip = 0.0;  // Re<T, Tbar> with T = Gamma w / nn
    for (long long f = threadIdx.x; f < S; f += blockDim.x) {
      long long r = f / p;
      double w = 1.0;
      for (int j = d - 1; j >= 0; --j) {
        const double l = lam[(long long)g.adj_edge[base + j] * chi + r % chi];
        r /= chi;
        w *= l > eps ? sqrt(l) : 0.0;
      }
      ip += (w / nn) * (t[f].re * tb[f].re + t[f].im * tb[f].im);
    }
    ip = block_sum(ip, red);
    for (long long f = threadIdx.x; f < S; f += blockDim.x) {
      long long r = f / p;
      int bj[kMaxDeg];
      double fj[kMaxDeg];
      double w = 1.0;
      for (int j = d - 1; j >= 0; --j) {
        bj[j] = (int)(r % chi);
        r /= chi;
        const double l = lam[(long long)g.adj_edge[base + j] * chi + bj[j]];
        fj[j] = l > eps ? sqrt(l) : 0.0;
        w *= fj[j];
      }
      // Ybar = (Tbar - ip T) / nn;  Gamma-bar += Ybar w
      const cd yb = (1.0 / nn) * (tb[f] - (ip * w / nn) * t[f]);
      gb[f] = gb[f] + w * yb;
      const double rr = yb.re * t[f].re + yb.im * t[f].im;  // Re conj(Ybar) Gamma
      if (rr != 0.0)
        for (int j = 0; j < d; ++j) {
          if (fj[j] == 0.0) continue;
          double others = 1.0;
          for (int i = 0; i < d; ++i)
            if (i != j) others *= fj[i];
          atomicAdd(&lacc[j * chi + bj[j]], rr * others * 0.5 / fj[j]);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < d * chi; q += blockDim.x)
      if (lacc[q] != 0.0) atomicAdd(&lam_bar[(long long)g.adj_edge[base + q / chi] * chi + q % chi], lacc[q]);
    __syncthreads();
  }
}

// lambdas2messages adjoint (symmetric_gauge.py:50-61): lam_bar[e][k] +=
// Re of the diagonal of every message gradient of edge e (n2e and e2n,
// both sides; either pointer may be NULL).
__global__ void k_lam2msg_vjp(DevGraph g, const cd* n2e_bar, const cd* e2n_bar, double* lam_bar) {
  const int chi = g.chi;
  const long long cc = (long long)chi * chi;
  const long long total = (long long)g.n_edges * chi;
  for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < total; f += (long long)gridDim.x * blockDim.x) {
    const long long e = f / chi;
    const int k = (int)(f % chi);
    double a = 0.0;
    for (int s = 0; s < 2; ++s) {
      const long long q = (2 * e + s) * cc + (long long)k * chi + k;
      if (n2e_bar) a += n2e_bar[q].re;
      if (e2n_bar) a += e2n_bar[q].re;
    }
    lam_bar[f] += a;
  }
}

// One-site density matrix adjoint with diagonal messages diag(lambda)
// (bp_quantum_emulator.py:194-205 feeds the symmetric tensors and
// diag(lambda)): rho = rt / tr(rt), rt[s][s'] = sum_a w_a T[a,s] conj(T[a,s']),
// w_a = prod_j lambda_j[a_j].  Accumulates Tbar; lam_bar by atomics.
__global__ void __launch_bounds__(256)
    k_rdm1_vjp(DevGraph g, const cd* T, const double* lam, const int* nodes, const cd* rho_bar, cd* Tbar,
               double* lam_bar) {
  __shared__ cd rt[64], rb[64];
  __shared__ double lacc[kMaxDeg * 64];
  const int chi = g.chi, p = g.phys;
  const int node = nodes[blockIdx.x];
  const int d = g.deg[node], base = g.adj_ptr[node];
  const long long rows = ipow(chi, d);
  const cd* t = T + g.toff[node];
  for (int q = threadIdx.x; q < p * p; q += blockDim.x) {
    const int s = q / p, sp = q % p;
    cd acc = mk<double>(0, 0);
    for (long long a = 0; a < rows; ++a) {
      long long r = a;
      double w = 1.0;
      for (int j = d - 1; j >= 0; --j) {
        w *= lam[(long long)g.adj_edge[base + j] * chi + r % chi];
        r /= chi;
      }
      fma_acc(acc, w * t[a * p + s], conj(t[a * p + sp]));
    }
    rt[q] = acc;
  }
  for (int q = threadIdx.x; q < d * chi; q += blockDim.x) lacc[q] = 0.0;
  __syncthreads();
  adj::trace_norm_vjp(rt, rho_bar + (long long)blockIdx.x * p * p, p, rb);
  cd* tb = Tbar + g.toff[node];
  for (long long a = threadIdx.x; a < rows; a += blockDim.x) {
    long long r = a;
    int bj[kMaxDeg];
    double lj[kMaxDeg];
    double w = 1.0;
    for (int j = d - 1; j >= 0; --j) {
      bj[j] = (int)(r % chi);
      r /= chi;
      lj[j] = lam[(long long)g.adj_edge[base + j] * chi + bj[j]];
      w *= lj[j];
    }
    double q2 = 0.0;  // Re sum conj(rb[s][s']) T[a,s] conj(T[a,s'])
    for (int s = 0; s < p; ++s) {
      cd v = mk<double>(0, 0);
      for (int sp = 0; sp < p; ++sp) {
        const cd x = t[a * p + sp];
        fma_acc(v, rb[s * p + sp] + conj(rb[sp * p + s]), x);
        const cd z = conj(rb[s * p + sp]) * t[a * p + s] * conj(x);
        q2 += z.re;
      }
      tb[a * p + s] = tb[a * p + s] + w * v;
    }
    if (q2 != 0.0)
      for (int j = 0; j < d; ++j) {
        double others = 1.0;
        for (int i = 0; i < d; ++i)
          if (i != j) others *= lj[i];
        atomicAdd(&lacc[j * chi + bj[j]], q2 * others);
      }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < d * chi; q += blockDim.x)
    if (lacc[q] != 0.0) atomicAdd(&lam_bar[(long long)g.adj_edge[base + q / chi] * chi + q % chi], lacc[q]);
}

// Two-site density matrix adjoint with diagonal messages on the outer bonds:
// rt = sum_{k,k'} E_a[(k s_a),(k' s_a')] E_b[(k s_b),(k' s_b')], rho = rt / tr,
// E_x[(k s),(k' s')] = sum_o w_x(o) T_x[o,k,s] conj(T_x[o,k',s']).  Tbar and
// lam_bar by atomics (several edges meet at a node).  ws: 4 (chi p)^2 cd.
__global__ void __launch_bounds__(256)
    k_rdm2_vjp(DevGraph g, const cd* T, const double* lam, const int* edges, const cd* rho_bar, cd* Tbar,
               double* lam_bar, char* ws, long long wsb) {
  __shared__ cd rt[256], rb[256];
  const int chi = g.chi, p = g.phys, cp = chi * p, pp = p * p;
  const int e = edges[blockIdx.x];
  const int nd[2] = {g.edge_node[2 * e], g.edge_node[2 * e + 1]};
  const int sl[2] = {find_slot(g, nd[0], e, 0), find_slot(g, nd[1], e, 1)};
  cd* E[2];
  cd* Eb[2];
  E[0] = (cd*)(ws + (long long)blockIdx.x * wsb);
  E[1] = E[0] + (long long)cp * cp;
  Eb[0] = E[1] + (long long)cp * cp;
  Eb[1] = Eb[0] + (long long)cp * cp;
  int str[2][kMaxDeg];
  for (int x = 0; x < 2; ++x) {
    const int d = g.deg[nd[x]];
    int mm = p;
    for (int j = d - 1; j >= 0; --j) {
      str[x][j] = mm;
      mm *= chi;
    }
  }
  for (int x = 0; x < 2; ++x) {
    const int nod = nd[x], d = g.deg[nod], base = g.adj_ptr[nod], slot = sl[x];
    const long long orows = ipow(chi, d - 1);
    const cd* t = T + g.toff[nod];
    for (int q = threadIdx.x; q < cp * cp; q += blockDim.x) {
      const int a = q / cp, b = q % cp;
      const int k = a / p, s = a % p, kp = b / p, sp = b % p;
      cd acc = mk<double>(0, 0);
      for (long long o = 0; o < orows; ++o) {
        long long r = o;
        double w = 1.0;
        for (int j = d - 1; j >= 0; --j) {
          if (j == slot) continue;
          w *= lam[(long long)g.adj_edge[base + j] * chi + r % chi];
          r /= chi;
        }
        fma_acc(acc, w * t[elem_off(str[x], d, slot, chi, (int)o, k, s)], conj(t[elem_off(str[x], d, slot, chi, (int)o, kp, sp)]));
      }
      E[x][q] = acc;
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < pp * pp; q += blockDim.x) {
    const int row = q / pp, col = q % pp;
    const int sa = row / p, sb = row % p, sap = col / p, sbp = col % p;
    cd v = mk<double>(0, 0);
    for (int k = 0; k < chi; ++k)
      for (int kp = 0; kp < chi; ++kp)
        fma_acc(v, E[0][(k * p + sa) * cp + kp * p + sap], E[1][(k * p + sb) * cp + kp * p + sbp]);
    rt[q] = v;
  }
  __syncthreads();
  adj::trace_norm_vjp(rt, rho_bar + (long long)blockIdx.x * pp * pp, pp, rb);
  // Ebar_a[(k sa),(k' sa')] = sum_{sb,sb'} rb[(sa sb),(sa' sb')] conj(E_b[(k sb),(k' sb')]), and vice versa
  for (int q = threadIdx.x; q < cp * cp; q += blockDim.x) {
    const int a = q / cp, b = q % cp;
    const int k = a / p, s = a % p, kp = b / p, sp = b % p;
    cd va = mk<double>(0, 0), vb = va;
    for (int u = 0; u < p; ++u)
      for (int up = 0; up < p; ++up) {
        fma_acc(va, rb[(s * p + u) * pp + sp * p + up], conj(E[1][(k * p + u) * cp + kp * p + up]));
        fma_acc(vb, rb[(u * p + s) * pp + up * p + sp], conj(E[0][(k * p + u) * cp + kp * p + up]));
      }
    Eb[0][q] = va;
    Eb[1][q] = vb;
  }
  __syncthreads();
  for (int x = 0; x < 2; ++x) {
    const int nod = nd[x], d = g.deg[nod], base = g.adj_ptr[nod], slot = sl[x];
    const long long orows = ipow(chi, d - 1);
    const cd* t = T + g.toff[nod];
    cd* tb = Tbar + g.toff[nod];
    const cd* Ex = Eb[x];
    for (long long f = threadIdx.x; f < orows * cp; f += blockDim.x) {
      const int o = (int)(f / cp), ks = (int)(f % cp);
      const int k = ks / p, s = ks % p;
      long long r = o;
      int bj[kMaxDeg];
      double lj[kMaxDeg];
      double w = 1.0;
      for (int j = d - 1; j >= 0; --j) {
        if (j == slot) continue;
        bj[j] = (int)(r % chi);
        r /= chi;
        lj[j] = lam[(long long)g.adj_edge[base + j] * chi + bj[j]];
        w *= lj[j];
      }
      cd v = mk<double>(0, 0);
      double q2 = 0.0;
      const cd tv = t[elem_off(str[x], d, slot, chi, o, k, s)];
      for (int k2 = 0; k2 < chi; ++k2)
        for (int s2 = 0; s2 < p; ++s2) {
          const int ks2 = k2 * p + s2;
          const cd x2 = t[elem_off(str[x], d, slot, chi, o, k2, s2)];
          fma_acc(v, Ex[(long long)ks * cp + ks2] + conj(Ex[(long long)ks2 * cp + ks]), x2);
          const cd z = conj(Ex[(long long)ks * cp + ks2]) * tv * conj(x2);
          q2 += z.re;
        }
      const long long off = elem_off(str[x], d, slot, chi, o, k, s);
      atomicAdd(&tb[off].re, w * v.re);
      atomicAdd(&tb[off].im, w * v.im);
      if (q2 != 0.0)
        for (int j = 0; j < d; ++j) {
          if (j == slot) continue;
          double others = 1.0;
          for (int i = 0; i < d; ++i)
            if (i != j && i != slot) others *= lj[i];
          atomicAdd(&lam_bar[(long long)g.adj_edge[base + j] * chi + bj[j]], q2 * others);
        }
    }
  }
}

// One-site gate adjoint: Gamma' = Y / |Y|, Y[a,o] = sum_i G[o][i] T[a,i]
// (tensor_graph.py:268-275).  Tbar turned in place from post- to pre-gate
// gradients; gbar per gate.
__global__ void __launch_bounds__(256)
    k_apply1_vjp(DevGraph g, const cd* Tpre, const int* nodes, const cd* gates, cd* Tbar, cd* gbar) {
  __shared__ double red[8];
  __shared__ cd gacc[64];
  const int chi = g.chi, p = g.phys;
  const int node = nodes[blockIdx.x];
  const long long rows = ipow(chi, g.deg[node]);
  const cd* t = Tpre + g.toff[node];
  cd* tb = Tbar + g.toff[node];
  const cd* G = gates + (long long)blockIdx.x * p * p;
  double nn = 0.0, ip = 0.0;
  for (long long a = threadIdx.x; a < rows; a += blockDim.x)
    for (int o = 0; o < p; ++o) {
      cd y = mk<double>(0, 0);
      for (int i = 0; i < p; ++i) fma_acc(y, G[o * p + i], t[a * p + i]);
      nn += norm2(y);
      const cd yb = tb[a * p + o];
      ip += y.re * yb.re + y.im * yb.im;
    }
  nn = sqrt(block_sum(nn, red));
  ip = block_sum(ip, red) / nn;
  for (int q = threadIdx.x; q < p * p; q += blockDim.x) gacc[q] = mk<double>(0, 0);
  __syncthreads();
  for (long long a = threadIdx.x; a < rows; a += blockDim.x) {
    cd yb[8], tin[8];
    for (int i = 0; i < p; ++i) tin[i] = t[a * p + i];
    for (int o = 0; o < p; ++o) {
      cd y = mk<double>(0, 0);
      for (int i = 0; i < p; ++i) fma_acc(y, G[o * p + i], tin[i]);
      yb[o] = (1.0 / nn) * (tb[a * p + o] - (ip / nn) * y);
    }
    for (int i = 0; i < p; ++i) {
      cd v = mk<double>(0, 0);
      for (int o = 0; o < p; ++o) fma_acc(v, conj(G[o * p + i]), yb[o]);
      tb[a * p + i] = v;
    }
    for (int o = 0; o < p; ++o)
      for (int i = 0; i < p; ++i) {
        const cd z = yb[o] * conj(tin[i]);
        atomicAdd(&gacc[o * p + i].re, z.re);
        atomicAdd(&gacc[o * p + i].im, z.im);
      }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < p * p; q += blockDim.x) gbar[(long long)blockIdx.x * p * p + q] = gacc[q];
}

// ------------------------------------------------------------- launchers
static inline int vjp_grid(long long items) { return (int)(items < 2048 ? (items > 0 ? items : 1) : 2048); }

long long bp_vjp_ws_bytes(int chi, int phys, int max_deg) {
  const long long S = ipow(chi, max_deg) * phys, cc = (long long)chi * chi;
  return ((max_deg + 2) * S + (3 + max_deg) * cc) * 16 + 256;
}
long long absorb_vjp_ws_bytes(int chi, int phys, int max_deg) {
  const long long S = ipow(chi, max_deg) * phys, cc = (long long)chi * chi;
  return ((max_deg + 2) * S + cc) * 16 + 256;
}
long long edge_svd_vjp_ws_bytes(int chi) { return 14LL * chi * chi * 16 + 256; }
long long rdm2_vjp_ws_bytes(int chi, int phys) { return 4LL * (chi * phys) * (chi * phys) * 16 + 256; }
int vjp_blocks(long long items) { return vjp_grid(items); }

hipError_t launch_bp_node_vjp_fast(const DevGraph& g, const void* T, const void* e2n, const void* obar, void* ebar,
                                   void* Tbar, hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bp_node_vjp_d3c16, dim3(g.n_nodes), dim3(64 * fast64::NW), 0, s, g, (const cd*)T,
                     (const cd*)e2n, (const cd*)obar, (cd*)ebar, (cd*)Tbar);
  return hipGetLastError();
}
hipError_t launch_bp_node_vjp(const DevGraph& g, const void* T, const void* e2n, const void* obar, void* ebar,
                              void* Tbar, void* ws, long long wsb, hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bp_node_vjp, dim3(vjp_grid(g.n_nodes)), dim3(256), 0, s, g, (const cd*)T, (const cd*)e2n,
                     (const cd*)obar, (cd*)ebar, (cd*)Tbar, (char*)ws, wsb);
  return hipGetLastError();
}
hipError_t launch_absorb_vjp(const DevGraph& g, const void* T, const void* X, const void* gbar, void* Tbar, void* Xbar,
                             void* ws, long long wsb, hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_absorb_vjp, dim3(vjp_grid(g.n_nodes)), dim3(256), 0, s, g, (const cd*)T, (const cd*)X,
                     (const cd*)gbar, (cd*)Tbar, (cd*)Xbar, (char*)ws, wsb);
  return hipGetLastError();
}
hipError_t launch_edge_svd_vjp(const DevGraph& g, const void* tape, double eps, double cut, const void* Xbar,
                               const double* lambar, void* n2e_bar, void* ws, long long wsb, hipStream_t s) {
  if (g.n_edges <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_edge_svd_vjp, dim3(vjp_grid(g.n_edges)), dim3(256), 0, s, g, (const char*)tape, eps, cut,
                     (const cd*)Xbar, lambar, (cd*)n2e_bar, (char*)ws, wsb);
  return hipGetLastError();
}
hipError_t launch_symmetric_vjp(const DevGraph& g, const void* gam, const double* lam, double eps, const void* tbar,
                                void* gam_bar, double* lam_bar, hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_symmetric_vjp, dim3(vjp_grid(g.n_nodes)), dim3(256), 0, s, g, (const cd*)gam, lam, eps,
                     (const cd*)tbar, (cd*)gam_bar, lam_bar);
  return hipGetLastError();
}
hipError_t launch_lam2msg_vjp(const DevGraph& g, const void* n2e_bar, const void* e2n_bar, double* lam_bar,
                              hipStream_t s) {
  if (g.n_edges <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lam2msg_vjp, dim3(grid_for((long long)g.n_edges * g.chi, 256)), dim3(256), 0, s, g,
                     (const cd*)n2e_bar, (const cd*)e2n_bar, lam_bar);
  return hipGetLastError();
}
hipError_t launch_rdm1_vjp(const DevGraph& g, const void* T, const double* lam, const int* nodes, int count,
                           const void* rho_bar, void* Tbar, double* lam_bar, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rdm1_vjp, dim3(count), dim3(256), 0, s, g, (const cd*)T, lam, nodes, (const cd*)rho_bar,
                     (cd*)Tbar, lam_bar);
  return hipGetLastError();
}
hipError_t launch_rdm2_vjp(const DevGraph& g, const void* T, const double* lam, const int* edges, int count,
                           const void* rho_bar, void* Tbar, double* lam_bar, void* ws, long long wsb, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rdm2_vjp, dim3(count), dim3(256), 0, s, g, (const cd*)T, lam, edges, (const cd*)rho_bar,
                     (cd*)Tbar, lam_bar, (char*)ws, wsb);
  return hipGetLastError();
}
hipError_t launch_apply1_vjp(const DevGraph& g, const void* Tpre, const int* nodes, int count, const void* gates,
                             void* Tbar, void* gbar, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply1_vjp, dim3(count), dim3(256), 0, s, g, (const cd*)Tpre, nodes, (const cd*)gates,
                     (cd*)Tbar, (cd*)gbar);
  return hipGetLastError();
}

}  // namespace dq

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
__device__ inline int find_slot(const DevGraph& g, int node, int e, int side) {
  const int b = g.adj_ptr[node], d = g.deg[node];
  for (int k = 0; k < d; ++k)
    if (g.adj_edge[b + k] == e && g.adj_side[b + k] == side) return k;
  return -1;
}

// Workspace of one gate block (bytes), for nodes of degree <= max_deg.
__host__ __device__ inline long long su_ws_bytes(int chi, int phys, int max_deg, int max_r) {
  const long long m = ipow(chi, max_deg - 1 > 0 ? max_deg - 1 : 0) * phys;
  const long long n = (long long)chi * max_r;
  const long long k = m < n ? m : n;
  const long long cdn = 2 * m * n      // A_a, A_b
                        + 2 * n        // beta_a, beta_b
                        + k * k        // G
                        + k * k        // U
                        + k * k        // V
                        + 2 * m * chi  // Y_a, Y_b
                        + 16;
  return cdn * (long long)sizeof(cd) + (4 * k + 64) * sizeof(double) + (k + 8) * sizeof(int) + 64;
}

// Build the normalised (others.., s_out) x (b * r + q) matrix of one side
// (column-major) from the node tensor, the lambdas of its other bonds and the
// gate half h[q][s_out][s_in].
template <typename R>
__device__ void su_build(const DevGraph& g, const cx<R>* t, int node, int slot, const double* lam, const cd* h, int r,
                         cd* A, long long m, int n, double* red) {
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const int base = g.adj_ptr[node];
  double acc = 0.0;
  for (long long f = threadIdx.x; f < m * n; f += blockDim.x) {
    const int col = (int)(f / m);
    const long long row = f % m;
    const int b = col / r, q = col % r;
    const int so = (int)(row % p);
    long long orest = row / p;
    // decode the other bonds (last one fastest) and build the tensor index
    long long tidx = 0, mul = p;
    double wt = 1.0;
    for (int j = d - 1; j >= 0; --j) {
      int bj;
      if (j == slot) {
        bj = b;
      } else {
        bj = (int)(orest % chi);
        orest /= chi;
        wt *= lam[(long long)g.adj_edge[base + j] * chi + bj];
      }
      tidx += bj * mul;
      mul *= chi;
    }
    cd v = mk<double>(0, 0);
    if (wt != 0.0)
      for (int si = 0; si < p; ++si) fma_acc(v, h[(q * p + so) * p + si], cvt<double>(t[tidx + si]));
    v = wt * v;
    A[f] = v;
    acc += norm2(v);
  }
  const double nrm = sqrt(block_sum(acc, red));
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  for (long long f = threadIdx.x; f < m * n; f += blockDim.x) A[f] = inv * A[f];
  __syncthreads();
}

// Write Gamma = Y (rows (others.., s), column = new bond) * prod safe_inv(lambda_other)
// into the padded node tensor with the new bond at `slot`.
template <typename R>
__device__ void su_store(const DevGraph& g, cx<R>* t, int node, int slot, const double* lam, double eps, const cd* Y,
                         long long m, int rank) {
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const int base = g.adj_ptr[node];
  const long long S = ipow(chi, d) * p;
  for (long long f = threadIdx.x; f < S; f += blockDim.x) {
    const int s = (int)(f % p);
    long long rr = f / p;
    long long row = 0, mul = p;
    int nb = 0;
    double wt = 1.0;
    for (int j = d - 1; j >= 0; --j) {
      const int bj = (int)(rr % chi);
      rr /= chi;
      if (j == slot) {
        nb = bj;
      } else {
        row += bj * mul;
        mul *= chi;
        wt *= su_safe_inverse(lam[(long long)g.adj_edge[base + j] * chi + bj], eps);
      }
    }
    row += s;
    cd v = mk<double>(0, 0);
    if (nb < rank) v = wt * Y[(long long)nb * m + row];
    t[f] = cvt<R>(v);
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_su(DevGraph g, cx<R>* T, double* lam, const SuGate* gates, const cd* halves, double eps, int max_rank,
         double* err_out, double* ent_out, char* ws, long long wsb) {
  __shared__ double red[8];
  __shared__ int sh_rank;
  const SuGate G = gates[blockIdx.x];
  const int e = G.edge, r = G.rank, chi = g.chi, p = g.phys;
  const int na = g.edge_node[2 * e + G.first], nb = g.edge_node[2 * e + 1 - G.first];
  const int ia = find_slot(g, na, e, G.first), ib = find_slot(g, nb, e, 1 - G.first);
  const int da = g.deg[na], db = g.deg[nb];
  const long long ma = ipow(chi, da - 1) * p, mb = ipow(chi, db - 1) * p;
  const int n = chi * r;
  const int ka = (int)(ma < n ? ma : n), kb = (int)(mb < n ? mb : n);
  const int kk = ka < kb ? ka : kb;
  const long long hstride = (long long)p * p * p * p;  // r <= p^2 rows of (p x p)
  const cd* h1 = halves + (long long)blockIdx.x * 2 * hstride;
  const cd* h2 = h1 + hstride;

  // carve the workspace
  char* w = ws + (long long)blockIdx.x * wsb;
  cd* Aa = (cd*)w;
  cd* Ab = Aa + ma * n;
  cd* ba = Ab + mb * n;
  cd* bb = ba + n;
  const int kmax = ka > kb ? ka : kb;
  cd* Gm = bb + n;
  cd* U = Gm + (long long)kmax * kmax;
  cd* V = U + (long long)kmax * kmax;
  cd* Ya = V + (long long)kmax * kmax;
  cd* Yb = Ya + ma * chi;
  double* sigma = (double*)(Yb + mb * chi);
  double* rot = sigma + kmax + 8;
  int* flag = (int*)(rot + 2 * kmax + 16);
  int* perm = flag + 2;

  cx<R>* ta = T + g.toff[na];
  cx<R>* tb = T + g.toff[nb];
  su_build<R>(g, ta, na, ia, lam, h1, r, Aa, ma, n, red);
  su_build<R>(g, tb, nb, ib, lam, h2, r, Ab, mb, n, red);
  householder_qr(Aa, (int)ma, (int)ma, n, ba, red);
  householder_qr(Ab, (int)mb, (int)mb, n, bb, red);
  // core = R_a diag(lambda_rep) R_b^T, stored column-major as the SVD input
  const double* le = lam + (long long)e * chi;
  const bool tall = ka >= kb;
  for (int f = threadIdx.x; f < ka * kb; f += blockDim.x) {
    const int i = f / kb, j = f % kb;
    cd acc = mk<double>(0, 0);
    for (int c = (i > j ? i : j); c < n; ++c) {
      const cd ra = c == i ? ba[i] : Aa[(long long)c * ma + i];
      const cd rb = c == j ? bb[j] : Ab[(long long)c * mb + j];
      fma_acc(acc, le[c / r] * ra, rb);
    }
    if (tall)
      Gm[(long long)j * ka + i] = acc;
    else
      Gm[(long long)i * kb + j] = acc;
  }
  __syncthreads();
  if (tall)
    jacobi_svd(Gm, ka, ka, kb, U, ka, V, sigma, rot, flag, perm);
  else
    jacobi_svd(Gm, kb, kb, ka, U, kb, V, sigma, rot, flag, perm);
  // threshold, rank, entropy, error
  if (threadIdx.x == 0) {
    double nrm = 0.0;
    for (int j = 0; j < kk; ++j) {
      if (!(sigma[j] > eps)) sigma[j] = 0.0;
      nrm += sigma[j] * sigma[j];
    }
    nrm = sqrt(nrm);
    int rank = kk < max_rank ? kk : max_rank;
    if (rank > chi) rank = chi;
    double ent = 0.0, err = 0.0;
    for (int j = 0; j < kk; ++j) {
      const double x = nrm > 0.0 ? (sigma[j] / nrm) * (sigma[j] / nrm) : 0.0;
      if (x > 0.0) ent -= x * log(x);
      if (j >= rank) err += x;
    }
    err_out[blockIdx.x] = sqrt(err);
    ent_out[blockIdx.x] = ent;
    sh_rank = rank;
  }
  __syncthreads();
  const int rank = sh_rank;
  // Y_a = [U_core[:, :rank]; 0], Y_b = [conj(V_core)[:, :rank]; 0]
  for (long long f = threadIdx.x; f < ma * chi; f += blockDim.x) {
    const int j = (int)(f / ma);
    const long long i = f % ma;
    cd v = mk<double>(0, 0);
    if (i < ka && j < rank) v = tall ? U[(long long)j * ka + i] : conj(V[(long long)perm[j] * ka + i]);
    Ya[f] = v;
  }
  for (long long f = threadIdx.x; f < mb * chi; f += blockDim.x) {
    const int j = (int)(f / mb);
    const long long i = f % mb;
    cd v = mk<double>(0, 0);
    if (i < kb && j < rank) v = tall ? conj(V[(long long)perm[j] * kb + i]) : U[(long long)j * kb + i];
    Yb[f] = v;
  }
  __syncthreads();
  householder_apply_q(Aa, (int)ma, (int)ma, ka, Ya, (int)ma, rank);
  householder_apply_q(Ab, (int)mb, (int)mb, kb, Yb, (int)mb, rank);
  su_store<R>(g, ta, na, ia, lam, eps, Ya, ma, rank);
  su_store<R>(g, tb, nb, ib, lam, eps, Yb, mb, rank);
  __syncthreads();
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
                     const cd* halves, double eps, int max_rank, double* err, double* entropy, void* ws, long long wsb,
                     hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_su<float>, dim3(count), dim3(256), 0, s, g, (cx<float>*)T, lam, gates, halves, eps,
                       max_rank, err, entropy, (char*)ws, wsb);
  else
    hipLaunchKernelGGL(k_su<double>, dim3(count), dim3(256), 0, s, g, (cx<double>*)T, lam, gates, halves, eps,
                       max_rank, err, entropy, (char*)ws, wsb);
  return hipGetLastError();
}

}  // namespace dq

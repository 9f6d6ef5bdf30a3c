// Observables and storage maintenance.
//
// rdm1   one-site density matrix from a node tensor and its incoming messages
//        (reference dequantizer/src/mappings/density_matrix.py:16-37).
// rdm2   two-site density matrix of an edge (no reference counterpart): the
//        shared bond is contracted ket-ket and bra-bra, every other bond
//        carries its incoming message; computed as
//        rho = sum_{k,k'} E_a[(k s_a),(k' s_a')] E_b[(k s_b),(k' s_b')]
//        with E_x = (T_x with messages) contracted with conj(T_x) over the
//        other bonds of x.
// repad  copy node tensors into storage with a larger padded bond dimension.
#include "dq_launch.hpp"

namespace dq {

__host__ __device__ inline long long rdm_ws_bytes(int chi, int phys, int max_deg) {
  const long long S = ipow(chi, max_deg) * phys;
  const long long e = (long long)chi * phys * chi * phys;
  return 2 * S * (long long)sizeof(cx<double>) + 2 * e * (long long)sizeof(cd) + 64;
}

// ket = T with the incoming message applied on every bond except `skip`.
template <typename R>
__device__ const cx<R>* apply_incoming(const DevGraph& g, const cx<R>* t, const cx<R>* e2n, int node, int skip,
                                       cx<R>* w0, cx<R>* w1) {
  const int d = g.deg[node], chi = g.chi, p = g.phys, base = g.adj_ptr[node];
  const long long cc = (long long)chi * chi;
  const cx<R>* src = t;
  int buf = 0;
  for (int j = 0; j < d; ++j) {
    if (j == skip) continue;
    const int e = g.adj_edge[base + j], s = g.adj_side[base + j];
    cx<R>* dst = buf ? w1 : w0;
    mode_product<R, R>(dst, src, e2n + (2LL * e + s) * cc, chi, d, p, j, false);
    __syncthreads();
    src = dst;
    buf ^= 1;
  }
  return src;
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_rdm1(DevGraph g, const cx<R>* T, const cx<R>* e2n, const int* nodes, cd* rho, char* ws, long long wsb) {
  const int node = nodes[blockIdx.x];
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const long long S = ipow(chi, d) * p;
  cx<R>* w0 = (cx<R>*)(ws + (long long)blockIdx.x * wsb);
  cx<R>* w1 = w0 + S;
  const cx<R>* t = T + g.toff[node];
  const cx<R>* ket = apply_incoming<R>(g, t, e2n, node, -1, w0, w1);
  cd* out = (cd*)(ws + (long long)blockIdx.x * wsb + 2 * S * (long long)sizeof(cx<double>));
  const long long rows = S / p;
  // rho[s][s'] = sum over rows of ket[row, s] conj(t[row, s']): every thread
  // sums a strided share of the rows, then one block reduction per entry
  __shared__ double red[8];
  for (int q = 0; q < p * p; ++q) {
    const int s = q / p, sp = q % p;
    double re = 0.0, im = 0.0;
    for (long long a = threadIdx.x; a < rows; a += blockDim.x) {
      const cx<R> x = ket[a * p + s], y = t[a * p + sp];
      re += (double)x.re * y.re + (double)x.im * y.im;
      im += (double)x.im * y.re - (double)x.re * y.im;
    }
    re = block_sum(re, red);
    im = block_sum(im, red);
    if (threadIdx.x == 0) out[q] = mk<double>(re, im);
  }
  __syncthreads();
  cd tr = mk<double>(0, 0);
  for (int s = 0; s < p; ++s) tr = tr + out[s * p + s];
  for (int q = threadIdx.x; q < p * p; q += blockDim.x) rho[(long long)blockIdx.x * p * p + q] = cdiv(out[q], tr);
}

// E[(k s), (k' s')] = sum over the bonds of `node` other than `slot` of
// ket[..k..s] * conj(t[..k'..s']).  Rows of the other bonds are staged through
// LDS in tiles (their offsets decoded once per element), each thread keeping
// up to 16 entries of E in registers (cp = chi * phys <= 64); wider shapes
// take the element loop.
template <typename R>
__device__ void half_env(const DevGraph& g, const cx<R>* ket, const cx<R>* t, int node, int slot, cd* E) {
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const long long inner = ipow(chi, d - 1 - slot) * p;  // stride of bond `slot`
  const long long orows = ipow(chi, d - 1);              // other-bond combinations
  const long long inner_b = inner / p;
  const int cp = chi * p;
  constexpr int kRows = 16, kMaxCp = 64, kPer = kMaxCp * kMaxCp / 256;
  if (cp <= kMaxCp && blockDim.x == 256) {
    __shared__ cd xs[kRows * kMaxCp], ys[kRows * kMaxCp];
    cd acc[kPer];
#pragma unroll
    for (int v = 0; v < kPer; ++v) acc[v] = mk<double>(0, 0);
    for (long long o0 = 0; o0 < orows; o0 += kRows) {
      const int nr = orows - o0 < kRows ? (int)(orows - o0) : kRows;
      for (int f = threadIdx.x; f < nr * cp; f += blockDim.x) {
        const int rr = f / cp, col = f - rr * cp;
        const int k = col / p, s = col - k * p;
        const long long o = o0 + rr;
        const long long ob = o / inner_b, ib = o - ob * inner_b;
        const long long idx = (ob * chi + k) * inner + ib * p + s;
        xs[rr * cp + col] = cvt<double>(ket[idx]);
        ys[rr * cp + col] = cvt<double>(t[idx]);
      }
      __syncthreads();
#pragma unroll
      for (int v = 0; v < kPer; ++v) {
        const int q = threadIdx.x + 256 * v;
        if (q < cp * cp) {
          const int a = q / cp, b = q - a * cp;
          double re = acc[v].re, im = acc[v].im;
          for (int rr = 0; rr < nr; ++rr) {
            const cd x = xs[rr * cp + a], y = ys[rr * cp + b];
            re += x.re * y.re + x.im * y.im;
            im += x.im * y.re - x.re * y.im;
          }
          acc[v] = mk<double>(re, im);
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < kPer; ++v) {
      const int q = threadIdx.x + 256 * v;
      if (q < cp * cp) E[q] = acc[v];
    }
    return;
  }
  for (int q = threadIdx.x; q < cp * cp; q += blockDim.x) {
    const int a = q / cp, b = q % cp;
    const int k = a / p, s = a % p, kp = b / p, sp = b % p;
    double re = 0.0, im = 0.0;
    for (long long o = 0; o < orows; ++o) {
      // o enumerates (outer bonds before slot, inner bonds after slot)
      const long long ob = o / inner_b, ib = o % inner_b;
      const long long i1 = (ob * chi + k) * inner + ib * p + s;
      const long long i2 = (ob * chi + kp) * inner + ib * p + sp;
      const cx<R> x = ket[i1], y = t[i2];
      re += (double)x.re * y.re + (double)x.im * y.im;
      im += (double)x.im * y.re - (double)x.re * y.im;
    }
    E[q] = mk<double>(re, im);
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_rdm2(DevGraph g, const cx<R>* T, const cx<R>* e2n, const int* edges, cd* rho, char* ws, long long wsb) {
  const int e = edges[blockIdx.x];
  const int chi = g.chi, p = g.phys;
  const int a = g.edge_node[2 * e], b = g.edge_node[2 * e + 1];
  int ia = -1, ib = -1;
  for (int k = 0; k < g.deg[a]; ++k)
    if (g.adj_edge[g.adj_ptr[a] + k] == e && g.adj_side[g.adj_ptr[a] + k] == 0) ia = k;
  for (int k = 0; k < g.deg[b]; ++k)
    if (g.adj_edge[g.adj_ptr[b] + k] == e && g.adj_side[g.adj_ptr[b] + k] == 1) ib = k;
  const long long Smax = ipow(chi, g.deg[a] > g.deg[b] ? g.deg[a] : g.deg[b]) * p;
  char* w = ws + (long long)blockIdx.x * wsb;
  cx<R>* w0 = (cx<R>*)w;
  cx<R>* w1 = w0 + Smax;
  const int cp = chi * p;
  cd* Ea = (cd*)(w + 2 * Smax * (long long)sizeof(cx<double>));
  cd* Eb = Ea + (long long)cp * cp;
  const cx<R>* ta = T + g.toff[a];
  const cx<R>* tb = T + g.toff[b];
  const cx<R>* ka = apply_incoming<R>(g, ta, e2n, a, ia, w0, w1);
  half_env<R>(g, ka, ta, a, ia, Ea);
  __syncthreads();
  const cx<R>* kb = apply_incoming<R>(g, tb, e2n, b, ib, w0, w1);
  half_env<R>(g, kb, tb, b, ib, Eb);
  __syncthreads();
  cd* out = rho + (long long)blockIdx.x * p * p * p * p;
  __shared__ cd acc[256];
  const int pp = p * p;
  for (int q = threadIdx.x; q < pp * pp; q += blockDim.x) {
    const int row = q / pp, col = q % pp;
    const int sa = row / p, sb = row % p, sap = col / p, sbp = col % p;
    cd v = mk<double>(0, 0);
    for (int k = 0; k < chi; ++k)
      for (int kp = 0; kp < chi; ++kp)
        fma_acc(v, Ea[(k * p + sa) * cp + kp * p + sap], Eb[(k * p + sb) * cp + kp * p + sbp]);
    acc[q] = v;
  }
  __syncthreads();
  cd tr = mk<double>(0, 0);
  for (int i = 0; i < pp; ++i) tr = tr + acc[i * pp + i];
  for (int q = threadIdx.x; q < pp * pp; q += blockDim.x) out[q] = cdiv(acc[q], tr);
}

template <typename R>
__global__ void k_repad(DevGraph g, const long long* toff_old, int chi_old, const cx<R>* T_old, cx<R>* T_new) {
  const int node = blockIdx.x;
  const int d = g.deg[node], chi = g.chi, p = g.phys;
  const long long S = ipow(chi, d) * p;
  const cx<R>* src = T_old + toff_old[node];
  cx<R>* dst = T_new + g.toff[node];
  for (long long f = threadIdx.x; f < S; f += blockDim.x) {
    const int s = (int)(f % p);
    long long r = f / p, oidx = 0, mul = p;
    bool inside = true;
    for (int j = d - 1; j >= 0; --j) {
      const int bj = (int)(r % chi);
      r /= chi;
      if (bj >= chi_old) inside = false;
      oidx += bj * mul;
      mul *= chi_old;
    }
    dst[f] = inside ? src[oidx + s] : mk<R>(0, 0);
  }
}

// ------------------------------------------------------------- launchers
hipError_t launch_rdm1(int rb, const DevGraph& g, const void* T, const void* e2n, const int* nodes, int count, cd* rho,
                       void* ws, long long wsb, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_rdm1<float>, dim3(count), dim3(256), 0, s, g, (const cx<float>*)T, (const cx<float>*)e2n,
                       nodes, rho, (char*)ws, wsb);
  else
    hipLaunchKernelGGL(k_rdm1<double>, dim3(count), dim3(256), 0, s, g, (const cx<double>*)T,
                       (const cx<double>*)e2n, nodes, rho, (char*)ws, wsb);
  return hipGetLastError();
}

hipError_t launch_rdm2(int rb, const DevGraph& g, const void* T, const void* e2n, const int* edges, int count, cd* rho,
                       void* ws, long long wsb, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_rdm2<float>, dim3(count), dim3(256), 0, s, g, (const cx<float>*)T, (const cx<float>*)e2n,
                       edges, rho, (char*)ws, wsb);
  else
    hipLaunchKernelGGL(k_rdm2<double>, dim3(count), dim3(256), 0, s, g, (const cx<double>*)T,
                       (const cx<double>*)e2n, edges, rho, (char*)ws, wsb);
  return hipGetLastError();
}

hipError_t launch_repad(int rb, const DevGraph& g, const long long* toff_old, int chi_old, const void* T_old,
                        void* T_new, hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_repad<float>, dim3(g.n_nodes), dim3(256), 0, s, g, toff_old, chi_old,
                       (const cx<float>*)T_old, (cx<float>*)T_new);
  else
    hipLaunchKernelGGL(k_repad<double>, dim3(g.n_nodes), dim3(256), 0, s, g, toff_old, chi_old,
                       (const cx<double>*)T_old, (cx<double>*)T_new);
  return hipGetLastError();
}

}  // namespace dq

// BP message passing: node update (pass_through_node), edge update
// (pass_through_edge), diag(lambda) messages and the mean Frobenius residual.
//
// Node update (reference: dequantizer/src/tensor_ops.py:36-60): for every
// bond idx of a node, U = T with the incoming message of every other bond
// applied on the ket side, out[k,l] = sum conj(T)[..k..] U[..l..] over all
// axes but idx, out /= trace(out).
#include "dq_launch.hpp"

namespace dq {

// Generic node update: one block per node of the level, any degree <= kMaxDeg,
// any chi/phys.  Intermediates live in a per-block global workspace.
template <typename R>
__global__ void __launch_bounds__(256)
    k_bp_nodes(DevGraph g, const int* nodes, const cx<R>* T, const cx<R>* n2e_old, const cx<R>* e2n_old,
               cx<R>* n2e_new, const uint8_t* node_src, char* ws, long long ws_bytes, const int* inst_active) {
  const int node = nodes[blockIdx.x];
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node], chi = g.chi, phys = g.phys;
  const long long cc = (long long)chi * chi;
  const long long S = ipow(chi, d) * phys;
  const int base = g.adj_ptr[node];
  const cx<R>* t = T + g.toff[node];
  char* wsb = ws + blockIdx.x * ws_bytes;
  cd* out = (cd*)wsb;
  cx<R>* w0 = (cx<R>*)(wsb + cc * sizeof(cd));
  cx<R>* w1 = w0 + S;
  const cx<R>* msg[kMaxDeg];
  for (int k = 0; k < d; ++k) {
    const int e = g.adj_edge[base + k], s = g.adj_side[base + k];
    const int code = node_src ? node_src[base + k] : 0;
    if (code == 0)
      msg[k] = e2n_old + (2LL * e + s) * cc;
    else
      msg[k] = (code == 1 ? n2e_new : n2e_old) + (2LL * e + 1 - s) * cc;
  }
  for (int idx = 0; idx < d; ++idx) {
    const cx<R>* src = t;
    int buf = 0;
    for (int j = 0; j < d; ++j) {
      if (j == idx) continue;
      cx<R>* dst = buf ? w1 : w0;
      mode_product<R, R>(dst, src, msg[j], chi, d, phys, j, false);
      __syncthreads();
      src = dst;
      buf ^= 1;
    }
    bond_gram<R>(out, t, src, chi, d, phys, idx, nullptr);
    __syncthreads();
    double tr_re = 0.0, tr_im = 0.0;
    for (int k = 0; k < chi; ++k) {
      tr_re += out[k * chi + k].re;
      tr_im += out[k * chi + k].im;
    }
    const cd tr = mk<double>(tr_re, tr_im);
    const int e = g.adj_edge[base + idx], s = g.adj_side[base + idx];
    cx<R>* dst = n2e_new + (2LL * e + s) * cc;
    for (long long q = threadIdx.x; q < cc; q += blockDim.x) dst[q] = cvt<R>(cdiv(out[q], tr));
    __syncthreads();
  }
}

template <typename R>
__global__ void k_bp_edges(DevGraph g, const cx<R>* n2e_new, const cx<R>* n2e_old, cx<R>* e2n_new,
                           const uint8_t* edge_src, const int* inst_active) {
  const long long cc = (long long)g.chi * g.chi;
  const long long total = 2LL * g.n_edges * cc;
  for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < total;
       f += (long long)gridDim.x * blockDim.x) {
    const long long es = f / cc, q = f % cc;
    const int e = (int)(es >> 1), s = (int)(es & 1);
    if (inst_active && !inst_active[g.edge_inst[e]]) continue;
    const int code = edge_src ? edge_src[es] : 1;
    const cx<R>* src = code == 0 ? n2e_new : n2e_old;
    e2n_new[f] = src[(2LL * e + 1 - s) * cc + q];
  }
}

template <typename R>
__global__ void k_lam2msg(DevGraph g, const double* lam, cx<R>* n2e, cx<R>* e2n) {
  const int chi = g.chi;
  const long long cc = (long long)chi * chi;
  const long long total = 2LL * g.n_edges * cc;
  for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < total;
       f += (long long)gridDim.x * blockDim.x) {
    const long long es = f / cc, q = f % cc;
    const int e = (int)(es >> 1);
    const int k = (int)(q / chi), l = (int)(q % chi);
    const cx<R> v = mk<R>(k == l ? (R)lam[(long long)e * chi + k] : (R)0, (R)0);
    n2e[f] = v;
    e2n[f] = v;
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_msg_distance(DevGraph g, const int* inst_edge_ptr, const cx<R>* a0, const cx<R>* a1, const cx<R>* b0,
                   const cx<R>* b1, double* out) {
  __shared__ double red[8];
  const int inst = blockIdx.x;
  const long long cc = (long long)g.chi * g.chi;
  const int e0 = inst_edge_ptr[inst], e1 = inst_edge_ptr[inst + 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  double acc = 0.0;
  // one wave per message: 4 messages per edge (n2e s=0,1; e2n s=0,1)
  for (long long m = (long long)e0 * 4 + wave; m < (long long)e1 * 4; m += nw) {
    const long long e = m >> 2;
    const int kind = (int)(m & 3);
    const cx<R>* A = ((kind >> 1) ? a1 : a0) + (2 * e + (kind & 1)) * cc;
    const cx<R>* B = ((kind >> 1) ? b1 : b0) + (2 * e + (kind & 1)) * cc;
    double s = 0.0;
    for (long long q = lane; q < cc; q += 64) {
      const double dr = (double)A[q].re - (double)B[q].re, di = (double)A[q].im - (double)B[q].im;
      s += dr * dr + di * di;
    }
    s = wave_sum(s);
    if (lane == 0) acc += sqrt(s);
  }
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) out[inst] = e1 > e0 ? tot / (4.0 * (e1 - e0)) : 0.0;
}

__global__ void k_update_active(const double* dist, double acc, int* iters, int max_iters, int* active, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !active[i]) return;
  const int it = iters[i] + 1;
  iters[i] = it;
  active[i] = (dist[i] > acc) && (max_iters < 0 || it < max_iters);
}

// ------------------------------------------------------------- launchers


hipError_t launch_bp_nodes(int rb, const DevGraph& g, const int* nodes, int count, const void* T, const void* n2e_old,
                           const void* e2n_old, void* n2e_new, const uint8_t* node_src, void* ws, long long wsb,
                           const int* inst_active, int min_deg, int max_deg, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (bp_fast_applicable(rb, g, min_deg, max_deg))
    return launch_bp_nodes_fast(g, nodes, count, T, n2e_old, e2n_old, n2e_new, node_src, inst_active, s);
  if (rb == 8 && g.chi == 16 && g.phys == 2 && min_deg == 3 && max_deg == 3)
    return launch_bp_nodes_fast64(g, nodes, count, T, n2e_old, e2n_old, n2e_new, node_src, inst_active, s);
  if (rb == 4)
    hipLaunchKernelGGL(k_bp_nodes<float>, dim3(count), dim3(256), 0, s, g, nodes, (const cx<float>*)T,
                       (const cx<float>*)n2e_old, (const cx<float>*)e2n_old, (cx<float>*)n2e_new, node_src,
                       (char*)ws, wsb, inst_active);
  else
    hipLaunchKernelGGL(k_bp_nodes<double>, dim3(count), dim3(256), 0, s, g, nodes, (const cx<double>*)T,
                       (const cx<double>*)n2e_old, (const cx<double>*)e2n_old, (cx<double>*)n2e_new, node_src,
                       (char*)ws, wsb, inst_active);
  return hipGetLastError();
}

hipError_t launch_bp_edges(int rb, const DevGraph& g, const void* n2e_new, const void* n2e_old, void* e2n_new,
                           const uint8_t* edge_src, const int* inst_active, hipStream_t s) {
  const long long total = 2LL * g.n_edges * g.chi * g.chi;
  if (total == 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_bp_edges<float>, dim3(grid_for(total, 256)), dim3(256), 0, s, g,
                       (const cx<float>*)n2e_new, (const cx<float>*)n2e_old, (cx<float>*)e2n_new, edge_src,
                       inst_active);
  else
    hipLaunchKernelGGL(k_bp_edges<double>, dim3(grid_for(total, 256)), dim3(256), 0, s, g,
                       (const cx<double>*)n2e_new, (const cx<double>*)n2e_old, (cx<double>*)e2n_new, edge_src,
                       inst_active);
  return hipGetLastError();
}

hipError_t launch_lambdas_to_messages(int rb, const DevGraph& g, const double* lam, void* n2e, void* e2n,
                                      hipStream_t s) {
  const long long total = 2LL * g.n_edges * g.chi * g.chi;
  if (total == 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_lam2msg<float>, dim3(grid_for(total, 256)), dim3(256), 0, s, g, lam, (cx<float>*)n2e,
                       (cx<float>*)e2n);
  else
    hipLaunchKernelGGL(k_lam2msg<double>, dim3(grid_for(total, 256)), dim3(256), 0, s, g, lam, (cx<double>*)n2e,
                       (cx<double>*)e2n);
  return hipGetLastError();
}

hipError_t launch_messages_distance(int rb, const DevGraph& g, const int* inst_edge_ptr, int n_inst,
                                    const void* n2e_a, const void* e2n_a, const void* n2e_b, const void* e2n_b,
                                    double* out, hipStream_t s) {
  if (n_inst <= 0) return hipSuccess;
  if (rb == 4)
    hipLaunchKernelGGL(k_msg_distance<float>, dim3(n_inst), dim3(256), 0, s, g, inst_edge_ptr,
                       (const cx<float>*)n2e_a, (const cx<float>*)e2n_a, (const cx<float>*)n2e_b,
                       (const cx<float>*)e2n_b, out);
  else
    hipLaunchKernelGGL(k_msg_distance<double>, dim3(n_inst), dim3(256), 0, s, g, inst_edge_ptr,
                       (const cx<double>*)n2e_a, (const cx<double>*)e2n_a, (const cx<double>*)n2e_b,
                       (const cx<double>*)e2n_b, out);
  return hipGetLastError();
}

hipError_t launch_update_active(const double* dist, double acc, int* iters, int max_iters, int* active, int n,
                                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_update_active, dim3((n + 255) / 256), dim3(256), 0, s, dist, acc, iters, max_iters, active,
                     n);
  return hipGetLastError();
}

}  // namespace dq

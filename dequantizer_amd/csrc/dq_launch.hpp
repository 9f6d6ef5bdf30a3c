// Host-side launchers of the libdq kernels (one per kernel family).  Each
// returns the hipError_t of its launch.  `real_bytes` selects the bulk
// arithmetic: 4 = complex64 (float accumulation), 8 = complex128.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dq_internal.hpp"

namespace dq {

// Grid size of a grid-stride loop over `total` items (capped at 65536 blocks).
// Compute units of the current device (cached per device).
inline int device_cu_count() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cache[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cache[dev] = v;
  }
  return cache[dev];
}

// Kernel variants picked by the autotuner, process-wide per (kernel, shape
// key); -1 = not tuned (the launcher's default).
enum TuneKernel { kTuneBpD3C16 = 0, kTuneSu = 1 };
int tune_get(int kernel, long long key);
void tune_set(int kernel, long long key, int variant);

inline int grid_for(long long total, int block) {
  long long gsz = (total + block - 1) / block;
  if (gsz > 65536) gsz = 65536;
  if (gsz < 1) gsz = 1;
  return (int)gsz;
}

// Per-level BP node update.  node_src codes per adjacency slot:
// 0 = e2n_old[e][side], 1 = n2e_new[e][1-side], 2 = n2e_old[e][1-side].
hipError_t launch_bp_nodes(int real_bytes, const DevGraph& g, const int* nodes, int count, const void* T,
                           const void* n2e_old, const void* e2n_old, void* n2e_new, const uint8_t* node_src,
                           void* ws, long long ws_bytes_per_block, const int* inst_active, int min_deg, int max_deg,
                           hipStream_t s);

// Matrix-core node update for graphs whose nodes all have degree 3, chi 16,
// phys 2, complex64 (dq_bp_fast.hip).
bool bp_fast_applicable(int real_bytes, const DevGraph& g, int min_deg, int max_deg);
hipError_t launch_absorb_fast64(const DevGraph& g, const void* T, const cd* X, void* gammas, const int* inst_active,
                                hipStream_t s);
hipError_t launch_bp_nodes_fast64(const DevGraph& g, const int* nodes, int count, const void* T, const void* n2e_old,
                                  const void* e2n_old, void* n2e_new, const uint8_t* node_src, const int* inst_active,
                                  hipStream_t s);
hipError_t launch_absorb_fast(const DevGraph& g, const void* T, const cd* X, void* gammas, const int* inst_active,
                              hipStream_t s);
// Large node tensors (chi in {16, 32}, chi^max_deg * phys >= 2^15): node
// updates and the absorb as HBM-streaming launches over all nodes at once.
bool big_applicable(const DevGraph& g, int max_deg);
long long big_ws_bytes(const DevGraph& g, int max_deg, int count, int real_bytes);
long long vidal_big_ws_bytes(const DevGraph& g, int max_deg, int real_bytes);
hipError_t launch_vidal_distance_big(int real_bytes, const DevGraph& g, int min_deg, int max_deg, const void* gam,
                                     const double* lam, double eps, const int* all_nodes, const uint8_t* zero_src,
                                     const int* inst_edge_ptr, int n_inst, double* inst_dist, void* ws,
                                     const int* inst_active, hipStream_t s);
hipError_t launch_bp_nodes_big(int real_bytes, const DevGraph& g, int min_deg, int max_deg, const int* nodes,
                               int count, const void* T, const void* n2e_old, const void* e2n_old, void* n2e_new,
                               const uint8_t* node_src, void* ws, const int* inst_active, hipStream_t s);
hipError_t launch_absorb_big(int real_bytes, const DevGraph& g, int max_deg, const int* nodes, int count,
                             const void* T, const cd* X, void* gammas, void* ws, const int* inst_active,
                             hipStream_t s);
// e2n_copy != NULL (plain synchronous schedule only): the edge update
// e2n_copy[e][s] = n2e_old[e][1 - s] is done by the node kernel as well.
hipError_t launch_bp_nodes_fast(const DevGraph& g, const int* nodes, int count, const void* T, const void* n2e_old,
                                const void* e2n_old, void* n2e_new, const uint8_t* node_src, const int* inst_active,
                                hipStream_t s, void* e2n_copy = nullptr);
// The same node update with diag(lam_e^2) messages on every bond (Vidal distance).
hipError_t launch_bp_nodes_lam2_fast(const DevGraph& g, const int* nodes, int count, const void* T,
                                     const double* lam, void* n2e_new, const int* inst_active, hipStream_t s);

// e2n_new[e][s] = (edge_src[2e+s] == 0 ? n2e_new : n2e_old)[e][1-s].
hipError_t launch_bp_edges(int real_bytes, const DevGraph& g, const void* n2e_new, const void* n2e_old, void* e2n_new,
                           const uint8_t* edge_src, const int* inst_active, hipStream_t s);

// n2e[e][s] = e2n[e][s] = diag(lam[e]).
hipError_t launch_lambdas_to_messages(int real_bytes, const DevGraph& g, const double* lam, void* n2e, void* e2n,
                                      hipStream_t s);

// out[i] = sum over the messages of instance i of |A - B|_F, divided by the
// message count (4 per edge): A/B = (n2e, e2n) pairs.
hipError_t launch_messages_distance(int real_bytes, const DevGraph& g, const int* inst_edge_ptr, int n_inst,
                                    const void* n2e_a, const void* e2n_a, const void* n2e_b, const void* e2n_b,
                                    double* out, hipStream_t s);

// Vidal gauge, edge part: orthogonalisers X[e][s] (complex double, chi x chi)
// and lambda[e] from the node->edge messages.
// Per-edge record of the Vidal edge decomposition for the adjoint:
// eigenvectors (row-major, column i = vector i) and eigenvalues of both
// messages, sorted U and V (row-major), singular values, and the scalars
// (singular-value norm, masked norms of U and V).
struct EdgeSvdTape {
  char* base;
  int n;
  __host__ __device__ EdgeSvdTape(char* b, int nn) : base(b), n(nn) {}
  __host__ __device__ cd* Qf() const { return (cd*)base; }
  __host__ __device__ cd* Qb() const { return Qf() + (long long)n * n; }
  __host__ __device__ cd* U() const { return Qb() + (long long)n * n; }
  __host__ __device__ cd* V() const { return U() + (long long)n * n; }
  __host__ __device__ double* wf() const { return (double*)(V() + (long long)n * n); }
  __host__ __device__ double* wb() const { return wf() + n; }
  __host__ __device__ double* sigma() const { return wb() + n; }
  __host__ __device__ double* scal() const { return sigma() + n; }
};
__host__ __device__ inline long long edge_svd_tape_bytes(int chi) {
  const long long n = chi;
  return (4 * n * n * 16 + 3 * n * 8 + 32 + 255) / 256 * 256;
}
hipError_t launch_edge_svd(int real_bytes, const DevGraph& g, const void* n2e, double eps, cd* X, double* lam,
                           void* ws, long long ws_bytes_per_block, const int* inst_active, void* tape, hipStream_t s);

long long edge_svd_ws_bytes_host(int chi);

// Vidal gauge, node part: Gamma = T x_i X_i^T normalised.
hipError_t launch_absorb(int real_bytes, const DevGraph& g, const void* T, const cd* X, void* gammas, void* ws,
                         long long ws_bytes_per_block, const int* inst_active, hipStream_t s);

// Per-node Vidal distance, then per-instance sum / (2 |E_i|).
hipError_t launch_vidal_distance(int real_bytes, const DevGraph& g, const void* gammas, const double* lam, double eps,
                                 double* node_dist, const int* inst_node_ptr, const int* inst_edge_ptr, int n_inst,
                                 double* inst_dist, void* ws, long long ws_bytes_per_block, const int* inst_active,
                                 hipStream_t s);

// Vidal distance on the matrix-core graphs (bp_fast_applicable): diag(lam^2)
// messages through the MFMA node kernel, then one wave per (edge, side).
// ws: vidal_fast_ws_bytes(g); all_nodes = 0..n-1, zero_src = zeros(sum deg).
long long vidal_fast_ws_bytes(const DevGraph& g);
hipError_t launch_vidal_distance_fast(const DevGraph& g, const void* gam, const double* lam, double eps,
                                      const int* all_nodes, const uint8_t* zero_src, const int* inst_edge_ptr,
                                      int n_inst, double* inst_dist, void* ws, const int* inst_active,
                                      hipStream_t s);

// Symmetric gauge: out = Gamma * prod_i sqrt(lam_i > eps) normalised.
hipError_t launch_symmetric_gauge(int real_bytes, const DevGraph& g, const void* gammas, const double* lam, double eps,
                                  void* out, hipStream_t s);

// One-site gates: T[nodes[i]] <- normalise(T . G_i^T), G_i = gates[i] (phys x phys complex double).
hipError_t launch_apply1(int real_bytes, const DevGraph& g, void* T, const int* nodes, int count, const cd* gates,
                         hipStream_t s);

// Simple update for `count` gates on pairwise disjoint edges.
struct SuGate {
  int edge;   // edge index
  int first;  // side of the controlling node in that edge
  int rank;   // operator-Schmidt rank r of the gate
  int pad;
};
// Per-gate record of the simple update for the adjoint (tape != NULL):
// R_a, R_b, U, V (n x n row-major, n = chi * r, slots sized for
// nmax = chi * max_gate_rank), singular values, traces of both
// environments, Cholesky pivots, kept rank.
struct SuTape {
  char* base;
  int nmax;
  __host__ __device__ SuTape(char* b, int nm) : base(b), nmax(nm) {}
  __host__ __device__ cd* Ra() const { return (cd*)base; }
  __host__ __device__ cd* Rb() const { return Ra() + (long long)nmax * nmax; }
  __host__ __device__ cd* U() const { return Rb() + (long long)nmax * nmax; }
  __host__ __device__ cd* V() const { return U() + (long long)nmax * nmax; }
  __host__ __device__ double* sigma() const { return (double*)(V() + (long long)nmax * nmax); }
  __host__ __device__ double* tr() const { return sigma() + nmax; }
  __host__ __device__ int* piv_a() const { return (int*)(tr() + 2); }
  __host__ __device__ int* piv_b() const { return piv_a() + nmax; }
  __host__ __device__ int* meta() const { return piv_b() + nmax; }  // kept rank, n
};
__host__ __device__ inline long long su_tape_bytes(int chi, int max_r) {
  const long long nm = (long long)chi * max_r;
  return (4 * nm * nm * 16 + nm * 8 + 16 + 2 * nm * 4 + 16 + 255) / 256 * 256;
}
hipError_t launch_su(int real_bytes, const DevGraph& g, void* T, double* lam, const SuGate* gates, int count,
                     const cd* halves, int max_gate_rank, double eps, int max_rank, double* err, double* entropy,
                     void* ws, long long wsb, void* tape, hipStream_t s);

// Per-block workspace of the simple update (bytes) and whether the small
// FP64 matrices fit the LDS variant.
long long su_ws_bytes(int chi, int phys, int max_r);
int su_mem_mode(int chi, int phys, int max_r);
int su_launch_mode(int real_bytes, int chi, int phys, int max_r);
long long su_tune_key(int real_bytes, int chi, int phys, int max_r);

// One-site density matrices with messages e2n: rho[i] (phys x phys).
hipError_t launch_rdm1(int real_bytes, const DevGraph& g, const void* T, const void* e2n, const int* nodes, int count,
                       cd* rho, void* ws, long long ws_bytes_per_block, hipStream_t s);

// Two-site density matrices on edges: rho[i] (phys^2 x phys^2), rows (s_side0, s_side1).
hipError_t launch_rdm2(int real_bytes, const DevGraph& g, const void* T, const void* e2n, const int* edges, int count,
                       cd* rho, void* ws, long long ws_bytes_per_block, hipStream_t s);

// Re-pad node tensors from chi_old to the graph's chi (zeros in new slots).
hipError_t launch_repad(int real_bytes, const DevGraph& g_new, const long long* toff_old, int chi_old,
                        const void* T_old, void* T_new, hipStream_t s);

// Gate half (r, out, in) on one unpadded tensor viewed as (lead, bond, trail,
// phys), fused bond b*r+q, then Frobenius normalisation (dq_tensor.hip).
// `partial` holds kTensorParts doubles.
constexpr int kTensorParts = 1024;
hipError_t launch_tensor_gate(int real_bytes, const void* t, long long lead, int bond, long long trail, int p,
                              const cd* half, int r, void* out, double* partial, hipStream_t s);

// Operator-Schmidt split of two-site gates (dq_gates.hip): halves
// [count][2][p^4], ranks[i * rank_stride].  phys 1..4.
hipError_t launch_decompose_gates(const cd* gates, int count, int p, double eps, cd* halves, int* ranks,
                                  int rank_stride, hipStream_t s);

// Adjoint of one simple-update launch (complex128, dq_adjoint.hip).  Tpre,
// lam: state before the launch; tape: the forward's record; Tbar, lambar:
// gradients of the state after the launch, turned in place into those of
// the state before it; hbar: per-gate gradients of the two halves.
hipError_t launch_su_vjp(const DevGraph& g, const void* Tpre, const double* lam, const SuGate* gates, int count,
                         const cd* halves, int max_gate_rank, double eps, double cut, const void* tape, void* Tbar,
                         double* lambar, void* hbar, void* ws, long long wsb, hipStream_t s);
long long su_vjp_ws_bytes_host(int chi, int phys, int max_r, int max_deg);

// Adjoints of the other forward kernels (complex128, dq_adjoint.hip).
long long bp_vjp_ws_bytes(int chi, int phys, int max_deg);
hipError_t launch_absorb_vjp_fast(const DevGraph& g, const void* T, const void* X, const void* gbar, void* Tbar,
                                  void* Xbar, void* ws, hipStream_t s);
hipError_t launch_bp_node_vjp_fast(const DevGraph& g, const void* T, const void* e2n, const void* obar, void* ebar,
                                   void* Tbar, hipStream_t s);
long long absorb_vjp_ws_bytes(int chi, int phys, int max_deg);
long long edge_svd_vjp_ws_bytes(int chi);
long long rdm2_vjp_ws_bytes(int chi, int phys);
int vjp_blocks(long long items);
hipError_t launch_bp_node_vjp(const DevGraph& g, const void* T, const void* e2n, const void* obar, void* ebar,
                              void* Tbar, void* ws, long long wsb, hipStream_t s);
hipError_t launch_absorb_vjp(const DevGraph& g, const void* T, const void* X, const void* gbar, void* Tbar, void* Xbar,
                             void* ws, long long wsb, hipStream_t s);
hipError_t launch_edge_svd_vjp(const DevGraph& g, const void* tape, double eps, double cut, const void* Xbar,
                               const double* lambar, void* n2e_bar, void* ws, long long wsb, hipStream_t s);
hipError_t launch_symmetric_vjp(const DevGraph& g, const void* gam, const double* lam, double eps, const void* tbar,
                                void* gam_bar, double* lam_bar, hipStream_t s);
hipError_t launch_lam2msg_vjp(const DevGraph& g, const void* n2e_bar, const void* e2n_bar, double* lam_bar,
                              hipStream_t s);
hipError_t launch_rdm1_vjp(const DevGraph& g, const void* T, const double* lam, const int* nodes, int count,
                           const void* rho_bar, void* Tbar, double* lam_bar, hipStream_t s);
hipError_t launch_rdm2_vjp(const DevGraph& g, const void* T, const double* lam, const int* edges, int count,
                           const void* rho_bar, void* Tbar, double* lam_bar, void* ws, long long wsb, hipStream_t s);
hipError_t launch_apply1_vjp(const DevGraph& g, const void* Tpre, const int* nodes, int count, const void* gates,
                             void* Tbar, void* gbar, hipStream_t s);

// Exact state vector (complex128, dq_sv.hip); bits are amplitude-index bits.
hipError_t launch_sv_basis(void* psi, int n, unsigned long long idx, hipStream_t st);
hipError_t launch_sv_apply1(void* psi, int n, int b, const void* g, hipStream_t st);
hipError_t launch_sv_apply2(void* psi, int n, int b1, int b2, const void* g, hipStream_t st);
hipError_t launch_sv_diag(void* psi, int n, int K, const int* tb, const void* td, void* work, hipStream_t st);
hipError_t launch_sv_layer1(void* psi, int n, unsigned long long targets, const void* gates, hipStream_t st);
hipError_t launch_sv_diag_layer1(void* psi, int n, int K, const int* tb, const void* td, unsigned long long targets,
                                 const void* gates, void* work, hipStream_t st);
hipError_t launch_sv_rdm1(const void* psi, int n, unsigned long long targets, void* rho, void* work, hipStream_t st);
long long sv_rdm_part_bytes();

// Per-instance convergence bookkeeping of the BP-to-Vidal loop.
hipError_t launch_update_active(const double* dist, double acc, int* iters, int max_iters, int* active, int n_inst,
                                hipStream_t s);

}  // namespace dq

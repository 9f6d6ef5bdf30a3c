/* libdq -- MI355X (gfx950) belief-propagation / Vidal-gauge / simple-update
 * engine for graph tensor networks.  C ABI: plain pointers and sizes.
 *
 * Conventions
 *   - Every compute entry point takes caller-owned DEVICE buffers, a stream
 *     (hipStream_t passed as void*, NULL = default stream) and returns a
 *     status (DQ_OK = 0).  On error, dq_last_error() returns a thread-local
 *     description.  Nothing is synchronised unless stated.
 *   - dtype: DQ_C64 (complex64 storage, float arithmetic for the bulk
 *     contractions) or DQ_C128 (complex128).  Eigen/singular decompositions,
 *     Cholesky factors, lambdas and residuals are always FP64.
 *   - A graph handle fixes the topology, a padded bond dimension chi and a
 *     uniform physical dimension phys.  Node tensors live in one buffer at
 *     offsets from dq_graph_tensor_offsets, row-major (bond_0..bond_{d-1},
 *     phys) with bonds in the node's adjacency order, each padded to chi.
 *     Messages: n2e/e2n buffers of 2*n_edges chi x chi matrices, index
 *     (edge*2 + side) where side is the node's position in the edge; entries
 *     M[bra][ket].  Lambdas: n_edges x chi doubles.
 *   - Several graphs may be packed into one handle as instances (nodes and
 *     edges of an instance contiguous); per-instance results are indexed by
 *     instance.  inst_active (device int[n_inst], may be NULL) masks work of
 *     inactive instances (their outputs are left untouched).
 *   - A handle owns its topology arrays and a scratch workspace; it is not
 *     thread-safe, distinct handles are.  The library never frees caller
 *     memory.
 */
#ifndef DQ_H
#define DQ_H

#ifdef __cplusplus
extern "C" {
#endif

#define DQ_OK 0
#define DQ_ERR_INVALID 1     /* bad argument / shape */
#define DQ_ERR_HIP 2         /* HIP runtime error */
#define DQ_ERR_UNSUPPORTED 3 /* valid request outside what is implemented */

#define DQ_C64 0
#define DQ_C128 1

typedef struct dq_graph dq_graph;

/* Library version (major*10000 + minor*100 + patch). */
int dq_version(void);

/* Description of the last error raised on the calling thread. */
const char* dq_last_error(void);

/* Topology (host arrays, copied):
 *   deg[n_nodes], adj_edge/adj_side[sum deg] (bond slots in adjacency order),
 *   edge_node[2*n_edges] (node at side 0, side 1), node_inst[n_nodes]
 *   (non-decreasing instance ids in [0, n_inst)).
 * Replaces the graph container of dequantizer/src/graph/tensor_graph.py:31-94
 * (TensorGraph.add_node/add_edge; bond order = insertion order, node.py:19-27). */
int dq_graph_create(dq_graph** out, int n_nodes, int n_edges, int chi, int phys, const int* deg,
                    const int* adj_edge, const int* adj_side, const int* edge_node, const int* node_inst,
                    int n_inst);
int dq_graph_destroy(dq_graph* g);

/* Total complex elements of the padded node-tensor buffer; per-node offsets
 * (host long long[n_nodes]). */
long long dq_graph_tensor_elems(const dq_graph* g);
int dq_graph_tensor_offsets(const dq_graph* g, long long* offsets);

/* BP sweep schedule (host arrays, copied).  Nodes are processed level by
 * level; node_src[slot] selects each incoming message: 0 = e2n_in of that
 * slot, 1 = this sweep's n2e_out of the partner node, 2 = n2e_in of the
 * partner; edge_src[2*e+s] selects the source of e2n_out[e][s]: 0 = n2e_out,
 * 1 = n2e_in (of the partner side).  All-zero node_src, all-one edge_src and
 * one level reproduce the synchronous sweep; the asynchronous traversal-order
 * sweep of dequantizer/src/mappings/belief_propagation.py:24-64 is expressed
 * by the levels of its dependency DAG. */
int dq_bp_set_schedule(dq_graph* g, int n_levels, const int* level_ptr, const int* level_nodes,
                       const unsigned char* node_src, const unsigned char* edge_src);

/* One BP sweep: replaces get_belief_propagation_map(...)(tensors, messages)
 * (dequantizer/src/mappings/belief_propagation.py:24-64), node update
 * pass_through_node (tensor_ops.py:36-60), edge update pass_through_edge
 * (tensor_ops.py:70-76).  Output buffers must not alias inputs. */
int dq_bp_sweep(dq_graph* g, int dtype, const void* tensors, const void* n2e_in, const void* e2n_in, void* n2e_out,
                void* e2n_out, const int* inst_active, void* stream);

/* n_sweeps consecutive sweeps (dq_bp_sweep) alternating between message set
 * A (n2e_a, e2n_a: the input) and B: sweep i reads A and writes B for even i,
 * the reverse for odd i, so the result is in B for odd n_sweeps and in A for
 * even.  The launches of the n sweeps are captured once into a HIP graph
 * (keyed by every pointer and the count) and the graph is replayed on
 * `stream`; when `stream` is itself being captured the sweeps are launched
 * directly into the caller's capture.  No counterpart in the reference,
 * whose BP loop runs one jitted sweep per host iteration
 * (dequantizer/src/quantum_emulator/bp_quantum_emulator.py:253-266). */
int dq_bp_sweeps(dq_graph* g, int dtype, const void* tensors, void* n2e_a, void* e2n_a, void* n2e_b, void* e2n_b,
                 int n_sweeps, const int* inst_active, void* stream);

/* Autotuner: times the kernel variants that apply to this handle's shape on
 * synthetic inputs of that shape (median of several runs, HIP events) and
 * keeps the fastest for every later launch of the same shape in this
 * process: the complex64 degree-3 chi=16 node update (4 or 8 waves per
 * block, persistent blocks or one block per node) and the simple-update
 * storage mode (compact LDS at two blocks per CU, or all-LDS at one) for
 * gates up to max_gate_rank.  Shapes already tuned are skipped unless
 * `force`.  A text report of the timings goes to `report` (may be NULL).
 * No counterpart in the reference. */
int dq_autotune(dq_graph* g, int dtype, int max_gate_rank, int force, void* stream, char* report, int report_len);

/* Sets (or with variant -1 clears) a tuned choice by hand, for tests and
 * experiments: kernel 0 = the complex64 degree-3 chi=16 node update (variant
 * bit 0: one block per node, bit 1: 4 waves per block); kernel 1 = the
 * simple-update storage mode for (dtype, chi, phys, max_gate_rank)
 * (1 = all-LDS, 2 = compact). */
int dq_tune_override(int kernel, int dtype, int chi, int phys, int max_gate_rank, int variant);

/* Mean Frobenius distance between two message sets, per instance (device
 * double[n_inst]).  Replaces messages_frob_distance
 * (dequantizer/src/mappings/messages_distance.py:14-22). */
int dq_messages_distance(dq_graph* g, int dtype, const void* n2e_a, const void* e2n_a, const void* n2e_b,
                         const void* e2n_b, double* out, void* stream);

/* n2e = e2n = diag(lambda) on every edge.  Replaces lambdas2messages
 * (dequantizer/src/mappings/symmetric_gauge.py:50-61). */
int dq_lambdas_to_messages(dq_graph* g, int dtype, const double* lam, void* n2e, void* e2n, void* stream);

/* Vidal gauge from converged node->edge messages: gammas (tensor buffer) and
 * lambdas.  Replaces get_vidal_gauge_fixing_map (mappings/vidal_gauge.py:21-63)
 * = edge_svd (tensor_ops.py:108-130) + absorb_by_node (tensor_ops.py:88-93). */
int dq_vidal_gauge(dq_graph* g, int dtype, const void* tensors, const void* n2e, double eps, void* gammas,
                   double* lam, const int* inst_active, void* stream);

/* Distance to the Vidal gauge per instance (device double[n_inst]).  Replaces
 * get_vidal_gauge_distance_map (mappings/vidal_distance.py:17-42) = vidal_dist
 * (tensor_ops.py:143-175). */
int dq_vidal_distance(dq_graph* g, int dtype, const void* gammas, const double* lam, double eps, double* out,
                      const int* inst_active, void* stream);

/* dq_vidal_distance without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments
 * and results as dq_vidal_distance (the same kernels), but the kernels'
 * scratch (bond Grams and per-bond distances) lives in a fixed array of
 * DQ_NOALLOC_SCRATCH_BYTES in device memory, one per GPU, held until the
 * process ends, instead of the handle's workspace.  Graphs whose scratch
 * would exceed it are rejected with DQ_ERR_INVALID (on the complex64 chi=16
 * degree-3 path: 2 * n_edges * 2056 + 256 bytes, about 16,300 edges).  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams or threads) race. */
#define DQ_NOALLOC_SCRATCH_BYTES (64LL << 20)
int dq_vidal_distance_noalloc(dq_graph* g, int dtype, const void* gammas, const double* lam, double eps, double* out,
                              const int* inst_active, void* stream);

/* Symmetric gauge.  Replaces get_symmetric_gauge_fixing_map
 * (mappings/symmetric_gauge.py:17-39) = canonicalize (tensor_ops.py:189-199). */
int dq_symmetric_gauge(dq_graph* g, int dtype, const void* gammas, const double* lam, double eps, void* out,
                       void* stream);

/* One-site gates on `count` nodes (device int nodes[count], device complex128
 * gates[count][phys][phys], out x in).  Replaces TensorGraph.apply1
 * (dequantizer/src/graph/tensor_graph.py:268-275). */
int dq_apply1(dq_graph* g, int dtype, void* tensors, const int* nodes, int count, const void* gates, void* stream);

/* Operator-Schmidt split of `count` two-site gates (device complex128
 * gates[count][phys][phys][phys][phys], axes (out1, out2, in1, in2)): SVD of
 * the (out1 in1) x (out2 in2) matrix in FP64, the r singular values above
 * eps kept in descending order.  halves (device complex128
 * [count][2][phys^4]): half 1 = (r, out1, in1), half 2 = (r, out2, in2), each
 * scaled by sqrt(s), zero beyond r*phys^2.  ranks[i * rank_stride] = r
 * (rank_stride 4 fills the rank field of dq_simple_update's gate records).
 * Replaces _decompose_gate (dequantizer/src/tensor_ops.py:202-218).
 * phys 1..4. */
int dq_decompose_gates(const void* gates, int count, int phys, double eps, void* halves, int* ranks, int rank_stride,
                       void* stream);

/* Simple update of `count` two-site gates on pairwise node-disjoint edges
 * (device int gates[count][4] = {edge, side of the controlling node, gate rank
 * r, 0}; device complex128 halves[count][2][phys^4], each half (r, out, in)
 * from the operator-Schmidt split; max_gate_rank = an upper bound of the r,
 * which selects the kernel variant; a gate whose rank is above it is skipped
 * and gets NaN error and entropy).  Updates tensors (Vidal Gammas) and
 * lambdas in place; per-gate truncation error and entropy (device double).
 * Replaces TensorGraph.apply2_to_vidal_canonical (tensor_graph.py:294-375) =
 * simple_update (tensor_ops.py:318-389) with accuracy None and
 * max_rank = max_rank (capped at chi). */
int dq_simple_update(dq_graph* g, int dtype, void* tensors, double* lam, const int* gates, int count,
                     const void* halves, int max_gate_rank, double eps, int max_rank, double* err, double* entropy,
                     void* stream);

/* Workspace of the single-tensor calls below: device double[DQ_TENSOR_WS]. */
#define DQ_TENSOR_WS 1024

/* One-site gate on one unpadded tensor t[rows][phys] (any bond shape,
 * flattened): out[x][o] = sum_i gate[o][i] t[x][i], then out /= |out|_F.
 * gate: device complex128 [phys][phys]; out must not alias t.  Replaces
 * TensorGraph.apply1 (dequantizer/src/graph/tensor_graph.py:268-275) of the
 * dictionary API. */
int dq_tensor_apply1(int dtype, const void* t, long long rows, int phys, const void* gate, void* out, double* ws,
                     void* stream);

/* Gate half on one unpadded tensor viewed as t[lead][bond][trail][phys]:
 * out[l][b*rank+q][tr][o] = sum_i half[q][o][i] t[l][b][tr][i], then
 * out /= |out|_F.  half: device complex128 [rank][phys][phys]; out has
 * lead*bond*rank*trail*phys elements and must not alias t.  Replaces
 * _apply_half_gate (dequantizer/src/tensor_ops.py:221-237), used by
 * TensorGraph.apply2 (tensor_graph.py:387-446). */
int dq_tensor_apply_half(int dtype, const void* t, long long lead, int bond, long long trail, int phys,
                         const void* half, int rank, void* out, double* ws, void* stream);

/* dq_simple_update with a record for the adjoint: tape is a device buffer
 * of count * dq_su_tape_bytes(g, max_gate_rank) bytes that receives, per
 * gate, the Cholesky factors and pivots of both environments, the core's
 * singular vectors and values, the traces and the kept rank.  Same results
 * as dq_simple_update.  No reference counterpart (the reference has no
 * gradients). */
long long dq_su_tape_bytes(const dq_graph* g, int max_gate_rank);
int dq_simple_update_tape(dq_graph* g, int dtype, void* tensors, double* lam, const int* gates, int count,
                          const void* halves, int max_gate_rank, double eps, int max_rank, double* err,
                          double* entropy, void* tape, void* stream);

/* Adjoint of one dq_simple_update_tape launch (DQ_C128 only).  T_pre, lam_pre:
 * tensors and lambdas before the launch; tape: that launch's record; T_bar,
 * lam_bar (device, same layouts as tensors / lambdas): gradients with
 * respect to the state after the launch on entry, with respect to the state
 * before it on return (only the gates' nodes and bonds change); halves_bar
 * (device complex128 [count][2][phys^4]): gradients of the gate halves.
 * Gradients follow dL = Re sum conj(zbar) dz.  cut in [0, 1): singular
 * values and Cholesky pivots below cut * the largest of their gate are
 * frozen (no gradient through them; 0 = the exact adjoint).  The adjoint of the
 * reference's simple_update (dequantizer/src/tensor_ops.py:318-389), which
 * has none. */
int dq_simple_update_vjp(dq_graph* g, int dtype, const void* T_pre, const double* lam_pre, const int* gates,
                         int count, const void* halves, int max_gate_rank, double eps, double cut, const void* tape,
                         void* T_bar, double* lam_bar, void* halves_bar, void* stream);

/* Vidal gauge with a record for the adjoint: as dq_vidal_gauge (all
 * instances), but the orthogonalisers go to the caller's X (device complex128
 * [2*n_edges][chi][chi], index edge*2+side) and each edge's decomposition is
 * recorded in tape (n_edges * dq_edge_svd_tape_bytes(g) bytes). */
long long dq_edge_svd_tape_bytes(const dq_graph* g);
int dq_vidal_gauge_tape(dq_graph* g, int dtype, const void* tensors, const void* n2e, double eps, void* gammas,
                        double* lam, void* X, void* tape, void* stream);

/* Adjoints (DQ_C128 only; dL = Re sum conj(zbar) dz).  "_bar" outputs marked
 * (acc) are accumulated into, the others overwritten.  None of these has a
 * counterpart in the reference, which has no gradients; each is the adjoint
 * of the forward entry point named.
 *
 * dq_bp_sweep_vjp: one synchronous sweep (dq_bp_sweep, all nodes reading
 * e2n_in): n2e_out_bar -> e2n_in_bar, tensors_bar (acc). */
int dq_bp_sweep_vjp(dq_graph* g, int dtype, const void* tensors, const void* e2n_in, const void* n2e_out_bar,
                    void* e2n_in_bar, void* tensors_bar, void* stream);

/* dq_vidal_gauge_tape: gammas_bar, lam_bar -> tensors_bar (acc), n2e_bar
 * (Hermitian part; both sides of every edge).  X_bar: device scratch of the
 * size of X.  cut in [0, 1): message eigen-directions and singular values
 * below cut * the largest on their edge are frozen (0 = the exact adjoint). */
int dq_vidal_gauge_vjp(dq_graph* g, int dtype, const void* tensors, const void* X, const void* tape, double eps,
                       double cut, const void* gammas_bar, const double* lam_bar, void* X_bar, void* tensors_bar,
                       void* n2e_bar, void* stream);

/* dq_symmetric_gauge: out_bar -> gammas_bar (acc), lam_bar (acc). */
int dq_symmetric_gauge_vjp(dq_graph* g, int dtype, const void* gammas, const double* lam, double eps,
                           const void* out_bar, void* gammas_bar, double* lam_bar, void* stream);

/* dq_lambdas_to_messages: message gradients (either may be NULL) -> lam_bar (acc). */
int dq_lambdas_to_messages_vjp(dq_graph* g, int dtype, const void* n2e_bar, const void* e2n_bar, double* lam_bar,
                               void* stream);

/* dq_rdm1 / dq_rdm2 with the diagonal messages diag(lam) (how the emulator
 * evaluates observables): rho_bar -> tensors_bar (acc), lam_bar (acc). */
int dq_rdm1_vjp(dq_graph* g, int dtype, const void* tensors, const double* lam, const int* nodes, int count,
                const void* rho_bar, void* tensors_bar, double* lam_bar, void* stream);
int dq_rdm2_vjp(dq_graph* g, int dtype, const void* tensors, const double* lam, const int* edges, int count,
                const void* rho_bar, void* tensors_bar, double* lam_bar, void* stream);

/* dq_apply1: tensors_pre = tensors before the gates; tensors_bar turned in
 * place from the gradient after to the gradient before; gates_bar
 * [count][phys][phys]. */
int dq_apply1_vjp(dq_graph* g, int dtype, const void* tensors_pre, const int* nodes, int count, const void* gates,
                  void* tensors_bar, void* gates_bar, void* stream);

/* One-site density matrices (device complex128 rho[count][phys][phys]) from
 * tensors and edge->node messages.  Replaces get_one_side_density_matrix
 * (dequantizer/src/mappings/density_matrix.py:16-37). */
int dq_rdm1(dq_graph* g, int dtype, const void* tensors, const void* e2n, const int* nodes, int count, void* rho,
            void* stream);

/* Two-site density matrices of edges (device complex128
 * rho[count][phys^2][phys^2], rows (s_side0, s_side1)).  No reference
 * counterpart (the reference exports only the one-site matrix,
 * dequantizer/src/mappings/__init__.py:1-6). */
int dq_rdm2(dq_graph* g, int dtype, const void* tensors, const void* e2n, const int* edges, int count, void* rho,
            void* stream);

/* Convergence bookkeeping of the BP-to-Vidal loop
 * (dequantizer/src/quantum_emulator/bp_quantum_emulator.py:253-266): for each
 * active instance, iters += 1 and active = dist > acc && iters < max_iters
 * (max_iters < 0: no cap).  All device arrays. */
int dq_update_active(const double* dist, double acc, int* iters, int max_iters, int* active, int n_inst,
                     void* stream);

/* Copy node tensors of handle `old` (smaller chi) into the padded layout of
 * `g` (same topology, chi >= old chi); new slots are zero. */
int dq_repad(const dq_graph* old, dq_graph* g, int dtype, const void* tensors_old, void* tensors_new, void* stream);

/* Exact state vector: device complex128 psi[2^n_qubits], 1 <= n_qubits <= 40
 * (the caller allocates it).  Bit b of an amplitude's index is one qubit; the
 * reference's qubit at position pos (its row-major reshape to n axes of 2,
 * dequantizer/src/quantum_emulator/exact_quantum_emulator.py:67-76) is bit
 * n_qubits - 1 - pos.  Gates are device complex128, row-major [out][in].
 * Replaces ExactQuantumEmulator (exact_quantum_emulator.py:11-147) and the
 * state object of the external `qem` package that the exact annealer drives
 * (examples/quantum_annealing/quantum_annealing/src/exact_quantum_annealer.py:84-122). */

/* Bytes of the `work` buffer of dq_sv_apply_diagonal and dq_sv_rdm1. */
long long dq_sv_work_bytes(void);

/* psi = the basis state `index` (set_to_product_state,
 * exact_quantum_emulator.py:37-49). */
int dq_sv_set_basis(void* psi, int n_qubits, long long index, void* stream);

/* 2x2 gate on one bit (apply_q1, exact_quantum_emulator.py:87-103). */
int dq_sv_apply1(void* psi, int n_qubits, int bit, const void* gate, void* stream);

/* 4x4 gate [o1][o2][i1][i2] with index 1 on bit1 and index 2 on bit2
 * (apply_q2, exact_quantum_emulator.py:51-85). */
int dq_sv_apply2(void* psi, int n_qubits, int bit1, int bit2, const void* gate, void* stream);

/* n_terms diagonal two-bit gates in one pass over psi: term k acts on bits
 * term_bits[2k], term_bits[2k+1] (distinct, < n_qubits; device int) with
 * diagonal term_diag[4k + 2 x1 + x2] (device complex128).  The coupling layer
 * of the annealers (exact_quantum_annealer.py:97-100, diagonal gates of
 * :33-58) as one launch. */
int dq_sv_apply_diagonal(void* psi, int n_qubits, int n_terms, const int* term_bits, const void* term_diag,
                         void* work, void* stream);

/* gates[b] (device complex128 [n_qubits][2][2]) applied to every bit b set in
 * target_mask: single-qubit layers (the mixing layer and the initial
 * Hadamards of exact_quantum_annealer.py:92-102), 12 bits per pass over psi. */
int dq_sv_apply1_layer(void* psi, int n_qubits, long long target_mask, const void* gates, void* stream);

/* dq_sv_apply_diagonal followed by dq_sv_apply1_layer: one annealing layer
 * (the coupling gates, then the mixing gates; exact_quantum_annealer.py:97-102).
 * With n_qubits >= 12 and n_terms <= 64 the diagonal terms are applied in the
 * first pass of the single-qubit layer, saving one pass over psi. */
int dq_sv_apply_diagonal_layer1(void* psi, int n_qubits, int n_terms, const int* term_bits, const void* term_diag,
                                long long target_mask, const void* gates, void* work, void* stream);

/* One-qubit density matrices rho[b] (device complex128 [n_qubits][2][2],
 * rows s, columns s') of every bit b set in target_mask; other entries are
 * left untouched.  dens_q1, exact_quantum_emulator.py:105-124. */
int dq_sv_rdm1(const void* psi, int n_qubits, long long target_mask, void* rho, void* work, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DQ_H */

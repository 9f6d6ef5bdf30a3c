// C ABI of libdq: graph handles, workspace management, argument checks and
// dispatch to the kernel launchers.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/dq.h"
#include "dq_launch.hpp"

using namespace dq;

static thread_local std::string g_last_error;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                             \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) return fail(DQ_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                                      __FILE__, __LINE__);                                        \
  } while (0)

struct dq_graph {
  DevGraph dev;
  int max_deg = 0;
  int min_deg = 0;
  int n_inst = 0;
  long long tensor_elems = 0;
  std::vector<long long> toff_host;
  // device topology
  int *d_deg = nullptr, *d_adj_ptr = nullptr, *d_adj_edge = nullptr, *d_adj_side = nullptr;
  long long* d_toff = nullptr;
  int *d_edge_node = nullptr, *d_node_inst = nullptr, *d_edge_inst = nullptr;
  int *d_inst_node_ptr = nullptr, *d_inst_edge_ptr = nullptr;
  // BP schedule
  int n_levels = 0;
  std::vector<int> level_ptr, level_nodes;
  int* d_level_nodes = nullptr;
  uint8_t *d_node_src = nullptr, *d_edge_src = nullptr;
  int max_level = 0;
  // one level, every node reading the incoming e2n and every edge passing
  // the previous n2e (the plain synchronous sweep)
  bool plain_sync = false;
  // every node once with all sources = incoming e2n (one synchronous level)
  int* d_all_nodes = nullptr;
  uint8_t* d_zero_src = nullptr;
  // captured multi-sweep graphs (dq_bp_sweeps), keyed by every pointer and
  // count they bake in; captured on a private stream, launched on the caller's
  struct SweepGraph {
    std::vector<const void*> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<SweepGraph> sweep_graphs;
  hipStream_t capture_stream = nullptr;
  // owned scratch
  void* ws = nullptr;
  size_t ws_bytes = 0;
  cd* X = nullptr;             // orthogonalisers 2E chi^2
  double* node_dist = nullptr;  // n_nodes
};

// Captured sweep graphs bake in the schedule arrays and the workspace: any
// change to either makes them stale.
static void drop_sweep_graphs(dq_graph* g) {
  for (auto& sg : g->sweep_graphs)
    if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
  g->sweep_graphs.clear();
}

static int ws_ensure(dq_graph* g, size_t bytes) {
  if (bytes <= g->ws_bytes) return DQ_OK;
  drop_sweep_graphs(g);
  if (g->ws) HIP_TRY(hipFree(g->ws));
  g->ws = nullptr;
  g->ws_bytes = 0;
  HIP_TRY(hipMalloc(&g->ws, bytes));
  g->ws_bytes = bytes;
  return DQ_OK;
}

template <typename T>
static int upload(T** dst, const T* src, size_t n) {
  if (n == 0) n = 1;
  HIP_TRY(hipMalloc((void**)dst, n * sizeof(T)));
  if (src) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return DQ_OK;
}

static int check_dtype(int dtype, int* rb) {
  if (dtype == DQ_C64) {
    *rb = 4;
    return DQ_OK;
  }
  if (dtype == DQ_C128) {
    *rb = 8;
    return DQ_OK;
  }
  return fail(DQ_ERR_INVALID, "unknown dtype %d", dtype);
}

static inline hipStream_t S(void* s) { return (hipStream_t)s; }

#define RET_LAUNCH(expr)                                                                              \
  do {                                                                                                \
    hipError_t _e = (expr);                                                                           \
    if (_e != hipSuccess) return fail(DQ_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(_e)); \
  } while (0)

extern "C" {

int dq_version(void) { return 100; }

const char* dq_last_error(void) { return g_last_error.c_str(); }

int dq_graph_create(dq_graph** out, int n_nodes, int n_edges, int chi, int phys, const int* deg,
                    const int* adj_edge, const int* adj_side, const int* edge_node, const int* node_inst,
                    int n_inst) {
  if (!out) return fail(DQ_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (n_nodes < 0 || n_edges < 0 || chi < 1 || phys < 1 || phys > 4 || n_inst < 1)
    return fail(DQ_ERR_INVALID, "bad sizes: n_nodes=%d n_edges=%d chi=%d phys=%d n_inst=%d", n_nodes, n_edges, chi,
                phys, n_inst);
  if (n_nodes > 0 && (!deg || !node_inst)) return fail(DQ_ERR_INVALID, "deg/node_inst NULL");
  if (n_edges > 0 && !edge_node) return fail(DQ_ERR_INVALID, "edge_node NULL");
  std::vector<int> adj_ptr(n_nodes + 1, 0);
  int max_deg = 0;
  for (int n = 0; n < n_nodes; ++n) {
    if (deg[n] < 0 || deg[n] > kMaxDeg) return fail(DQ_ERR_UNSUPPORTED, "node %d degree %d > %d", n, deg[n], kMaxDeg);
    adj_ptr[n + 1] = adj_ptr[n] + deg[n];
    if (deg[n] > max_deg) max_deg = deg[n];
  }
  const int nadj = adj_ptr[n_nodes];
  if (nadj != 2 * n_edges) return fail(DQ_ERR_UNSUPPORTED, "sum of degrees %d != 2 * edges (%d): hyperedges?", nadj,
                                       2 * n_edges);
  std::vector<int> seen(2 * (size_t)n_edges, 0);
  for (int n = 0; n < n_nodes; ++n) {
    if (node_inst[n] < 0 || node_inst[n] >= n_inst || (n > 0 && node_inst[n] < node_inst[n - 1]))
      return fail(DQ_ERR_INVALID, "node_inst must be non-decreasing in [0, n_inst)");
    for (int k = adj_ptr[n]; k < adj_ptr[n + 1]; ++k) {
      const int e = adj_edge[k], s = adj_side[k];
      if (e < 0 || e >= n_edges || s < 0 || s > 1) return fail(DQ_ERR_INVALID, "bad adjacency slot %d", k);
      if (edge_node[2 * e + s] != n) return fail(DQ_ERR_INVALID, "adjacency slot %d disagrees with edge_node", k);
      if (seen[2 * e + s]++) return fail(DQ_ERR_INVALID, "edge %d side %d listed twice", e, s);
    }
  }
  std::vector<int> edge_inst(n_edges);
  for (int e = 0; e < n_edges; ++e) {
    const int a = edge_node[2 * e], b = edge_node[2 * e + 1];
    if (node_inst[a] != node_inst[b]) return fail(DQ_ERR_INVALID, "edge %d joins two instances", e);
    edge_inst[e] = node_inst[a];
    if (e > 0 && edge_inst[e] < edge_inst[e - 1]) return fail(DQ_ERR_INVALID, "edges must be grouped by instance");
  }
  std::vector<int> inode(n_inst + 1, 0), iedge(n_inst + 1, 0);
  for (int n = 0; n < n_nodes; ++n) inode[node_inst[n] + 1]++;
  for (int e = 0; e < n_edges; ++e) iedge[edge_inst[e] + 1]++;
  for (int i = 0; i < n_inst; ++i) {
    inode[i + 1] += inode[i];
    iedge[i + 1] += iedge[i];
  }
  dq_graph* g = new dq_graph();
  g->max_deg = max_deg;
  g->min_deg = max_deg;
  for (int n = 0; n < n_nodes; ++n)
    if (deg[n] < g->min_deg) g->min_deg = deg[n];
  g->n_inst = n_inst;
  g->toff_host.resize(n_nodes);
  long long off = 0;
  for (int n = 0; n < n_nodes; ++n) {
    g->toff_host[n] = off;
    off += ipow(chi, deg[n]) * phys;
  }
  g->tensor_elems = off;
  int rc;
#define UP(dst, src, n)                   \
  if ((rc = upload(dst, src, n)) != DQ_OK) { \
    dq_graph_destroy(g);                 \
    return rc;                           \
  }
  UP(&g->d_deg, deg, n_nodes);
  UP(&g->d_adj_ptr, adj_ptr.data(), n_nodes + 1);
  UP(&g->d_adj_edge, adj_edge, nadj);
  UP(&g->d_adj_side, adj_side, nadj);
  UP(&g->d_toff, g->toff_host.data(), n_nodes);
  UP(&g->d_edge_node, edge_node, 2 * (size_t)n_edges);
  UP(&g->d_node_inst, node_inst, n_nodes);
  UP(&g->d_edge_inst, edge_inst.data(), n_edges);
  UP(&g->d_inst_node_ptr, inode.data(), n_inst + 1);
  UP(&g->d_inst_edge_ptr, iedge.data(), n_inst + 1);
  UP(&g->X, (const cd*)nullptr, 2 * (size_t)n_edges * chi * chi);
  UP(&g->node_dist, (const double*)nullptr, n_nodes);
#undef UP
  DevGraph& d = g->dev;
  d.n_nodes = n_nodes;
  d.n_edges = n_edges;
  d.chi = chi;
  d.phys = phys;
  d.deg = g->d_deg;
  d.adj_ptr = g->d_adj_ptr;
  d.adj_edge = g->d_adj_edge;
  d.adj_side = g->d_adj_side;
  d.toff = g->d_toff;
  d.edge_node = g->d_edge_node;
  d.node_inst = g->d_node_inst;
  d.edge_inst = g->d_edge_inst;
  // default schedule: synchronous, one level with every node
  std::vector<int> lp = {0, n_nodes}, ln(n_nodes);
  for (int n = 0; n < n_nodes; ++n) ln[n] = n;
  std::vector<unsigned char> ns(nadj, 0), es(2 * (size_t)n_edges, 1);
  rc = dq_bp_set_schedule(g, 1, lp.data(), ln.data(), ns.data(), es.data());
  if (rc == DQ_OK) rc = upload(&g->d_all_nodes, ln.data(), (size_t)n_nodes);
  if (rc == DQ_OK) rc = upload(&g->d_zero_src, ns.data(), (size_t)nadj);
  if (rc != DQ_OK) {
    dq_graph_destroy(g);
    return rc;
  }
  *out = g;
  return DQ_OK;
}

int dq_graph_destroy(dq_graph* g) {
  if (!g) return DQ_OK;
  drop_sweep_graphs(g);
  if (g->capture_stream) (void)hipStreamDestroy(g->capture_stream);
  void* ptrs[] = {g->d_deg,       g->d_adj_ptr,       g->d_adj_edge,      g->d_adj_side,   g->d_toff,
                  g->d_edge_node, g->d_node_inst,     g->d_edge_inst,     g->d_inst_node_ptr,
                  g->d_inst_edge_ptr, g->d_level_nodes, g->d_node_src, g->d_edge_src, g->ws, g->X, g->node_dist,
                  g->d_all_nodes, g->d_zero_src};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete g;
  return DQ_OK;
}

long long dq_graph_tensor_elems(const dq_graph* g) { return g ? g->tensor_elems : -1; }

int dq_graph_tensor_offsets(const dq_graph* g, long long* offsets) {
  if (!g || !offsets) return fail(DQ_ERR_INVALID, "NULL argument");
  memcpy(offsets, g->toff_host.data(), g->toff_host.size() * sizeof(long long));
  return DQ_OK;
}

int dq_bp_set_schedule(dq_graph* g, int n_levels, const int* level_ptr, const int* level_nodes,
                       const unsigned char* node_src, const unsigned char* edge_src) {
  if (!g || n_levels < 1 || !level_ptr || !level_nodes || !node_src || !edge_src)
    return fail(DQ_ERR_INVALID, "bad schedule arguments");
  const int nn = g->dev.n_nodes;
  if (level_ptr[0] != 0 || level_ptr[n_levels] != nn)
    return fail(DQ_ERR_INVALID, "schedule must list every node exactly once");
  std::vector<int> count(nn, 0);
  int max_level = 0;
  for (int l = 0; l < n_levels; ++l) {
    if (level_ptr[l + 1] < level_ptr[l]) return fail(DQ_ERR_INVALID, "level_ptr not monotone");
    const int sz = level_ptr[l + 1] - level_ptr[l];
    if (sz > max_level) max_level = sz;
  }
  for (int i = 0; i < nn; ++i) {
    if (level_nodes[i] < 0 || level_nodes[i] >= nn || count[level_nodes[i]]++)
      return fail(DQ_ERR_INVALID, "schedule must list every node exactly once");
  }
  const int nadj = 2 * g->dev.n_edges;
  for (int i = 0; i < nadj; ++i)
    if (node_src[i] > 2 || edge_src[i] > 1) return fail(DQ_ERR_INVALID, "bad source code");
  // graphs replayed after this call must not read the freed schedule arrays
  drop_sweep_graphs(g);
  if (g->d_level_nodes) HIP_TRY(hipFree(g->d_level_nodes));
  if (g->d_node_src) HIP_TRY(hipFree(g->d_node_src));
  if (g->d_edge_src) HIP_TRY(hipFree(g->d_edge_src));
  g->d_level_nodes = nullptr;
  g->d_node_src = g->d_edge_src = nullptr;
  int rc;
  if ((rc = upload(&g->d_level_nodes, level_nodes, nn)) != DQ_OK) return rc;
  if ((rc = upload(&g->d_node_src, node_src, nadj)) != DQ_OK) return rc;
  if ((rc = upload(&g->d_edge_src, edge_src, nadj)) != DQ_OK) return rc;
  g->n_levels = n_levels;
  g->level_ptr.assign(level_ptr, level_ptr + n_levels + 1);
  g->max_level = max_level;
  bool plain = n_levels == 1;
  for (int i = 0; i < nadj && plain; ++i) plain = node_src[i] == 0 && edge_src[i] == 1;
  g->plain_sync = plain;
  return DQ_OK;
}

// complex128 degree-3 chi=16 phys=2 graphs: the FP64 matrix-core kernels
static bool fast64_applicable(int rb, const dq_graph* g) {
  return rb == 8 && g->dev.chi == 16 && g->dev.phys == 2 && g->min_deg == 3 && g->max_deg == 3;
}

static long long node_ws_bytes(const dq_graph* g, int rb) {
  const long long S = ipow(g->dev.chi, g->max_deg) * g->dev.phys;
  return 2 * S * 2 * rb + (long long)g->dev.chi * g->dev.chi * (long long)sizeof(cd) + 256;
}

// Workspace for one sweep of this handle (grown before any launch, so that
// launches can be captured).
static int bp_sweep_ws(dq_graph* g, int rb) {
  if (big_applicable(g->dev, g->max_deg))
    return ws_ensure(g, (size_t)big_ws_bytes(g->dev, g->max_deg, g->max_level > 0 ? g->max_level : 1, rb));
  const long long wsb = (node_ws_bytes(g, rb) + 255) / 256 * 256;
  return ws_ensure(g, (size_t)wsb * (g->max_level > 0 ? g->max_level : 1));
}

static hipError_t bp_sweep_launch(dq_graph* g, int rb, const void* T, const void* n2e_in, const void* e2n_in,
                                  void* n2e_out, void* e2n_out, const int* inst_active, hipStream_t s) {
  hipError_t e = hipSuccess;
  const bool big = big_applicable(g->dev, g->max_deg);
  if (!big && g->plain_sync && bp_fast_applicable(rb, g->dev, g->min_deg, g->max_deg))
    // the matrix-core node kernel also does the synchronous edge update
    return launch_bp_nodes_fast(g->dev, g->d_level_nodes, g->level_ptr[1], T, n2e_in, e2n_in, n2e_out,
                                g->d_node_src, inst_active, s, e2n_out);
  const long long wsb = (node_ws_bytes(g, rb) + 255) / 256 * 256;
  for (int l = 0; l < g->n_levels && e == hipSuccess; ++l) {
    const int b = g->level_ptr[l], cnt = g->level_ptr[l + 1] - b;
    e = big ? launch_bp_nodes_big(rb, g->dev, g->min_deg, g->max_deg, g->d_level_nodes + b, cnt, T, n2e_in, e2n_in,
                                  n2e_out, g->d_node_src, g->ws, inst_active, s)
            : launch_bp_nodes(rb, g->dev, g->d_level_nodes + b, cnt, T, n2e_in, e2n_in, n2e_out, g->d_node_src, g->ws,
                              wsb, inst_active, g->min_deg, g->max_deg, s);
  }
  if (e == hipSuccess) e = launch_bp_edges(rb, g->dev, n2e_out, n2e_in, e2n_out, g->d_edge_src, inst_active, s);
  return e;
}

int dq_bp_sweep(dq_graph* g, int dtype, const void* T, const void* n2e_in, const void* e2n_in, void* n2e_out,
                void* e2n_out, const int* inst_active, void* stream) {
  int rb, rc;
  if (!g) return fail(DQ_ERR_INVALID, "NULL graph");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (n2e_out == n2e_in || e2n_out == e2n_in || n2e_out == e2n_in || e2n_out == n2e_in)
    return fail(DQ_ERR_INVALID, "bp sweep outputs must not alias inputs");
  if ((rc = bp_sweep_ws(g, rb)) != DQ_OK) return rc;
  RET_LAUNCH(bp_sweep_launch(g, rb, T, n2e_in, e2n_in, n2e_out, e2n_out, inst_active, S(stream)));
  return DQ_OK;
}

static hipError_t bp_sweeps_direct(dq_graph* g, int rb, const void* T, void* const* buf, int n_sweeps,
                                   const int* inst_active, hipStream_t s) {
  hipError_t e = hipSuccess;
  for (int i = 0; i < n_sweeps && e == hipSuccess; ++i) {
    const int c = 2 * (i & 1), o = 2 - c;
    e = bp_sweep_launch(g, rb, T, buf[c], buf[c + 1], buf[o], buf[o + 1], inst_active, s);
  }
  return e;
}

int dq_bp_sweeps(dq_graph* g, int dtype, const void* T, void* n2e_a, void* e2n_a, void* n2e_b, void* e2n_b,
                 int n_sweeps, const int* inst_active, void* stream) {
  int rb, rc;
  if (!g || !T || !n2e_a || !e2n_a || !n2e_b || !e2n_b || n_sweeps < 0) return fail(DQ_ERR_INVALID, "bad argument");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (n2e_a == n2e_b || e2n_a == e2n_b || n2e_a == e2n_b || e2n_a == n2e_b || n2e_a == e2n_a || n2e_b == e2n_b)
    return fail(DQ_ERR_INVALID, "the two message sets must be distinct buffers");
  if (n_sweeps == 0) return DQ_OK;
  if ((rc = bp_sweep_ws(g, rb)) != DQ_OK) return rc;
  void* const buf[4] = {n2e_a, e2n_a, n2e_b, e2n_b};
  hipStream_t s = S(stream);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    // the caller is capturing this stream: the launches join its graph
    RET_LAUNCH(bp_sweeps_direct(g, rb, T, buf, n_sweeps, inst_active, s));
    return DQ_OK;
  }
  const std::vector<const void*> key = {T, n2e_a, e2n_a, n2e_b, e2n_b, inst_active, g->ws,
                                        (const void*)(intptr_t)n_sweeps, (const void*)(intptr_t)rb};
  hipGraphExec_t exec = nullptr;
  for (auto& sg : g->sweep_graphs)
    if (sg.key == key) exec = sg.exec;
  if (!exec) {
    if (!g->capture_stream) HIP_TRY(hipStreamCreateWithFlags(&g->capture_stream, hipStreamNonBlocking));
    (void)device_cu_count();  // device queries before the capture, not inside it
    HIP_TRY(hipStreamBeginCapture(g->capture_stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = bp_sweeps_direct(g, rb, T, buf, n_sweeps, inst_active, g->capture_stream);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(g->capture_stream, &graph);
    if (le != hipSuccess || ce != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(DQ_ERR_HIP, "sweep graph capture failed: %s", hipGetErrorString(le != hipSuccess ? le : ce));
    }
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(DQ_ERR_HIP, "sweep graph instantiation failed: %s", hipGetErrorString(ie));
    if (g->sweep_graphs.size() >= 4) {
      (void)hipGraphExecDestroy(g->sweep_graphs.front().exec);
      g->sweep_graphs.erase(g->sweep_graphs.begin());
    }
    g->sweep_graphs.push_back({key, exec});
  }
  RET_LAUNCH(hipGraphLaunch(exec, s));
  return DQ_OK;
}

int dq_messages_distance(dq_graph* g, int dtype, const void* n2e_a, const void* e2n_a, const void* n2e_b,
                         const void* e2n_b, double* out, void* stream) {
  int rb, rc;
  if (!g || !out) return fail(DQ_ERR_INVALID, "NULL argument");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  RET_LAUNCH(launch_messages_distance(rb, g->dev, g->d_inst_edge_ptr, g->n_inst, n2e_a, e2n_a, n2e_b, e2n_b, out,
                                      S(stream)));
  return DQ_OK;
}

int dq_lambdas_to_messages(dq_graph* g, int dtype, const double* lam, void* n2e, void* e2n, void* stream) {
  int rb, rc;
  if (!g) return fail(DQ_ERR_INVALID, "NULL graph");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  RET_LAUNCH(launch_lambdas_to_messages(rb, g->dev, lam, n2e, e2n, S(stream)));
  return DQ_OK;
}

int dq_vidal_gauge(dq_graph* g, int dtype, const void* T, const void* n2e, double eps, void* gammas, double* lam,
                   const int* inst_active, void* stream) {
  int rb, rc;
  if (!g) return fail(DQ_ERR_INVALID, "NULL graph");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (gammas == T) return fail(DQ_ERR_INVALID, "gammas must not alias tensors");
  const long long wse = (edge_svd_ws_bytes_host(g->dev.chi) + 255) / 256 * 256;
  const long long wsn = (node_ws_bytes(g, rb) + 255) / 256 * 256;
  const size_t need_e = (size_t)wse * (g->dev.n_edges > 0 ? g->dev.n_edges : 1);
  const size_t need_n = (size_t)wsn * (g->dev.n_nodes > 0 ? g->dev.n_nodes : 1);
  const bool big = big_applicable(g->dev, g->max_deg);
  const size_t need_b = big ? (size_t)big_ws_bytes(g->dev, g->max_deg, g->dev.n_nodes, rb) : 0;
  if ((rc = ws_ensure(g, need_e > need_n ? (need_e > need_b ? need_e : need_b) : (need_n > need_b ? need_n : need_b))) !=
      DQ_OK)
    return rc;
  RET_LAUNCH(launch_edge_svd(rb, g->dev, n2e, eps, g->X, lam, g->ws, wse, inst_active, nullptr, S(stream)));
  if (big)
    RET_LAUNCH(launch_absorb_big(rb, g->dev, g->max_deg, g->d_all_nodes, g->dev.n_nodes, T, g->X, gammas, g->ws,
                                 inst_active, S(stream)));
  else if (bp_fast_applicable(rb, g->dev, g->min_deg, g->max_deg))
    RET_LAUNCH(launch_absorb_fast(g->dev, T, g->X, gammas, inst_active, S(stream)));
  else if (fast64_applicable(rb, g))
    RET_LAUNCH(launch_absorb_fast64(g->dev, T, g->X, gammas, inst_active, S(stream)));
  else
    RET_LAUNCH(launch_absorb(rb, g->dev, T, g->X, gammas, g->ws, wsn, inst_active, S(stream)));
  return DQ_OK;
}

long long dq_edge_svd_tape_bytes(const dq_graph* g) { return g ? edge_svd_tape_bytes(g->dev.chi) : -1; }

int dq_vidal_gauge_tape(dq_graph* g, int dtype, const void* T, const void* n2e, double eps, void* gammas, double* lam,
                        void* X, void* tape, void* stream) {
  int rb, rc;
  if (!g || !X || !tape) return fail(DQ_ERR_INVALID, "NULL argument");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (gammas == T) return fail(DQ_ERR_INVALID, "gammas must not alias tensors");
  const long long wse = (edge_svd_ws_bytes_host(g->dev.chi) + 255) / 256 * 256;
  const long long wsn = (node_ws_bytes(g, rb) + 255) / 256 * 256;
  const size_t need_e = (size_t)wse * (g->dev.n_edges > 0 ? g->dev.n_edges : 1);
  const size_t need_n = (size_t)wsn * (g->dev.n_nodes > 0 ? g->dev.n_nodes : 1);
  const bool big = big_applicable(g->dev, g->max_deg);
  const size_t need_b = big ? (size_t)big_ws_bytes(g->dev, g->max_deg, g->dev.n_nodes, rb) : 0;
  if ((rc = ws_ensure(g, need_e > need_n ? (need_e > need_b ? need_e : need_b) : (need_n > need_b ? need_n : need_b))) !=
      DQ_OK)
    return rc;
  RET_LAUNCH(launch_edge_svd(rb, g->dev, n2e, eps, (cd*)X, lam, g->ws, wse, nullptr, tape, S(stream)));
  if (big)
    RET_LAUNCH(launch_absorb_big(rb, g->dev, g->max_deg, g->d_all_nodes, g->dev.n_nodes, T, (const cd*)X, gammas,
                                 g->ws, nullptr, S(stream)));
  else if (bp_fast_applicable(rb, g->dev, g->min_deg, g->max_deg))
    RET_LAUNCH(launch_absorb_fast(g->dev, T, (const cd*)X, gammas, nullptr, S(stream)));
  else if (fast64_applicable(rb, g))
    RET_LAUNCH(launch_absorb_fast64(g->dev, T, (const cd*)X, gammas, nullptr, S(stream)));
  else
    RET_LAUNCH(launch_absorb(rb, g->dev, T, (const cd*)X, gammas, g->ws, wsn, nullptr, S(stream)));
  return DQ_OK;
}

// Per-node workspace stride of the generic Vidal-distance kernels.
static long long vidal_generic_ws_stride(const dq_graph* g) {
  const long long cc = (long long)g->dev.chi * g->dev.chi;
  return ((5 * cc * (long long)sizeof(cd) + ((long long)g->max_deg * g->dev.chi + 4 * g->dev.chi + 16) * 8 + 16) +
          255) /
         256 * 256;
}

// Workspace bytes of one Vidal-distance call on this handle, by kernel path.
static size_t vidal_distance_ws_bytes(const dq_graph* g, int rb) {
  if (bp_fast_applicable(rb, g->dev, g->min_deg, g->max_deg)) return (size_t)vidal_fast_ws_bytes(g->dev);
  if (big_applicable(g->dev, g->max_deg)) return (size_t)vidal_big_ws_bytes(g->dev, g->max_deg, rb);
  return (size_t)vidal_generic_ws_stride(g) * (g->dev.n_nodes > 0 ? g->dev.n_nodes : 1);
}

static hipError_t vidal_distance_launch(dq_graph* g, int rb, const void* gammas, const double* lam, double eps,
                                        double* out, const int* inst_active, void* ws, hipStream_t s) {
  if (bp_fast_applicable(rb, g->dev, g->min_deg, g->max_deg))
    return launch_vidal_distance_fast(g->dev, gammas, lam, eps, g->d_all_nodes, g->d_zero_src, g->d_inst_edge_ptr,
                                      g->n_inst, out, ws, inst_active, s);
  if (big_applicable(g->dev, g->max_deg))
    return launch_vidal_distance_big(rb, g->dev, g->min_deg, g->max_deg, gammas, lam, eps, g->d_all_nodes,
                                     g->d_zero_src, g->d_inst_edge_ptr, g->n_inst, out, ws, inst_active, s);
  return launch_vidal_distance(rb, g->dev, gammas, lam, eps, g->node_dist, g->d_inst_node_ptr, g->d_inst_edge_ptr,
                               g->n_inst, out, ws, vidal_generic_ws_stride(g), inst_active, s);
}

int dq_vidal_distance(dq_graph* g, int dtype, const void* gammas, const double* lam, double eps, double* out,
                      const int* inst_active, void* stream) {
  int rb, rc;
  if (!g || !out) return fail(DQ_ERR_INVALID, "NULL argument");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if ((rc = ws_ensure(g, vidal_distance_ws_bytes(g, rb))) != DQ_OK) return rc;
  RET_LAUNCH(vidal_distance_launch(g, rb, gammas, lam, eps, out, inst_active, g->ws, S(stream)));
  return DQ_OK;
}

// ---------------------------------------------------------------------------
// Scratch for the entry points that allocate nothing: a fixed array in device
// memory, one copy per GPU (held until the process ends).  Host code never
// passes the array's name to a kernel: its device address on the current GPU
// comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[DQ_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mutex;
static std::map<int, void*> g_scratch_ptr;

static int scratch_on_current_device(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(DQ_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    it = g_scratch_ptr.emplace(dev, ptr).first;
  }
  *out = it->second;
  return DQ_OK;
}

int dq_vidal_distance_noalloc(dq_graph* g, int dtype, const void* gammas, const double* lam, double eps, double* out,
                              const int* inst_active, void* stream) {
  int rb, rc;
  if (!g || !out || !gammas || !lam) return fail(DQ_ERR_INVALID, "NULL argument");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  const size_t need = vidal_distance_ws_bytes(g, rb);
  if (need > (size_t)DQ_NOALLOC_SCRATCH_BYTES)
    return fail(DQ_ERR_INVALID,
                "dq_vidal_distance_noalloc: this graph needs %zu bytes of scratch, above the fixed %lld "
                "(use dq_vidal_distance)",
                need, (long long)DQ_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  if ((rc = scratch_on_current_device(&ws)) != DQ_OK) return rc;
  RET_LAUNCH(vidal_distance_launch(g, rb, gammas, lam, eps, out, inst_active, ws, S(stream)));
  return DQ_OK;
}

int dq_symmetric_gauge(dq_graph* g, int dtype, const void* gammas, const double* lam, double eps, void* out,
                       void* stream) {
  int rb, rc;
  if (!g) return fail(DQ_ERR_INVALID, "NULL graph");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  RET_LAUNCH(launch_symmetric_gauge(rb, g->dev, gammas, lam, eps, out, S(stream)));
  return DQ_OK;
}

int dq_apply1(dq_graph* g, int dtype, void* T, const int* nodes, int count, const void* gates, void* stream) {
  int rb, rc;
  if (!g || count < 0) return fail(DQ_ERR_INVALID, "bad argument");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  RET_LAUNCH(launch_apply1(rb, g->dev, T, nodes, count, (const cd*)gates, S(stream)));
  return DQ_OK;
}

int dq_simple_update_vjp(dq_graph* g, int dtype, const void* T_pre, const double* lam_pre, const int* gates,
                         int count, const void* halves, int max_gate_rank, double eps, double cut, const void* tape,
                         void* T_bar, double* lam_bar, void* halves_bar, void* stream) {
  if (!g || count < 0) return fail(DQ_ERR_INVALID, "bad argument");
  if (!(cut >= 0.0 && cut < 1.0)) return fail(DQ_ERR_INVALID, "cut %g outside [0, 1)", cut);
  if (dtype != DQ_C128) return fail(DQ_ERR_UNSUPPORTED, "the adjoint runs in complex128 only");
  if (max_gate_rank < 1 || max_gate_rank > g->dev.phys * g->dev.phys)
    return fail(DQ_ERR_INVALID, "max_gate_rank %d outside [1, phys^2]", max_gate_rank);
  if (count == 0) return DQ_OK;
  if (!T_pre || !lam_pre || !gates || !halves || !tape || !T_bar || !lam_bar || !halves_bar)
    return fail(DQ_ERR_INVALID, "NULL argument");
  const int chi = g->dev.chi, p = g->dev.phys;
  if (chi * p > 64 || chi * p * p > 128)
    return fail(DQ_ERR_UNSUPPORTED, "simple update supports chi*phys <= 64 and chi*phys^2 <= 128 (chi=%d, phys=%d)",
                chi, p);
  const long long wsb = su_vjp_ws_bytes_host(chi, p, max_gate_rank, g->max_deg);
  int rc;
  if ((rc = ws_ensure(g, (size_t)wsb * count)) != DQ_OK) return rc;
  RET_LAUNCH(launch_su_vjp(g->dev, T_pre, lam_pre, (const SuGate*)gates, count, (const cd*)halves, max_gate_rank, eps,
                           cut, tape, T_bar, lam_bar, halves_bar, g->ws, wsb, S(stream)));
  return DQ_OK;
}

static int need_c128(int dtype) {
  return dtype == DQ_C128 ? DQ_OK : fail(DQ_ERR_UNSUPPORTED, "the adjoint runs in complex128 only");
}

int dq_bp_sweep_vjp(dq_graph* g, int dtype, const void* T, const void* e2n_in, const void* n2e_out_bar,
                    void* e2n_in_bar, void* T_bar, void* stream) {
  int rc;
  if (!g || !T || !e2n_in || !n2e_out_bar || !e2n_in_bar || !T_bar) return fail(DQ_ERR_INVALID, "NULL argument");
  if ((rc = need_c128(dtype)) != DQ_OK) return rc;
  if (fast64_applicable(8, g)) {
    RET_LAUNCH(launch_bp_node_vjp_fast(g->dev, T, e2n_in, n2e_out_bar, e2n_in_bar, T_bar, S(stream)));
    return DQ_OK;
  }
  const long long wsb = (bp_vjp_ws_bytes(g->dev.chi, g->dev.phys, g->max_deg) + 255) / 256 * 256;
  if ((rc = ws_ensure(g, (size_t)wsb * vjp_blocks(g->dev.n_nodes))) != DQ_OK) return rc;
  RET_LAUNCH(launch_bp_node_vjp(g->dev, T, e2n_in, n2e_out_bar, e2n_in_bar, T_bar, g->ws, wsb, S(stream)));
  return DQ_OK;
}

int dq_vidal_gauge_vjp(dq_graph* g, int dtype, const void* T, const void* X, const void* tape, double eps,
                       double cut, const void* gammas_bar, const double* lam_bar, void* X_bar, void* T_bar,
                       void* n2e_bar, void* stream) {
  int rc;
  if (!(cut >= 0.0 && cut < 1.0)) return fail(DQ_ERR_INVALID, "cut %g outside [0, 1)", cut);
  if (!g || !T || !X || !tape || !gammas_bar || !lam_bar || !X_bar || !T_bar || !n2e_bar)
    return fail(DQ_ERR_INVALID, "NULL argument");
  if ((rc = need_c128(dtype)) != DQ_OK) return rc;
  const long long wa = (absorb_vjp_ws_bytes(g->dev.chi, g->dev.phys, g->max_deg) + 255) / 256 * 256;
  const long long we = (edge_svd_vjp_ws_bytes(g->dev.chi) + 255) / 256 * 256;
  const bool fast = fast64_applicable(8, g);
  const size_t na = fast ? (size_t)g->dev.n_nodes * 8192 * sizeof(cd) : (size_t)wa * vjp_blocks(g->dev.n_nodes);
  const size_t ne = (size_t)we * vjp_blocks(g->dev.n_edges);
  if ((rc = ws_ensure(g, na > ne ? na : ne)) != DQ_OK) return rc;
  if (fast)
    RET_LAUNCH(launch_absorb_vjp_fast(g->dev, T, X, gammas_bar, T_bar, X_bar, g->ws, S(stream)));
  else
    RET_LAUNCH(launch_absorb_vjp(g->dev, T, X, gammas_bar, T_bar, X_bar, g->ws, wa, S(stream)));
  RET_LAUNCH(launch_edge_svd_vjp(g->dev, tape, eps, cut, X_bar, lam_bar, n2e_bar, g->ws, we, S(stream)));
  return DQ_OK;
}

int dq_symmetric_gauge_vjp(dq_graph* g, int dtype, const void* gammas, const double* lam, double eps,
                           const void* out_bar, void* gammas_bar, double* lam_bar, void* stream) {
  int rc;
  if (!g || !gammas || !lam || !out_bar || !gammas_bar || !lam_bar) return fail(DQ_ERR_INVALID, "NULL argument");
  if ((rc = need_c128(dtype)) != DQ_OK) return rc;
  RET_LAUNCH(launch_symmetric_vjp(g->dev, gammas, lam, eps, out_bar, gammas_bar, lam_bar, S(stream)));
  return DQ_OK;
}

int dq_lambdas_to_messages_vjp(dq_graph* g, int dtype, const void* n2e_bar, const void* e2n_bar, double* lam_bar,
                               void* stream) {
  int rc;
  if (!g || !lam_bar) return fail(DQ_ERR_INVALID, "NULL argument");
  if ((rc = need_c128(dtype)) != DQ_OK) return rc;
  RET_LAUNCH(launch_lam2msg_vjp(g->dev, n2e_bar, e2n_bar, lam_bar, S(stream)));
  return DQ_OK;
}

int dq_rdm1_vjp(dq_graph* g, int dtype, const void* T, const double* lam, const int* nodes, int count,
                const void* rho_bar, void* T_bar, double* lam_bar, void* stream) {
  int rc;
  if (!g || count < 0 || (count > 0 && (!T || !lam || !nodes || !rho_bar || !T_bar || !lam_bar)))
    return fail(DQ_ERR_INVALID, "bad argument");
  if ((rc = need_c128(dtype)) != DQ_OK) return rc;
  if (g->dev.phys > 8) return fail(DQ_ERR_UNSUPPORTED, "density-matrix adjoints support phys <= 8");
  RET_LAUNCH(launch_rdm1_vjp(g->dev, T, lam, nodes, count, rho_bar, T_bar, lam_bar, S(stream)));
  return DQ_OK;
}

int dq_rdm2_vjp(dq_graph* g, int dtype, const void* T, const double* lam, const int* edges, int count,
                const void* rho_bar, void* T_bar, double* lam_bar, void* stream) {
  int rc;
  if (!g || count < 0 || (count > 0 && (!T || !lam || !edges || !rho_bar || !T_bar || !lam_bar)))
    return fail(DQ_ERR_INVALID, "bad argument");
  if ((rc = need_c128(dtype)) != DQ_OK) return rc;
  if (g->dev.phys > 4) return fail(DQ_ERR_UNSUPPORTED, "two-site density-matrix adjoint supports phys <= 4");
  const long long wsb = (rdm2_vjp_ws_bytes(g->dev.chi, g->dev.phys) + 255) / 256 * 256;
  if ((rc = ws_ensure(g, (size_t)wsb * (count > 0 ? count : 1))) != DQ_OK) return rc;
  RET_LAUNCH(launch_rdm2_vjp(g->dev, T, lam, edges, count, rho_bar, T_bar, lam_bar, g->ws, wsb, S(stream)));
  return DQ_OK;
}

int dq_apply1_vjp(dq_graph* g, int dtype, const void* T_pre, const int* nodes, int count, const void* gates,
                  void* T_bar, void* gates_bar, void* stream) {
  int rc;
  if (!g || count < 0 || (count > 0 && (!T_pre || !nodes || !gates || !T_bar || !gates_bar)))
    return fail(DQ_ERR_INVALID, "bad argument");
  if ((rc = need_c128(dtype)) != DQ_OK) return rc;
  if (g->dev.phys > 8) return fail(DQ_ERR_UNSUPPORTED, "one-site gate adjoint supports phys <= 8");
  RET_LAUNCH(launch_apply1_vjp(g->dev, T_pre, nodes, count, gates, T_bar, gates_bar, S(stream)));
  return DQ_OK;
}

int dq_decompose_gates(const void* gates, int count, int phys, double eps, void* halves, int* ranks, int rank_stride,
                       void* stream) {
  if (count < 0 || rank_stride < 1) return fail(DQ_ERR_INVALID, "bad argument");
  if (phys < 1 || phys > 4) return fail(DQ_ERR_UNSUPPORTED, "gate split supports phys 1..4 (phys=%d)", phys);
  if (count > 0 && (!gates || !halves || !ranks)) return fail(DQ_ERR_INVALID, "null pointer");
  RET_LAUNCH(launch_decompose_gates((const cd*)gates, count, phys, eps, (cd*)halves, ranks, rank_stride, S(stream)));
  return DQ_OK;
}

int dq_tensor_apply1(int dtype, const void* t, long long rows, int phys, const void* gate, void* out, double* ws,
                     void* stream) {
  int rb, rc;
  if (rows < 0 || phys < 1 || (rows > 0 && (!t || !gate || !out || !ws))) return fail(DQ_ERR_INVALID, "bad argument");
  if (t == out) return fail(DQ_ERR_INVALID, "dq_tensor_apply1: out must not alias the input");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  RET_LAUNCH(launch_tensor_gate(rb, t, rows, 1, 1, phys, (const cd*)gate, 1, out, ws, S(stream)));
  return DQ_OK;
}

int dq_tensor_apply_half(int dtype, const void* t, long long lead, int bond, long long trail, int phys,
                         const void* half, int rank, void* out, double* ws, void* stream) {
  int rb, rc;
  if (lead < 0 || bond < 0 || trail < 0 || phys < 1 || rank < 1) return fail(DQ_ERR_INVALID, "bad argument");
  if (lead * bond * trail > 0 && (!t || !half || !out || !ws)) return fail(DQ_ERR_INVALID, "null pointer");
  if (t == out) return fail(DQ_ERR_INVALID, "dq_tensor_apply_half: out must not alias the input");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  RET_LAUNCH(launch_tensor_gate(rb, t, lead, bond, trail, phys, (const cd*)half, rank, out, ws, S(stream)));
  return DQ_OK;
}

static int simple_update_impl(dq_graph* g, int dtype, void* T, double* lam, const int* gates, int count,
                              const void* halves, int max_gate_rank, double eps, int max_rank, double* err,
                              double* entropy, void* tape, void* stream) {
  int rb, rc;
  if (!g || count < 0 || max_rank < 1) return fail(DQ_ERR_INVALID, "bad argument");
  if (max_gate_rank < 1 || max_gate_rank > g->dev.phys * g->dev.phys)
    return fail(DQ_ERR_INVALID, "max_gate_rank %d outside [1, phys^2]", max_gate_rank);
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (count == 0) return DQ_OK;
  const int chi = g->dev.chi, p = g->dev.phys;
  if (chi * p > 64 || chi * p * p > 128)
    return fail(DQ_ERR_UNSUPPORTED, "simple update supports chi*phys <= 64 and chi*phys^2 <= 128 (chi=%d, phys=%d)",
                chi, p);
  const long long wsb = su_ws_bytes(chi, p, max_gate_rank);
  if ((rc = ws_ensure(g, (size_t)wsb * count)) != DQ_OK) return rc;
  RET_LAUNCH(launch_su(rb, g->dev, T, lam, (const SuGate*)gates, count, (const cd*)halves, max_gate_rank, eps,
                       max_rank, err, entropy, g->ws, wsb, tape, S(stream)));
  return DQ_OK;
}

int dq_simple_update(dq_graph* g, int dtype, void* T, double* lam, const int* gates, int count, const void* halves,
                     int max_gate_rank, double eps, int max_rank, double* err, double* entropy, void* stream) {
  return simple_update_impl(g, dtype, T, lam, gates, count, halves, max_gate_rank, eps, max_rank, err, entropy,
                            nullptr, stream);
}

long long dq_su_tape_bytes(const dq_graph* g, int max_gate_rank) {
  return g ? su_tape_bytes(g->dev.chi, max_gate_rank) : -1;
}

int dq_simple_update_tape(dq_graph* g, int dtype, void* T, double* lam, const int* gates, int count,
                          const void* halves, int max_gate_rank, double eps, int max_rank, double* err,
                          double* entropy, void* tape, void* stream) {
  if (!tape && count > 0) return fail(DQ_ERR_INVALID, "dq_simple_update_tape: NULL tape");
  return simple_update_impl(g, dtype, T, lam, gates, count, halves, max_gate_rank, eps, max_rank, err, entropy, tape,
                            stream);
}

int dq_rdm1(dq_graph* g, int dtype, const void* T, const void* e2n, const int* nodes, int count, void* rho,
            void* stream) {
  int rb, rc;
  if (!g || count < 0) return fail(DQ_ERR_INVALID, "bad argument");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (count == 0) return DQ_OK;
  const int chi = g->dev.chi, p = g->dev.phys;
  const long long Sm = ipow(chi, g->max_deg) * p;
  const long long wsb = (2 * Sm * 16 + 2LL * chi * p * chi * p * 16 + 64 + 255) / 256 * 256;
  if ((rc = ws_ensure(g, (size_t)wsb * count)) != DQ_OK) return rc;
  RET_LAUNCH(launch_rdm1(rb, g->dev, T, e2n, nodes, count, (cd*)rho, g->ws, wsb, S(stream)));
  return DQ_OK;
}

int dq_rdm2(dq_graph* g, int dtype, const void* T, const void* e2n, const int* edges, int count, void* rho,
            void* stream) {
  int rb, rc;
  if (!g || count < 0) return fail(DQ_ERR_INVALID, "bad argument");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (count == 0) return DQ_OK;
  const int chi = g->dev.chi, p = g->dev.phys;
  const long long Sm = ipow(chi, g->max_deg) * p;
  const long long wsb = (2 * Sm * 16 + 2LL * chi * p * chi * p * 16 + 64 + 255) / 256 * 256;
  if ((rc = ws_ensure(g, (size_t)wsb * count)) != DQ_OK) return rc;
  RET_LAUNCH(launch_rdm2(rb, g->dev, T, e2n, edges, count, (cd*)rho, g->ws, wsb, S(stream)));
  return DQ_OK;
}

int dq_update_active(const double* dist, double acc, int* iters, int max_iters, int* active, int n_inst,
                     void* stream) {
  RET_LAUNCH(launch_update_active(dist, acc, iters, max_iters, active, n_inst, S(stream)));
  return DQ_OK;
}

static int sv_check(const void* psi, int n, const char* what) {
  if (!psi) return fail(DQ_ERR_INVALID, "%s: psi is NULL", what);
  if (n < 1 || n > 40) return fail(DQ_ERR_INVALID, "%s: n_qubits %d outside [1, 40]", what, n);
  return DQ_OK;
}

long long dq_sv_work_bytes(void) {
  const long long dlow = 4096LL * 16;
  const long long part = sv_rdm_part_bytes();
  return part > dlow ? part : dlow;
}

int dq_sv_set_basis(void* psi, int n_qubits, long long index, void* stream) {
  int rc = sv_check(psi, n_qubits, "dq_sv_set_basis");
  if (rc) return rc;
  if (index < 0 || (unsigned long long)index >= (1ULL << n_qubits))
    return fail(DQ_ERR_INVALID, "dq_sv_set_basis: index %lld outside [0, 2^%d)", index, n_qubits);
  RET_LAUNCH(launch_sv_basis(psi, n_qubits, (unsigned long long)index, S(stream)));
  return DQ_OK;
}

int dq_sv_apply1(void* psi, int n_qubits, int bit, const void* gate, void* stream) {
  int rc = sv_check(psi, n_qubits, "dq_sv_apply1");
  if (rc) return rc;
  if (!gate) return fail(DQ_ERR_INVALID, "dq_sv_apply1: gate is NULL");
  if (bit < 0 || bit >= n_qubits) return fail(DQ_ERR_INVALID, "dq_sv_apply1: bit %d outside [0, %d)", bit, n_qubits);
  RET_LAUNCH(launch_sv_apply1(psi, n_qubits, bit, gate, S(stream)));
  return DQ_OK;
}

int dq_sv_apply2(void* psi, int n_qubits, int bit1, int bit2, const void* gate, void* stream) {
  int rc = sv_check(psi, n_qubits, "dq_sv_apply2");
  if (rc) return rc;
  if (!gate) return fail(DQ_ERR_INVALID, "dq_sv_apply2: gate is NULL");
  if (bit1 < 0 || bit1 >= n_qubits || bit2 < 0 || bit2 >= n_qubits || bit1 == bit2)
    return fail(DQ_ERR_INVALID, "dq_sv_apply2: bits (%d, %d) must be distinct and in [0, %d)", bit1, bit2, n_qubits);
  RET_LAUNCH(launch_sv_apply2(psi, n_qubits, bit1, bit2, gate, S(stream)));
  return DQ_OK;
}

int dq_sv_apply_diagonal(void* psi, int n_qubits, int n_terms, const int* term_bits, const void* term_diag,
                         void* work, void* stream) {
  int rc = sv_check(psi, n_qubits, "dq_sv_apply_diagonal");
  if (rc) return rc;
  if (n_terms < 0) return fail(DQ_ERR_INVALID, "dq_sv_apply_diagonal: n_terms %d < 0", n_terms);
  if (n_terms == 0) return DQ_OK;
  if (!term_bits || !term_diag || !work) return fail(DQ_ERR_INVALID, "dq_sv_apply_diagonal: NULL argument");
  RET_LAUNCH(launch_sv_diag(psi, n_qubits, n_terms, term_bits, term_diag, work, S(stream)));
  return DQ_OK;
}

static int sv_mask_check(int n, long long mask, const char* what) {
  if (n < 64 && ((unsigned long long)mask >> n) != 0ULL)
    return fail(DQ_ERR_INVALID, "%s: target_mask has bits at or above n_qubits = %d", what, n);
  return DQ_OK;
}

int dq_sv_apply1_layer(void* psi, int n_qubits, long long target_mask, const void* gates, void* stream) {
  int rc = sv_check(psi, n_qubits, "dq_sv_apply1_layer");
  if (rc) return rc;
  if ((rc = sv_mask_check(n_qubits, target_mask, "dq_sv_apply1_layer"))) return rc;
  if (target_mask == 0) return DQ_OK;
  if (!gates) return fail(DQ_ERR_INVALID, "dq_sv_apply1_layer: gates is NULL");
  RET_LAUNCH(launch_sv_layer1(psi, n_qubits, (unsigned long long)target_mask, gates, S(stream)));
  return DQ_OK;
}

int dq_sv_apply_diagonal_layer1(void* psi, int n_qubits, int n_terms, const int* term_bits, const void* term_diag,
                                long long target_mask, const void* gates, void* work, void* stream) {
  int rc = sv_check(psi, n_qubits, "dq_sv_apply_diagonal_layer1");
  if (rc) return rc;
  if ((rc = sv_mask_check(n_qubits, target_mask, "dq_sv_apply_diagonal_layer1"))) return rc;
  if (n_terms < 0) return fail(DQ_ERR_INVALID, "dq_sv_apply_diagonal_layer1: n_terms %d < 0", n_terms);
  if (n_terms > 0 && (!term_bits || !term_diag || !work))
    return fail(DQ_ERR_INVALID, "dq_sv_apply_diagonal_layer1: NULL term argument");
  if (target_mask != 0 && !gates) return fail(DQ_ERR_INVALID, "dq_sv_apply_diagonal_layer1: gates is NULL");
  RET_LAUNCH(launch_sv_diag_layer1(psi, n_qubits, n_terms, term_bits, term_diag, (unsigned long long)target_mask,
                                   gates, work, S(stream)));
  return DQ_OK;
}

int dq_sv_rdm1(const void* psi, int n_qubits, long long target_mask, void* rho, void* work, void* stream) {
  int rc = sv_check(psi, n_qubits, "dq_sv_rdm1");
  if (rc) return rc;
  if ((rc = sv_mask_check(n_qubits, target_mask, "dq_sv_rdm1"))) return rc;
  if (target_mask == 0) return DQ_OK;
  if (!rho || !work) return fail(DQ_ERR_INVALID, "dq_sv_rdm1: NULL argument");
  RET_LAUNCH(launch_sv_rdm1(psi, n_qubits, (unsigned long long)target_mask, rho, work, S(stream)));
  return DQ_OK;
}

int dq_repad(const dq_graph* old, dq_graph* g, int dtype, const void* T_old, void* T_new, void* stream) {
  int rb, rc;
  if (!old || !g) return fail(DQ_ERR_INVALID, "NULL graph");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (old->dev.n_nodes != g->dev.n_nodes || old->dev.chi > g->dev.chi || old->dev.phys != g->dev.phys)
    return fail(DQ_ERR_INVALID, "repad needs the same topology and a chi that does not shrink");
  RET_LAUNCH(launch_repad(rb, g->dev, old->d_toff, old->dev.chi, T_old, T_new, S(stream)));
  return DQ_OK;
}

}  // extern "C"

// ------------------------------------------------------------- autotuner
namespace dq {
static std::mutex g_tune_mutex;
static std::map<std::pair<int, long long>, int> g_tune;

int tune_get(int kernel, long long key) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  const auto it = g_tune.find({kernel, key});
  return it == g_tune.end() ? -1 : it->second;
}
void tune_set(int kernel, long long key, int variant) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  if (variant < 0)
    g_tune.erase({kernel, key});
  else
    g_tune[{kernel, key}] = variant;
}

// Deterministic values in [lo, lo + 1) for timing inputs.
__global__ void k_tune_fill(double* x, long long n, double lo, unsigned seed) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    unsigned h = (unsigned)i * 2654435761u ^ seed;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    x[i] = lo + (h & 0xffffff) / 16777216.0;
  }
}
__global__ void k_tune_fill_f(float* x, long long n, float lo, unsigned seed) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    unsigned h = (unsigned)i * 2654435761u ^ seed;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    x[i] = lo + (h & 0xffffff) / 16777216.0f;
  }
}
}  // namespace dq

namespace {
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int fill_real(void* p, long long count, int rb, double lo, unsigned seed, hipStream_t s) {
  if (rb == 8)
    hipLaunchKernelGGL(dq::k_tune_fill, dim3(dq::grid_for(count, 256)), dim3(256), 0, s, (double*)p, count, lo, seed);
  else
    hipLaunchKernelGGL(dq::k_tune_fill_f, dim3(dq::grid_for(count, 256)), dim3(256), 0, s, (float*)p, count,
                       (float)lo, seed);
  return hipGetLastError() == hipSuccess ? DQ_OK : fail(DQ_ERR_HIP, "autotune fill failed");
}

// Median of `reps` timed runs of `run` (after one untimed run).
template <typename F>
int time_median(hipStream_t s, int reps, F run, float* ms) {
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  std::vector<float> t;
  int rc = DQ_OK;
  for (int i = 0; i <= reps && rc == DQ_OK; ++i) {
    (void)hipEventRecord(a, s);
    if (run() != hipSuccess) rc = fail(DQ_ERR_HIP, "autotune launch failed");
    (void)hipEventRecord(b, s);
    (void)hipEventSynchronize(b);
    float e = 0.f;
    (void)hipEventElapsedTime(&e, a, b);
    if (i > 0) t.push_back(e);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc != DQ_OK) return rc;
  std::sort(t.begin(), t.end());
  *ms = t[t.size() / 2];
  return DQ_OK;
}
}  // namespace

int dq_tune_override(int kernel, int dtype, int chi, int phys, int max_gate_rank, int variant) {
  int rb, rc;
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  if (kernel == 0) {
    if (variant < -1 || variant > 3) return fail(DQ_ERR_INVALID, "node-update variant %d outside [-1, 3]", variant);
    tune_set(kTuneBpD3C16, 0, variant);
  } else if (kernel == 1) {
    if (variant != -1 && variant != 1 && variant != 2)
      return fail(DQ_ERR_INVALID, "simple-update mode %d is not -1, 1 (all-LDS) or 2 (compact)", variant);
    tune_set(kTuneSu, su_tune_key(rb, chi, phys, max_gate_rank), variant);
  } else {
    return fail(DQ_ERR_INVALID, "unknown kernel %d", kernel);
  }
  return DQ_OK;
}

int dq_autotune(dq_graph* g, int dtype, int max_gate_rank, int force, void* stream, char* report, int report_len) {
  int rb, rc;
  if (!g) return fail(DQ_ERR_INVALID, "NULL graph");
  if ((rc = check_dtype(dtype, &rb)) != DQ_OK) return rc;
  hipStream_t s = S(stream);
  std::string rep;
  char line[256];
  const int chi = g->dev.chi, p = g->dev.phys;
  const long long cc = (long long)chi * chi, E = g->dev.n_edges, nm = 2 * E * cc;
  // BP node kernel variants (complex64, degree 3, chi 16)
  if (bp_fast_applicable(rb, g->dev, g->min_deg, g->max_deg) && (force || tune_get(kTuneBpD3C16, 0) < 0)) {
    DevBuf T, m[4];
    HIP_TRY(hipMalloc(&T.p, (size_t)g->tensor_elems * 2 * rb));
    for (auto& b : m) HIP_TRY(hipMalloc(&b.p, (size_t)nm * 2 * rb));
    if ((rc = fill_real(T.p, 2 * g->tensor_elems, rb, -0.5, 1u, s)) != DQ_OK) return rc;
    for (int i = 0; i < 4; ++i)
      if ((rc = fill_real(m[i].p, 2 * nm, rb, 0.0, 2u + i, s)) != DQ_OK) return rc;
    if ((rc = bp_sweep_ws(g, rb)) != DQ_OK) return rc;
    const char* names[4] = {"8 waves, persistent", "8 waves, block per node", "4 waves, persistent",
                            "4 waves, block per node"};
    int best = 0;
    float best_ms = 1e30f;
    for (int v = 0; v < 4; ++v) {
      tune_set(kTuneBpD3C16, 0, v);
      float ms = 0.f;
      if ((rc = time_median(s, 5, [&] { return bp_sweep_launch(g, rb, T.p, m[0].p, m[1].p, m[2].p, m[3].p, nullptr, s); },
                            &ms)) != DQ_OK) {
        tune_set(kTuneBpD3C16, 0, -1);
        return rc;
      }
      snprintf(line, sizeof line, "bp node update d3 chi16 c64: %s %.1f us\n", names[v], 1e3 * ms);
      rep += line;
      if (ms < best_ms) {
        best_ms = ms;
        best = v;
      }
    }
    tune_set(kTuneBpD3C16, 0, best);
    snprintf(line, sizeof line, "bp node update d3 chi16 c64: chose %s\n", names[best]);
    rep += line;
  }
  // simple-update storage mode where the compact and all-LDS layouts both apply
  const long long skey = su_tune_key(rb, chi, p, max_gate_rank);
  if (max_gate_rank >= 1 && su_mem_mode(chi, p, max_gate_rank) == 2 && (force || tune_get(kTuneSu, skey) < 0) &&
      E > 0 && g->max_deg <= 4) {
    // a greedy matching of the edges: gates that share no node, one launch
    std::vector<int> en(2 * E);
    HIP_TRY(hipMemcpy(en.data(), g->d_edge_node, en.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<char> used(g->dev.n_nodes, 0);
    std::vector<SuGate> gates;
    for (int e = 0; e < E; ++e)
      if (!used[en[2 * e]] && !used[en[2 * e + 1]]) {
        used[en[2 * e]] = used[en[2 * e + 1]] = 1;
        gates.push_back({e, 0, max_gate_rank, 0});
      }
    const int cnt = (int)gates.size();
    const long long hst = (long long)p * p * p * p;
    DevBuf T0, T, lam0, lam, dg, hv, err, ent;
    HIP_TRY(hipMalloc(&T0.p, (size_t)g->tensor_elems * 2 * rb));
    HIP_TRY(hipMalloc(&T.p, (size_t)g->tensor_elems * 2 * rb));
    HIP_TRY(hipMalloc(&lam0.p, (size_t)E * chi * sizeof(double)));
    HIP_TRY(hipMalloc(&lam.p, (size_t)E * chi * sizeof(double)));
    HIP_TRY(hipMalloc(&dg.p, (size_t)cnt * sizeof(SuGate)));
    HIP_TRY(hipMalloc(&hv.p, (size_t)cnt * 2 * hst * sizeof(cd)));
    HIP_TRY(hipMalloc(&err.p, (size_t)cnt * sizeof(double)));
    HIP_TRY(hipMalloc(&ent.p, (size_t)cnt * sizeof(double)));
    HIP_TRY(hipMemcpy(dg.p, gates.data(), cnt * sizeof(SuGate), hipMemcpyHostToDevice));
    if ((rc = fill_real(T0.p, 2 * g->tensor_elems, rb, -0.5, 7u, s)) != DQ_OK) return rc;
    if ((rc = fill_real(lam0.p, E * chi, 8, 0.1, 8u, s)) != DQ_OK) return rc;
    if ((rc = fill_real(hv.p, (long long)cnt * 2 * hst * 2, 8, -0.5, 9u, s)) != DQ_OK) return rc;
    const long long wsb = su_ws_bytes(chi, p, max_gate_rank);
    if ((rc = ws_ensure(g, (size_t)wsb * cnt)) != DQ_OK) return rc;
    const int modes[2] = {2, 1};
    const char* names[2] = {"compact (two blocks per CU)", "all-LDS (one block per CU)"};
    int best = 2;
    float best_ms = 1e30f;
    for (int i = 0; i < 2; ++i) {
      tune_set(kTuneSu, skey, modes[i]);
      float ms = 0.f;
      auto run = [&] {
        hipError_t e = hipMemcpyAsync(T.p, T0.p, (size_t)g->tensor_elems * 2 * rb, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(lam.p, lam0.p, (size_t)E * chi * sizeof(double), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
          e = launch_su(rb, g->dev, T.p, (double*)lam.p, (const SuGate*)dg.p, cnt, (const cd*)hv.p, max_gate_rank,
                        1e-10, chi, (double*)err.p, (double*)ent.p, g->ws, wsb, nullptr, s);
        return e;
      };
      if ((rc = time_median(s, 3, run, &ms)) != DQ_OK) {
        tune_set(kTuneSu, skey, -1);
        return rc;
      }
      snprintf(line, sizeof line, "simple update chi%d p%d r%d %s: %s %.1f us (%d gates, incl. input reset)\n", chi, p,
               max_gate_rank, rb == 4 ? "c64" : "c128", names[i], 1e3 * ms, cnt);
      rep += line;
      if (ms < best_ms) {
        best_ms = ms;
        best = modes[i];
      }
    }
    tune_set(kTuneSu, skey, best);
    snprintf(line, sizeof line, "simple update chi%d p%d r%d %s: chose %s\n", chi, p, max_gate_rank,
             rb == 4 ? "c64" : "c128", names[best == 2 ? 0 : 1]);
    rep += line;
  }
  HIP_TRY(hipStreamSynchronize(s));
  if (report && report_len > 0) {
    strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return DQ_OK;
}

// Node updates and the Vidal absorb for large node tensors: S = chi^d p of
// 2^15 elements and more (degree 3 at chi = 32: 64 Ki elements; degree 6 at
// chi = 16: 33.5 M elements, 256 MiB in complex64), where one workgroup per
// node leaves the chip idle and the tensors are beyond LDS.  Same arithmetic as the generic kernels
// (pass_through_node, dequantizer/src/tensor_ops.py:36-60; absorb,
// tensor_ops.py:88-93), split into HBM-streaming launches over all nodes of a
// level at once:
//   mode product   dst = src x_axis M: one thread per column (all indices but
//                  the axis), the chi values of the column in registers, M in
//                  LDS; grid = column blocks x nodes.  The products of a node
//                  update follow a plan that shares partial products between
//                  output slots (big_plan: 16 products at degree 6 where each
//                  slot on its own takes 5), in per-node buffers.
//   Gram partial   conj(T) . U over a chunk of columns per block, columns
//                  staged through LDS in tiles of 64, FP64 accumulation.
//   Gram reduce    per node: sum of the chunks, division by the trace.
// Each element of T and of the intermediates is read once per pass: the
// launches are bound by HBM bandwidth.
#include <vector>

#include "dq_launch.hpp"

namespace dq {

constexpr int kBigThreads = 256;
constexpr int kBigTile = 64;  // Gram columns per LDS tile

// First tile of block b of n in a node's grid-stride tile loop, dealt so that
// neighbouring tiles run on one XCD: workgroups go to the 8 XCDs in turn
// (blocks b and b + 8 share an L2), and with 64-byte runs per tile row two
// neighbouring tiles read and write the two halves of each 128-byte line.
__device__ inline unsigned big_xcd_tile(unsigned b, unsigned n) { return n % 8 == 0 ? (b % 8) * (n / 8) + b / 8 : b; }

// Absorb product t of a node (axis t): dst = src x_t X^T, ping-ponging
// between the node's first two buffers (stride nb buffers per node).
template <typename R, int CHI>
__global__ void __launch_bounds__(kBigThreads)
    k_big_absorb_product(DevGraph g, const int* nodes, int t, const cx<R>* T, const cd* X, cx<R>* buf, long long smax,
                         int nb, const int* inst_active) {
  constexpr int CC = CHI * CHI;
  __shared__ cx<R> Ms[CC];
  const int k = blockIdx.y;
  const int node = nodes[k];
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node];
  const int axis = t;
  if (axis >= d) return;
  const int base = g.adj_ptr[node];
  const long long S = ipow(CHI, d) * g.phys;
  const long long inner = ipow(CHI, d - 1 - axis) * g.phys;
  const long long cols = S / CHI;
  const int e = g.adj_edge[base + axis], sd = g.adj_side[base + axis];
  // Ms[a][b] = X[b][a]: the orthogonaliser's column index is the new bond
  const cd* x = X + (2LL * e + sd) * CC;
  for (int q = threadIdx.x; q < CC; q += blockDim.x) Ms[q] = cvt<R>(x[(q % CHI) * CHI + q / CHI]);
  __syncthreads();
  const cx<R>* src = t == 0 ? T + g.toff[node] : buf + ((long long)nb * k + ((t - 1) & 1)) * smax;
  cx<R>* dst = buf + ((long long)nb * k + (t & 1)) * smax;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < cols;
       c += (long long)gridDim.x * blockDim.x) {
    const long long o = c / inner, i = c - o * inner;
    const long long b0 = o * CHI * inner + i;
    cx<R> in[CHI];
#pragma unroll
    for (int b = 0; b < CHI; ++b) in[b] = src[b0 + b * inner];
#pragma unroll 2
    for (int a = 0; a < CHI; ++a) {
      cx<R> acc = mk<R>(0, 0);
#pragma unroll
      for (int b = 0; b < CHI; ++b) fma_acc(acc, Ms[a * CHI + b], in[b]);
      dst[b0 + a * inner] = acc;
    }
  }
}

// One product of a node-update plan: buffer dst = (buffer src, or T when
// src < 0) x_axis M_axis, buffers nb per node.  For a node of degree <= axis
// the product is the identity (a copy): one plan serves every degree of a
// level.
template <typename R, int CHI>
__global__ void __launch_bounds__(kBigThreads)
    k_big_plan_product(DevGraph g, const int* nodes, int axis, int src_id, int dst_id, int nb, const cx<R>* T,
                       const cx<R>* n2e_old, const cx<R>* e2n_old, const cx<R>* n2e_new, const uint8_t* node_src,
                       cx<R>* buf, long long smax, const int* inst_active) {
  constexpr int CC = CHI * CHI;
  __shared__ cx<R> Ms[CC];
  const int k = blockIdx.y;
  const int node = nodes[k];
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node];
  const long long S = ipow(CHI, d) * g.phys;
  const cx<R>* src = src_id < 0 ? T + g.toff[node] : buf + ((long long)nb * k + src_id) * smax;
  cx<R>* dst = buf + ((long long)nb * k + dst_id) * smax;
  if (axis >= d) {
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < S;
         f += (long long)gridDim.x * blockDim.x)
      dst[f] = src[f];
    return;
  }
  const int base = g.adj_ptr[node];
  const long long inner = ipow(CHI, d - 1 - axis) * g.phys;
  const long long cols = S / CHI;
  const int e = g.adj_edge[base + axis], sd = g.adj_side[base + axis];
  // the incoming message, applied on the ket side: out[a] = sum_b M[a][b] in[b]
  const int code = node_src ? node_src[base + axis] : 0;
  const cx<R>* m = code == 0 ? e2n_old + (2LL * e + sd) * CC
                             : (code == 1 ? n2e_new : n2e_old) + (2LL * e + 1 - sd) * CC;
  for (int q = threadIdx.x; q < CC; q += blockDim.x) Ms[q] = m[q];
  __syncthreads();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < cols;
       c += (long long)gridDim.x * blockDim.x) {
    const long long o = c / inner, i = c - o * inner;
    const long long b0 = o * CHI * inner + i;
    cx<R> in[CHI];
#pragma unroll
    for (int b = 0; b < CHI; ++b) in[b] = src[b0 + b * inner];
#pragma unroll 2
    for (int a = 0; a < CHI; ++a) {
      cx<R> acc = mk<R>(0, 0);
#pragma unroll
      for (int b = 0; b < CHI; ++b) fma_acc(acc, Ms[a * CHI + b], in[b]);
      dst[b0 + a * inner] = acc;
    }
  }
}

// Partial Gram of output slot idx over this block's columns:
// part[(k * chunks + chunk)][kk * CHI + l] = sum conj(T[..kk..]) U[..l..],
// U in buffer src_id (T itself when src_id < 0).  Columns are staged through
// LDS in tiles of 64 (loads along the columns, coalesced); the products run
// on FP64 matrix cores (v_mfma_f64_16x16x4f64: exact products and FP64
// sums of the values converted from R, as the FMA form), four real MFMAs per
// complex k-step: Re += Tr Ur + Ti Ui, Im += Tr Ui - Ti Ur.  Operand layout:
// lane l holds A[l % 16][l / 16] = conj T[column l / 16][kk = m0 + l % 16]
// and B[l / 16][l % 16] = U[column l / 16][n0 + l % 16]; the result
// D[l / 16 + 4 v][l % 16] stays in four accumulators.  chi 32: wave w owns the
// 16 x 16 tile (w / 2, w % 2) over all columns; chi 16: the one tile, the
// four waves taking a quarter of each column tile each, summed at the end.
typedef double big_d4 __attribute__((ext_vector_type(4)));
typedef float big_f4 __attribute__((ext_vector_type(4)));

template <typename R, int CHI>
__global__ void __launch_bounds__(kBigThreads)
    k_big_gram(DevGraph g, const int* nodes, int idx, int src_id, int nb, const cx<R>* T, const cx<R>* buf,
               long long smax, cd* part, const int* inst_active) {
  static_assert(CHI == 16 || CHI == 32, "matrix-core Gram for chi 16 and 32");
  constexpr int CC = CHI * CHI;
  __shared__ cx<R> Tt[kBigTile][CHI + 1], Ut[kBigTile][CHI + 1];
  const int k = blockIdx.y;
  const int node = nodes[k];
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node];
  if (idx >= d) return;
  const cx<R>* t = T + g.toff[node];
  const cx<R>* u = src_id < 0 ? t : buf + ((long long)nb * k + src_id) * smax;
  const long long S = ipow(CHI, d) * g.phys;
  const long long inner = ipow(CHI, d - 1 - idx) * g.phys;
  const long long cols = S / CHI;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = CHI == 32 ? 16 * (w >> 1) : 0, n0 = CHI == 32 ? 16 * (w & 1) : 0;
  // k-steps of a column tile this wave takes: all 16 (chi 32) or 4 (chi 16)
  constexpr int KS = CHI == 32 ? kBigTile / 4 : kBigTile / 16;
  const int ks0 = CHI == 32 ? 0 : w * KS;
  const int mi = lane & 15, kq = lane >> 4;
  big_d4 gr = {0.0, 0.0, 0.0, 0.0}, gi = gr;
  for (long long c0 = (long long)blockIdx.x * kBigTile; c0 < cols; c0 += (long long)gridDim.x * kBigTile) {
    for (int q = threadIdx.x; q < kBigTile * CHI; q += blockDim.x) {
      const int r = q / kBigTile, cc = q % kBigTile;
      const long long c = c0 + cc;
      cx<R> a = mk<R>(0, 0), b = a;
      if (c < cols) {
        const long long o = c / inner, i = c - o * inner;
        const long long off = o * CHI * inner + i + r * inner;
        a = t[off];
        b = u[off];
      }
      Tt[cc][r] = a;
      Ut[cc][r] = b;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int c = 4 * (ks0 + ks) + kq;
      const cx<R> a = Tt[c][m0 + mi], b = Ut[c][n0 + mi];
      const double ar = (double)a.re, ai = (double)a.im, br = (double)b.re, bi = (double)b.im;
      gr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, gr, 0, 0, 0);
      gr = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, gr, 0, 0, 0);
      gi = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, gi, 0, 0, 0);
      gi = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, br, gi, 0, 0, 0);
    }
    __syncthreads();
  }
  cd* pp = part + ((long long)k * gridDim.x + blockIdx.x) * CC;
  if (CHI == 32) {
#pragma unroll
    for (int v = 0; v < 4; ++v) pp[(m0 + kq + 4 * v) * CHI + n0 + mi] = mk<double>(gr[v], gi[v]);
  } else {
    // the four waves' partial tiles, summed through LDS
    __shared__ cd red[4 * 256];
#pragma unroll
    for (int v = 0; v < 4; ++v) red[w * 256 + (kq + 4 * v) * 16 + mi] = mk<double>(gr[v], gi[v]);
    __syncthreads();
    for (int q = threadIdx.x; q < 256; q += blockDim.x)
      pp[q] = red[q] + red[256 + q] + red[512 + q] + red[768 + q];
  }
}

// The last product of an output slot fused with its Gram, chi = 16, for
// levels whose nodes all have the plan's degree:
//   G[k][l] += sum_{x_j} conj(T[.., x_j, .., k@idx, ..]) (M_j P)[.., x_j, .., l@idx, ..]
// summed over every other index c'.  A tile holds the 16 x 16 (x_j, x_idx)
// slices of P and T for W = 8 consecutive c' (c' enumerates the remaining
// indices, the contiguous ones fastest, so the loads run along memory) in
// LDS; wave w takes slices 2w and 2w + 1.  Both products run on matrix
// cores: U = M_j P_w in FP32 (complex64 data, as the separate product), whose
// result registers (row 4 (l / 16) + v, column l % 16) are exactly the B
// operands of the Gram's four k-steps over a (rows v, v + 4, v + 8, v + 12),
// the Gram in FP64 with A = conj(T_w)^T.  U never goes to memory: one pass
// reads P and T where the separate product and Gram read P, write U, then
// read U and T.  complex64 only (the complex128 levels keep the separate
// launches, which run faster there).
template <typename R>
__global__ void __launch_bounds__(kBigThreads, 3)
    k_big_prod_gram16(DevGraph g, const int* nodes, int j, int idx, int src_id, int nb, const cx<R>* T,
                      const cx<R>* n2e_old, const cx<R>* e2n_old, const cx<R>* n2e_new, const uint8_t* node_src,
                      const cx<R>* buf, long long smax, cd* part, const int* inst_active) {
  constexpr int CHI = 16, CC = CHI * CHI, W = 8;
  __shared__ cx<R> Pt[W][CHI][CHI + 1], Tt[W][CHI][CHI + 1];
  __shared__ cd red[4 * CC];
  const int k = blockIdx.y;
  const int node = nodes[k];
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, mi = lane & 15, kq = lane >> 4;
  const long long S = ipow(CHI, d) * g.phys;
  const int a_hi = j < idx ? j : idx, a_lo = j < idx ? idx : j;
  const long long n_bot = ipow(CHI, d - 1 - a_lo) * g.phys;  // contiguous indices after both axes
  const long long n_mid = ipow(CHI, a_lo - a_hi - 1);
  const long long st_lo = n_bot, st_mid = CHI * n_bot, st_hi = n_mid * CHI * n_bot, st_top = CHI * st_hi;
  const long long ncp = S / CC, ntiles = ncp / W;
  const cx<R>* t = T + g.toff[node];
  const cx<R>* P = src_id < 0 ? t : buf + ((long long)nb * k + src_id) * smax;
  // M_j[a][b] as the A operand of the product's four k-steps: lane l holds
  // M[l % 16][4 ks + l / 16]
  float mvr[4], mvi[4];
  {
    const int base = g.adj_ptr[node];
    const int e = g.adj_edge[base + j], sd = g.adj_side[base + j];
    const int code = node_src ? node_src[base + j] : 0;
    const cx<R>* m = code == 0 ? e2n_old + (2LL * e + sd) * CC
                               : (code == 1 ? n2e_new : n2e_old) + (2LL * e + 1 - sd) * CC;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const cx<R> mm = m[mi * CHI + 4 * ks + kq];
      mvr[ks] = (float)mm.re;
      mvi[ks] = (float)mm.im;
    }
  }
  // the next tile's elements are loaded into registers while the current
  // tile is multiplied (8 of P and 8 of T per thread)
  // Index arithmetic in 32 bits (the host fuses only tensors below 2^31
  // elements): the (x_hi, x_lo) part of each of the thread's PER elements is
  // fixed; when the contiguous run n_bot holds whole tiles the c' part is one
  // offset per tile.  Element u of thread tid is q = tid + 256 u of the tile:
  // slice ww = q % W and x_lo = (q / W) % CHI are the thread's own, x_hi =
  // q / (W CHI) steps by two per u (one offset and a uniform stride).
  constexpr int PER = W * CC / kBigThreads;
  static_assert(kBigThreads % (W * CHI) == 0, "element u of a thread keeps its slice and x_lo");
  constexpr int XH_STEP = kBigThreads / (W * CHI);
  cx<R> np[PER], nt[PER];
  const int ww = tid % W, xl = (tid / W) % CHI, xh0 = tid / (W * CHI);
  const unsigned offx = (unsigned)(xh0 * st_hi + xl * st_lo), ostep = (unsigned)(XH_STEP * st_hi);
  const unsigned nbot = (unsigned)n_bot, nmid = (unsigned)n_mid, smid = (unsigned)st_mid, stop = (unsigned)st_top;
  const bool whole = nbot % W == 0;
  auto load_tile = [&](long long tile) {
    const unsigned cp0 = (unsigned)(tile * W);
    if (whole) {
      const unsigned t2 = cp0 / nbot, base = (t2 / nmid) * stop + (t2 % nmid) * smid + cp0 % nbot;
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const unsigned off = base + offx + u * ostep + ww;
        np[u] = P[off];
        nt[u] = t[off];
      }
    } else {
      const unsigned cp = cp0 + ww;
      const unsigned t2 = cp / nbot;
      const unsigned base = (t2 / nmid) * stop + (t2 % nmid) * smid + cp % nbot + offx;
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const unsigned off = base + u * ostep;
        np[u] = P[off];
        nt[u] = t[off];
      }
    }
  };
  // two slices per wave, each with its own accumulators (four independent
  // chains of matrix-core steps in each phase)
  big_d4 gr0 = {0.0, 0.0, 0.0, 0.0}, gi0 = gr0, gr1 = gr0, gi1 = gr0;
  const unsigned tb = big_xcd_tile(blockIdx.x, gridDim.x);
  if (tb < ntiles) load_tile(tb);
  for (long long tile = tb; tile < ntiles; tile += gridDim.x) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int xh = xh0 + XH_STEP * u;
      const int xj = j < idx ? xh : xl, xi = j < idx ? xl : xh;
      Pt[ww][xj][xi] = np[u];
      Tt[ww][xj][xi] = nt[u];
    }
    __syncthreads();
    if (tile + gridDim.x < ntiles) load_tile(tile + gridDim.x);
    const int s0 = 2 * w, s1 = 2 * w + 1;
    // U = M_j P_w in single precision (v_mfma_f32_16x16x4f32, as the
    // separate product's FP32 FMAs): lane l gets rows 4 (l / 16) + v of U,
    // column l % 16 -- for each v a Gram k-step over rows {v, v + 4, v + 8,
    // v + 12}, with A = conj(T_w[4 (l / 16) + v][l % 16]).
    big_f4 ur0 = {0.f, 0.f, 0.f, 0.f}, ui0 = ur0, ur1 = ur0, ui1 = ur0;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const cx<R> p0 = Pt[s0][4 * ks + kq][mi], p1 = Pt[s1][4 * ks + kq][mi];
      const float mr = mvr[ks], mim = mvi[ks];
      const float p0r = (float)p0.re, p0i = (float)p0.im, p1r = (float)p1.re, p1i = (float)p1.im;
      ur0 = __builtin_amdgcn_mfma_f32_16x16x4f32(mr, p0r, ur0, 0, 0, 0);
      ur1 = __builtin_amdgcn_mfma_f32_16x16x4f32(mr, p1r, ur1, 0, 0, 0);
      ui0 = __builtin_amdgcn_mfma_f32_16x16x4f32(mr, p0i, ui0, 0, 0, 0);
      ui1 = __builtin_amdgcn_mfma_f32_16x16x4f32(mr, p1i, ui1, 0, 0, 0);
      ur0 = __builtin_amdgcn_mfma_f32_16x16x4f32(-mim, p0i, ur0, 0, 0, 0);
      ur1 = __builtin_amdgcn_mfma_f32_16x16x4f32(-mim, p1i, ur1, 0, 0, 0);
      ui0 = __builtin_amdgcn_mfma_f32_16x16x4f32(mim, p0r, ui0, 0, 0, 0);
      ui1 = __builtin_amdgcn_mfma_f32_16x16x4f32(mim, p1r, ui1, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double ur0v = ur0[v], ui0v = ui0[v], ur1v = ur1[v], ui1v = ui1[v];
      const cx<R> t0 = Tt[s0][4 * kq + v][mi], t1 = Tt[s1][4 * kq + v][mi];
      const double tr0 = (double)t0.re, ti0 = (double)t0.im, tr1 = (double)t1.re, ti1 = (double)t1.im;
      gr0 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr0, ur0v, gr0, 0, 0, 0);
      gr1 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr1, ur1v, gr1, 0, 0, 0);
      gi0 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr0, ui0v, gi0, 0, 0, 0);
      gi1 = __builtin_amdgcn_mfma_f64_16x16x4f64(tr1, ui1v, gi1, 0, 0, 0);
      gr0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ti0, ui0v, gr0, 0, 0, 0);
      gr1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ti1, ui1v, gr1, 0, 0, 0);
      gi0 = __builtin_amdgcn_mfma_f64_16x16x4f64(-ti0, ur0v, gi0, 0, 0, 0);
      gi1 = __builtin_amdgcn_mfma_f64_16x16x4f64(-ti1, ur1v, gi1, 0, 0, 0);
    }
    __syncthreads();
  }
  const big_d4 gr = gr0 + gr1, gi = gi0 + gi1;
#pragma unroll
  for (int v = 0; v < 4; ++v) red[w * CC + (kq + 4 * v) * CHI + mi] = mk<double>(gr[v], gi[v]);
  __syncthreads();
  cd* pp = part + ((long long)k * gridDim.x + blockIdx.x) * CC;
  for (int q = tid; q < CC; q += blockDim.x) pp[q] = red[q] + red[CC + q] + red[2 * CC + q] + red[3 * CC + q];
}

// Two products of a node update in one pass, chi = 16, complex64 data:
//   dst = src x_a M_a x_b M_b   (a != b),
// over tiles as in k_big_prod_gram16 but twice as wide: the 16 x 16 (x_a,
// x_b) slices S of src for W = 16 consecutive c' (the remaining indices,
// contiguous ones fastest; whole 128-byte lines where the tile rows are runs
// in memory, so no write covers half a line), 512 threads, wave w taking
// slices 2w and 2w + 1.  Per slice on FP32 matrix cores (as the
// separate products' FP32 FMAs): U1 = S M_b^T with A = S (lane l: S[l % 16]
// [4 ks + l / 16]) and B = M_b^T (lane l: M_b[l % 16][4 ks + l / 16]); the
// result registers hold U1 rows 4 (l / 16) + v, column l % 16, which are the
// B operands of U = M_a U1 when the k-steps run over those rows (step v:
// A = M_a[l % 16][4 (l / 16) + v]).  U goes back over the slice in LDS and
// out along memory: one HBM pass where two separate products take two.
constexpr int kProd2Threads = 512;
template <typename R>
__global__ void __launch_bounds__(kProd2Threads)
    k_big_prod2_16(DevGraph g, const int* nodes, int ax_a, int ax_b, int src_id, int dst_id, int nb, const cx<R>* T,
                   const cx<R>* n2e_old, const cx<R>* e2n_old, const cx<R>* n2e_new, const uint8_t* node_src,
                   cx<R>* buf, long long smax, const int* inst_active) {
  constexpr int CHI = 16, CC = CHI * CHI, W = 16;
  __shared__ cx<R> St[W][CHI][CHI + 1];  // [slice][x_a][x_b]
  const int k = blockIdx.y;
  const int node = nodes[k];
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, mi = lane & 15, kq = lane >> 4;
  const long long S = ipow(CHI, d) * g.phys;
  const int a_hi = ax_a < ax_b ? ax_a : ax_b, a_lo = ax_a < ax_b ? ax_b : ax_a;
  const long long n_bot = ipow(CHI, d - 1 - a_lo) * g.phys;  // contiguous indices after both axes
  const long long n_mid = ipow(CHI, a_lo - a_hi - 1);
  const long long st_lo = n_bot, st_mid = CHI * n_bot, st_hi = n_mid * CHI * n_bot, st_top = CHI * st_hi;
  const long long ncp = S / CC, ntiles = ncp / W;
  const cx<R>* src = src_id < 0 ? T + g.toff[node] : buf + ((long long)nb * k + src_id) * smax;
  cx<R>* dst = buf + ((long long)nb * k + dst_id) * smax;
  // the incoming messages of both axes (ket side, as k_big_plan_product):
  // mb[ks] = M_b[l % 16][4 ks + l / 16], ma[v] = M_a[l % 16][4 (l / 16) + v]
  cd ma[4], mb[4];
  {
    const int base = g.adj_ptr[node];
    auto msg = [&](int axis) {
      const int e = g.adj_edge[base + axis], sd = g.adj_side[base + axis];
      const int code = node_src ? node_src[base + axis] : 0;
      return code == 0 ? e2n_old + (2LL * e + sd) * CC : (code == 1 ? n2e_new : n2e_old) + (2LL * e + 1 - sd) * CC;
    };
    const cx<R>* m_a = msg(ax_a);
    const cx<R>* m_b = msg(ax_b);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      mb[v] = cvt<double>(m_b[mi * CHI + 4 * v + kq]);
      ma[v] = cvt<double>(m_a[mi * CHI + 4 * kq + v]);
    }
  }
  // each thread's PER elements of a tile: slice q % W, (x_hi, x_lo) from q / W
  // (as k_big_prod_gram16; 32-bit offsets: the host fuses only tensors below
  // 2^31 elements)
  // (element u: x_hi = q / (W CHI) steps by two per u, slice and x_lo fixed)
  constexpr int PER = W * CC / kProd2Threads;
  static_assert(kProd2Threads % (W * CHI) == 0, "element u of a thread keeps its slice and x_lo");
  static_assert(kProd2Threads / 64 * 2 == W, "two slices per wave");
  constexpr int XH_STEP = kProd2Threads / (W * CHI);
  cx<R> np[PER];
  const int ww = tid % W, xl = (tid / W) % CHI, xh0 = tid / (W * CHI);
  const bool a_first = ax_a == a_hi;
  const unsigned offx = (unsigned)(xh0 * st_hi + xl * st_lo), ostep = (unsigned)(XH_STEP * st_hi);
  const unsigned nbot = (unsigned)n_bot, nmid = (unsigned)n_mid, smid = (unsigned)st_mid, stop = (unsigned)st_top;
  // offset of the thread's element 0 of a tile (element u: + u ostep)
  auto tile_off = [&](long long tile) -> unsigned {
    const unsigned cp = (unsigned)(tile * W) + ww;
    const unsigned t2 = cp / nbot;
    return (t2 / nmid) * stop + (t2 % nmid) * smid + cp % nbot + offx;
  };
  // the thread's element u in the slice tile St[ww][x_a][x_b]
  auto st_at = [&](int u) -> cx<R>& {
    const int xh = xh0 + XH_STEP * u;
    return St[ww][a_first ? xh : xl][a_first ? xl : xh];
  };
  const unsigned tb = big_xcd_tile(blockIdx.x, gridDim.x);
  if (tb < ntiles) {
    const unsigned o = tile_off(tb);
#pragma unroll
    for (int u = 0; u < PER; ++u) np[u] = src[o + u * ostep];
  }
  const int s0 = 2 * w, s1 = 2 * w + 1;
  for (long long tile = tb; tile < ntiles; tile += gridDim.x) {
#pragma unroll
    for (int u = 0; u < PER; ++u) st_at(u) = np[u];
    __syncthreads();
    if (tile + gridDim.x < ntiles) {
      const unsigned o = tile_off(tile + gridDim.x);
#pragma unroll
      for (int u = 0; u < PER; ++u) np[u] = src[o + u * ostep];
    }
    // U1 = S M_b^T
    big_f4 ur0 = {0.f, 0.f, 0.f, 0.f}, ui0 = ur0, ur1 = ur0, ui1 = ur0;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const cx<R> p0 = St[s0][mi][4 * ks + kq], p1 = St[s1][mi][4 * ks + kq];
      const float br = (float)mb[ks].re, bi = (float)mb[ks].im;
      const float p0r = (float)p0.re, p0i = (float)p0.im, p1r = (float)p1.re, p1i = (float)p1.im;
      ur0 = __builtin_amdgcn_mfma_f32_16x16x4f32(p0r, br, ur0, 0, 0, 0);
      ur1 = __builtin_amdgcn_mfma_f32_16x16x4f32(p1r, br, ur1, 0, 0, 0);
      ui0 = __builtin_amdgcn_mfma_f32_16x16x4f32(p0r, bi, ui0, 0, 0, 0);
      ui1 = __builtin_amdgcn_mfma_f32_16x16x4f32(p1r, bi, ui1, 0, 0, 0);
      ur0 = __builtin_amdgcn_mfma_f32_16x16x4f32(-p0i, bi, ur0, 0, 0, 0);
      ur1 = __builtin_amdgcn_mfma_f32_16x16x4f32(-p1i, bi, ur1, 0, 0, 0);
      ui0 = __builtin_amdgcn_mfma_f32_16x16x4f32(p0i, br, ui0, 0, 0, 0);
      ui1 = __builtin_amdgcn_mfma_f32_16x16x4f32(p1i, br, ui1, 0, 0, 0);
    }
    // U = M_a U1, k-steps over the U1 rows 4 (l / 16) + v each lane holds
    big_f4 vr0 = {0.f, 0.f, 0.f, 0.f}, vi0 = vr0, vr1 = vr0, vi1 = vr0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float ar = (float)ma[v].re, ai = (float)ma[v].im;
      vr0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, ur0[v], vr0, 0, 0, 0);
      vr1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, ur1[v], vr1, 0, 0, 0);
      vi0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, ui0[v], vi0, 0, 0, 0);
      vi1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, ui1[v], vi1, 0, 0, 0);
      vr0 = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, ui0[v], vr0, 0, 0, 0);
      vr1 = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, ui1[v], vr1, 0, 0, 0);
      vi0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, ur0[v], vi0, 0, 0, 0);
      vi1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, ur1[v], vi1, 0, 0, 0);
    }
    // over this wave's own slices (every read of them above is complete: the
    // products depend on it), then out along memory once every slice is written
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      St[s0][4 * kq + v][mi] = mk<R>((R)vr0[v], (R)vi0[v]);
      St[s1][4 * kq + v][mi] = mk<R>((R)vr1[v], (R)vi1[v]);
    }
    __syncthreads();
    const unsigned o = tile_off(tile);
#pragma unroll
    for (int u = 0; u < PER; ++u) dst[o + u * ostep] = st_at(u);
    __syncthreads();
  }
}

template <typename R, int CHI>
__global__ void __launch_bounds__(kBigThreads)
    k_big_gram_reduce(DevGraph g, const int* nodes, int idx, const cd* part, int chunks, cx<R>* n2e_new,
                      const int* inst_active) {
  constexpr int CC = CHI * CHI;
  __shared__ cd G[CC];
  __shared__ cd tr;
  const int k = blockIdx.x;
  const int node = nodes[k];
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  if (idx >= g.deg[node]) return;
  for (int q = threadIdx.x; q < CC; q += blockDim.x) {
    cd s = mk<double>(0, 0);
    for (int c = 0; c < chunks; ++c) s = s + part[((long long)k * chunks + c) * CC + q];
    G[q] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    cd s = mk<double>(0, 0);
    for (int a = 0; a < CHI; ++a) s = s + G[a * CHI + a];
    tr = s;
  }
  __syncthreads();
  const int base = g.adj_ptr[node];
  const int e = g.adj_edge[base + idx], sd = g.adj_side[base + idx];
  cx<R>* dst = n2e_new + (2LL * e + sd) * CC;
  for (int q = threadIdx.x; q < CC; q += blockDim.x) dst[q] = cvt<R>(cdiv(G[q], tr));
}

// Absorb epilogue: per-chunk sums of |x|^2 of the final product, then
// Gamma = x / |x| written to the node's slot.
template <typename R>
__global__ void __launch_bounds__(kBigThreads)
    k_big_norm(DevGraph g, const int* nodes, const cx<R>* buf, long long smax, int chi, double* part) {
  __shared__ double red[kBigThreads / 64];
  const int k = blockIdx.y;
  const int d = g.deg[nodes[k]];
  const long long S = ipow(chi, d) * g.phys;
  const cx<R>* x = buf + (2LL * k + ((d - 1) & 1)) * smax;
  double acc = 0.0;
  if (d >= 1)
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < S; f += (long long)gridDim.x * blockDim.x)
      acc += (double)norm2(x[f]);
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[(long long)k * gridDim.x + blockIdx.x] = acc;
}

template <typename R>
__global__ void __launch_bounds__(kBigThreads)
    k_big_scale_out(DevGraph g, const int* nodes, const cx<R>* buf, long long smax, int chi, const double* part,
                    int chunks, cx<R>* gam, const int* inst_active) {
  const int k = blockIdx.y;
  const int node = nodes[k];
  if (inst_active && !inst_active[g.node_inst[node]]) return;
  const int d = g.deg[node];
  if (d < 1) return;
  // the node's squared norm: the chunk sums split over the block, then one
  // block reduction (uniform per block: the returns above depend on the node)
  __shared__ double red[kBigThreads / 64];
  double s = 0.0;
  for (int c = threadIdx.x; c < chunks; c += blockDim.x) s += part[(long long)k * chunks + c];
  s = block_sum(s, red);
  const R inv = (R)(1.0 / sqrt(s));
  const long long S = ipow(chi, d) * g.phys;
  const cx<R>* x = buf + (2LL * k + ((d - 1) & 1)) * smax;
  cx<R>* out = gam + g.toff[node];
  for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < S; f += (long long)gridDim.x * blockDim.x)
    out[f] = inv * x[f];
}

// ------------------------------------------------------------- launchers
// Blocks per node of the absorb's norm and scaling passes: at least 8192
// elements per block (one block per 256 elements left 256 blocks and as many
// partial sums per degree-3 node at chi 32).
static int big_norm_blocks(long long smax) {
  long long b = (smax + 8191) / 8192;
  return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}
static int big_col_blocks(long long cols) {
  long long b = (cols + kBigThreads - 1) / kBigThreads;
  return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}
// Gram partials per node: at least 512 columns per chunk, so that small node
// tensors (degree 3 at chi 32: 2048 columns) write a few partial Grams
// instead of one per 64-column tile; at most 256 chunks.
constexpr int kBigGramCols = 512;
static int big_gram_chunks(long long cols) {
  long long b = (cols + kBigGramCols - 1) / kBigGramCols;
  return (int)(b < 1 ? 1 : (b > 256 ? 256 : b));
}

bool big_applicable(const DevGraph& g, int max_deg) {
  return (g.chi == 16 || g.chi == 32) && ipow(g.chi, max_deg) * g.phys >= (1LL << 15);
}

// The node update of a degree-d node as a plan of products and Grams that
// shares partial products between output slots: the slots are split in two
// halves, the other half's messages are applied once for each, recursively.
// Output slot idx still receives T with every other message applied, so the
// result is that of pass_through_node (tensor_ops.py:36-60) with the products
// in another order.  d = 6: 16 products (instead of 6 x 5), three buffers;
// d = 3: 5 products (instead of 6), two buffers.
namespace {
struct BigOp {
  bool gram;  // false: dst = src x_axis M; true: Gram of slot `axis` from src
  int src, dst, axis;
};
struct BigPlanner {
  std::vector<BigOp> ops;
  std::vector<int> free_list;
  int nbuf = 0;
  int alloc() {
    if (!free_list.empty()) {
      const int b = free_list.back();
      free_list.pop_back();
      return b;
    }
    return nbuf++;
  }
  void release(int b) {
    if (b >= 0) free_list.push_back(b);
  }
  int chain(int src, const std::vector<int>& axes) {
    int cur = src;
    for (int a : axes) {
      const int dst = alloc();
      ops.push_back({false, cur, dst, a});
      if (cur != src) release(cur);
      cur = dst;
    }
    return cur;
  }
  void solve(int src, const std::vector<int>& S) {
    if (S.size() == 1) {
      ops.push_back({true, src, -1, S[0]});
      return;
    }
    const size_t h = S.size() / 2;
    const std::vector<int> S1(S.begin(), S.begin() + h), S2(S.begin() + h, S.end());
    const int x1 = chain(src, S2);
    solve(x1, S1);
    if (x1 != src) release(x1);
    const int x2 = chain(src, S1);
    solve(x2, S2);
    if (x2 != src) release(x2);
  }
};
BigPlanner big_plan(int d) {
  BigPlanner p;
  std::vector<int> S;
  for (int i = 0; i < d; ++i) S.push_back(i);
  if (d >= 1) p.solve(-1, S);
  return p;
}
}  // namespace

long long big_ws_bytes(const DevGraph& g, int max_deg, int count, int rb) {
  const long long smax = ipow(g.chi, max_deg) * g.phys;
  const long long cols = smax / g.chi;
  const long long part = (long long)count * big_gram_chunks(cols) * g.chi * g.chi * (long long)sizeof(cd);
  const long long norm = (long long)count * big_norm_blocks(smax) * (long long)sizeof(double);
  // node-update plan buffers (the absorb uses two)
  int nb = big_plan(max_deg).nbuf;
  if (nb < 2) nb = 2;
  return (long long)nb * count * smax * 2 * rb + (part > norm ? part : norm) + 256;
}

template <typename R, int CHI>
static hipError_t big_bp_level(const DevGraph& g, int min_deg, int max_deg, const int* nodes, int count,
                               const void* T, const void* n2e_old, const void* e2n_old, void* n2e_new,
                               const uint8_t* node_src, void* ws, const int* inst_active, hipStream_t s) {
  const long long smax = ipow(CHI, max_deg) * g.phys;
  const long long cols = smax / CHI;
  const BigPlanner plan = big_plan(max_deg);
  const int nb = plan.nbuf < 2 ? 2 : plan.nbuf;
  cx<R>* buf = (cx<R>*)ws;
  cd* part = (cd*)((char*)ws + (long long)nb * count * smax * sizeof(cx<R>));
  const int chunks = big_gram_chunks(cols);
  const int pblocks = big_col_blocks(cols);
  // chi 16, every node of the plan's degree: a product whose result only
  // feeds the next Gram runs fused with it
  const bool fuse = sizeof(R) == 4 && CHI == 16 && min_deg == max_deg && max_deg >= 4 && smax < (1LL << 31);
  if (fuse && (max_deg == 6 || max_deg == 4)) {
    // The messages in pairs {0,1}, {2,3}, {4,5}: slots a, b of a pair take
    // Q = T x (every other pair), then their partner's message fused with the
    // Gram.  Degree 6: Q_{01} = (T x23) x45, Q_{23} = (T x01) x45,
    // Q_{45} = (T x01) x23 -- five two-product passes and six product-Gram
    // passes (22 tensor reads and writes per node against 32 for the plan
    // above, same 16 products); degree 4: Q_{01} = T x23, Q_{23} = T x01.
    auto prod2 = [&](int a, int b, int src, int dst) {
      hipLaunchKernelGGL((k_big_prod2_16<R>), dim3(chunks, count), dim3(kProd2Threads), 0, s, g, nodes, a, b, src, dst,
                         nb, (const cx<R>*)T, (const cx<R>*)n2e_old, (const cx<R>*)e2n_old, (const cx<R>*)n2e_new,
                         node_src, buf, smax, inst_active);
    };
    auto pgram = [&](int j, int idx, int src) {
      hipLaunchKernelGGL((k_big_prod_gram16<R>), dim3(chunks, count), dim3(kBigThreads), 0, s, g, nodes, j, idx, src,
                         nb, (const cx<R>*)T, (const cx<R>*)n2e_old, (const cx<R>*)e2n_old, (const cx<R>*)n2e_new,
                         node_src, (const cx<R>*)buf, smax, part, inst_active);
      hipLaunchKernelGGL((k_big_gram_reduce<R, CHI>), dim3(count), dim3(kBigThreads), 0, s, g, nodes, idx,
                         (const cd*)part, chunks, (cx<R>*)n2e_new, inst_active);
    };
    if (max_deg == 6) {
      prod2(0, 1, -1, 0);  // buffer 0: T x01
      prod2(2, 3, -1, 1);  // buffer 1: T x23
      prod2(4, 5, 1, 2);   // buffer 2: T x2345
      pgram(1, 0, 2);
      pgram(0, 1, 2);
      prod2(4, 5, 0, 1);  // buffer 1: T x0145
      pgram(3, 2, 1);
      pgram(2, 3, 1);
      prod2(2, 3, 0, 2);  // buffer 2: T x0123
      pgram(5, 4, 2);
      pgram(4, 5, 2);
    } else {
      prod2(2, 3, -1, 0);  // T x23
      pgram(1, 0, 0);
      pgram(0, 1, 0);
      prod2(0, 1, -1, 0);  // T x01
      pgram(3, 2, 0);
      pgram(2, 3, 0);
    }
    return hipGetLastError();
  }
  for (size_t i = 0; i < plan.ops.size(); ++i) {
    const BigOp& op = plan.ops[i];
    if (fuse && !op.gram && i + 1 < plan.ops.size() && plan.ops[i + 1].gram && plan.ops[i + 1].src == op.dst) {
      const int idx = plan.ops[i + 1].axis;
      hipLaunchKernelGGL((k_big_prod_gram16<R>), dim3(chunks, count), dim3(kBigThreads), 0, s, g, nodes, op.axis,
                         idx, op.src, nb, (const cx<R>*)T, (const cx<R>*)n2e_old, (const cx<R>*)e2n_old,
                         (const cx<R>*)n2e_new, node_src, (const cx<R>*)buf, smax, part, inst_active);
      hipLaunchKernelGGL((k_big_gram_reduce<R, CHI>), dim3(count), dim3(kBigThreads), 0, s, g, nodes, idx,
                         (const cd*)part, chunks, (cx<R>*)n2e_new, inst_active);
      ++i;
      continue;
    }
    if (!op.gram) {
      hipLaunchKernelGGL((k_big_plan_product<R, CHI>), dim3(pblocks, count), dim3(kBigThreads), 0, s, g, nodes,
                         op.axis, op.src, op.dst, nb, (const cx<R>*)T, (const cx<R>*)n2e_old, (const cx<R>*)e2n_old,
                         (const cx<R>*)n2e_new, node_src, buf, smax, inst_active);
    } else {
      hipLaunchKernelGGL((k_big_gram<R, CHI>), dim3(chunks, count), dim3(kBigThreads), 0, s, g, nodes, op.axis,
                         op.src, nb, (const cx<R>*)T, (const cx<R>*)buf, smax, part, inst_active);
      hipLaunchKernelGGL((k_big_gram_reduce<R, CHI>), dim3(count), dim3(kBigThreads), 0, s, g, nodes, op.axis,
                         (const cd*)part, chunks, (cx<R>*)n2e_new, inst_active);
    }
  }
  return hipGetLastError();
}

hipError_t launch_bp_nodes_big(int rb, const DevGraph& g, int min_deg, int max_deg, const int* nodes, int count,
                               const void* T, const void* n2e_old, const void* e2n_old, void* n2e_new,
                               const uint8_t* node_src, void* ws, const int* inst_active, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (rb == 4)
    return g.chi == 16 ? big_bp_level<float, 16>(g, min_deg, max_deg, nodes, count, T, n2e_old, e2n_old, n2e_new, node_src,
                                                 ws, inst_active, s)
                       : big_bp_level<float, 32>(g, min_deg, max_deg, nodes, count, T, n2e_old, e2n_old, n2e_new, node_src,
                                                 ws, inst_active, s);
  return g.chi == 16 ? big_bp_level<double, 16>(g, min_deg, max_deg, nodes, count, T, n2e_old, e2n_old, n2e_new, node_src, ws,
                                                inst_active, s)
                     : big_bp_level<double, 32>(g, min_deg, max_deg, nodes, count, T, n2e_old, e2n_old, n2e_new, node_src, ws,
                                                inst_active, s);
}

template <typename R, int CHI>
static hipError_t big_absorb(const DevGraph& g, int max_deg, const int* nodes, int count, const void* T,
                             const cd* X, void* gam, void* ws, const int* inst_active, hipStream_t s) {
  const long long smax = ipow(CHI, max_deg) * g.phys;
  const long long cols = smax / CHI;
  cx<R>* buf = (cx<R>*)ws;
  double* part = (double*)((char*)ws + 2LL * count * smax * sizeof(cx<R>));
  for (int t = 0; t < max_deg; ++t)
    hipLaunchKernelGGL((k_big_absorb_product<R, CHI>), dim3(big_col_blocks(cols), count), dim3(kBigThreads), 0, s, g,
                       nodes, t, (const cx<R>*)T, X, buf, smax, 2, inst_active);
  const int nb = big_norm_blocks(smax);
  hipLaunchKernelGGL((k_big_norm<R>), dim3(nb, count), dim3(kBigThreads), 0, s, g, nodes, (const cx<R>*)buf, smax,
                     CHI, part);
  hipLaunchKernelGGL((k_big_scale_out<R>), dim3(nb, count), dim3(kBigThreads), 0, s, g, nodes, (const cx<R>*)buf,
                     smax, CHI, (const double*)part, nb, (cx<R>*)gam, inst_active);
  return hipGetLastError();
}

hipError_t launch_absorb_big(int rb, const DevGraph& g, int max_deg, const int* nodes, int count, const void* T,
                             const cd* X, void* gam, void* ws, const int* inst_active, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (rb == 4)
    return g.chi == 16 ? big_absorb<float, 16>(g, max_deg, nodes, count, T, X, gam, ws, inst_active, s)
                       : big_absorb<float, 32>(g, max_deg, nodes, count, T, X, gam, ws, inst_active, s);
  return g.chi == 16 ? big_absorb<double, 16>(g, max_deg, nodes, count, T, X, gam, ws, inst_active, s)
                     : big_absorb<double, 32>(g, max_deg, nodes, count, T, X, gam, ws, inst_active, s);
}

}  // namespace dq

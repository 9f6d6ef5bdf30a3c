// BP node update on CDNA4 matrix cores for degree-3 nodes, chi = 16,
// phys = 2, complex64 (FP32 accumulation).
//
// Same arithmetic as the reference pass_through_node
// (dequantizer/src/tensor_ops.py:36-60): for every bond idx,
// U = T x_{j != idx} M_j (message j applied on the ket side), out[k,l] =
// sum conj(T)[..k..] U[..l..] over the other axes, out /= trace(out).
//
// Persistent 512-thread blocks (8 waves, two per SIMD so that one wave's LDS
// fetches overlap the other's MFMAs), one per CU, each looping over its
// nodes.  T (16x16x16x2 complex) and one work tensor live in LDS with a
// padded layout: element (b0, b1, b2, s) at b0*529 + b1*33 + b2*2 + s
// (complex), so that the 16 lanes of an MFMA operand fetch that walk along
// any bond axis hit distinct banks.  The two tensors take 135 KiB of the
// CU's 160 KiB (with the incoming messages and the block's node metadata,
// 142 KiB).
//   mode product  X x_j M: a GEMM (16 x 16) x (16 x 512) split in 32 column
//                 tiles; per tile 4 k-steps of V_MFMA_F32_16X16X4_F32, each
//                 complex product = 4 real MFMAs;
//   Gram          (16 x 512) x (512 x 16), K split over the 8 waves and
//                 reduced through LDS.
// Per node: 5 mode products + 3 Grams.  X = T x_2 M2 is stored once and
// serves U_0 = X x_1 M1 and U_1 = X x_0 M0; then X = T x_0 M0 serves
// U_2 = X x_1 M1.  The second product of each U is fused with its Gram: the
// 16 x 16 product tile leaves the matrix cores in the layout the Gram's
// B operand needs (lane l holds row 4(l/16)+v, column l%16), so U never goes
// through LDS.
#include "dq_fast64.hpp"
#include "dq_launch.hpp"

namespace dq {

typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace fast3 {
constexpr int RP = 33;              // pitch of a (b0, b1) fiber of 32 complex values
constexpr int B0P = 16 * RP + 1;    // pitch of b0
constexpr int NPAD = 16 * B0P;      // padded complex elements of one tensor
constexpr int S = 16 * 16 * 16 * 2;  // unpadded complex elements

__device__ __forceinline__ int axis_stride(int j) { return j == 0 ? B0P : (j == 1 ? RP : 2); }

// padded offset of column c (0..511) of the matrix view of mode j: columns
// enumerate the remaining axes in row-major order.
__device__ __forceinline__ int col_base(int j, int c) {
  if (j == 0) return (c >> 5) * RP + (c & 31);
  if (j == 1) return (c >> 5) * B0P + (c & 31);
  return (c >> 5) * B0P + ((c >> 1) & 15) * RP + (c & 1);
}

__device__ __forceinline__ floatx4 mfma(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// dst = src x_j M (may be in place).  m[ks] = M[lane%16][4*ks + lane/16].
template <int W>
__device__ __forceinline__ void mode_product(float2* dst, const float2* src, const float2 (&m)[4], int j, int wave,
                                             int lane) {
  const int st = axis_stride(j);
  const int col = lane & 15, kq = lane >> 4;
  for (int tile = wave; tile < 32; tile += W) {
    const int cb = col_base(j, tile * 16 + col);
    floatx4 ar = {0.f, 0.f, 0.f, 0.f}, ai = {0.f, 0.f, 0.f, 0.f};
    float2 x[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) x[ks] = src[cb + (4 * ks + kq) * st];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      ar = mfma(m[ks].x, x[ks].x, ar);
      ar = mfma(-m[ks].y, x[ks].y, ar);
      ai = mfma(m[ks].x, x[ks].y, ai);
      ai = mfma(m[ks].y, x[ks].x, ai);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[cb + (4 * kq + v) * st] = make_float2(ar[v], ai[v]);
  }
}

// U = X x_j M fused with the partial Gram of this wave over axis idx
// (accumulated into gr, gi): G[k][l] += sum conj(T[k@idx, ...]) U[l@idx, ...].
// Tiles enumerate (b_tax, s) for the third bond axis tax; within a tile lane
// l%16 runs along axis idx.  The product's result D[4(l/16)+v][l%16] =
// U[b_j = 4(l/16)+v][b_idx = l%16] is, for each v, the B operand of a Gram
// k-step over b_j in {v, 4+v, 8+v, 12+v}; the A operand is conj(T) at the
// same (b_j, b_tax, s) with b_idx = l%16.
template <int W>
__device__ __forceinline__ void mode_gram(const float2* t, const float2* x, const float2 (&m)[4], int j, int idx,
                                          int tax, int wave, int lane, floatx4& gr, floatx4& gi) {
  const int sj = axis_stride(j);
  const int y = lane & 15, kq = lane >> 4;
  const int lane_base = y * axis_stride(idx);
#pragma unroll
  for (int tile = wave; tile < 32; tile += W) {
    const int cb = lane_base + (tile >> 1) * axis_stride(tax) + (tile & 1);
    float2 xv[4], tv[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) xv[ks] = x[cb + (4 * ks + kq) * sj];
#pragma unroll
    for (int v = 0; v < 4; ++v) tv[v] = t[cb + (4 * kq + v) * sj];
    floatx4 ur = {0.f, 0.f, 0.f, 0.f}, ui = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      ur = mfma(m[ks].x, xv[ks].x, ur);
      ur = mfma(-m[ks].y, xv[ks].y, ur);
      ui = mfma(m[ks].x, xv[ks].y, ui);
      ui = mfma(m[ks].y, xv[ks].x, ui);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      // conj(a) * u
      gr = mfma(tv[v].x, ur[v], gr);
      gr = mfma(tv[v].y, ui[v], gr);
      gi = mfma(tv[v].x, ui[v], gi);
      gi = mfma(-tv[v].y, ur[v], gi);
    }
  }
}

}  // namespace fast3

template <int NW>
__global__ void __launch_bounds__(64 * NW)
    k_bp_nodes_d3c16(DevGraph g, const int* nodes, int count, const float2* T, const float2* n2e_old,
                     const float2* e2n_old, float2* n2e_new, const uint8_t* node_src, const int* inst_active,
                     float2* e2n_copy) {
  using namespace fast3;
  constexpr int NT = 64 * NW;   // threads
  constexpr int PER = S / NT;   // tensor elements staged per thread
  __shared__ float2 tl[NPAD];
  __shared__ float2 wk[NPAD];
  __shared__ float2 diag[3 * 16];
  __shared__ float2 mb[3 * 16 * 17];  // incoming messages, row pitch 17
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // Persistent blocks (one per CU: the two tensors take most of the LDS):
  // block b updates nodes[b], nodes[b + gridDim.x], ...  The next node's
  // tensor is loaded into registers while the current one is processed, and
  // its three incoming messages while the current outputs are reduced, so
  // their HBM latency is hidden behind the matrix-core work.
  // Graph metadata of this block's first MT nodes (node j is list entry
  // blockIdx.x + j * gridDim.x) is gathered once into LDS at the start, so
  // that the per-node chain of dependent metadata reads (list entry, node
  // offset, adjacency) leaves the critical path; later nodes read it through
  // scalar loads.
  constexpr int MT = 160;
  static_assert(MT <= NT, "one thread per gathered node");
  __shared__ long long mt_toff[MT];
  __shared__ int mt_valid[MT], mt_e[MT][3], mt_s[MT][3], mt_code[MT][3];
  const int nmine = count > (int)blockIdx.x ? (count - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x : 0;
  if (tid < MT) {
    const int j = tid;
    int valid = 0;
    if (j < nmine) {
      const int node = nodes[blockIdx.x + (long long)j * gridDim.x];
      valid = !inst_active || inst_active[g.node_inst[node]];
      const int base = g.adj_ptr[node];
      mt_toff[j] = g.toff[node];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        mt_e[j][k] = g.adj_edge[base + k];
        mt_s[j][k] = g.adj_side[base + k];
        mt_code[j][k] = node_src ? node_src[base + k] : 0;
      }
    }
    mt_valid[j] = valid;
  }
  __syncthreads();
  auto node_of = [&](int j) { return ldc(nodes, blockIdx.x + (long long)j * gridDim.x); };
  auto valid_of = [&](int j) {
    if (j < MT) return mt_valid[j] != 0;
    return !inst_active || ldc(inst_active, ldc(g.node_inst, node_of(j))) != 0;
  };
  auto toff_of = [&](int j) { return j < MT ? mt_toff[j] : ldc(g.toff, node_of(j)); };
  auto adj_of = [&](int j, int (&e)[3], int (&sd)[3], int (&code)[3]) {
    if (j < MT) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        e[k] = mt_e[j][k];
        sd[k] = mt_s[j][k];
        code[k] = mt_code[j][k];
      }
    } else {
      const int base = ldc(g.adj_ptr, node_of(j));
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        e[k] = ldc(g.adj_edge, base + k);
        sd[k] = ldc(g.adj_side, base + k);
        code[k] = node_src ? ldc(node_src, base + k) : 0;
      }
    }
  };
  constexpr int QM = (768 + NT - 1) / NT;  // message elements per thread
  auto fetch_msgs = [&](int j, float2 (&r)[QM]) {
    int e[3], sd[3], code[3];
    adj_of(j, e, sd, code);
    const float2* src[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      src[k] = code[k] == 0 ? e2n_old + (2LL * e[k] + sd[k]) * 256
                            : (code[k] == 1 ? n2e_new : n2e_old) + (2LL * e[k] + 1 - sd[k]) * 256;
#pragma unroll
    for (int qi = 0; qi < QM; ++qi) {
      const int q = tid + qi * NT;
      if (q < 768) {
        const int k = q >> 8;
        r[qi] = (k == 0 ? src[0] : (k == 1 ? src[1] : src[2]))[q & 255];
      }
    }
  };
  auto store_msgs = [&](const float2 (&r)[QM]) {
#pragma unroll
    for (int qi = 0; qi < QM; ++qi) {
      const int q = tid + qi * NT;
      if (q < 768) mb[(q >> 8) * 272 + ((q >> 4) & 15) * 17 + (q & 15)] = r[qi];
    }
  };
  int j = 0;
  while (j < nmine && !valid_of(j)) ++j;
  if (j >= nmine) return;
  // the node tensor moves as 16-byte pairs (elements 2i, 2i+1: same b0, b1)
  float4 pre[PER / 2];
  {
    const float4* tg = (const float4*)(T + toff_of(j));
#pragma unroll
    for (int k = 0; k < PER / 2; ++k) pre[k] = tg[tid + k * NT];
    float2 mr[QM];
    fetch_msgs(j, mr);
    store_msgs(mr);
  }
  while (true) {
    // stage T into LDS (padded); the previous node's last reads of tl are
    // behind the barrier that precedes its partial writes
#pragma unroll
    for (int k = 0; k < PER / 2; ++k) {
      const int f = 2 * (tid + k * NT);
      const int r = f & 31, b1 = (f >> 5) & 15, b0 = f >> 9;
      float2* d = tl + b0 * B0P + b1 * RP + r;
      d[0] = make_float2(pre[k].x, pre[k].y);
      d[1] = make_float2(pre[k].z, pre[k].w);
    }
    int oe[3], os[3];
    {
      int code[3];
      adj_of(j, oe, os, code);
    }
    int next = j + 1;
    while (next < nmine && !valid_of(next)) ++next;
    const float4* tg_n = next < nmine ? (const float4*)(T + toff_of(next)) : nullptr;
    __syncthreads();
    // the next node's tensor is requested in four parts, one after each of
    // the next four barriers, so that the issue of its loads (64 KiB through
    // the CU's address path) overlaps matrix-core work instead of stalling
    // one phase
    constexpr int PQ = PER / 8;  // float4 loads per part
    auto prefetch = [&](int part) {
      if (tg_n) {
#pragma unroll
        for (int k = part * PQ; k < (part + 1) * PQ; ++k) pre[k] = tg_n[tid + k * NT];
      }
    };
    prefetch(0);
    // synchronous edge update fused in (e2n_copy != NULL): this node's three
    // edge->node messages of the next sweep are the opposite sides' inputs
    // of this one, e2n_copy[e][s] = n2e_old[e][1 - s]
    float2 cp[QM];
    if (e2n_copy) {
#pragma unroll
      for (int qi = 0; qi < QM; ++qi) {
        const int q = tid + qi * NT;
        if (q < 768) {
          const int k = q >> 8;
          const int e = k == 0 ? oe[0] : (k == 1 ? oe[1] : oe[2]), sd = k == 0 ? os[0] : (k == 1 ? os[1] : os[2]);
          cp[qi] = n2e_old[(2LL * e + 1 - sd) * 256 + (q & 255)];
        }
      }
    }
    // message operands: m[j][ks] = M_j[lane%16][4 ks + lane/16]
    float2 m[3][4];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) m[k][ks] = mb[k * 272 + (lane & 15) * 17 + 4 * ks + (lane >> 4)];
    floatx4 gr[3], gi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) gr[k] = gi[k] = floatx4{0.f, 0.f, 0.f, 0.f};
    mode_product<NW>(wk, tl, m[2], 2, wave, lane);  // X = T x_2 M2
    __syncthreads();
    prefetch(1);
    mode_gram<NW>(tl, wk, m[1], 1, 0, 2, wave, lane, gr[0], gi[0]);  // U_0 = X x_1 M1
    mode_gram<NW>(tl, wk, m[0], 0, 1, 2, wave, lane, gr[1], gi[1]);  // U_1 = X x_0 M0
    __syncthreads();
    prefetch(2);
    mode_product<NW>(wk, tl, m[0], 0, wave, lane);  // X = T x_0 M0
    __syncthreads();
    prefetch(3);
    mode_gram<NW>(tl, wk, m[1], 1, 2, 0, wave, lane, gr[2], gi[2]);  // U_2 = X x_1 M1
    float2 mr[QM];
    if (next < nmine) fetch_msgs(next, mr);
    __syncthreads();
    // wave partials of the three Grams into the work tensor, summed over the
    // waves, divided by their traces
    float2* part = wk;  // [3][NW][256]
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        part[(k * NW + wave) * 256 + (4 * (lane >> 4) + v) * 16 + (lane & 15)] = make_float2(gr[k][v], gi[k][v]);
    __syncthreads();
    constexpr int QN = (768 + NT - 1) / NT;
    float2 sum[QN];
#pragma unroll
    for (int qi = 0; qi < QN; ++qi) {
      const int q = tid + qi * NT;
      sum[qi] = make_float2(0.f, 0.f);
      if (q < 768) {
        const int k = q >> 8, e = q & 255;
        for (int w = 0; w < NW; ++w) {
          sum[qi].x += part[(k * NW + w) * 256 + e].x;
          sum[qi].y += part[(k * NW + w) * 256 + e].y;
        }
        if (e % 17 == 0) diag[k * 16 + e / 17] = sum[qi];
      }
    }
    __syncthreads();
#pragma unroll
    for (int qi = 0; qi < QN; ++qi) {
      const int q = tid + qi * NT;
      if (q < 768) {
        const int k = q >> 8, e = q & 255;
        float tr_r = 0.f, tr_i = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          tr_r += diag[k * 16 + c].x;
          tr_i += diag[k * 16 + c].y;
        }
        const float den = tr_r * tr_r + tr_i * tr_i;
        float2 o;
        o.x = (sum[qi].x * tr_r + sum[qi].y * tr_i) / den;
        o.y = (sum[qi].y * tr_r - sum[qi].x * tr_i) / den;
        n2e_new[(2LL * oe[k] + os[k]) * 256 + e] = o;
      }
    }
    if (e2n_copy) {
#pragma unroll
      for (int qi = 0; qi < QM; ++qi) {
        const int q = tid + qi * NT;
        if (q < 768) {
          const int k = q >> 8;
          const int e = k == 0 ? oe[0] : (k == 1 ? oe[1] : oe[2]), sd = k == 0 ? os[0] : (k == 1 ? os[1] : os[2]);
          e2n_copy[(2LL * e + sd) * 256 + (q & 255)] = cp[qi];
        }
      }
    }
    // the next node's staging writes only tl, last read before the barrier
    // that precedes the partial writes; its first write of wk follows the
    // staging barrier, which every thread reaches after its reads of part;
    // mb was last read right after the previous staging barrier
    if (next >= nmine) break;
    store_msgs(mr);
    j = next;
  }
}

// The same node update in complex128 on FP64 matrix cores
// (V_MFMA_F64_16X16X4F64: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
// the result D[l / 16 + 4 v][l % 16] stays in four accumulators).  Messages act
// on bond axes only and the Gram sums over the physical index, so the two
// physical components are processed one after the other: per component the
// 16^3 tensor and one work tensor (complex128, element (b0, b1, b2) at
// b0*273 + b1*17 + b2, bank-spread for 16-byte elements along every axis)
// take 137 KiB of LDS.  Gram partials stay in registers across both passes
// and are reduced over the waves at the end.

__global__ void __launch_bounds__(64 * fast64::NW)
    k_bp_nodes_d3c16_f64(DevGraph g, const int* nodes, const cd* T, const cd* n2e_old, const cd* e2n_old,
                         cd* n2e_new, const uint8_t* node_src, const int* inst_active) {
  using namespace fast64;
  __shared__ cd tl[NPAD];
  __shared__ cd wk[NPAD];
  __shared__ cd diag[16];
  const int node = ldc(nodes, blockIdx.x);
  if (inst_active && !ldc(inst_active, ldc(g.node_inst, node))) return;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int base = ldc(g.adj_ptr, node);
  const cd* tg = T + ldc(g.toff, node);
  cd m[3][4];
  int oe[3], os[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int e = ldc(g.adj_edge, base + k), sd = ldc(g.adj_side, base + k);
    oe[k] = e;
    os[k] = sd;
    const int code = node_src ? ldc(node_src, base + k) : 0;
    const cd* src = code == 0 ? e2n_old + (2LL * e + sd) * 256
                              : (code == 1 ? n2e_new : n2e_old) + (2LL * e + 1 - sd) * 256;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) m[k][ks] = src[(lane & 15) * 16 + 4 * ks + (lane >> 4)];
  }
  d4 gr[3], gi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) gr[k] = gi[k] = d4{0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < 2; ++s) {
    for (int f = tid; f < 4096; f += 64 * NW) tl[(f >> 8) * B0P + ((f >> 4) & 15) * RP + (f & 15)] = tg[2 * f + s];
    __syncthreads();
    // X = T x_2 M2 serves U_0 = X x_1 M1 and U_1 = X x_0 M0; X = T x_0 M0
    // serves U_2 = X x_1 M1; each U's second product is fused with its Gram
    mode_product(wk, tl, m[2], 2, wave, lane);
    __syncthreads();
    mode_gram(tl, wk, m[1], 1, 0, 2, wave, lane, gr[0], gi[0]);
    mode_gram(tl, wk, m[0], 0, 1, 2, wave, lane, gr[1], gi[1]);
    __syncthreads();
    mode_product(wk, tl, m[0], 0, wave, lane);
    __syncthreads();
    mode_gram(tl, wk, m[1], 1, 2, 0, wave, lane, gr[2], gi[2]);
    __syncthreads();
  }
  // reduce the wave partials through LDS (the work tensor is free), one
  // output at a time, then divide by the trace
  cd* part = wk;  // NW x 256
#pragma unroll
  for (int idx = 0; idx < 3; ++idx) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
      part[wave * 256 + ((lane >> 4) + 4 * v) * 16 + (lane & 15)] = mk<double>(gr[idx][v], gi[idx][v]);
    __syncthreads();
    cd sum = mk<double>(0.0, 0.0);
    if (tid < 256) {
      for (int w = 0; w < NW; ++w) sum = sum + part[w * 256 + tid];
      if (tid % 17 == 0) diag[tid / 17] = sum;
    }
    __syncthreads();
    if (tid < 256) {
      cd tr = mk<double>(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 16; ++k) tr = tr + diag[k];
      n2e_new[(2LL * oe[idx] + os[idx]) * 256 + tid] = cdiv(sum, tr);
    }
    __syncthreads();
  }
}

hipError_t launch_bp_nodes_fast64(const DevGraph& g, const int* nodes, int count, const void* T, const void* n2e_old,
                                  const void* e2n_old, void* n2e_new, const uint8_t* node_src, const int* inst_active,
                                  hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bp_nodes_d3c16_f64, dim3(count), dim3(64 * fast64::NW), 0, s, g, nodes, (const cd*)T,
                     (const cd*)n2e_old, (const cd*)e2n_old, (cd*)n2e_new, node_src, inst_active);
  return hipGetLastError();
}

// Vidal absorb in complex128 on FP64 matrix cores for the same node shape:
// per physical component, three in-place products with X_j^T in LDS, the
// unnormalised result written to Gamma; after both components Gamma is
// scaled by 1 / |Y| (re-read from global memory by the block that wrote it).
__global__ void __launch_bounds__(64 * fast64::NW)
    k_absorb_d3c16_f64(DevGraph g, const cd* T, const cd* X, cd* gam, const int* inst_active) {
  using namespace fast64;
  __shared__ cd tl[NPAD];
  __shared__ double red[NW];
  const int node = blockIdx.x;
  if (inst_active && !ldc(inst_active, ldc(g.node_inst, node))) return;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int base = ldc(g.adj_ptr, node);
  const cd* tg = T + ldc(g.toff, node);
  cd* out = gam + ldc(g.toff, node);
  cd m[3][4];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    load_operand(X + (2LL * ldc(g.adj_edge, base + k) + ldc(g.adj_side, base + k)) * 256, 2, lane, m[k]);
  double nn = 0.0;
  for (int s = 0; s < 2; ++s) {
    stage_component(tl, tg, s);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      mode_product(tl, tl, m[j], j, wave, lane);
      __syncthreads();
    }
    for (int f = tid; f < 4096; f += blockDim.x) {
      const cd v = tl[(f >> 8) * B0P + ((f >> 4) & 15) * RP + (f & 15)];
      out[2 * f + s] = v;
      nn += norm2(v);
    }
    __syncthreads();
  }
  const double inv = 1.0 / sqrt(block_sum(nn, red));
  for (int f = tid; f < 8192; f += blockDim.x) out[f] = inv * out[f];
}

hipError_t launch_absorb_fast64(const DevGraph& g, const void* T, const cd* X, void* gam, const int* inst_active,
                                hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_absorb_d3c16_f64, dim3(g.n_nodes), dim3(64 * fast64::NW), 0, s, g, (const cd*)T, X, (cd*)gam,
                     inst_active);
  return hipGetLastError();
}

// Vidal absorb (tensor_ops.py:88-93) on matrix cores for the same shape:
// Gamma = T x_0 X_0^T x_1 X_1^T x_2 X_2^T / |.|, X_j the orthogonaliser of the
// node's side of bond j (complex128, rounded to complex64 operands, FP32
// accumulation like the generic complex64 kernel).  256-thread blocks with a
// single padded tensor in LDS (68 KiB, two blocks per CU), persistent over
// the nodes with the next node's tensor and orthogonalisers loaded into
// registers while the current node is processed: the three mode products run
// in place (each column tile reads and rewrites only its own columns).
__global__ void __launch_bounds__(256, 2)
    k_absorb_d3c16(DevGraph g, const float2* T, const cd* X, float2* gam, const int* inst_active) {
  using namespace fast3;
  constexpr int PER = S / 256;
  __shared__ float2 tl[NPAD];
  __shared__ double red[8];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  auto active = [&](int node) { return !inst_active || ldc(inst_active, ldc(g.node_inst, node)); };
  // m[j][ks] = X_j^T[lane%16][4 ks + lane/16] = X_j[4 ks + lane/16][lane%16]
  auto load = [&](int node, float2 (&pre)[PER], float2 (&m)[3][4]) {
    const float2* tg = T + ldc(g.toff, node);
#pragma unroll
    for (int k = 0; k < PER; ++k) pre[k] = tg[tid + k * 256];
    const int base = ldc(g.adj_ptr, node);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const cd* x = X + (2LL * ldc(g.adj_edge, base + k) + ldc(g.adj_side, base + k)) * 256;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const cd v = x[(4 * ks + (lane >> 4)) * 16 + (lane & 15)];
        m[k][ks] = make_float2((float)v.re, (float)v.im);
      }
    }
  };
  int node = blockIdx.x;
  while (node < g.n_nodes && !active(node)) node += gridDim.x;
  if (node >= g.n_nodes) return;
  float2 pre[PER], mn[3][4];
  load(node, pre, mn);
  while (true) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int f = tid + k * 256;
      const int r = f & 31, b1 = (f >> 5) & 15, b0 = f >> 9;
      tl[b0 * B0P + b1 * RP + r] = pre[k];
    }
    float2 m[3][4];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) m[k][ks] = mn[k][ks];
    int next = node + gridDim.x;
    while (next < g.n_nodes && !active(next)) next += gridDim.x;
    if (next < g.n_nodes) load(next, pre, mn);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      mode_product<4>(tl, tl, m[j], j, wave, lane);
      __syncthreads();
    }
    double acc = 0.0;
    for (int f = tid; f < S; f += 256) {
      const int r = f & 31, b1 = (f >> 5) & 15, b0 = f >> 9;
      const float2 v = tl[b0 * B0P + b1 * RP + r];
      acc += (double)(v.x * v.x + v.y * v.y);
    }
    const float inv = (float)(1.0 / sqrt(block_sum(acc, red)));
    float2* out = gam + ldc(g.toff, node);
    for (int f = tid; f < S; f += 256) {
      const int r = f & 31, b1 = (f >> 5) & 15, b0 = f >> 9;
      const float2 v = tl[b0 * B0P + b1 * RP + r];
      out[f] = make_float2(inv * v.x, inv * v.y);
    }
    if (next >= g.n_nodes) break;
    __syncthreads();  // the next staging overwrites tl
    node = next;
  }
}

hipError_t launch_absorb_fast(const DevGraph& g, const void* T, const cd* X, void* gam, const int* inst_active,
                              hipStream_t s) {
  if (g.n_nodes <= 0) return hipSuccess;
  const int slots = 2 * device_cu_count();
  hipLaunchKernelGGL(k_absorb_d3c16, dim3(g.n_nodes < slots ? g.n_nodes : slots), dim3(256), 0, s, g,
                     (const float2*)T, X, (float2*)gam, inst_active);
  return hipGetLastError();
}

bool bp_fast_applicable(int rb, const DevGraph& g, int min_deg, int max_deg) {
  return rb == 4 && g.chi == 16 && g.phys == 2 && min_deg == 3 && max_deg == 3;
}

// The same node update with diagonal messages diag(lambda_e^2) on every bond
// (the lambda^2-weighted bond Grams of the Vidal distance): a mode product
// with a diagonal matrix scales the tensor along that axis, so each output is
// one Gram, out_idx[k][l] = sum conj(T[k@idx, c]) w(c) T[l@idx, c] with
// w(c) the product of the two other bonds' lambda^2 at column c, and the
// matrix cores do only the three Grams.  Weights are rounded to float as the
// diagonal messages of the general path are.  Persistent blocks as
// k_bp_nodes_d3c16.
__global__ void __launch_bounds__(512)
    k_bp_nodes_lam2_d3c16(DevGraph g, const int* nodes, int count, const float2* T, const double* lam,
                          float2* n2e_new, const int* inst_active) {
  using namespace fast3;
  constexpr int NW = 8, NT = 512, PER = S / NT;
  __shared__ float2 tl[NPAD];
  __shared__ float2 part[3 * NW * 256];
  __shared__ float wl[3][16];
  __shared__ float2 diag[3 * 16];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  auto active = [&](int it) { return !inst_active || ldc(inst_active, ldc(g.node_inst, ldc(nodes, it))); };
  int it = blockIdx.x;
  while (it < count && !active(it)) it += gridDim.x;
  if (it >= count) return;
  float2 pre[PER];
  {
    const float2* tg = T + ldc(g.toff, ldc(nodes, it));
#pragma unroll
    for (int k = 0; k < PER; ++k) pre[k] = tg[tid + k * NT];
  }
  while (true) {
    const int node = ldc(nodes, it);
    const int base = ldc(g.adj_ptr, node);
    // tl and wl were last read before the barrier that precedes the partial
    // writes of the previous node
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int f = tid + k * NT;
      const int r = f & 31, b1 = (f >> 5) & 15, b0 = f >> 9;
      tl[b0 * B0P + b1 * RP + r] = pre[k];
    }
    int oe[3], os[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      oe[k] = ldc(g.adj_edge, base + k);
      os[k] = ldc(g.adj_side, base + k);
    }
    if (tid < 48) {
      const int k = tid >> 4;
      const double v = lam[(long long)(k == 0 ? oe[0] : (k == 1 ? oe[1] : oe[2])) * 16 + (tid & 15)];
      wl[k][tid & 15] = (float)(v * v);
    }
    int next = it + gridDim.x;
    while (next < count && !active(next)) next += gridDim.x;
    if (next < count) {
      const float2* tg = T + ldc(g.toff, ldc(nodes, next));
#pragma unroll
      for (int k = 0; k < PER; ++k) pre[k] = tg[tid + k * NT];
    }
    __syncthreads();
    const int row = lane & 15, kq = lane >> 4;
    constexpr int per = 128 / NW;
#pragma unroll
    for (int idx = 0; idx < 3; ++idx) {
      // columns c enumerate the other two bonds (lower axis first) and s
      const int jlo = idx == 0 ? 1 : 0, jhi = idx == 2 ? 1 : 2;
      const int st = axis_stride(idx);
      floatx4 gr = {0.f, 0.f, 0.f, 0.f}, gi = {0.f, 0.f, 0.f, 0.f};
      for (int ks = wave * per; ks < wave * per + per; ++ks) {
        const int c = 4 * ks + kq;
        const float2 a = tl[col_base(idx, c) + row * st];
        const float w = wl[jlo][c >> 5] * wl[jhi][(c >> 1) & 15];
        const float2 b = make_float2(a.x * w, a.y * w);
        gr = mfma(a.x, b.x, gr);
        gr = mfma(a.y, b.y, gr);
        gi = mfma(a.x, b.y, gi);
        gi = mfma(-a.y, b.x, gi);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
        part[(idx * NW + wave) * 256 + (4 * (lane >> 4) + v) * 16 + (lane & 15)] = make_float2(gr[v], gi[v]);
    }
    __syncthreads();
    constexpr int QN = (768 + NT - 1) / NT;
    float2 sum[QN];
#pragma unroll
    for (int qi = 0; qi < QN; ++qi) {
      const int q = tid + qi * NT;
      sum[qi] = make_float2(0.f, 0.f);
      if (q < 768) {
        const int k = q >> 8, e = q & 255;
        for (int w = 0; w < NW; ++w) {
          sum[qi].x += part[(k * NW + w) * 256 + e].x;
          sum[qi].y += part[(k * NW + w) * 256 + e].y;
        }
        if (e % 17 == 0) diag[k * 16 + e / 17] = sum[qi];
      }
    }
    __syncthreads();
#pragma unroll
    for (int qi = 0; qi < QN; ++qi) {
      const int q = tid + qi * NT;
      if (q < 768) {
        const int k = q >> 8, e = q & 255;
        float tr_r = 0.f, tr_i = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          tr_r += diag[k * 16 + c].x;
          tr_i += diag[k * 16 + c].y;
        }
        const float den = tr_r * tr_r + tr_i * tr_i;
        float2 o;
        o.x = (sum[qi].x * tr_r + sum[qi].y * tr_i) / den;
        o.y = (sum[qi].y * tr_r - sum[qi].x * tr_i) / den;
        n2e_new[(2LL * oe[k] + os[k]) * 256 + e] = o;
      }
    }
    // part and diag are next written after the next node's staging barrier
    if (next >= count) break;
    it = next;
  }
}

hipError_t launch_bp_nodes_lam2_fast(const DevGraph& g, const int* nodes, int count, const void* T,
                                     const double* lam, void* n2e_new, const int* inst_active, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int cus = device_cu_count();
  const int grid = count < cus ? count : cus;
  hipLaunchKernelGGL(k_bp_nodes_lam2_d3c16, dim3(grid), dim3(512), 0, s, g, nodes, count, (const float2*)T, lam,
                     (float2*)n2e_new, inst_active);
  return hipGetLastError();
}

hipError_t launch_bp_nodes_fast(const DevGraph& g, const int* nodes, int count, const void* T, const void* n2e_old,
                                const void* e2n_old, void* n2e_new, const uint8_t* node_src, const int* inst_active,
                                hipStream_t s, void* e2n_copy) {
  if (count <= 0) return hipSuccess;
  // variant (autotuner): bit 0 = one block per node instead of persistent
  // blocks, bit 1 = 4 waves per block instead of 8
  int v = tune_get(kTuneBpD3C16, 0);
  if (v < 0) v = 0;
  const int cus = device_cu_count();
  const int grid = (v & 1) ? count : (count < cus ? count : cus);
  if (v & 2)
    hipLaunchKernelGGL(k_bp_nodes_d3c16<4>, dim3(grid), dim3(256), 0, s, g, nodes, count, (const float2*)T,
                       (const float2*)n2e_old, (const float2*)e2n_old, (float2*)n2e_new, node_src, inst_active,
                       (float2*)e2n_copy);
  else
    hipLaunchKernelGGL(k_bp_nodes_d3c16<8>, dim3(grid), dim3(512), 0, s, g, nodes, count, (const float2*)T,
                       (const float2*)n2e_old, (const float2*)e2n_old, (float2*)n2e_new, node_src, inst_active,
                       (float2*)e2n_copy);
  return hipGetLastError();
}

}  // namespace dq

// Exact state-vector kernels (complex128 amplitudes, 2^n of them).
//
// Bit b of an amplitude's index is one qubit; the host maps the reference's
// qubit positions (row-major reshape of the vector to n axes of 2,
// dequantizer/src/quantum_emulator/exact_quantum_emulator.py:67-76) to bits
// as b = n - 1 - pos.  Each kernel below but the small table builders
// streams the vector through HBM once:
//   k_sv_basis      write a computational basis state;
//   k_sv_apply1     one 2x2 gate on one bit (pairs of amplitudes);
//   k_sv_apply2     one 4x4 gate on two bits (quads of amplitudes);
//   k_sv_diag       up to 256 diagonal two-bit gates in one pass: the
//                   phase of an amplitude factors into a chunk constant, two
//                   64-entry tables over the chunk's low bits and a table
//                   over the low 12 bits shared by all chunks;
//   k_sv_diag_tile  up to 64 diagonal terms fused into the low tile pass of
//                   a single-qubit layer (a coupling layer plus the low 12
//                   bits of a mixing layer in one pass);
//   k_sv_tile_*     a tile of 2^12 amplitudes spanning up to 12 target
//                   bits (the 12 lowest, or 8 higher ones with 16-amplitude
//                   runs of the lowest 4 for coalescing): applies one 2x2
//                   gate per target bit, or accumulates each target bit's
//                   one-qubit density matrix into per-block partial sums;
//                   16 amplitudes per thread, four target bits at a time in
//                   registers, LDS transposes between them (the *_small
//                   variants: vectors of fewer than 2^12 amplitudes).
#include <algorithm>

#include "dq_launch.hpp"
#include "dq_internal.hpp"

namespace dq {

namespace {

constexpr int kTileBits = 12;
constexpr int kTile = 1 << kTileBits;
constexpr int kThreads = 256;
// threads of the diagonal pass (4 amplitudes each per chunk of 2^12) and
// terms per launch of it (longer lists take several passes)
constexpr int kDiagThreads = 1024;
constexpr int kDiagTerms = 256;
// diagonal terms of the pass fused with a single-qubit layer
constexpr int kFusedTerms = 64;
// blocks of the density-matrix pass (one partial sum each)
constexpr int kSvRdmBlocks = 2048;

__device__ __forceinline__ cd ldcd(const cd* p, long long i) {
  const double* q = (const double*)p;
  return mk<double>(ldc(q, 2 * i), ldc(q, 2 * i + 1));
}

// insert a zero bit at position b into x
__device__ __forceinline__ unsigned long long ins0(unsigned long long x, int b) {
  const unsigned long long lo = x & ((1ULL << b) - 1ULL);
  return ((x >> b) << (b + 1)) | lo;
}

__global__ void __launch_bounds__(kThreads) k_sv_basis(cd* psi, unsigned long long total, unsigned long long idx) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)kThreads + threadIdx.x; i < total;
       i += (unsigned long long)gridDim.x * kThreads)
    psi[i] = mk<double>(i == idx ? 1.0 : 0.0, 0.0);
}

// out[o] = sum_i g[o][i] a[i]; g row-major 2x2 in device memory
__global__ void __launch_bounds__(kThreads) k_sv_apply1(cd* psi, unsigned long long pairs, int b, const cd* g) {
  const cd g00 = ldcd(g, 0), g01 = ldcd(g, 1), g10 = ldcd(g, 2), g11 = ldcd(g, 3);
  const unsigned long long bit = 1ULL << b;
  for (unsigned long long p = blockIdx.x * (unsigned long long)kThreads + threadIdx.x; p < pairs;
       p += (unsigned long long)gridDim.x * kThreads) {
    const unsigned long long i0 = ins0(p, b), i1 = i0 | bit;
    const cd a0 = psi[i0], a1 = psi[i1];
    psi[i0] = g00 * a0 + g01 * a1;
    psi[i1] = g10 * a0 + g11 * a1;
  }
}

// out[o1 o2] = sum g[o1][o2][i1][i2] a[i1 i2] with x1 on bit b1, x2 on bit b2
__global__ void __launch_bounds__(kThreads) k_sv_apply2(cd* psi, unsigned long long quads, int b1, int b2,
                                                        const cd* g) {
  cd gm[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) gm[k] = ldcd(g, k);
  const int lo = b1 < b2 ? b1 : b2, hi = b1 < b2 ? b2 : b1;
  const unsigned long long m1 = 1ULL << b1, m2 = 1ULL << b2;
  for (unsigned long long q = blockIdx.x * (unsigned long long)kThreads + threadIdx.x; q < quads;
       q += (unsigned long long)gridDim.x * kThreads) {
    const unsigned long long base = ins0(ins0(q, lo), hi);
    cd a[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) a[x] = psi[base | ((x >> 1) ? m1 : 0ULL) | ((x & 1) ? m2 : 0ULL)];
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      cd acc = mk<double>(0.0, 0.0);
#pragma unroll
      for (int x = 0; x < 4; ++x) fma_acc(acc, gm[o * 4 + x], a[x]);
      psi[base | ((o >> 1) ? m1 : 0ULL) | ((o & 1) ? m2 : 0ULL)] = acc;
    }
  }
}

// ------------------------------------------------------------ diagonal layer
// Term k: bits (tb[2k], tb[2k+1]), diagonal td[4k + 2 x1 + x2] where x1, x2
// are the values of the two bits.  Chunks of 2^lb amplitudes (lb = min(n,
// 12)); bits below lb are "low".

// dlow[l] = product over terms with both bits low of their entry at l
__global__ void __launch_bounds__(kThreads) k_sv_diag_low(cd* dlow, int lb, int K, const int* tb, const cd* td) {
  for (int l = threadIdx.x; l < (1 << lb); l += kThreads) {
    cd f = mk<double>(1.0, 0.0);
    for (int k = 0; k < K; ++k) {
      const int b1 = ldc(tb, 2 * k), b2 = ldc(tb, 2 * k + 1);
      if (b1 < lb && b2 < lb) f = f * ldcd(td, 4LL * k + 2 * ((l >> b1) & 1) + ((l >> b2) & 1));
    }
    dlow[l] = f;
  }
}

// LDS state of the diagonal terms in a block: the terms (at most KMAX), and
// per low bit q < lb the list of terms with low bit q and one high bit
// (entry: k | high bit << 8 | (low bit first) << 16); list lb: the terms
// with both bits high (k | b1 << 8 | b2 << 16).  fq[q][x]: product of the
// entries of list q at low bit value x; c: product over list lb; ta, tbh: the
// chunk's phase factors over its low la bits and its next lh bits.
template <int KMAX>
struct DiagLds {
  cd sd[KMAX][4];
  unsigned int ent[KMAX];
  int beg[kTileBits + 2];
  cd fq[kTileBits][2];
  cd c;
  cd ta[64], tbh[64];
};

// Stage the K terms and build the lists (thread q <= lb counts its terms,
// thread 0 turns the counts into offsets, thread q writes its list in term
// order).  Ends with a barrier.
template <int KMAX, int NT>
__device__ void diag_setup(DiagLds<KMAX>& d, int lb, int K, const int* tb, const cd* td) {
  const int t = threadIdx.x;
  for (int i = t; i < 4 * K; i += NT) d.sd[i >> 2][i & 3] = td[i];
  int mine = 0;
  if (t <= lb) {
    for (int k = 0; k < K; ++k) {
      const int b1 = ldc(tb, 2 * k), b2 = ldc(tb, 2 * k + 1);
      const bool l1 = b1 < lb, l2 = b2 < lb;
      mine += (t == lb) ? (!l1 && !l2) : (l1 != l2 && (l1 ? b1 : b2) == t);
    }
    d.beg[t + 1] = mine;
  }
  __syncthreads();
  if (t == 0) {
    d.beg[0] = 0;
    for (int q = 0; q <= lb; ++q) d.beg[q + 1] += d.beg[q];
  }
  __syncthreads();
  if (t <= lb) {
    int m = d.beg[t];
    for (int k = 0; k < K; ++k) {
      const int b1 = ldc(tb, 2 * k), b2 = ldc(tb, 2 * k + 1);
      const bool l1 = b1 < lb, l2 = b2 < lb;
      if (t == lb) {
        if (!l1 && !l2) d.ent[m++] = (unsigned)k | ((unsigned)b1 << 8) | ((unsigned)b2 << 16);
      } else if (l1 != l2 && (l1 ? b1 : b2) == t) {
        d.ent[m++] = (unsigned)k | ((unsigned)(l1 ? b2 : b1) << 8) | ((unsigned)l1 << 16);
      }
    }
  }
  __syncthreads();
}

// The phase tables ta, tbh of the chunk whose high bits are `high`.  Needs
// at least 128 threads; ends with a barrier.
template <int KMAX>
__device__ void diag_tables(DiagLds<KMAX>& d, int lb, unsigned long long high) {
  const int t = threadIdx.x;
  const int la = lb / 2, lh = lb - la;
  if (t <= lb) {
    cd f0 = mk<double>(1.0, 0.0), f1 = f0;
    const int e1 = d.beg[t + 1];
    for (int m = d.beg[t]; m < e1; ++m) {
      const unsigned e = d.ent[m];
      const int k = (int)(e & 255u);
      if (t == lb) {
        const int b1 = (int)((e >> 8) & 255u), b2 = (int)(e >> 16);
        f0 = f0 * d.sd[k][2 * (int)((high >> b1) & 1ULL) + (int)((high >> b2) & 1ULL)];
      } else {
        const int xh = (int)((high >> ((e >> 8) & 255u)) & 1ULL);
        if ((e >> 16) & 1u) {  // low bit first: entry 2 x_low + x_high
          f0 = f0 * d.sd[k][xh];
          f1 = f1 * d.sd[k][2 + xh];
        } else {  // entry 2 x_high + x_low
          f0 = f0 * d.sd[k][2 * xh];
          f1 = f1 * d.sd[k][2 * xh + 1];
        }
      }
    }
    if (t == lb) d.c = f0;
    else {
      d.fq[t][0] = f0;
      d.fq[t][1] = f1;
    }
  }
  __syncthreads();
  if (t < (1 << la)) {
    cd f = d.c;
    for (int q = 0; q < la; ++q) f = f * d.fq[q][(t >> q) & 1];
    d.ta[t] = f;
  } else if (t >= 64 && t - 64 < (1 << lh)) {
    const int y = t - 64;
    cd f = mk<double>(1.0, 0.0);
    for (int q = 0; q < lh; ++q) f = f * d.fq[la + q][(y >> q) & 1];
    d.tbh[y] = f;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kDiagThreads) k_sv_diag(cd* psi, unsigned long long chunks, int lb, int K,
                                                             const int* tb, const cd* td, const cd* dlow) {
  __shared__ DiagLds<kDiagTerms> d;
  const int la = lb / 2;
  const int t = threadIdx.x;
  diag_setup<kDiagTerms, kDiagThreads>(d, lb, K, tb, td);
  constexpr int kPer = kTile / kDiagThreads;
  for (unsigned long long ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const unsigned long long high = ch << lb;
    // the chunk's amplitudes are requested first, so that their loads are in
    // flight while the tables are built
    cd v[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int l = t + r * kDiagThreads;
      if (l < (1 << lb)) v[r] = psi[high | (unsigned long long)l];
    }
    diag_tables(d, lb, high);
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int l = t + r * kDiagThreads;
      if (l < (1 << lb)) {
        const cd ph = dlow[l] * (d.ta[l & ((1 << la) - 1)] * d.tbh[l >> la]);
        psi[high | (unsigned long long)l] = ph * v[r];
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------- tile passes
// A tile: 2^lo amplitudes contiguous in the lowest lo bits times 2^h
// combinations of the high bits hb[0..h-1] (ascending, all >= lo), with
// lo + h = 12 (or n when n < 12).  Tile index t = (combination << lo) | low.
// Target local bits: tmask bit j (j < lo: index bit j; j >= lo: hb[j - lo]).
struct TileSpec {
  int n, lo, h;
  int hb[8];
  int tmask;
};

__device__ __forceinline__ unsigned long long tile_base(const TileSpec& s, unsigned long long blk) {
  // the block id enumerates the bits that are neither low nor in hb
  unsigned long long x = blk << s.lo;
  for (int j = 0; j < s.h; ++j) x = ins0(x, s.hb[j]);
  return x;
}
__device__ __forceinline__ unsigned long long tile_offset(const TileSpec& s, int t) {
  unsigned long long off = (unsigned long long)(t & ((1 << s.lo) - 1));
  const int comb = t >> s.lo;
  for (int j = 0; j < s.h; ++j)
    if ((comb >> j) & 1) off |= 1ULL << s.hb[j];
  return off;
}
__device__ __forceinline__ int global_bit(const TileSpec& s, int j) { return j < s.lo ? j : s.hb[j - s.lo]; }

// gates[b]: 2x2 gate of bit b (device, cd[4] per bit)
__global__ void __launch_bounds__(kThreads) k_sv_tile_apply_small(cd* psi, unsigned long long blocks, TileSpec s,
                                                            const cd* gates) {
  __shared__ cd tile[kTile];
  const int tsz = 1 << (s.lo + s.h), nb = s.lo + s.h;
  for (unsigned long long blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    const unsigned long long base = tile_base(s, blk);
    for (int t = threadIdx.x; t < tsz; t += kThreads) tile[t] = psi[base | tile_offset(s, t)];
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
      if (!((s.tmask >> j) & 1)) continue;
      const long long gb = 4LL * global_bit(s, j);
      const cd g00 = ldcd(gates, gb), g01 = ldcd(gates, gb + 1), g10 = ldcd(gates, gb + 2),
               g11 = ldcd(gates, gb + 3);
      for (int p = threadIdx.x; p < tsz / 2; p += kThreads) {
        const int t0 = (int)ins0((unsigned long long)p, j), t1 = t0 | (1 << j);
        const cd a0 = tile[t0], a1 = tile[t1];
        tile[t0] = g00 * a0 + g01 * a1;
        tile[t1] = g10 * a0 + g11 * a1;
      }
      __syncthreads();
    }
    for (int t = threadIdx.x; t < tsz; t += kThreads) psi[base | tile_offset(s, t)] = tile[t];
    __syncthreads();
  }
}

// part[(blockIdx.x * 12 + j) * 4 + {0: |a0|^2, 1: |a1|^2, 2: re, 3: im of
// a0 conj(a1)}] for every local bit j < lo + h (all twelve written).
__global__ void __launch_bounds__(kThreads) k_sv_tile_rdm_small(const cd* psi, unsigned long long blocks, TileSpec s,
                                                          double* part) {
  __shared__ cd tile[kTile];
  __shared__ double red[kThreads / 64][kTileBits * 4];
  const int tsz = 1 << (s.lo + s.h), nb = s.lo + s.h;
  double acc[kTileBits][4];
#pragma unroll
  for (int j = 0; j < kTileBits; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[j][v] = 0.0;
  for (unsigned long long blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    const unsigned long long base = tile_base(s, blk);
    for (int t = threadIdx.x; t < tsz; t += kThreads) tile[t] = psi[base | tile_offset(s, t)];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTileBits; ++j) {
      if (j < nb && ((s.tmask >> j) & 1)) {
        for (int p = threadIdx.x; p < tsz / 2; p += kThreads) {
          const int t0 = (int)ins0((unsigned long long)p, j), t1 = t0 | (1 << j);
          const cd a0 = tile[t0], a1 = tile[t1];
          acc[j][0] += norm2(a0);
          acc[j][1] += norm2(a1);
          acc[j][2] += a0.re * a1.re + a0.im * a1.im;
          acc[j][3] += a0.im * a1.re - a0.re * a1.im;
        }
      }
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kTileBits; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double x = wave_sum(acc[j][v]);
      if (lane == 0) red[w][j * 4 + v] = x;
    }
  __syncthreads();
  if (threadIdx.x < kTileBits * 4) {
    double x = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) x += red[i][threadIdx.x];
    part[(long long)blockIdx.x * kTileBits * 4 + threadIdx.x] = x;
  }
}

// Full tiles (lo + h = 12) with register staging.  Each thread holds 16
// amplitudes of the tile; three layouts put four local bits into the
// register index r (element r of thread tid):
//   L2: t = (r << 8) | tid                              (bits 8..11 in r)
//   L1: t = ((tid >> 4) << 8) | (r << 4) | (tid & 15)   (bits 4..7 in r)
//   L0: t = (tid << 4) | r                              (bits 0..3 in r)
// Global loads and stores use L2 (lanes along the tile's low bits, which are
// the contiguous runs of the vector); a stage's gates act on registers only,
// and the tile goes through LDS (padded t + t/16 against bank conflicts)
// between stages.  Stages without target bits are skipped.
constexpr int kTilePad = kTile + kTile / 16;
__device__ __forceinline__ int padt(int t) { return t + (t >> 4); }

template <int S>
__device__ __forceinline__ int lay(int tid, int r) {
  if (S == 2) return (r << 8) | tid;
  if (S == 1) return ((tid >> 4) << 8) | (r << 4) | (tid & 15);
  return (tid << 4) | r;
}

template <int S>
__device__ __forceinline__ void gate_stage(cd (&a)[16], const TileSpec& s, const cd* gates) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = 4 * S + q;
    if (!((s.tmask >> j) & 1)) continue;
    const long long gb = 4LL * global_bit(s, j);
    const cd g00 = ldcd(gates, gb), g01 = ldcd(gates, gb + 1), g10 = ldcd(gates, gb + 2), g11 = ldcd(gates, gb + 3);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if ((r >> q) & 1) continue;
      const cd a0 = a[r], a1 = a[r | (1 << q)];
      a[r] = g00 * a0 + g01 * a1;
      a[r | (1 << q)] = g10 * a0 + g11 * a1;
    }
  }
}

template <int S>
__device__ __forceinline__ void rdm_stage(const cd (&a)[16], const TileSpec& s, double (&acc)[kTileBits][4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = 4 * S + q;
    if (!((s.tmask >> j) & 1)) continue;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if ((r >> q) & 1) continue;
      const cd a0 = a[r], a1 = a[r | (1 << q)];
      acc[j][0] += norm2(a0);
      acc[j][1] += norm2(a1);
      acc[j][2] += a0.re * a1.re + a0.im * a1.im;
      acc[j][3] += a0.im * a1.re - a0.re * a1.im;
    }
  }
}

template <int S>
__device__ __forceinline__ void lds_read(cd (&a)[16], const cd* tile, int tid) {
#pragma unroll
  for (int r = 0; r < 16; ++r) a[r] = tile[padt(lay<S>(tid, r))];
}
template <int S>
__device__ __forceinline__ void lds_write(cd* tile, const cd (&a)[16], int tid) {
#pragma unroll
  for (int r = 0; r < 16; ++r) tile[padt(lay<S>(tid, r))] = a[r];
}

// per-thread and per-element parts of the global offsets of layout L2
__device__ __forceinline__ void l2_offsets(const TileSpec& s, int tid, unsigned long long& ot,
                                           unsigned long long (&orr)[16]) {
  ot = tile_offset(s, tid);
#pragma unroll
  for (int r = 0; r < 16; ++r) orr[r] = tile_offset(s, r << 8);
}

__global__ void __launch_bounds__(kThreads, 2) k_sv_tile_apply(cd* psi, unsigned long long blocks, TileSpec s,
                                                               const cd* gates) {
  __shared__ cd tile[kTilePad];
  const int tid = threadIdx.x;
  unsigned long long ot, orr[16];
  l2_offsets(s, tid, ot, orr);
  const bool need1 = (s.tmask >> 4) & 15, need0 = s.tmask & 15;
  cd cur[16];
  for (unsigned long long blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    const unsigned long long base = tile_base(s, blk) | ot;
#pragma unroll
    for (int r = 0; r < 16; ++r) cur[r] = psi[base | orr[r]];
    gate_stage<2>(cur, s, gates);
    if (need1 || need0) {
      lds_write<2>(tile, cur, tid);
      __syncthreads();
      if (need1) {
        lds_read<1>(cur, tile, tid);
        gate_stage<1>(cur, s, gates);
        lds_write<1>(tile, cur, tid);
        __syncthreads();
      }
      if (need0) {
        lds_read<0>(cur, tile, tid);
        gate_stage<0>(cur, s, gates);
        lds_write<0>(tile, cur, tid);
        __syncthreads();
      }
      lds_read<2>(cur, tile, tid);
      // every thread has read the tile before the next one is written
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) psi[base | orr[r]] = cur[r];
  }
}

// The low tile pass (lo = 12, h = 0) of a single-qubit layer with up to
// kFusedTerms diagonal two-bit terms applied first: a coupling layer and the
// low 12 bits of the following mixing layer in one pass over the vector.
// A tile is a chunk of the diagonal pass (lb = 12); element r of thread tid
// in layout L2 is l = (r << 8) | tid, so its phase factors are
// dlow[l] * ta[tid & 63] * tbh[(r << 2) | (tid >> 6)].
__global__ void __launch_bounds__(kThreads, 2) k_sv_diag_tile(cd* psi, unsigned long long blocks, TileSpec s,
                                                              const cd* gates, int K, const int* tb, const cd* td,
                                                              const cd* dlow) {
  __shared__ cd tile[kTilePad];
  __shared__ DiagLds<kFusedTerms> d;
  const int tid = threadIdx.x;
  diag_setup<kFusedTerms, kThreads>(d, kTileBits, K, tb, td);
  const bool need1 = (s.tmask >> 4) & 15, need0 = s.tmask & 15;
  cd cur[16];
  for (unsigned long long blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    const unsigned long long base = blk << kTileBits;
#pragma unroll
    for (int r = 0; r < 16; ++r) cur[r] = psi[base | (unsigned long long)((r << 8) | tid)];
    diag_tables(d, kTileBits, base);
    const cd fa = d.ta[tid & 63];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int l = (r << 8) | tid;
      cur[r] = (dlow[l] * (fa * d.tbh[(r << 2) | (tid >> 6)])) * cur[r];
    }
    gate_stage<2>(cur, s, gates);
    if (need1 || need0) {
      lds_write<2>(tile, cur, tid);
      __syncthreads();
      if (need1) {
        lds_read<1>(cur, tile, tid);
        gate_stage<1>(cur, s, gates);
        lds_write<1>(tile, cur, tid);
        __syncthreads();
      }
      if (need0) {
        lds_read<0>(cur, tile, tid);
        gate_stage<0>(cur, s, gates);
        lds_write<0>(tile, cur, tid);
        __syncthreads();
      }
      lds_read<2>(cur, tile, tid);
    }
    // every thread is done with the tile and the phase tables before the
    // next ones are written
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) psi[base | (unsigned long long)((r << 8) | tid)] = cur[r];
  }
}

__global__ void __launch_bounds__(kThreads) k_sv_tile_rdm(const cd* psi, unsigned long long blocks, TileSpec s,
                                                          double* part) {
  __shared__ cd tile[kTilePad];
  __shared__ double red[kThreads / 64][kTileBits * 4];
  const int tid = threadIdx.x;
  unsigned long long ot, orr[16];
  l2_offsets(s, tid, ot, orr);
  const bool need1 = (s.tmask >> 4) & 15, need0 = s.tmask & 15;
  double acc[kTileBits][4];
#pragma unroll
  for (int j = 0; j < kTileBits; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[j][v] = 0.0;
  cd a[16];
  for (unsigned long long blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    const unsigned long long base = tile_base(s, blk) | ot;
#pragma unroll
    for (int r = 0; r < 16; ++r) a[r] = psi[base | orr[r]];
    rdm_stage<2>(a, s, acc);
    if (need1 || need0) {
      lds_write<2>(tile, a, tid);
      __syncthreads();
      if (need1) {
        lds_read<1>(a, tile, tid);
        rdm_stage<1>(a, s, acc);
      }
      if (need0) {
        lds_read<0>(a, tile, tid);
        rdm_stage<0>(a, s, acc);
      }
      __syncthreads();
    }
  }
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int j = 0; j < kTileBits; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double x = wave_sum(acc[j][v]);
      if (lane == 0) red[w][j * 4 + v] = x;
    }
  __syncthreads();
  if (tid < kTileBits * 4) {
    double x = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) x += red[i][tid];
    part[(long long)blockIdx.x * kTileBits * 4 + tid] = x;
  }
}

// Targets of one tile pass: local bit j[r] of the tile, written to rho of
// global bit gb[r].
struct RdmTargets {
  int count;
  int j[kTileBits];
  int gb[kTileBits];
};

// rho[gb] (2x2 complex, row-major) for target r = blockIdx.x: the partials of
// its local bit summed over nblk blocks in a fixed order.
__global__ void __launch_bounds__(kThreads) k_sv_rdm_reduce(const double* part, int nblk, RdmTargets tg, cd* rho) {
  __shared__ double red[kThreads / 64][4];
  const int r = blockIdx.x;
  if (r >= tg.count) return;
  const int j = tg.j[r];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += kThreads)
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] += part[((long long)b * kTileBits + j) * 4 + q];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double x = wave_sum(v[q]);
    if (lane == 0) red[w][q] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < kThreads / 64; ++i)
      for (int q = 0; q < 4; ++q) s[q] += red[i][q];
    cd* o = rho + 4LL * tg.gb[r];
    o[0] = mk<double>(s[0], 0.0);
    o[1] = mk<double>(s[2], s[3]);
    o[2] = mk<double>(s[2], -s[3]);
    o[3] = mk<double>(s[1], 0.0);
  }
}

int grid_for(unsigned long long work, int cus) {
  const unsigned long long cap = (unsigned long long)cus * 16ULL;
  const unsigned long long need = (work + kThreads - 1) / kThreads;
  return (int)(need < cap ? (need > 0 ? need : 1) : cap);
}

}  // namespace

int sv_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    cus = n;
  }
  return cus;
}

hipError_t launch_sv_basis(void* psi, int n, unsigned long long idx, hipStream_t st) {
  const unsigned long long total = 1ULL << n;
  hipLaunchKernelGGL(k_sv_basis, dim3(grid_for(total, sv_cus())), dim3(kThreads), 0, st, (cd*)psi, total, idx);
  return hipGetLastError();
}

hipError_t launch_sv_apply1(void* psi, int n, int b, const void* g, hipStream_t st) {
  const unsigned long long pairs = 1ULL << (n - 1);
  hipLaunchKernelGGL(k_sv_apply1, dim3(grid_for(pairs, sv_cus())), dim3(kThreads), 0, st, (cd*)psi, pairs, b,
                     (const cd*)g);
  return hipGetLastError();
}

hipError_t launch_sv_apply2(void* psi, int n, int b1, int b2, const void* g, hipStream_t st) {
  const unsigned long long quads = 1ULL << (n - 2);
  hipLaunchKernelGGL(k_sv_apply2, dim3(grid_for(quads, sv_cus())), dim3(kThreads), 0, st, (cd*)psi, quads, b1, b2,
                     (const cd*)g);
  return hipGetLastError();
}

hipError_t launch_sv_diag(void* psi, int n, int K, const int* tb, const void* td, void* work, hipStream_t st) {
  const int lb = n < kTileBits ? n : kTileBits;
  cd* dlow = (cd*)work;
  const unsigned long long chunks = 1ULL << (n - lb);
  // one 1024-thread block per CU, persistent over the chunks
  const unsigned long long cap = (unsigned long long)sv_cus();
  const int grid = (int)(chunks < cap ? chunks : cap);
  for (int k0 = 0; k0 < K; k0 += kDiagTerms) {
    const int kb = K - k0 < kDiagTerms ? K - k0 : kDiagTerms;
    const int* tbk = tb + 2LL * k0;
    const cd* tdk = (const cd*)td + 4LL * k0;
    hipLaunchKernelGGL(k_sv_diag_low, dim3(1), dim3(kThreads), 0, st, dlow, lb, kb, tbk, tdk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sv_diag, dim3(grid), dim3(kDiagThreads), 0, st, (cd*)psi, chunks, lb, kb, tbk, tdk,
                       (const cd*)dlow);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Passes over the target bits (bitmask over [0, n)): bits below 12 in one
// pass with lo = min(n, 12); higher bits in groups of 8 with lo = 4.
namespace {
template <typename F>
hipError_t for_each_pass(int n, unsigned long long targets, F&& f) {
  const int lo0 = n < kTileBits ? n : kTileBits;
  const unsigned long long lowmask = (lo0 >= 64) ? ~0ULL : ((1ULL << lo0) - 1ULL);
  if (targets & lowmask) {
    TileSpec s{};
    s.n = n;
    s.lo = lo0;
    s.h = 0;
    s.tmask = (int)(targets & lowmask);
    hipError_t e = f(s);
    if (e != hipSuccess) return e;
  }
  int hbits[64], nh = 0;
  for (int b = lo0; b < n; ++b)
    if ((targets >> b) & 1ULL) hbits[nh++] = b;
  for (int g0 = 0; g0 < nh; g0 += 8) {
    TileSpec s{};
    s.n = n;
    s.lo = 4;
    s.h = 0;
    // up to 8 target bits; fill to 8 with non-target bits >= 4 so that the
    // tile always holds 2^12 amplitudes
    int chosen[64], nc = 0;
    for (int k = g0; k < nh && k < g0 + 8; ++k) chosen[nc++] = hbits[k];
    const int ntarget = nc;
    for (int b = 4; b < n && nc < 8; ++b) {
      bool used = false;
      for (int k = 0; k < nc; ++k) used |= chosen[k] == b;
      if (!used) chosen[nc++] = b;
    }
    // ascending order, remembering which are targets
    int tm = 0;
    std::sort(chosen, chosen + nc);
    for (int k = 0; k < nc; ++k) {
      s.hb[k] = chosen[k];
      for (int q = g0; q < g0 + ntarget; ++q)
        if (hbits[q] == chosen[k]) tm |= 1 << (4 + k);
    }
    s.h = nc;
    s.tmask = tm;
    hipError_t e = f(s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
}  // namespace

hipError_t launch_sv_layer1(void* psi, int n, unsigned long long targets, const void* gates, hipStream_t st) {
  return for_each_pass(n, targets, [&](const TileSpec& s) {
    const unsigned long long blocks = 1ULL << (n - s.lo - s.h);
    if (s.lo + s.h == kTileBits) {
      // two 68 KiB tiles per CU, persistent
      const unsigned long long cap = (unsigned long long)sv_cus() * 2ULL;
      const int grid = (int)(blocks < cap ? blocks : cap);
      hipLaunchKernelGGL(k_sv_tile_apply, dim3(grid), dim3(kThreads), 0, st, (cd*)psi, blocks, s,
                         (const cd*)gates);
    } else {
      hipLaunchKernelGGL(k_sv_tile_apply_small, dim3(1), dim3(kThreads), 0, st, (cd*)psi, blocks, s,
                         (const cd*)gates);
    }
    return hipGetLastError();
  });
}

// The diagonal terms, then the single-qubit gates of `targets`: with at
// least 12 qubits and at most kFusedTerms terms the terms ride on the low
// tile pass, otherwise they take a pass of their own.
hipError_t launch_sv_diag_layer1(void* psi, int n, int K, const int* tb, const void* td, unsigned long long targets,
                                 const void* gates, void* work, hipStream_t st) {
  if (K == 0) return targets ? launch_sv_layer1(psi, n, targets, gates, st) : hipSuccess;
  if (n < kTileBits || K > kFusedTerms) {
    hipError_t e = launch_sv_diag(psi, n, K, tb, td, work, st);
    if (e != hipSuccess || !targets) return e;
    return launch_sv_layer1(psi, n, targets, gates, st);
  }
  cd* dlow = (cd*)work;
  hipLaunchKernelGGL(k_sv_diag_low, dim3(1), dim3(kThreads), 0, st, dlow, kTileBits, K, tb, (const cd*)td);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  TileSpec s{};
  s.n = n;
  s.lo = kTileBits;
  s.h = 0;
  s.tmask = (int)(targets & ((1ULL << kTileBits) - 1ULL));
  const unsigned long long blocks = 1ULL << (n - kTileBits);
  const unsigned long long cap = (unsigned long long)sv_cus() * 2ULL;
  const int grid = (int)(blocks < cap ? blocks : cap);
  hipLaunchKernelGGL(k_sv_diag_tile, dim3(grid), dim3(kThreads), 0, st, (cd*)psi, blocks, s, (const cd*)gates, K,
                     tb, (const cd*)td, (const cd*)dlow);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned long long rest = targets & ~((1ULL << kTileBits) - 1ULL);
  return rest ? launch_sv_layer1(psi, n, rest, gates, st) : hipSuccess;
}

long long sv_rdm_part_bytes() { return (long long)kSvRdmBlocks * kTileBits * 4 * sizeof(double); }

// rho[b] (4 complex per bit, indexed by bit) for every bit in targets.
hipError_t launch_sv_rdm1(const void* psi, int n, unsigned long long targets, void* rho, void* work,
                          hipStream_t st) {
  double* part = (double*)work;
  return for_each_pass(n, targets, [&](const TileSpec& s) {
    const unsigned long long blocks = 1ULL << (n - s.lo - s.h);
    int grid = 1;
    if (s.lo + s.h == kTileBits) {
      unsigned long long cap = (unsigned long long)sv_cus() * 2ULL;
      if (cap > (unsigned long long)kSvRdmBlocks) cap = kSvRdmBlocks;
      grid = (int)(blocks < cap ? blocks : cap);
      hipLaunchKernelGGL(k_sv_tile_rdm, dim3(grid), dim3(kThreads), 0, st, (const cd*)psi, blocks, s, part);
    } else {
      hipLaunchKernelGGL(k_sv_tile_rdm_small, dim3(1), dim3(kThreads), 0, st, (const cd*)psi, blocks, s, part);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    RdmTargets tg{};
    for (int j = 0; j < s.lo + s.h; ++j) {
      if (!((s.tmask >> j) & 1)) continue;
      tg.j[tg.count] = j;
      tg.gb[tg.count] = j < s.lo ? j : s.hb[j - s.lo];
      tg.count++;
    }
    hipLaunchKernelGGL(k_sv_rdm_reduce, dim3(tg.count), dim3(kThreads), 0, st, (const double*)part, grid, tg,
                       (cd*)rho);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipSuccess;
  });
}

}  // namespace dq

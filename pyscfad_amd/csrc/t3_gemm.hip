// Grouped, batched, strided MFMA GEMM for gfx950 with three arithmetic modes:
//   PREC_F64    : v_mfma_f64_16x16x4f64, FP64 operands and accumulation
//   PREC_F32    : v_mfma_f32_16x16x4f32, FP32 operands and accumulation
//   PREC_BF16X3 : FP32 operands split on the way into LDS as x = hi + lo with
//                 hi, lo in bf16; three v_mfma_f32_16x16x32_bf16 per K-step
//                 (hi*hi + hi*lo + lo*hi), FP32 accumulation (Henry, Tang and
//                 Heinecke, ARITH 2019, "bf16x3").
//
// Every contraction of the (T) energy and its VJP is expressed as a list of
// problems C[b][m][n] (op)= alpha * sum_seg scale_seg * sum_k A_seg[b][m][k] B_seg[b][k][n]
// whose operands are strided views of the resident tensors (see t3_common.h).
// A block owns one (problem, batch, m-tile, n-tile, k-split) output tile, so
// every output element has exactly one writer: no atomics anywhere, and the
// result does not depend on scheduling.
//
// Tiles are staged through LDS with double buffering (one barrier per K-chunk);
// global loads of chunk c+1 are in flight while chunk c is multiplied.
#include "t3_common.h"

namespace t3 {

enum LoadMode : int { LOAD_ROWCONT = 0, LOAD_KCONT = 1, LOAD_SCALAR = 2 };

template <int PREC> struct PrecTraits;
template <> struct PrecTraits<PREC_F64> { using T = double; using Acc = double; static constexpr int BK = 16; static constexpr int VEC = 2; };
template <> struct PrecTraits<PREC_F32> { using T = float; using Acc = float; static constexpr int BK = 32; static constexpr int VEC = 4; };
template <> struct PrecTraits<PREC_BF16X3> { using T = float; using Acc = float; static constexpr int BK = 32; static constexpr int VEC = 4; };

template <class T, int VEC> struct VecT;
template <> struct VecT<double, 2> { using type = double2; };
template <> struct VecT<float, 4> { using type = float4; };

// Register stage of one operand tile (ROWS x BK) loaded from global memory.
template <class T, int ROWS, int BK, int VEC>
struct Stage {
  static constexpr int NE = ROWS * BK / 256;   // elements per thread
  static constexpr int NV = NE / VEC;          // vectors per thread
  // tiles too small for one vector per thread are always loaded element-wise
  static constexpr bool CANVEC = NE >= VEC && (NE % VEC) == 0;
  static_assert(NE >= 1 && (ROWS * BK) % 256 == 0, "tile must give each thread an element");
  T r[NE];

  __device__ __forceinline__ void load(int mode, const T* __restrict__ base, long long s_row,
                                       long long s_k, int rows_valid, float scale) {
    const int t = threadIdx.x;
    if (!CANVEC) mode = LOAD_SCALAR;
    if (mode == LOAD_ROWCONT) {
      using V = typename VecT<T, VEC>::type;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int id = t + 256 * q;
        const int k = id / (ROWS / VEC), rv = id % (ROWS / VEC);
        V x;
        if (rv * VEC < rows_valid) x = *reinterpret_cast<const V*>(base + k * s_k + rv * VEC);
        else x = V{};
        const T* px = reinterpret_cast<const T*>(&x);
#pragma unroll
        for (int e = 0; e < VEC; e++) r[q * VEC + e] = px[e] * (T)scale;
      }
    } else if (mode == LOAD_KCONT) {
      using V = typename VecT<T, VEC>::type;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int id = t + 256 * q;
        const int kv = id % (BK / VEC), row = id / (BK / VEC);
        V x;
        if (row < rows_valid) x = *reinterpret_cast<const V*>(base + row * s_row + kv * VEC);
        else x = V{};
        const T* px = reinterpret_cast<const T*>(&x);
#pragma unroll
        for (int e = 0; e < VEC; e++) r[q * VEC + e] = px[e] * (T)scale;
      }
    } else {
#pragma unroll
      for (int q = 0; q < NE; q++) {
        const int id = t + 256 * q;
        const int row = id % ROWS, k = id / ROWS;
        r[q] = (row < rows_valid) ? base[row * s_row + k * s_k] * (T)scale : (T)0;
      }
    }
  }

  // LDS layout [k][row] (f64 / f32 MFMA operands)
  template <int LD>
  __device__ __forceinline__ void store_krow(int mode, T* __restrict__ lds) const {
    const int t = threadIdx.x;
    if (!CANVEC) mode = LOAD_SCALAR;
    if (mode == LOAD_ROWCONT) {
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int id = t + 256 * q;
        const int k = id / (ROWS / VEC), rv = id % (ROWS / VEC);
#pragma unroll
        for (int e = 0; e < VEC; e++) lds[k * LD + rv * VEC + e] = r[q * VEC + e];
      }
    } else if (mode == LOAD_KCONT) {
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int id = t + 256 * q;
        const int kv = id % (BK / VEC), row = id / (BK / VEC);
#pragma unroll
        for (int e = 0; e < VEC; e++) lds[(kv * VEC + e) * LD + row] = r[q * VEC + e];
      }
    } else {
#pragma unroll
      for (int q = 0; q < NE; q++) {
        const int id = t + 256 * q;
        const int row = id % ROWS, k = id / ROWS;
        lds[k * LD + row] = r[q];
      }
    }
  }

  // K-contiguous tiles go to LDS as [row][k] (rows of LDK elements); all
  // other tiles as [k][row] (rows of ROWS elements, XOR-swizzled by
  // swz(k) on 16-element groups), so that neither store needs a transpose.
  template <int LDK>
  __device__ __forceinline__ void store_rk(T* __restrict__ lds) const {
    const int t = threadIdx.x;
    using V = typename VecT<T, VEC>::type;
#pragma unroll
    for (int q = 0; q < NV; q++) {
      const int id = t + 256 * q;
      const int kv = id % (BK / VEC), row = id / (BK / VEC);
      *reinterpret_cast<V*>(&lds[row * LDK + kv * VEC]) = *reinterpret_cast<const V*>(&r[q * VEC]);
    }
  }
  __device__ __forceinline__ static int swz(int k) {
    // f32 [k][row] reads touch rows k = g*KG + ks for the four lane groups g;
    // flipping column bit 4 on odd g spreads them over both bank halves
    return (sizeof(T) == 4) ? (((k / (BK / 4)) & 1) << 4) : 0;
  }
  __device__ __forceinline__ void store_kr(int mode, T* __restrict__ lds) const {
    const int t = threadIdx.x;
    if (mode == LOAD_ROWCONT && CANVEC) {
      using V = typename VecT<T, VEC>::type;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int id = t + 256 * q;
        const int k = id / (ROWS / VEC), rv = id % (ROWS / VEC);
        *reinterpret_cast<V*>(&lds[k * ROWS + ((rv * VEC) ^ swz(k))]) = *reinterpret_cast<const V*>(&r[q * VEC]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < NE; q++) {
        const int id = t + 256 * q;
        const int row = id % ROWS, k = id / ROWS;
        lds[k * ROWS + (row ^ swz(k))] = r[q];
      }
    }
  }

  // LDS layout [row][k] as split bf16 pairs (bf16x3 MFMA operands)
  template <int LDK>
  __device__ __forceinline__ void store_rowk_split(int mode, __bf16* __restrict__ hi,
                                                   __bf16* __restrict__ lo) const {
    const int t = threadIdx.x;
    if (!CANVEC) mode = LOAD_SCALAR;
    auto put = [&](int row, int k, float x) {
      const __bf16 h = (__bf16)x;
      const __bf16 l = (__bf16)(x - (float)h);
      hi[row * LDK + k] = h;
      lo[row * LDK + k] = l;
    };
    if (mode == LOAD_ROWCONT) {
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int id = t + 256 * q;
        const int k = id / (ROWS / VEC), rv = id % (ROWS / VEC);
#pragma unroll
        for (int e = 0; e < VEC; e++) put(rv * VEC + e, k, (float)r[q * VEC + e]);
      }
    } else if (mode == LOAD_KCONT) {
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int id = t + 256 * q;
        const int kv = id % (BK / VEC), row = id / (BK / VEC);
#pragma unroll
        for (int e = 0; e < VEC; e++) put(row, kv * VEC + e, (float)r[q * VEC + e]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < NE; q++) {
        const int id = t + 256 * q;
        const int row = id % ROWS, k = id / ROWS;
        put(row, k, (float)r[q]);
      }
    }
  }
};

template <class Acc> struct AccFrag;
template <> struct AccFrag<double> { using type = d4v; };
template <> struct AccFrag<float> { using type = f4v; };

// Locate the block's problem and tile.
struct TileCoord {
  int p, b, tm, tn, split;
};

__device__ __forceinline__ TileCoord find_tile(const GemmProb* __restrict__ probs, int nprob,
                                               long long bid) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {   // last problem with block0 <= bid
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].block0 <= bid) lo = mid; else hi = mid - 1;
  }
  const GemmProb& P = probs[lo];
  long long local = bid - P.block0;
  TileCoord tc;
  tc.p = lo;
  tc.split = (int)(local % P.ksplit); local /= P.ksplit;
  tc.tn = (int)(local % P.tiles_n); local /= P.tiles_n;
  tc.tm = (int)(local % P.tiles_m); local /= P.tiles_m;
  tc.b = (int)local;
  return tc;
}

template <class Acc>
__device__ __forceinline__ void write_out(const GemmProb& P, const TileCoord& tc, int m, int n,
                                          Acc v) {
  if (m >= P.M || n >= P.N) return;
  if (P.mode == GEMM_PARTIAL) {
    Acc* out = reinterpret_cast<Acc*>(P.P);
    out[((long long)(tc.split * P.nb + tc.b) * P.M + m) * P.N + n] = v;
    return;
  }
  const long long off = (long long)tc.b * P.c_sb + (long long)m * P.c_sm + (long long)n * P.c_sn;
  const double x = P.alpha * (double)v;
  if (P.c_is_f64) {
    double* c = reinterpret_cast<double*>(P.C) + off;
    *c = (P.mode == GEMM_ACCUM) ? (*c + x) : x;
  } else {
    float* c = reinterpret_cast<float*>(P.C) + off;
    *c = (P.mode == GEMM_ACCUM) ? (float)((double)*c + x) : (float)x;
  }
}

// BM x BN tile, 4 waves arranged WGM x (4/WGM).
template <int PREC, int BM, int BN, int WGM>
__global__ void __launch_bounds__(256)
gemm_grouped(const GemmProb* __restrict__ probs, int nprob, const GemmSeg* __restrict__ segs) {
  using TR = PrecTraits<PREC>;
  using T = typename TR::T;
  using Acc = typename TR::Acc;
  using Frag = typename AccFrag<Acc>::type;
  constexpr int BK = TR::BK, VEC = TR::VEC;
  constexpr int WGN = 4 / WGM;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int FM = WM / 16, FN = WN / 16;
  static_assert(FM >= 1 && FN >= 1, "wave tile");

  // Blocks are dealt round-robin to the 8 XCDs (each with its own L2); map
  // them so that every XCD works through a contiguous range of logical tiles,
  // which share operand rows.
  long long bid = blockIdx.x;
  {
    const long long nb = gridDim.x, q = nb / 8;
    if (bid < 8 * q) bid = (bid % 8) * q + bid / 8;
  }
  const TileCoord tc = find_tile(probs, nprob, bid);
  const GemmProb& P = probs[tc.p];
  const int m0 = tc.tm * BM, n0 = tc.tn * BN;
  const int rows_a = min(BM, P.M - m0), rows_b = min(BN, P.N - n0);
  const int c_begin = (int)((long long)P.nchunks * tc.split / P.ksplit);
  const int c_end = (int)((long long)P.nchunks * (tc.split + 1) / P.ksplit);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm0 = (wave % WGM) * WM, wn0 = (wave / WGM) * WN;

  Frag acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++) acc[i][j] = Frag{0, 0, 0, 0};

  Stage<T, BM, BK, VEC> sa;
  Stage<T, BN, BK, VEC> sb;

  // segment cursor
  int sidx = P.seg0;
  auto seg_of = [&](int c) -> const GemmSeg& {
    while (sidx + 1 < P.seg0 + P.nseg && segs[sidx + 1].chunk0 <= c) sidx++;
    return segs[sidx];
  };
  auto load_chunk = [&](int c, int& amode, int& bmode) {
    const GemmSeg& S = seg_of(c);
    const long long kk = (long long)(c - S.chunk0) * BK;
    const long long k1 = kk / S.K2, k2 = kk % S.K2;
    const T* A = reinterpret_cast<const T*>(S.A) + tc.b * S.a_sb + m0 * S.a_sm + k1 * S.a_sk1 + k2 * S.a_sk2;
    const T* B = reinterpret_cast<const T*>(S.B) + tc.b * S.b_sb + n0 * S.b_sn + k1 * S.b_sk1 + k2 * S.b_sk2;
    amode = S.a_sm == 1 ? LOAD_ROWCONT : (S.a_sk2 == 1 ? LOAD_KCONT : LOAD_SCALAR);
    bmode = (S.b_sn == 1 && (P.N % VEC) == 0) ? LOAD_ROWCONT : (S.b_sk2 == 1 ? LOAD_KCONT : LOAD_SCALAR);
    if (amode == LOAD_ROWCONT && (P.M % VEC) != 0) amode = (S.a_sk2 == 1 ? LOAD_KCONT : LOAD_SCALAR);
    sa.load(amode, A, S.a_sm, S.a_sk2, rows_a, S.scale);
    sb.load(bmode, B, S.b_sn, S.b_sk2, rows_b, 1.0f);
  };

  if constexpr (PREC != PREC_BF16X3) {
    // K is consumed in a permuted order: in step ks the lanes of group
    // g = lane/16 multiply k = g*KG + ks (the sum over k is order-free), so a
    // lane's KG values of a [row][k] tile are contiguous (two 16-byte reads).
    constexpr int KG = BK / 4;
    constexpr int LDK = BK + 16 / (int)sizeof(T);   // [row][k] row pitch, staggered by 16 B
    constexpr int SZA = (BM * LDK > BK * BM) ? BM * LDK : BK * BM;
    constexpr int SZB = (BN * LDK > BK * BN) ? BN * LDK : BK * BN;
    __shared__ __attribute__((aligned(16))) T lds_a[2][SZA];
    __shared__ __attribute__((aligned(16))) T lds_b[2][SZB];
    using SA = Stage<T, BM, BK, VEC>;
    using SB = Stage<T, BN, BK, VEC>;
    int amode = 0, bmode = 0;
    bool rk_a[2] = {false, false}, rk_b[2] = {false, false};
    auto store = [&](int buf) {
      rk_a[buf] = (amode == LOAD_KCONT) && SA::CANVEC;
      rk_b[buf] = (bmode == LOAD_KCONT) && SB::CANVEC;
      if (rk_a[buf]) sa.template store_rk<LDK>(lds_a[buf]); else sa.store_kr(amode, lds_a[buf]);
      if (rk_b[buf]) sb.template store_rk<LDK>(lds_b[buf]); else sb.store_kr(bmode, lds_b[buf]);
    };
    if (c_begin < c_end) {
      load_chunk(c_begin, amode, bmode);
      store(0);
    }
    __syncthreads();
    int cur = 0;
    const int g = lane >> 4, l16 = lane & 15;
    for (int c = c_begin; c < c_end; c++) {
      const bool more = c + 1 < c_end;
      if (more) load_chunk(c + 1, amode, bmode);
      const T* la = lds_a[cur];
      const T* lb = lds_b[cur];
      T av[FM][KG], bv[FN][KG];
      if (rk_a[cur]) {
#pragma unroll
        for (int i = 0; i < FM; i++) {
          const T* src = la + (wm0 + i * 16 + l16) * LDK + g * KG;
#pragma unroll
          for (int ks = 0; ks < KG; ks++) av[i][ks] = src[ks];
        }
      } else {
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int ks = 0; ks < KG; ks++) {
            const int k = g * KG + ks;
            av[i][ks] = la[k * BM + ((wm0 + i * 16 + l16) ^ SA::swz(k))];
          }
      }
      if (rk_b[cur]) {
#pragma unroll
        for (int j = 0; j < FN; j++) {
          const T* src = lb + (wn0 + j * 16 + l16) * LDK + g * KG;
#pragma unroll
          for (int ks = 0; ks < KG; ks++) bv[j][ks] = src[ks];
        }
      } else {
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
          for (int ks = 0; ks < KG; ks++) {
            const int k = g * KG + ks;
            bv[j][ks] = lb[k * BN + ((wn0 + j * 16 + l16) ^ SB::swz(k))];
          }
      }
#pragma unroll
      for (int ks = 0; ks < KG; ks++)
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++) {
            if constexpr (PREC == PREC_F64) acc[i][j] = mfma_f64(av[i][ks], bv[j][ks], acc[i][j]);
            else acc[i][j] = mfma_f32(av[i][ks], bv[j][ks], acc[i][j]);
          }
      if (more) store(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
  } else {
    constexpr int LDK = BK + 8;
    __shared__ __bf16 lah[2][BM * LDK], lal[2][BM * LDK];
    __shared__ __bf16 lbh[2][BN * LDK], lbl[2][BN * LDK];
    int amode = 0, bmode = 0;
    if (c_begin < c_end) {
      load_chunk(c_begin, amode, bmode);
      sa.template store_rowk_split<LDK>(amode, lah[0], lal[0]);
      sb.template store_rowk_split<LDK>(bmode, lbh[0], lbl[0]);
    }
    __syncthreads();
    int cur = 0;
    for (int c = c_begin; c < c_end; c++) {
      const bool more = c + 1 < c_end;
      if (more) load_chunk(c + 1, amode, bmode);
      const int ko = (lane >> 4) * 8;
      bf8v ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; i++) {
        const int row = wm0 + i * 16 + (lane & 15);
        ah[i] = *reinterpret_cast<const bf8v*>(&lah[cur][row * LDK + ko]);
        al[i] = *reinterpret_cast<const bf8v*>(&lal[cur][row * LDK + ko]);
      }
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int row = wn0 + j * 16 + (lane & 15);
        bh[j] = *reinterpret_cast<const bf8v*>(&lbh[cur][row * LDK + ko]);
        bl[j] = *reinterpret_cast<const bf8v*>(&lbl[cur][row * LDK + ko]);
      }
#pragma unroll
      for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) {
          acc[i][j] = mfma_bf16(al[i], bh[j], acc[i][j]);
          acc[i][j] = mfma_bf16(ah[i], bl[j], acc[i][j]);
          acc[i][j] = mfma_bf16(ah[i], bh[j], acc[i][j]);
        }
      if (more) {
        sa.template store_rowk_split<LDK>(amode, lah[cur ^ 1], lal[cur ^ 1]);
        sb.template store_rowk_split<LDK>(bmode, lbh[cur ^ 1], lbl[cur ^ 1]);
      }
      __syncthreads();
      cur ^= 1;
    }
  }

  // epilogue: 16x16 fragment layout of gfx950 MFMA results
  //   f32 / bf16 : row = 4*(lane/16) + r, col = lane%16
  //   f64        : row = lane/16 + 4*r,   col = lane%16
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = (PREC == PREC_F64) ? ((lane >> 4) + 4 * r) : (4 * (lane >> 4) + r);
        const int m = m0 + wm0 + i * 16 + row;
        const int n = n0 + wn0 + j * 16 + (lane & 15);
        write_out<Acc>(P, tc, m, n, acc[i][j][r]);
      }
}

// C (+)= alpha * sum_s P[s], summed in FP64 in fixed split order.
template <class Acc>
__global__ void __launch_bounds__(256)
reduce_partials(const ReduceProb* __restrict__ probs, int nprob, long long total) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].elem0 <= e) lo = mid; else hi = mid - 1;
  }
  const ReduceProb& R = probs[lo];
  const long long local = e - R.elem0;
  const long long per = (long long)R.nb * R.M * R.N;
  if (local >= per) return;
  const int n = (int)(local % R.N);
  const int m = (int)((local / R.N) % R.M);
  const int b = (int)(local / ((long long)R.N * R.M));
  const Acc* P = reinterpret_cast<const Acc*>(R.P);
  double s = 0.0;
  for (int sp = 0; sp < R.ksplit; sp++) s += (double)P[sp * per + local];
  s *= R.alpha;
  const long long off = (long long)b * R.c_sb + (long long)m * R.c_sm + (long long)n * R.c_sn;
  if (R.c_is_f64) {
    double* c = reinterpret_cast<double*>(R.C) + off;
    *c = R.accumulate ? *c + s : s;
  } else {
    float* c = reinterpret_cast<float*>(R.C) + off;
    *c = R.accumulate ? (float)((double)*c + s) : (float)s;
  }
}

// Host-side launch helpers (C linkage inside the library only).
#define T3_GEMM_INST(PREC, BM, BN, WGM)                                                        \
  template __global__ void gemm_grouped<PREC, BM, BN, WGM>(const GemmProb*, int, const GemmSeg*);

}  // namespace t3

using namespace t3;

// tile config ids: 0 -> 64x64 (2x2 waves), 1 -> 128x128 (2x2), 2 -> 64x16 (4x1),
//                  3 -> 64x32 (2x2), 4 -> 32x64 (2x2)
extern "C" int t3_gemm_tile_m(int cfg) { return cfg == 1 ? 128 : (cfg == 4 ? 32 : 64); }
extern "C" int t3_gemm_tile_n(int cfg) {
  return cfg == 1 ? 128 : (cfg == 2 ? 16 : (cfg == 3 ? 32 : 64));
}
extern "C" int t3_gemm_chunk(int prec) { return prec == PREC_F64 ? 16 : 32; }

extern "C" hipError_t t3_launch_gemm(int prec, int cfg, const GemmProb* probs, int nprob,
                                     const GemmSeg* segs, long long nblocks, hipStream_t stream) {
  if (nblocks <= 0) return hipSuccess;
  dim3 grid((unsigned)nblocks), block(256);
#define T3_DISPATCH(PR)                                                                                    \
  if (cfg == 0) hipLaunchKernelGGL((gemm_grouped<PR, 64, 64, 2>), grid, block, 0, stream, probs, nprob, segs); \
  else if (cfg == 1) hipLaunchKernelGGL((gemm_grouped<PR, 128, 128, 2>), grid, block, 0, stream, probs, nprob, segs); \
  else if (cfg == 3) hipLaunchKernelGGL((gemm_grouped<PR, 64, 32, 2>), grid, block, 0, stream, probs, nprob, segs); \
  else if (cfg == 4) hipLaunchKernelGGL((gemm_grouped<PR, 32, 64, 2>), grid, block, 0, stream, probs, nprob, segs); \
  else hipLaunchKernelGGL((gemm_grouped<PR, 64, 16, 4>), grid, block, 0, stream, probs, nprob, segs);
  if (prec == PREC_F64) { T3_DISPATCH(PREC_F64) }
  else if (prec == PREC_F32) { T3_DISPATCH(PREC_F32) }
  else { T3_DISPATCH(PREC_BF16X3) }
#undef T3_DISPATCH
  return hipGetLastError();
}

extern "C" hipError_t t3_launch_reduce(int prec, const ReduceProb* probs, int nprob, long long total,
                                       hipStream_t stream) {
  if (total <= 0) return hipSuccess;
  dim3 grid((unsigned)((total + 255) / 256)), block(256);
  if (prec == PREC_F64) hipLaunchKernelGGL(reduce_partials<double>, grid, block, 0, stream, probs, nprob, total);
  else hipLaunchKernelGGL(reduce_partials<float>, grid, block, 0, stream, probs, nprob, total);
  return hipGetLastError();
}

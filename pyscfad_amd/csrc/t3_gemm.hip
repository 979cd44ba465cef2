// Grouped, batched, strided MFMA GEMM for gfx950 with three arithmetic modes:
//   PREC_F64    : v_mfma_f64_16x16x4f64, FP64 operands and accumulation
//   PREC_F32    : v_mfma_f32_16x16x4f32, FP32 operands and accumulation
//   PREC_BF16X3 : FP32 operands split on the way into LDS as x = hi + lo with
//                 hi, lo in bf16; three v_mfma_f32_16x16x32_bf16 per K-step
//                 (hi*hi, lo*hi, hi*lo), FP32 accumulation (Henry, Tang and
//                 Heinecke, ARITH 2019, "bf16x3").
//   PREC_BF16X6 : the same with a three-way split x = hi + mid + lo (24
//                 significant bits; |mid| <= 2^-9 |x|, |lo| <= 2^-18 |x|) and
//                 six products per K-step (hi*hi, hi*mid, mid*hi, hi*lo,
//                 lo*hi, mid*mid; the dropped mid*lo, lo*mid are at most
//                 2^-27 relative, lo*lo 2^-36), FP32 accumulation: FP32-level accuracy
//                 from bf16 matrix cores (Henry, Tang and Heinecke, "bf16x6").
//
// Every contraction of the (T) energy and its VJP is expressed as a list of
// problems C[b][m][n] (op)= alpha * sum_seg scale_seg * sum_k A_seg[b][m][k] B_seg[b][k][n]
// whose operands are strided views of the resident tensors (see t3_common.h).
// A block owns one (problem, batch, m-tile, n-tile, k-split) output tile, so
// every output element has exactly one writer: no atomics anywhere, and the
// result does not depend on scheduling.
//
// Staging: a K-chunk of each operand tile is read from HBM with 16-byte
// global loads into registers (rows past the edge are clamped to the last
// valid row, so the loads carry no branches), and written to LDS during the
// multiply phase of the previous chunk, negated for segments of scale -1 (the
// clamped rows are left as loaded: they reach only outputs that are not
// written).  The LDS layout is always [row][k] in 16-byte units permuted per
// row (lds_off); row-contiguous tiles are transposed in registers (each thread
// holds a VEC-row block).  Two LDS buffers and one barrier per chunk; the
// global loads of chunk c+2 are in flight while chunk c is multiplied (one
// register stage; bf16x6 keeps two, and its tiles with a 128-row side keep one
// LDS buffer and two barriers per chunk, for occupancy).
#include <type_traits>

#include "t3_common.h"

#define T3_GLOBAL __attribute__((address_space(1)))

namespace t3 {

template <int PREC> struct PrecTraits;
template <> struct PrecTraits<PREC_F64> { using T = double; using Acc = double; using V = d2v; static constexpr int BK = 16; static constexpr int VEC = 2; };
template <> struct PrecTraits<PREC_F32> { using T = float; using Acc = float; using V = f4v; static constexpr int BK = 32; static constexpr int VEC = 4; };
template <> struct PrecTraits<PREC_BF16X3> { using T = float; using Acc = float; using V = f4v; static constexpr int BK = 32; static constexpr int VEC = 4; };
template <> struct PrecTraits<PREC_BF16X6> { using T = float; using Acc = float; using V = f4v; static constexpr int BK = 32; static constexpr int VEC = 4; };

// instruction classes of __builtin_amdgcn_sched_group_barrier
constexpr int SG_VALU = 0x002, SG_MFMA = 0x008;

typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef uint32_t u2v __attribute__((ext_vector_type(2)));

// LDS tiles are [row][k] with a row pitch of LDK elements, split into 16-byte
// units of UE elements; unit u of row r is stored at position u ^ s(r).  The
// permutations s follow the gfx950 LDS bank structure (64 banks of 4 bytes):
//  * ds_read_b128 serves 16 lanes per pass; with l16 = lane & 15 and
//    g = lane >> 4, a pass takes groups g and g ^ 1 of a lane pair (0, 1) or
//    (2, 3), choosing g where l16 is in 0-3 or 12-15 and g ^ 1 where l16 is in
//    4-11, and is conflict-free when its 16 addresses fall in distinct 16-byte
//    bank groups modulo 256 bytes;
//  * ds_write_b128 serves 8 consecutive lanes per pass over 128 bytes of banks,
//    ds_write_b64 16 consecutive lanes.
// Fragment reads have lane (l16, g) read row l16 at unit 2g + u (FP32, FP64)
// or g (bf16); k-contiguous stores write whole rows; row-contiguous stores
// write one unit of rows 4 (FP32, bf16) or 2 (FP64) apart.
//  * FP32: 128-byte rows, s = (b2, b1^b3, b3^b4) (bit j of s listed j-th, bk =
//    bit k of r).  Reads and 16-byte stores are conflict-free; the 8-byte
//    row-contiguous stores of 64-row tiles take two passes, their minimum.
//  * FP64: 128-byte rows, s = (b1, b2, b2^b3): conflict-free reads and stores
//    of rows 2 apart; s depends on row bits 0-3 only, so the fragment reads at
//    16-row offsets share one address register (the 64x64 tile has none to
//    spare).
//  * bf16: 80-byte rows (4 units + 16 bytes of padding), s = (b2^b3, b4):
//    conflict-free reads; the 8-byte stores take two passes.
enum { SWZ_F32 = 0, SWZ_F64 = 1, SWZ_BF16 = 2 };
template <int LDK, int UE, int KIND>
__device__ __forceinline__ int lds_off(int row, int k) {
  int s;
  if constexpr (KIND == SWZ_F32) s = ((row >> 2) & 1) | (((row >> 1) ^ (row >> 3)) & 1) << 1 | (((row >> 3) ^ (row >> 4)) & 1) << 2;
  else if constexpr (KIND == SWZ_F64) s = ((row >> 1) & 1) | ((row >> 2) & 1) << 1 | (((row >> 2) ^ (row >> 3)) & 1) << 2;
  else s = (((row >> 2) ^ (row >> 3)) & 1) | ((row >> 4) & 1) << 1;
  return row * LDK + (((k / UE) ^ s) * UE) + (k % UE);
}

// LDS offsets of a wave's fragment reads: fragment i covers rows row0 + 16 i
// (row0 a multiple of 16).  The permutations use row bits below 5, so
// fragments i and i + 2 sit a fixed 32-row stride apart (FP64: bits below 4,
// a 16-row stride), and the reads of all fragments share one or two address
// registers with immediate offsets.
template <int LDK, int UE, int KIND>
struct FragOff {
  int o0, o1;
  __device__ __forceinline__ FragOff(int row, int k)
      : o0(lds_off<LDK, UE, KIND>(row, k)), o1(lds_off<LDK, UE, KIND>(row + 16, k)) {}
  __device__ __forceinline__ int operator()(int i) const {
    if constexpr (KIND == SWZ_F64) return o0 + i * 16 * LDK;
    else return ((i & 1) ? o1 : o0) + (i >> 1) * 32 * LDK;
  }
};

// Register stage of one operand tile (ROWS x BK elements, NE per thread).
//   KCONT  : thread holds NV vectors, each VEC consecutive k of one row
//            (thread t: k-vector t % (BK/VEC), rows t / (BK/VEC) + q * RSTEP)
//   ROWCONT: thread holds a VEC-row x NV-k block, one vector per k
//            (thread t: row block t % RB, k block t / RB)
template <class T, class V, int ROWS, int BK, int VEC>
struct Stage {
  static constexpr int NE = ROWS * BK / 256;
  static constexpr int NV = NE / VEC;
  static constexpr int KV = BK / VEC;         // k-vectors per row (KCONT)
  static constexpr int RSTEP = 256 / KV;      // rows per q step (KCONT)
  static constexpr int RB = ROWS / VEC;       // row blocks (ROWCONT)
  static constexpr int SWZ = sizeof(T) == 8 ? SWZ_F64 : SWZ_F32;   // LDS unit permutation
  static_assert(NV >= 1 && NE % VEC == 0 && RB * (BK / NV) == 256, "tile must give each thread whole vectors");
  V r[NV];
  int mode;
  bool neg;   // segment scale -1 (the only scales are +1 and -1)

  // Byte offsets of this thread's NV vectors from the tile origin.  Rows past
  // the edge are clamped to the last valid row, so every load reads real
  // data; those rows feed only output rows or columns past M or N, which are
  // never written, so they are not zeroed.  The offsets depend only on the
  // segment's strides and are computed once per segment.
  static __device__ __forceinline__ void offsets(int mode, long long s_row, long long s_k, int rows_valid,
                                                 uint32_t* off) {
    const int t = threadIdx.x;
    if (mode == LOAD_ROWCONT) {
      const int rb = min(t % RB, (rows_valid - 1) / VEC), kb = t / RB;
#pragma unroll
      for (int j = 0; j < NV; j++)
        off[j] = (uint32_t)(((long long)rb * VEC + (long long)(kb * NV + j) * s_k) * (long long)sizeof(T));
    } else {
      const int kv = t % KV, row0 = t / KV;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int row = min(row0 + q * RSTEP, rows_valid - 1);
        off[q] = (uint32_t)(((long long)row * s_row + kv * VEC) * (long long)sizeof(T));
      }
    }
  }

  __device__ __forceinline__ void load(int mode_, const char* base, const uint32_t* off, bool neg_) {
    mode = mode_;
    neg = neg_;
    const T3_GLOBAL char* g = (const T3_GLOBAL char*)base;
#pragma unroll
    for (int q = 0; q < NV; q++) r[q] = *(const T3_GLOBAL V*)(g + off[q]);
  }

  // LDS [row][k], row pitch LDK elements
  template <int LDK>
  __device__ __forceinline__ void store(T* __restrict__ lds) const {
    const int t = threadIdx.x;
    const T s = neg ? (T)-1 : (T)1;
    if (mode == LOAD_ROWCONT) {
      const int rb = t % RB, kb = t / RB;
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        const int row = rb * VEC + e;
        if constexpr (NV * sizeof(T) <= 16 && NV > 1) {   // within one 16-byte unit
          typedef T tnv __attribute__((ext_vector_type(NV)));
          tnv x;
#pragma unroll
          for (int j = 0; j < NV; j++) x[j] = r[j][e];
          *reinterpret_cast<tnv*>(lds + lds_off<LDK, VEC, SWZ>(row, kb * NV)) = x * s;
        } else {
#pragma unroll
          for (int j = 0; j < NV; j++) lds[lds_off<LDK, VEC, SWZ>(row, kb * NV + j)] = r[j][e] * s;
        }
      }
    } else {
      const int kv = t % KV, row0 = t / KV;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int row = row0 + q * RSTEP;
        *reinterpret_cast<V*>(lds + lds_off<LDK, VEC, SWZ>(row, kv * VEC)) = r[q] * s;
      }
    }
  }

  // LDS [row][k] as split bf16 pairs x = hi + lo (FP32 operands only).  The
  // sign is a uniform branch (scale +1 segments spend nothing on it; for -1
  // the negation folds into source modifiers).  Per pair of values: one
  // packed RNE conversion gives hi, a shift and a mask give hi back as FP32,
  // two subtractions the remainders, one packed conversion lo.
  template <int LDK>
  __device__ __forceinline__ void store_split(__bf16* __restrict__ hi, __bf16* __restrict__ lo) const {
    if (neg) store_split_as<LDK, true>(hi, lo);
    else store_split_as<LDK, false>(hi, lo);
  }
  template <int LDK, bool NEG>
  __device__ __forceinline__ void store_split_as(__bf16* __restrict__ hi, __bf16* __restrict__ lo) const {
    const int t = threadIdx.x;
    auto split2 = [](float x0, float x1, uint32_t& h, uint32_t& l) {
      if constexpr (NEG) { x0 = -x0; x1 = -x1; }
      const bf2v hb = __builtin_convertvector(f2v{x0, x1}, bf2v);
      h = __builtin_bit_cast(uint32_t, hb);
      const float r0 = x0 - __uint_as_float(h << 16);
      const float r1 = x1 - __uint_as_float(h & 0xffff0000u);
      l = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v{r0, r1}, bf2v));
    };
    auto split4 = [&](f4v x, u2v& h, u2v& l) {
      uint32_t h0, l0, h1, l1;
      split2(x[0], x[1], h0, l0);
      split2(x[2], x[3], h1, l1);
      h = u2v{h0, h1};
      l = u2v{l0, l1};
    };
    if (mode == LOAD_ROWCONT) {
      const int rb = t % RB, kb = t / RB;
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        const int off = lds_off<LDK, 8, SWZ_BF16>(rb * VEC + e, kb * NV);
        if constexpr (NV == 4) {
          u2v h, l;
          split4(f4v{r[0][e], r[1][e], r[2][e], r[3][e]}, h, l);
          *reinterpret_cast<u2v*>(hi + off) = h;
          *reinterpret_cast<u2v*>(lo + off) = l;
        } else if constexpr (NV == 2) {
          uint32_t h, l;
          split2(r[0][e], r[1][e], h, l);
          *reinterpret_cast<uint32_t*>(hi + off) = h;
          *reinterpret_cast<uint32_t*>(lo + off) = l;
        } else {
          static_assert(NV == 1, "split store handles 1, 2 or 4 k values per thread");
          uint32_t h, l;
          split2(r[0][e], 0.0f, h, l);
          hi[off] = __builtin_bit_cast(__bf16, (uint16_t)h);
          lo[off] = __builtin_bit_cast(__bf16, (uint16_t)l);
        }
      }
    } else {
      const int kv = t % KV, row0 = t / KV;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int row = row0 + q * RSTEP;
        u2v h, l;
        split4(r[q], h, l);
        const int off = lds_off<LDK, 8, SWZ_BF16>(row, kv * VEC);
        *reinterpret_cast<u2v*>(hi + off) = h;
        *reinterpret_cast<u2v*>(lo + off) = l;
      }
    }
  }

  // LDS [row][k] as three bf16 parts x = hi + mid + lo (FP32 operands only).
  // Each remainder is exact in FP32; per pair of values: three packed RNE
  // conversions, four shifts / masks, four subtractions.
  template <int LDK>
  __device__ __forceinline__ void store_split3(__bf16* __restrict__ hi, __bf16* __restrict__ mid,
                                               __bf16* __restrict__ lo) const {
    if (neg) store_split3_as<LDK, true>(hi, mid, lo);
    else store_split3_as<LDK, false>(hi, mid, lo);
  }
  // The same split without branches (128- and 64-row tiles, NV = 4 or 2):
  // the sign of the segment is an XOR of the sign bits, the load mode a
  // per-element select between the row-contiguous (transposed) and the
  // k-contiguous element sets and offsets.  Branch-free, the store shares the
  // basic block of the MFMAs it is meant to be scheduled between.  With
  // NV = 2 every mode writes four pairs of k per plane (4-byte stores).
  template <int LDK>
  __device__ __forceinline__ void store_split3_flat(__bf16* __restrict__ hi, __bf16* __restrict__ mid,
                                                    __bf16* __restrict__ lo) const {
    static_assert((NV == 4 || NV == 2) && VEC == 4, "the branch-free split store takes 2 or 4 4-vectors per thread");
    const int t = threadIdx.x;
    const bool rc = mode == LOAD_ROWCONT;
    const uint32_t sg = neg ? 0x80000000u : 0u;
    auto part = [](float& x0, float& x1) {
      const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v{x0, x1}, bf2v));
      x0 -= __uint_as_float(h << 16);
      x1 -= __uint_as_float(h & 0xffff0000u);
      return h;
    };
    auto split2 = [&](float x0, float x1, uint32_t& h, uint32_t& m, uint32_t& l) {
      x0 = __uint_as_float(__float_as_uint(x0) ^ sg);
      x1 = __uint_as_float(__float_as_uint(x1) ^ sg);
      h = part(x0, x1);
      m = part(x0, x1);
      l = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v{x0, x1}, bf2v));
    };
    const int rb = t % RB, kb = t / RB;        // row-contiguous: row block, k block
    const int kv = t % KV, row0 = t / KV;      // k-contiguous: k vector, first row
    if constexpr (NV == 2) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        // row-contiguous: k pair (kb*2, +1) of row rb*4 + q; k-contiguous:
        // k pair kv*4 + 2(q&1) of row row0 + (q>>1) RSTEP
        const float x0 = rc ? r[0][q] : r[q >> 1][2 * (q & 1)];
        const float x1 = rc ? r[1][q] : r[q >> 1][2 * (q & 1) + 1];
        const int off = rc ? lds_off<LDK, 8, SWZ_BF16>(rb * VEC + q, kb * NV)
                           : lds_off<LDK, 8, SWZ_BF16>(row0 + (q >> 1) * RSTEP, kv * VEC + 2 * (q & 1));
        uint32_t h, m, l;
        split2(x0, x1, h, m, l);
        *reinterpret_cast<uint32_t*>(hi + off) = h;
        *reinterpret_cast<uint32_t*>(mid + off) = m;
        *reinterpret_cast<uint32_t*>(lo + off) = l;
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const f4v x = rc ? f4v{r[0][q], r[1][q], r[2][q], r[3][q]} : r[q];
      const int off = rc ? lds_off<LDK, 8, SWZ_BF16>(rb * VEC + q, kb * NV)
                         : lds_off<LDK, 8, SWZ_BF16>(row0 + q * RSTEP, kv * VEC);
      uint32_t h0, m0, l0, h1, m1, l1;
      split2(x[0], x[1], h0, m0, l0);
      split2(x[2], x[3], h1, m1, l1);
      *reinterpret_cast<u2v*>(hi + off) = u2v{h0, h1};
      *reinterpret_cast<u2v*>(mid + off) = u2v{m0, m1};
      *reinterpret_cast<u2v*>(lo + off) = u2v{l0, l1};
    }
  }
  template <int LDK, bool NEG>
  __device__ __forceinline__ void store_split3_as(__bf16* __restrict__ hi, __bf16* __restrict__ mid,
                                                  __bf16* __restrict__ lo) const {
    const int t = threadIdx.x;
    auto part = [](float& x0, float& x1) {   // bf16 pair of (x0, x1); x0, x1 <- the remainders
      const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v{x0, x1}, bf2v));
      x0 -= __uint_as_float(h << 16);
      x1 -= __uint_as_float(h & 0xffff0000u);
      return h;
    };
    auto split2 = [&](float x0, float x1, uint32_t& h, uint32_t& m, uint32_t& l) {
      if constexpr (NEG) { x0 = -x0; x1 = -x1; }
      h = part(x0, x1);
      m = part(x0, x1);
      l = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v{x0, x1}, bf2v));
    };
    auto split4 = [&](f4v x, u2v& h, u2v& m, u2v& l) {
      uint32_t h0, m0, l0, h1, m1, l1;
      split2(x[0], x[1], h0, m0, l0);
      split2(x[2], x[3], h1, m1, l1);
      h = u2v{h0, h1};
      m = u2v{m0, m1};
      l = u2v{l0, l1};
    };
    if (mode == LOAD_ROWCONT) {
      const int rb = t % RB, kb = t / RB;
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        const int off = lds_off<LDK, 8, SWZ_BF16>(rb * VEC + e, kb * NV);
        if constexpr (NV == 4) {
          u2v h, m, l;
          split4(f4v{r[0][e], r[1][e], r[2][e], r[3][e]}, h, m, l);
          *reinterpret_cast<u2v*>(hi + off) = h;
          *reinterpret_cast<u2v*>(mid + off) = m;
          *reinterpret_cast<u2v*>(lo + off) = l;
        } else if constexpr (NV == 2) {
          uint32_t h, m, l;
          split2(r[0][e], r[1][e], h, m, l);
          *reinterpret_cast<uint32_t*>(hi + off) = h;
          *reinterpret_cast<uint32_t*>(mid + off) = m;
          *reinterpret_cast<uint32_t*>(lo + off) = l;
        } else {
          static_assert(NV == 1, "split store handles 1, 2 or 4 k values per thread");
          uint32_t h, m, l;
          split2(r[0][e], 0.0f, h, m, l);
          hi[off] = __builtin_bit_cast(__bf16, (uint16_t)h);
          mid[off] = __builtin_bit_cast(__bf16, (uint16_t)m);
          lo[off] = __builtin_bit_cast(__bf16, (uint16_t)l);
        }
      }
    } else {
      const int kv = t % KV, row0 = t / KV;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const int row = row0 + q * RSTEP;
        u2v h, m, l;
        split4(r[q], h, m, l);
        const int off = lds_off<LDK, 8, SWZ_BF16>(row, kv * VEC);
        *reinterpret_cast<u2v*>(hi + off) = h;
        *reinterpret_cast<u2v*>(mid + off) = m;
        *reinterpret_cast<u2v*>(lo + off) = l;
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
  int lo = 0;
  long long local;
  if (probs[0].interleave) {   // problems of one tile grid, their tiles interleaved
    lo = (int)((unsigned)bid % (unsigned)nprob);
    local = (unsigned)bid / (unsigned)nprob;
  } else {
    int hi = nprob - 1;
    while (lo < hi) {   // last problem with block0 <= bid
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].block0 <= bid) lo = mid; else hi = mid - 1;
    }
    local = bid - probs[lo].block0;
  }
  const GemmProb& P = probs[lo];
  // a launch has fewer than 2^31 tiles (checked by the plan): 32-bit arithmetic.
  // The output tiles of one k-split are consecutive (n, then m, then split):
  // the blocks that run together on an XCD read the same K range of the A
  // rows and B columns, which their L2 then serves once.  (With the split
  // fastest, the o=64/v=512 t2 particle adjoint fetched every A panel once
  // per n-tile and every B panel once per m-tile: 38.7 GB per triple.)
  unsigned u = (unsigned)local;
  TileCoord tc;
  tc.p = lo;
  tc.tn = (int)(u % (unsigned)P.tiles_n); u /= (unsigned)P.tiles_n;
  tc.tm = (int)(u % (unsigned)P.tiles_m); u /= (unsigned)P.tiles_m;
  tc.split = (int)(u % (unsigned)P.ksplit); u /= (unsigned)P.ksplit;
  tc.b = (int)u;
  return tc;
}

template <class Acc>
__device__ __forceinline__ void write_out(const GemmProb& P, const TileCoord& tc, int m, int n,
                                          Acc v) {
  if (P.mode == GEMM_PARTIAL) {
    if (m >= P.M || n >= P.N) return;
    T3_GLOBAL Acc* out = (T3_GLOBAL Acc*)P.P;
    out[((long long)(tc.split * P.nb + tc.b) * P.M + m) * P.N + n] = v;
    return;
  }
  if (m >= P.Mv || n >= P.Nv) return;
  const long long off = (long long)tc.b * P.c_sb + (long long)m * P.c_sm + (long long)n * P.c_sn;
  const double x = P.alpha * (double)v;
  if (P.c_is_f64) {
    T3_GLOBAL double* c = (T3_GLOBAL double*)P.C + off;
    *c = (P.mode == GEMM_ACCUM) ? (*c + x) : x;
  } else {
    T3_GLOBAL float* c = (T3_GLOBAL float*)P.C + off;
    *c = (P.mode == GEMM_ACCUM) ? (float)((double)*c + x) : (float)x;
  }
}

// BM x BN tile, 4 waves arranged WGM x (4/WGM).  Register budget: four waves
// per SIMD for tiles up to 64x64 (four blocks per CU fit in LDS), two above;
// bf16x6, whose three-part LDS tiles (double-buffered up to 64x64, single
// above) fit two blocks per CU at every shape, gets two.
template <int PREC, int BM, int BN> struct MinBlocks {
  static constexpr int value = PREC == PREC_BF16X6 ? 2 : ((BM * BN <= 64 * 64) ? 4 : 2);
};

template <int PREC, int BM, int BN, int WGM>
__global__ void __launch_bounds__(256, (MinBlocks<PREC, BM, BN>::value))
gemm_grouped(const GemmProb* __restrict__ probs, int nprob, const GemmSeg* __restrict__ segs) {
  using TR = PrecTraits<PREC>;
  using T = typename TR::T;
  using Acc = typename TR::Acc;
  using V = typename TR::V;
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
  const int g = lane >> 4, l16 = lane & 15;

  Frag acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++) acc[i][j] = Frag{0, 0, 0, 0};

  // FP64, FP32, bf16x3: one register stage per operand.  When step c starts
  // (LDS buffer c & 1 holds chunk c), the stage holds chunk c+1; after the
  // step's fragment reads are issued it is written to the other buffer and
  // refilled with chunk c+2, whose loads then have the whole multiply phase
  // to land.  The MFMAs touch no memory, so the compiler may move the second
  // half of them past the closing barrier, overlapping them with the next
  // step's reads.
  // bf16x6 keeps two stages (chunk c+2 loaded while c is multiplied, c+1
  // split and written between its MFMAs): its three-way split is long VALU
  // work, which the scheduler interleaves with the MFMAs.
  constexpr int NSTAGE = PREC == PREC_BF16X6 ? 2 : 1;
  using SA = Stage<T, V, BM, BK, VEC>;
  using SB = Stage<T, V, BN, BK, VEC>;
  SA sa[NSTAGE];
  SB sb[NSTAGE];

  // Segment walker: the chunks of this block are loaded in increasing order;
  // the tile origins advance by a fixed step per chunk (and a carry when k2
  // wraps into the next k1), the per-thread offsets change only between
  // segments.
  int si = P.seg0;
  while (si + 1 < P.seg0 + P.nseg && segs[si + 1].chunk0 <= c_begin) si++;
  const char *pa = nullptr, *pb = nullptr;
  long long da2 = 0, db2 = 0, da1 = 0, db1 = 0;
  int k2 = 0, K2 = 0, seg_end = 0, amode = 0, bmode = 0;
  bool neg = false;
  uint32_t oa[SA::NV], ob[SB::NV];
  auto seg_init = [&](int c) {
    const GemmSeg& S = segs[si];
    // K1 * K2 < 2^31 for every segment (checked by the plan): 32-bit division
    const unsigned kk = (unsigned)(c - S.chunk0) * BK;
    const long long k1 = kk / (unsigned)S.K2;
    k2 = (int)(kk % (unsigned)S.K2);
    K2 = S.K2;
    constexpr long long E = sizeof(T);
    pa = reinterpret_cast<const char*>(S.A) + (tc.b * S.a_sb + m0 * S.a_sm + k1 * S.a_sk1 + k2 * S.a_sk2) * E;
    pb = reinterpret_cast<const char*>(S.B) + (tc.b * S.b_sb + n0 * S.b_sn + k1 * S.b_sk1 + k2 * S.b_sk2) * E;
    da2 = BK * S.a_sk2 * E;
    db2 = BK * S.b_sk2 * E;
    da1 = (S.a_sk1 - (long long)S.K2 * S.a_sk2) * E;
    db1 = (S.b_sk1 - (long long)S.K2 * S.b_sk2) * E;
    amode = S.a_mode;
    bmode = S.b_mode;
    neg = S.scale < 0.0f;
    seg_end = (si + 1 < P.seg0 + P.nseg) ? segs[si + 1].chunk0 : P.nchunks;
    SA::offsets(amode, S.a_sm, S.a_sk2, rows_a, oa);
    SB::offsets(bmode, S.b_sn, S.b_sk2, rows_b, ob);
  };
  int lc = c_begin;   // chunk the walker points at
  if (c_begin < c_end) seg_init(c_begin);
  // load the walker's chunk, then step to the next one (the last chunk is
  // loaded again at the tail, which keeps the loop free of branches)
  auto load_next = [&](SA& xa, SB& xb) {
    xa.load(amode, pa, oa, neg);
    xb.load(bmode, pb, ob, false);
    if (lc + 1 < c_end) {
      lc++;
      if (lc == seg_end) {
        si++;
        seg_init(lc);
      } else {
        k2 += BK;
        pa += da2;
        pb += db2;
        if (k2 == K2) {
          k2 = 0;
          pa += da1;
          pb += db1;
        }
      }
    }
  };
  const int nch = c_end - c_begin;

  if constexpr (PREC == PREC_F64 || PREC == PREC_F32) {
    // K is consumed in a permuted order: in step ks the lanes of group
    // g = lane/16 multiply k = g*KG + ks (the sum over k is order-free), so a
    // lane's KG values of a row are contiguous in LDS (two 16-byte reads).
    constexpr int KG = BK / 4;
    constexpr int LDK = BK;   // 128-byte rows, unit-permuted
    constexpr int SWZ = SA::SWZ;
    __shared__ __attribute__((aligned(16))) T lds_a[2][BM * LDK];
    __shared__ __attribute__((aligned(16))) T lds_b[2][BN * LDK];
    if (nch > 0) {
      load_next(sa[0], sb[0]);
      sa[0].template store<LDK>(lds_a[0]);
      sb[0].template store<LDK>(lds_b[0]);
      load_next(sa[0], sb[0]);
    }
    __syncthreads();
    using FO = FragOff<LDK, VEC, SWZ>;
    FO fa[KG / VEC] = {FO(wm0 + l16, g * KG), FO(wm0 + l16, g * KG + VEC)};
    FO fb[KG / VEC] = {FO(wn0 + l16, g * KG), FO(wn0 + l16, g * KG + VEC)};
    static_assert(KG / VEC == 2, "two k-vectors per lane group");
    // chunk c_begin + it sits in LDS buffer it & 1
    auto step = [&](int it, auto par) {
      constexpr int p = decltype(par)::value;
      // reads in the order the MFMAs consume them (k-vector u outermost), so the
      // first MFMAs wait only for the first half of the reads
      V av[FM][KG / VEC], bv[FN][KG / VEC];
#pragma unroll
      for (int u = 0; u < KG / VEC; u++) {
#pragma unroll
        for (int i = 0; i < FM; i++) av[i][u] = *reinterpret_cast<const V*>(lds_a[p] + fa[u](i));
#pragma unroll
        for (int j = 0; j < FN; j++) bv[j][u] = *reinterpret_cast<const V*>(lds_b[p] + fb[u](j));
        __builtin_amdgcn_sched_barrier(0);   // keep the reads in k-vector order
      }
      // Tiles up to 64x64 store and refill unconditionally (the last step
      // writes the buffer no later step reads, and the walker reloads its
      // last chunk): without the branch their schedule measured faster,
      // while the larger tiles' was slower.
      if (BM * BN <= 64 * 64 || it + 1 < nch) {
        sa[0].template store<LDK>(lds_a[p ^ 1]);
        sb[0].template store<LDK>(lds_b[p ^ 1]);
        load_next(sa[0], sb[0]);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int ks = 0; ks < KG; ks++) {
        // the MFMAs of the second k-vector stay after those of the first, so
        // the first half waits only for its own reads
        if (ks == VEC) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++) {
            const T a = av[i][ks / VEC][ks % VEC], b = bv[j][ks / VEC][ks % VEC];
            if constexpr (PREC == PREC_F64) acc[i][j] = mfma_f64(a, b, acc[i][j]);
            else acc[i][j] = mfma_f32(a, b, acc[i][j]);
          }
      }
      __syncthreads();
    };
    // straight-line pairs keep the LDS buffer index a constant
    int it = 0;
    for (; it + 1 < nch; it += 2) {
      step(it, std::integral_constant<int, 0>{});
      step(it + 1, std::integral_constant<int, 1>{});
    }
    if (it < nch) step(it, std::integral_constant<int, 0>{});
  } else if constexpr (PREC == PREC_BF16X6) {
    constexpr int LDK = BK + 8;   // bf16 elements; 80-byte rows keep 16-byte reads aligned
    // Tiles whose double-buffered three-part planes would exceed 80 KB (a
    // 128-row side: 92-123 KB, one block per CU) keep one LDS buffer, so that
    // two or three blocks share a CU and hide each other's barrier and load
    // waits; a step then reads its fragments, waits until every wave has, and
    // writes the next chunk into the same buffer.
    constexpr int NBUF = 3 * 2 * (BM + BN) * LDK * 2 > 80 * 1024 ? 1 : 2;
    __shared__ __attribute__((aligned(16))) __bf16 la[3][NBUF][BM * LDK];   // [hi/mid/lo][buffer]
    __shared__ __attribute__((aligned(16))) __bf16 lb[3][NBUF][BN * LDK];
    if (nch > 0) {
      load_next(sa[0], sb[0]);
      load_next(sa[NSTAGE - 1], sb[NSTAGE - 1]);
      sa[0].template store_split3<LDK>(la[0][0], la[1][0], la[2][0]);
      sb[0].template store_split3<LDK>(lb[0][0], lb[1][0], lb[2][0]);
    }
    __syncthreads();
    const FragOff<LDK, 8, SWZ_BF16> fa(wm0 + l16, g * 8), fb(wn0 + l16, g * 8);
    // 128- and 64-row operand tiles take the branch-free split store
    auto store3 = [&](auto& st, __bf16* h, __bf16* m, __bf16* l) {
      constexpr int nv = std::remove_reference_t<decltype(st)>::NV;
      if constexpr (nv == 4 || nv == 2) st.template store_split3_flat<LDK>(h, m, l);
      else st.template store_split3<LDK>(h, m, l);
    };
    auto step = [&](int it, auto par) {
      constexpr int p = decltype(par)::value;
      constexpr int rbuf = NBUF == 2 ? p : 0, wbuf = NBUF == 2 ? p ^ 1 : 0;   // buffer read / written
      load_next(sa[p % NSTAGE], sb[p % NSTAGE]);
      bf8v ah[FM], am[FM], al[FM], bh[FN], bm[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; i++) {
        const int o = fa(i);
        ah[i] = *reinterpret_cast<const bf8v*>(&la[0][rbuf][o]);
        am[i] = *reinterpret_cast<const bf8v*>(&la[1][rbuf][o]);
        al[i] = *reinterpret_cast<const bf8v*>(&la[2][rbuf][o]);
      }
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int o = fb(j);
        bh[j] = *reinterpret_cast<const bf8v*>(&lb[0][rbuf][o]);
        bm[j] = *reinterpret_cast<const bf8v*>(&lb[1][rbuf][o]);
        bl[j] = *reinterpret_cast<const bf8v*>(&lb[2][rbuf][o]);
      }
      // one buffer: every wave's fragments are in registers before it is rewritten
      if constexpr (NBUF == 1) __syncthreads();
      // smallest products first
#pragma unroll
      for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) {
          acc[i][j] = mfma_bf16(am[i], bm[j], acc[i][j]);
          acc[i][j] = mfma_bf16(al[i], bh[j], acc[i][j]);
          acc[i][j] = mfma_bf16(ah[i], bl[j], acc[i][j]);
          acc[i][j] = mfma_bf16(am[i], bh[j], acc[i][j]);
          acc[i][j] = mfma_bf16(ah[i], bm[j], acc[i][j]);
          acc[i][j] = mfma_bf16(ah[i], bh[j], acc[i][j]);
        }
      // The split of the next chunk is written unconditionally (the last
      // step writes the buffer no later step reads; the walker reloads its
      // last chunk), so it shares the MFMAs' basic block, and its VALU work
      // is issued between the MFMAs, two instructions after each: they run
      // in the shadow of the 16-cycle bf16 MFMAs instead of after them.
      store3(sa[(p ^ 1) % NSTAGE], la[0][wbuf], la[1][wbuf], la[2][wbuf]);
      store3(sb[(p ^ 1) % NSTAGE], lb[0][wbuf], lb[1][wbuf], lb[2][wbuf]);
#pragma unroll
      for (int q = 0; q < FM * FN * 6; q++) {
        __builtin_amdgcn_sched_group_barrier(SG_MFMA, 1, 0);
        __builtin_amdgcn_sched_group_barrier(SG_VALU, 2, 0);
      }
      __syncthreads();
    };
    int it = 0;
    for (; it + 1 < nch; it += 2) {
      step(it, std::integral_constant<int, 0>{});
      step(it + 1, std::integral_constant<int, 1>{});
    }
    if (it < nch) step(it, std::integral_constant<int, 0>{});
  } else {
    constexpr int LDK = BK + 8;   // bf16 elements; 80-byte rows keep 16-byte reads aligned
    __shared__ __attribute__((aligned(16))) __bf16 lah[2][BM * LDK], lal[2][BM * LDK];
    __shared__ __attribute__((aligned(16))) __bf16 lbh[2][BN * LDK], lbl[2][BN * LDK];
    if (nch > 0) {
      load_next(sa[0], sb[0]);
      sa[0].template store_split<LDK>(lah[0], lal[0]);
      sb[0].template store_split<LDK>(lbh[0], lbl[0]);
      load_next(sa[0], sb[0]);
    }
    __syncthreads();
    const FragOff<LDK, 8, SWZ_BF16> fa(wm0 + l16, g * 8), fb(wn0 + l16, g * 8);
    auto step = [&](int it, auto par) {
      constexpr int p = decltype(par)::value;
      // Tiles up to 64x64: hi parts first, so the hi*hi products wait only
      // for them.  Larger tiles keep the compiler's interleaving (the pinned
      // order would keep all 4 fragment sets live and spill).
      constexpr bool PHASED = BM * BN <= 64 * 64;
      bf8v ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; i++) ah[i] = *reinterpret_cast<const bf8v*>(&lah[p][fa(i)]);
#pragma unroll
      for (int j = 0; j < FN; j++) bh[j] = *reinterpret_cast<const bf8v*>(&lbh[p][fb(j)]);
      if constexpr (PHASED) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < FM; i++) al[i] = *reinterpret_cast<const bf8v*>(&lal[p][fa(i)]);
#pragma unroll
      for (int j = 0; j < FN; j++) bl[j] = *reinterpret_cast<const bf8v*>(&lbl[p][fb(j)]);
      if (it + 1 < nch) {
        sa[0].template store_split<LDK>(lah[p ^ 1], lal[p ^ 1]);
        sb[0].template store_split<LDK>(lbh[p ^ 1], lbl[p ^ 1]);
        load_next(sa[0], sb[0]);
      }
      if constexpr (PHASED) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++) acc[i][j] = mfma_bf16(ah[i], bh[j], acc[i][j]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++) {
            acc[i][j] = mfma_bf16(al[i], bh[j], acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bl[j], acc[i][j]);
          }
      } else {
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++) {
            acc[i][j] = mfma_bf16(al[i], bh[j], acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bl[j], acc[i][j]);
            acc[i][j] = mfma_bf16(ah[i], bh[j], acc[i][j]);
          }
      }
      __syncthreads();
    };
    // straight-line pairs keep the LDS buffer index a constant
    int it = 0;
    for (; it + 1 < nch; it += 2) {
      step(it, std::integral_constant<int, 0>{});
      step(it + 1, std::integral_constant<int, 1>{});
    }
    if (it < nch) step(it, std::integral_constant<int, 0>{});
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
        const int row = (PREC == PREC_F64) ? (g + 4 * r) : (4 * g + r);
        const int m = m0 + wm0 + i * 16 + row;
        const int n = n0 + wn0 + j * 16 + l16;
        write_out<Acc>(P, tc, m, n, acc[i][j][r]);
      }
}

// C (+)= alpha * sum_s P[s], summed in FP64 in fixed split order.
// VE consecutive elements of a row per thread (VE > 1: every problem's N is a
// multiple of VE and nb * M * N < 2^31, so the partials load as 16-byte
// vectors and the index arithmetic is 32-bit); each element keeps its own sum.
template <class Acc, int VE>
__global__ void __launch_bounds__(256)
reduce_partials(const ReduceProb* __restrict__ probs, int nprob, long long total) {
  const long long e = ((long long)blockIdx.x * 256 + threadIdx.x) * VE;
  if (e >= total) return;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].elem0 <= e) lo = mid; else hi = mid - 1;
  }
  const ReduceProb& R = probs[lo];
  const T3_GLOBAL Acc* P = (const T3_GLOBAL Acc*)R.P;
  double s[VE];
  int b, m, n0;
  long long local;
  if constexpr (VE == 1) {
    local = e - R.elem0;
    const long long per = (long long)R.nb * R.M * R.N;
    if (local >= per) return;
    n0 = (int)(local % R.N);
    m = (int)((local / R.N) % R.M);
    b = (int)(local / ((long long)R.N * R.M));
    if (m >= R.Mv || n0 >= R.Nv) return;
    s[0] = 0.0;
    for (int sp = 0; sp < R.ksplit; sp++) s[0] += (double)P[sp * per + local];
  } else {
    typedef Acc accv __attribute__((ext_vector_type(VE)));
    const unsigned ul = (unsigned)(e - R.elem0), N = (unsigned)R.N, M = (unsigned)R.M;
    const unsigned per = (unsigned)R.nb * M * N;
    if (ul >= per) return;
    local = ul;
    n0 = (int)(ul % N);
    const unsigned mb = ul / N;
    m = (int)(mb % M);
    b = (int)(mb / M);
    if (m >= R.Mv || n0 >= R.Nv) return;
#pragma unroll
    for (int j = 0; j < VE; j++) s[j] = 0.0;
    const T3_GLOBAL accv* Pv = (const T3_GLOBAL accv*)(P + ul);
    const unsigned pstride = per / VE;
#pragma unroll 4
    for (int sp = 0; sp < R.ksplit; sp++) {
      const accv x = Pv[(unsigned)sp * pstride];
#pragma unroll
      for (int j = 0; j < VE; j++) s[j] += (double)x[j];
    }
  }
#pragma unroll
  for (int j = 0; j < VE; j++) {
    const int n = n0 + j;
    if (n >= R.Nv) break;
    const double x = s[j] * R.alpha;
    const long long off = (long long)b * R.c_sb + (long long)m * R.c_sm + (long long)n * R.c_sn;
    if (R.c_is_f64) {
      T3_GLOBAL double* c = (T3_GLOBAL double*)R.C + off;
      *c = R.accumulate ? *c + x : x;
    } else {
      T3_GLOBAL float* c = (T3_GLOBAL float*)R.C + off;
      *c = R.accumulate ? (float)((double)*c + x) : (float)x;
    }
  }
}

}  // namespace t3

using namespace t3;

// Tile configurations (4 waves each):
//   0: 64x64   1: 128x128 (FP32 / bf16x3 only: FP64 would spill)   3: 64x32
//   4: 32x64   5: 128x64   6: 64x128
extern "C" int t3_gemm_tile_m(int cfg) {
  switch (cfg) { case 1: case 5: return 128; case 4: return 32; default: return 64; }
}
extern "C" int t3_gemm_tile_n(int cfg) {
  switch (cfg) { case 1: case 6: return 128; case 3: return 32; default: return 64; }
}
extern "C" int t3_gemm_cfg_valid(int prec, int cfg) {
  return (cfg == 0 || cfg == 3 || cfg == 4 || cfg == 5 || cfg == 6 || (cfg == 1 && prec != PREC_F64)) ? 1 : 0;
}
extern "C" int t3_gemm_chunk(int prec) { return prec == PREC_F64 ? 16 : 32; }
extern "C" int t3_gemm_vec(int prec) { return prec == PREC_F64 ? 2 : 4; }

template <int PR>
static void launch_cfg(int cfg, dim3 grid, hipStream_t stream, const GemmProb* probs, int nprob,
                       const GemmSeg* segs) {
  dim3 block(256);
  switch (cfg) {
    case 0: hipLaunchKernelGGL((gemm_grouped<PR, 64, 64, 2>), grid, block, 0, stream, probs, nprob, segs); break;
    case 1:
      if constexpr (PR != PREC_F64)
        hipLaunchKernelGGL((gemm_grouped<PR, 128, 128, 2>), grid, block, 0, stream, probs, nprob, segs);
      break;
    case 3: hipLaunchKernelGGL((gemm_grouped<PR, 64, 32, 2>), grid, block, 0, stream, probs, nprob, segs); break;
    case 4: hipLaunchKernelGGL((gemm_grouped<PR, 32, 64, 2>), grid, block, 0, stream, probs, nprob, segs); break;
    case 5: hipLaunchKernelGGL((gemm_grouped<PR, 128, 64, 2>), grid, block, 0, stream, probs, nprob, segs); break;
    default: hipLaunchKernelGGL((gemm_grouped<PR, 64, 128, 2>), grid, block, 0, stream, probs, nprob, segs); break;
  }
}

extern "C" hipError_t t3_launch_gemm(int prec, int cfg, const GemmProb* probs, int nprob,
                                     const GemmSeg* segs, long long nblocks, hipStream_t stream) {
  if (nblocks <= 0) return hipSuccess;
  if (!t3_gemm_cfg_valid(prec, cfg) || nblocks >= (1LL << 31)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)nblocks);
  if (prec == PREC_F64) launch_cfg<PREC_F64>(cfg, grid, stream, probs, nprob, segs);
  else if (prec == PREC_F32) launch_cfg<PREC_F32>(cfg, grid, stream, probs, nprob, segs);
  else if (prec == PREC_BF16X3) launch_cfg<PREC_BF16X3>(cfg, grid, stream, probs, nprob, segs);
  else launch_cfg<PREC_BF16X6>(cfg, grid, stream, probs, nprob, segs);
  return hipGetLastError();
}

// vec: the plan found every problem's N a multiple of the vector width
// (4 FP32 / 2 FP64 partials) and every problem below 2^31 elements
extern "C" hipError_t t3_launch_reduce(int prec, const ReduceProb* probs, int nprob, long long total, int vec,
                                       hipStream_t stream) {
  if (total <= 0) return hipSuccess;
  const int ve = vec ? (prec == PREC_F64 ? 2 : 4) : 1;
  const long long threads = (total + ve - 1) / ve;
  dim3 grid((unsigned)((threads + 255) / 256)), block(256);
  if (prec == PREC_F64) {
    if (ve > 1) hipLaunchKernelGGL((reduce_partials<double, 2>), grid, block, 0, stream, probs, nprob, total);
    else hipLaunchKernelGGL((reduce_partials<double, 1>), grid, block, 0, stream, probs, nprob, total);
  } else {
    if (ve > 1) hipLaunchKernelGGL((reduce_partials<float, 4>), grid, block, 0, stream, probs, nprob, total);
    else hipLaunchKernelGGL((reduce_partials<float, 1>), grid, block, 0, stream, probs, nprob, total);
  }
  return hipGetLastError();
}

// Shared definitions of the gfx950 (T) kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace t3 {

// Arithmetic of the W/V contractions.
enum Prec : int { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16X3 = 2, PREC_BF16X6 = 3 };

typedef double d2v __attribute__((ext_vector_type(2)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

// One K-segment of a grouped GEMM problem:  C[b][m][n] += scale * sum_k A[b][m][k] * B[b][k][n]
// with k = (k1, k2), 0 <= k1 < K1, 0 <= k2 < K2 and K2 a multiple of the K-chunk.
// Element addresses (in elements):
//   A: b*a_sb + m*a_sm + k1*a_sk1 + k2*a_sk2
//   B: b*b_sb + k1*b_sk1 + k2*b_sk2 + n*b_sn
// Every operand is either row-contiguous (a_sm == 1 / b_sn == 1) or
// k-contiguous (a_sk2 == 1 / b_sk2 == 1); the host sets the modes and checks
// the vector alignment of every stride and base.
enum LoadMode : int { LOAD_ROWCONT = 0, LOAD_KCONT = 1 };

struct GemmSeg {
  const void* A;
  const void* B;
  long long a_sb, a_sm, a_sk1, a_sk2;
  long long b_sb, b_sn, b_sk1, b_sk2;
  int K1, K2;
  float scale;        // +1 or -1 (checked when the plan is built)
  int chunk0;   // first K-chunk of this segment in the problem's concatenated K
  int a_mode, b_mode;
};

enum GemmMode : int {
  GEMM_STORE = 0,     // C = alpha * acc
  GEMM_ACCUM = 1,     // C += alpha * acc
  GEMM_PARTIAL = 2,   // P[split][b][m][n] = acc   (reduced later in fixed split order)
};

struct GemmProb {
  int nb, M, N;
  int Mv, Nv;         // rows / columns of C that exist (<= M, N; tiles and operand loads use M, N)
  int seg0, nseg;
  int nchunks;        // total K-chunks over all segments
  int ksplit;
  int tiles_m, tiles_n;
  long long block0;   // first block of this problem in the launch
  int interleave;     // same in every problem of a launch: 1 when all have the same tile grid and
                      // block t is tile t / nprob of problem t % nprob (block0 then unused)
  void* C;
  long long c_sb, c_sm, c_sn;
  int mode;
  int c_is_f64;       // output element type: 1 double, 0 float
  double alpha;
  void* P;            // GEMM_PARTIAL scratch, acc type, [ksplit][nb][M][N]
};

// Reduction of split-K partials into an output:  C (+)= alpha * sum_s P[s]
struct ReduceProb {
  const void* P;
  int ksplit, nb, M, N;
  int Mv, Nv;         // rows / columns of C that exist
  void* C;
  long long c_sb, c_sm, c_sn;
  int accumulate;
  int c_is_f64;
  double alpha;
  long long elem0;    // first element index of this problem in the launch
};


// ---- orbit-cube epilogue (t3_epilogue.hip) ----
constexpr int EPI_T = 8;          // tile edge
constexpr int EPI_PSTRIDE = 32;   // Q, Rsum, 3 x 8 vir partials

struct EpiTriple {
  int i, j, k;
  double fo;          // 1, 1/2, 1/6
  const void* w0;     // V^3 cube
  void* G;            // V^3 (backward)
  void* H;            // V^3 (backward; unused in the LNO form, where H = G)
  double* part;       // [ncube][EPI_PSTRIDE]
  const double* cm;   // [d][CM/CN][V][V] pair matrices of the V term (vcm_build)
  const void* s;      // LNO form: V^3 cube of the slot-projected W
  void* GS;           // LNO form, backward: V^3 adjoint of s
};

struct EpiArgs {
  int nocc, nvir, V;
  int ntile, ncube;
  int backward;
  int lno;                 // 1: the LNO energy form (the triples carry s, GS)
  double et_bar;
  const double* eps_occ;   // (o)
  const double* eps_vir;   // (V), padded entries large and finite
  const double* t1;        // (o, v) reference layout
  const double* fvoP;      // (o, V): fvoP[r][z] = fock[o+z, r], zero padded
  const int* cube_tab;     // (ncube): tile triple A | B << 10 | C << 20 of each cube id
};

// Per-batch reduction of the epilogue partials (one block, fixed order):
//   E += sum_t 2 f_t sum_cube Q;  eps_bar[occ of t] -= sum_cube Rsum;
//   eps_bar[o + x] += sum of the vir partials of x.
struct EpiReduceArgs {
  int ntriple, ncube, ntile, nocc, V;
  const EpiTriple* triples;
  double* energy;      // scalar accumulator
  double* eps_bar;     // (o + V)
  int backward;
};

// ---- V-term adjoint kernels (t3_prep.hip) ----
struct VpTriple {
  int i, j, k;
  const void* H;
  double* out;     // 19 V^2: [0,12) contractions, [12,18) CM_d / CN_d, [18] 6 vector sums
};
constexpr int VP_OUT = 0, VP_CM = 12, VP_VSUM = 18, VP_TOTAL = 19;   // offsets in V^2 units

struct VpSlab {
  int p, q;
  int c0, nc;       // contributions [c0, c0+nc) of the contribution list
};

struct VpContrib {
  int slot;         // triple slot in the batch
  int d;
  int transposed;
};

__device__ __forceinline__ d4v mfma_f64(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma_f32(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma_bf16(bf8v a, bf8v b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace t3

// Kernels of the LNO (T) form (occupied projector mat, lno_ccsd_t.c,
// lno_ccsd_t_vjp.c, pyscfad/lno/ccsd_t.py:70-237): layout conversions between
// the arrays that entry point takes and the canonical ones, projections of one
// occupied slot by mat, and the mat cotangent.  All FP64, no atomics; every
// sum runs in a fixed order.
//
// A tensor is viewed as [A][o][B] around the occupied slot it is projected on.
//   project : out[a][i][b] (+)= sum_l K(l, i) in[a][l][b]
//             K(l, i) = mat[l][i] (forward: the slot takes mat's column), or
//             mat[i][l] (back-projection of a cotangent onto the input slot)
//   matbar  : C[l][i] (+)= sum_{a,b} X[a][l][b] Y[a][i][b], per-block partials
//             over fixed position ranges, then summed over blocks in order
#include <algorithm>

#include "t3_common.h"

namespace t3 {

constexpr int LNO_TT = 32;          // transpose tile edge
constexpr int LNO_PROJ_OUT = 16;    // projected outputs per thread
constexpr int LNO_MB_POS = 32;      // positions per LDS stage of the mat cotangent
constexpr int LNO_MB_BLOCKS = 512;  // partial blocks of the mat cotangent

__device__ __forceinline__ long long swap_pair(long long x, int o, int swap) {
  return swap ? (x % o) * o + x / o : x;
}

// out[pc(c)][pr(r)] = in[r][c] for an (rows x cols) input; pr / pc swap the two
// occupied indices of a row / column index of o*o values when asked.
__global__ void __launch_bounds__(256) lno_transpose(double* __restrict__ out, const double* __restrict__ in,
                                                     long long rows, long long cols, int o, int swap_r,
                                                     int swap_c) {
  __shared__ double tile[LNO_TT][LNO_TT + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const long long tiles_c = (cols + LNO_TT - 1) / LNO_TT;
  const long long tiles = tiles_c * ((rows + LNO_TT - 1) / LNO_TT);
  for (long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    const long long r0 = (t / tiles_c) * LNO_TT, c0 = (t % tiles_c) * LNO_TT;
    for (int q = 0; q < 4; q++) {
      const long long r = r0 + ty + 8 * q, c = c0 + tx;
      tile[ty + 8 * q][tx] = (r < rows && c < cols) ? in[r * cols + c] : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < 4; q++) {
      const long long c = c0 + ty + 8 * q, r = r0 + tx;
      if (r < rows && c < cols) out[swap_pair(c, o, swap_c) * rows + swap_pair(r, o, swap_r)] = tile[tx][ty + 8 * q];
    }
    __syncthreads();
  }
}

// fock (n, n) from fvo (v, o): fock[o + a][i] = fvo[a][i], zero elsewhere (to_fock);
// fvo (v, o) from fock (n, n): fvo[a][i] = fock[o + a][i] (otherwise).
__global__ void lno_fock(double* __restrict__ out, const double* __restrict__ in, int o, int v, int to_fock) {
  const int n = o + v;
  const long long total = to_fock ? (long long)n * n : (long long)v * o;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    if (to_fock) {
      const int r = (int)(e / n), c = (int)(e % n);
      out[e] = (r >= o && c < o) ? in[(long long)(r - o) * o + c] : 0.0;
    } else {
      const long long a = e / o, i = e % o;
      out[e] = in[(o + a) * n + i];
    }
  }
}

// out[a][i][b] (+)= sum_l K(l, i) in[a][l][b]; blockIdx.y picks LNO_PROJ_OUT
// consecutive i.  in and out must not overlap.
__global__ void __launch_bounds__(256) lno_project(double* __restrict__ out, const double* __restrict__ in,
                                                   const double* __restrict__ mat, long long A, int o, long long B,
                                                   int transK, int accumulate) {
  const int i0 = blockIdx.y * LNO_PROJ_OUT;
  const int ni = min(LNO_PROJ_OUT, o - i0);
  const long long npos = A * B;
  for (long long pos = (long long)blockIdx.x * blockDim.x + threadIdx.x; pos < npos;
       pos += (long long)gridDim.x * blockDim.x) {
    const long long a = pos / B, b = pos % B;
    const double* src = in + a * o * B + b;
    double acc[LNO_PROJ_OUT];
#pragma unroll
    for (int q = 0; q < LNO_PROJ_OUT; q++) acc[q] = 0.0;
    for (int l = 0; l < o; l++) {
      const double x = src[(long long)l * B];
#pragma unroll
      for (int q = 0; q < LNO_PROJ_OUT; q++) {
        if (q < ni) {
          const int i = i0 + q;
          acc[q] += (transK ? mat[(long long)i * o + l] : mat[(long long)l * o + i]) * x;
        }
      }
    }
    double* dst = out + a * o * B + b;
#pragma unroll
    for (int q = 0; q < LNO_PROJ_OUT; q++) {
      if (q < ni) {
        double* p = dst + (long long)(i0 + q) * B;
        *p = accumulate ? *p + acc[q] : acc[q];
      }
    }
  }
}

// Partials of C[l][i] = sum_{a,b} X[a][l][b] Y[a][i][b]: block k takes the
// positions (a, b) (flattened a * B + b) of its fixed range, stages
// LNO_MB_POS of them at a time for every l through LDS and accumulates its
// entries (l, i) = (e / o, e % o), e = tid + 256 m, in registers.
// part[k][l][i]; o <= 128.
__global__ void __launch_bounds__(256) lno_matbar_partial(double* __restrict__ part, const double* __restrict__ X,
                                                          const double* __restrict__ Y, long long A, int o,
                                                          long long B) {
  __shared__ double xs[128][LNO_MB_POS + 1];
  __shared__ double ys[128][LNO_MB_POS + 1];
  constexpr int MAXE = 128 * 128 / 256;
  const int tid = threadIdx.x;
  const long long npos = A * B;
  const long long per = (npos + gridDim.x - 1) / gridDim.x;
  const long long p0 = (long long)blockIdx.x * per, p1 = min(npos, p0 + per);
  const int ne = o * o;
  double acc[MAXE];
#pragma unroll
  for (int m = 0; m < MAXE; m++) acc[m] = 0.0;
  for (long long s0 = p0; s0 < p1; s0 += LNO_MB_POS) {
    const int ns = (int)min((long long)LNO_MB_POS, p1 - s0);
    for (int e = tid; e < o * LNO_MB_POS; e += 256) {
      const int l = e / LNO_MB_POS, t = e % LNO_MB_POS;
      double x = 0.0, y = 0.0;
      if (t < ns) {
        const long long pos = s0 + t, a = pos / B, b = pos % B;
        const long long idx = (a * o + l) * B + b;
        x = X[idx];
        y = Y[idx];
      }
      xs[l][t] = x;
      ys[l][t] = y;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MAXE; m++) {
      const int e = tid + 256 * m;
      if (e < ne) {
        const int l = e / o, i = e % o;
        double s = acc[m];
        for (int t = 0; t < LNO_MB_POS; t++) s += xs[l][t] * ys[i][t];
        acc[m] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < MAXE; m++) {
    const int e = tid + 256 * m;
    if (e < ne) part[(long long)blockIdx.x * ne + e] = acc[m];
  }
}

// C[e] (+)= sum_k part[k][e], k in increasing order.
__global__ void lno_matbar_reduce(double* __restrict__ C, const double* __restrict__ part, int nblocks, int ne,
                                  int accumulate) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double s = 0.0;
  for (int k = 0; k < nblocks; k++) s += part[(long long)k * ne + e];
  C[e] = accumulate ? C[e] + s : s;
}

}  // namespace t3

using namespace t3;

static inline unsigned lno_blocks(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536LL * 8) b = 65536LL * 8;
  return (unsigned)(b < 1 ? 1 : b);
}

extern "C" int t3_lno_max_nocc(void) { return 128; }
extern "C" int t3_lno_matbar_blocks(void) { return LNO_MB_BLOCKS; }

extern "C" hipError_t t3_lno_transpose(double* out, const double* in, long long rows, long long cols, int o,
                                       int swap_r, int swap_c, hipStream_t s) {
  const long long tiles = ((rows + LNO_TT - 1) / LNO_TT) * ((cols + LNO_TT - 1) / LNO_TT);
  const unsigned g = (unsigned)std::max(1LL, std::min(tiles, 65536LL * 4));
  hipLaunchKernelGGL(lno_transpose, dim3(g), dim3(256), 0, s, out, in, rows, cols, o, swap_r, swap_c);
  return hipGetLastError();
}

extern "C" hipError_t t3_lno_fock(double* out, const double* in, int o, int v, int to_fock, hipStream_t s) {
  const long long n = o + v;
  hipLaunchKernelGGL(lno_fock, dim3(lno_blocks(to_fock ? n * n : (long long)v * o)), dim3(256), 0, s, out, in, o, v,
                     to_fock);
  return hipGetLastError();
}

extern "C" hipError_t t3_lno_project(double* out, const double* in, const double* mat, long long A, int o,
                                     long long B, int transK, int accumulate, hipStream_t s) {
  const unsigned gy = (unsigned)((o + LNO_PROJ_OUT - 1) / LNO_PROJ_OUT);
  hipLaunchKernelGGL(lno_project, dim3(lno_blocks(A * B), gy), dim3(256), 0, s, out, in, mat, A, o, B, transK,
                     accumulate);
  return hipGetLastError();
}

// C (+)= sum_{a,b} X[a][l][b] Y[a][i][b]; part: LNO_MB_BLOCKS * o * o doubles of scratch
extern "C" hipError_t t3_lno_matbar(double* C, double* part, const double* X, const double* Y, long long A, int o,
                                    long long B, int accumulate, hipStream_t s) {
  if (o > 128) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lno_matbar_partial, dim3(LNO_MB_BLOCKS), dim3(256), 0, s, part, X, Y, A, o, B);
  const int ne = o * o;
  hipLaunchKernelGGL(lno_matbar_reduce, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, C, part, LNO_MB_BLOCKS,
                     ne, accumulate);
  return hipGetLastError();
}

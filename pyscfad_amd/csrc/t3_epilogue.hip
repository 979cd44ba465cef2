// Orbit-cube epilogue of the occupied-triple (T) algorithm.
//
// For one unique occupied triple (i >= j >= k) the W-build GEMMs leave the
// symmetrised cube w0[a][b][c] = w0^{abc}_{ijk} (all a,b,c) in HBM.  The (T)
// summand S = w0 * r3(w0 + v0) / D is invariant under simultaneous
// permutations of (a,i),(b,j),(c,k), so
//     E(T) = 2 sum_{i>=j>=k} f_ijk sum_{a>=b>=c} f_abc sum_pi W[pi] Z[pi] / D
// where, for a canonical point s = (a,b,c), W[pi] = w0[s_pi0][s_pi1][s_pi2] over
// the six position permutations pi, U = W + V, and r3 acting on the occupied
// orderings becomes, at fixed (i,j,k), the virtual-index permutation
//     Z[pi] = 3 U[pi] + sum_{same parity} U - 2 sum_{other parity} U
// (coefficients 4 / 1 / -2 for identity / 3-cycles / transpositions, as in
// the reference's r3, ccsd_t.c:98-119).  This is the same sum as the
// reference's a>=b>=c job loop (ccsd_t.c:6-72, lno_ccsd_t.c:6-29) taken in a
// different order.
//
// A block takes one unique tile triple TA >= TB >= TC (8 virtual indices per
// tile), loads the six permuted 8^3 sub-cubes of w0 into LDS, builds v0 of the
// orbit points from ovov, t2, t1 and fock (FP64), and computes in FP64:
//   forward : the block's share of Q_ijk = sum f_abc sum_pi W Z / D
//   backward: G = s (Z + C W) / D and H = s (C W) / D at every orbit point
//             (the adjoints of w0 and v0, s = 2 f_ijk et_bar, C the r3 matrix,
//             ccsd_t_vjp.c:55-96), and the block's share of the mo_energy
//             adjoint (ccsd_t_vjp.c:98-117).
// Per-block partial sums go to a [triple][cube][32] FP64 table reduced later
// in a fixed order; nothing is accumulated with atomics.
#include "t3_common.h"

namespace t3 {

// Position permutations in the order (012),(021),(102),(120),(201),(210).
// constexpr so that, inside fully unrolled loops, every lookup is a constant.
constexpr int c_perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
constexpr int c_parity[6] = {0, 1, 1, 0, 0, 1};

__device__ __forceinline__ int sel3(int i, int a, int b, int c) { return i == 0 ? a : (i == 1 ? b : c); }




__device__ __forceinline__ int perm_index(int p0, int p1) {
  // index of the permutation (p0, p1, 3-p0-p1) in c_perm order
  return 2 * p0 + (p1 > 3 - p0 - p1 ? 1 : 0);
}

__device__ __forceinline__ void decode_cube(int m, int& A, int& B, int& C) {
  int a = 0;
  while ((a + 1) * (a + 2) * (a + 3) / 6 <= m) a++;
  m -= a * (a + 1) * (a + 2) / 6;
  int b = 0;
  while ((b + 1) * (b + 2) / 2 <= m) b++;
  m -= b * (b + 1) / 2;
  A = a; B = b; C = m;
}

template <class TP, bool BW>
__global__ void __launch_bounds__(256)
orbit_epilogue(EpiArgs args, const EpiTriple* __restrict__ triples) {
  constexpr int T = EPI_T, T3 = T * T * T;
  const EpiTriple tr = triples[blockIdx.y];
  const int cube = blockIdx.x;
  int TA, TB, TC;
  decode_cube(cube, TA, TB, TC);
  const int V = args.V, v = args.nvir, o = args.nocc;
  const int tid = threadIdx.x;

  __shared__ TP sc[6][T3];                 // w0 sub-cubes, later G
  __shared__ TP sv[6][T3];                 // v0 sub-cubes, later H
  __shared__ double rbuf[2][T3];           // per canonical point: Q and R contributions
  __shared__ double red[256];

  // --- load the six permuted sub-cubes of w0 and v0 ---
  const TP* w0 = reinterpret_cast<const TP*>(tr.w0);
  const TP* v0 = reinterpret_cast<const TP*>(tr.v0);
  for (int e = tid; e < 6 * T3; e += 256) {
    const int p = e / T3, r = e % T3;
    const int x0 = r / (T * T), x1 = (r / T) % T, x2 = r % T;
    int q0, q1, q2;
    switch (p) {
      case 0: q0 = TA; q1 = TB; q2 = TC; break;
      case 1: q0 = TA; q1 = TC; q2 = TB; break;
      case 2: q0 = TB; q1 = TA; q2 = TC; break;
      case 3: q0 = TB; q1 = TC; q2 = TA; break;
      case 4: q0 = TC; q1 = TA; q2 = TB; break;
      default: q0 = TC; q1 = TB; q2 = TA; break;
    }
    const long long g = (((long long)q0 * T + x0) * V + (q1 * T + x1)) * V + (q2 * T + x2);
    sc[p][r] = w0[g];
    sv[p][r] = v0[g];
  }
  __syncthreads();

  const double eijk = args.eps_occ[tr.i] + args.eps_occ[tr.j] + args.eps_occ[tr.k];
  const double s_bw = 2.0 * tr.fo * args.et_bar;
  const bool eqAB = (TA == TB), eqBC = (TB == TC);

  double gval[2][6], hval[2][6];
  int npt = 0;
  for (int pt = tid; pt < T3; pt += 256, npt++) {
    const int al[3] = {pt / (T * T), (pt / T) % T, pt % T};
    rbuf[0][pt] = 0.0;
    rbuf[1][pt] = 0.0;
    if ((eqAB && al[0] < al[1]) || (eqBC && al[1] < al[2])) continue;
    const int s0 = TA * T + al[0], s1 = TB * T + al[1], s2 = TC * T + al[2];
    const double fp = (s0 == s2) ? (1.0 / 6.0) : ((s0 == s1 || s1 == s2) ? 0.5 : 1.0);
    const double D = eijk - args.eps_vir[s0] - args.eps_vir[s1] - args.eps_vir[s2];
    double W[6], U[6];
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const int x0 = al[c_perm[p][0]], x1 = al[c_perm[p][1]], x2 = al[c_perm[p][2]];
      W[p] = (double)sc[p][(x0 * T + x1) * T + x2];
      U[p] = W[p] + (double)sv[p][(x0 * T + x1) * T + x2];
    }
    double sU[2] = {0.0, 0.0}, sW[2] = {0.0, 0.0};
#pragma unroll
    for (int p = 0; p < 6; p++) { sU[c_parity[p]] += U[p]; sW[c_parity[p]] += W[p]; }
    double wz = 0.0;
    const double invD = 1.0 / D;
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const int par = c_parity[p];
      const double Z = 3.0 * U[p] + sU[par] - 2.0 * sU[par ^ 1];
      const double CW = 3.0 * W[p] + sW[par] - 2.0 * sW[par ^ 1];
      wz += W[p] * Z;
      if constexpr (BW) {
        gval[npt][p] = s_bw * (Z + CW) * invD;
        hval[npt][p] = s_bw * CW * invD;
      }
    }
    rbuf[0][pt] = fp * wz * invD;                    // Q contribution
    rbuf[1][pt] = s_bw * fp * wz * invD * invD;      // mo_energy adjoint weight
  }
  __syncthreads();

  double* part = tr.part + (long long)cube * EPI_PSTRIDE;
  // Q: fixed-order reduction (thread t sums points t, t+256; then a tree)
  red[tid] = rbuf[0][tid] + rbuf[0][tid + 256];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) part[0] = red[0];
  if constexpr (!BW) return;

  // mo_energy adjoint partials
  __syncthreads();
  red[tid] = rbuf[1][tid] + rbuf[1][tid + 256];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) part[1] = red[0];
  if (tid < 3 * T) {
    const int pos = tid / T, x = tid % T;
    double s = 0.0;
    for (int u = 0; u < T; u++)
      for (int w2 = 0; w2 < T; w2++) {
        int idx;
        if (pos == 0) idx = (x * T + u) * T + w2;
        else if (pos == 1) idx = (u * T + x) * T + w2;
        else idx = (u * T + w2) * T + x;
        s += rbuf[1][idx];
      }
    part[2 + tid] = s;
  }

  // G and H into LDS, over the w0 and v0 sub-cubes
  npt = 0;
  for (int pt = tid; pt < T3; pt += 256, npt++) {
    const int al[3] = {pt / (T * T), (pt / T) % T, pt % T};
    if ((eqAB && al[0] < al[1]) || (eqBC && al[1] < al[2])) continue;
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const int x0 = al[c_perm[p][0]], x1 = al[c_perm[p][1]], x2 = al[c_perm[p][2]];
      const int r = (x0 * T + x1) * T + x2;
      sc[p][r] = (TP)gval[npt][p];
      sv[p][r] = (TP)hval[npt][p];
    }
  }
  __syncthreads();

  // write-out: an entry of sub-cube p is written only if it came from a
  // canonical point and no lower-numbered sub-cube covering the same global
  // block also holds it (one writer per element).
  TP* G = reinterpret_cast<TP*>(tr.G);
  TP* H = reinterpret_cast<TP*>(tr.H);
  const int tt[3] = {TA, TB, TC};
#pragma unroll
  for (int p = 0; p < 6; p++) {
    const long long b0 = (long long)tt[c_perm[p][0]] * T, b1 = (long long)tt[c_perm[p][1]] * T,
                    b2 = (long long)tt[c_perm[p][2]] * T;
    for (int r = tid; r < T3; r += 256) {
      const int x[3] = {r / (T * T), (r / T) % T, r % T};
      bool write = false, owned = false;
#pragma unroll
      for (int q = 0; q <= p; q++) {
        const bool same = tt[c_perm[q][0]] == tt[c_perm[p][0]] && tt[c_perm[q][1]] == tt[c_perm[p][1]] &&
                          tt[c_perm[q][2]] == tt[c_perm[p][2]];
        // local canonical coordinates of this entry as seen from sub-cube q
        int al[3];
        al[c_perm[q][0]] = x[0];
        al[c_perm[q][1]] = x[1];
        al[c_perm[q][2]] = x[2];
        const bool canon = !((eqAB && al[0] < al[1]) || (eqBC && al[1] < al[2]));
        if (same && canon && !owned) {
          if (q == p) write = true;
          else owned = true;   // a lower sub-cube owns it
        }
      }
      if (!write) continue;
      const long long off = ((b0 + x[0]) * V + (b1 + x[1])) * V + (b2 + x[2]);
      G[off] = sc[p][r];
      H[off] = sv[p][r];
    }
  }
}

__device__ __forceinline__ int cube_id(int A, int B, int C) {
  return A * (A + 1) * (A + 2) / 6 + B * (B + 1) / 2 + C;
}

__global__ void __launch_bounds__(256) epi_reduce(EpiReduceArgs a) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  if (blockIdx.x == 0) {
    // energy
    double s = 0.0;
    for (int t = 0; t < a.ntriple; t++) {
      const EpiTriple& tr = a.triples[t];
      double q = 0.0;
      for (int c = tid; c < a.ncube; c += 256) q += tr.part[(long long)c * EPI_PSTRIDE];
      s += 2.0 * tr.fo * q;
    }
    red[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    if (tid == 0) *a.energy += red[0];
    if (!a.backward) return;
    __syncthreads();
    // occupied adjoint
    for (int t = 0; t < a.ntriple; t++) {
      const EpiTriple& tr = a.triples[t];
      double r = 0.0;
      for (int c = tid; c < a.ncube; c += 256) r += tr.part[(long long)c * EPI_PSTRIDE + 1];
      red[tid] = r;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
      }
      if (tid == 0) {
        a.eps_bar[tr.i] -= red[0];
        a.eps_bar[tr.j] -= red[0];
        a.eps_bar[tr.k] -= red[0];
      }
      __syncthreads();
    }
    return;
  }
  if (!a.backward) return;
  // virtual adjoint: one thread per padded virtual index
  const int x = (blockIdx.x - 1) * 256 + tid;
  if (x >= a.V) return;
  const int TX = x / EPI_T, lx = x % EPI_T;
  double s = 0.0;
  for (int t = 0; t < a.ntriple; t++) {
    const double* part = a.triples[t].part;
    for (int X = 0; X < a.ntile; X++)
      for (int Y = 0; Y <= X; Y++) {
        int u0 = TX, u1 = X, u2 = Y;   // sort descending
        if (u1 > u0) { int tmp = u0; u0 = u1; u1 = tmp; }
        if (u2 > u1) { int tmp = u1; u1 = u2; u2 = tmp; }
        if (u1 > u0) { int tmp = u0; u0 = u1; u1 = tmp; }
        const double* pc = part + (long long)cube_id(u0, u1, u2) * EPI_PSTRIDE;
        if (u0 == TX) s += pc[2 + lx];
        if (u1 == TX) s += pc[2 + EPI_T + lx];
        if (u2 == TX) s += pc[2 + 2 * EPI_T + lx];
      }
  }
  a.eps_bar[a.nocc + x] += s;
}

}  // namespace t3

using namespace t3;

extern "C" hipError_t t3_launch_epilogue(int tp_is_f64, const EpiArgs* args, const EpiTriple* triples,
                                         int ntriple, hipStream_t stream) {
  if (ntriple <= 0) return hipSuccess;
  dim3 grid(args->ncube, ntriple), block(256);
  if (args->backward) {
    if (tp_is_f64) hipLaunchKernelGGL((orbit_epilogue<double, true>), grid, block, 0, stream, *args, triples);
    else hipLaunchKernelGGL((orbit_epilogue<float, true>), grid, block, 0, stream, *args, triples);
  } else {
    if (tp_is_f64) hipLaunchKernelGGL((orbit_epilogue<double, false>), grid, block, 0, stream, *args, triples);
    else hipLaunchKernelGGL((orbit_epilogue<float, false>), grid, block, 0, stream, *args, triples);
  }
  return hipGetLastError();
}

extern "C" hipError_t t3_launch_epi_reduce(const EpiReduceArgs* a, hipStream_t stream) {
  const int nb = a->backward ? 1 + (a->V + 255) / 256 : 1;
  hipLaunchKernelGGL(epi_reduce, dim3(nb), dim3(256), 0, stream, *a);
  return hipGetLastError();
}

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
// The V term (ccsd_t.c:149-150 summed over the six orderings) is built here,
// in FP64, from three pair matrices per triple:
//     v0[x] = 1/2 sum_d ( CM_d[x_e1][x_e2] t1[t_d, x_d] + CN_d[x_e1][x_e2] fvo[x_d, t_d] )
// with e1 < e2 the positions other than d and
//     CM_d[x][y] = ovov[t_e1,x,t_e2,y] + ovov[t_e2,y,t_e1,x],
//     CN_d[x][y] = t2[t_e2,t_e1,y,x] + t2[t_e1,t_e2,x,y]      (vcm_build).
//
// A block takes one unique tile triple TA >= TB >= TC (8 virtual indices per
// tile) and computes in FP64:
//   forward : the block's share of Q_ijk = sum f_abc sum_pi W Z / D
//   backward: G = s (Z + C W) / D and H = s (C W) / D at every orbit point
//             (the adjoints of w0 and v0, s = 2 f_ijk et_bar, C the r3 matrix,
//             ccsd_t_vjp.c:55-96), and the block's share of the mo_energy
//             adjoint (ccsd_t_vjp.c:98-117).
// Per-block partial sums go to a [triple][cube][32] FP64 table reduced later
// in a fixed order; nothing is accumulated with atomics.
//
// LNO form (the occupied projector of lno_ccsd_t.c:6-29): the triple also
// carries S = (P1 + P2 + P3) w0, w0 with each occupied slot in turn projected
// by mat (built by the GEMMs from projected operands).  Its energy is
//     E = 2/3 sum_{i>=j>=k} f_ijk sum_{abc} S Z / D,
// which for mat = I (S = 3 w0) is E(T).  The summand pairs S with Z at the
// same orbit points, so the block computes Q_ijk = 1/3 sum f_abc sum_pi S Z / D
// and, backward, with s = 2/3 f_ijk et_bar,
//     GS = s Z / D (the adjoint of S),  G = H = s (C S) / D
// (the adjoints of w0 and v0 through Z only); H is not written, its
// consumers read G.
#include "t3_common.h"

namespace t3 {

// Position permutations in the order (012),(021),(102),(120),(201),(210).
// constexpr so that, inside fully unrolled loops, every lookup is a constant.
constexpr int c_perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
constexpr int c_parity[6] = {0, 1, 1, 0, 0, 1};
// the two positions other than d, in increasing order
constexpr int c_oth[3][2] = {{1, 2}, {0, 2}, {0, 1}};

__device__ __forceinline__ int sel3(int i, int a, int b, int c) { return i == 0 ? a : (i == 1 ? b : c); }

__host__ __device__ constexpr int perm_index(int p0, int p1) {
  // index of the permutation (p0, p1, 3-p0-p1) in c_perm order
  return 2 * p0 + (p1 > 3 - p0 - p1 ? 1 : 0);
}


// Sum over the 64 lanes of a wave with a fixed butterfly (every lane gets the
// same, schedule-independent result).
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}
// Sum over the lanes that differ only in the bits of `mask` (fixed order).
template <int MASK>
__device__ __forceinline__ double lane_group_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    if (MASK & off) x += __shfl_xor(x, off, 64);
  return x;
}

// LDS position of entry (y0, y1, y2) of an 8x8x8 sub-cube: each row y0
// XOR-permutes its 8-entry groups and the entries within them.  A wave reads
// the sub-cubes transposed, its lanes spanning two of the three indices in any
// of the six orders; with this map every such pattern hits 64 distinct banks
// (the plain [y0][y1][y2] order puts 8 lanes on a bank whenever the lanes vary
// y0).  A VW-aligned run of y2 stays one aligned vector, its elements permuted
// by XOR with y0 % VW (xor_perm).
__device__ __forceinline__ int sub_pos(int y0, int y1, int y2) {
  return (y0 * EPI_T + (y1 ^ y0)) * EPI_T + (y2 ^ y0);
}
template <class VT, int VW>
__device__ __forceinline__ VT xor_perm(VT v, int k) {
  if constexpr (VW == 4) {
    const VT a = (k & 1) ? v.yxwz : v;
    return (k & 2) ? a.zwxy : a;
  } else {
    static_assert(VW == 2, "four floats or two doubles per vector");
    return (k & 1) ? v.yx : v;
  }
}

// 512 threads per block, one point of the 8^3 tile per thread: the kernel is
// bound by the latency of its operand loads, and with the LDS of five blocks
// per CU this keeps 32 waves per CU in flight instead of 20.
constexpr int EPI_THREADS = 512, EPI_WAVES = EPI_THREADS / 64;

template <class TP, bool BW, bool LNO>
__global__ void __launch_bounds__(EPI_THREADS)
orbit_epilogue(EpiArgs args, const EpiTriple* __restrict__ triples) {
  constexpr int T = EPI_T, T3 = T * T * T, NT = EPI_THREADS, NW = EPI_WAVES;
  static_assert(T == 8 && T3 == NT, "one point per thread, wave = first index, lane bits = the other two");
  constexpr int NCM = 3 * 6 * 2 * T * T;                   // CM/CN blocks
  constexpr int NH = (6 * T3 * (int)sizeof(TP) + 7) / 8;   // H sub-cubes, in doubles
  constexpr int NU = NCM > NH ? NCM : NH;
  constexpr int VW = 16 / (int)sizeof(TP);                 // elements per 16-byte load
  typedef TP tpv __attribute__((ext_vector_type(VW)));
  typedef double dv2 __attribute__((ext_vector_type(2)));
  const EpiTriple tr = triples[blockIdx.y];
  // Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
  // range of cube ids, so that neighbouring tiles (which share 128-byte lines
  // of w0, G and H along the last index) meet in the same L2.
  int cube = blockIdx.x;
  {
    const int nb = gridDim.x, q = nb / 8;
    if (cube < 8 * q) cube = (cube % 8) * q + cube / 8;
  }
  // tile triple of the cube from the host-built table (a scalar load; the
  // closed-form decode is a serial scalar search)
  const int tcode = args.cube_tab[cube];
  const int TA = tcode & 1023, TB = (tcode >> 10) & 1023, TC = tcode >> 20;
  const int V = args.V, v = args.nvir;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  __shared__ __attribute__((aligned(16))) TP sc[6][T3];   // w0 sub-cubes, later G
  // LNO form: the S sub-cubes
  __shared__ __attribute__((aligned(16))) TP ss[LNO ? 6 : 1][LNO ? T3 : 1];
  __shared__ __attribute__((aligned(16))) double cmh[NU]; // CM/CN tile-pair blocks [d][pair][M/N][8x8], later H
  __shared__ double vec[2][3][3][T];                      // [t1/fvo][d][tile pos][8]
  __shared__ double wred[NW][20];                         // per-wave partial sums

  // --- the six permuted sub-cubes of w0, 16-byte loads along the rows ---
  const TP* w0 = reinterpret_cast<const TP*>(tr.w0);
#pragma unroll
  for (int it = 0; it < (6 * T3 / VW + NT - 1) / NT; it++) {
    const int e = tid + NT * it;
    if (e >= 6 * T3 / VW) break;
    const int p = e / (T3 / VW), r = (e % (T3 / VW)) * VW;
    const int x0 = r / (T * T), x1 = (r / T) % T, x2 = r % T;
    // tiles of sub-cube p (c_perm order), as selects: p is a lane value
    const int q0 = p < 2 ? TA : (p < 4 ? TB : TC);
    const int q1 = (p == 0 || p == 5) ? TB : ((p == 1 || p == 3) ? TC : TA);
    const int q2 = (p == 1 || p == 4) ? TB : ((p == 3 || p == 5) ? TA : TC);
    const long long off = (((long long)q0 * T + x0) * V + (q1 * T + x1)) * V + (q2 * T + x2);
    const tpv x = *reinterpret_cast<const tpv*>(&w0[off]);
    *reinterpret_cast<tpv*>(&sc[p][sub_pos(x0, x1, x2 ^ (x0 & (VW - 1)))]) = xor_perm<tpv, VW>(x, x0 & (VW - 1));
    if constexpr (LNO) {
      const tpv y = *reinterpret_cast<const tpv*>(&reinterpret_cast<const TP*>(tr.s)[off]);
      *reinterpret_cast<tpv*>(&ss[p][sub_pos(x0, x1, x2 ^ (x0 & (VW - 1)))]) = xor_perm<tpv, VW>(y, x0 & (VW - 1));
    }
  }
  // --- CM_d / CN_d on the ordered tile pairs: block [d][pair mu][which][8x8] ---
  {
    const long long VV = (long long)V * V;
#pragma unroll
    for (int it = 0; it < (NCM / 2 + NT - 1) / NT; it++) {
      const int e2 = tid + NT * it;
      if (e2 >= NCM / 2) break;
      const int e = 2 * e2;
      const int xy = e % (T * T);
      const int which = (e / (T * T)) % 2;
      const int mu = (e / (2 * T * T)) % 6;
      const int d = e / (12 * T * T);
      // first two tiles of the ordered pair mu (c_perm order), as selects
      const int u0 = mu < 2 ? TA : (mu < 4 ? TB : TC);
      const int u1 = (mu == 0 || mu == 5) ? TB : ((mu == 1 || mu == 3) ? TC : TA);
      const int x = u0 * T + xy / T;
      const int y = u1 * T + xy % T;
      *reinterpret_cast<dv2*>(&cmh[e]) = *reinterpret_cast<const dv2*>(&tr.cm[(2 * d + which) * VV + (long long)x * V + y]);
    }
  }
  static_assert(2 * 3 * 3 * T <= NT, "one round");
  if (tid < 2 * 3 * 3 * T) {
    const int e = tid;
    const int which = e / (9 * T), r = e % (9 * T);
    const int d = r / (3 * T), tp = (r / T) % 3, z = sel3(tp, TA, TB, TC) * T + r % T;
    const int td = sel3(d, tr.i, tr.j, tr.k);
    double val = 0.0;
    if (which == 0) val = z < v ? args.t1[(long long)td * v + z] : 0.0;
    else val = args.fvoP[(long long)td * V + z];
    vec[which][d][tp][r % T] = val;
  }
  __syncthreads();

  const double eijk = args.eps_occ[tr.i] + args.eps_occ[tr.j] + args.eps_occ[tr.k];
  const double s_bw = (LNO ? 2.0 / 3.0 : 2.0) * tr.fo * args.et_bar;
  const bool eqAB = (TA == TB), eqBC = (TB == TC);

  // one point per thread: pt = tid, i.e. al0 = wave, al1 = lane bits 3-5, al2 = lane bits 0-2
  constexpr int NPT = T3 / NT;
  double gval[NPT][6], hval[NPT][6], qpt[NPT], rpt[NPT];
#pragma unroll
  for (int h = 0; h < NPT; h++) {
    const int pt = tid + NT * h;
    const int al[3] = {pt / (T * T), (pt / T) % T, pt % T};
    qpt[h] = 0.0;
    rpt[h] = 0.0;
#pragma unroll
    for (int p = 0; p < 6; p++) { gval[h][p] = 0.0; hval[h][p] = 0.0; }
    if ((eqAB && al[0] < al[1]) || (eqBC && al[1] < al[2])) continue;
    const int s0 = TA * T + al[0], s1 = TB * T + al[1], s2 = TC * T + al[2];
    const double fp = (s0 == s2) ? (1.0 / 6.0) : ((s0 == s1 || s1 == s2) ? 0.5 : 1.0);
    const double D = eijk - args.eps_vir[s0] - args.eps_vir[s1] - args.eps_vir[s2];
    double W[6], U[6], Sv[6];
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const int x0 = al[c_perm[p][0]], x1 = al[c_perm[p][1]], x2 = al[c_perm[p][2]];
      W[p] = (double)sc[p][sub_pos(x0, x1, x2)];
      if constexpr (LNO) Sv[p] = (double)ss[p][sub_pos(x0, x1, x2)];
      // v0 at the orbit point pi(s): position m of the point holds s_{pi_m}
      double vv = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int m1 = c_perm[p][c_oth[d][0]], m2 = c_perm[p][c_oth[d][1]], md = c_perm[p][d];
        const int blk = ((d * 6 + perm_index(m1, m2)) * 2) * T * T + al[m1] * T + al[m2];
        vv += cmh[blk] * vec[0][d][md][al[md]] + cmh[blk + T * T] * vec[1][d][md][al[md]];
      }
      U[p] = W[p] + 0.5 * vv;
    }
    // X: the first factor of the summand (w0, or S in the LNO form)
    double sU[2] = {0.0, 0.0}, sX[2] = {0.0, 0.0};
#pragma unroll
    for (int p = 0; p < 6; p++) {
      sU[c_parity[p]] += U[p];
      sX[c_parity[p]] += LNO ? Sv[p] : W[p];
    }
    double wz = 0.0;
    const double invD = 1.0 / D;
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const int par = c_parity[p];
      const double X = LNO ? Sv[p] : W[p];
      const double Z = 3.0 * U[p] + sU[par] - 2.0 * sU[par ^ 1];
      const double CX = 3.0 * X + sX[par] - 2.0 * sX[par ^ 1];
      wz += X * Z;
      if constexpr (BW) {
        if constexpr (LNO) {
          gval[h][p] = s_bw * CX * invD;   // G (= H)
          hval[h][p] = s_bw * Z * invD;    // GS
        } else {
          gval[h][p] = s_bw * (Z + CX) * invD;
          hval[h][p] = s_bw * CX * invD;
        }
      }
    }
    qpt[h] = (LNO ? 1.0 / 3.0 : 1.0) * fp * wz * invD;   // Q contribution
    rpt[h] = s_bw * fp * wz * invD * invD;      // mo_energy adjoint weight
  }

  // wave partials, fixed shuffle order:
  //   [0] Q, [1] R (= the vir pos-0 partial of x = wave), [4,12) pos 1, [12,20) pos 2
  static_assert(NPT == 1 && NW == T, "one point per thread, one wave per value of the first index");
  {
    const double q = wave_sum(qpt[0]);
    if (lane == 0) wred[wv][0] = q;
  }
  if constexpr (BW) {
    const double ra = wave_sum(rpt[0]);
    const double r1 = lane_group_sum<7>(rpt[0]);    // same al1 = lane bits 3-5
    const double r2 = lane_group_sum<56>(rpt[0]);   // same al2 = lane bits 0-2
    if (lane == 0) wred[wv][1] = ra;
    if ((lane & 7) == 0) wred[wv][4 + (lane >> 3)] = r1;
    if (lane < 8) wred[wv][12 + lane] = r2;
  }
  __syncthreads();   // every read of sc / cmh is done; wave partials are in place

  // sums over the waves in wave order
  auto wsum = [&](int k) {
    double x = wred[0][k];
#pragma unroll
    for (int w = 1; w < NW; w++) x += wred[w][k];
    return x;
  };
  double* part = tr.part + (long long)cube * EPI_PSTRIDE;
  if (tid == 0) part[0] = wsum(0);
  if constexpr (!BW) return;
  if (tid == 1) part[1] = wsum(1);
  if (tid >= 2 && tid < 2 + T) {
    const int x = tid - 2;   // pos 0: x is the wave
    part[2 + x] = wred[x][1];
  } else if (tid >= 2 + T && tid < 2 + 3 * T) {
    const int k = tid - 2 - T;   // 0..15: pos 1 then pos 2
    part[2 + T + k] = wsum(4 + k);
  }

  // G and H into LDS, over the w0 sub-cubes and the CM/CN area
  TP* hl = reinterpret_cast<TP*>(cmh);
#pragma unroll
  for (int h = 0; h < NPT; h++) {
    const int pt = tid + NT * h;
    const int al[3] = {pt / (T * T), (pt / T) % T, pt % T};
    if ((eqAB && al[0] < al[1]) || (eqBC && al[1] < al[2])) continue;
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const int x0 = al[c_perm[p][0]], x1 = al[c_perm[p][1]], x2 = al[c_perm[p][2]];
      const int r = sub_pos(x0, x1, x2);
      sc[p][r] = (TP)gval[h][p];
      hl[p * T3 + r] = (TP)hval[h][p];
    }
  }
  __syncthreads();

  // write-out: an entry of sub-cube p is written only if it came from a
  // canonical point and no lower-numbered sub-cube covering the same global
  // block also holds it (one writer per element).
  TP* G = reinterpret_cast<TP*>(tr.G);
  TP* H = reinterpret_cast<TP*>(LNO ? tr.GS : tr.H);   // the second LDS set: H, or GS in the LNO form
  const int tt[3] = {TA, TB, TC};
  const bool distinct = !eqAB && !eqBC;
#pragma unroll
  for (int p = 0; p < 6; p++) {
    const long long b0 = (long long)tt[c_perm[p][0]] * T, b1 = (long long)tt[c_perm[p][1]] * T,
                    b2 = (long long)tt[c_perm[p][2]] * T;
    if (distinct) {   // every sub-cube is a different global block and all its points are canonical
      static_assert(T3 / VW <= NT, "one round");
      if (tid < T3 / VW) {
        const int e = tid;
        const int r = e * VW;
        const int x0 = r / (T * T), x1 = (r / T) % T, x2 = r % T;
        const long long off = ((b0 + x0) * V + (b1 + x1)) * V + (b2 + x2);
        const int k = x0 & (VW - 1), rs = sub_pos(x0, x1, x2 ^ k);
        *reinterpret_cast<tpv*>(&G[off]) = xor_perm<tpv, VW>(*reinterpret_cast<const tpv*>(&sc[p][rs]), k);
        *reinterpret_cast<tpv*>(&H[off]) = xor_perm<tpv, VW>(*reinterpret_cast<const tpv*>(&hl[p * T3 + rs]), k);
      }
      continue;
    }
    static_assert(T3 % NT == 0, "whole rounds");
#pragma unroll
    for (int rr = 0; rr < T3 / NT; rr++) {
      const int r = tid + NT * rr;
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
      const int rs = sub_pos(x[0], x[1], x[2]);
      G[off] = sc[p][rs];
      H[off] = hl[p * T3 + rs];
    }
  }
}

__device__ __forceinline__ int cube_id(int A, int B, int C) {
  return A * (A + 1) * (A + 2) / 6 + B * (B + 1) / 2 + C;
}

// Per-batch reduction of the epilogue partials, every sum in a fixed order.
// Block `ntile` : E += sum_t 2 f_t sum_cube Q;  eps_bar[occ of t] -= sum_cube Rsum.
// Block X < ntile: eps_bar[o + 8X + l] += the vir partials of every cube that
//                  contains tile X (backward only).
__global__ void __launch_bounds__(256) epi_reduce(EpiReduceArgs a) {
  __shared__ double wpart[4][EPI_T];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // block sum in a fixed order: a butterfly within each wave, then the four
  // wave sums in order
  auto block_sum = [&](double x) -> double {
    x = wave_sum(x);
    if (lane == 0) wpart[wv][0] = x;
    __syncthreads();
    const double r = ((wpart[0][0] + wpart[1][0]) + wpart[2][0]) + wpart[3][0];
    __syncthreads();
    return r;
  };
  if ((int)blockIdx.x == a.ntile) {
    double e = 0.0;
    for (int t = 0; t < a.ntriple; t++) {
      const EpiTriple& tr = a.triples[t];
      double q = 0.0;
      // unrolled: eight loads in flight, the sum keeps its order
#pragma unroll 8
      for (int c = tid; c < a.ncube; c += 256) q += tr.part[(long long)c * EPI_PSTRIDE];
      e += 2.0 * tr.fo * block_sum(q);
    }
    if (tid == 0) *a.energy += e;
    if (!a.backward) return;
    for (int t = 0; t < a.ntriple; t++) {
      const EpiTriple& tr = a.triples[t];
      double r = 0.0;
#pragma unroll 8
      for (int c = tid; c < a.ncube; c += 256) r += tr.part[(long long)c * EPI_PSTRIDE + 1];
      r = block_sum(r);
      if (tid == 0) {
        a.eps_bar[tr.i] -= r;
        a.eps_bar[tr.j] -= r;
        a.eps_bar[tr.k] -= r;
      }
    }
    return;
  }
  if (!a.backward) return;
  const int TX = blockIdx.x;
  const int npair = a.ntile * (a.ntile + 1) / 2;
  double total[EPI_T];
  for (int l = 0; l < EPI_T; l++) total[l] = 0.0;
  for (int t = 0; t < a.ntriple; t++) {
    const double* part = a.triples[t].part;
    double acc[EPI_T];
    for (int l = 0; l < EPI_T; l++) acc[l] = 0.0;
#pragma unroll 2
    for (int pr = tid; pr < npair; pr += 256) {
      // pair pr = (X, Y), Y <= X, in the order X(X+1)/2 + Y
      int X = (int)((sqrtf(8.0f * (float)pr + 1.0f) - 1.0f) * 0.5f);
      while (X > 0 && X * (X + 1) / 2 > pr) X--;
      while ((X + 1) * (X + 2) / 2 <= pr) X++;
      const int Y = pr - X * (X + 1) / 2;
      int u0 = TX, u1 = X, u2 = Y;   // sort descending
      if (u1 > u0) { int tmp = u0; u0 = u1; u1 = tmp; }
      if (u2 > u1) { int tmp = u1; u1 = u2; u2 = tmp; }
      if (u1 > u0) { int tmp = u0; u0 = u1; u1 = tmp; }
      const double* pc = part + (long long)cube_id(u0, u1, u2) * EPI_PSTRIDE;
      for (int l = 0; l < EPI_T; l++) {
        if (u0 == TX) acc[l] += pc[2 + l];
        if (u1 == TX) acc[l] += pc[2 + EPI_T + l];
        if (u2 == TX) acc[l] += pc[2 + 2 * EPI_T + l];
      }
    }
#pragma unroll
    for (int l = 0; l < EPI_T; l++) acc[l] = wave_sum(acc[l]);
    if (lane < EPI_T) {
      double x = acc[0];
#pragma unroll
      for (int l = 1; l < EPI_T; l++) x = lane == l ? acc[l] : x;
      wpart[wv][lane] = x;
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < EPI_T; l++) total[l] += ((wpart[0][l] + wpart[1][l]) + wpart[2][l]) + wpart[3][l];
    __syncthreads();
  }
  if (tid < EPI_T) {
    double x = total[0];
#pragma unroll
    for (int l = 1; l < EPI_T; l++) x = tid == l ? total[l] : x;
    a.eps_bar[a.nocc + TX * EPI_T + tid] += x;
  }
}

}  // namespace t3

using namespace t3;

extern "C" hipError_t t3_launch_epilogue(int tp_is_f64, const EpiArgs* args, const EpiTriple* triples,
                                         int ntriple, hipStream_t stream) {
  if (ntriple <= 0) return hipSuccess;
  dim3 grid(args->ncube, ntriple), block(EPI_THREADS);
#define T3_EPI(TP, BW, LNO) hipLaunchKernelGGL((orbit_epilogue<TP, BW, LNO>), grid, block, 0, stream, *args, triples)
  if (args->lno) {
    if (args->backward) {
      if (tp_is_f64) T3_EPI(double, true, true);
      else T3_EPI(float, true, true);
    } else {
      if (tp_is_f64) T3_EPI(double, false, true);
      else T3_EPI(float, false, true);
    }
  } else if (args->backward) {
    if (tp_is_f64) T3_EPI(double, true, false);
    else T3_EPI(float, true, false);
  } else {
    if (tp_is_f64) T3_EPI(double, false, false);
    else T3_EPI(float, false, false);
  }
#undef T3_EPI
  return hipGetLastError();
}

extern "C" hipError_t t3_launch_epi_reduce(const EpiReduceArgs* a, hipStream_t stream) {
  const int nb = a->ntile + 1;
  hipLaunchKernelGGL(epi_reduce, dim3(nb), dim3(256), 0, stream, *a);
  return hipGetLastError();
}

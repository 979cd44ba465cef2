// Layout preparation, V-term adjoints and cotangent finalisation.
//
// Device layouts used by the contractions (V = nvir padded to a multiple of
// 32, Op = nocc padded to a multiple of the K-chunk, all padding zero):
//   ovvvT [o ][V ][V][V]  = ovvv[p,x,f,y]           (OVVV[p][x][y][f])
//   t2T   [Op][Op][V][V]  = t2[r,q,z,f]
//   t2H   [Op][Op][V][V]  = t2[r,m,z,y] at [r][m][y][z]
//   ovooT [o ][o ][V][Op] = ovoo[p,x,m,q] at [p][q][x][m]
// The reference builds the analogous host arrays at pyscfad/cc/ccsd_t.py:88-106.
#include "t3_common.h"

namespace t3 {

__device__ __forceinline__ long long tri_index(long long f, long long y) {
  return f >= y ? f * (f + 1) / 2 + y : y * (y + 1) / 2 + f;
}

// ---------------- tiled re-layouts (32 x 32 tiles through LDS) ----------------
// A block of 256 threads takes one 32 x 32 tile of the (row, col) plane of a
// run of TR_SLABS consecutive slabs; thread (tx = t % 32, ty = t / 32) moves
// the elements (ty + 8 q, tx), q = 0..3, so every global access is 32
// consecutive elements of one row.  Tiles whose source is the transpose of
// their destination go through LDS (row pitch 33: conflict-free both ways).
// Block b takes tile b % ntiles of slab run b / ntiles.
constexpr int TR_T = 32, TR_SLABS = 8;

// ovvvT[p][x][y][f] = ovvv[p,x,f,y] for one slab s = p * V + x (zero past v):
// packed input pair (f, y) at tri(max, min) of the slab's v(v+1)/2 pairs.
template <class TP>
__global__ void __launch_bounds__(256) prep_ovvv(TP* __restrict__ out, const double* __restrict__ ovvv, int o,
                                                 int v, int V, int packed) {
  __shared__ double tile[TR_T][TR_T + 1];
  const int nt = V / TR_T, ntiles = nt * nt;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int tile_id = (int)(blockIdx.x % (unsigned)ntiles);
  const long long run = blockIdx.x / (unsigned)ntiles;
  const int Y0 = (tile_id / nt) * TR_T, F0 = (tile_id % nt) * TR_T;
  const long long npair = (long long)v * (v + 1) / 2;
  const long long nslab = (long long)o * V;
  for (int sl = 0; sl < TR_SLABS; sl++) {
    const long long slab = run * TR_SLABS + sl;
    if (slab >= nslab) break;
    const int x = (int)(slab % V);
    const long long p = slab / V;
    TP* dst = out + slab * V * V;
    if (x >= v) {
      for (int q = 0; q < 4; q++) dst[(long long)(Y0 + ty + 8 * q) * V + F0 + tx] = (TP)0.0;
      continue;
    }
    const double* src = ovvv + (p * v + x) * (packed ? npair : (long long)v * v);
    if (packed && F0 < Y0) {
      // below the diagonal (f < y): pair (y, f), consecutive in f
      for (int q = 0; q < 4; q++) {
        const int y = Y0 + ty + 8 * q, f = F0 + tx;
        const double val = (y < v && f < v) ? src[(long long)y * (y + 1) / 2 + f] : 0.0;
        dst[(long long)y * V + f] = (TP)val;
      }
      continue;
    }
    // tile[fi][yi] = ovvv[p,x,F0+fi,Y0+yi]: pair (f, y) (f >= y above the
    // diagonal, consecutive in y) or dense [f][y]
    for (int q = 0; q < 4; q++) {
      const int f = F0 + ty + 8 * q, y = Y0 + tx;
      double val = 0.0;
      if (f < v && y < v) val = packed ? src[tri_index(f, y)] : src[(long long)f * v + y];
      tile[ty + 8 * q][tx] = val;
    }
    __syncthreads();
    for (int q = 0; q < 4; q++) dst[(long long)(Y0 + ty + 8 * q) * V + F0 + tx] = (TP)tile[tx][ty + 8 * q];
    __syncthreads();
  }
}

// t2T[r][q][z][f] = t2[r,q,z,f];  t2H[r][m][y][z] = t2[r,m,z,y]  (per (r, q) slab, zero padded)
template <class TP>
__global__ void __launch_bounds__(256) prep_t2(TP* __restrict__ t2T, TP* __restrict__ t2H,
                                               const double* __restrict__ t2, int o, int v, int Op, int V) {
  __shared__ double tile[TR_T][TR_T + 1];
  const int nt = V / TR_T, ntiles = nt * nt;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int tile_id = (int)(blockIdx.x % (unsigned)ntiles);
  const long long run = blockIdx.x / (unsigned)ntiles;
  const int R0 = (tile_id / nt) * TR_T, C0 = (tile_id % nt) * TR_T;
  const long long nslab = (long long)Op * Op;
  for (int sl = 0; sl < TR_SLABS; sl++) {
    const long long slab = run * TR_SLABS + sl;
    if (slab >= nslab) break;
    const int q = (int)(slab % Op), r = (int)(slab / Op);
    const bool live = r < o && q < o;
    const double* src = t2 + ((long long)r * o + q) * v * v;
    TP* dT = t2T + slab * V * V;
    TP* dH = t2H + slab * V * V;
    // tile[zi][fi] = t2[r,q,R0+zi,C0+fi]
    for (int k = 0; k < 4; k++) {
      const int z = R0 + ty + 8 * k, f = C0 + tx;
      const double val = (live && z < v && f < v) ? src[(long long)z * v + f] : 0.0;
      tile[ty + 8 * k][tx] = val;
      dT[(long long)z * V + f] = (TP)val;
    }
    __syncthreads();
    // t2H row y = C0 + yi, column z = R0 + zi
    for (int k = 0; k < 4; k++) dH[(long long)(C0 + ty + 8 * k) * V + R0 + tx] = (TP)tile[tx][ty + 8 * k];
    __syncthreads();
  }
}

template <class TP>
__global__ void prep_ovoo(TP* __restrict__ out, const double* __restrict__ ovoo, int o, int v, int V,
                          int Op) {
  const long long total = (long long)o * o * V * Op;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(e % Op);
    const int x = (int)((e / Op) % V);
    const int q = (int)((e / ((long long)Op * V)) % o);
    const int p = (int)(e / ((long long)Op * V * o));
    double val = 0.0;
    if (m < o && x < v) val = ovoo[(((long long)p * v + x) * o + m) * o + q];
    out[e] = (TP)val;
  }
}

__global__ void prep_eps(double* __restrict__ eps_occ, double* __restrict__ eps_vir,
                         double* __restrict__ fvoP, const double* __restrict__ mo_energy,
                         const double* __restrict__ fock, int o, int v, int V) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < o * V) {
    const int r = e / V, z = e % V;
    fvoP[e] = z < v ? fock[(long long)(o + z) * (o + v) + r] : 0.0;
  }
  if (e < o) eps_occ[e] = mo_energy[e];
  // padded virtuals get a large orbital energy: every padded term has w0 = v0 = 0,
  // this only keeps its denominator finite and away from zero
  if (e < V) eps_vir[e] = e < v ? mo_energy[o + e] : 1.0e4;
}

// ---------------- V-term adjoints (ccsd_t_vjp.c:171-204, 228-245) -----------------
// Per triple (i,j,k) and contracted position d of the H cube:
//   S_d[e1][e2] = sum_{s_d} H[s] t1[t_d, s_d]            (task d)
//   F_d[e1][e2] = sum_{s_d} H[s] fock[o + s_d, t_d]      (task d)
//   P_d[*][*]   : partial sums of sum_{e1,e2} H[s] CM_d[s_e1][s_e2]   (tasks 3..5)
// with e1 < e2 the two other positions, CM_d / CN_d the two-ordering sums of
//   M_{pq}[x][y] = ovov[p,x,q,y]  and  N_{pq}[x][y] = t2[q,p,y,x].
// Output per triple: 12 V x V FP64 arrays: [S0 F0 S1 F1 S2 F2 PM0 PN0 PM1 PN1 PM2 PN2].

// Three passes over H per triple, one per contracted position, each with
// coalesced reads of H:
//   pass 0 (contract a): 256 (b,c) per block, 4 per lane, 4 waves split a  -> S0, F0
//   pass 1 (contract b): 256 (a,c) per block, 4 per lane, 4 waves split b  -> S1, F1, PM2, PN2
//   pass 2 (contract c): 16 lanes per (a,b), 4 c per lane step -> S2, F2, PM1, PN1, PM0, PN0
// The wave partials of passes 0/1 are added in wave order.  PM_d / PN_d are
// partial sums (over the position that is neither d nor the row), summed to
// vectors by vpart_vecsum.  blockIdx.y = npass * triple + (pass - pass0).
template <class TP>
__global__ void __launch_bounds__(256)
vpart_contract(const VpTriple* __restrict__ trs, int o, int v, int V, const double* __restrict__ t1,
               const double* __restrict__ fvoP, int npass, int pass0) {
  const VpTriple tr = trs[blockIdx.y / npass];
  const int pass = pass0 + (int)(blockIdx.y % npass);
  const TP* H = reinterpret_cast<const TP*>(tr.H);
  const int occ[3] = {tr.i, tr.j, tr.k};
  const long long VV = (long long)V * V;
  double* out = tr.out;
  // passes 0, 1: per-wave partial sums; pass 2: staged rows (red[0])
  __shared__ __attribute__((aligned(32))) double red[4][4][256];
  if (pass < 2) {
    // four consecutive elements per lane, read with one vector load
    typedef TP tp4 __attribute__((ext_vector_type(4)));
    const int el = threadIdx.x & 63, sp = threadIdx.x >> 6;
    const long long e0 = ((long long)blockIdx.x * 64 + el) * 4;
    const bool live = e0 < VV;   // VV is a multiple of 4: all four or none
    const long long ec = live ? e0 : 0;
    const int u = (int)(ec / V), w = (int)(ec % V);   // V is a multiple of 4: one u for all four
    const int td = occ[pass];
    double s[4] = {0, 0, 0, 0}, f[4] = {0, 0, 0, 0}, pm[4] = {0, 0, 0, 0}, pn[4] = {0, 0, 0, 0};
    // unrolled so that eight H loads per lane are in flight (each sum keeps
    // its order: x ascending within the wave)
    if (pass == 0) {        // e = (b, c)
#pragma unroll 8
      for (int x = sp; x < v; x += 4) {
        const tp4 hv = *reinterpret_cast<const tp4*>(H + (long long)x * VV + ec);
        const double t = t1[(long long)td * v + x], fv = fvoP[(long long)td * V + x];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          s[q] += (double)hv[q] * t;
          f[q] += (double)hv[q] * fv;
        }
      }
    } else {                // e = (a, c)
      const double* CM2 = out + (VP_CM + 4) * VV + (long long)u * V;
      const double* CN2 = out + (VP_CM + 5) * VV + (long long)u * V;
#pragma unroll 8
      for (int x = sp; x < v; x += 4) {
        const tp4 hv = *reinterpret_cast<const tp4*>(H + (long long)u * VV + (long long)x * V + w);
        const double t = t1[(long long)td * v + x], fv = fvoP[(long long)td * V + x];
        const double cm = CM2[x], cn = CN2[x];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          s[q] += (double)hv[q] * t;
          f[q] += (double)hv[q] * fv;
          pm[q] += (double)hv[q] * cm;
          pn[q] += (double)hv[q] * cn;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      red[0][sp][el * 4 + q] = s[q];
      red[1][sp][el * 4 + q] = f[q];
      red[2][sp][el * 4 + q] = pm[q];
      red[3][sp][el * 4 + q] = pn[q];
    }
    __syncthreads();
    if (sp != 0 || !live) return;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = el * 4 + q;
      const long long e = e0 + q;
      double r[4];
#pragma unroll
      for (int m = 0; m < 4; m++) r[m] = ((red[m][0][k] + red[m][1][k]) + red[m][2][k]) + red[m][3][k];
      if (pass == 0) {
        out[0 * VV + e] = r[0];
        out[1 * VV + e] = r[1];
      } else {
        out[2 * VV + e] = r[0];
        out[3 * VV + e] = r[1];
        out[(6 + 4) * VV + e] = r[2];   // PM2[a][c]
        out[(7 + 4) * VV + e] = r[3];   // PN2[a][c]
      }
    }
    return;
  }
  // pass 2: 16 lanes per (a, b) row, four rows per wave; a lane takes four
  // consecutive c per step with one vector load of H.  The block's 16 rows
  // share a (V is a multiple of 32), so the rows every one of them needs
  // (CM1[a], CN1[a], fvo, t1) are staged once per 256 c in LDS.
  typedef TP tp4 __attribute__((ext_vector_type(4)));
  double (*srow)[256] = red[0];
  const int lane = threadIdx.x & 63, l16 = lane & 15;
  const long long row = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4 + (lane >> 4);
  const bool live = row < VV;
  const long long rw = live ? row : 0;
  const int b = (int)(rw % V);
  const int ablk = (int)(((long long)blockIdx.x * 16) / V);   // a of all 16 rows
  const int t2o = occ[2];
  const TP* h = H + rw * V;
  const double* CM1 = out + (VP_CM + 2) * VV + (long long)ablk * V;
  const double* CN1 = out + (VP_CM + 3) * VV + (long long)ablk * V;
  const double* CM0 = out + (VP_CM + 0) * VV + (long long)b * V;
  const double* CN0 = out + (VP_CM + 1) * VV + (long long)b * V;
  const double* t1r = t1 + (long long)t2o * v;
  const double* fvr = fvoP + (long long)t2o * V;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  // the pair-matrix rows are V doubles (a multiple of 32) from 256-byte
  // aligned bases: 32-byte vector loads
  typedef double dv4 __attribute__((ext_vector_type(4)));
  for (int cc = 0; cc < v; cc += 256) {
    __syncthreads();   // the previous chunk's LDS reads are done
    {
      const int c = cc + (int)threadIdx.x;
      srow[0][threadIdx.x] = c < V ? CM1[c] : 0.0;
      srow[1][threadIdx.x] = c < V ? CN1[c] : 0.0;
      srow[2][threadIdx.x] = c < V ? fvr[c] : 0.0;
      srow[3][threadIdx.x] = c < v ? t1r[c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int cl = l16 * 4 + 64 * k, c0 = cc + cl;   // c0 + 3 < V: V is a multiple of 32
      if (c0 >= v) break;
      const tp4 hv = *reinterpret_cast<const tp4*>(h + c0);
      const dv4 m0 = *reinterpret_cast<const dv4*>(CM0 + c0), n0 = *reinterpret_cast<const dv4*>(CN0 + c0);
      const dv4 m1 = *reinterpret_cast<const dv4*>(&srow[0][cl]), n1 = *reinterpret_cast<const dv4*>(&srow[1][cl]);
      const dv4 fv = *reinterpret_cast<const dv4*>(&srow[2][cl]), tv = *reinterpret_cast<const dv4*>(&srow[3][cl]);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int c = c0 + q;
        if (c < v) {
          const double x = (double)hv[q];
          acc[0] += x * tv[q];
          acc[1] += x * fv[q];
          acc[2] += x * m1[q];
          acc[3] += x * n1[q];
          acc[4] += x * m0[q];
          acc[5] += x * n0[q];
        }
      }
    }
  }
  // fixed butterfly over the row's 16 lanes: every lane of the group ends with
  // the same sum, independent of scheduling
#pragma unroll
  for (int q = 0; q < 6; q++)
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off, 64);
  if (l16 == 0 && live) {
    out[4 * VV + row] = acc[0];            // S2[a][b]
    out[5 * VV + row] = acc[1];            // F2[a][b]
    out[(6 + 2) * VV + row] = acc[2];      // PM1[a][b]
    out[(7 + 2) * VV + row] = acc[3];      // PN1[a][b]
    out[(6 + 0) * VV + row] = acc[4];      // PM0[a][b]
    out[(7 + 0) * VV + row] = acc[5];      // PN0[a][b]
  }
}

// Vector sums of the partial matrices, per triple and task = 2 d + (M/N):
//   d = 0: row sums    vsum[z] = sum_b P0[z][b]
//   d = 1: column sums vsum[z] = sum_a P1[a][z]
//   d = 2: column sums vsum[z] = sum_a P2[a][z]
// into out[VP_VSUM][task][z]; every sum in a fixed order.
__global__ void __launch_bounds__(256)
vpart_vecsum(const VpTriple* __restrict__ trs, int V) {
  __shared__ double red[4][64];
  const VpTriple tr = trs[blockIdx.y / 6];
  const int task = blockIdx.y % 6, d = task / 2;
  const long long VV = (long long)V * V;
  const double* P = tr.out + (6 + task) * VV;
  double* vs = tr.out + (long long)VP_VSUM * VV + (long long)task * V;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (d == 0) {
    // 16 rows per wave, lanes over the row (unrolled: the loads of several
    // rows in flight together; every sum keeps its order)
#pragma unroll 4
    for (int rr = 0; rr < 16; rr++) {
      const int z = blockIdx.x * 64 + wv * 16 + rr;
      if (z >= V) break;
      double sacc = 0.0;
#pragma unroll 4
      for (int b = lane; b < V; b += 64) sacc += P[(long long)z * V + b];
      for (int off = 32; off > 0; off >>= 1) sacc += __shfl_xor(sacc, off, 64);
      if (lane == 0) vs[z] = sacc;
    }
    return;
  }
  const int z = blockIdx.x * 64 + lane;
  double sacc = 0.0;
  if (z < V)
#pragma unroll 16
    for (int a = wv; a < V; a += 4) sacc += P[(long long)a * V + z];
  red[wv][lane] = sacc;
  __syncthreads();
  if (wv == 0 && z < V) vs[z] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Pair matrices of the V term of one triple (FP64), task = 2 d + (M/N):
//   CM_d[x][y] = M_{p1 p2}[x][y] + M_{p2 p1}[y][x],  M_pq[x][y] = ovov[p,x,q,y]
//   CN_d[x][y] = N_{p1 p2}[x][y] + N_{p2 p1}[y][x],  N_pq[x][y] = t2[q,p,y,x]
// (p1, p2) = (t_e1, t_e2), e1 < e2 the positions other than d.  They sum the
// two orderings with the same contracted position in ccsd_t.c:149-150 and
// ccsd_t_vjp.c:171-204, 228-245.
// A block builds a 32 x 32 (x, y) tile: the term read transposed (contiguous
// in x) goes through LDS, so both reads and the write are coalesced.
__global__ void __launch_bounds__(256)
vcm_build(const VpTriple* __restrict__ trs, int o, int v, int V, const double* __restrict__ ovov,
          const double* __restrict__ t2) {
  __shared__ double tt[32][33];
  const VpTriple tr = trs[blockIdx.y / 6];
  const int task = blockIdx.y % 6;
  const long long VV = (long long)V * V;
  const int tiles = (V + 31) / 32;
  const int x0 = (int)(blockIdx.x / tiles) * 32, y0 = (int)(blockIdx.x % tiles) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
  const int d = task / 2, which = task % 2;
  const int occ[3] = {tr.i, tr.j, tr.k};
  const int p1 = occ[d == 0 ? 1 : 0], p2 = occ[d == 2 ? 1 : 2];
  // transposed term at (x, y) = x0 + tx, y0 + ty + 8 k, stored as tt[y - y0][x - x0]
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int x = x0 + tx, y = y0 + ty + 8 * k;
    double b = 0.0;
    if (x < v && y < v)
      b = which == 0 ? ovov[(((long long)p2 * v + y) * o + p1) * v + x] : t2[(((long long)p2 * o + p1) * v + y) * v + x];
    tt[ty + 8 * k][tx] = b;
  }
  __syncthreads();
  // direct term and write at (x, y) = x0 + ty + 8 k, y0 + tx
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int x = x0 + ty + 8 * k, y = y0 + tx;
    if (x >= V || y >= V) continue;
    double val = 0.0;
    if (x < v && y < v)
      val = (which == 0 ? ovov[(((long long)p1 * v + x) * o + p2) * v + y]
                        : t2[(((long long)p1 * o + p2) * v + x) * v + y]) +
            tt[tx][ty + 8 * k];
    tr.out[(VP_CM + task) * VV + (long long)x * V + y] = val;
  }
}

// Scatter of the V-term adjoints of one batch.  Slab tasks (blockIdx.y <
// nslab): element (x, y) of the slab (p, q) adds 1/2 of every (triple, d,
// transposed) contribution, in list order, to ovov_bar[p, x, q, y] (in place)
// and to t2h_acc[q][p][x][y] (the [i][j][b][a] accumulator of t2_bar, shared
// with the hole adjoint).  Vector tasks (blockIdx.y == nslab): t1_acc[r][z],
// fvo_acc[r][z].

__global__ void __launch_bounds__(256)
vpart_scatter(const VpSlab* __restrict__ slabs, int nslab, const VpContrib* __restrict__ contribs,
              const VpTriple* __restrict__ trs, int ntriple, int o, int v, int Op, int V,
              double* __restrict__ ovov_bar, double* __restrict__ t2h_acc, double* __restrict__ t1_acc,
              double* __restrict__ fvo_acc) {
  const long long VV = (long long)V * V;
  if ((int)blockIdx.y < nslab) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= VV) return;
    const int x = (int)(e / V), y = (int)(e % V);
    if (x >= v || y >= v) return;
    const VpSlab sl = slabs[blockIdx.y];
    double am = 0.0, af = 0.0;
    for (int c = sl.c0; c < sl.c0 + sl.nc; c++) {
      const VpContrib ct = contribs[c];
      const double* out = trs[ct.slot].out;
      const long long idx = ct.transposed ? (long long)y * V + x : e;
      am += out[(2 * ct.d) * VV + idx];
      af += out[(2 * ct.d + 1) * VV + idx];
    }
    ovov_bar[(((long long)sl.p * v + x) * o + sl.q) * v + y] += 0.5 * am;
    t2h_acc[((long long)sl.q * Op + sl.p) * VV + e] += 0.5 * af;
    return;
  }
  // vectors: thread per (r, z), the batch's vector sums in triple and d order
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)o * V) return;
  const int r = (int)(e / V), z = (int)(e % V);
  double sm = 0.0, sn = 0.0;
  for (int t = 0; t < ntriple; t++) {
    const VpTriple& tr = trs[t];
    const int occ[3] = {tr.i, tr.j, tr.k};
    const double* vs = tr.out + (long long)VP_VSUM * VV;
    for (int d = 0; d < 3; d++) {
      if (occ[d] != r) continue;
      sm += vs[(2 * d) * V + z];
      sn += vs[(2 * d + 1) * V + z];
    }
  }
  t1_acc[e] += 0.5 * sm;
  fvo_acc[e] += 0.5 * sn;
}

// ---------------- finalisation into the reference's cotangent layouts --------------
// (ccsd_t.py:179-192)
__global__ void __launch_bounds__(256) fin_ovvv(double* __restrict__ out, const double* __restrict__ acc, int o,
                                                int v, int V, int packed) {
  // packed: out[p,x, pair (f, y)] = acc[p][x][y][f] + acc[p][x][f][y] (f > y) or
  // acc[p][x][f][f]; tiles (F0 rows, Y0 columns) on or below the diagonal only.
  // dense: out[p,x,f,y] = acc[p][x][y][f].
  __shared__ double tile[TR_T][TR_T + 1];
  const int nt = (v + TR_T - 1) / TR_T;
  const int ntiles = packed ? nt * (nt + 1) / 2 : nt * nt;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int tile_id = (int)(blockIdx.x % (unsigned)ntiles);
  const long long run = blockIdx.x / (unsigned)ntiles;
  int tf, tyy;
  if (packed) {
    tf = 0;
    while ((tf + 1) * (tf + 2) / 2 <= tile_id) tf++;
    tyy = tile_id - tf * (tf + 1) / 2;
  } else {
    tf = tile_id / nt;
    tyy = tile_id % nt;
  }
  const int F0 = tf * TR_T, Y0 = tyy * TR_T;
  const long long npair = (long long)v * (v + 1) / 2;
  const long long nslab = (long long)o * v;
  for (int sl = 0; sl < TR_SLABS; sl++) {
    const long long slab = run * TR_SLABS + sl;
    if (slab >= nslab) break;
    const int x = (int)(slab % v);
    const long long p = slab / v;
    const double* a = acc + (p * V + x) * V * V;
    // tile[yi][fi] = acc[p][x][Y0+yi][F0+fi]; packed: the other triangle's
    // element of each output pair is read in the same round (one round trip)
    double lower[4];
    for (int q = 0; q < 4; q++) {
      const int y = Y0 + ty + 8 * q, f = F0 + tx;
      tile[ty + 8 * q][tx] = (y < v && f < v) ? a[(long long)y * V + f] : 0.0;
      const int fo = F0 + ty + 8 * q, yo = Y0 + tx;
      lower[q] = (packed && fo < v && yo < fo) ? a[(long long)fo * V + yo] : 0.0;
    }
    __syncthreads();
    if (packed) {
      double* dst = out + slab * npair;
      for (int q = 0; q < 4; q++) {
        const int f = F0 + ty + 8 * q, y = Y0 + tx;
        if (f < v && y <= f) dst[(long long)f * (f + 1) / 2 + y] = tile[tx][ty + 8 * q] + lower[q];
      }
    } else {
      double* dst = out + slab * v * v;
      for (int q = 0; q < 4; q++) {
        const int f = F0 + ty + 8 * q, y = Y0 + tx;
        if (f < v && y < v) dst[(long long)f * v + y] = tile[tx][ty + 8 * q];
      }
    }
    __syncthreads();
  }
}

// t2_bar[i,j,a,b] += t2hacc[i][j][b][a] (the rest of t2_bar was accumulated in
// place), per (i, j) slab
__global__ void __launch_bounds__(256) fin_t2(double* __restrict__ out, const double* __restrict__ t2hacc, int o,
                                              int v, int Op, int V) {
  __shared__ double tile[TR_T][TR_T + 1];
  const int nt = (v + TR_T - 1) / TR_T, ntiles = nt * nt;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int tile_id = (int)(blockIdx.x % (unsigned)ntiles);
  const long long run = blockIdx.x / (unsigned)ntiles;
  const int A0 = (tile_id / nt) * TR_T, B0 = (tile_id % nt) * TR_T;
  const long long nslab = (long long)o * o;
  for (int sl = 0; sl < TR_SLABS; sl++) {
    const long long slab = run * TR_SLABS + sl;
    if (slab >= nslab) break;
    const int j = (int)(slab % o), i = (int)(slab / o);
    const double* h = t2hacc + ((long long)i * Op + j) * V * V;
    // tile[bi][ai] = t2hacc[i][j][B0+bi][A0+ai]
    for (int q = 0; q < 4; q++) {
      const int b = B0 + ty + 8 * q, a = A0 + tx;
      tile[ty + 8 * q][tx] = (a < v && b < v) ? h[(long long)b * V + a] : 0.0;
    }
    __syncthreads();
    double* dst = out + slab * v * v;
    for (int q = 0; q < 4; q++) {
      const int a = A0 + ty + 8 * q, b = B0 + tx;
      if (a < v && b < v) dst[(long long)a * v + b] += tile[tx][ty + 8 * q];
    }
    __syncthreads();
  }
}

// ovoo_bar[p,x,m,q] = acc[p][q][x][m]; t1_bar, fock_bar and mo_energy_bar from
// their accumulators (ovov_bar is accumulated in place).  With t1_bar null,
// ovoo_bar only.
__global__ void fin_small(double* __restrict__ ovoo_bar, double* __restrict__ t1_bar,
                          double* __restrict__ fock_bar, double* __restrict__ mo_energy_bar,
                          const double* __restrict__ ovoo_acc, const double* __restrict__ t1_acc,
                          const double* __restrict__ fvo_acc, const double* __restrict__ eps_bar, int o,
                          int v, int Op, int V) {
  const int n = o + v;
  const long long n1 = (long long)o * v * o * o, n3 = (long long)o * v, n4 = (long long)n * n, n5 = n;
  const long long total = t1_bar ? n1 + n3 + n4 + n5 : n1;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    long long r = e;
    if (r < n1) {
      const int q = (int)(r % o), m = (int)((r / o) % o), x = (int)((r / ((long long)o * o)) % v);
      const int p = (int)(r / ((long long)o * o * v));
      ovoo_bar[r] = ovoo_acc[(((long long)p * o + q) * V + x) * Op + m];
      continue;
    }
    r -= n1;
    if (r < n3) {
      const int z = (int)(r % v), i = (int)(r / v);
      t1_bar[r] = t1_acc[(long long)i * V + z];
      continue;
    }
    r -= n3;
    if (r < n4) {
      const int col = (int)(r % n), row = (int)(r / n);
      fock_bar[r] = (row >= o && col < o) ? fvo_acc[(long long)col * V + (row - o)] : 0.0;
      continue;
    }
    r -= n4;
    mo_energy_bar[r] = r < o ? eps_bar[r] : eps_bar[o + (r - o)];
  }
}

}  // namespace t3

using namespace t3;

static inline dim3 grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536LL * 8) b = 65536LL * 8;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

// grid of the tiled re-layouts: every tile of every run of TR_SLABS slabs
static inline dim3 tile_grid(long long ntiles, long long nslab) {
  const long long runs = (nslab + TR_SLABS - 1) / TR_SLABS;
  return dim3((unsigned)std::max(1LL, ntiles * runs));
}

extern "C" hipError_t t3_prep_inputs(int tp_is_f64, void* ovvvT, void* t2T, void* t2H, void* ovooT,
                                     double* eps_occ, double* eps_vir, double* fvoP, const double* ovvv,
                                     int packed, const double* t2, const double* ovoo, const double* mo_energy,
                                     const double* fock, int o, int v, int Op, int V, hipStream_t s) {
  const long long n_ovoo = (long long)o * o * V * Op;
  const dim3 g_ovvv = tile_grid((long long)(V / TR_T) * (V / TR_T), (long long)o * V);
  const dim3 g_t2 = tile_grid((long long)(V / TR_T) * (V / TR_T), (long long)Op * Op);
  if (tp_is_f64) {
    if (ovvv) hipLaunchKernelGGL(prep_ovvv<double>, g_ovvv, dim3(256), 0, s, (double*)ovvvT, ovvv, o, v, V, packed);
    hipLaunchKernelGGL(prep_t2<double>, g_t2, dim3(256), 0, s, (double*)t2T, (double*)t2H, t2, o, v, Op, V);
    hipLaunchKernelGGL(prep_ovoo<double>, grid_for(n_ovoo), dim3(256), 0, s, (double*)ovooT, ovoo, o, v, V, Op);
  } else {
    if (ovvv) hipLaunchKernelGGL(prep_ovvv<float>, g_ovvv, dim3(256), 0, s, (float*)ovvvT, ovvv, o, v, V, packed);
    hipLaunchKernelGGL(prep_t2<float>, g_t2, dim3(256), 0, s, (float*)t2T, (float*)t2H, t2, o, v, Op, V);
    hipLaunchKernelGGL(prep_ovoo<float>, grid_for(n_ovoo), dim3(256), 0, s, (float*)ovooT, ovoo, o, v, V, Op);
  }
  const long long ne = (long long)o * V > V ? (long long)o * V : V;
  hipLaunchKernelGGL(prep_eps, grid_for(ne), dim3(256), 0, s, eps_occ, eps_vir, fvoP, mo_energy, fock, o, v, V);
  return hipGetLastError();
}

// Device layouts of the slot-projected operands of the LNO form, from their
// reference-layout FP64 forms (any of them may be null):
//   ovvvT_P <- ovvvP;  t2T_Q <- t2Q;  t2T_P0, t2H_P0 <- t2P0;  ovooT_Q <- ovooQ.
// prep_t2 writes both t2 layouts: t2Q's t2H goes to t2H_P0 first and is
// overwritten by t2P0's (stream order), so no buffer is needed for it.
extern "C" hipError_t t3_prep_projected(int tp_is_f64, void* ovvvT_P, void* t2T_Q, void* t2T_P0, void* t2H_P0,
                                        void* ovooT_Q, const double* ovvvP, int packed, const double* t2Q,
                                        const double* t2P0, const double* ovooQ, int o, int v, int Op, int V,
                                        hipStream_t s) {
  const long long n_ovoo = (long long)o * o * V * Op;
  const dim3 g_ovvv = tile_grid((long long)(V / TR_T) * (V / TR_T), (long long)o * V);
  const dim3 g_t2 = tile_grid((long long)(V / TR_T) * (V / TR_T), (long long)Op * Op);
  if (tp_is_f64) {
    if (ovvvP) hipLaunchKernelGGL(prep_ovvv<double>, g_ovvv, dim3(256), 0, s, (double*)ovvvT_P, ovvvP, o, v, V, packed);
    if (t2Q) hipLaunchKernelGGL(prep_t2<double>, g_t2, dim3(256), 0, s, (double*)t2T_Q, (double*)t2H_P0, t2Q, o, v, Op, V);
    if (t2P0) hipLaunchKernelGGL(prep_t2<double>, g_t2, dim3(256), 0, s, (double*)t2T_P0, (double*)t2H_P0, t2P0, o, v, Op, V);
    if (ovooQ) hipLaunchKernelGGL(prep_ovoo<double>, grid_for(n_ovoo), dim3(256), 0, s, (double*)ovooT_Q, ovooQ, o, v, V, Op);
  } else {
    if (ovvvP) hipLaunchKernelGGL(prep_ovvv<float>, g_ovvv, dim3(256), 0, s, (float*)ovvvT_P, ovvvP, o, v, V, packed);
    if (t2Q) hipLaunchKernelGGL(prep_t2<float>, g_t2, dim3(256), 0, s, (float*)t2T_Q, (float*)t2H_P0, t2Q, o, v, Op, V);
    if (t2P0) hipLaunchKernelGGL(prep_t2<float>, g_t2, dim3(256), 0, s, (float*)t2T_P0, (float*)t2H_P0, t2P0, o, v, Op, V);
    if (ovooQ) hipLaunchKernelGGL(prep_ovoo<float>, grid_for(n_ovoo), dim3(256), 0, s, (float*)ovooT_Q, ovooQ, o, v, V, Op);
  }
  return hipGetLastError();
}

// Reference-layout FP64 cotangents of the slot-projected LNO operands from
// their device accumulators (any output may be null):
//   ovvvP_bar <- ovvvP_acc (packed like the input);  t2P0_bar += t2P0hacc^T
//   (the rest of t2P0_bar accumulated in place);  ovooQ_bar <- ovooQ_acc.
extern "C" hipError_t t3_finalize_projected(double* ovvvP_bar, int packed, const double* ovvvP_acc,
                                            double* t2P0_bar, const double* t2P0hacc, double* ovooQ_bar,
                                            const double* ovooQ_acc, int o, int v, int Op, int V, hipStream_t s) {
  const long long ntv = (v + TR_T - 1) / TR_T;
  if (ovvvP_bar)
    hipLaunchKernelGGL(fin_ovvv, tile_grid(packed ? ntv * (ntv + 1) / 2 : ntv * ntv, (long long)o * v), dim3(256), 0, s,
                       ovvvP_bar, ovvvP_acc, o, v, V, packed);
  if (t2P0_bar)
    hipLaunchKernelGGL(fin_t2, tile_grid(ntv * ntv, (long long)o * o), dim3(256), 0, s, t2P0_bar, t2P0hacc, o, v, Op, V);
  if (ovooQ_bar)
    hipLaunchKernelGGL(fin_small, grid_for((long long)o * v * o * o), dim3(256), 0, s, ovooQ_bar, (double*)nullptr,
                       (double*)nullptr, (double*)nullptr, ovooQ_acc, (const double*)nullptr, (const double*)nullptr,
                       (const double*)nullptr, o, v, Op, V);
  return hipGetLastError();
}

extern "C" hipError_t t3_launch_vpart(int tp_is_f64, const VpTriple* trs, int ntriple, const VpSlab* slabs,
                                      int nslab, const VpContrib* contribs, int o, int v, int Op, int V,
                                      const double* t1, const double* fvoP, double* ovov_bar, double* t2h_acc,
                                      double* t1_acc, double* fvo_acc, hipStream_t s) {
  if (ntriple <= 0) return hipSuccess;
  const long long VV = (long long)V * V;
  // passes 0/1: four output elements per lane (256 per block); pass 2: 16 lanes per row (16 rows per block)
  dim3 g1((unsigned)((VV + 255) / 256), 2 * ntriple), g2p((unsigned)((VV + 15) / 16), ntriple);
  if (tp_is_f64) {
    hipLaunchKernelGGL(vpart_contract<double>, g1, dim3(256), 0, s, trs, o, v, V, t1, fvoP, 2, 0);
    hipLaunchKernelGGL(vpart_contract<double>, g2p, dim3(256), 0, s, trs, o, v, V, t1, fvoP, 1, 2);
  } else {
    hipLaunchKernelGGL(vpart_contract<float>, g1, dim3(256), 0, s, trs, o, v, V, t1, fvoP, 2, 0);
    hipLaunchKernelGGL(vpart_contract<float>, g2p, dim3(256), 0, s, trs, o, v, V, t1, fvoP, 1, 2);
  }
  hipLaunchKernelGGL(vpart_vecsum, dim3((unsigned)((V + 63) / 64), 6 * ntriple), dim3(256), 0, s, trs, V);
  const long long nx = VV > (long long)o * V ? VV : (long long)o * V;
  dim3 g2((unsigned)((nx + 255) / 256), nslab + 1);
  hipLaunchKernelGGL(vpart_scatter, g2, dim3(256), 0, s, slabs, nslab, contribs, trs, ntriple, o, v, Op, V,
                     ovov_bar, t2h_acc, t1_acc, fvo_acc);
  return hipGetLastError();
}

extern "C" hipError_t t3_launch_vcm(const VpTriple* trs, int ntriple, int o, int v, int V, const double* ovov,
                                    const double* t2, hipStream_t s) {
  if (ntriple <= 0) return hipSuccess;
  const int tiles = (V + 31) / 32;
  dim3 g((unsigned)(tiles * tiles), 6 * ntriple);
  hipLaunchKernelGGL(vcm_build, g, dim3(256), 0, s, trs, o, v, V, ovov, t2);
  return hipGetLastError();
}

extern "C" hipError_t t3_finalize(double* t1_bar, double* t2_bar, double* fock_bar, double* mo_energy_bar,
                                  double* ovoo_bar, double* ovvv_bar, int packed, const double* ovvv_acc,
                                  const double* t2hacc, const double* ovoo_acc, const double* t1_acc,
                                  const double* fvo_acc, const double* eps_bar, int o, int v, int Op, int V,
                                  hipStream_t s) {
  const long long ntv = (v + TR_T - 1) / TR_T;
  if (ovvv_bar)
    hipLaunchKernelGGL(fin_ovvv, tile_grid(packed ? ntv * (ntv + 1) / 2 : ntv * ntv, (long long)o * v), dim3(256), 0, s,
                       ovvv_bar, ovvv_acc, o, v, V, packed);
  hipLaunchKernelGGL(fin_t2, tile_grid(ntv * ntv, (long long)o * o), dim3(256), 0, s, t2_bar, t2hacc, o, v, Op, V);
  const long long n = o + v;
  const long long nsmall = (long long)o * v * o * o + (long long)o * v + n * n + n;
  hipLaunchKernelGGL(fin_small, grid_for(nsmall), dim3(256), 0, s, ovoo_bar, t1_bar, fock_bar, mo_energy_bar,
                     ovoo_acc, t1_acc, fvo_acc, eps_bar, o, v, Op, V);
  return hipGetLastError();
}

// ---------------- density-fitted ovvv (slabs built on the device) ----------------
// The DF factors of the reference's DF-CCSD (pyscfad/cc/dfccsd.py:81-91):
//   Lov (naux, o, v),  Lvv (naux, v(v+1)/2) lower-triangle packed over (y, f),
//   ovvv[p, x, :] = sum_P Lov[P, p, x] Lvv[P, :]   (dfccsd.py:91)
// are re-laid out with the auxiliary index P innermost and padded to NP, so
// that a slab build and both cotangent folds are GEMMs whose operands are
// k-contiguous or row-contiguous strided views:
//   LovP[p][x][P]  (o x V x NP),   LvvF[y][f][P]  (V x V x NP, both triangles).
// Rows x, y, f >= v and P >= naux stay zero.  64 x 64 tiles through LDS keep
// both the reads and the writes coalesced.

// row-major lower-triangle pair index -> (y, f), f <= y
__device__ __forceinline__ void tri_decode(long long pair, int& y, int& f) {
  int r = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
  while ((long long)r * (r + 1) / 2 > pair) r--;
  while ((long long)(r + 1) * (r + 2) / 2 <= pair) r++;
  y = r;
  f = (int)(pair - (long long)r * (r + 1) / 2);
}

// grid (ceil(o*v / 64), ceil(NP / 64))
__global__ void __launch_bounds__(256) df_lov_in(double* __restrict__ LovP, const double* __restrict__ Lov,
                                                 int naux, int o, int v, int V, int NP) {
  __shared__ double tile[64][65];
  const long long ncol = (long long)o * v;
  const long long c0 = (long long)blockIdx.x * 64;
  const int P0 = blockIdx.y * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll 4
  for (int i = 0; i < 16; i++) {
    const int r = ty + 4 * i, P = P0 + r;
    const long long c = c0 + tx;
    tile[r][tx] = (P < naux && c < ncol) ? Lov[(long long)P * ncol + c] : 0.0;
  }
  __syncthreads();
#pragma unroll 4
  for (int i = 0; i < 16; i++) {
    const int r = ty + 4 * i, P = P0 + tx;
    const long long c = c0 + r;
    if (c < ncol && P < NP) LovP[((c / v) * V + c % v) * NP + P] = tile[tx][r];
  }
}

// grid (ceil(v(v+1)/2 / 64), ceil(NP / 64))
__global__ void __launch_bounds__(256) df_lvv_in(double* __restrict__ LvvF, const double* __restrict__ Lvv,
                                                 int naux, int v, int V, int NP) {
  __shared__ double tile[64][65];
  const long long npair = (long long)v * (v + 1) / 2;
  const long long c0 = (long long)blockIdx.x * 64;
  const int P0 = blockIdx.y * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll 4
  for (int i = 0; i < 16; i++) {
    const int r = ty + 4 * i, P = P0 + r;
    const long long c = c0 + tx;
    tile[r][tx] = (P < naux && c < npair) ? Lvv[(long long)P * npair + c] : 0.0;
  }
  __syncthreads();
#pragma unroll 4
  for (int i = 0; i < 16; i++) {
    const int r = ty + 4 * i, P = P0 + tx;
    const long long c = c0 + r;
    if (c < npair && P < NP) {
      int y, f;
      tri_decode(c, y, f);
      const double x = tile[tx][r];
      LvvF[((long long)y * V + f) * NP + P] = x;
      if (y != f) LvvF[((long long)f * V + y) * NP + P] = x;
    }
  }
}

// Lov_bar[P, p, x] = LovBar[p][x][P];  grid (ceil(o*v / 64), ceil(naux / 64))
__global__ void __launch_bounds__(256) df_lov_out(double* __restrict__ Lov_bar, const double* __restrict__ LovBar,
                                                  int naux, int o, int v, int V, int NP) {
  __shared__ double tile[64][65];
  const long long ncol = (long long)o * v;
  const long long c0 = (long long)blockIdx.x * 64;
  const int P0 = blockIdx.y * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll 4
  for (int i = 0; i < 16; i++) {
    const int r = ty + 4 * i, P = P0 + tx;
    const long long c = c0 + r;
    tile[r][tx] = (c < ncol && P < naux) ? LovBar[((c / v) * V + c % v) * NP + P] : 0.0;
  }
  __syncthreads();
#pragma unroll 4
  for (int i = 0; i < 16; i++) {
    const int r = ty + 4 * i, P = P0 + r;
    const long long c = c0 + tx;
    if (P < naux && c < ncol) Lov_bar[(long long)P * ncol + c] = tile[tx][r];
  }
}

// Lvv_bar[P, pair(y, f)] = LvvBar[y][f][P] + LvvBar[f][y][P] (y != f), LvvBar[y][y][P]
// (the adjoint of the packed-to-symmetric unpacking, pyscfad/lib/_numpy_helper_opt.py:32-55);
// grid (ceil(v(v+1)/2 / 64), ceil(naux / 64))
__global__ void __launch_bounds__(256) df_lvv_out(double* __restrict__ Lvv_bar, const double* __restrict__ LvvBar,
                                                  int naux, int v, int V, int NP) {
  __shared__ double tile[64][65];
  const long long npair = (long long)v * (v + 1) / 2;
  const long long c0 = (long long)blockIdx.x * 64;
  const int P0 = blockIdx.y * 64, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll 4
  for (int i = 0; i < 16; i++) {
    const int r = ty + 4 * i, P = P0 + tx;
    const long long c = c0 + r;
    double x = 0.0;
    if (c < npair && P < naux) {
      int y, f;
      tri_decode(c, y, f);
      x = LvvBar[((long long)y * V + f) * NP + P];
      if (y != f) x += LvvBar[((long long)f * V + y) * NP + P];
    }
    tile[r][tx] = x;
  }
  __syncthreads();
#pragma unroll 4
  for (int i = 0; i < 16; i++) {
    const int r = ty + 4 * i, P = P0 + r;
    const long long c = c0 + tx;
    if (P < naux && c < npair) Lvv_bar[(long long)P * npair + c] = tile[tx][r];
  }
}

extern "C" hipError_t t3_zero(void* p, size_t bytes, hipStream_t s);

extern "C" hipError_t t3_prep_df(double* LovP, double* LvvF, const double* Lov, const double* Lvv, int naux, int o,
                                 int v, int V, int NP, hipStream_t s) {
  hipError_t e;
  if ((e = t3_zero(LovP, (size_t)o * V * NP * 8, s)) != hipSuccess) return e;
  if ((e = t3_zero(LvvF, (size_t)V * V * NP * 8, s)) != hipSuccess) return e;
  const long long npair = (long long)v * (v + 1) / 2;
  const unsigned pt = (unsigned)((NP + 63) / 64);
  hipLaunchKernelGGL(df_lov_in, dim3((unsigned)(((long long)o * v + 63) / 64), pt), dim3(256), 0, s, LovP, Lov,
                     naux, o, v, V, NP);
  hipLaunchKernelGGL(df_lvv_in, dim3((unsigned)((npair + 63) / 64), pt), dim3(256), 0, s, LvvF, Lvv, naux, v, V,
                     NP);
  return hipGetLastError();
}

extern "C" hipError_t t3_finalize_df(double* Lov_bar, double* Lvv_bar, const double* LovBar, const double* LvvBar,
                                     int naux, int o, int v, int V, int NP, hipStream_t s) {
  const long long npair = (long long)v * (v + 1) / 2;
  const unsigned pt = (unsigned)((naux + 63) / 64);
  hipLaunchKernelGGL(df_lov_out, dim3((unsigned)(((long long)o * v + 63) / 64), pt), dim3(256), 0, s, Lov_bar,
                     LovBar, naux, o, v, V, NP);
  hipLaunchKernelGGL(df_lvv_out, dim3((unsigned)((npair + 63) / 64), pt), dim3(256), 0, s, Lvv_bar, LvvBar, naux,
                     v, V, NP);
  return hipGetLastError();
}

// ---------------- stream-ordered clears and copies as kernels ----------------
// (kernels rather than hipMemsetAsync / hipMemcpyAsync so that every node of a
// captured graph is a kernel node)
__global__ void zero_fill(uint4* __restrict__ p, long long n16) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n16;
       e += (long long)gridDim.x * blockDim.x)
    p[e] = uint4{0u, 0u, 0u, 0u};
}

__global__ void zero_f64(double* __restrict__ p, long long n) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    p[e] = 0.0;
}

// x[e] *= *scale, the factor read from device memory (a traced cotangent)
__global__ void scale_f64(double* __restrict__ x, long long n, const double* __restrict__ scale) {
  const double f = *scale;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    x[e] *= f;
}

__global__ void copy_f64(double* __restrict__ dst, const double* __restrict__ src, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) dst[e] = src[e];
}

// bytes must be a multiple of 16 and p 16-byte aligned
extern "C" hipError_t t3_zero(void* p, size_t bytes, hipStream_t s) {
  const long long n16 = (long long)(bytes / 16);
  if (n16 <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_fill, grid_for(n16), dim3(256), 0, s, reinterpret_cast<uint4*>(p), n16);
  return hipGetLastError();
}

// any 8-byte aligned array of n doubles (caller-owned outputs)
extern "C" hipError_t t3_zero_f64(double* p, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_f64, grid_for(n), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

extern "C" hipError_t t3_scale_f64(double* x, long long n, const double* scale, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(scale_f64, grid_for(n), dim3(256), 0, s, x, n, scale);
  return hipGetLastError();
}

extern "C" hipError_t t3_copy_f64(double* dst, const double* src, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(copy_f64, dim3((n + 255) / 256), dim3(256), 0, s, dst, src, n);
  return hipGetLastError();
}

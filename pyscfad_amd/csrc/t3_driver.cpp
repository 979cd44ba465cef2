// Host driver of the gfx950 CCSD(T) (T) energy and VJP: workspace layout,
// the static launch plan over unique occupied triples, and the C ABI
// declared in include/pyscfad_amd.h.
//
// Algorithm (unique occupied triples i >= j >= k, in batches of B triples):
//   1. W-build: two grouped GEMM launches produce, per triple, the cube
//      w0[a][b][c] = sum over the six simultaneous orderings rho of
//        X_{t_rho}[s_rho],  X_{pqr}[x,y,z] = sum_f ovvv[p,x,f,y] t2[r,q,z,f]
//                                            - sum_m ovoo[p,x,m,q] t2[r,m,z,y]
//      (the W of ccsd_t.c:121-152 for every virtual triple at once).
//   2. Orbit epilogue (t3_epilogue.hip): V term, r3, FP64 denominators, E(T)
//      partials; backward also G = dE/dw0, H = dE/dv0 and mo_energy partials.
//   3. Backward only: the adjoint contractions of ccsd_t_vjp.c:139-246 over G as
//      grouped GEMMs into FP64 accumulators, and over H as memory-bound kernels.
//   4. Backward only: finalisation into the reference's cotangent layouts
//      (ccsd_t.py:179-192).
// Every output element has one writer per launch and launches are stream
// ordered, so results are deterministic and no atomics are used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "t3_common.h"
#include "../../include/pyscfad_amd.h"

using namespace t3;

extern "C" {
int t3_gemm_tile_m(int cfg);
int t3_gemm_tile_n(int cfg);
int t3_gemm_vec(int prec);
int t3_gemm_cfg_valid(int prec, int cfg);
int t3_gemm_chunk(int prec);
hipError_t t3_launch_gemm(int prec, int cfg, const GemmProb* probs, int nprob, const GemmSeg* segs,
                          long long nblocks, hipStream_t stream);
hipError_t t3_launch_reduce(int prec, const ReduceProb* probs, int nprob, long long total, int vec,
                            hipStream_t stream);
hipError_t t3_launch_epilogue(int tp_is_f64, const EpiArgs* args, const EpiTriple* triples, int ntriple,
                              hipStream_t stream);
hipError_t t3_launch_epi_reduce(const EpiReduceArgs* a, hipStream_t stream);
hipError_t t3_prep_inputs(int tp_is_f64, void* ovvvT, void* t2T, void* t2H, void* ovooT, double* eps_occ,
                          double* eps_vir, double* fvoP, const double* ovvv, int packed, const double* t2,
                          const double* ovoo, const double* mo_energy, const double* fock, int o, int v, int Op,
                          int V, hipStream_t s);
hipError_t t3_launch_vpart(int tp_is_f64, const VpTriple* trs, int ntriple, const VpSlab* slabs, int nslab,
                           const VpContrib* contribs, int o, int v, int Op, int V, const double* t1,
                           const double* fvoP, double* ovov_bar, double* t2h_acc, double* t1_acc, double* fvo_acc,
                           hipStream_t s);
hipError_t t3_launch_vcm(const VpTriple* trs, int ntriple, int o, int v, int V, const double* ovov,
                         const double* t2, hipStream_t s);
hipError_t t3_zero(void* p, size_t bytes, hipStream_t s);
hipError_t t3_prep_df(double* LovP, double* LvvF, const double* Lov, const double* Lvv, int naux, int o, int v,
                      int V, int NP, hipStream_t s);
hipError_t t3_finalize_df(double* Lov_bar, double* Lvv_bar, const double* LovBar, const double* LvvBar, int naux,
                          int o, int v, int V, int NP, hipStream_t s);
hipError_t t3_copy_f64(double* dst, const double* src, int n, hipStream_t s);
hipError_t t3_scale_f64(double* x, long long n, const double* scale, hipStream_t s);
hipError_t t3_noalloc_scratch(void** ptr, size_t* bytes);
size_t t3_noalloc_scratch_bytes(void);
hipError_t t3_zero_f64(double* p, long long n, hipStream_t s);
hipError_t t3_prep_projected(int tp_is_f64, void* ovvvT_P, void* t2T_Q, void* t2T_P0, void* t2H_P0, void* ovooT_Q,
                             const double* ovvvP, int packed, const double* t2Q, const double* t2P0,
                             const double* ovooQ, int o, int v, int Op, int V, hipStream_t s);
hipError_t t3_finalize_projected(double* ovvvP_bar, int packed, const double* ovvvP_acc, double* t2P0_bar,
                                 const double* t2P0hacc, double* ovooQ_bar, const double* ovooQ_acc, int o, int v,
                                 int Op, int V, hipStream_t s);
int t3_lno_max_nocc(void);
int t3_lno_matbar_blocks(void);
hipError_t t3_lno_transpose(double* out, const double* in, long long rows, long long cols, int o, int swap_r,
                            int swap_c, hipStream_t s);
hipError_t t3_lno_fock(double* out, const double* in, int o, int v, int to_fock, hipStream_t s);
hipError_t t3_lno_project(double* out, const double* in, const double* mat, long long A, int o, long long B,
                          int transK, int accumulate, hipStream_t s);
hipError_t t3_lno_matbar(double* C, double* part, const double* X, const double* Y, long long A, int o, long long B,
                         int accumulate, hipStream_t s);
hipError_t t3_finalize(double* t1_bar, double* t2_bar, double* fock_bar, double* mo_energy_bar, double* ovoo_bar,
                       double* ovvv_bar, int packed, const double* ovvv_acc, const double* t2hacc,
                       const double* ovoo_acc, const double* t1_acc, const double* fvo_acc, const double* eps_bar,
                       int o, int v, int Op, int V, hipStream_t s);
}

namespace {

constexpr int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
constexpr int GEMM_TARGET_BLOCKS = 2048;
constexpr int MAX_KSPLIT = 64;
// FP32 / bf16x3 GEMMs accumulate a split's K in FP32 registers; splits are
// also cut so that none sums more than ACC_CHUNKS K-chunks (2048 products),
// up to MAX_KSPLIT_ACC splits, before the FP64 split-order reduction
constexpr int ACC_CHUNKS = 64;
constexpr int MAX_KSPLIT_ACC = 256;
constexpr int DF_KSPLIT = 8;   // split-K of the Lov cotangent folds (K = V^2)

thread_local std::string g_last_error;
// per-family device time (ms) of the last call, recorded with hipEvents
// families 5..9 split the GEMM total (family 0) by contraction: W build
// (forward and backward recompute), then the four adjoint contractions
enum Family {
  FAM_GEMM = 0, FAM_EPILOGUE = 1, FAM_VPART = 2, FAM_REDUCE = 3, FAM_PREP = 4,
  FAM_GEMM_W = 5, FAM_GEMM_B1 = 6, FAM_GEMM_B2 = 7, FAM_GEMM_B3 = 8, FAM_GEMM_B4 = 9,
  FAM_GEMM_DF = 10,   // density-fitted mode: ovvv slab builds and cotangent folds
  FAM_COUNT = 11
};
thread_local int g_timing = 0;
thread_local double g_family_ms[FAM_COUNT];
thread_local long long g_family_launches[FAM_COUNT];

// Autotuned tile configuration per (precision, contraction family, V, Op),
// filled by the autotuner, read by every plan built afterwards.
std::mutex g_tune_mu;
std::map<std::vector<int>, int> g_tuned;

std::map<std::vector<int>, int> g_tuned_batch;   // triples per batch, same key without the family

int tuned_cfg(int prec, int fam, int V, int Op) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tuned.find({prec, fam, V, Op});
  return it == g_tuned.end() ? -1 : it->second;
}

int tuned_batch(int prec, int V, int Op) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tuned_batch.find({prec, V, Op});
  return it == g_tuned_batch.end() ? -1 : it->second;
}

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

inline long long round_up(long long x, long long m) { return (x + m - 1) / m * m; }

struct Dims {
  int o, v, n, V, Op, prec, BK;
  size_t elt;      // bytes of a contraction operand (and GEMM accumulator) element
  int tp_is_f64;
  int ntile, ncube;
  // density-fitted mode (naux > 0): ovvv slabs are built on the device from
  // the DF factors into a cache of `dfslots` block slots of `dfblk` slabs
  int naux = 0, NP = 0, dfblk = 0, dfslots = 0;
  // LNO form (occupied projector): a second cube S per triple from the
  // slot-projected operands, and their cotangents
  int lno = 0;
};

Dims make_dims(int o, int v, int prec) {
  Dims d;
  d.o = o;
  d.v = v;
  d.n = o + v;
  d.prec = prec;
  d.BK = t3_gemm_chunk(prec);
  d.V = (int)round_up(v, 32);
  d.Op = (int)round_up(o, d.BK);
  d.tp_is_f64 = prec == PREC_F64;
  d.elt = d.tp_is_f64 ? 8 : 4;
  d.ntile = d.V / EPI_T;
  d.ncube = d.ntile * (d.ntile + 1) * (d.ntile + 2) / 6;
  return d;
}

// Density-fitted mode: the occupied indices are cut into blocks of dfblk;
// the unique triples are enumerated block triple by block triple
// (I >= J >= K), so at most three blocks of ovvv slabs are needed at a time.
// A block of slabs is built when first needed and (backward) its FP64 slab
// cotangents are folded into the DF factors' cotangents when it is evicted.
// The slab cache (and, backward, its FP64 accumulators) is held to about
// 40 GB; PYSCFAD_AMD_DF_BLOCK (1..nocc) overrides the block size.
void set_df(Dims& d, int naux, bool backward, int blk_override = 0) {
  d.naux = naux;
  d.NP = (int)round_up(naux, 32);
  const double slab = (double)d.V * d.V * d.V * (double)(d.elt + (backward ? 8 : 0));
  long long blk = (long long)(40.0e9 / (3.0 * slab));
  if (const char* env = std::getenv("PYSCFAD_AMD_DF_BLOCK")) {
    const long long x = std::atoll(env);
    if (x >= 1) blk = x;
  }
  if (blk_override > 0) blk = blk_override;
  d.dfblk = (int)std::max<long long>(1, std::min<long long>(blk, d.o));
  const int nblk = (d.o + d.dfblk - 1) / d.dfblk;
  d.dfslots = std::min(3, nblk);
}

Dims dims_for(int o, int v, int prec, int naux, bool backward, int dfblk = 0, int lno = 0) {
  Dims d = make_dims(o, v, prec);
  if (naux > 0) set_df(d, naux, backward, dfblk);
  d.lno = lno;
  return d;
}

// splits a K = V^2 problem needs to keep each FP32 partial within ACC_CHUNKS
int acc_ksplit(const Dims& d) {
  const long long chunks = (long long)d.V * d.V / d.BK;
  return d.prec == PREC_F64 ? 1 : (int)std::min<long long>(MAX_KSPLIT_ACC, (chunks + ACC_CHUNKS - 1) / ACC_CHUNKS);
}

// ----------------------------------------------------------------------------
// Workspace: one device allocation carved into named regions.
struct Layout {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off = (size_t)round_up((long long)(off + bytes), 256);
    return o;
  }
};

struct Regions {
  size_t ovvvT, t2T, t2H, ovooT, eps_occ, eps_vir, fvoP, energy;
  size_t ovvv_acc, t2hacc, ovoo_acc, t1_acc, fvo_acc, eps_bar;   // (ovov_bar and most of t2_bar accumulate in place)
  size_t LovP, LvvF, LovBar, LvvBar;   // density-fitted mode only
  size_t acc_end;                      // end of the FP64 accumulators cleared at setup
  size_t dfsplit, dfsplit_bytes;       // split-K scratch of the Lov cotangent folds
  size_t slots, slot_bytes;
  size_t slot_w0, slot_H, slot_part, slot_vp;
  size_t gslots, gslot_bytes;   // G cubes: B x S of them (S batches per ovvv-adjoint group)
  int S;
  // LNO form.  Device layouts of the slot-projected operands (P: slot p of
  // ovvv; Q: slots r and q of t2, p and q of ovoo; P0: slot r of t2):
  size_t ovvvT_P, t2T_Q, t2T_P0, t2H_P0, ovooT_Q;
  // reference-layout FP64 work arrays: the canonical t1, t2, fock formed from
  // t1T, t2T, fvo, and the projected inputs
  size_t lt1, lt2, lfock, lovvvP, lt2Q, lt2P0, lovooQ;
  // backward: FP64 accumulators of the projected operands (cleared with the
  // others), reference-layout cotangents, the mat cotangent's partials
  size_t ovvvP_acc, ovooQ_acc, t2P0hacc, lt2Q_bar, lt2P0_bar;
  size_t lt1_bar, lt2_bar, lfock_bar, lovvvP_bar, lovooQ_bar, matpart;
  size_t slot_s;              // S cube of a batch slot
  size_t gsslots;             // adjoints of S, laid out like the G cubes
  size_t split, split_bytes;
  size_t desc, desc_bytes;
  size_t total;
};

// Triples per batch.  Larger batches merge the adjoint GEMM problems that
// share an output slab (consecutive triples share two occupied indices) and
// give every launch more blocks; smaller ones keep the batch's cubes (w0, G,
// H) resident in the 256 MB last-level cache while the adjoint contractions
// re-read them.  The batch's cubes are held to about 96 MB: one triple at
// v = 256, dozens for small instances.
// The autotuner may replace it with a measured choice; PYSCFAD_AMD_BATCH
// (1..64) overrides both, for experiments.  The workspace size depends on it:
// query the size after tuning.
int default_batch(const Dims& d) {
  const double cube = (double)d.V * d.V * d.V * (double)d.elt;
  const long long b = (long long)(96.0e6 / cube);
  return (int)std::max<long long>(1, std::min<long long>(b, 64));
}

int choose_batch(const Dims& d) {
  if (const char* env = std::getenv("PYSCFAD_AMD_BATCH")) {
    const long long x = std::atoll(env);
    if (x >= 1 && x <= 64) return (int)x;
  }
  const int t = tuned_batch(d.prec, d.V, d.Op);
  return t > 0 ? t : default_batch(d);
}

// Batches per ovvv-adjoint group (backward only): the ovvv adjoint of S
// consecutive batches runs as one launch, whose problems merge the triples
// that share an output slab, so the FP64 slab is read and written once per
// group instead of once per batch.  Their G cubes stay resident meanwhile;
// the extra cubes are held to about 512 MB, so that they stay in the
// last-level cache.  A batch whose cubes alone exceed that is not cache
// resident anyway: it takes groups of up to 4, with the group's cubes held to
// about 4.3 GB (at o=64/v=512 FP32, four batches of two triples: the ovvv
// adjoint went from 475 to 446 ms per 64 triples; FP64 takes two).
// PYSCFAD_AMD_B1_GROUP (1..8) overrides it, for experiments.
int b1_group(const Dims& d, int B) {
  if (const char* env = std::getenv("PYSCFAD_AMD_B1_GROUP")) {   // experiments: 1..8
    const long long x = std::atoll(env);
    if (x >= 1 && x <= 8) return (int)x;
  }
  const double cube = (double)d.V * d.V * d.V * (double)d.elt;
  if (cube * B > 512.0e6) return (int)std::max<long long>(1, std::min<long long>(4, (long long)(4.3e9 / (cube * B))));
  const long long g = (long long)(512.0e6 / (cube * B));
  return (int)std::max<long long>(1, std::min<long long>(g, 4));
}

Regions plan_regions(const Dims& d, int B, bool backward) {
  Layout L;
  Regions r{};
  r.S = backward ? b1_group(d, B) : 1;
  const size_t V = d.V, o = d.o, Op = d.Op;
  // ovvv slabs: all o of them, or the density-fitted block cache
  const size_t nslab = d.naux > 0 ? (size_t)d.dfslots * d.dfblk : o;
  r.ovvvT = L.take(nslab * V * V * V * d.elt);
  r.t2T = L.take(Op * Op * V * V * d.elt);
  r.t2H = L.take(Op * Op * V * V * d.elt);
  r.ovooT = L.take(o * o * V * Op * d.elt);
  r.eps_occ = L.take(o * 8);
  r.eps_vir = L.take(V * 8);
  r.fvoP = L.take(o * V * 8);
  r.energy = L.take(8);
  if (d.naux > 0) {
    r.LovP = L.take(o * V * d.NP * 8);
    r.LvvF = L.take(V * V * d.NP * 8);
  }
  const size_t v = d.v, n = d.n;
  if (d.lno) {
    r.ovvvT_P = L.take(o * V * V * V * d.elt);
    r.t2T_Q = L.take(Op * Op * V * V * d.elt);
    r.t2T_P0 = L.take(Op * Op * V * V * d.elt);
    r.t2H_P0 = L.take(Op * Op * V * V * d.elt);
    r.ovooT_Q = L.take(o * o * V * Op * d.elt);
    r.lt1 = L.take(o * v * 8);
    r.lt2 = L.take(o * o * v * v * 8);
    r.lfock = L.take(n * n * 8);
    r.lovvvP = L.take(o * v * v * v * 8);   // room for the dense form
    r.lt2Q = L.take(o * o * v * v * 8);
    r.lt2P0 = L.take(o * o * v * v * 8);
    r.lovooQ = L.take(o * v * o * o * 8);
  }
  if (backward) {
    r.ovvv_acc = L.take(nslab * V * V * V * 8);
    r.t2hacc = L.take(Op * Op * V * V * 8);
    r.ovoo_acc = L.take(o * o * V * Op * 8);
    r.t1_acc = L.take(o * V * 8);
    r.fvo_acc = L.take(o * V * 8);
    r.eps_bar = L.take((o + V) * 8);
    if (d.naux > 0) {
      r.LovBar = L.take(o * V * d.NP * 8);
      r.LvvBar = L.take(V * V * d.NP * 8);
    }
    if (d.lno) {
      r.ovvvP_acc = L.take(o * V * V * V * 8);
      r.ovooQ_acc = L.take(o * o * V * Op * 8);
      r.t2P0hacc = L.take(Op * Op * V * V * 8);
      r.lt2Q_bar = L.take(o * o * v * v * 8);
      r.lt2P0_bar = L.take(o * o * v * v * 8);
    }
    r.acc_end = L.off;
    if (d.lno) {
      r.lt1_bar = L.take(o * v * 8);
      r.lt2_bar = L.take(o * o * v * v * 8);
      r.lfock_bar = L.take(n * n * 8);
      r.lovvvP_bar = L.take(o * v * v * v * 8);
      r.lovooQ_bar = L.take(o * v * o * o * 8);
      r.matpart = L.take((size_t)t3_lno_matbar_blocks() * o * o * 8);
    }
  }
  Layout S;
  r.slot_w0 = S.take(V * V * V * d.elt);
  if (d.lno) r.slot_s = S.take(V * V * V * d.elt);
  if (backward && !d.lno) r.slot_H = S.take(V * V * V * d.elt);   // the LNO form's H is its G
  r.slot_vp = S.take((size_t)VP_TOTAL * V * V * 8);
  r.slot_part = S.take((size_t)d.ncube * EPI_PSTRIDE * 8);
  r.slot_bytes = S.off;
  r.slots = L.take(r.slot_bytes * B);
  r.gslot_bytes = backward ? (size_t)round_up((long long)(V * V * V * d.elt), 256) : 0;
  r.gslots = L.take(r.gslot_bytes * B * r.S);
  if (d.lno) r.gsslots = L.take(r.gslot_bytes * B * r.S);
  // split-K scratch: B2 (V x V outputs) and B3 (V x Op) problems, at most 6
  // per triple (12 in the LNO form, whose projected operands have outputs of
  // their own)
  const size_t ks = (size_t)std::max(MAX_KSPLIT, acc_ksplit(d));
  const size_t npt = d.lno ? 12 : 6;
  r.split_bytes = backward ? npt * B * ks * V * std::max(V, Op) * d.elt + npt * B * 256 : 0;
  r.split = L.take(r.split_bytes);
  r.dfsplit_bytes = (backward && d.naux > 0) ? (size_t)d.dfblk * DF_KSPLIT * V * d.NP * 8 + (size_t)d.dfblk * 256 : 0;
  r.dfsplit = L.take(r.dfsplit_bytes);
  r.desc = L.off;
  return r;
}

// ----------------------------------------------------------------------------
// Launch plan
enum LaunchKind { L_GEMM, L_REDUCE, L_EPI, L_EPIRED, L_VPART, L_VMN, L_ZERO };

struct Launch {
  LaunchKind kind;
  int cfg = 0;
  size_t off_a = 0, off_b = 0, off_c = 0;   // descriptor offsets in the blob
  int n_a = 0, n_b = 0, n_c = 0;
  long long work = 0;                        // GEMM blocks / reduce elements
  int aux = -1;                              // index of host-side argument struct
  int sub = -1;                              // GEMM contraction family (FAM_GEMM_*)
  int prec = -1;                             // GEMM / reduce arithmetic (-1: the plan's)
  // L_ZERO: clears `work` bytes at workspace offset off_a
};

struct Blob {
  std::vector<char> data;
  template <class T>
  size_t put(const std::vector<T>& v) {
    size_t off = (size_t)round_up((long long)data.size(), 256);
    data.resize(off + std::max<size_t>(v.size() * sizeof(T), 1));
    if (!v.empty()) std::memcpy(data.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

struct Triple {
  int t[3];
  double fo;
};

std::vector<Triple> enumerate_triples(int o) {
  std::vector<Triple> out;
  for (int i = 0; i < o; i++)
    for (int j = 0; j <= i; j++)
      for (int k = 0; k <= j; k++) {
        Triple tr;
        tr.t[0] = i; tr.t[1] = j; tr.t[2] = k;
        tr.fo = (i == k) ? 1.0 / 6.0 : ((i == j || j == k) ? 0.5 : 1.0);
        out.push_back(tr);
      }
  return out;
}

// The unique triples i >= j >= k with i in block I, j in block J, k in block
// K of `blk` occupied indices (I >= J >= K), in the order of
// enumerate_triples.  Concatenated over I, J, K in nested ascending order they
// list every unique triple once.
std::vector<Triple> block_triples(int o, int blk, int I, int J, int K) {
  std::vector<Triple> out;
  const int i1 = std::min(o, (I + 1) * blk), j1 = std::min(o, (J + 1) * blk), k1 = std::min(o, (K + 1) * blk);
  for (int i = I * blk; i < i1; i++)
    for (int j = J * blk; j < j1 && j <= i; j++)
      for (int k = K * blk; k < k1 && k <= j; k++) {
        Triple tr;
        tr.t[0] = i; tr.t[1] = j; tr.t[2] = k;
        tr.fo = (i == k) ? 1.0 / 6.0 : ((i == j || j == k) ? 0.5 : 1.0);
        out.push_back(tr);
      }
  return out;
}

// Grouped GEMM under construction: problems keyed by their output slab.
struct GemmBuild {
  int prec, BK;
  std::map<long long, int> index;
  std::vector<GemmProb> probs;
  std::vector<std::vector<GemmSeg>> segs;

  explicit GemmBuild(int prec_) : prec(prec_), BK(t3_gemm_chunk(prec_)) {}
  int problem(long long key, const GemmProb& tmpl) {
    auto it = index.find(key);
    if (it != index.end()) return it->second;
    int p = (int)probs.size();
    index[key] = p;
    probs.push_back(tmpl);
    segs.emplace_back();
    return p;
  }
  void add(int p, const GemmSeg& s) { segs[p].push_back(s); }
  bool empty() const { return probs.empty(); }
};

struct Ctx {
  Dims d;
  Regions r;
  char* ws = nullptr;   // device workspace base (may be null while sizing)
  int B = 1;
  bool backward = false;
  double et_bar = 1.0;
  long long t_begin = 0, t_end = -1;   // slice of the unique-triple list (-1: to the end)
  // inputs (device, reference layouts)
  const double *mo_energy = nullptr, *t1 = nullptr, *t2 = nullptr, *fock = nullptr, *ovoo = nullptr,
               *ovov = nullptr, *ovvv = nullptr;
  int packed = 1;
  // cotangent outputs accumulated in place by the plan (device, reference layouts)
  double *t2_out = nullptr, *ovov_out = nullptr;
  // LNO form: the reference-layout cotangents of the projected t2 operands
  double *t2Q_out = nullptr, *t2P0_out = nullptr;
  char* outp(double* base, size_t elems) const { return reinterpret_cast<char*>(base) + elems * 8; }
  Blob blob;
  std::vector<Launch> launches;
  std::vector<EpiArgs> epi_args;
  size_t cube_tab = (size_t)-1;   // blob offset of the epilogue's cube table (one per plan)
  std::vector<EpiReduceArgs> epired_args;
  std::string plan_error;   // set when a contraction does not fit the GEMM's operand modes
  int force_cfg[FAM_COUNT] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};   // autotuner trials

  // tile configuration of a contraction: a trial's, else the one
  // PYSCFAD_AMD_TILE names (every family, for experiments and tests; ignored
  // where the precision has no such tile), else the tuned one, else the default
  int cfg_for(int fam, int dflt) const {
    if (force_cfg[fam] >= 0) return force_cfg[fam];
    if (const char* env = std::getenv("PYSCFAD_AMD_TILE")) {
      const int x = std::atoi(env);
      if (env[0] >= '0' && env[0] <= '9' && t3_gemm_cfg_valid(d.prec, x)) return x;
    }
    const int t = tuned_cfg(d.prec, fam, d.V, d.Op);
    return t >= 0 ? t : dflt;
  }

  template <class T = char>
  T* at(size_t off) const { return reinterpret_cast<T*>(ws + off); }
  char* slot(int s, size_t field) const { return ws + r.slots + (size_t)s * r.slot_bytes + field; }
  int gbase = 0;   // first G cube of the current batch within its ovvv-adjoint group
  char* gslot(int s) const { return ws + r.gslots + (size_t)(gbase + s) * r.gslot_bytes; }
  char* gsslot(int s) const { return ws + r.gsslots + (size_t)(gbase + s) * r.gslot_bytes; }
  // adjoint cube read by a backward contraction: G, or (LNO form) the adjoint of S
  char* gcube(int which, int s) const { return which ? gsslot(s) : gslot(s); }
  // the H cube of a batch slot (the LNO form's H is its G)
  char* hcube(int s) const { return d.lno ? gslot(s) : slot(s, r.slot_H); }
  long long str(int pos) const {
    return pos == 0 ? (long long)d.V * d.V : (pos == 1 ? (long long)d.V : 1LL);
  }
  const char* el(size_t region, long long elems) const { return ws + region + elems * (long long)d.elt; }
  const double* dl(size_t region, long long elems) const {
    return reinterpret_cast<const double*>(ws + region) + elems;
  }

  // Defaults: 128x128 tiles for large FP32 / bf16x3 problems; FP64 stays on
  // 64x64 (its 128x128 variant would spill).
  int pick_cfg(const GemmBuild& gb) const {
    if (gb.prec == PREC_F64) return 0;
    for (const auto& P : gb.probs)
      if (P.M < 128 || P.N < 128) return 0;
    return 1;
  }

  // Flatten a GemmBuild into launches.
  void emit_gemm(GemmBuild& gb, int cfg, int mode, int c_is_f64, bool allow_split, int sub,
                 int max_ksplit = MAX_KSPLIT, size_t scratch_region = (size_t)-1, size_t scratch_bytes = 0) {
    if (gb.empty()) return;
    if (cfg < 0) cfg = pick_cfg(gb);
    // Unsplit FP32 / bf16x3 adjoint launches: a problem that merges many
    // segments (all the triples of a batch or group that write one output)
    // would sum all of them in FP32 registers.  Whole segments are packed into
    // successive launches of at most ACC_CHUNKS K-chunks per problem; the
    // later launches accumulate into the same output in stream order.  (A W
    // build problem is one triple's cube: at most 4 particle and 4 hole
    // segments per pass.)
    if (!allow_split && gb.prec != PREC_F64 && sub != FAM_GEMM_W) {
      std::vector<std::vector<std::vector<GemmSeg>>> passes;   // [pass][problem] -> segments
      for (size_t p = 0; p < gb.probs.size(); p++) {
        int pass = 0, chunks = 0;
        for (const auto& sg : gb.segs[p]) {
          const int c = (int)((long long)sg.K1 * sg.K2 / gb.BK);
          if (chunks > 0 && chunks + c > ACC_CHUNKS) { pass++; chunks = 0; }
          if ((int)passes.size() <= pass) passes.resize(pass + 1, std::vector<std::vector<GemmSeg>>(gb.probs.size()));
          passes[pass][p].push_back(sg);
          chunks += c;
        }
      }
      if (passes.size() > 1) {
        for (size_t k = 0; k < passes.size(); k++) {
          GemmBuild part(gb.prec);
          for (size_t p = 0; p < gb.probs.size(); p++) {
            if (passes[k][p].empty()) continue;
            const int q = part.problem((long long)p, gb.probs[p]);
            for (const auto& sg : passes[k][p]) part.add(q, sg);
          }
          emit_gemm_launch(part, cfg, k == 0 ? mode : GEMM_ACCUM, c_is_f64, false, sub);
        }
        return;
      }
    }
    emit_gemm_launch(gb, cfg, mode, c_is_f64, allow_split, sub, max_ksplit, scratch_region, scratch_bytes);
  }

  void emit_gemm_launch(GemmBuild& gb, int cfg, int mode, int c_is_f64, bool allow_split, int sub,
                        int max_ksplit = MAX_KSPLIT, size_t scratch_region = (size_t)-1, size_t scratch_bytes = 0) {
    if (scratch_region == (size_t)-1) {
      scratch_region = r.split;
      scratch_bytes = r.split_bytes;
    }
    const size_t elt = gb.prec == PREC_F64 ? 8 : 4;   // operand and accumulator bytes
    const int BM = t3_gemm_tile_m(cfg), BN = t3_gemm_tile_n(cfg);
    long long tiles = 0;
    int min_chunks = 1 << 30;
    for (size_t p = 0; p < gb.probs.size(); p++) {
      GemmProb& P = gb.probs[p];
      if (P.Mv <= 0) P.Mv = P.M;
      if (P.Nv <= 0) P.Nv = P.N;
      P.tiles_m = (P.M + BM - 1) / BM;
      P.tiles_n = (P.N + BN - 1) / BN;
      tiles += (long long)P.nb * P.tiles_m * P.tiles_n;
      int c = 0;
      for (auto& s : gb.segs[p]) {
        s.chunk0 = c;
        c += (int)((long long)s.K1 * s.K2 / gb.BK);
      }
      P.nchunks = c;
      min_chunks = std::min(min_chunks, c);
    }
    int ksplit = 1;
    if (allow_split) {
      ksplit = (int)std::max<long long>(1, (GEMM_TARGET_BLOCKS + tiles - 1) / tiles);
      ksplit = std::min(ksplit, max_ksplit);
      if (gb.prec != PREC_F64) {   // bound the FP32 accumulation length (ACC_CHUNKS)
        int max_chunks = 0;
        for (const auto& P : gb.probs) max_chunks = std::max(max_chunks, P.nchunks);
        ksplit = std::max(ksplit, std::min(MAX_KSPLIT_ACC, (max_chunks + ACC_CHUNKS - 1) / ACC_CHUNKS));
      }
      ksplit = std::min(ksplit, std::max(1, min_chunks));
      // the partials of every problem must fit the scratch region (problems
      // merge all the triples of a batch that share an output, so their K,
      // and the split the accumulation bound asks for, vary)
      size_t per_split = 0;
      for (const auto& P : gb.probs) per_split += (size_t)P.nb * P.M * P.N * elt;
      const size_t room = scratch_bytes > 256 * gb.probs.size() ? scratch_bytes - 256 * gb.probs.size() : 0;
      const long long cap = per_split ? (long long)(room / per_split) : ksplit;
      ksplit = (int)std::max<long long>(1, std::min<long long>(ksplit, cap));
    }
    // Problems merge all the triples of a batch that write one output, so
    // their K differ: each problem is split into about as many chunks per
    // split as the largest one gets, so that the blocks carry equal work (a
    // uniform split count would leave the long problems' blocks as the tail).
    // The hole-side adjoints keep one split count when all their problems
    // share one tile grid: their tiles are then interleaved (below), which
    // measured worth more than the balance.
    int max_chunks_all = 1;
    for (const auto& P : gb.probs) max_chunks_all = std::max(max_chunks_all, P.nchunks);
    bool one_grid = (sub == FAM_GEMM_B3 || sub == FAM_GEMM_B4) && gb.probs.size() > 1;
    for (const auto& P : gb.probs)
      one_grid = one_grid && P.nb == gb.probs[0].nb && P.tiles_m == gb.probs[0].tiles_m &&
                 P.tiles_n == gb.probs[0].tiles_n;
    const int cps = one_grid ? max_chunks_all : (max_chunks_all + ksplit - 1) / ksplit;
    // operand load modes; every base and stride must keep VEC-element vectors aligned
    const long long vec = t3_gemm_vec(gb.prec);
    const long long vbytes = vec * (long long)elt;
    auto aligned = [&](const void* p) { return ((uintptr_t)p % (uintptr_t)vbytes) == 0; };
    for (size_t p = 0; p < gb.probs.size(); p++) {
      const GemmProb& P = gb.probs[p];
      for (auto& s : gb.segs[p]) {
        s.a_mode = s.a_sm == 1 ? LOAD_ROWCONT : LOAD_KCONT;
        s.b_mode = s.b_sn == 1 ? LOAD_ROWCONT : LOAD_KCONT;
        const long long a_row = s.a_mode == LOAD_ROWCONT ? s.a_sk2 : s.a_sm;
        const long long b_row = s.b_mode == LOAD_ROWCONT ? s.b_sk2 : s.b_sn;
        const bool ok = (s.a_mode == LOAD_ROWCONT ? P.M % vec == 0 : s.a_sk2 == 1) &&
                        (s.b_mode == LOAD_ROWCONT ? P.N % vec == 0 : s.b_sk2 == 1) &&
                        a_row % vec == 0 && b_row % vec == 0 && s.a_sb % vec == 0 && s.b_sb % vec == 0 &&
                        s.a_sk1 % vec == 0 && s.b_sk1 % vec == 0 && s.K2 % gb.BK == 0 &&
                        aligned(s.A) && aligned(s.B) &&
                        // per-thread tile offsets are 32-bit byte offsets
                        ((BM - 1) * s.a_sm + (gb.BK - 1) * s.a_sk2 + vec) * (long long)elt < (1LL << 32) &&
                        ((BN - 1) * s.b_sn + (gb.BK - 1) * s.b_sk2 + vec) * (long long)elt < (1LL << 32);
        if (!ok && plan_error.empty()) plan_error = "GEMM operand is neither row- nor k-contiguous, or misaligned";
        if (s.scale != 1.0f && s.scale != -1.0f && plan_error.empty()) plan_error = "GEMM segment scale must be +1 or -1";
        // the kernel's chunk and tile index arithmetic is 32-bit
        if ((long long)s.K1 * s.K2 >= (1LL << 31) && plan_error.empty()) plan_error = "GEMM segment K exceeds 2^31";
      }
    }
    std::vector<GemmSeg> flat;
    std::vector<ReduceProb> red;
    long long nblocks = 0, relem = 0;
    size_t scratch = 0;
    for (size_t p = 0; p < gb.probs.size(); p++) {
      GemmProb& P = gb.probs[p];
      P.seg0 = (int)flat.size();
      P.nseg = (int)gb.segs[p].size();
      for (auto& s : gb.segs[p]) flat.push_back(s);
      P.c_is_f64 = c_is_f64;
      P.ksplit = !allow_split ? 1 : one_grid ? ksplit : std::max(1, std::min(ksplit, (P.nchunks + cps - 1) / cps));
      P.block0 = nblocks;
      nblocks += (long long)P.nb * P.tiles_m * P.tiles_n * P.ksplit;
      if (nblocks >= (1LL << 31) && plan_error.empty()) plan_error = "GEMM launch exceeds 2^31 tiles";
      if (allow_split) {
        const size_t bytes = (size_t)P.ksplit * P.nb * P.M * P.N * elt;
        P.mode = GEMM_PARTIAL;
        P.P = ws ? ws + scratch_region + scratch : nullptr;
        ReduceProb R{};
        R.P = P.P;
        R.ksplit = P.ksplit;
        R.nb = P.nb;
        R.M = P.M;
        R.N = P.N;
        R.Mv = P.Mv;
        R.Nv = P.Nv;
        R.C = P.C;
        R.c_sb = P.c_sb;
        R.c_sm = P.c_sm;
        R.c_sn = P.c_sn;
        R.accumulate = mode == GEMM_ACCUM;
        R.c_is_f64 = c_is_f64;
        R.alpha = P.alpha;
        R.elem0 = relem;
        relem += (long long)P.nb * P.M * P.N;
        red.push_back(R);
        scratch = (size_t)round_up((long long)(scratch + bytes), 256);
        if (scratch > scratch_bytes && plan_error.empty()) plan_error = "split-K partials exceed their scratch region";
      } else {
        P.mode = mode;
        P.P = nullptr;
      }
    }
    // The hole-side adjoints read the G cube once per permutation, and the
    // problems of a permutation pair read it through the same view: with one
    // tile grid for all problems, their tiles are interleaved, so both
    // readers of a G block run on one XCD at about the same time.
    bool inter = (sub == FAM_GEMM_B3 || sub == FAM_GEMM_B4) && gb.probs.size() > 1;
    for (const auto& P : gb.probs) {
      const GemmProb& Q = gb.probs[0];
      inter = inter && P.nb == Q.nb && P.tiles_m == Q.tiles_m && P.tiles_n == Q.tiles_n && P.ksplit == Q.ksplit;
    }
    for (auto& P : gb.probs) P.interleave = inter ? 1 : 0;
    Launch L;
    L.kind = L_GEMM;
    L.cfg = cfg;
    L.off_a = blob.put(gb.probs);
    L.n_a = (int)gb.probs.size();
    L.off_b = blob.put(flat);
    L.n_b = (int)flat.size();
    L.work = nblocks;
    L.sub = sub;
    L.prec = gb.prec;
    launches.push_back(L);
    if (allow_split) {
      Launch R;
      R.kind = L_REDUCE;
      R.prec = gb.prec;
      // vector partial loads: every N a multiple of the vector width, 32-bit element indices
      const int ve = gb.prec == PREC_F64 ? 2 : 4;
      bool vec = true;
      for (const auto& Q : red) vec = vec && Q.N % ve == 0 && (long long)Q.nb * Q.M * Q.N < (1LL << 31);
      R.cfg = vec ? 1 : 0;
      R.off_a = blob.put(red);
      R.n_a = (int)red.size();
      R.work = relem;
      launches.push_back(R);
    }
  }

  // ---- ovvv slabs ----
  // Slab of occupied index x in r.ovvvT (and its FP64 cotangent accumulator in
  // r.ovvv_acc): all o are resident in the dense mode; the density-fitted mode
  // maps x through the block slots of the cache.
  std::vector<int> df_pos;        // x -> slab position in the cache, -1 if not resident
  int df_res[3] = {-1, -1, -1};   // block held by each cache slot
  long long slab_pos(int x) const { return d.naux > 0 ? (long long)df_pos[x] : (long long)x; }

  void emit_zero(size_t off, size_t bytes) {
    Launch L;
    L.kind = L_ZERO;
    L.off_a = off;
    L.work = (long long)bytes;
    L.sub = FAM_GEMM_DF;
    launches.push_back(L);
  }

  // Build the slabs of block blk into cache slot `slot`, in FP64 arithmetic
  // (rounded to the contraction precision on store, as the dense mode rounds
  // its ovvv input):  ovvvT[x][a][y][f] = sum_P LovP[x][a][P] LvvF[y][f][P],
  // one problem per index (batch y, rows a, columns f, k = P).
  void emit_df_build(int blk, int slot) {
    const long long V = d.V, NP = d.NP;
    const int x0 = blk * d.dfblk, x1 = std::min(d.o, x0 + d.dfblk);
    GemmBuild gb(PREC_F64);
    for (int x = x0; x < x1; x++) {
      GemmProb P{};
      P.nb = d.V; P.M = d.V; P.N = d.V;
      P.C = ws + r.ovvvT + (size_t)((long long)slot * d.dfblk + (x - x0)) * V * V * V * d.elt;
      P.c_sb = V; P.c_sm = V * V; P.c_sn = 1; P.alpha = 1.0;
      const int q = gb.problem(x, P);
      GemmSeg sg{};
      sg.A = ws + r.LovP + (size_t)x * V * NP * 8;
      sg.a_sb = 0; sg.a_sm = NP; sg.a_sk2 = 1;
      sg.B = ws + r.LvvF;
      sg.b_sb = V * NP; sg.b_sn = NP; sg.b_sk2 = 1;
      sg.K1 = 1; sg.K2 = d.NP; sg.scale = 1.0f;
      gb.add(q, sg);
    }
    emit_gemm(gb, 0, GEMM_STORE, d.tp_is_f64, false, FAM_GEMM_DF);
  }

  // Fold the FP64 slab cotangents Sbar of the block in cache slot `slot` into
  // the DF factors' cotangents (the chain rule through the slab build):
  //   LovBar[x][a][P] += sum_{y,f} Sbar_x[a][y][f] LvvF[y][f][P]   (per x, split-K over (y, f))
  //   LvvBar[y][f][P] += sum_x sum_a Sbar_x[a][y][f] LovP[x][a][P] (one problem, one segment per x)
  void emit_df_fold(int blk, int slot) {
    const long long V = d.V, NP = d.NP;
    const int x0 = blk * d.dfblk, x1 = std::min(d.o, x0 + d.dfblk);
    auto sbar = [&](int x) {
      return ws + r.ovvv_acc + (size_t)((long long)slot * d.dfblk + (x - x0)) * V * V * V * 8;
    };
    {
      GemmBuild gb(PREC_F64);
      for (int x = x0; x < x1; x++) {
        GemmProb P{};
        P.nb = 1; P.M = d.V; P.N = d.NP;
        P.C = ws + r.LovBar + (size_t)x * V * NP * 8;
        P.c_sm = NP; P.c_sn = 1; P.alpha = 1.0;
        const int q = gb.problem(x, P);
        GemmSeg sg{};
        sg.A = sbar(x); sg.a_sm = V * V; sg.a_sk1 = V; sg.a_sk2 = 1;
        sg.B = ws + r.LvvF; sg.b_sk1 = V * NP; sg.b_sk2 = NP; sg.b_sn = 1;
        sg.K1 = d.V; sg.K2 = d.V; sg.scale = 1.0f;
        gb.add(q, sg);
      }
      emit_gemm(gb, 0, GEMM_ACCUM, 1, true, FAM_GEMM_DF, DF_KSPLIT, r.dfsplit, r.dfsplit_bytes);
    }
    {
      GemmBuild gb(PREC_F64);
      GemmProb P{};
      P.nb = d.V; P.M = d.V; P.N = d.NP;
      P.C = ws + r.LvvBar;
      P.c_sb = V * NP; P.c_sm = NP; P.c_sn = 1; P.alpha = 1.0;
      const int q = gb.problem(0, P);
      for (int x = x0; x < x1; x++) {
        GemmSeg sg{};
        sg.A = sbar(x); sg.a_sb = V; sg.a_sm = 1; sg.a_sk2 = V * V;
        sg.B = ws + r.LovP + (size_t)x * V * NP * 8; sg.b_sb = 0; sg.b_sk2 = NP; sg.b_sn = 1;
        sg.K1 = 1; sg.K2 = d.V; sg.scale = 1.0f;
        gb.add(q, sg);
      }
      emit_gemm(gb, 0, GEMM_ACCUM, 1, false, FAM_GEMM_DF);
    }
  }

  // Make blocks need[0..2] resident: each missing block takes a slot holding
  // no needed block (one always exists: at most three distinct blocks are
  // needed and there are min(3, number of blocks) slots); the evicted block's
  // cotangents are folded first, and the slot's accumulators cleared after.
  void df_ensure(const int need[3]) {
    const size_t acc_slot = (size_t)d.dfblk * d.V * d.V * d.V * 8;
    for (int t = 0; t < 3; t++) {
      bool have = false;
      for (int u = 0; u < d.dfslots; u++) have = have || df_res[u] == need[t];
      if (have) continue;
      int slot = -1;
      for (int u = 0; u < d.dfslots && slot < 0; u++)
        if (df_res[u] != need[0] && df_res[u] != need[1] && df_res[u] != need[2]) slot = u;
      if (slot < 0) {
        if (plan_error.empty()) plan_error = "density-fitted slab cache has no free slot";
        return;
      }
      if (df_res[slot] >= 0) {
        if (backward) emit_df_fold(df_res[slot], slot);
        set_block_pos(df_res[slot], -1);
      }
      emit_df_build(need[t], slot);
      if (backward) emit_zero(r.ovvv_acc + (size_t)slot * acc_slot, acc_slot);
      df_res[slot] = need[t];
      set_block_pos(need[t], slot);
    }
  }
  void set_block_pos(int blk, int slot) {
    const int x0 = blk * d.dfblk, x1 = std::min(d.o, x0 + d.dfblk);
    for (int x = x0; x < x1; x++) df_pos[x] = slot < 0 ? -1 : slot * d.dfblk + (x - x0);
  }

  // ---- W-build segments (output layout: batch position ib, row position im, col position 2) ----
  // ovvv_reg / t2_reg: the operand regions (the plain ones, or for the LNO
  // form's S their slot-projected forms)
  GemmSeg particle_seg(const Triple& tr, int rho, int ib, int im, size_t ovvv_reg, size_t t2_reg) const {
    const int* p = PERM[rho];
    const long long V = d.V, Op = d.Op;
    GemmSeg s{};
    const char* L = el(ovvv_reg, slab_pos(tr.t[p[0]]) * V * V * V);
    const char* R = el(t2_reg, ((long long)tr.t[p[2]] * Op + tr.t[p[1]]) * V * V);
    if (p[2] == 2) {
      s.A = L; s.a_sb = (p[0] == ib) ? V * V : V; s.a_sm = (p[0] == im) ? V * V : V; s.a_sk2 = 1;
      s.B = R; s.b_sb = 0; s.b_sn = V; s.b_sk2 = 1;
    } else {   // p[2] == im
      s.A = R; s.a_sb = 0; s.a_sm = V; s.a_sk2 = 1;
      s.B = L; s.b_sb = (p[0] == ib) ? V * V : V; s.b_sn = (p[0] == 2) ? V * V : V; s.b_sk2 = 1;
    }
    s.K1 = 1; s.K2 = d.V; s.scale = 1.0f;
    return s;
  }
  GemmSeg hole_seg(const Triple& tr, int rho, int ib, int im, size_t ovoo_reg, size_t t2_reg, size_t t2h_reg) const {
    const int* p = PERM[rho];
    const long long V = d.V, Op = d.Op;
    GemmSeg s{};
    const char* Lp = el(ovoo_reg, ((long long)tr.t[p[0]] * d.o + tr.t[p[1]]) * V * Op);
    const char* T2 = el(t2_reg, (long long)tr.t[p[2]] * Op * V * V);   // [m][s_p2][s_p1]
    const char* T2H = el(t2h_reg, (long long)tr.t[p[2]] * Op * V * V);  // [m][s_p1][s_p2]
    if (p[0] == 2) {
      s.B = Lp; s.b_sb = 0; s.b_sn = Op; s.b_sk2 = 1;
      s.A = (p[1] == im) ? T2 : T2H;
      s.a_sm = 1; s.a_sb = V; s.a_sk2 = V * V;
    } else {   // p[0] == im
      s.A = Lp; s.a_sb = 0; s.a_sm = Op; s.a_sk2 = 1;
      s.B = (p[1] == 2) ? T2 : T2H;
      s.b_sn = 1; s.b_sb = V; s.b_sk2 = V * V;
    }
    s.K1 = 1; s.K2 = d.Op; s.scale = -1.0f;
    return s;
  }

  // Operand sets of a cube build: particle terms (ovvv, t2) and hole terms
  // (ovoo, t2, t2H), as region offsets; every set contributes the W terms of
  // ccsd_t.c:137-142 with its operands.
  struct CubeTerms {
    std::vector<std::array<size_t, 2>> part;
    std::vector<std::array<size_t, 3>> hole;
  };

  void emit_cube_build(const std::vector<Triple>& bt, size_t field, const CubeTerms& ct) {
    const long long V = d.V;
    static const int P1[4] = {0, 2, 3, 5}, H1[4] = {0, 1, 4, 5};   // batch b (pos 1), rows a
    static const int P2[2] = {1, 4}, H2[2] = {2, 3};               // batch a (pos 0), rows b
    for (int pass = 0; pass < 2; pass++) {
      GemmBuild gb(d.prec);
      for (size_t s = 0; s < bt.size(); s++) {
        GemmProb P{};
        P.nb = d.V; P.M = d.V; P.N = d.V;
        P.C = slot((int)s, field);
        P.c_sn = 1; P.alpha = 1.0;
        if (pass == 0) { P.c_sb = V; P.c_sm = V * V; }
        else { P.c_sb = V * V; P.c_sm = V; }
        const int p = gb.problem((long long)s, P);
        const int ib = pass == 0 ? 1 : 0, im = pass == 0 ? 0 : 1;
        const int np = pass == 0 ? 4 : 2;
        for (const auto& t : ct.part)
          for (int q = 0; q < np; q++) gb.add(p, particle_seg(bt[s], pass == 0 ? P1[q] : P2[q], ib, im, t[0], t[1]));
        for (const auto& t : ct.hole)
          for (int q = 0; q < np; q++)
            gb.add(p, hole_seg(bt[s], pass == 0 ? H1[q] : H2[q], ib, im, t[0], t[1], t[2]));
      }
      emit_gemm(gb, cfg_for(FAM_GEMM_W, pick_cfg(gb)), pass == 0 ? GEMM_STORE : GEMM_ACCUM, d.tp_is_f64, false,
                FAM_GEMM_W);
    }
  }

  void emit_wbuild(const std::vector<Triple>& bt) {
    emit_cube_build(bt, r.slot_w0, CubeTerms{{{r.ovvvT, r.t2T}}, {{r.ovooT, r.t2T, r.t2H}}});
  }

  // LNO form: S = (P1 + P2 + P3) w0, each occupied slot of every term in turn
  // projected by mat.  A particle term sum_f ovvv[p,x,f,y] t2[r,q,z,f] has
  // slot p on ovvv and q, r on t2; a hole term sum_m ovoo[p,x,m,q] t2[r,m,z,y]
  // has p, q on ovoo and r on t2.  So S is the W build over the operand sets
  // (P ovvv, t2) + (ovvv, Q t2) and (Q ovoo, t2) + (ovoo, P0 t2).
  void emit_sbuild(const std::vector<Triple>& bt) {
    emit_cube_build(bt, r.slot_s,
                    CubeTerms{{{r.ovvvT_P, r.t2T}, {r.ovvvT, r.t2T_Q}},
                              {{r.ovooT_Q, r.t2T, r.t2H}, {r.ovooT, r.t2T_P0, r.t2H_P0}}});
  }

  void emit_epilogue(const std::vector<Triple>& bt, bool bw) {
    std::vector<EpiTriple> et;
    for (size_t s = 0; s < bt.size(); s++) {
      EpiTriple e{};
      e.i = bt[s].t[0]; e.j = bt[s].t[1]; e.k = bt[s].t[2];
      e.fo = bt[s].fo;
      e.w0 = slot((int)s, r.slot_w0);
      e.G = bw ? gslot((int)s) : nullptr;
      e.H = (bw && !d.lno) ? slot((int)s, r.slot_H) : nullptr;
      e.s = d.lno ? slot((int)s, r.slot_s) : nullptr;
      e.GS = (bw && d.lno) ? gsslot((int)s) : nullptr;
      e.part = reinterpret_cast<double*>(slot((int)s, r.slot_part));
      e.cm = reinterpret_cast<const double*>(slot((int)s, r.slot_vp)) + (size_t)VP_CM * d.V * d.V;
      et.push_back(e);
    }
    // compact V-term operands of the batch (and the backward's CM/CN)
    std::vector<VpTriple> vt;
    for (size_t s = 0; s < bt.size(); s++) {
      VpTriple t{};
      t.i = bt[s].t[0]; t.j = bt[s].t[1]; t.k = bt[s].t[2];
      t.H = bw ? hcube((int)s) : nullptr;
      t.out = reinterpret_cast<double*>(slot((int)s, r.slot_vp));
      vt.push_back(t);
    }
    Launch M;
    M.kind = L_VMN;
    M.off_a = blob.put(vt);
    M.n_a = (int)vt.size();
    M.aux = bw ? 1 : 0;
    launches.push_back(M);
    if (cube_tab == (size_t)-1) {
      // tile triple of every cube id, in the order cube = A(A+1)(A+2)/6 + B(B+1)/2 + C, A >= B >= C;
      // 10 bits a tile (nvir < 4096 keeps ntile <= 512)
      if (d.ntile > 1023) plan_error = "cube table: more than 1023 tiles per virtual edge";
      std::vector<int> tab;
      tab.reserve(d.ncube);
      for (int A = 0; A < d.ntile; A++)
        for (int B = 0; B <= A; B++)
          for (int C = 0; C <= B; C++) tab.push_back(A | B << 10 | C << 20);
      cube_tab = blob.put(tab);
    }
    EpiArgs a{};
    a.nocc = d.o; a.nvir = d.v; a.V = d.V; a.ntile = d.ntile; a.ncube = d.ncube;
    a.backward = bw; a.lno = d.lno; a.et_bar = et_bar;
    a.eps_occ = dl(r.eps_occ, 0); a.eps_vir = dl(r.eps_vir, 0);
    a.t1 = t1; a.fvoP = dl(r.fvoP, 0);
    epi_args.push_back(a);
    Launch L;
    L.kind = L_EPI;
    L.off_a = blob.put(et);
    L.n_a = (int)et.size();
    L.aux = (int)epi_args.size() - 1;
    launches.push_back(L);
    EpiReduceArgs ra{};
    ra.ntriple = (int)et.size(); ra.ncube = d.ncube; ra.ntile = d.ntile; ra.nocc = d.o; ra.V = d.V;
    ra.triples = nullptr;   // patched at execution: descriptor offset L.off_a
    ra.energy = reinterpret_cast<double*>(ws ? ws + r.energy : nullptr);
    ra.eps_bar = bw ? reinterpret_cast<double*>(ws + r.eps_bar) : nullptr;
    ra.backward = bw;
    epired_args.push_back(ra);
    Launch R;
    R.kind = L_EPIRED;
    R.off_a = L.off_a;
    R.aux = (int)epired_args.size() - 1;
    launches.push_back(R);
  }

  // ---- backward contractions over G ----
  // Each family runs over terms (adjoint cube, operand regions, output): the
  // canonical term reads G with the plain operands.  In the LNO form two
  // terms per family read the adjoint of S: one with the other factor's
  // projected operand into the plain output (same problems, more K
  // segments), one with the plain operand into the projected operand's own
  // output (more problems, keyed apart).
  struct B1Term { int g; size_t t2; size_t out; long long key; };
  struct B2Term { int g; size_t ovvv; double* out; long long key; };
  struct B3Term { int g; size_t t2, t2h, out; long long key; };
  struct B4Term { int g; size_t ovoo; double* out; size_t hacc; long long key; };
  static constexpr long long KEY_PROJ = 1LL << 40;

  std::vector<B1Term> b1_terms() const {
    std::vector<B1Term> t{{0, r.t2T, r.ovvv_acc, 0}};
    if (d.lno) {
      t.push_back({1, r.t2T_Q, r.ovvv_acc, 0});
      t.push_back({1, r.t2T, r.ovvvP_acc, KEY_PROJ});
    }
    return t;
  }
  std::vector<B2Term> b2_terms() const {
    std::vector<B2Term> t{{0, r.ovvvT, t2_out, 0}};
    if (d.lno) {
      t.push_back({1, r.ovvvT_P, t2_out, 0});
      t.push_back({1, r.ovvvT, t2Q_out, KEY_PROJ});
    }
    return t;
  }
  std::vector<B3Term> b3_terms() const {
    std::vector<B3Term> t{{0, r.t2T, r.t2H, r.ovoo_acc, 0}};
    if (d.lno) {
      t.push_back({1, r.t2T_P0, r.t2H_P0, r.ovoo_acc, 0});
      t.push_back({1, r.t2T, r.t2H, r.ovooQ_acc, KEY_PROJ});
    }
    return t;
  }
  std::vector<B4Term> b4_terms() const {
    std::vector<B4Term> t{{0, r.ovooT, t2_out, r.t2hacc, 0}};
    if (d.lno) {
      t.push_back({1, r.ovooT_Q, t2_out, r.t2hacc, 0});
      t.push_back({1, r.ovooT, t2P0_out, r.t2P0hacc, KEY_PROJ});
    }
    return t;
  }

  void emit_b1(const std::vector<Triple>& bt) {
    const long long V = d.V, Op = d.Op;
    const std::vector<B1Term> terms = b1_terms();
    for (int orient = 0; orient < 2; orient++) {
      GemmBuild gb(d.prec);
      for (const B1Term& T : terms)
        for (size_t s = 0; s < bt.size(); s++)
          for (int rho = 0; rho < 6; rho++) {
            const int* p = PERM[rho];
            if ((p[0] == 2) != (orient == 1)) continue;
            const Triple& tr = bt[s];
            GemmProb P{};
            P.nb = d.V; P.M = d.V; P.N = d.V;
            P.C = ws ? ws + T.out + (size_t)slab_pos(tr.t[p[0]]) * V * V * V * 8 : nullptr;
            P.c_sn = 1; P.alpha = 1.0;
            if (orient == 0) { P.c_sb = V * V; P.c_sm = V; } else { P.c_sb = V; P.c_sm = V * V; }
            const int q = gb.problem(T.key + tr.t[p[0]], P);
            GemmSeg sg{};
            sg.A = gcube(T.g, (int)s);
            if (orient == 0) { sg.a_sb = str(p[0]); sg.a_sm = str(p[1]); }
            else { sg.a_sb = str(p[1]); sg.a_sm = str(p[0]); }
            sg.a_sk2 = str(p[2]);
            sg.B = el(T.t2, ((long long)tr.t[p[2]] * Op + tr.t[p[1]]) * V * V);
            sg.b_sb = 0; sg.b_sk2 = V; sg.b_sn = 1;
            sg.K1 = 1; sg.K2 = d.V; sg.scale = 1.0f;
            gb.add(q, sg);
          }
      emit_gemm(gb, cfg_for(FAM_GEMM_B1, pick_cfg(gb)), GEMM_ACCUM, 1, false, FAM_GEMM_B1);
    }
  }

  void emit_b2(const std::vector<Triple>& bt) {
    const long long V = d.V;
    GemmBuild gb(d.prec);
    for (const B2Term& T : b2_terms())
      for (size_t s = 0; s < bt.size(); s++)
        for (int rho = 0; rho < 6; rho++) {
          const int* p = PERM[rho];
          const Triple& tr = bt[s];
          // t2_bar[t_p2][t_p1][m][n], accumulated in place (rows and columns past v not written)
          GemmProb P{};
          P.nb = 1; P.M = d.V; P.N = d.V; P.Mv = d.v; P.Nv = d.v;
          P.C = outp(T.out, ((size_t)tr.t[p[2]] * d.o + tr.t[p[1]]) * d.v * d.v);
          P.c_sm = d.v; P.c_sn = 1; P.alpha = 1.0;
          const int q = gb.problem(T.key + (long long)tr.t[p[2]] * 4096 + tr.t[p[1]], P);
          GemmSeg sg{};
          sg.A = gcube(T.g, (int)s);
          sg.a_sm = str(p[2]);
          sg.B = el(T.ovvv, slab_pos(tr.t[p[0]]) * V * V * V);
          sg.b_sn = 1;
          if (p[0] != 2) {
            sg.a_sk1 = str(p[0]); sg.a_sk2 = str(p[1]); sg.b_sk1 = V * V; sg.b_sk2 = V;
          } else {
            sg.a_sk1 = str(p[1]); sg.a_sk2 = str(p[0]); sg.b_sk1 = V; sg.b_sk2 = V * V;
          }
          sg.K1 = d.V; sg.K2 = d.V; sg.scale = 1.0f;
          gb.add(q, sg);
        }
    // FP64 partials of 64 splits cost more in the reduction than the extra blocks gain
    emit_gemm(gb, cfg_for(FAM_GEMM_B2, pick_cfg(gb)), GEMM_ACCUM, 1, true, FAM_GEMM_B2,
              d.prec == PREC_F64 ? 16 : MAX_KSPLIT);
  }

  void emit_b3(const std::vector<Triple>& bt) {
    const long long V = d.V, Op = d.Op;
    GemmBuild gb(d.prec);
    for (const B3Term& T : b3_terms())
      for (size_t s = 0; s < bt.size(); s++)
        for (int rho = 0; rho < 6; rho++) {
          const int* p = PERM[rho];
          const Triple& tr = bt[s];
          GemmProb P{};
          P.nb = 1; P.M = d.V; P.N = d.Op;
          P.C = ws ? ws + T.out + ((size_t)tr.t[p[0]] * d.o + tr.t[p[1]]) * V * Op * 8 : nullptr;
          P.c_sm = Op; P.c_sn = 1; P.alpha = 1.0;
          const int q = gb.problem(T.key + (long long)tr.t[p[0]] * 4096 + tr.t[p[1]], P);
          GemmSeg sg{};
          sg.A = gcube(T.g, (int)s);
          sg.a_sm = str(p[0]);
          sg.b_sn = V * V;
          // k runs over the two contracted cube positions in increasing order,
          // so the two permutations with the same free position read G alike
          if (p[1] < p[2]) {   // k = (s_p1, s_p2), B from t2H[t_p2][m][s_p1][s_p2]
            sg.a_sk1 = str(p[1]); sg.a_sk2 = str(p[2]);
            sg.B = el(T.t2h, (long long)tr.t[p[2]] * Op * V * V);
          } else {             // k = (s_p2, s_p1), B from t2T[t_p2][m][s_p2][s_p1]
            sg.a_sk1 = str(p[2]); sg.a_sk2 = str(p[1]);
            sg.B = el(T.t2, (long long)tr.t[p[2]] * Op * V * V);
          }
          sg.b_sk1 = V; sg.b_sk2 = 1;
          sg.K1 = d.V; sg.K2 = d.V; sg.scale = -1.0f;
          gb.add(q, sg);
        }
    // default: 32-wide tiles on the occupied side unless it is a multiple of 64
    emit_gemm(gb, cfg_for(FAM_GEMM_B3, d.Op % 64 ? 3 : 0), GEMM_ACCUM, 1, true, FAM_GEMM_B3);
  }

  void emit_b4(const std::vector<Triple>& bt) {
    const long long V = d.V, Op = d.Op;
    GemmBuild gb(d.prec);
    for (const B4Term& T : b4_terms())
      for (size_t s = 0; s < bt.size(); s++)
        for (int rho = 0; rho < 6; rho++) {
          const int* p = PERM[rho];
          const Triple& tr = bt[s];
          const bool t2layout = (p[1] == 2);
          // t2 layout (the cube's c is t_p1's partner): t2_bar[t_p2][m][b][n],
          // accumulated in place (batches and columns: the v that exist, rows:
          // the o that exist); otherwise the [i][j][b][a] accumulator t2hacc,
          // added to t2_bar transposed at the end (writing t2_bar directly
          // would take uncoalesced stores: the cube's contiguous index is not
          // t2_bar's)
          GemmProb P{};
          if (t2layout) {
            P.nb = d.v; P.M = d.Op; P.N = d.V; P.Mv = d.o; P.Nv = d.v;
            P.C = outp(T.out, (size_t)tr.t[p[2]] * d.o * d.v * d.v);
            P.c_sb = d.v; P.c_sm = (long long)d.v * d.v; P.c_sn = 1;
          } else {
            P.nb = d.V; P.M = d.Op; P.N = d.V;
            P.C = ws ? ws + T.hacc + (size_t)tr.t[p[2]] * Op * V * V * 8 : nullptr;
            P.c_sb = V; P.c_sm = V * V; P.c_sn = 1;
          }
          P.alpha = 1.0;
          const int q = gb.problem(T.key + (t2layout ? 0 : 1) * 4096LL + tr.t[p[2]], P);
          GemmSeg sg{};
          sg.A = el(T.ovoo, ((long long)tr.t[p[0]] * d.o + tr.t[p[1]]) * V * Op);
          sg.a_sb = 0; sg.a_sm = 1; sg.a_sk2 = Op;
          sg.B = gcube(T.g, (int)s);
          sg.b_sk2 = str(p[0]);
          if (t2layout) { sg.b_sb = str(p[2]); sg.b_sn = str(p[1]); }
          else { sg.b_sb = str(p[1]); sg.b_sn = str(p[2]); }
          sg.K1 = 1; sg.K2 = d.V; sg.scale = -1.0f;
          gb.add(q, sg);
        }
    emit_gemm(gb, cfg_for(FAM_GEMM_B4, d.Op % 64 ? 4 : 0), GEMM_ACCUM, 1, false, FAM_GEMM_B4);
  }

  void emit_vpart(const std::vector<Triple>& bt) {
    std::vector<VpTriple> trs;
    std::map<long long, std::vector<VpContrib>> by_slab;
    std::vector<long long> order;
    for (size_t s = 0; s < bt.size(); s++) {
      VpTriple t{};
      t.i = bt[s].t[0]; t.j = bt[s].t[1]; t.k = bt[s].t[2];
      t.H = hcube((int)s);
      t.out = reinterpret_cast<double*>(slot((int)s, r.slot_vp));
      trs.push_back(t);
      for (int rho = 0; rho < 6; rho++) {
        const int* p = PERM[rho];
        const long long key = (long long)bt[s].t[p[0]] * 4096 + bt[s].t[p[1]];
        if (!by_slab.count(key)) order.push_back(key);
        by_slab[key].push_back(VpContrib{(int)s, p[2], p[0] > p[1] ? 1 : 0});
      }
    }
    std::vector<VpSlab> slabs;
    std::vector<VpContrib> contribs;
    for (long long key : order) {
      VpSlab sl{};
      sl.p = (int)(key / 4096); sl.q = (int)(key % 4096);
      sl.c0 = (int)contribs.size();
      for (auto& c : by_slab[key]) contribs.push_back(c);
      sl.nc = (int)contribs.size() - sl.c0;
      slabs.push_back(sl);
    }
    Launch L;
    L.kind = L_VPART;
    L.off_a = blob.put(trs);
    L.n_a = (int)trs.size();
    L.off_b = blob.put(slabs);
    L.n_b = (int)slabs.size();
    L.off_c = blob.put(contribs);
    L.n_c = (int)contribs.size();
    launches.push_back(L);
  }

  void build_plan() {
    if (d.naux > 0) {
      build_plan_df();
      return;
    }
    std::vector<Triple> all = enumerate_triples(d.o);
    const long long n = (long long)all.size();
    const long long e = (t_end < 0 || t_end > n) ? n : t_end;
    const long long b = std::min(std::max(0LL, t_begin), e);
    emit_triples(std::vector<Triple>(all.begin() + b, all.begin() + e));
  }

  // Density-fitted plan: block triples I >= J >= K in order, each preceded
  // by the slab builds (and evictions) it needs; the slice [t_begin, t_end)
  // refers to this order.
  void build_plan_df() {
    df_pos.assign(d.o, -1);
    for (auto& x : df_res) x = -1;
    const int nb = (d.o + d.dfblk - 1) / d.dfblk;
    const long long n = (long long)d.o * (d.o + 1) * (d.o + 2) / 6;
    const long long e = (t_end < 0 || t_end > n) ? n : t_end;
    const long long b = std::min(std::max(0LL, t_begin), e);
    long long idx = 0;
    for (int I = 0; I < nb; I++)
      for (int J = 0; J <= I; J++)
        for (int K = 0; K <= J; K++) {
          const std::vector<Triple> seg = block_triples(d.o, d.dfblk, I, J, K);
          const long long s0 = idx, s1 = idx + (long long)seg.size();
          idx = s1;
          const long long lo = std::max(s0, b), hi = std::min(s1, e);
          if (lo >= hi) continue;
          const int need[3] = {I, J, K};
          df_ensure(need);
          emit_triples(std::vector<Triple>(seg.begin() + (lo - s0), seg.begin() + (hi - s0)));
        }
    if (backward)
      for (int u = 0; u < d.dfslots; u++)
        if (df_res[u] >= 0) emit_df_fold(df_res[u], u);
  }

  // Batches (and, backward, ovvv-adjoint groups of S batches) over a run of
  // triples whose slabs are resident.
  void emit_triples(const std::vector<Triple>& all) {
    const size_t group = (size_t)B * (backward ? r.S : 1);
    for (size_t g0 = 0; g0 < all.size(); g0 += group) {
      const size_t g1 = std::min(all.size(), g0 + group);
      for (size_t b0 = g0; b0 < g1; b0 += B) {
        std::vector<Triple> bt(all.begin() + b0, all.begin() + std::min(g1, b0 + B));
        gbase = (int)(b0 - g0);
        emit_wbuild(bt);
        if (d.lno) emit_sbuild(bt);
        emit_epilogue(bt, backward);
        if (backward) {
          emit_b2(bt);
          emit_b3(bt);
          emit_b4(bt);
          emit_vpart(bt);
        }
      }
      if (backward) {   // the group's G cubes are gslot 0 .. g1-g0-1
        gbase = 0;
        emit_b1(std::vector<Triple>(all.begin() + g0, all.begin() + g1));
      }
    }
    gbase = 0;
  }

  // kernels enqueued by one Launch (the V-term launch runs four)
  static int kernels_of(const Launch& L) { return L.kind == L_VPART ? 4 : 1; }

  // Enqueue launches [from, to) of the plan (all of it by default).
  hipError_t execute(hipStream_t st, size_t from = 0, size_t to = (size_t)-1) {
    char* db = ws + r.desc;
    hipError_t e = hipSuccess;
    std::vector<hipEvent_t> ev;
    std::vector<int> fam, sub;
    to = std::min(to, launches.size());
    for (size_t li = from; li < to; li++) {
      const Launch& L = launches[li];
      if (g_timing) {
        hipEvent_t a, b;
        if ((e = hipEventCreate(&a)) != hipSuccess) return e;
        if ((e = hipEventCreate(&b)) != hipSuccess) return e;
        ev.push_back(a);
        ev.push_back(b);
        fam.push_back(L.kind == L_GEMM ? FAM_GEMM
                      : (L.kind == L_EPI || L.kind == L_EPIRED) ? FAM_EPILOGUE
                      : (L.kind == L_VPART || L.kind == L_VMN) ? FAM_VPART : FAM_REDUCE);
        sub.push_back((L.kind == L_GEMM || L.kind == L_ZERO) ? L.sub : -1);
        if ((e = hipEventRecord(a, st)) != hipSuccess) return e;
      }
      switch (L.kind) {
        case L_GEMM:
          e = t3_launch_gemm(L.prec >= 0 ? L.prec : d.prec, L.cfg, reinterpret_cast<const GemmProb*>(db + L.off_a),
                             L.n_a, reinterpret_cast<const GemmSeg*>(db + L.off_b), L.work, st);
          break;
        case L_REDUCE:
          e = t3_launch_reduce(L.prec >= 0 ? L.prec : d.prec, reinterpret_cast<const ReduceProb*>(db + L.off_a),
                               L.n_a, L.work, L.cfg, st);
          break;
        case L_ZERO:
          e = t3_zero(ws + L.off_a, (size_t)L.work, st);
          break;
        case L_EPI: {
          EpiArgs a = epi_args[L.aux];
          a.cube_tab = reinterpret_cast<const int*>(db + cube_tab);
          e = t3_launch_epilogue(d.tp_is_f64, &a, reinterpret_cast<const EpiTriple*>(db + L.off_a), L.n_a, st);
          break;
        }
        case L_EPIRED: {
          EpiReduceArgs a = epired_args[L.aux];
          a.triples = reinterpret_cast<const EpiTriple*>(db + L.off_a);
          e = t3_launch_epi_reduce(&a, st);
          break;
        }
        case L_VMN:
          e = t3_launch_vcm(reinterpret_cast<const VpTriple*>(db + L.off_a), L.n_a, d.o, d.v, d.V, ovov, t2, st);
          break;
        case L_VPART:
          e = t3_launch_vpart(d.tp_is_f64, reinterpret_cast<const VpTriple*>(db + L.off_a), L.n_a,
                              reinterpret_cast<const VpSlab*>(db + L.off_b), L.n_b,
                              reinterpret_cast<const VpContrib*>(db + L.off_c), d.o, d.v, d.Op, d.V, t1,
                              at<double>(r.fvoP), ovov_out, at<double>(r.t2hacc),
                              at<double>(r.t1_acc), at<double>(r.fvo_acc), st);
          break;
      }
      if (e != hipSuccess) return e;
      if (g_timing && (e = hipEventRecord(ev.back(), st)) != hipSuccess) return e;
    }
    if (g_timing) {
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
      for (size_t x = 0; x < fam.size(); x++) {
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, ev[2 * x], ev[2 * x + 1])) != hipSuccess) return e;
        g_family_ms[fam[x]] += ms;
        g_family_launches[fam[x]]++;
        if (sub[x] >= 0) {
          g_family_ms[sub[x]] += ms;
          g_family_launches[sub[x]]++;
        }
      }
      for (auto& x : ev) (void)hipEventDestroy(x);
    }
    return hipSuccess;
  }
};

int check_args(int nocc, int nvir, int precision) {
  if (nocc < 1 || nvir < 1) return fail(PYSCFAD_AMD_ERR_ARGS, "nocc and nvir must be >= 1");
  if (nocc > 4095 || nvir > 4095) return fail(PYSCFAD_AMD_ERR_ARGS, "nocc and nvir must be < 4096");
  if (precision < PYSCFAD_AMD_PREC_F64 || precision > PYSCFAD_AMD_PREC_BF16X6)
    return fail(PYSCFAD_AMD_ERR_PRECISION, "precision must be 0 (F64), 1 (F32), 2 (BF16X3) or 3 (BF16X6)");
  return PYSCFAD_AMD_OK;
}

// Size the descriptor blob by building the plan against a null workspace.
size_t plan_bytes(int nocc, int nvir, int precision, bool backward, Regions* out_r, int* out_B,
                  long long t_begin = 0, long long t_end = -1, int batch = -1, int naux = 0, int dfblk = 0,
                  int lno = 0) {
  Ctx c;
  c.d = dims_for(nocc, nvir, precision, naux, backward, dfblk, lno);
  c.B = batch > 0 ? batch : choose_batch(c.d);
  c.backward = backward;
  c.t_begin = t_begin;
  c.t_end = t_end;
  c.r = plan_regions(c.d, c.B, backward);
  c.build_plan();
  c.r.desc_bytes = c.blob.data.size();
  c.r.total = c.r.desc + (size_t)round_up((long long)c.r.desc_bytes, 256);
  if (out_r) *out_r = c.r;
  if (out_B) *out_B = c.B;
  return c.r.total;
}

// Inputs and outputs of one (T) problem, in the reference layouts.
struct Problem {
  int nocc = 0, nvir = 0, packed = 1;
  bool backward = false;
  double et_bar = 1.0;
  const double *mo_energy = nullptr, *t1 = nullptr, *t2 = nullptr, *fock = nullptr, *ovoo = nullptr,
               *ovov = nullptr, *ovvv = nullptr;
  double *t1_bar = nullptr, *t2_bar = nullptr, *fock_bar = nullptr, *mo_energy_bar = nullptr,
         *ovoo_bar = nullptr, *ovov_bar = nullptr, *ovvv_bar = nullptr, *e_out = nullptr;
  // density-fitted mode (naux > 0): Lov, Lvv replace ovvv, Lov_bar, Lvv_bar replace ovvv_bar
  int naux = 0;
  const double *Lov = nullptr, *Lvv = nullptr;
  double *Lov_bar = nullptr, *Lvv_bar = nullptr;
  // LNO form (pyscfad/lno/ccsd_t.py:70): mat (o,o), t1T (v,o), t2T (v,v,o,o)
  // and fvo (v,o) replace t1, t2 and fock, and the cotangents follow them
  bool lno = false;
  const double *mat = nullptr, *t1T = nullptr, *t2T = nullptr, *fvo = nullptr;
  double *mat_bar = nullptr, *t1T_bar = nullptr, *t2T_bar = nullptr, *fvo_bar = nullptr;
};

int check_problem(const Problem& p, int precision) {
  int rc = check_args(p.nocc, p.nvir, precision);
  if (rc) return rc;
  if (p.lno) {
    if (p.nocc > t3_lno_max_nocc())
      return fail(PYSCFAD_AMD_ERR_ARGS, "the LNO form takes nocc <= " + std::to_string(t3_lno_max_nocc()));
    if (p.packed != 0 && p.packed != 1) return fail(PYSCFAD_AMD_ERR_ARGS, "ovvv_packed must be 0 or 1");
    if (!p.mat || !p.t1T || !p.t2T || !p.mo_energy || !p.fvo || !p.ovoo || !p.ovov || !p.ovvv)
      return fail(PYSCFAD_AMD_ERR_ARGS, "null input pointer");
    if (p.backward && (!p.mat_bar || !p.t1T_bar || !p.t2T_bar || !p.mo_energy_bar || !p.fvo_bar || !p.ovoo_bar ||
                       !p.ovov_bar || !p.ovvv_bar))
      return fail(PYSCFAD_AMD_ERR_ARGS, "null output pointer");
    if (!p.backward && !p.e_out) return fail(PYSCFAD_AMD_ERR_ARGS, "null e_out");
    return PYSCFAD_AMD_OK;
  }
  const bool df = p.naux > 0;
  if (p.naux < 0 || p.naux > 1000000) return fail(PYSCFAD_AMD_ERR_ARGS, "naux must be in 1..1000000");
  if (!p.mo_energy || !p.t1 || !p.t2 || !p.fock || !p.ovoo || !p.ovov || (df ? (!p.Lov || !p.Lvv) : !p.ovvv))
    return fail(PYSCFAD_AMD_ERR_ARGS, "null input pointer");
  if (p.packed != 0 && p.packed != 1) return fail(PYSCFAD_AMD_ERR_ARGS, "ovvv_packed must be 0 or 1");
  if (p.backward && (!p.t1_bar || !p.t2_bar || !p.fock_bar || !p.mo_energy_bar || !p.ovoo_bar || !p.ovov_bar ||
                     (df ? (!p.Lov_bar || !p.Lvv_bar) : !p.ovvv_bar)))
    return fail(PYSCFAD_AMD_ERR_ARGS, "null output pointer");
  if (!p.backward && !p.e_out) return fail(PYSCFAD_AMD_ERR_ARGS, "null e_out");
  return PYSCFAD_AMD_OK;
}

void bind(Ctx& c, const Problem& p) {
  c.mo_energy = p.mo_energy; c.t1 = p.t1; c.t2 = p.t2; c.fock = p.fock; c.ovoo = p.ovoo; c.ovov = p.ovov;
  c.ovvv = p.ovvv;
  c.packed = p.packed;
  c.backward = p.backward;
  c.t2_out = p.t2_bar;
  c.ovov_out = p.ovov_bar;
  c.et_bar = p.et_bar;
}

// All device work of a planned problem whose descriptor blob is in place at
// ws + r.desc: accumulator clears and input re-layout (setup), the launch
// plan, the cotangent write-back and the energy copy (finish).  Kernels only
// (no memset, copy, allocation or synchronisation), so the sequence can be
// captured into a graph.  sev: 4 events bracketing setup / plan / finish, or
// null.
// kernels enqueued by enqueue_setup / enqueue_finish (upper bounds)
constexpr int SETUP_KERNELS = 12, FINISH_KERNELS = 6;

hipError_t enqueue_setup(Ctx& c, const Problem& p, hipStream_t st) {
  hipError_t e;
  if ((e = t3_zero(c.ws + c.r.energy, 16, st)) != hipSuccess) return e;
  if (p.backward) {
    // G and H cubes are fully rewritten per triple; the split scratch is fully
    // rewritten before each reduction
    const size_t acc0 = c.r.ovvv_acc, acc1 = c.r.acc_end;
    if ((e = t3_zero(c.ws + acc0, (size_t)round_up((long long)(acc1 - acc0), 16), st)) != hipSuccess) return e;
    // t2_bar and ovov_bar are accumulated in place
    const long long n22 = (long long)c.d.o * c.d.o * c.d.v * c.d.v;
    if ((e = t3_zero_f64(p.t2_bar, n22, st)) != hipSuccess) return e;
    if ((e = t3_zero_f64(p.ovov_bar, n22, st)) != hipSuccess) return e;
  }
  e = t3_prep_inputs(c.d.tp_is_f64, c.ws + c.r.ovvvT, c.ws + c.r.t2T, c.ws + c.r.t2H, c.ws + c.r.ovooT,
                     c.at<double>(c.r.eps_occ), c.at<double>(c.r.eps_vir), c.at<double>(c.r.fvoP), p.ovvv,
                     p.packed, p.t2, p.ovoo, p.mo_energy, p.fock, c.d.o, c.d.v, c.d.Op, c.d.V, st);
  if (e == hipSuccess && p.naux > 0)
    e = t3_prep_df(c.at<double>(c.r.LovP), c.at<double>(c.r.LvvF), p.Lov, p.Lvv, p.naux, c.d.o, c.d.v, c.d.V,
                   c.d.NP, st);
  return e;
}

hipError_t enqueue_finish(Ctx& c, const Problem& p, hipStream_t st) {
  hipError_t e;
  if (p.backward) {
    e = t3_finalize(p.t1_bar, p.t2_bar, p.fock_bar, p.mo_energy_bar, p.ovoo_bar, p.ovvv_bar, p.packed,
                    c.at<double>(c.r.ovvv_acc), c.at<double>(c.r.t2hacc), c.at<double>(c.r.ovoo_acc), c.at<double>(c.r.t1_acc),
                    c.at<double>(c.r.fvo_acc), c.at<double>(c.r.eps_bar), c.d.o, c.d.v, c.d.Op, c.d.V, st);
    if (e != hipSuccess) return e;
    if (p.naux > 0 && (e = t3_finalize_df(p.Lov_bar, p.Lvv_bar, c.at<double>(c.r.LovBar), c.at<double>(c.r.LvvBar),
                                          p.naux, c.d.o, c.d.v, c.d.V, c.d.NP, st)) != hipSuccess)
      return e;
  }
  if (p.e_out && (e = t3_copy_f64(p.e_out, c.at<double>(c.r.energy), 1, st)) != hipSuccess) return e;
  return hipSuccess;
}

// LNO form: the canonical problem the plan runs on.  Its t1, t2, fock (and
// their cotangents) are workspace arrays formed from (and folded back into)
// the caller's t1T, t2T, fvo; the other inputs and outputs are the caller's.
Problem lno_inner(const Ctx& c, const Problem& p) {
  Problem q = p;
  q.lno = false;
  q.t1 = c.at<double>(c.r.lt1);
  q.t2 = c.at<double>(c.r.lt2);
  q.fock = c.at<double>(c.r.lfock);
  if (p.backward) {
    q.t1_bar = c.at<double>(c.r.lt1_bar);
    q.t2_bar = c.at<double>(c.r.lt2_bar);
    q.fock_bar = c.at<double>(c.r.lfock_bar);
  }
  return q;
}

// Sizes of the projected tensors viewed as [A][o][B] around the slot.
struct LnoViews {
  long long ovvv_B;                 // ovvv (o, v, pairs): slot 0
  long long t2_B0, t2_A1, t2_B1;    // t2 (o, o, v, v): slot 0, slot 1
  long long ovoo_B0, ovoo_A3;       // ovoo (o, v, o, o): slot 0, slot 3
};
LnoViews lno_views(const Problem& p) {
  const long long o = p.nocc, v = p.nvir;
  return LnoViews{v * (p.packed ? v * (v + 1) / 2 : v * v), o * v * v, o, v * v, v * o * o, o * v * o};
}

// Before the canonical setup: t1, t2, fock from t1T, t2T, fvo, then the
// slot-projected inputs (mat's column on the slot, lno_ccsd_t.c:16-18) in the
// reference layouts, then their device layouts.
hipError_t lno_setup(Ctx& c, const Problem& p, hipStream_t st) {
  const int o = p.nocc, v = p.nvir;
  const long long vv = (long long)v * v, oo = (long long)o * o;
  const LnoViews w = lno_views(p);
  double *t1 = c.at<double>(c.r.lt1), *t2 = c.at<double>(c.r.lt2);
  double *ovvvP = c.at<double>(c.r.lovvvP), *t2Q = c.at<double>(c.r.lt2Q), *t2P0 = c.at<double>(c.r.lt2P0);
  double* ovooQ = c.at<double>(c.r.lovooQ);
  hipError_t e;
  if ((e = t3_lno_transpose(t1, p.t1T, v, o, o, 0, 0, st)) != hipSuccess) return e;        // t1[i,a] = t1T[a,i]
  if ((e = t3_lno_transpose(t2, p.t2T, vv, oo, o, 0, 1, st)) != hipSuccess) return e;      // t2[j,i,a,b] = t2T[a,b,i,j]
  if ((e = t3_lno_fock(c.at<double>(c.r.lfock), p.fvo, o, v, 1, st)) != hipSuccess) return e;
  if ((e = t3_lno_project(ovvvP, p.ovvv, p.mat, 1, o, w.ovvv_B, 0, 0, st)) != hipSuccess) return e;
  if ((e = t3_lno_project(t2Q, t2, p.mat, 1, o, w.t2_B0, 0, 0, st)) != hipSuccess) return e;
  if ((e = t3_lno_project(t2Q, t2, p.mat, w.t2_A1, o, w.t2_B1, 0, 1, st)) != hipSuccess) return e;
  if ((e = t3_lno_project(t2P0, t2, p.mat, 1, o, w.t2_B0, 0, 0, st)) != hipSuccess) return e;
  if ((e = t3_lno_project(ovooQ, p.ovoo, p.mat, 1, o, w.ovoo_B0, 0, 0, st)) != hipSuccess) return e;
  if ((e = t3_lno_project(ovooQ, p.ovoo, p.mat, w.ovoo_A3, o, 1, 0, 1, st)) != hipSuccess) return e;
  return t3_prep_projected(c.d.tp_is_f64, c.ws + c.r.ovvvT_P, c.ws + c.r.t2T_Q, c.ws + c.r.t2T_P0, c.ws + c.r.t2H_P0,
                           c.ws + c.r.ovooT_Q, ovvvP, p.packed, t2Q, t2P0, ovooQ, o, v, c.d.Op, c.d.V, st);
}

// After the canonical write-back (backward): the projected operands'
// cotangents in the reference layouts; mat_bar[l][i] = sum over each
// projected slot of input[.., l, ..] * cotangent[.., i, ..]
// (lno_ccsd_t_vjp.c:104-106 for the cube, here through the operands); the
// back-projections onto the inputs' own cotangents; t1T_bar, t2T_bar, fvo_bar.
hipError_t lno_finish(Ctx& c, const Problem& p, const Problem& q, hipStream_t st) {
  const int o = p.nocc, v = p.nvir;
  const long long vv = (long long)v * v, oo = (long long)o * o;
  const LnoViews w = lno_views(p);
  double *ovvvPb = c.at<double>(c.r.lovvvP_bar), *t2Qb = c.at<double>(c.r.lt2Q_bar);
  double *t2P0b = c.at<double>(c.r.lt2P0_bar), *ovooQb = c.at<double>(c.r.lovooQ_bar);
  double* part = c.at<double>(c.r.matpart);
  const double* t2 = q.t2;
  hipError_t e;
  if ((e = t3_finalize_projected(ovvvPb, p.packed, c.at<double>(c.r.ovvvP_acc), t2P0b, c.at<double>(c.r.t2P0hacc),
                                 ovooQb, c.at<double>(c.r.ovooQ_acc), o, v, c.d.Op, c.d.V, st)) != hipSuccess)
    return e;
  struct Term { const double* x; const double* y; double* xbar; long long A, B; };
  const Term terms[6] = {
      {p.ovvv, ovvvPb, p.ovvv_bar, 1, w.ovvv_B},
      {p.ovoo, ovooQb, p.ovoo_bar, 1, w.ovoo_B0},
      {p.ovoo, ovooQb, p.ovoo_bar, w.ovoo_A3, 1},
      {t2, t2Qb, q.t2_bar, 1, w.t2_B0},
      {t2, t2Qb, q.t2_bar, w.t2_A1, w.t2_B1},
      {t2, t2P0b, q.t2_bar, 1, w.t2_B0},
  };
  for (int k = 0; k < 6; k++) {
    const Term& T = terms[k];
    if ((e = t3_lno_matbar(p.mat_bar, part, T.x, T.y, T.A, o, T.B, k > 0, st)) != hipSuccess) return e;
    if ((e = t3_lno_project(T.xbar, T.y, p.mat, T.A, o, T.B, 1, 1, st)) != hipSuccess) return e;
  }
  if ((e = t3_lno_transpose(p.t1T_bar, q.t1_bar, o, v, o, 0, 0, st)) != hipSuccess) return e;
  if ((e = t3_lno_transpose(p.t2T_bar, q.t2_bar, oo, vv, o, 1, 0, st)) != hipSuccess) return e;
  return t3_lno_fock(p.fvo_bar, q.fock_bar, o, v, 0, st);
}

hipError_t enqueue(Ctx& c, const Problem& p, hipStream_t st, hipEvent_t* sev) {
  hipError_t e;
  auto mark = [&](int i) { return sev ? hipEventRecord(sev[i], st) : hipSuccess; };
  const Problem q = p.lno ? lno_inner(c, p) : p;
  if ((e = mark(0)) != hipSuccess) return e;
  if (p.lno && (e = lno_setup(c, p, st)) != hipSuccess) return e;
  if ((e = enqueue_setup(c, q, st)) != hipSuccess) return e;
  if ((e = mark(1)) != hipSuccess) return e;
  if ((e = c.execute(st)) != hipSuccess) return e;
  if ((e = mark(2)) != hipSuccess) return e;
  if ((e = enqueue_finish(c, q, st)) != hipSuccess) return e;
  if (p.lno && p.backward && (e = lno_finish(c, p, q, st)) != hipSuccess) return e;
  return mark(3);
}

void reset_timing() {
  for (int f = 0; f < FAM_COUNT; f++) { g_family_ms[f] = 0.0; g_family_launches[f] = 0; }
}

// Add the setup and write-back intervals of one enqueue() to family 4.
hipError_t add_setup_time(hipEvent_t* sev) {
  float a = 0.f, b = 0.f;
  hipError_t e;
  if ((e = hipEventElapsedTime(&a, sev[0], sev[1])) != hipSuccess) return e;
  if ((e = hipEventElapsedTime(&b, sev[2], sev[3])) != hipSuccess) return e;
  g_family_ms[FAM_PREP] += (double)a + (double)b;
  g_family_launches[FAM_PREP] += 1;
  return hipSuccess;
}

struct Events4 {
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipError_t create() {
    for (auto& x : ev) {
      hipError_t e = hipEventCreate(&x);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  ~Events4() {
    for (auto& x : ev)
      if (x) (void)hipEventDestroy(x);
  }
};

int autotune(int nocc, int nvir, int precision, int* cfg_out, hipStream_t st);
bool autotune_enabled();

// batch > 0 fixes the triples per batch (else the tuned or default choice);
// allow_tune = false never starts the autotuner (which allocates).
int run(const Problem& p, int precision, void* workspace, size_t workspace_bytes, void* stream,
        long long t_begin = 0, long long t_end = -1, int batch = -1, bool allow_tune = true) {
  g_last_error.clear();
  int rc = check_problem(p, precision);
  if (rc) return rc;
  const long long ntrip = (long long)p.nocc * (p.nocc + 1) * (p.nocc + 2) / 6;
  if (t_end < 0) t_end = ntrip;
  if (t_begin < 0 || t_begin > t_end || t_end > ntrip)
    return fail(PYSCFAD_AMD_ERR_ARGS, "triple range must satisfy 0 <= begin <= end <= nocc(nocc+1)(nocc+2)/6");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (allow_tune && autotune_enabled()) {
    const Dims d = make_dims(p.nocc, p.nvir, precision);
    if (tuned_cfg(precision, FAM_GEMM_W, d.V, d.Op) < 0 && (rc = autotune(p.nocc, p.nvir, precision, nullptr, st)))
      return rc;
  }

  Ctx c;
  c.d = dims_for(p.nocc, p.nvir, precision, p.naux, p.backward, 0, p.lno ? 1 : 0);
  c.t_begin = t_begin;
  c.t_end = t_end;
  bind(c, p);
  const size_t need = plan_bytes(p.nocc, p.nvir, precision, p.backward, &c.r, &c.B, t_begin, t_end, batch, p.naux,
                                 0, p.lno ? 1 : 0);
  bool own = false;
  if (workspace) {
    if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
      return fail(PYSCFAD_AMD_ERR_ARGS, "workspace must be 256-byte aligned");
    if (workspace_bytes < need)
      return fail(PYSCFAD_AMD_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
    c.ws = reinterpret_cast<char*>(workspace);
  } else {
    if (hipMalloc(&c.ws, need) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PYSCFAD_AMD_ERR_OOM, "hipMalloc of " + std::to_string(need) + " bytes failed");
    }
    own = true;
  }
  if (p.lno) {   // the plan reads the workspace copies of t1, t2 and writes the projected t2 cotangents there
    bind(c, lno_inner(c, p));
    c.t2Q_out = p.backward ? c.at<double>(c.r.lt2Q_bar) : nullptr;
    c.t2P0_out = p.backward ? c.at<double>(c.r.lt2P0_bar) : nullptr;
  }
  c.build_plan();
  if (!c.plan_error.empty()) {
    if (own) (void)hipFree(c.ws);
    return fail(PYSCFAD_AMD_ERR_ARGS, "internal plan error: " + c.plan_error);
  }
  auto cleanup = [&](int code) {
    if (own) {
      (void)hipStreamSynchronize(st);
      (void)hipFree(c.ws);
    }
    return code;
  };
  auto hip_fail = [&](hipError_t e, const char* where) {
    return cleanup(fail(PYSCFAD_AMD_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e)));
  };
  reset_timing();
  hipError_t e;
  e = hipMemcpyAsync(c.ws + c.r.desc, c.blob.data.data(), c.blob.data.size(), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "descriptor upload");
  // the upload reads host memory that must stay alive until it completes
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "descriptor upload");
  // with timing on, the per-call setup (accumulator clears, input re-layout)
  // and the cotangent write-back are recorded as family 4
  Events4 ev;
  if (g_timing && (e = ev.create()) != hipSuccess) return hip_fail(e, "event");
  e = enqueue(c, p, st, g_timing ? ev.ev : nullptr);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  if (g_timing) {
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "timing");
    if ((e = add_setup_time(ev.ev)) != hipSuccess) return hip_fail(e, "timing");
  }
  return cleanup(PYSCFAD_AMD_OK);
}

// ---------------------------------------------------------------------------
// Autotuner: for every contraction family, time each valid tile configuration
// on one batch of triples of an (nocc, nvir) problem (operands: a zeroed
// scratch workspace; only GEMM and split-K reduction launches run, so no
// input is read) and keep the fastest in the tuning table, keyed by
// (precision, family, V, Op).  Plans built afterwards use the table.
const int TUNE_FAMILIES[5] = {FAM_GEMM_W, FAM_GEMM_B1, FAM_GEMM_B2, FAM_GEMM_B3, FAM_GEMM_B4};
const int TUNE_CFGS[7] = {0, 1, 3, 4, 5, 6, -1};

// Batch size: the default and twice it, timed on the same 2x-default
// triples through their GEMM and reduction launches (the launches whose
// speed the batch size moves: block counts and cache reuse of the cubes).
int tune_batch(int nocc, int nvir, int precision, hipStream_t st) {
  const Dims d = make_dims(nocc, nvir, precision);
  const long long ntrip = (long long)nocc * (nocc + 1) * (nocc + 2) / 6;
  const int b0 = default_batch(d);
  if (2 * b0 > 64 || ntrip < 2 * b0) return b0;
  const long long t0 = ntrip / 2 - std::min<long long>(ntrip / 2, b0), t1 = t0 + 2 * b0;
  const int cand[2] = {b0, 2 * b0};
  Regions rr[2];
  size_t need = 0;
  for (int x = 0; x < 2; x++) {
    int B = cand[x];
    need = std::max(need, plan_bytes(nocc, nvir, precision, true, &rr[x], &B, t0, t1, cand[x]));
  }
  // the trial plans accumulate t2_bar in place: a scratch one after the workspace
  const size_t t2_bytes = (size_t)round_up((long long)nocc * nocc * nvir * nvir * 8, 256);
  char* ws = nullptr;
  if (hipMalloc(&ws, need + t2_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return b0;
  }
  static const double dummy[4] = {0, 0, 0, 0};
  Ctx plans[2];
  hipError_t e = t3_zero(ws, (need + t2_bytes) / 16 * 16, st);
  for (int x = 0; x < 2; x++) {
    Ctx& c = plans[x];
    c.d = d;
    c.B = cand[x];
    c.r = rr[x];
    c.t_begin = t0;
    c.t_end = t1;
    c.backward = true;
    c.ws = ws;
    c.mo_energy = c.t1 = c.t2 = c.fock = c.ovoo = c.ovov = c.ovvv = dummy;
    c.t2_out = reinterpret_cast<double*>(ws + need);
    c.build_plan();
  }
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  double best[2] = {1e300, 1e300};
  for (int round = 0; round < 4 && e == hipSuccess; round++) {
    for (int x = 0; x < 2 && e == hipSuccess; x++) {
      Ctx& c = plans[x];
      if (!c.plan_error.empty()) continue;
      e = hipMemcpy(ws + c.r.desc, c.blob.data.data(), c.blob.data.size(), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipEventRecord(ev0, st);
      for (size_t li = 0; li < c.launches.size() && e == hipSuccess; li++)
        if (c.launches[li].kind == L_GEMM || c.launches[li].kind == L_REDUCE) e = c.execute(st, li, li + 1);
      if (e == hipSuccess) e = hipEventRecord(ev1, st);
      if (e == hipSuccess) e = hipEventSynchronize(ev1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
      if (round > 0 && e == hipSuccess) best[x] = std::min(best[x], (double)ms);
    }
  }
  (void)hipStreamSynchronize(st);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  (void)hipFree(ws);
  (void)hipGetLastError();
  return (e == hipSuccess && best[1] < best[0]) ? cand[1] : b0;
}

int autotune(int nocc, int nvir, int precision, int* cfg_out, hipStream_t st) {
  int rc = check_args(nocc, nvir, precision);
  if (rc) return rc;
  const long long ntrip = (long long)nocc * (nocc + 1) * (nocc + 2) / 6;
  {
    const int timing = g_timing;
    g_timing = 0;
    const Dims d0 = make_dims(nocc, nvir, precision);
    const int bt = tune_batch(nocc, nvir, precision, st);
    g_timing = timing;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned_batch[{precision, d0.V, d0.Op}] = bt;
  }
  Regions r;
  int B = 1;
  plan_bytes(nocc, nvir, precision, true, &r, &B);
  const long long t0 = ntrip / 2, t1 = std::min(ntrip, t0 + B);
  // Trial plans, in order of preference: the dense backward plan; the
  // density-fitted backward plan with one-index blocks (its contractions
  // have the same shapes, and its slab cache holds 3 slabs instead of o, so
  // it fits where the dense one does not; its slab builds are never run);
  // the forward plan (W build only).
  bool bw = true;
  Dims d = make_dims(nocc, nvir, precision);
  // the backward trial plans accumulate t2_bar in place: a scratch one after the workspace
  const size_t t2_bytes = (size_t)round_up((long long)nocc * nocc * nvir * nvir * 8, 256);
  size_t need = plan_bytes(nocc, nvir, precision, true, &r, &B, t0, t1) + t2_bytes;
  char* ws = nullptr;
  if (hipMalloc(&ws, need) != hipSuccess) {
    (void)hipGetLastError();
    constexpr int TUNE_NAUX = 32;
    need = plan_bytes(nocc, nvir, precision, true, &r, &B, t0, t1, -1, TUNE_NAUX, 1) + t2_bytes;
    if (hipMalloc(&ws, need) == hipSuccess) {
      d = dims_for(nocc, nvir, precision, TUNE_NAUX, true, 1);
    } else {
      (void)hipGetLastError();
      bw = false;
      need = plan_bytes(nocc, nvir, precision, false, &r, &B, t0, t1);
      if (hipMalloc(&ws, need) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PYSCFAD_AMD_ERR_OOM, "autotune workspace of " + std::to_string(need) + " bytes");
      }
    }
  }
  const int timing = g_timing;   // trial launches are timed here, not per family
  g_timing = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipError_t e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  if (e == hipSuccess) e = t3_zero(ws, need / 16 * 16, st);
  // placeholder input pointers: the timed launches never read the inputs
  static const double dummy[4] = {0, 0, 0, 0};
  for (int f = 0; f < 5 && e == hipSuccess; f++) {
    const int fam = TUNE_FAMILIES[f];
    if (!bw && fam != FAM_GEMM_W) {   // untuned: the defaults stay
      if (cfg_out) cfg_out[f] = -1;
      continue;
    }
    // one plan per candidate; the candidates are timed round-robin and each
    // keeps its best round, so clock ramps and noise hit them alike
    std::vector<int> cands;
    std::vector<Ctx> plans;
    std::vector<std::vector<size_t>> runs;
    for (int ci = 0; TUNE_CFGS[ci] >= 0; ci++) {
      const int cfg = TUNE_CFGS[ci];
      if (!t3_gemm_cfg_valid(precision, cfg)) continue;
      Ctx c;
      c.d = d;
      c.B = B;
      c.r = r;
      c.t_begin = t0;
      c.t_end = t1;
      c.backward = bw;
      c.ws = ws;
      c.mo_energy = c.t1 = c.t2 = c.fock = c.ovoo = c.ovov = c.ovvv = dummy;
      if (bw) c.t2_out = reinterpret_cast<double*>(ws + need - t2_bytes);
      c.force_cfg[fam] = cfg;
      c.build_plan();
      if (!c.plan_error.empty()) continue;
      std::vector<size_t> idx;   // the family's GEMM launches and their reductions
      for (size_t li = 0; li < c.launches.size(); li++) {
        const Launch& L = c.launches[li];
        if (L.kind == L_GEMM && L.sub == fam) {
          idx.push_back(li);
          if (li + 1 < c.launches.size() && c.launches[li + 1].kind == L_REDUCE) idx.push_back(li + 1);
        }
      }
      if (idx.empty()) continue;
      cands.push_back(cfg);
      plans.push_back(std::move(c));
      runs.push_back(std::move(idx));
    }
    std::vector<double> best_ms(cands.size(), 1e300);
    const int ROUNDS = 5;   // round 0 warms up
    for (int round = 0; round < ROUNDS && e == hipSuccess; round++) {
      for (size_t x = 0; x < cands.size() && e == hipSuccess; x++) {
        Ctx& c = plans[x];
        e = hipMemcpy(ws + c.r.desc, c.blob.data.data(), c.blob.data.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipEventRecord(ev0, st);
        for (size_t li : runs[x])
          if (e == hipSuccess) e = c.execute(st, li, li + 1);
        if (e == hipSuccess) e = hipEventRecord(ev1, st);
        if (e == hipSuccess) e = hipEventSynchronize(ev1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
        if (round > 0 && e == hipSuccess) best_ms[x] = std::min(best_ms[x], (double)ms);
      }
    }
    int best = -1;
    double bms = 1e300;
    for (size_t x = 0; x < cands.size(); x++)
      if (best_ms[x] < bms) { bms = best_ms[x]; best = cands[x]; }
    if (best >= 0 && e == hipSuccess) {
      std::lock_guard<std::mutex> lock(g_tune_mu);
      g_tuned[{precision, fam, d.V, d.Op}] = best;
    }
    if (cfg_out) cfg_out[f] = best;
  }
  if (cfg_out) cfg_out[5] = B;
  (void)hipStreamSynchronize(st);
  g_timing = timing;
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  (void)hipFree(ws);
  if (e != hipSuccess) return fail(PYSCFAD_AMD_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
  return PYSCFAD_AMD_OK;
}

bool autotune_enabled() {
  const char* env = std::getenv("PYSCFAD_AMD_AUTOTUNE");
  return env && env[0] == '1';
}

// ---------------------------------------------------------------------------
// Batched problems.  The instances are dealt to up to MAX_BATCH_STREAMS
// groups, largest F_E first to the group with the least work so far; a
// group's instances run one after another in stream order on one region part
// of the workspace (sized for the largest instance), and the groups run on
// streams of their own, so small instances' launches overlap on the GPU.
// Every instance's descriptor blob has its own place after the region parts.
// Each group's sequence of kernels is captured once into hipGraphs; a call
// uploads the descriptors from pinned host memory on the caller's stream,
// forks the group streams from it, replays each group's graphs on its stream
// and joins them back.  Graphs are cached by the exact call (shapes,
// pointers, precision, workspace).
constexpr int MAX_BATCH_STREAMS = 4;

int batch_streams(int n) {
  int g = MAX_BATCH_STREAMS;
  if (const char* env = std::getenv("PYSCFAD_AMD_BATCH_STREAMS")) {
    const int x = std::atoi(env);
    if (x >= 1 && x <= 16) g = x;
  }
  return std::max(1, std::min(g, n));
}

struct BatchPlan {
  std::vector<Problem> probs;
  std::vector<Ctx> ctx;
  std::vector<int> group;           // group of each instance
  std::vector<size_t> blob_off;     // descriptor blob of each instance, from the workspace base
  int ngroups = 1;
  size_t region_bytes = 0, blob_bytes = 0;
  size_t need() const { return region_bytes * ngroups + blob_bytes; }
};

int plan_batch(int n, const PyscfadAmdInstance* inst, int precision, int with_vjp, BatchPlan& bp) {
  if (n < 1 || !inst) return fail(PYSCFAD_AMD_ERR_ARGS, "need at least one instance");
  bp.probs.resize(n);
  bp.ctx.resize(n);
  bp.group.assign(n, 0);
  bp.blob_off.assign(n, 0);
  std::vector<size_t> blob(n);
  std::vector<double> fe(n);
  for (int x = 0; x < n; x++) {
    const PyscfadAmdInstance& I = inst[x];
    Problem& p = bp.probs[x];
    p.nocc = I.nocc; p.nvir = I.nvir; p.packed = I.ovvv_packed; p.backward = with_vjp != 0; p.et_bar = I.et_bar;
    p.mo_energy = I.mo_energy; p.t1 = I.t1; p.t2 = I.t2; p.fock = I.fock; p.ovoo = I.ovoo; p.ovov = I.ovov;
    p.ovvv = I.ovvv;
    p.t1_bar = I.t1_bar; p.t2_bar = I.t2_bar; p.fock_bar = I.fock_bar; p.mo_energy_bar = I.mo_energy_bar;
    p.ovoo_bar = I.ovoo_bar; p.ovov_bar = I.ovov_bar; p.ovvv_bar = I.ovvv_bar; p.e_out = I.e_out;
    int rc = check_problem(p, precision);
    if (rc) return fail(rc, "instance " + std::to_string(x) + ": " + g_last_error);
    Regions r;
    int B = 1;
    plan_bytes(p.nocc, p.nvir, precision, p.backward, &r, &B);
    bp.region_bytes = std::max(bp.region_bytes, r.desc);
    blob[x] = (size_t)round_up((long long)r.desc_bytes, 256);
    Ctx& c = bp.ctx[x];
    c.d = make_dims(p.nocc, p.nvir, precision);
    c.B = B;
    c.r = r;
    bind(c, p);
    fe[x] = 2.0 * (double)p.nocc * p.nocc * p.nocc * (double)p.nvir * p.nvir * p.nvir * (p.nvir + p.nocc);
  }
  bp.region_bytes = (size_t)round_up((long long)bp.region_bytes, 256);
  bp.ngroups = batch_streams(n);
  // largest first to the least loaded group (ties: the lower group)
  std::vector<int> order(n);
  for (int x = 0; x < n; x++) order[x] = x;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return fe[a] > fe[b]; });
  std::vector<double> load(bp.ngroups, 0.0);
  for (int x : order) {
    const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    bp.group[x] = g;
    load[g] += fe[x];
  }
  size_t off = bp.region_bytes * bp.ngroups;
  for (int x = 0; x < n; x++) {
    // the plan of instance x is built against its group's region part; its
    // descriptor offset is taken from that base
    bp.blob_off[x] = off;
    bp.ctx[x].r.desc = off - bp.region_bytes * bp.group[x];
    off += blob[x];
  }
  bp.blob_bytes = off - bp.region_bytes * bp.ngroups;
  return PYSCFAD_AMD_OK;
}

// A captured call: per group, a chain of graphs, each of at most
// MAX_GRAPH_NODES kernel nodes (a single graph for small groups), launched
// back to back on the group's stream.
struct GraphEntry {
  std::vector<char> key;
  std::vector<std::vector<hipGraph_t>> graphs;        // per group
  std::vector<std::vector<hipGraphExec_t>> execs;     // per group
  std::vector<hipStream_t> streams;                   // one per group (default priority)
  std::vector<hipEvent_t> events;                     // [0] fork, [1 + g] join of group g
  void* host_blobs = nullptr;   // pinned; uploaded on the caller's stream before the fork
  size_t blob_at = 0, blob_bytes = 0;
};
constexpr int MAX_GRAPH_NODES = 2048;

std::mutex g_graph_mu;
std::vector<GraphEntry> g_graphs;   // least recently used first
long long g_graph_replays = 0, g_graph_captures = 0;   // guarded by g_graph_mu
constexpr size_t MAX_GRAPHS = 4;

void destroy_entry(GraphEntry& g) {
  for (auto& v : g.execs)
    for (auto x : v) (void)hipGraphExecDestroy(x);
  for (auto& v : g.graphs)
    for (auto x : v) (void)hipGraphDestroy(x);
  for (auto x : g.events) (void)hipEventDestroy(x);
  for (auto x : g.streams) (void)hipStreamDestroy(x);
  if (g.host_blobs) (void)hipHostFree(g.host_blobs);
  g = GraphEntry{};
}

hipError_t launch_entry(const GraphEntry& g, char* ws, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(ws + g.blob_at, g.host_blobs, g.blob_bytes, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  if ((e = hipEventRecord(g.events[0], st)) != hipSuccess) return e;
  for (size_t gi = 0; gi < g.streams.size(); gi++) {
    if ((e = hipStreamWaitEvent(g.streams[gi], g.events[0], 0)) != hipSuccess) return e;
    for (auto x : g.execs[gi])
      if ((e = hipGraphLaunch(x, g.streams[gi])) != hipSuccess) return e;
    if ((e = hipEventRecord(g.events[1 + gi], g.streams[gi])) != hipSuccess) return e;
  }
  for (size_t gi = 0; gi < g.streams.size(); gi++)
    if ((e = hipStreamWaitEvent(st, g.events[1 + gi], 0)) != hipSuccess) return e;
  return hipSuccess;
}

int run_batched(int n, const PyscfadAmdInstance* inst, int precision, int with_vjp, void* workspace,
                size_t workspace_bytes, void* stream) {
  g_last_error.clear();
  if (autotune_enabled() && n > 0 && inst) {
    // one tuning run per size class not tuned yet (the largest nocc of the class)
    std::map<std::pair<int, int>, int> cls;
    for (int x = 0; x < n; x++) {
      if (check_args(inst[x].nocc, inst[x].nvir, precision)) continue;
      const Dims d = make_dims(inst[x].nocc, inst[x].nvir, precision);
      if (tuned_cfg(precision, FAM_GEMM_W, d.V, d.Op) >= 0) continue;
      auto key = std::make_pair(d.V, d.Op);
      if (!cls.count(key) || inst[cls[key]].nocc < inst[x].nocc) cls[key] = x;
    }
    for (auto& kv : cls) {
      const int rc = autotune(inst[kv.second].nocc, inst[kv.second].nvir, precision, nullptr,
                              reinterpret_cast<hipStream_t>(stream));
      if (rc) return rc;
    }
  }
  BatchPlan bp;
  int rc = plan_batch(n, inst, precision, with_vjp, bp);
  if (rc) return rc;
  if (!workspace) return fail(PYSCFAD_AMD_ERR_ARGS, "the batched entry point needs a caller workspace");
  if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
    return fail(PYSCFAD_AMD_ERR_ARGS, "workspace must be 256-byte aligned");
  if (workspace_bytes < bp.need())
    return fail(PYSCFAD_AMD_ERR_WORKSPACE, "workspace too small: need " + std::to_string(bp.need()) + " bytes");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* ws = reinterpret_cast<char*>(workspace);
  reset_timing();

  // the cache key holds every field of the call, written one by one: the
  // struct's padding bytes (whatever the caller's memory held) never enter it
  std::vector<char> key;
  auto put = [&key](const void* src, size_t nb) {
    const char* b = reinterpret_cast<const char*>(src);
    key.insert(key.end(), b, b + nb);
  };
  {
    const int head[4] = {n, precision, with_vjp, bp.ngroups};
    put(head, sizeof(head));
    put(&workspace, sizeof(void*));
    put(&workspace_bytes, sizeof(size_t));
    for (int x = 0; x < n; x++) {
      const PyscfadAmdInstance& I = inst[x];
      const int dims[3] = {I.nocc, I.nvir, I.ovvv_packed};
      put(dims, sizeof(dims));
      put(&I.et_bar, sizeof(double));
      const void* ptrs[15] = {I.mo_energy, I.t1, I.t2, I.fock, I.ovoo, I.ovov, I.ovvv, I.t1_bar, I.t2_bar,
                              I.fock_bar, I.mo_energy_bar, I.ovoo_bar, I.ovov_bar, I.ovvv_bar, I.e_out};
      put(ptrs, sizeof(ptrs));
    }
  }
  std::lock_guard<std::mutex> lock(g_graph_mu);
  if (!g_timing) {
    for (size_t x = 0; x < g_graphs.size(); x++) {
      if (g_graphs[x].key != key) continue;
      GraphEntry hit = g_graphs[x];
      g_graphs.erase(g_graphs.begin() + x);
      g_graphs.push_back(hit);
      const hipError_t e = launch_entry(hit, ws, st);
      if (e != hipSuccess) return fail(PYSCFAD_AMD_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(e));
      g_graph_replays++;
      return PYSCFAD_AMD_OK;
    }
  }

  // plans against the real workspace (each instance on its group's region
  // part), blobs into one pinned buffer
  GraphEntry g;
  g.key = key;
  g.blob_at = bp.region_bytes * bp.ngroups;
  g.blob_bytes = bp.blob_bytes;
  if (hipHostMalloc(&g.host_blobs, std::max<size_t>(bp.blob_bytes, 256), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PYSCFAD_AMD_ERR_OOM, "pinned host buffer for the descriptors");
  }
  for (int x = 0; x < n; x++) {
    Ctx& c = bp.ctx[x];
    c.ws = ws + bp.region_bytes * bp.group[x];
    c.build_plan();
    if (!c.plan_error.empty()) {
      destroy_entry(g);
      return fail(PYSCFAD_AMD_ERR_ARGS, "internal plan error: " + c.plan_error);
    }
    std::memcpy(static_cast<char*>(g.host_blobs) + (bp.blob_off[x] - g.blob_at), c.blob.data.data(),
                c.blob.data.size());
  }
  auto hip_fail = [&](hipError_t e, const char* where) {
    destroy_entry(g);
    return fail(PYSCFAD_AMD_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  if (g_timing) {
    // eager and in order on the caller's stream, with per-launch events, for measurement
    e = hipMemcpyAsync(ws + g.blob_at, g.host_blobs, bp.blob_bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "descriptor upload");
    for (int x = 0; x < n; x++) {
      Events4 ev;
      if ((e = ev.create()) != hipSuccess) return hip_fail(e, "event");
      if ((e = enqueue(bp.ctx[x], bp.probs[x], st, ev.ev)) != hipSuccess) return hip_fail(e, "kernel launch");
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "timing");
      if ((e = add_setup_time(ev.ev)) != hipSuccess) return hip_fail(e, "timing");
    }
    destroy_entry(g);
    return PYSCFAD_AMD_OK;
  }
  // Per group: its instances as a list of pieces (instance setup, runs of
  // plan launches, instance finish), cut into graphs of at most
  // MAX_GRAPH_NODES kernels: the graph runtime walks a graph's dependency
  // chain, and one chain per call (hundreds of thousands of nodes for 64
  // instances) is kept out of it.
  struct Piece { int inst, kind; size_t a, b; int nodes; };   // kind 0 setup, 1 launches [a,b), 2 finish
  g.graphs.resize(bp.ngroups);
  g.execs.resize(bp.ngroups);
  // capture on a private stream (the caller's may be the null stream, which
  // cannot be captured); the graphs are launched on the group streams
  hipStream_t cs = nullptr;
  if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "stream");
  for (int gr = 0; gr < bp.ngroups && e == hipSuccess; gr++) {
    std::vector<std::vector<Piece>> pieces(1);
    int in_graph = 0;
    auto add = [&](const Piece& pc) {
      if (in_graph + pc.nodes > MAX_GRAPH_NODES && !pieces.back().empty()) {
        pieces.emplace_back();
        in_graph = 0;
      }
      pieces.back().push_back(pc);
      in_graph += pc.nodes;
    };
    for (int x = 0; x < n; x++) {
      if (bp.group[x] != gr) continue;
      add(Piece{x, 0, 0, 0, SETUP_KERNELS});
      const auto& ls = bp.ctx[x].launches;
      size_t a = 0;
      while (a < ls.size()) {
        size_t b = a;
        int k = 0;
        const int room = std::max(MAX_GRAPH_NODES - in_graph, 0);
        const int cap = room > 0 ? room : MAX_GRAPH_NODES;
        while (b < ls.size() && k + Ctx::kernels_of(ls[b]) <= cap) k += Ctx::kernels_of(ls[b++]);
        if (b == a) k += Ctx::kernels_of(ls[b++]);
        add(Piece{x, 1, a, b, k});
        a = b;
      }
      add(Piece{x, 2, 0, 0, FINISH_KERNELS});
    }
    for (size_t gi = 0; gi < pieces.size() && e == hipSuccess; gi++) {
      if (pieces[gi].empty()) continue;
      e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) break;
      for (const Piece& pc : pieces[gi]) {
        if (e != hipSuccess) break;
        Ctx& c = bp.ctx[pc.inst];
        const Problem& p = bp.probs[pc.inst];
        if (pc.kind == 0) e = enqueue_setup(c, p, cs);
        else if (pc.kind == 1) e = c.execute(cs, pc.a, pc.b);
        else e = enqueue_finish(c, p, cs);
      }
      hipGraph_t graph = nullptr;
      const hipError_t e2 = hipStreamEndCapture(cs, &graph);
      if (e == hipSuccess) e = e2;
      if (graph) g.graphs[gr].push_back(graph);
      if (e != hipSuccess) break;
      hipGraphExec_t exec = nullptr;
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e == hipSuccess) g.execs[gr].push_back(exec);
    }
  }
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) return hip_fail(e, "graph capture");
  for (int gr = 0; gr < bp.ngroups && e == hipSuccess; gr++) {
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) g.streams.push_back(s);
  }
  for (int x = 0; x <= bp.ngroups && e == hipSuccess; x++) {
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) g.events.push_back(ev);
  }
  if (e != hipSuccess) return hip_fail(e, "streams and events");
  if ((e = launch_entry(g, ws, st)) != hipSuccess) return hip_fail(e, "graph launch");
  if (g_graphs.size() >= MAX_GRAPHS) {
    // the evicted graph may still be running: its pinned upload source must outlive it
    (void)hipDeviceSynchronize();
    destroy_entry(g_graphs.front());
    g_graphs.erase(g_graphs.begin());
  }
  g_graphs.push_back(g);
  g_graph_captures++;
  return PYSCFAD_AMD_OK;
}

}  // namespace

namespace {
Problem make_problem(int nocc, int nvir, const double* mo_energy, const double* t1, const double* t2,
                     const double* fock, const double* ovoo, const double* ovov, const double* ovvv, int packed,
                     bool backward, double et_bar, double* t1_bar, double* t2_bar, double* fock_bar,
                     double* mo_energy_bar, double* ovoo_bar, double* ovov_bar, double* ovvv_bar, double* e_out) {
  Problem p;
  p.nocc = nocc; p.nvir = nvir; p.packed = packed; p.backward = backward; p.et_bar = et_bar;
  p.mo_energy = mo_energy; p.t1 = t1; p.t2 = t2; p.fock = fock; p.ovoo = ovoo; p.ovov = ovov; p.ovvv = ovvv;
  p.t1_bar = t1_bar; p.t2_bar = t2_bar; p.fock_bar = fock_bar; p.mo_energy_bar = mo_energy_bar;
  p.ovoo_bar = ovoo_bar; p.ovov_bar = ovov_bar; p.ovvv_bar = ovvv_bar; p.e_out = e_out;
  return p;
}
}  // namespace

extern "C" {

int pyscfad_amd_ccsd_t_energy(int nocc, int nvir, const double* mo_energy, const double* t1, const double* t2,
                              const double* fock, const double* ovoo, const double* ovov, const double* ovvv,
                              int ovvv_packed, int precision, double* e_out, void* workspace,
                              size_t workspace_bytes, void* stream) {
  return run(make_problem(nocc, nvir, mo_energy, t1, t2, fock, ovoo, ovov, ovvv, ovvv_packed, false, 1.0, nullptr,
                          nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, e_out),
             precision, workspace, workspace_bytes, stream);
}

int pyscfad_amd_ccsd_t_energy_vjp(int nocc, int nvir, const double* mo_energy, const double* t1, const double* t2,
                                  const double* fock, const double* ovoo, const double* ovov, const double* ovvv,
                                  int ovvv_packed, double et_bar, int precision, double* t1_bar, double* t2_bar,
                                  double* fock_bar, double* mo_energy_bar, double* ovoo_bar, double* ovov_bar,
                                  double* ovvv_bar, double* e_out, void* workspace, size_t workspace_bytes,
                                  void* stream) {
  return run(make_problem(nocc, nvir, mo_energy, t1, t2, fock, ovoo, ovov, ovvv, ovvv_packed, true, et_bar, t1_bar,
                          t2_bar, fock_bar, mo_energy_bar, ovoo_bar, ovov_bar, ovvv_bar, e_out),
             precision, workspace, workspace_bytes, stream);
}

int pyscfad_amd_ccsd_t_energy_vjp_noalloc(int nocc, int nvir, const double* mo_energy, const double* t1,
                                          const double* t2, const double* fock, const double* ovoo,
                                          const double* ovov, const double* ovvv, int ovvv_packed, double et_bar,
                                          int precision, double* t1_bar, double* t2_bar, double* fock_bar,
                                          double* mo_energy_bar, double* ovoo_bar, double* ovov_bar,
                                          double* ovvv_bar, double* e_out, void* stream) {
  g_last_error.clear();
  const Problem p = make_problem(nocc, nvir, mo_energy, t1, t2, fock, ovoo, ovov, ovvv, ovvv_packed, true, et_bar,
                                 t1_bar, t2_bar, fock_bar, mo_energy_bar, ovoo_bar, ovov_bar, ovvv_bar, e_out);
  int rc = check_problem(p, precision);
  if (rc) return rc;
  // the largest batch (up to the usual choice, halving) whose plan fits the scratch
  const size_t cap = t3_noalloc_scratch_bytes();
  int B = choose_batch(make_dims(nocc, nvir, precision));
  size_t need = 0;
  for (;; B /= 2) {
    need = plan_bytes(nocc, nvir, precision, true, nullptr, nullptr, 0, -1, B);
    if (need <= cap || B == 1) break;
  }
  if (need > cap)
    return fail(PYSCFAD_AMD_ERR_WORKSPACE, "nocc=" + std::to_string(nocc) + ", nvir=" + std::to_string(nvir) +
                                               " needs " + std::to_string(need) +
                                               " bytes of scratch even at one triple per batch; the allocation-free "
                                               "entry point holds " + std::to_string(cap));
  void* scratch = nullptr;
  size_t bytes = 0;
  const hipError_t e = t3_noalloc_scratch(&scratch, &bytes);
  if (e != hipSuccess) return fail(PYSCFAD_AMD_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
  return run(p, precision, scratch, bytes, stream, 0, -1, B, false);
}

int pyscfad_amd_ccsd_t_autotune(int nocc, int nvir, int precision, int* cfg_out, void* stream) {
  g_last_error.clear();
  return autotune(nocc, nvir, precision, cfg_out, reinterpret_cast<hipStream_t>(stream));
}

size_t pyscfad_amd_ccsd_t_batched_workspace_bytes(int n, const PyscfadAmdInstance* instances, int precision,
                                                  int with_vjp) {
  BatchPlan bp;
  if (plan_batch(n, instances, precision, with_vjp, bp)) return 0;
  return bp.need();
}

int pyscfad_amd_ccsd_t_batched(int n, const PyscfadAmdInstance* instances, int precision, int with_vjp,
                               void* workspace, size_t workspace_bytes, void* stream) {
  return run_batched(n, instances, precision, with_vjp, workspace, workspace_bytes, stream);
}

long long pyscfad_amd_ccsd_t_num_triples(int nocc) {
  return nocc < 1 ? 0 : (long long)nocc * (nocc + 1) * (nocc + 2) / 6;
}

int pyscfad_amd_ccsd_t_energy_slice(int nocc, int nvir, const double* mo_energy, const double* t1, const double* t2,
                                    const double* fock, const double* ovoo, const double* ovov, const double* ovvv,
                                    int ovvv_packed, int precision, long long triple_begin, long long triple_end,
                                    double* e_out, void* workspace, size_t workspace_bytes, void* stream) {
  return run(make_problem(nocc, nvir, mo_energy, t1, t2, fock, ovoo, ovov, ovvv, ovvv_packed, false, 1.0, nullptr,
                          nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, e_out),
             precision, workspace, workspace_bytes, stream, triple_begin, triple_end);
}

int pyscfad_amd_ccsd_t_energy_vjp_slice(int nocc, int nvir, const double* mo_energy, const double* t1,
                                        const double* t2, const double* fock, const double* ovoo,
                                        const double* ovov, const double* ovvv, int ovvv_packed, double et_bar,
                                        int precision, long long triple_begin, long long triple_end,
                                        double* t1_bar, double* t2_bar, double* fock_bar, double* mo_energy_bar,
                                        double* ovoo_bar, double* ovov_bar, double* ovvv_bar, double* e_out,
                                        void* workspace, size_t workspace_bytes, void* stream) {
  return run(make_problem(nocc, nvir, mo_energy, t1, t2, fock, ovoo, ovov, ovvv, ovvv_packed, true, et_bar, t1_bar,
                          t2_bar, fock_bar, mo_energy_bar, ovoo_bar, ovov_bar, ovvv_bar, e_out),
             precision, workspace, workspace_bytes, stream, triple_begin, triple_end);
}

size_t pyscfad_amd_ccsd_t_workspace_bytes(int nocc, int nvir, int precision, int with_vjp) {
  if (check_args(nocc, nvir, precision)) return 0;
  return plan_bytes(nocc, nvir, precision, with_vjp != 0, nullptr, nullptr);
}

size_t pyscfad_amd_ccsd_t_df_workspace_bytes(int nocc, int nvir, int naux, int precision, int with_vjp) {
  if (check_args(nocc, nvir, precision) || naux < 1 || naux > 1000000) return 0;
  return plan_bytes(nocc, nvir, precision, with_vjp != 0, nullptr, nullptr, 0, -1, -1, naux);
}

int pyscfad_amd_ccsd_t_df_energy(int nocc, int nvir, int naux, const double* mo_energy, const double* t1,
                                 const double* t2, const double* fock, const double* ovoo, const double* ovov,
                                 const double* Lov, const double* Lvv, int precision, long long triple_begin,
                                 long long triple_end, double* e_out, void* workspace, size_t workspace_bytes,
                                 void* stream) {
  Problem p = make_problem(nocc, nvir, mo_energy, t1, t2, fock, ovoo, ovov, nullptr, 1, false, 1.0, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, e_out);
  p.naux = naux < 1 ? -1 : naux;
  p.Lov = Lov;
  p.Lvv = Lvv;
  return run(p, precision, workspace, workspace_bytes, stream, triple_begin, triple_end);
}

int pyscfad_amd_ccsd_t_df_energy_vjp(int nocc, int nvir, int naux, const double* mo_energy, const double* t1,
                                     const double* t2, const double* fock, const double* ovoo, const double* ovov,
                                     const double* Lov, const double* Lvv, double et_bar, int precision,
                                     long long triple_begin, long long triple_end, double* t1_bar, double* t2_bar,
                                     double* fock_bar, double* mo_energy_bar, double* ovoo_bar, double* ovov_bar,
                                     double* Lov_bar, double* Lvv_bar, double* e_out, void* workspace,
                                     size_t workspace_bytes, void* stream) {
  Problem p = make_problem(nocc, nvir, mo_energy, t1, t2, fock, ovoo, ovov, nullptr, 1, true, et_bar, t1_bar, t2_bar,
                           fock_bar, mo_energy_bar, ovoo_bar, ovov_bar, nullptr, e_out);
  p.naux = naux < 1 ? -1 : naux;
  p.Lov = Lov;
  p.Lvv = Lvv;
  p.Lov_bar = Lov_bar;
  p.Lvv_bar = Lvv_bar;
  return run(p, precision, workspace, workspace_bytes, stream, triple_begin, triple_end);
}

size_t pyscfad_amd_lno_ccsd_t_workspace_bytes(int nocc, int nvir, int precision, int with_vjp) {
  if (check_args(nocc, nvir, precision) || nocc > t3_lno_max_nocc()) return 0;
  return plan_bytes(nocc, nvir, precision, with_vjp != 0, nullptr, nullptr, 0, -1, -1, 0, 0, 1);
}

int pyscfad_amd_lno_ccsd_t_energy(int nocc, int nvir, const double* mat, const double* t1T, const double* t2T,
                                  const double* mo_energy, const double* fvo, const double* ovoo, const double* ovov,
                                  const double* ovvv, int ovvv_packed, int precision, double* e_out, void* workspace,
                                  size_t workspace_bytes, void* stream) {
  Problem p;
  p.lno = true;
  p.nocc = nocc; p.nvir = nvir; p.packed = ovvv_packed; p.backward = false;
  p.mat = mat; p.t1T = t1T; p.t2T = t2T; p.mo_energy = mo_energy; p.fvo = fvo;
  p.ovoo = ovoo; p.ovov = ovov; p.ovvv = ovvv; p.e_out = e_out;
  return run(p, precision, workspace, workspace_bytes, stream);
}

int pyscfad_amd_lno_ccsd_t_energy_vjp(int nocc, int nvir, const double* mat, const double* t1T, const double* t2T,
                                      const double* mo_energy, const double* fvo, const double* ovoo,
                                      const double* ovov, const double* ovvv, int ovvv_packed, double et_bar,
                                      int precision, double* mat_bar, double* t1T_bar, double* t2T_bar,
                                      double* mo_energy_bar, double* fvo_bar, double* ovoo_bar, double* ovov_bar,
                                      double* ovvv_bar, double* e_out, void* workspace, size_t workspace_bytes,
                                      void* stream) {
  Problem p;
  p.lno = true;
  p.nocc = nocc; p.nvir = nvir; p.packed = ovvv_packed; p.backward = true; p.et_bar = et_bar;
  p.mat = mat; p.t1T = t1T; p.t2T = t2T; p.mo_energy = mo_energy; p.fvo = fvo;
  p.ovoo = ovoo; p.ovov = ovov; p.ovvv = ovvv; p.e_out = e_out;
  p.mat_bar = mat_bar; p.t1T_bar = t1T_bar; p.t2T_bar = t2T_bar; p.mo_energy_bar = mo_energy_bar;
  p.fvo_bar = fvo_bar; p.ovoo_bar = ovoo_bar; p.ovov_bar = ovov_bar; p.ovvv_bar = ovvv_bar;
  return run(p, precision, workspace, workspace_bytes, stream);
}

const char* pyscfad_amd_last_error(void) { return g_last_error.c_str(); }

// XLA legacy custom-call entry points (api_version 0, platform ROCM): the
// handler signature needs no XLA headers.  buffers[] holds the operands then
// the results, opaque a PyscfadAmdXlaOpaque.  The legacy ABI has no error
// channel: a failure is reported on stderr and the results are left unset.
namespace {
bool xla_opaque(const char* opaque, size_t opaque_len, PyscfadAmdXlaOpaque& op, const char* who) {
  if (!opaque || opaque_len < sizeof(op)) {
    std::fprintf(stderr, "%s: opaque of %zu bytes, need %zu\n", who, opaque_len, sizeof(op));
    return false;
  }
  std::memcpy(&op, opaque, sizeof(op));
  return true;
}
}  // namespace

void pyscfad_amd_xla_ccsd_t_energy(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
  PyscfadAmdXlaOpaque op{};
  if (!xla_opaque(opaque, opaque_len, op, "pyscfad_amd_xla_ccsd_t_energy")) return;
  const double* in[7];
  for (int x = 0; x < 7; x++) in[x] = reinterpret_cast<const double*>(buffers[x]);
  void* ws = op.workspace_bytes ? buffers[8] : nullptr;
  const int rc = pyscfad_amd_ccsd_t_energy(op.nocc, op.nvir, in[0], in[1], in[2], in[3], in[4], in[5], in[6],
                                           op.ovvv_packed, op.precision, reinterpret_cast<double*>(buffers[7]),
                                           ws, (size_t)op.workspace_bytes, stream);
  if (rc) std::fprintf(stderr, "pyscfad_amd_xla_ccsd_t_energy: error %d: %s\n", rc, g_last_error.c_str());
}

// The cotangent of E(T) is the device scalar buffers[7] (traced inside a jitted
// backward): the VJP is linear in it, so the cotangents are formed for 1 and
// then scaled by it on the stream.
void pyscfad_amd_xla_ccsd_t_energy_vjp(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
  PyscfadAmdXlaOpaque op{};
  if (!xla_opaque(opaque, opaque_len, op, "pyscfad_amd_xla_ccsd_t_energy_vjp")) return;
  const double* in[7];
  double* out[7];
  for (int x = 0; x < 7; x++) in[x] = reinterpret_cast<const double*>(buffers[x]);
  const double* et_bar = reinterpret_cast<const double*>(buffers[7]);
  for (int x = 0; x < 7; x++) out[x] = reinterpret_cast<double*>(buffers[8 + x]);
  void* ws = op.workspace_bytes ? buffers[15] : nullptr;
  int rc = pyscfad_amd_ccsd_t_energy_vjp(op.nocc, op.nvir, in[0], in[1], in[2], in[3], in[4], in[5], in[6],
                                         op.ovvv_packed, 1.0, op.precision, out[0], out[1], out[2], out[3], out[4],
                                         out[5], out[6], nullptr, ws, (size_t)op.workspace_bytes, stream);
  if (!rc) {
    const long long o = op.nocc, v = op.nvir, n = o + v;
    const long long len[7] = {o * v, o * o * v * v, n * n, n, o * v * o * o, o * v * o * v,
                              op.ovvv_packed ? o * v * (v * (v + 1) / 2) : o * v * v * v};
    for (int x = 0; x < 7 && !rc; x++) {
      const hipError_t e = t3_scale_f64(out[x], len[x], et_bar, reinterpret_cast<hipStream_t>(stream));
      if (e != hipSuccess) rc = fail(PYSCFAD_AMD_ERR_HIP, std::string("scale: ") + hipGetErrorString(e));
    }
  }
  if (rc) std::fprintf(stderr, "pyscfad_amd_xla_ccsd_t_energy_vjp: error %d: %s\n", rc, g_last_error.c_str());
}

void pyscfad_amd_set_timing(int enable) { g_timing = enable != 0; }

int pyscfad_amd_last_timing(double* ms, long long* launches, int n) {
  const int m = n < FAM_COUNT ? n : FAM_COUNT;
  for (int f = 0; f < m; f++) {
    if (ms) ms[f] = g_family_ms[f];
    if (launches) launches[f] = g_family_launches[f];
  }
  return m;
}

void pyscfad_amd_graph_cache_stats(long long* replays, long long* captures) {
  std::lock_guard<std::mutex> lock(g_graph_mu);
  if (replays) *replays = g_graph_replays;
  if (captures) *captures = g_graph_captures;
}

int pyscfad_amd_abi_version(void) { return 2; }

}  // extern "C"

// Host-side bounds check of the launch plan (test infrastructure, no GPU):
// builds the plan of an (nocc, nvir, precision[, naux]) backward call with
// the library's own code (t3_driver.cpp is included) against a fake workspace
// base, and checks, for every GEMM problem and segment, that the furthest
// element its loads can reach (rows clamped as the kernel clamps them, whole
// row vectors for row-contiguous operands) and the furthest element its
// output or split partials can reach lie inside the workspace region or the
// caller's in-place outputs; likewise every split-K reduction (and, for the
// vectorised reductions, the alignment and size they assume).  Also checks
// that interleaved launches have one tile grid and that no split is empty,
// and counts the interleaved launches and the problems with their own split
// count.
// Usage: plan_check nocc nvir precision [naux [plant | lno]]
//   plant: moves the split-K scratch to the end of the workspace, so that a
//   correct checker must report out-of-bounds partials.
//   lno: the plan of the LNO form (second cube S from the slot-projected
//   operands, the adjoint families' extra terms, cotangents of t2 and of the
//   projected t2 operands in the workspace).
// Usage: plan_check batched precision nocc1 nvir1 [nocc2 nvir2 ...]
//   the workspace layout of a batched backward call (plan_batch): every
//   instance's plan, built against its stream group's region part, stays
//   inside the part of the workspace that its own regions occupy (so that
//   groups never touch each other's memory), within its own in-place
//   outputs; each instance's descriptor blob lies after the region parts, at
//   the address its launches read descriptors from, without overlapping
//   another instance's blob, and inside the size the entry point reports.
//   batched_plant (in place of batched): shrinks the region parts by 256
//   bytes before the plans are built, so that neighbouring groups would
//   overlap and a correct checker must fail.
#include DRIVER
#include <algorithm>
#include <cstdio>
static const uintptr_t WS = 1ull << 40;   // fake workspace base (never dereferenced)
static const uintptr_t OUT_T2 = 1ull << 44, OUT_OVOV = 2ull << 44, OUT_STRIDE = 1ull << 38;

struct Stats {
  int nl = 0, ninter = 0, nsplit = 0, nvaried = 0;
  unsigned cfgs = 0;   // bit c: some GEMM launch uses tile configuration c
};

// Checks every GEMM and reduction access of the built plan of `c` against
// [ws_lo, ws_lo + ws_size) and the two in-place outputs (t2_size bytes each);
// returns the number of violations.
static int check_plan(const Ctx& c, uintptr_t ws_lo, size_t ws_size, uintptr_t t2_lo, uintptr_t ovov_lo,
                      size_t t2_size, Stats& st) {
  int bad = 0;
  auto check = [&](uintptr_t lo, uintptr_t hi, const char* what, size_t li, int p) {
    bool ok = (lo >= ws_lo && hi <= ws_lo + ws_size) || (lo >= t2_lo && hi <= t2_lo + t2_size) ||
              (lo >= ovov_lo && hi <= ovov_lo + t2_size);
    if (!ok && bad++ < 20)
      printf("OOB launch %zu prob %d %s: [%#lx, %#lx) ws [%#lx, %#lx)\n", li, p, what, lo, hi, ws_lo,
             ws_lo + ws_size);
  };
  for (size_t li = 0; li < c.launches.size(); li++) {
    const Launch& L = c.launches[li];
    if (L.kind == L_REDUCE) {
      const int pr = L.prec >= 0 ? L.prec : c.d.prec;
      const size_t e = pr == PREC_F64 ? 8 : 4;
      const ReduceProb* R = reinterpret_cast<const ReduceProb*>(c.blob.data.data() + L.off_a);
      long long total = 0;
      for (int p = 0; p < L.n_a; p++) {
        const ReduceProb& Q = R[p];
        if (Q.elem0 != total) { printf("reduce %zu prob %d: elem0 %lld != %lld\n", li, p, Q.elem0, total); bad++; }
        total += (long long)Q.nb * Q.M * Q.N;
        check((uintptr_t)Q.P, (uintptr_t)Q.P + (size_t)Q.ksplit * Q.nb * Q.M * Q.N * e, "reduce P", li, p);
        // vector reduction: 16-byte partial vectors, 32-bit element indices
        if (L.cfg && ((uintptr_t)Q.P % 16 != 0 || Q.N % (16 / e) != 0 || (long long)Q.nb * Q.M * Q.N >= (1LL << 31))) {
          printf("reduce %zu prob %d: vector reduction on an unaligned or oversized problem\n", li, p);
          bad++;
        }
        const size_t ce = Q.c_is_f64 ? 8 : 4;
        const long long cmax = (long long)(Q.nb - 1) * Q.c_sb + (long long)(Q.Mv - 1) * Q.c_sm + (long long)(Q.Nv - 1) * Q.c_sn;
        check((uintptr_t)Q.C, (uintptr_t)Q.C + (cmax + 1) * ce, "reduce C", li, p);
      }
      if (total != L.work) { printf("reduce %zu: %lld elements != work %lld\n", li, total, L.work); bad++; }
      continue;
    }
    if (L.kind != L_GEMM) continue;
    st.nl++;
    st.cfgs |= 1u << L.cfg;
    const int pr = L.prec >= 0 ? L.prec : c.d.prec;
    const size_t e = pr == PREC_F64 ? 8 : 4;
    const int vec = t3_gemm_vec(pr);
    const GemmProb* P = reinterpret_cast<const GemmProb*>(c.blob.data.data() + L.off_a);
    const GemmSeg* S = reinterpret_cast<const GemmSeg*>(c.blob.data.data() + L.off_b);
    long long blocks = 0;
    if (P[0].interleave) st.ninter++;
    for (int p = 0; p < L.n_a; p++) {
      const GemmProb& Q = P[p];
      blocks += (long long)Q.nb * Q.tiles_m * Q.tiles_n * Q.ksplit;
      // interleaved tiles (block t = tile t / nprob of problem t % nprob) need one tile grid
      if (Q.interleave != P[0].interleave ||
          (P[0].interleave && (Q.nb != P[0].nb || Q.tiles_m != P[0].tiles_m || Q.tiles_n != P[0].tiles_n ||
                               Q.ksplit != P[0].ksplit))) {
        printf("launch %zu prob %d: interleaved problems with different tile grids\n", li, p);
        bad++;
      }
      // every split of a problem holds at least one K-chunk
      if (Q.ksplit < 1 || (Q.nchunks > 0 && Q.ksplit > Q.nchunks)) {
        printf("launch %zu prob %d: %d splits of %d chunks\n", li, p, Q.ksplit, Q.nchunks);
        bad++;
      }
      if (Q.ksplit > 1) {
        st.nsplit++;
        if (Q.ksplit != P[0].ksplit) st.nvaried++;
      }
      for (int si = Q.seg0; si < Q.seg0 + Q.nseg; si++) {
        const GemmSeg& g = S[si];
        // rows: loads clamp to M-1 (KCONT) or to the last full vector block (ROWCONT)
        const long long mrow = g.a_mode == LOAD_ROWCONT ? ((Q.M - 1) / vec) * vec + vec - 1 : Q.M - 1;
        const long long nrow = g.b_mode == LOAD_ROWCONT ? ((Q.N - 1) / vec) * vec + vec - 1 : Q.N - 1;
        const long long amax = (long long)(Q.nb - 1) * g.a_sb + mrow * g.a_sm + (long long)(g.K1 - 1) * g.a_sk1 +
                               (long long)(g.K2 - 1) * g.a_sk2;
        const long long bmax = (long long)(Q.nb - 1) * g.b_sb + nrow * g.b_sn + (long long)(g.K1 - 1) * g.b_sk1 +
                               (long long)(g.K2 - 1) * g.b_sk2;
        check((uintptr_t)g.A, (uintptr_t)g.A + (amax + 1) * e, "A", li, p);
        check((uintptr_t)g.B, (uintptr_t)g.B + (bmax + 1) * e, "B", li, p);
        if (g.a_sb < 0 || g.a_sm < 0 || g.b_sb < 0 || g.b_sn < 0) { printf("negative stride\n"); bad++; }
      }
      if (Q.mode == GEMM_PARTIAL) {
        const size_t bytes = (size_t)Q.ksplit * Q.nb * Q.M * Q.N * e;
        check((uintptr_t)Q.P, (uintptr_t)Q.P + bytes, "P", li, p);
      } else {
        const size_t ce = Q.c_is_f64 ? 8 : 4;
        const long long cmax = (long long)(Q.nb - 1) * Q.c_sb + (long long)(Q.Mv - 1) * Q.c_sm + (long long)(Q.Nv - 1) * Q.c_sn;
        check((uintptr_t)Q.C, (uintptr_t)Q.C + (cmax + 1) * ce, "C", li, p);
      }
    }
    if (blocks != L.work) { printf("launch %zu: block count %lld != work %lld\n", li, blocks, L.work); bad++; }
  }
  return bad;
}

static void print_cfgs(const Stats& st) {
  printf(" cfgs=");
  for (int k = 0; k < 32; k++)
    if (st.cfgs >> k & 1) printf("%d,", k);
  printf("\n");
}

static int run_batched(int argc, char** argv, bool plant) {
  const int prec = atoi(argv[2]);
  const int n = (argc - 3) / 2;
  if (n < 1 || argc != 3 + 2 * n) {
    printf("usage: plan_check batched precision nocc nvir [nocc nvir ...]\n");
    return 2;
  }
  // fake, distinct, non-null caller pointers: instance x's t2 and ovov
  // cotangents at OUT_T2 / OUT_OVOV + x * OUT_STRIDE, its inputs elsewhere
  std::vector<PyscfadAmdInstance> inst(n);
  for (int x = 0; x < n; x++) {
    PyscfadAmdInstance& I = inst[x];
    std::memset(&I, 0, sizeof(I));
    I.nocc = atoi(argv[3 + 2 * x]);
    I.nvir = atoi(argv[4 + 2 * x]);
    I.ovvv_packed = 1;
    I.et_bar = 1.0;
    const double* in = reinterpret_cast<const double*>((3ull << 44) + x * OUT_STRIDE);
    I.mo_energy = I.t1 = I.t2 = I.fock = I.ovoo = I.ovov = I.ovvv = in;
    double* out = reinterpret_cast<double*>((4ull << 44) + x * OUT_STRIDE);
    I.t1_bar = I.fock_bar = I.mo_energy_bar = I.ovoo_bar = I.ovvv_bar = out;
    I.t2_bar = reinterpret_cast<double*>(OUT_T2 + x * OUT_STRIDE);
    I.ovov_bar = reinterpret_cast<double*>(OUT_OVOV + x * OUT_STRIDE);
  }
  BatchPlan bp;
  if (plan_batch(n, inst.data(), prec, 1, bp) != PYSCFAD_AMD_OK) {
    printf("plan_batch failed: %s\n", g_last_error.c_str());
    return 1;
  }
  const size_t need = pyscfad_amd_ccsd_t_batched_workspace_bytes(n, inst.data(), prec, 1);
  int bad = 0;
  if (need != bp.need()) { printf("workspace query %zu != plan %zu\n", need, bp.need()); bad++; }
  if (plant) bp.region_bytes -= 256;
  Stats st;
  std::vector<std::pair<size_t, size_t>> blobs;   // [offset, end) of each instance's descriptors
  for (int x = 0; x < n; x++) {
    Ctx& c = bp.ctx[x];
    const int o = inst[x].nocc, v = inst[x].nvir, g = bp.group[x];
    if (g < 0 || g >= bp.ngroups) { printf("instance %d: group %d of %d\n", x, g, bp.ngroups); bad++; continue; }
    // the instance's own regions, as a single call would size them
    Regions r0;
    int B0 = 1;
    plan_bytes(o, v, prec, true, &r0, &B0);
    if (B0 != c.B) { printf("instance %d: batch %d != %d\n", x, c.B, B0); bad++; }
    if (r0.desc > bp.region_bytes) {
      printf("instance %d: regions %zu exceed the part %zu\n", x, r0.desc, bp.region_bytes);
      bad++;
    }
    c.ws = reinterpret_cast<char*>(WS + bp.region_bytes * g);
    c.build_plan();
    if (!c.plan_error.empty()) { printf("instance %d: plan error '%s'\n", x, c.plan_error.c_str()); bad++; }
    // the launches read descriptors at ws + r.desc: that must be this instance's blob
    if (c.r.desc + bp.region_bytes * g != bp.blob_off[x]) {
      printf("instance %d: descriptors read at %zu, blob at %zu\n", x, c.r.desc + bp.region_bytes * g,
             bp.blob_off[x]);
      bad++;
    }
    if (bp.blob_off[x] < bp.region_bytes * bp.ngroups || bp.blob_off[x] % 256 != 0) {
      printf("instance %d: blob offset %zu inside the region parts or unaligned\n", x, bp.blob_off[x]);
      bad++;
    }
    blobs.emplace_back(bp.blob_off[x], bp.blob_off[x] + c.blob.data.size());
    const size_t t2_size = (size_t)o * o * v * v * 8;
    const int b = check_plan(c, WS + bp.region_bytes * g, r0.desc, OUT_T2 + x * OUT_STRIDE,
                             OUT_OVOV + x * OUT_STRIDE, t2_size, st);
    if (b) printf("instance %d (o=%d v=%d group %d): %d out of bounds\n", x, o, v, g, b);
    bad += b;
  }
  std::sort(blobs.begin(), blobs.end());
  for (size_t k = 0; k < blobs.size(); k++) {
    if (k + 1 < blobs.size() && blobs[k].second > blobs[k + 1].first) {
      printf("descriptor blobs overlap at %zu\n", blobs[k + 1].first);
      bad++;
    }
    if (blobs[k].second > bp.need()) { printf("descriptor blob ends at %zu past %zu\n", blobs[k].second, bp.need()); bad++; }
  }
  printf("batched n=%d prec=%d groups=%d region part=%zu need=%zu launches=%d gemm, interleaved=%d, bad=%d", n, prec,
         bp.ngroups, bp.region_bytes, bp.need(), st.nl, st.ninter, bad);
  print_cfgs(st);
  return bad ? 1 : 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::string(argv[1]) == "batched") return run_batched(argc, argv, false);
  if (argc > 1 && std::string(argv[1]) == "batched_plant") return run_batched(argc, argv, true);
  int o = atoi(argv[1]), v = atoi(argv[2]), prec = atoi(argv[3]);
  int naux = argc > 4 ? atoi(argv[4]) : 0;
  const bool plant = argc > 5 && std::string(argv[5]) == "plant";
  const bool lno = argc > 5 && std::string(argv[5]) == "lno";
  Ctx c;
  c.d = dims_for(o, v, prec, naux, true, 0, lno ? 1 : 0);
  c.B = choose_batch(c.d);
  c.backward = true;
  c.r = plan_regions(c.d, c.B, true);
  if (plant) c.r.split = c.r.desc - 4096;
  c.ws = reinterpret_cast<char*>(WS);
  c.t2_out = lno ? c.at<double>(c.r.lt2_bar) : reinterpret_cast<double*>(OUT_T2);
  c.ovov_out = reinterpret_cast<double*>(OUT_OVOV);
  if (lno) {
    c.t2Q_out = c.at<double>(c.r.lt2Q_bar);
    c.t2P0_out = c.at<double>(c.r.lt2P0_bar);
  }
  c.build_plan();
  const size_t ws_size = c.r.desc;   // regions end where descriptors start
  const size_t t2_size = (size_t)o * o * v * v * 8;
  Stats st;
  const int bad = check_plan(c, WS, ws_size, OUT_T2, OUT_OVOV, t2_size, st);
  printf("o=%d v=%d prec=%d naux=%d B=%d S=%d cube=%zu launches=%d gemm, interleaved=%d, split problems=%d "
         "(%d with their own count), bad=%d, plan_error='%s'", o, v, prec, naux, c.B, c.r.S,
         (size_t)c.d.V * c.d.V * c.d.V * c.d.elt, st.nl, st.ninter, st.nsplit, st.nvaried, bad,
         c.plan_error.c_str());
  print_cfgs(st);
  return bad ? 1 : 0;
}

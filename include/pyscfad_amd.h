/*
 * C ABI of the MI355X (gfx950) CCSD(T) perturbative-triples energy and VJP.
 *
 * All tensor arguments are DEVICE pointers to C-contiguous FP64 arrays in the
 * reference's own conventions (pyscfad/cc/ccsd.py:470-498, ccsd_t.py:88-106):
 *     mo_energy (n)            t1 (o,v)              t2 (o,o,v,v)
 *     fock (n,n)               ovoo (o,v,o,o)        ovov (o,v,o,v)
 *     ovvv (o,v,v(v+1)/2) lower-triangle packed over the last pair
 *          (ovvv_packed = 1), or (o,v,v,v) (ovvv_packed = 0)
 * with o = nocc, v = nvir, n = o + v.  Work is enqueued on `stream` (a
 * hipStream_t passed as void*; NULL is the default stream).  A call first
 * uploads its launch descriptors from host memory on `stream` and waits for
 * that upload (so for the work already on `stream`), then enqueues its
 * kernels and returns; a call that allocates its own workspace also
 * synchronises the stream before freeing it.  The batched call uploads from
 * pinned memory that it keeps and does not wait.
 *
 * Return value: PYSCFAD_AMD_OK or an error code; pyscfad_amd_last_error()
 * describes the last failure of the calling thread.
 */
#ifndef PYSCFAD_AMD_H
#define PYSCFAD_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYSCFAD_AMD_OK 0
#define PYSCFAD_AMD_ERR_ARGS 1        /* bad dimensions, null pointer, bad flag */
#define PYSCFAD_AMD_ERR_PRECISION 2   /* unknown precision mode */
#define PYSCFAD_AMD_ERR_OOM 3         /* device allocation failed */
#define PYSCFAD_AMD_ERR_WORKSPACE 4   /* caller workspace smaller than required */
#define PYSCFAD_AMD_ERR_HIP 5         /* a HIP runtime or launch error */

/* Arithmetic of the W/V contractions.  Denominators, the E(T) sum and every
 * cross-block cotangent accumulation are FP64 in all modes. */
#define PYSCFAD_AMD_PREC_F64 0        /* FP64 MFMA throughout: the drop-in path */
#define PYSCFAD_AMD_PREC_F32 1        /* FP32 MFMA contractions */
#define PYSCFAD_AMD_PREC_BF16X3 2     /* split-bf16 (hi*hi + hi*lo + lo*hi) MFMA, FP32 accumulate */
#define PYSCFAD_AMD_PREC_BF16X6 3     /* three-part split-bf16, six products, FP32 accumulate: FP32-level
                                         accuracy from the bf16 matrix cores */

/*
 * E(T) of one (t1, t2, eris) instance, written to e_out (DEVICE pointer to one
 * double).  Replaces the forward of the custom_vjp, pyscf's ccsd_t.kernel
 * called at pyscfad/cc/ccsd_t.py:41 (pyscf libcc CCsd_t_contract).
 * workspace may be NULL (the library allocates and frees it); otherwise it
 * must hold pyscfad_amd_ccsd_t_workspace_bytes(nocc, nvir, precision, 0) bytes.
 */
int pyscfad_amd_ccsd_t_energy(int nocc, int nvir, const double *mo_energy, const double *t1,
                              const double *t2, const double *fock, const double *ovoo,
                              const double *ovov, const double *ovvv, int ovvv_packed,
                              int precision, double *e_out, void *workspace,
                              size_t workspace_bytes, void *stream);

/*
 * Cotangents of E(T) for the upstream cotangent et_bar.  Replaces the
 * backward of the custom_vjp: the host preparation, block loop and
 * back-transformation of _ccsd_t_energy_vjp (pyscfad/cc/ccsd_t.py:82-192)
 * together with the C kernel it calls, ccsd_t_energy_vjp
 * (pyscfadlib/pyscfadlib/vjp/cc/ccsd_t_vjp.c:332-346).
 * Outputs (DEVICE, FP64, overwritten) have the reference's shapes and element
 * conventions (ccsd_t.py:179-192):
 *     t1_bar (o,v), t2_bar (o,o,v,v), fock_bar (n,n) (only the [o:, :o] block
 *     nonzero), mo_energy_bar (n), ovoo_bar (o,v,o,o), ovov_bar (o,v,o,v),
 *     ovvv_bar packed (o,v,v(v+1)/2) if ovvv_packed else (o,v,v,v).
 * e_out may be NULL; otherwise it receives E(T) of the same instance.
 */
int pyscfad_amd_ccsd_t_energy_vjp(int nocc, int nvir, const double *mo_energy, const double *t1,
                                  const double *t2, const double *fock, const double *ovoo,
                                  const double *ovov, const double *ovvv, int ovvv_packed,
                                  double et_bar, int precision, double *t1_bar, double *t2_bar,
                                  double *fock_bar, double *mo_energy_bar, double *ovoo_bar,
                                  double *ovov_bar, double *ovvv_bar, double *e_out,
                                  void *workspace, size_t workspace_bytes, void *stream);

/*
 * pyscfad_amd_ccsd_t_energy_vjp without a workspace argument and without any
 * device allocation.  No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Its scratch is one
 * file-scope device array of 1 GiB that each GPU the library runs on holds
 * until the process ends; the call picks the largest number of triples per
 * batch (up to the usual choice) whose plan fits it, and rejects an instance
 * that does not fit even at one triple per batch with
 * PYSCFAD_AMD_ERR_WORKSPACE (FP64 energy+VJP fits up to about nocc=16,
 * nvir=128).  Arguments, outputs and results are those of
 * pyscfad_amd_ccsd_t_energy_vjp (equal to rounding: the batch size may
 * differ).  All calls on one GPU share its array: calls that overlap on one
 * GPU (other host threads, other streams) race and must be serialised by
 * the caller.  The autotuner is never started from here.
 */
int pyscfad_amd_ccsd_t_energy_vjp_noalloc(int nocc, int nvir, const double *mo_energy, const double *t1,
                                          const double *t2, const double *fock, const double *ovoo,
                                          const double *ovov, const double *ovvv, int ovvv_packed,
                                          double et_bar, int precision, double *t1_bar, double *t2_bar,
                                          double *fock_bar, double *mo_energy_bar, double *ovoo_bar,
                                          double *ovov_bar, double *ovvv_bar, double *e_out, void *stream);

/* Workspace needed by the calls above (with_vjp = 0: energy, 1: energy_vjp);
 * 0 for invalid arguments.  No counterpart in the reference, whose C kernel
 * mallocs its own scratch (ccsd_t_vjp.c:352-444); this follows the XLA FFI
 * scratch-allocator pattern of pyscfadlib/pyscfadlib/ffi_helpers.h:40-50. */
size_t pyscfad_amd_ccsd_t_workspace_bytes(int nocc, int nvir, int precision, int with_vjp);

/*
 * Many independent instances of uneven size in one call.  The reference has
 * no batched form: its callers loop over molecules and call
 * pyscfad/cc/ccsd_t.py:28 (kernel) and ccsd_t.py:82 (_ccsd_t_energy_vjp) once
 * per instance; each field below carries the meaning of the single-instance
 * arguments above (device pointers, reference layouts).  With with_vjp = 0
 * the cotangent pointers are ignored and e_out is required.
 */
typedef struct PyscfadAmdInstance {
  int nocc, nvir, ovvv_packed;
  double et_bar;
  const double *mo_energy, *t1, *t2, *fock, *ovoo, *ovov, *ovvv;
  double *t1_bar, *t2_bar, *fock_bar, *mo_energy_bar, *ovoo_bar, *ovov_bar, *ovvv_bar;
  double *e_out; /* may be NULL when with_vjp = 1 */
} PyscfadAmdInstance;

/* Autotuner (no counterpart in the reference): picks the number of triples
 * per batch (the default or twice it), then times every tile configuration of
 * each contraction family (W build; the ovvv, t2-particle, ovoo and t2-hole
 * adjoints) on one batch of triples of an (nocc, nvir) problem in a scratch
 * workspace, and keeps the fastest for every later call of the same padded
 * size and precision in this process.  When the backward workspace does not
 * fit in device memory, only the W build is tuned (the others report -1).
 * cfg_out (may be NULL) receives 6 ints: the 5 chosen configuration ids and
 * the batch size.  The workspace sizes
 * above follow the tuned batch size: query them after tuning.  Setting the
 * environment variable PYSCFAD_AMD_AUTOTUNE=1 makes every call tune sizes
 * not seen yet. */
int pyscfad_amd_ccsd_t_autotune(int nocc, int nvir, int precision, int *cfg_out, void *stream);

/* Workspace of pyscfad_amd_ccsd_t_batched for these instances; 0 if invalid. */
size_t pyscfad_amd_ccsd_t_batched_workspace_bytes(int n, const PyscfadAmdInstance *instances,
                                                  int precision, int with_vjp);

/* Runs every instance as captured hipGraphs of at most 2048 kernel nodes
 * each.  The instances are dealt to up to four groups (largest F_E first, to
 * the group with the least work so far); a group's instances run one after
 * another on a library-created stream of its own (default priority), which is
 * forked from `stream` after the descriptor upload and joined back into it,
 * so the call is ordered on `stream` as a whole.  The graphs are cached and
 * replayed when the same call (same instance array contents, precision and
 * workspace) comes again, so new values in the same buffers cost only the
 * upload and the graph launches.  workspace is required (256-byte aligned, at
 * least the size above) and must not be used by anything else while the call
 * runs. */
int pyscfad_amd_ccsd_t_batched(int n, const PyscfadAmdInstance *instances, int precision,
                               int with_vjp, void *workspace, size_t workspace_bytes, void *stream);

/*
 * Work sharding over unique occupied triples (no counterpart in the
 * reference, whose unit of work is the virtual-triple block of
 * ccsd_t.py:145-177).  The unique triples i >= j >= k are enumerated with i
 * outermost, k innermost; pyscfad_amd_ccsd_t_num_triples(nocc) =
 * nocc(nocc+1)(nocc+2)/6 of them.  The _slice calls compute the share of
 * E(T) and of every cotangent contributed by triples [triple_begin,
 * triple_end); the shares of a partition of the list sum to the full result.
 * This is the unit one GPU of a multi-GPU run takes, and what a benchmark
 * times on a bounded sample of a large instance.
 */
long long pyscfad_amd_ccsd_t_num_triples(int nocc);
int pyscfad_amd_ccsd_t_energy_slice(int nocc, int nvir, const double *mo_energy, const double *t1,
                                    const double *t2, const double *fock, const double *ovoo,
                                    const double *ovov, const double *ovvv, int ovvv_packed,
                                    int precision, long long triple_begin, long long triple_end,
                                    double *e_out, void *workspace, size_t workspace_bytes,
                                    void *stream);
int pyscfad_amd_ccsd_t_energy_vjp_slice(int nocc, int nvir, const double *mo_energy, const double *t1,
                                        const double *t2, const double *fock, const double *ovoo,
                                        const double *ovov, const double *ovvv, int ovvv_packed,
                                        double et_bar, int precision, long long triple_begin,
                                        long long triple_end, double *t1_bar, double *t2_bar,
                                        double *fock_bar, double *mo_energy_bar, double *ovoo_bar,
                                        double *ovov_bar, double *ovvv_bar, double *e_out,
                                        void *workspace, size_t workspace_bytes, void *stream);

/*
 * Density-fitted variant: ovvv is not an input.  Its slabs are built on the
 * device, one occupied index at a time, from the DF factors of the reference's
 * DF-CCSD (pyscfad/cc/dfccsd.py:81-91):
 *     Lov (naux, o, v)            = Lpq[:, :o, o:]
 *     Lvv (naux, v(v+1)/2)        = pack_tril(Lpq[:, o:, o:])  (row-major lower triangle)
 *     ovvv[i, a, :] = sum_P Lov[P, i, a] Lvv[P, :]               (dfccsd.py:91)
 * so that the o*v^3 integrals and their FP64 cotangent are never stored: the
 * unique triples are processed in blocks of occupied indices and a block's
 * slabs live in a cache while its triples run.  The energy is that of
 * pyscfad_amd_ccsd_t_energy on ovvv = Lov^T Lvv (pyscfad/cc/ccsd_t.py:41 with
 * the DF eris).  The VJP returns, in place of ovvv_bar, the cotangents of Lov
 * and Lvv that the reference obtains by differentiating dfccsd.py:91 after
 * _ccsd_t_energy_vjp (ccsd_t.py:82-192); all other outputs are those of
 * pyscfad_amd_ccsd_t_energy_vjp.  Lov_bar (naux, o, v) and Lvv_bar (naux,
 * v(v+1)/2) are overwritten.  [triple_begin, triple_end) selects a slice of
 * the unique triples in this variant's blocked order (triple_end < 0: to the
 * end); the shares of a partition sum to the full result.  The environment
 * variable PYSCFAD_AMD_DF_BLOCK sets the block size (default: a slab cache of
 * about 40 GB).
 */
size_t pyscfad_amd_ccsd_t_df_workspace_bytes(int nocc, int nvir, int naux, int precision, int with_vjp);
int pyscfad_amd_ccsd_t_df_energy(int nocc, int nvir, int naux, const double *mo_energy, const double *t1,
                                 const double *t2, const double *fock, const double *ovoo, const double *ovov,
                                 const double *Lov, const double *Lvv, int precision, long long triple_begin,
                                 long long triple_end, double *e_out, void *workspace, size_t workspace_bytes,
                                 void *stream);
int pyscfad_amd_ccsd_t_df_energy_vjp(int nocc, int nvir, int naux, const double *mo_energy, const double *t1,
                                     const double *t2, const double *fock, const double *ovoo, const double *ovov,
                                     const double *Lov, const double *Lvv, double et_bar, int precision,
                                     long long triple_begin, long long triple_end, double *t1_bar, double *t2_bar,
                                     double *fock_bar, double *mo_energy_bar, double *ovoo_bar, double *ovov_bar,
                                     double *Lov_bar, double *Lvv_bar, double *e_out, void *workspace,
                                     size_t workspace_bytes, void *stream);

/*
 * LNO (T): the impurity (T) correction of the LNO method, with an occupied
 * projector.  Replaces the custom_vjp _ccsd_t_energy of pyscfad/lno/ccsd_t.py:
 * the energy replaces its forward (lno/ccsd_t.py:70-113, the C kernel
 * lnoccsdt_contract, pyscfadlib/pyscfadlib/vjp/cc/lno_ccsd_t.c:96-138, and
 * the factor 2/3 at lno/ccsd_t.py:111); the VJP replaces its backward
 * (lno/ccsd_t.py:121-235 with lnoccsdt_energy_vjp, lno_ccsd_t_vjp.c:400-566).
 * Inputs are the arguments of _ccsd_t_energy, in its order and layouts:
 *     mat (o,o)   t1T (v,o)   t2T (v,v,o,o)   mo_energy (n)   fvo (v,o)
 *     ovoo (o,v,o,o)   ovov (o,v,o,v)   ovvv packed (o,v,v(v+1)/2) if ovvv_packed
 *     (as lno/ccsd_t.py:62-68 reads it), else (o,v,v,v)
 * E = 2/3 sum_{a, b>=c} g_bc sum_ijk (sum_l mat[l,i] w0[l,j,k]) z0[i,j,k] / D
 * (g_bc = 1/2 for b == c, lno_ccsd_t.c:16-28, 65-69); with mat = I it is the
 * E(T) of pyscfad_amd_ccsd_t_energy.  mat is any o x o matrix (the reference
 * forms ulo^T ulo, lno/ccsd_t.py:42); mat_bar is not symmetrised.
 * Outputs (DEVICE, FP64, overwritten) are the tuple the backward returns
 * (lno/ccsd_t.py:235): mat_bar (o,o), t1T_bar (v,o), t2T_bar (v,v,o,o),
 * mo_energy_bar (n), fvo_bar (v,o), ovoo_bar, ovov_bar, ovvv_bar (packed like
 * the input).  e_out: the energy (may be NULL for the VJP).  nocc <= 128.
 * The workspace (NULL: the library allocates) must hold
 * pyscfad_amd_lno_ccsd_t_workspace_bytes(nocc, nvir, precision, with_vjp)
 * bytes; that function returns 0 for invalid arguments.
 */
size_t pyscfad_amd_lno_ccsd_t_workspace_bytes(int nocc, int nvir, int precision, int with_vjp);
int pyscfad_amd_lno_ccsd_t_energy(int nocc, int nvir, const double *mat, const double *t1T, const double *t2T,
                                  const double *mo_energy, const double *fvo, const double *ovoo,
                                  const double *ovov, const double *ovvv, int ovvv_packed, int precision,
                                  double *e_out, void *workspace, size_t workspace_bytes, void *stream);
int pyscfad_amd_lno_ccsd_t_energy_vjp(int nocc, int nvir, const double *mat, const double *t1T, const double *t2T,
                                      const double *mo_energy, const double *fvo, const double *ovoo,
                                      const double *ovov, const double *ovvv, int ovvv_packed, double et_bar,
                                      int precision, double *mat_bar, double *t1T_bar, double *t2T_bar,
                                      double *mo_energy_bar, double *fvo_bar, double *ovoo_bar, double *ovov_bar,
                                      double *ovvv_bar, double *e_out, void *workspace, size_t workspace_bytes,
                                      void *stream);

/* Text of the last error of the calling thread ("" if none). */
const char *pyscfad_amd_last_error(void);

/* Kernel timing of the calling thread's calls (no counterpart in the
 * reference; for benchmarks and profiling).  With timing enabled every launch
 * is bracketed by hipEvents on the call's stream and the call synchronises
 * that stream before returning.  pyscfad_amd_last_timing fills up to n
 * entries of device milliseconds and launch counts of the last call, by
 * kernel family: 0 grouped MFMA GEMM, 1 orbit epilogue (+ its reduction),
 * 2 V-term adjoint kernels, 3 split-K reductions, 4 per-call setup (accumulator
 * clears, input re-layout, cotangent write-back), and family 0
 * split by contraction: 5 W build, 6 ovvv adjoint, 7 t2 adjoint (particle
 * term), 8 ovoo adjoint, 9 t2 adjoint (hole term), 10 density-fitted slab
 * builds, cotangent folds and slab clears (also counted in 0 and 3); returns
 * the count filled. */
void pyscfad_amd_set_timing(int enable);
int pyscfad_amd_last_timing(double *ms, long long *launches, int n);

/*
 * XLA custom-call handlers for jax on ROCm, legacy form (api_version 0):
 *     void handler(hipStream_t stream, void **buffers, const char *opaque, size_t opaque_len)
 * registered with jax.ffi.register_ffi_target(..., platform="ROCM", api_version=0).
 * They replace the two calls of the reference's custom_vjp: the forward
 * (pyscfad/cc/ccsd_t.py:41) and the backward's ctypes call of
 * ccsd_t_energy_vjp (ccsd_t.py:107-133, ccsd_t_vjp.c:332-346).  The
 * reference registers its GPU handlers through typed XLA FFI
 * (pyscfadlib/pyscfadlib/cuda/solver_kernels.cc:144-157); the legacy form is
 * used here because it needs no XLA headers.
 *   energy:     buffers = {mo_energy, t1, t2, fock, ovoo, ovov, ovvv, e_out[, workspace]}
 *   energy_vjp: buffers = {mo_energy, t1, t2, fock, ovoo, ovov, ovvv, et_bar,
 *                          t1_bar, t2_bar, fock_bar, mo_energy_bar, ovoo_bar, ovov_bar, ovvv_bar[, workspace]}
 * et_bar is a device FP64 scalar operand (the backward's traced cotangent,
 * applied on the stream).  With workspace_bytes > 0 the last buffer is a
 * caller scratch result of that size (at least
 * pyscfad_amd_ccsd_t_workspace_bytes, 256-byte aligned), so a call makes no
 * device allocation (it still waits for its descriptor upload); with 0 the
 * library allocates and frees its own, which synchronises the stream.
 * The legacy ABI has no error channel: failures are printed on stderr.
 */
typedef struct {
    int nocc, nvir, precision, ovvv_packed;
    unsigned long long workspace_bytes;
} PyscfadAmdXlaOpaque;

void pyscfad_amd_xla_ccsd_t_energy(void *stream, void **buffers, const char *opaque, size_t opaque_len);
void pyscfad_amd_xla_ccsd_t_energy_vjp(void *stream, void **buffers, const char *opaque, size_t opaque_len);

/* Counters of the batched entry point's graph cache since the library was
 * loaded (no counterpart in the reference; for tests and diagnostics): calls
 * served by replaying a cached graph, and calls that captured new graphs.
 * Either pointer may be NULL. */
void pyscfad_amd_graph_cache_stats(long long *replays, long long *captures);

/* ABI version, currently 2 (2: et_bar operand and workspace result of the XLA handlers). */
int pyscfad_amd_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PYSCFAD_AMD_H */

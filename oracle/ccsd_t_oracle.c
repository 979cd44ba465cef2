/*
 * CPU oracle for the CCSD(T) perturbative-triples energy and its VJP.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the FP64 checker that the GPU
 * kernels are compared against, and the timed CPU baseline of the bench.
 * The product path (pyscfad_amd) never links or calls it.
 *
 * It restates, in plain C99 + OpenMP with its own loops (no BLAS), the
 * algorithm of the reference:
 *   - virtual-triple jobs a >= b >= c          (pyscfadlib/.../vjp/cc/ccsd_t.c:6-72)
 *   - the six (p,q,r) orderings of (a,b,c) and the matching occupied index
 *     permutations sigma_s                     (ccsd_t_vjp.c:32-37, ccsd_t.c:74-96)
 *   - W and V of one ordering                  (ccsd_t.c:121-152, get_wv)
 *   - z0 = r3(w0 + v0)                         (ccsd_t.c:98-119, add_and_permute)
 *   - d3 = f/D with f = 1/6, 1/2, 1            (ccsd_t_vjp.c:42-53, 298-304)
 *   - E(T) = 2 * sum w0 z0 d3                  (lno_ccsd_t.c:6-29 energy form;
 *                                               factor 2 as in ccsd_t_slow.py:124)
 *   - the reverse pass                         (ccsd_t_vjp.c:55-117, 139-269, 271-329)
 *
 * Array conventions are the reference's host arrays (ccsd_t.py:88-106):
 *   t1T[a,i]        = t1[i,a]                       shape (v,o)
 *   t2T[a,b,i,j]    = t2[j,i,a,b]                   shape (v,v,o,o)
 *   vooo[a,i,j,k]   = ovoo[i,a,k,j]                 shape (v,o,o,o)
 *   fvo[a,i]        = fock[o+a,i]                   shape (v,o)
 *   vvop[a,b,i,j]   = ovov[i,a,j,b]   (j <  o)      shape (v,v,o,o+v)
 *   vvop[a,b,i,o+f] = ovvv[i,a,f,b]   (unpacked)
 * Outputs are accumulated (+=) into caller-zeroed buffers of the same shapes,
 * as the reference driver does (ccsd_t.py:89-106).
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* (p,q,r) for s = 0..5 as positions into (a,b,c): (abc),(acb),(bac),(bca),(cab),(cba) */
static const int PERM_ABC[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

/* flat index of sigma_s(i,j,k): (ijk),(ikj),(jik),(kij),(jki),(kji) */
static inline size_t sigma_index(int s, int i, int j, int k, int o)
{
    int x, y, z;
    switch (s) {
    case 0: x = i; y = j; z = k; break;
    case 1: x = i; y = k; z = j; break;
    case 2: x = j; y = i; z = k; break;
    case 3: x = k; y = i; z = j; break;
    case 4: x = j; y = k; z = i; break;
    default: x = k; y = j; z = i; break;
    }
    return ((size_t)x * o + y) * o + z;
}

/* r3(x)[ijk] = 4x[ijk] + x[jki] + x[kij] - 2x[kji] - 2x[ikj] - 2x[jik]; r3 is self-adjoint */
static void r3_apply(double *out, const double *x, int o)
{
    size_t oo = (size_t)o * o;
    for (int i = 0; i < o; i++)
        for (int j = 0; j < o; j++)
            for (int k = 0; k < o; k++) {
                out[i * oo + j * o + k] = 4.0 * x[i * oo + j * o + k]
                                        + x[j * oo + k * o + i]
                                        + x[k * oo + i * o + j]
                                        - 2.0 * x[k * oo + j * o + i]
                                        - 2.0 * x[i * oo + k * o + j]
                                        - 2.0 * x[j * oo + i * o + k];
            }
}

typedef struct {
    int o, v, n;
    const double *mo_energy, *t1T, *t2T, *vooo, *fvo, *vvop;
} T3In;

/* W_s and V_s of one ordering (p,q,r), scattered into w0/v0 through sigma_s.
 * W[i,j,k] = sum_f vvop[p,q,i,o+f] t2T[r,f,j,k] - sum_m vooo[p,i,j,m] t2T[r,q,m,k]
 * V[i,j,k] = 1/2 vvop[p,q,i,j] t1T[r,k] + 1/2 t2T[q,p,i,j] fvo[r,k]          */
static void build_wv(const T3In *in, int p, int q, int r, int s,
                     double *w0, double *v0, double *wtmp)
{
    const int o = in->o, v = in->v, n = in->n;
    const size_t oo = (size_t)o * o, ooo = oo * o, voo = (size_t)v * oo;
    const double *pq = in->vvop + ((size_t)p * v + q) * o * n;   /* vvop[p,q] : (o, n) */
    const double *t2r = in->t2T + (size_t)r * voo;                /* t2T[r] : (v, o, o) */
    const double *t2rq = t2r + (size_t)q * oo;                    /* t2T[r,q] : (o, o) */
    const double *voop = in->vooo + (size_t)p * ooo;             /* vooo[p] : (o,o,o) */
    memset(wtmp, 0, sizeof(double) * ooo);
    for (int i = 0; i < o; i++) {
        double *wi = wtmp + i * oo;
        for (int f = 0; f < v; f++) {
            const double x = pq[(size_t)i * n + o + f];
            const double *t = t2r + (size_t)f * oo;
            for (size_t jk = 0; jk < oo; jk++) wi[jk] += x * t[jk];
        }
        for (int j = 0; j < o; j++) {
            double *wij = wi + j * o;
            for (int m = 0; m < o; m++) {
                const double x = voop[(i * o + j) * o + m];
                const double *t = t2rq + (size_t)m * o;
                for (int k = 0; k < o; k++) wij[k] -= x * t[k];
            }
        }
    }
    const double *t2qp = in->t2T + ((size_t)q * v + p) * oo;     /* t2T[q,p] : (o,o) */
    const double *t1r = in->t1T + (size_t)r * o;
    const double *fr = in->fvo + (size_t)r * o;
    for (int i = 0; i < o; i++)
        for (int j = 0; j < o; j++)
            for (int k = 0; k < o; k++) {
                size_t d = sigma_index(s, i, j, k, o);
                w0[d] += wtmp[(i * o + j) * o + k];
                v0[d] += 0.5 * pq[(size_t)i * n + j] * t1r[k] + 0.5 * t2qp[i * o + j] * fr[k];
            }
}

static inline double triple_weight(int a, int b, int c)
{
    if (a == c) return 1.0 / 6.0;
    if (a == b || b == c) return 0.5;
    return 1.0;
}

/* w0 and z0 = r3(w0+v0) of the virtual triple (a,b,c), a >= b >= c. */
static void job_w_z(const T3In *in, int a, int b, int c,
                    double *w0, double *z0, double *buf)
{
    const int o = in->o;
    const size_t ooo = (size_t)o * o * o;
    const int abc[3] = {a, b, c};
    double *v0 = buf, *wtmp = buf + ooo;
    memset(w0, 0, sizeof(double) * ooo);
    memset(v0, 0, sizeof(double) * ooo);
    for (int s = 0; s < 6; s++)
        build_wv(in, abc[PERM_ABC[s][0]], abc[PERM_ABC[s][1]], abc[PERM_ABC[s][2]], s, w0, v0, wtmp);
    for (size_t x = 0; x < ooo; x++) v0[x] += w0[x];
    r3_apply(z0, v0, o);
}

/* number of jobs a >= b >= c with a < v */
static long njobs_total(int v) { return (long)v * (v + 1) * (v + 2) / 6; }

static void job_decode(long m, int *pa, int *pb, int *pc)
{
    /* jobs are ordered by a, then b <= a, then c <= b */
    int a = 0;
    while ((long)(a + 1) * (a + 2) * (a + 3) / 6 <= m) a++;
    m -= (long)a * (a + 1) * (a + 2) / 6;
    int b = 0;
    while ((long)(b + 1) * (b + 2) / 2 <= m) b++;
    m -= (long)b * (b + 1) / 2;
    *pa = a; *pb = b; *pc = (int)m;
}

/*
 * E(T) over the job range [job0, job1) of the a>=b>=c enumeration (the full
 * energy is job0 = 0, job1 = v(v+1)(v+2)/6).  A sub-range is what the bench
 * times as a bounded CPU sample.  Returns the partial E(T) (already x2).
 */
double oracle_ccsd_t_energy_range(int nocc, int nvir, const double *mo_energy,
                                  const double *t1T, const double *t2T,
                                  const double *vooo, const double *fvo,
                                  const double *vvop, long job0, long job1)
{
    T3In in = {nocc, nvir, nocc + nvir, mo_energy, t1T, t2T, vooo, fvo, vvop};
    const int o = nocc;
    const size_t ooo = (size_t)o * o * o;
    if (job1 > njobs_total(nvir)) job1 = njobs_total(nvir);
    double et = 0.0;
#pragma omp parallel reduction(+ : et)
    {
        double *w0 = malloc(sizeof(double) * ooo * 4);
        double *z0 = w0 + ooo, *buf = z0 + ooo;
#pragma omp for schedule(dynamic, 4)
        for (long m = job0; m < job1; m++) {
            int a, b, c;
            job_decode(m, &a, &b, &c);
            job_w_z(&in, a, b, c, w0, z0, buf);
            const double eabc = mo_energy[o + a] + mo_energy[o + b] + mo_energy[o + c];
            const double fac = triple_weight(a, b, c);
            double e = 0.0;
            for (int i = 0; i < o; i++)
                for (int j = 0; j < o; j++)
                    for (int k = 0; k < o; k++) {
                        size_t x = ((size_t)i * o + j) * o + k;
                        double d3 = fac / (mo_energy[i] + mo_energy[j] + mo_energy[k] - eabc);
                        e += w0[x] * z0[x] * d3;
                    }
            et += e;
        }
        free(w0);
    }
    return 2.0 * et;
}

double oracle_ccsd_t_energy(int nocc, int nvir, const double *mo_energy,
                            const double *t1T, const double *t2T,
                            const double *vooo, const double *fvo, const double *vvop)
{
    return oracle_ccsd_t_energy_range(nocc, nvir, mo_energy, t1T, t2T, vooo, fvo, vvop,
                                      0, njobs_total(nvir));
}

/*
 * Reverse pass.  et_bar is the cotangent of E(T); the reference doubles it
 * first (ccsd_t.py:86), which is done here too.  All outputs are accumulated.
 * Small cotangents (mo_energy, t1T, fvo, vooo) use per-thread private copies
 * reduced at the end, as the reference does (ccsd_t_vjp.c:379-493).  The two
 * large ones (vvop_bar, t2T_bar) are accumulated in place under one lock per
 * (p,q) block of vvop and one per leading index of t2T, which keeps memory
 * bounded at the sizes the bench samples.
 */
void oracle_ccsd_t_energy_vjp_range(int nocc, int nvir, const double *mo_energy,
                                    const double *t1T, const double *t2T,
                                    const double *vooo, const double *fvo, const double *vvop,
                                    double et_bar,
                                    double *mo_energy_bar, double *t1T_bar, double *t2T_bar,
                                    double *vooo_bar, double *fvo_bar, double *vvop_bar,
                                    long job0, long job1)
{
    T3In in = {nocc, nvir, nocc + nvir, mo_energy, t1T, t2T, vooo, fvo, vvop};
    const int o = nocc, v = nvir, n = nocc + nvir;
    const size_t oo = (size_t)o * o, ooo = oo * o, voo = (size_t)v * oo;
    const size_t len_me = n, len_t1 = (size_t)v * o;
    const size_t len_vooo = (size_t)v * ooo, len_fvo = (size_t)v * o;
    const size_t len_priv = len_me + len_t1 + len_vooo + len_fvo;
    const double ebar = 2.0 * et_bar;
    if (job1 > njobs_total(v)) job1 = njobs_total(v);
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
    omp_lock_t *lk_pq = malloc(sizeof(omp_lock_t) * (size_t)v * v);
    omp_lock_t *lk_r = malloc(sizeof(omp_lock_t) * (size_t)v);
    for (size_t x = 0; x < (size_t)v * v; x++) omp_init_lock(&lk_pq[x]);
    for (int x = 0; x < v; x++) omp_init_lock(&lk_r[x]);
#define LOCK(l) omp_set_lock(&(l))
#define UNLOCK(l) omp_unset_lock(&(l))
#else
#define LOCK(l)
#define UNLOCK(l)
#endif
    double **priv = calloc(nthreads, sizeof(double *));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *acc = calloc(len_priv, sizeof(double));
        priv[tid] = acc;
        double *me_b = acc, *t1_b = me_b + len_me;
        double *vooo_b = t1_b + len_t1, *fvo_b = vooo_b + len_vooo;
        double *w0 = malloc(sizeof(double) * ooo * 8);
        double *z0 = w0 + ooo, *d3 = z0 + ooo, *zb = d3 + ooo, *wb0 = zb + ooo, *vb0 = wb0 + ooo;
        double *wbar = vb0 + ooo, *vbar = wbar + ooo;
        double *buf = malloc(sizeof(double) * ooo * 2);
        double *pqb_loc = malloc(sizeof(double) * o * n);
        double *t2b_loc = malloc(sizeof(double) * voo);
        double *t2qp_loc = malloc(sizeof(double) * oo);
#pragma omp for schedule(dynamic, 4)
        for (long m = job0; m < job1; m++) {
            int a, b, c;
            job_decode(m, &a, &b, &c);
            const int abc[3] = {a, b, c};
            job_w_z(&in, a, b, c, w0, z0, buf);
            const double eabc = mo_energy[o + a] + mo_energy[o + b] + mo_energy[o + c];
            const double fac = triple_weight(a, b, c);
            double tsum = 0.0;
            for (int i = 0; i < o; i++)
                for (int j = 0; j < o; j++)
                    for (int k = 0; k < o; k++) {
                        size_t x = ((size_t)i * o + j) * o + k;
                        double d = fac / (mo_energy[i] + mo_energy[j] + mo_energy[k] - eabc);
                        d3[x] = d;
                        zb[x] = ebar * w0[x] * d;
                        double d3b = ebar * w0[x] * z0[x];
                        double t = d3b * d * d / fac;
                        me_b[i] -= t; me_b[j] -= t; me_b[k] -= t;
                        tsum += t;
                    }
            me_b[o + a] += tsum; me_b[o + b] += tsum; me_b[o + c] += tsum;
            r3_apply(vb0, zb, o);                                   /* v0_bar = r3(z_bar) */
            for (size_t x = 0; x < ooo; x++) wb0[x] = ebar * z0[x] * d3[x] + vb0[x];
            for (int s = 0; s < 6; s++) {
                const int p = abc[PERM_ABC[s][0]], q = abc[PERM_ABC[s][1]], r = abc[PERM_ABC[s][2]];
                for (int i = 0; i < o; i++)
                    for (int j = 0; j < o; j++)
                        for (int k = 0; k < o; k++) {
                            size_t d = sigma_index(s, i, j, k, o);
                            wbar[(i * o + j) * o + k] = wb0[d];
                            vbar[(i * o + j) * o + k] = vb0[d];
                        }
                const double *pq = vvop + ((size_t)p * v + q) * o * n;
                const double *t2r = t2T + (size_t)r * voo;
                const double *t2rq = t2r + (size_t)q * oo;
                const double *voop = vooo + (size_t)p * ooo;
                double *voop_b = vooo_b + (size_t)p * ooo;
                memset(pqb_loc, 0, sizeof(double) * o * n);
                memset(t2b_loc, 0, sizeof(double) * voo);
                memset(t2qp_loc, 0, sizeof(double) * oo);
                /* particle: vvop_bar[p,q,i,o+f] += sum_jk t2T[r,f,j,k] Wb[i,j,k]
                 *           t2T_bar[r,f,j,k]   += sum_i Wb[i,j,k] vvop[p,q,i,o+f] */
                for (int i = 0; i < o; i++) {
                    const double *wi = wbar + i * oo;
                    for (int f = 0; f < v; f++) {
                        const double *t = t2r + (size_t)f * oo;
                        double *tb = t2b_loc + (size_t)f * oo;
                        const double x = pq[(size_t)i * n + o + f];
                        double acc1 = 0.0;
                        for (size_t jk = 0; jk < oo; jk++) {
                            acc1 += t[jk] * wi[jk];
                            tb[jk] += x * wi[jk];
                        }
                        pqb_loc[(size_t)i * n + o + f] += acc1;
                    }
                }
                /* hole: vooo_bar[p,i,j,m] -= sum_k t2T[r,q,m,k] Wb[i,j,k]
                 *       t2T_bar[r,q,m,k]  -= sum_ij Wb[i,j,k] vooo[p,i,j,m] */
                double *t2rq_loc = t2b_loc + (size_t)q * oo;
                for (int i = 0; i < o; i++)
                    for (int j = 0; j < o; j++) {
                        const double *wij = wbar + (i * o + j) * o;
                        for (int mm = 0; mm < o; mm++) {
                            const double *t = t2rq + (size_t)mm * o;
                            double *tb = t2rq_loc + (size_t)mm * o;
                            const double x = voop[(i * o + j) * o + mm];
                            double acc1 = 0.0;
                            for (int k = 0; k < o; k++) {
                                acc1 += t[k] * wij[k];
                                tb[k] -= x * wij[k];
                            }
                            voop_b[(i * o + j) * o + mm] -= acc1;
                        }
                    }
                /* V part */
                const double *t1r = t1T + (size_t)r * o;
                const double *fr = fvo + (size_t)r * o;
                const double *t2qp = t2T + ((size_t)q * v + p) * oo;
                double *t1r_b = t1_b + (size_t)r * o;
                double *fr_b = fvo_b + (size_t)r * o;
                for (int i = 0; i < o; i++)
                    for (int j = 0; j < o; j++) {
                        const double *vij = vbar + (i * o + j) * o;
                        double s1 = 0.0, s2 = 0.0;
                        const double x1 = pq[(size_t)i * n + j], x2 = t2qp[i * o + j];
                        for (int k = 0; k < o; k++) {
                            s1 += vij[k] * t1r[k];
                            s2 += vij[k] * fr[k];
                            t1r_b[k] += 0.5 * vij[k] * x1;
                            fr_b[k] += 0.5 * vij[k] * x2;
                        }
                        pqb_loc[(size_t)i * n + j] += 0.5 * s1;
                        t2qp_loc[i * o + j] += 0.5 * s2;
                    }
                /* flush the job-local blocks into the shared cotangents */
                double *pq_b = vvop_bar + ((size_t)p * v + q) * o * n;
                LOCK(lk_pq[(size_t)p * v + q]);
                for (size_t x = 0; x < (size_t)o * n; x++) pq_b[x] += pqb_loc[x];
                UNLOCK(lk_pq[(size_t)p * v + q]);
                double *t2r_b = t2T_bar + (size_t)r * voo;
                LOCK(lk_r[r]);
                for (size_t x = 0; x < voo; x++) t2r_b[x] += t2b_loc[x];
                UNLOCK(lk_r[r]);
                double *t2qp_b = t2T_bar + ((size_t)q * v + p) * oo;
                LOCK(lk_r[q]);
                for (size_t x = 0; x < oo; x++) t2qp_b[x] += t2qp_loc[x];
                UNLOCK(lk_r[q]);
            }
        }
        free(t2qp_loc);
        free(t2b_loc);
        free(pqb_loc);
        free(buf);
        free(w0);
    }
    /* ordered reduction of the per-thread accumulators into the caller's buffers */
    double *outs[4] = {mo_energy_bar, t1T_bar, vooo_bar, fvo_bar};
    size_t lens[4] = {len_me, len_t1, len_vooo, len_fvo};
    size_t off = 0;
    for (int x = 0; x < 4; x++) {
        double *dst = outs[x];
        const size_t len = lens[x];
#pragma omp parallel for schedule(static)
        for (size_t e = 0; e < len; e++) {
            double sacc = 0.0;
            for (int t = 0; t < nthreads; t++)
                if (priv[t]) sacc += priv[t][off + e];
            dst[e] += sacc;
        }
        off += len;
    }
    for (int t = 0; t < nthreads; t++) free(priv[t]);
    free(priv);
#ifdef _OPENMP
    for (size_t x = 0; x < (size_t)v * v; x++) omp_destroy_lock(&lk_pq[x]);
    for (int x = 0; x < v; x++) omp_destroy_lock(&lk_r[x]);
    free(lk_pq);
    free(lk_r);
#endif
#undef LOCK
#undef UNLOCK
}

void oracle_ccsd_t_energy_vjp(int nocc, int nvir, const double *mo_energy,
                              const double *t1T, const double *t2T,
                              const double *vooo, const double *fvo, const double *vvop,
                              double et_bar,
                              double *mo_energy_bar, double *t1T_bar, double *t2T_bar,
                              double *vooo_bar, double *fvo_bar, double *vvop_bar)
{
    oracle_ccsd_t_energy_vjp_range(nocc, nvir, mo_energy, t1T, t2T, vooo, fvo, vvop, et_bar,
                                   mo_energy_bar, t1T_bar, t2T_bar, vooo_bar, fvo_bar, vvop_bar,
                                   0, njobs_total(nvir));
}

long oracle_njobs(int nvir) { return njobs_total(nvir); }

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/*
 * dqmc_oracle.c -- CPU restatement of the reference's finite-temperature DQMC sweep.
 *
 * TEST INFRASTRUCTURE ONLY (see dqmc_oracle.h).  Plain C99, single thread, complex128.
 * Each routine names the reference file:line it restates; paths are relative to the
 * YuanyueChen/sunhubbardmy source tree.  LAPACK calls of the reference (zgeqp3, zgeqrf,
 * zungqr, zgetrf, ztrtri, dznrm2, idamax, izamax) are restated from their published
 * algorithms (LAPACK 3.x reference implementation) without the blocking, which changes
 * rounding only.
 */
#include "dqmc_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ------------------------------------------------------------------------------------ */
/* SFMT19937: lib/m_spring.f90:187-379 (Saito & Matsumoto's SIMD-oriented Fast MT).       */
/* 156 128-bit words = 624 uint32 lanes; lane 0 is the least significant word.           */
/* ------------------------------------------------------------------------------------ */
#define SFMT_N32 624
#define SFMT_NW 156
#define SFMT_POS1 122

typedef struct { uint32_t s[SFMT_N32]; int idx; } sfmt_t;

static const uint32_t sfmt_parity[4] = {0x00000001u, 0x00000000u, 0x00000000u, 0x13c9e684u};
static const uint32_t sfmt_msk[4] = {0xdfffffefu, 0xddfecb7fu, 0xbffaffffu, 0xbffffff6u};

/* m_spring.f90:187-243: Knuth-style linear init, then period certification */
static void sfmt_init(sfmt_t *g, int32_t seed)
{
    g->s[0] = (uint32_t)seed;
    for (int i = 1; i < SFMT_N32; i++)
        g->s[i] = 1812433253u * (g->s[i - 1] ^ (g->s[i - 1] >> 30)) + (uint32_t)i;
    uint32_t inner = 0;
    for (int i = 0; i < 4; i++) inner ^= g->s[i] & sfmt_parity[i];
    for (int sh = 16; sh > 0; sh >>= 1) inner ^= inner >> sh;
    if ((inner & 1u) != 1u) {
        int done = 0;
        for (int i = 0; i < 4 && !done; i++) {
            uint32_t work = 1;
            for (int j = 0; j < 32; j++) {
                if (work & sfmt_parity[i]) { g->s[i] ^= work; done = 1; break; }
                work <<= 1;
            }
        }
    }
    g->idx = SFMT_N32;
}

/* 128-bit shifts by whole bytes, lanes little-endian */
static void sfmt_lsh8(uint32_t o[4], const uint32_t a[4])
{
    uint64_t hi = ((uint64_t)a[3] << 32) | a[2], lo = ((uint64_t)a[1] << 32) | a[0];
    uint64_t oh = (hi << 8) | (lo >> 56), ol = lo << 8;
    o[0] = (uint32_t)ol; o[1] = (uint32_t)(ol >> 32); o[2] = (uint32_t)oh; o[3] = (uint32_t)(oh >> 32);
}
static void sfmt_rsh8(uint32_t o[4], const uint32_t a[4])
{
    uint64_t hi = ((uint64_t)a[3] << 32) | a[2], lo = ((uint64_t)a[1] << 32) | a[0];
    uint64_t ol = (lo >> 8) | (hi << 56), oh = hi >> 8;
    o[0] = (uint32_t)ol; o[1] = (uint32_t)(ol >> 32); o[2] = (uint32_t)oh; o[3] = (uint32_t)(oh >> 32);
}

/* m_spring.f90:290-379: one refill of the whole state (recursion r = a ^ a<<8 ^ (b>>11 & msk) ^ c>>8 ^ d<<18) */
static void sfmt_refill(sfmt_t *g)
{
    uint32_t *w = g->s;
    const uint32_t *r1 = w + 4 * (SFMT_NW - 2), *r2 = w + 4 * (SFMT_NW - 1);
    for (int i = 0; i < SFMT_NW; i++) {
        uint32_t *a = w + 4 * i;
        const uint32_t *b = w + 4 * ((i + SFMT_POS1) % SFMT_NW);
        uint32_t x[4], y[4];
        sfmt_lsh8(x, a);
        sfmt_rsh8(y, r1);
        for (int l = 0; l < 4; l++)
            a[l] = a[l] ^ x[l] ^ ((b[l] >> 11) & sfmt_msk[l]) ^ y[l] ^ (r2[l] << 18);
        r1 = r2;
        r2 = a;
    }
}

static uint32_t sfmt_u32(sfmt_t *g)
{
    if (g->idx >= SFMT_N32) { sfmt_refill(g); g->idx = 0; }
    return g->s[g->idx++];
}

/* m_spring.f90:245-254: (u + 0.5) / 2^32, in (0,1) */
static double sfmt_stream(sfmt_t *g) { return ((double)sfmt_u32(g) + 0.5) * (1.0 / 4294967296.0); }

void orc_sfmt_u32(int32_t seed, int n, uint32_t *out)
{
    sfmt_t g; sfmt_init(&g, seed);
    for (int i = 0; i < n; i++) out[i] = sfmt_u32(&g);
}
void orc_sfmt_stream(int32_t seed, int n, double *out)
{
    sfmt_t g; sfmt_init(&g, seed);
    for (int i = 0; i < n; i++) out[i] = sfmt_stream(&g);
}

/* src/dqmc_util.f90:70-72: abs(0-(irank*1981+2008)*951049) in default (32-bit) integer;
 * the product wraps in two's complement before abs */
int32_t orc_chain_seed(int32_t chain)
{
    uint32_t a = (uint32_t)chain * 1981u + 2008u;
    uint32_t p = a * 951049u;
    int32_t v = (int32_t)p;
    if (v < 0) v = (int32_t)(0u - (uint32_t)v);
    return v;
}

/* ------------------------------------------------------------------------------------ */
/* dense complex helpers                                                                 */
/* ------------------------------------------------------------------------------------ */
static zc *zalloc(size_t n) { return (zc *)calloc(n ? n : 1, sizeof(zc)); }

/* C(m x n) = op(A) * op(B) (+ beta*C), op: 'N', 'T', 'C'; A, B, C column-major */
static void zgemm(char ta, char tb, int m, int n, int k, const zc *A, int lda, const zc *B, int ldb,
                  zc beta, zc *C, int ldc)
{
    zc *Ae = NULL, *Be = NULL;
    const zc *Ap = A, *Bp = B;
    int la = lda, lb = ldb;
    if (ta != 'N') {
        Ae = zalloc((size_t)m * k);
        for (int j = 0; j < k; j++)
            for (int i = 0; i < m; i++) {
                zc v = A[IDX(j, i, lda)];
                Ae[IDX(i, j, m)] = (ta == 'C') ? conj(v) : v;
            }
        Ap = Ae; la = m;
    }
    if (tb != 'N') {
        Be = zalloc((size_t)k * n);
        for (int j = 0; j < n; j++)
            for (int i = 0; i < k; i++) {
                zc v = B[IDX(j, i, ldb)];
                Be[IDX(i, j, k)] = (tb == 'C') ? conj(v) : v;
            }
        Bp = Be; lb = k;
    }
    for (int j = 0; j < n; j++) {
        zc *c = C + (size_t)j * ldc;
        if (beta == 0) for (int i = 0; i < m; i++) c[i] = 0;
        else if (beta != 1) for (int i = 0; i < m; i++) c[i] *= beta;
        for (int l = 0; l < k; l++) {
            zc bl = Bp[IDX(l, j, lb)];
            if (bl == 0) continue;
            const zc *a = Ap + (size_t)l * la;
            for (int i = 0; i < m; i++) c[i] += a[i] * bl;
        }
    }
    free(Ae);
    free(Be);
}

/* BLAS dznrm2 (scaled sum of squares) */
static double dznrm2(int n, const zc *x, int incx)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        double v[2] = {creal(x[(size_t)i * incx]), cimag(x[(size_t)i * incx])};
        for (int t = 0; t < 2; t++) {
            if (v[t] != 0.0) {
                double a = fabs(v[t]);
                if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
                else ssq += (a / scale) * (a / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

static double sgn(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }

/* LAPACK zlarfg: alpha <- beta (real), x <- v(2:n), returns tau; H = I - tau v v^H,
 * H^H [alpha; x] = [beta; 0] */
static zc zlarfg(int n, zc *alpha, zc *x)
{
    if (n <= 0) return 0;
    double xnorm = dznrm2(n - 1, x, 1);
    double ar = creal(*alpha), ai = cimag(*alpha);
    if (xnorm == 0.0 && ai == 0.0) return 0;
    double beta = -sgn(hypot(hypot(ar, ai), xnorm), ar);
    zc tau = CMPLX((beta - ar) / beta, -ai / beta);
    zc scal = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; i++) x[i] *= scal;
    *alpha = beta;
    return tau;
}

/* apply H^H = I - conj(tau) v v^H from the left to A(m x n) (zlarf 'Left' with conj(tau)) */
static void apply_hh_left(int m, int n, const zc *v, zc tauc, zc *A, int lda)
{
    if (tauc == 0) return;
    for (int j = 0; j < n; j++) {
        zc *a = A + (size_t)j * lda;
        zc w = 0;
        for (int i = 0; i < m; i++) w += conj(v[i]) * a[i];
        w *= tauc;
        for (int i = 0; i < m; i++) a[i] -= v[i] * w;
    }
}

/*
 * Householder QR of A (m x n, m >= n), optionally with column pivoting.
 *   pivot = 1: LAPACK zgeqp3 / zlaqp2 (max partial column norm, downdated with the
 *              tol3z = sqrt(eps) recompute rule); as called by s_zgeQRPT_logdetQ
 *              (lib/s_matrix.f90:1222-1341).
 *   pivot = 0: zgeqrf as called by s_zgeQR_logdetQ (lib/s_matrix.f90:1998-2109).
 * Returns the reflectors in A, tau[n], jpvt (1-based).  logdetQ = sum_i log(1 - tau_i 2Re(tau_i)/|tau_i|^2)
 * for tau_i != 0 (s_matrix.f90:1279-1289).
 */
static void zgeqr_core(int m, int n, zc *A, int lda, zc *tau, int32_t *jpvt, int pivot)
{
    double *vn1 = NULL, *vn2 = NULL;
    if (jpvt) for (int j = 0; j < n; j++) jpvt[j] = j + 1;
    if (pivot) {
        vn1 = (double *)malloc(sizeof(double) * n);
        vn2 = (double *)malloc(sizeof(double) * n);
        for (int j = 0; j < n; j++) vn1[j] = vn2[j] = dznrm2(m, A + (size_t)j * lda, 1);
    }
    const double tol3z = sqrt(2.220446049250313e-16);
    int kmax = m < n ? m : n;
    zc *v = zalloc(m);
    for (int i = 0; i < kmax; i++) {
        if (pivot) {
            int p = i;
            double best = fabs(vn1[i]);
            for (int j = i + 1; j < n; j++)
                if (fabs(vn1[j]) > best) { best = fabs(vn1[j]); p = j; }
            if (p != i) {
                for (int r = 0; r < m; r++) {
                    zc t = A[IDX(r, p, lda)]; A[IDX(r, p, lda)] = A[IDX(r, i, lda)]; A[IDX(r, i, lda)] = t;
                }
                int32_t it = jpvt[p]; jpvt[p] = jpvt[i]; jpvt[i] = it;
                vn1[p] = vn1[i];
                vn2[p] = vn2[i];
            }
        }
        zc *col = A + IDX(i, i, lda);
        tau[i] = zlarfg(m - i, col, col + 1);
        if (i < n - 1) {
            v[0] = 1;
            for (int r = 1; r < m - i; r++) v[r] = col[r];
            apply_hh_left(m - i, n - i - 1, v, conj(tau[i]), A + IDX(i, i + 1, lda), lda);
        }
        if (pivot) {
            for (int j = i + 1; j < n; j++) {
                if (vn1[j] != 0.0) {
                    double t = cabs(A[IDX(i, j, lda)]) / vn1[j];
                    t = 1.0 - t * t;
                    if (t < 0) t = 0;
                    double t2 = t * (vn1[j] / vn2[j]) * (vn1[j] / vn2[j]);
                    if (t2 <= tol3z) {
                        if (i < m - 1) {
                            vn1[j] = dznrm2(m - i - 1, A + IDX(i + 1, j, lda), 1);
                            vn2[j] = vn1[j];
                        } else {
                            vn1[j] = 0.0;
                            vn2[j] = 0.0;
                        }
                    } else {
                        vn1[j] *= sqrt(t);
                    }
                }
            }
        }
    }
    free(v);
    free(vn1);
    free(vn2);
}

static zc logdet_q(int k, const zc *tau)
{
    zc s = 0;
    for (int i = 0; i < k; i++)
        if (cabs(tau[i]) != 0.0)
            s += clog(1.0 - tau[i] * (2.0 * creal(tau[i])) / (tau[i] * conj(tau[i])));
    return s;
}

/* zung2r: Q = H_1 ... H_k formed in place from the reflectors in A (m x k) */
static void zungqr(int m, int k, zc *A, int lda, const zc *tau)
{
    zc *v = zalloc(m);
    for (int i = k - 1; i >= 0; i--) {
        /* apply H_i to A(i:m, i+1:k) from the left */
        if (i < k - 1) {
            v[0] = 1;
            for (int r = 1; r < m - i; r++) v[r] = A[IDX(i + r, i, lda)];
            /* H_i = I - tau v v^H : zlarf with tau (not conj) */
            for (int j = i + 1; j < k; j++) {
                zc *a = A + IDX(i, j, lda);
                zc w = 0;
                for (int r = 0; r < m - i; r++) w += conj(v[r]) * a[r];
                w *= tau[i];
                for (int r = 0; r < m - i; r++) a[r] -= v[r] * w;
            }
        }
        if (i < m - 1)
            for (int r = i + 1; r < m; r++) A[IDX(r, i, lda)] *= -tau[i];
        A[IDX(i, i, lda)] = 1.0 - tau[i];
        for (int r = 0; r < i; r++) A[IDX(r, i, lda)] = 0;
    }
    free(v);
}

/* QR of a square or tall matrix: Q (m x kmax), R (kmax x n) upper triangular */
static void qr_full(int m, int n, const zc *a, zc *q, zc *r, int32_t *jpvt, zc *ldq, int pivot)
{
    int k = m < n ? m : n;
    zc *w = zalloc((size_t)m * n), *tau = zalloc(n);
    memcpy(w, a, sizeof(zc) * (size_t)m * n);
    zgeqr_core(m, n, w, m, tau, jpvt, pivot);
    *ldq = logdet_q(k, tau);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < k; i++) r[IDX(i, j, k)] = (i <= j) ? w[IDX(i, j, m)] : 0;
    for (int j = 0; j < k; j++)
        for (int i = 0; i < m; i++) q[IDX(i, j, m)] = (i > j) ? w[IDX(i, j, m)] : (i == j ? 1 : 0);
    zungqr(m, k, q, m, tau);
    free(w);
    free(tau);
}

void orc_qr(int m, int n, const double *a, double *q, double *r, int32_t *jpvt, double *logdetq, int pivot)
{
    zc ldq;
    qr_full(m, n, (const zc *)a, (zc *)q, (zc *)r, jpvt, &ldq, pivot);
    logdetq[0] = creal(ldq);
    logdetq[1] = cimag(ldq);
}

/*
 * s_inv_logdet_lu_z (lib/s_matrix.f90:2689-2820): zgetrf with partial pivoting (pivot =
 * izamax, i.e. the largest |Re|+|Im|), log det = sum log(+-U_ii) with the imaginary part
 * folded by nint(Im/2pi), then A^-1.  The inverse is formed by triangular solves, which is
 * the same matrix as U^-1 L^-1 P^T up to rounding.
 */
static int zinv_logdet(int n, zc *a, zc *logdet)
{
    int32_t *ipiv = (int32_t *)malloc(sizeof(int32_t) * n);
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = -1;
        for (int i = k; i < n; i++) {
            double v = fabs(creal(a[IDX(i, k, n)])) + fabs(cimag(a[IDX(i, k, n)]));
            if (v > best) { best = v; p = i; }
        }
        ipiv[k] = p;
        if (a[IDX(p, k, n)] == 0) { if (!info) info = k + 1; continue; }
        if (p != k)
            for (int j = 0; j < n; j++) {
                zc t = a[IDX(p, j, n)]; a[IDX(p, j, n)] = a[IDX(k, j, n)]; a[IDX(k, j, n)] = t;
            }
        zc piv = 1.0 / a[IDX(k, k, n)];
        for (int i = k + 1; i < n; i++) a[IDX(i, k, n)] *= piv;
        for (int j = k + 1; j < n; j++) {
            zc akj = a[IDX(k, j, n)];
            if (akj == 0) continue;
            zc *cj = a + (size_t)j * n;
            const zc *ck = a + (size_t)k * n;
            for (int i = k + 1; i < n; i++) cj[i] -= ck[i] * akj;
        }
    }
    zc ld = 0;
    for (int i = 0; i < n; i++) ld += (ipiv[i] == i) ? clog(a[IDX(i, i, n)]) : clog(-a[IDX(i, i, n)]);
    double im = cimag(ld);
    double k2 = round(im / 2.0 / M_PI);
    ld = CMPLX(creal(ld), im - 2.0 * M_PI * k2);
    *logdet = ld;
    if (info) { free(ipiv); return info; }
    /* inverse: solve A X = I with P A = L U */
    zc *x = zalloc((size_t)n * n);
    for (int i = 0; i < n; i++) x[IDX(i, i, n)] = 1;
    for (int k = 0; k < n; k++)
        if (ipiv[k] != k)
            for (int j = 0; j < n; j++) {
                zc t = x[IDX(ipiv[k], j, n)]; x[IDX(ipiv[k], j, n)] = x[IDX(k, j, n)]; x[IDX(k, j, n)] = t;
            }
    for (int j = 0; j < n; j++) {
        zc *xj = x + (size_t)j * n;
        for (int k = 0; k < n; k++) {          /* L y = b, unit lower */
            zc v = xj[k];
            if (v == 0) continue;
            const zc *ck = a + (size_t)k * n;
            for (int i = k + 1; i < n; i++) xj[i] -= ck[i] * v;
        }
        for (int k = n - 1; k >= 0; k--) {     /* U x = y */
            xj[k] /= a[IDX(k, k, n)];
            zc v = xj[k];
            const zc *ck = a + (size_t)k * n;
            for (int i = 0; i < k; i++) xj[i] -= ck[i] * v;
        }
    }
    memcpy(a, x, sizeof(zc) * (size_t)n * n);
    free(x);
    free(ipiv);
    return 0;
}

int orc_inv_logdet(int n, double *a, double *logdet)
{
    zc ld;
    int info = zinv_logdet(n, (zc *)a, &ld);
    logdet[0] = creal(ld);
    logdet[1] = cimag(ld);
    return info;
}

/* ------------------------------------------------------------------------------------ */
/* context                                                                               */
/* ------------------------------------------------------------------------------------ */
struct orc_ctx {
    orc_params p;
    int ndim, lq, nsub, ltrot, nst, nwrap, nublock;
    int nn_nf, nn_lf;
    int *nnlf;          /* [nn_lf] A sites, 0-based */
    int *nnlist;        /* [nsites*nn_nf] */
    int *cx, *cy, *cz, *sub; /* unit-cell coordinates and sublattice of each site */
    int *iwrap_nt;      /* [ltrot+1] */
    int *wrap_step;     /* [2*nst], 1-based slice numbers */
    /* kinetic */
    int lwrapT, lwrapu;
    zc *urt, *urtm1;    /* checkerboard: [ist*4 + (i-1) + 2*(j-1)]; dense: N*N */
    /* HS */
    int lcomp;
    double alpha, etal[4], gaml[4];
    zc bmat[4], bmat_inv[4], phase_tab[4], delta_bmat[3 * 4], phase_ratio[3 * 4];
    int32_t *conf;      /* [ltrot][nsites] */
    sfmt_t rng;
    /* sweep state */
    zc *G, *Gtmp;
    zc **Ust, **Vst;
    double **Dst;
    zc *logdetQst;
    zc phase, logweightf;
    double max_wrap_error, max_phasediff, nchecks, sumlog10err, accepted, tried;
    double nobs;
    zc obs[5];
    double *trace;
    int trace_cap, trace_n;
    /* learned block proposals (no counterpart in the reference; see orc_set_proposal) */
    int prop_on, prop_local, prop_nprop, prop_z;
    int32_t *prop_nbr;
    float *prop_w;
    double prop_norm[10];
    uint64_t prop_rng;
    double prop_stats[5];
    /* time-displaced measurement (ltau; see orc_set_dyn) */
    int ltau, ntdm;
    zc *gt0, *g0t, *g00, *znn;
    double xmax_dyn, dyncount, sumlog10dyn;
    double *dtrace;
    int dtrace_cap, dtrace_n;
    /* the correlation bins and extra energy terms of equaltime_measure (see orc_set_eqt) */
    int eqt, zplq;
    double theta;
    zc *eqb;
    int *ca, *cnn, *csn, *cjj;  /* per unit cell: A site, its nnlist(.,1), snlist(.,1), current partner */
    int *plq;                   /* [lq][zplq] plaquette site cycles */
    /* the bond field of the t-V model (vconf with lprojv = .false.) */
    int lwrapv;
    zc vbp[2], vbm[2], vbpi[2], vbmi[2], vdp[2], vdm[2];
    int32_t *confv;             /* [ltrot][nn_nf][ndim] */
    double vacc[2];
    /* the plaquette field (hconf with lprojplqu = .false.) */
    int lwrapp, zp, lqp;
    int *pcord;                 /* [lqp][zp] */
    zc pb[4][36], pbi[4][36], pdl[4][3][36], ppr[4][3], pph[4];
    double pgaml[4];
    int32_t *confp;             /* [ltrot][lqp] */
    double pacc[2];
};

/* src/lattice.f90:792-961 (square), :611-776 (honeycomb), :240-316 (chain), :1278-1400 (cubic);
 * 0-based site indices */
static void setup_lattice(orc_ctx *c)
{
    int l1 = c->p.la, l2 = c->p.lb;
#define NPBC(v, l) ((v) > (l) ? (v) - (l) : ((v) < 1 ? (v) + (l) : (v)))
    if (c->p.lattice == 2) {
        c->nsub = 1; c->lq = l1; c->ndim = l1;
        c->nn_nf = 2; c->nn_lf = l1 / 2;
        c->nnlf = (int *)calloc(c->nn_lf > 0 ? c->nn_lf : 1, sizeof(int));
        c->nnlist = (int *)calloc((size_t)c->ndim * 2, sizeof(int));
        c->cx = (int *)calloc(c->ndim, sizeof(int));
        c->cy = (int *)calloc(c->ndim, sizeof(int));
        c->sub = (int *)calloc(c->ndim, sizeof(int));
        int ind = 0;
        for (int i1 = 1; i1 <= l1; i1++) {
            c->cx[i1 - 1] = i1; c->cy[i1 - 1] = 1;
            if (i1 % 2 == 0 && ind < c->nn_lf) c->nnlf[ind++] = i1 - 1;
        }
        for (int s = 0; s < c->ndim; s++) {
            c->nnlist[s * 2 + 0] = NPBC(c->cx[s] + 1, l1) - 1;
            c->nnlist[s * 2 + 1] = NPBC(c->cx[s] - 1, l1) - 1;
        }
    } else if (c->p.lattice == 3) {
        int l3 = c->p.lc;
        c->nsub = 1; c->lq = l1 * l2 * l3; c->ndim = c->lq;
        c->nn_nf = 6; c->nn_lf = c->ndim / 2;
        c->nnlf = (int *)calloc(c->nn_lf > 0 ? c->nn_lf : 1, sizeof(int));
        c->nnlist = (int *)calloc((size_t)c->ndim * 6, sizeof(int));
        c->cx = (int *)calloc(c->ndim, sizeof(int));
        c->cy = (int *)calloc(c->ndim, sizeof(int));
        c->cz = (int *)calloc(c->ndim, sizeof(int));
        c->sub = (int *)calloc(c->ndim, sizeof(int));
#define CSITE(x, y, z) (((z) - 1) * l1 * l2 + ((x) - 1) * l2 + (y) - 1)
        int ind = 0;
        for (int i1 = 1; i1 <= l1; i1++)
            for (int i2 = 1; i2 <= l2; i2++)
                for (int i3 = 1; i3 <= l3; i3++) {
                    int s = CSITE(i1, i2, i3);
                    c->cx[s] = i1; c->cy[s] = i2; c->cz[s] = i3;
                    if ((i1 + i2 + i3) % 2 == 0 && ind < c->nn_lf) c->nnlf[ind++] = s;
                }
        for (int s = 0; s < c->ndim; s++) {
            int x = c->cx[s], y = c->cy[s], z = c->cz[s];
            int *nl = c->nnlist + s * 6;
            nl[0] = CSITE(x, y, NPBC(z + 1, l3));
            nl[1] = CSITE(NPBC(x + 1, l1), y, z);
            nl[2] = CSITE(x, NPBC(y + 1, l2), z);
            nl[3] = CSITE(x, y, NPBC(z - 1, l3));
            nl[4] = CSITE(NPBC(x - 1, l1), y, z);
            nl[5] = CSITE(x, NPBC(y - 1, l2), z);
        }
#undef CSITE
    } else if (c->p.lattice == 0) {
        c->nsub = 1; c->lq = l1 * l2; c->ndim = c->lq;
        c->nn_nf = 4; c->nn_lf = l1 * l2 / 2;
        c->nnlf = (int *)calloc(c->nn_lf > 0 ? c->nn_lf : 1, sizeof(int));
        c->nnlist = (int *)calloc((size_t)c->ndim * 4, sizeof(int));
        c->cx = (int *)calloc(c->ndim, sizeof(int));
        c->cy = (int *)calloc(c->ndim, sizeof(int));
        c->sub = (int *)calloc(c->ndim, sizeof(int));
        int ind = 0;
        for (int i1 = 1; i1 <= l1; i1++)
            for (int i2 = 1; i2 <= l2; i2++) {
                int s = i2 + (i1 - 1) * l2 - 1;
                c->cx[s] = i1; c->cy[s] = i2; c->sub[s] = 0;
                if ((i1 + i2) % 2 == 0 && ind < c->nn_lf) c->nnlf[ind++] = s;
            }
        for (int s = 0; s < c->ndim; s++) {
            int x = c->cx[s], y = c->cy[s];
            c->nnlist[s * 4 + 0] = (NPBC(x + 1, l1) - 1) * l2 + NPBC(y, l2) - 1;
            c->nnlist[s * 4 + 1] = (NPBC(x, l1) - 1) * l2 + NPBC(y + 1, l2) - 1;
            c->nnlist[s * 4 + 2] = (NPBC(x - 1, l1) - 1) * l2 + NPBC(y, l2) - 1;
            c->nnlist[s * 4 + 3] = (NPBC(x, l1) - 1) * l2 + NPBC(y - 1, l2) - 1;
        }
    } else {
        c->nsub = 2; c->lq = l1 * l2; c->ndim = 2 * c->lq;
        c->nn_nf = 3; c->nn_lf = l1 * l2;
        c->nnlf = (int *)calloc(c->nn_lf, sizeof(int));
        c->nnlist = (int *)calloc((size_t)c->ndim * 3, sizeof(int));
        c->cx = (int *)calloc(c->ndim, sizeof(int));
        c->cy = (int *)calloc(c->ndim, sizeof(int));
        c->sub = (int *)calloc(c->ndim, sizeof(int));
        for (int i = 0; i < c->nn_lf; i++) c->nnlf[i] = 2 * i;
        int ind = 0;
        for (int i1 = 1; i1 <= l1; i1++)
            for (int i2 = 1; i2 <= l2; i2++)
                for (int o = 0; o < 2; o++) { c->cx[ind] = i1; c->cy[ind] = i2; c->sub[ind] = o; ind++; }
        for (int s = 0; s < c->ndim; s++) {
            int x = c->cx[s], y = c->cy[s];
            if (c->sub[s] == 0) {
                c->nnlist[s * 3 + 0] = s + 1;
                c->nnlist[s * 3 + 1] = ((NPBC(x + 1, l1) - 1) * l2 + NPBC(y, l2)) * 2 - 1;
                c->nnlist[s * 3 + 2] = ((NPBC(x, l1) - 1) * l2 + NPBC(y - 1, l2)) * 2 - 1;
            } else {
                c->nnlist[s * 3 + 0] = s - 1;
                c->nnlist[s * 3 + 1] = ((NPBC(x - 1, l1) - 1) * l2 + NPBC(y, l2)) * 2 - 2;
                c->nnlist[s * 3 + 2] = ((NPBC(x, l1) - 1) * l2 + NPBC(y + 1, l2)) * 2 - 2;
            }
        }
    }
    if (!c->cz) {
        c->cz = (int *)calloc(c->ndim, sizeof(int));
        for (int s = 0; s < c->ndim; s++) c->cz[s] = 1;
    }
#undef NPBC
}

/* src/model_para.f90:421-453: nst, iwrap_nt(0:ltrot), wrap_step(2,nst) */
static void setup_schedule(orc_ctx *c)
{
    int lt = c->ltrot, nw = c->p.nwrap;
    c->iwrap_nt = (int *)calloc(lt + 1, sizeof(int));
    if (lt < nw) nw = lt;
    c->nwrap = nw;
    if (lt % nw == 0) {
        c->nst = lt / nw;
        c->wrap_step = (int *)calloc(2 * c->nst, sizeof(int));
        for (int i = 1; i <= c->nst; i++) {
            c->iwrap_nt[i * nw] = i;
            c->wrap_step[2 * (i - 1)] = 1 + (i - 1) * nw;
            c->wrap_step[2 * (i - 1) + 1] = i * nw;
        }
    } else {
        c->nst = lt / nw + 1;
        c->wrap_step = (int *)calloc(2 * c->nst, sizeof(int));
        for (int i = 1; i <= c->nst - 1; i++) {
            c->iwrap_nt[i * nw] = i;
            c->wrap_step[2 * (i - 1)] = 1 + (i - 1) * nw;
            c->wrap_step[2 * (i - 1) + 1] = i * nw;
        }
        int i = c->nst;
        c->iwrap_nt[lt] = i;
        c->wrap_step[2 * (i - 1)] = (i - 1) * nw + 1;
        c->wrap_step[2 * (i - 1) + 1] = lt;
    }
}

/* src/dqmc_config_u/dqmc_set_conf.f90:1-176 (repulsive and attractive U, no projection) */
static void setup_hs(orc_ctx *c)
{
    double u = c->p.rhub, dtau = c->p.dtau, nu = c->p.nu;
    int nflr = c->p.nflr;
    int exact = c->p.exact_hs;
    int general = 1;
    if (exact && nflr == 2) {
        c->lcomp = 2;
        c->etal[0] = -1.0; c->etal[1] = 1.0;
        c->alpha = (u >= 0) ? acos(exp(-dtau * fabs(u) * 0.5)) : acosh(exp(dtau * fabs(u) * 0.5));
        c->gaml[0] = c->gaml[1] = 1.0;
        general = 0;
    } else if (exact && (nflr == 4 || nflr == 6)) {
        c->lcomp = 4;
        double a = exp(-0.5 * dtau * u);
        double d = sqrt(8.0 + a * a * (3.0 + a * a) * (3.0 + a * a));
        double xp = (a + 2.0 * a * a * a + pow(a, 5) + (a * a - 1.0) * d) / 4.0;
        double xm = (a + 2.0 * a * a * a + pow(a, 5) - (a * a - 1.0) * d) / 4.0;
        if (u >= 0) {
            c->etal[0] = -acos(xm); c->etal[1] = -acos(xp); c->etal[2] = acos(xp); c->etal[3] = acos(xm);
        } else {
            c->etal[0] = -acosh(xm); c->etal[1] = -acosh(xp); c->etal[2] = acosh(xp); c->etal[3] = acosh(xm);
        }
        c->alpha = 1.0;
        double t = a * (3.0 + a * a);
        c->gaml[0] = (t + d) / d; c->gaml[1] = (-t + d) / d; c->gaml[2] = (-t + d) / d; c->gaml[3] = (t + d) / d;
        general = 0;
    }
    if (general) {
        c->lcomp = 4;
        double s6 = sqrt(6.0);
        c->etal[0] = -sqrt(2.0 * (3.0 + s6));
        c->etal[1] = -sqrt(2.0 * (3.0 - s6));
        c->etal[2] = sqrt(2.0 * (3.0 - s6));
        c->etal[3] = sqrt(2.0 * (3.0 + s6));
        c->alpha = sqrt(dtau * fabs(u) * 0.5);
        c->gaml[0] = 1.0 - s6 / 3.0; c->gaml[1] = 1.0 + s6 / 3.0;
        c->gaml[2] = 1.0 + s6 / 3.0; c->gaml[3] = 1.0 - s6 / 3.0;
    }
    int L = c->lcomp;
    double al = c->alpha;
    for (int is = 0; is < L; is++) {
        if (u >= 0) {
            c->phase_tab[is] = cexp(CMPLX(0.0, al * (-c->etal[is]) * 0.5 * ((double)nflr + nu)));
            c->bmat[is] = cexp(CMPLX(0.0, al * c->etal[is]));
            c->bmat_inv[is] = cexp(CMPLX(0.0, -al * c->etal[is]));
        } else {
            c->phase_tab[is] = cexp(CMPLX(al * (-c->etal[is]) * 0.5 * ((double)nflr + nu), 0.0));
            c->bmat[is] = cexp(CMPLX(al * c->etal[is], 0.0));
            c->bmat_inv[is] = cexp(CMPLX(-al * c->etal[is], 0.0));
        }
    }
    for (int is = 0; is < L; is++)
        for (int fl = 1; fl <= L - 1; fl++) {
            int isp = (is + fl) % L; /* mod(is+iflip-1,lcomp)+1 with 1-based is */
            int k = (fl - 1) + (L - 1) * is;
            if (u >= 0) {
                c->phase_ratio[k] = cexp(CMPLX(0.0, al * (c->etal[is] - c->etal[isp]) * 0.5 * ((double)nflr + nu)));
                c->delta_bmat[k] = cexp(CMPLX(0.0, al * (c->etal[isp] - c->etal[is]))) - 1.0;
            } else {
                c->phase_ratio[k] = cexp(CMPLX(al * (c->etal[is] - c->etal[isp]) * 0.5 * ((double)nflr + nu), 0.0));
                c->delta_bmat[k] = cexp(CMPLX(al * (c->etal[isp] - c->etal[is]), 0.0)) - 1.0;
            }
        }
}

/* exp(-s H) for H = [[d, z],[conj z, d]] (2x2 bond block of dqmc_set_h0conf.f90:77-123) */
static void bond_exp(zc z, double d, double s, zc out[4])
{
    double r = cabs(z);
    double ch = cosh(s * r), sh_r = (r > 0) ? sinh(s * r) / r : s;
    double e = exp(-s * d);
    out[0] = e * ch;                /* (1,1) */
    out[1] = -e * sh_r * conj(z);   /* (2,1) */
    out[2] = -e * sh_r * z;         /* (1,2) */
    out[3] = e * ch;                /* (2,2) */
}

/* src/dqmc_config_h0/seth0.f90:1-28 with expar = 1 (no field, flux or dimerisation) */
static void build_h0(const orc_ctx *c, zc *H)
{
    int N = c->ndim;
    memset(H, 0, sizeof(zc) * (size_t)N * N);
    for (int b = 0; b < c->nn_lf; b++) {
        int i0 = c->nnlf[b];
        for (int nf = 0; nf < c->nn_nf; nf++) {
            int in = c->nnlist[i0 * c->nn_nf + nf];
            H[IDX(i0, in, N)] += -c->p.rt;
            H[IDX(in, i0, N)] += -c->p.rt;
        }
    }
    for (int i = 0; i < N; i++) H[IDX(i, i, N)] -= c->p.mu;
}

/* exp(s H) by a Taylor series with scaling and squaring (same matrix as the reference's
 * s_eig_he route in dqmc_set_h0conf.f90:484-510, up to rounding) */
static void expm_herm(int N, const zc *H, double s, zc *out)
{
    double nrm = 0;
    for (int j = 0; j < N; j++) {
        double cs = 0;
        for (int i = 0; i < N; i++) cs += cabs(H[IDX(i, j, N)]);
        if (cs > nrm) nrm = cs;
    }
    int sq = 0;
    double sc = fabs(s) * nrm;
    while (sc > 0.25) { sc *= 0.5; sq++; }
    double f = s / (double)(1 << sq);
    zc *A = zalloc((size_t)N * N), *T = zalloc((size_t)N * N), *T2 = zalloc((size_t)N * N);
    for (size_t k = 0; k < (size_t)N * N; k++) A[k] = f * H[k];
    memset(out, 0, sizeof(zc) * (size_t)N * N);
    for (int i = 0; i < N; i++) { out[IDX(i, i, N)] = 1; T[IDX(i, i, N)] = 1; }
    for (int term = 1; term <= 30; term++) {
        zgemm('N', 'N', N, N, N, A, N, T, N, 0, T2, N);
        double mx = 0;
        for (size_t k = 0; k < (size_t)N * N; k++) { T2[k] /= (double)term; if (cabs(T2[k]) > mx) mx = cabs(T2[k]); out[k] += T2[k]; }
        zc *sw = T; T = T2; T2 = sw;
        if (mx < 1e-18) break;
    }
    for (int q = 0; q < sq; q++) {
        zgemm('N', 'N', N, N, N, out, N, out, N, 0, T2, N);
        memcpy(out, T2, sizeof(zc) * (size_t)N * N);
    }
    free(A); free(T); free(T2);
}

static void setup_kinetic(orc_ctx *c)
{
    int N = c->ndim;
    c->lwrapT = (c->p.rt != 0.0);
    if (c->p.kmode == 0) {
        int nb = c->nn_nf * c->nn_lf;
        c->urt = zalloc((size_t)nb * 4);
        c->urtm1 = zalloc((size_t)nb * 4);
        for (int nf = 0; nf < c->nn_nf; nf++)
            for (int b = 0; b < c->nn_lf; b++) {
                int ist = b + nf * c->nn_lf;
                zc z = -c->p.rt;
                double d = -c->p.mu / (double)c->nn_nf;
                bond_exp(z, d, 0.5 * c->p.dtau, c->urt + 4 * ist);
                bond_exp(z, d, -0.5 * c->p.dtau, c->urtm1 + 4 * ist);
            }
    } else {
        zc *H = zalloc((size_t)N * N);
        build_h0(c, H);
        c->urt = zalloc((size_t)N * N);
        c->urtm1 = zalloc((size_t)N * N);
        expm_herm(N, H, -0.5 * c->p.dtau, c->urt);
        expm_herm(N, H, 0.5 * c->p.dtau, c->urtm1);
        free(H);
    }
}

static void setup_vfield(orc_ctx *c);
static void setup_pfield(orc_ctx *c);

orc_ctx *orc_create(const orc_params *p)
{
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->p = *p;
    setup_lattice(c);
    c->ltrot = (int)lround(p->beta / p->dtau);
    setup_schedule(c);
    c->nublock = p->nublock > 0 ? p->nublock : 1;
    c->lwrapu = (p->rhub != 0.0);
    setup_hs(c);
    setup_kinetic(c);
    sfmt_init(&c->rng, p->seed);
    c->prop_local = 1;
    int N = c->ndim;
    c->conf = (int32_t *)calloc((size_t)N * c->ltrot, sizeof(int32_t));
    c->lwrapv = (p->rv != 0.0);
    if (c->lwrapv) {
        setup_vfield(c);
        c->confv = (int32_t *)malloc(sizeof(int32_t) * (size_t)N * c->nn_nf * c->ltrot);
        for (size_t k = 0; k < (size_t)N * c->nn_nf * c->ltrot; k++) c->confv[k] = 1;
    }
    for (size_t k = 0; k < (size_t)N * c->ltrot; k++) c->conf[k] = 1;
    c->lwrapp = (p->rhub_plq != 0.0);
    if (c->lwrapp) {
        setup_pfield(c);
        c->confp = (int32_t *)malloc(sizeof(int32_t) * (size_t)c->lqp * c->ltrot);
        for (size_t k = 0; k < (size_t)c->lqp * c->ltrot; k++) c->confp[k] = 1;
    }
    c->G = zalloc((size_t)N * N);
    c->Gtmp = zalloc((size_t)N * N);
    c->Ust = (zc **)calloc(c->nst + 1, sizeof(zc *));
    c->Vst = (zc **)calloc(c->nst + 1, sizeof(zc *));
    c->Dst = (double **)calloc(c->nst + 1, sizeof(double *));
    c->logdetQst = zalloc(c->nst + 1);
    for (int n = 0; n <= c->nst; n++) {
        c->Ust[n] = zalloc((size_t)N * N);
        c->Vst[n] = zalloc((size_t)N * N);
        c->Dst[n] = (double *)calloc(N, sizeof(double));
    }
    c->phase = 1.0;
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    for (int n = 0; n <= c->nst; n++) { free(c->Ust[n]); free(c->Vst[n]); free(c->Dst[n]); }
    free(c->Ust); free(c->Vst); free(c->Dst); free(c->logdetQst);
    free(c->G); free(c->Gtmp); free(c->conf);
    free(c->urt); free(c->urtm1);
    free(c->nnlf); free(c->nnlist); free(c->cx); free(c->cy); free(c->cz); free(c->sub);
    free(c->eqb); free(c->ca); free(c->cnn); free(c->csn); free(c->cjj); free(c->plq);
    free(c->confv); free(c->confp); free(c->pcord);
    free(c->iwrap_nt); free(c->wrap_step);
    free(c->prop_nbr); free(c->prop_w);
    free(c->gt0); free(c->g0t); free(c->g00); free(c->znn);
    free(c);
}

void orc_sizes(const orc_ctx *c, int *out)
{
    out[0] = c->ndim; out[1] = c->lq; out[2] = c->ltrot; out[3] = c->nst;
    out[4] = c->lcomp; out[5] = c->nn_nf; out[6] = c->nn_lf; out[7] = c->nublock;
}

/* dqmc_input_conf.f90:55-63 (single rank): conf(i,nt) = ceiling(u*lcomp), nt outer, i inner */
void orc_init_conf(orc_ctx *c)
{
    /* hconf%input_plqconf first (src/dqmc_ft_ctrl.f90:19-23; dqmc_input_plqconf.f90:55-62):
     * conf_plqu(i, nt) = ceiling(u * 4), nt outer, plaquette inner */
    if (c->lwrapp)
        for (size_t k = 0; k < (size_t)c->lqp * c->ltrot; k++)
            c->confp[k] = (int32_t)ceil(sfmt_stream(&c->rng) * 4);
    if (c->lwrapu)
        for (int nt = 0; nt < c->ltrot; nt++)
            for (int i = 0; i < c->ndim; i++)
                c->conf[(size_t)nt * c->ndim + i] = (int32_t)ceil(sfmt_stream(&c->rng) * c->lcomp);
    /* v0conf%input_conf after u0conf%input_conf (src/dqmc_ft_ctrl.f90:25-35): conf(i, nf, nt),
     * i fastest, two components */
    if (c->lwrapv)
        for (size_t k = 0; k < (size_t)c->ndim * c->nn_nf * c->ltrot; k++)
            c->confv[k] = (int32_t)ceil(sfmt_stream(&c->rng) * 2);
}
void orc_get_confv(const orc_ctx *c, int32_t *conf)
{
    if (c->confv) memcpy(conf, c->confv, sizeof(int32_t) * (size_t)c->ndim * c->nn_nf * c->ltrot);
}
void orc_set_confv(orc_ctx *c, const int32_t *conf)
{
    if (c->confv) memcpy(c->confv, conf, sizeof(int32_t) * (size_t)c->ndim * c->nn_nf * c->ltrot);
}
void orc_get_vacc(const orc_ctx *c, double *acc) { acc[0] = c->vacc[0]; acc[1] = c->vacc[1]; }
void orc_get_confp(const orc_ctx *c, int32_t *conf)
{
    if (c->confp) memcpy(conf, c->confp, sizeof(int32_t) * (size_t)c->lqp * c->ltrot);
}
void orc_set_confp(orc_ctx *c, const int32_t *conf)
{
    if (c->confp) memcpy(c->confp, conf, sizeof(int32_t) * (size_t)c->lqp * c->ltrot);
}
void orc_get_pacc(const orc_ctx *c, double *acc) { acc[0] = c->pacc[0]; acc[1] = c->pacc[1]; }
int orc_lqp(const orc_ctx *c) { return c->lqp; }

/* ---- the bond field of the finite-temperature t-V model ------------------------------- */
/* vconf%set_conf with lprojv = .false. (src/dqmc_config_v/dqmc_set_conf.f90:14-77): alpha =
 * acosh(exp(dtau V / 2)), eta = (-1, +1), bmat_p = exp(alpha eta), bmat_m = exp(-alpha eta),
 * delta for the one flip iflip = 1 */
static void setup_vfield(orc_ctx *c)
{
    double al = acosh(exp(c->p.dtau * c->p.rv * 0.5));
    const double eta[2] = {-1.0, 1.0};
    for (int is = 0; is < 2; is++) {
        int isp = (is + 1) % 2;
        c->vbp[is] = exp(al * eta[is]);
        c->vbpi[is] = exp(-al * eta[is]);
        c->vbm[is] = exp(-al * eta[is]);
        c->vbmi[is] = exp(al * eta[is]);
        c->vdp[is] = exp(al * (eta[isp] - eta[is])) - 1.0;
        c->vdm[is] = exp(-al * (eta[isp] - eta[is])) - 1.0;
    }
}
#define VS 0.70710678118654752440   /* u = ut = [[1, 1], [1, -1]] / sqrt(2) (dqmc_config_v.f90:19-20) */
static int fv_is(const orc_ctx *c, int nt, int nf, int q) { return c->confv[((size_t)(nt - 1) * c->nn_nf + nf) * c->ndim + q] - 1; }
/* rows i1, i2 of every bond of family nf:
 *   mode 0: left_forward nflag 2 (dqmc_config_v/dqmc_left_forward_prop.f90:24-38): rows <- diag(bp, bm) ut rows
 *   mode 1: left_forward nflag 1 / left_backward nflag 2: rows <- u rows
 *   mode 2: left_backward nflag 1 (dqmc_left_backward_prop.f90:39-55): rows <- u diag(bpi, bmi) rows */
static void fv_left(orc_ctx *c, int nt, int nf, int mode, zc *M)
{
    int N = c->ndim;
    for (int q = 0; q < c->nn_lf; q++) {
        int i1 = c->nnlf[q], i2 = c->nnlist[i1 * c->nn_nf + nf], is = fv_is(c, nt, nf, q);
        for (int j = 0; j < N; j++) {
            zc x = M[IDX(i1, j, N)], y = M[IDX(i2, j, N)];
            if (mode == 2) { x *= c->vbpi[is]; y *= c->vbmi[is]; }
            zc v1 = VS * (x + y), v2 = VS * (x - y);
            if (mode == 0) { v1 *= c->vbp[is]; v2 *= c->vbm[is]; }
            M[IDX(i1, j, N)] = v1;
            M[IDX(i2, j, N)] = v2;
        }
    }
}
/* columns i1, i2:
 *   mode 0: right_backward nflag 2 (dqmc_right_backward_prop.f90:24-38): cols <- (cols u) diag(bpi, bmi)
 *   mode 1: right_backward nflag 1 / right_forward nflag 2: cols <- cols u
 *   mode 2: right_forward nflag 1 (dqmc_right_forward_prop.f90:40-56): cols <- (cols diag(bp, bm)) ut */
static void fv_right(orc_ctx *c, int nt, int nf, int mode, zc *M)
{
    int N = c->ndim;
    for (int q = 0; q < c->nn_lf; q++) {
        int i1 = c->nnlf[q], i2 = c->nnlist[i1 * c->nn_nf + nf], is = fv_is(c, nt, nf, q);
        for (int j = 0; j < N; j++) {
            zc x = M[IDX(j, i1, N)], y = M[IDX(j, i2, N)];
            if (mode == 2) { x *= c->vbp[is]; y *= c->vbm[is]; }
            zc v1 = VS * (x + y), v2 = VS * (x - y);
            if (mode == 0) { v1 *= c->vbpi[is]; v2 *= c->vbmi[is]; }
            M[IDX(j, i1, N)] = v1;
            M[IDX(j, i2, N)] = v2;
        }
    }
}
/* vconf%update (src/dqmc_config_v/dqmc_update.f90:35-96): per bond of family nf in order, a flip
 * with ratio det(1 + (1 - G)_bb diag(dp, dm)) (no flavour power, no phase factor), accepted on
 * |Re(ratio phase)| / |Re(phase)| > u; G <- G - G_{:,b} diag(dp, dm) M^-1 (E_b - G_{b,:}) */
static void fv_update(orc_ctx *c, int nt, int nf)
{
    int N = c->ndim;
    zc *G = c->G;
    for (int q = 0; q < c->nn_lf; q++) {
        int i1 = c->nnlf[q], i2 = c->nnlist[i1 * c->nn_nf + nf], is = fv_is(c, nt, nf, q);
        int iflip = (int)ceil(sfmt_stream(&c->rng) * 1.0);
        int isp = (is + iflip) % 2;
        zc dp = c->vdp[is], dm = c->vdm[is];
        zc m11 = 1.0 + (1.0 - G[IDX(i1, i1, N)]) * dp, m12 = -G[IDX(i1, i2, N)] * dm;
        zc m21 = -G[IDX(i2, i1, N)] * dp, m22 = 1.0 + (1.0 - G[IDX(i2, i2, N)]) * dm;
        zc ratio = m11 * m22 - m12 * m21;
        double rr = creal(ratio * c->phase) / creal(c->phase);
        double random = sfmt_stream(&c->rng);
        c->vacc[1] += 1.0;
        if (fabs(rr) > random) {
            c->vacc[0] += 1.0;
            c->phase *= ratio / cabs(ratio);
            /* S = diag(dp, dm) M^-1 */
            zc s11 = dp * m22 / ratio, s12 = -dp * m12 / ratio, s21 = -dm * m21 / ratio, s22 = dm * m11 / ratio;
            zc *r1 = zalloc(N), *r2 = zalloc(N), *c1 = zalloc(N), *c2 = zalloc(N);
            for (int j = 0; j < N; j++) {
                zc a = (j == i1 ? 1.0 : 0.0) - G[IDX(i1, j, N)], b = (j == i2 ? 1.0 : 0.0) - G[IDX(i2, j, N)];
                r1[j] = -(s11 * a + s12 * b);
                r2[j] = -(s21 * a + s22 * b);
                c1[j] = G[IDX(j, i1, N)];
                c2[j] = G[IDX(j, i2, N)];
            }
            for (int j = 0; j < N; j++)
                for (int i = 0; i < N; i++) G[IDX(i, j, N)] += c1[i] * r1[j] + c2[i] * r2[j];
            free(r1); free(r2); free(c1); free(c2);
            c->confv[((size_t)(nt - 1) * c->nn_nf + nf) * N + q] = isp + 1;
        }
    }
}
/* ---- the plaquette field (hconf, lprojplqu = .false.) ---------------------------------- */
static void expm_herm(int N, const zc *H, double s, zc *out);
/* latt%plq_cord (src/lattice.f90:898-958 square, :735-773 honeycomb), 0-based: the square's
 * plaquettes A (i1, i2 odd), B (i1 even, i2 odd), C (both even), D (i1 odd, i2 even), each class
 * in (i1, i2) order, anticlockwise from the lower-left corner; the honeycomb's A (mod(i1+i2, 3) =
 * 2), B (= 1), C (= 0), clockwise from orbital 2 of cell (i1, i2).  set_plqconf
 * (src/dqmc_config_plqu/dqmc_set_plqconf.f90:57-199): V = the plaquette ring with 1/nn_nf on the
 * diagonal and alpha e^{-+i theta}, a = sqrt(dtau rhub_plq), bmat(is) = exp(i a eta_is V) (the
 * reference's U exp(i a eta eig) U^H from zheev is the same matrix function), its inverse,
 * delta(iflip, is) = exp(i a (eta_isp - eta_is) V) - 1, phase(is) = exp(-i a eta_is f (nflr +
 * nu)), phase_ratio(iflip, is) = exp(i a (eta_is - eta_isp) f (nflr + nu)), f = 1/2 on the
 * square and 1 on the honeycomb; eta and gamma of dqmc_config_plqu.f90:42-51 */
static void setup_pfield(orc_ctx *c)
{
    int l1 = c->p.la, l2 = c->p.lb, z = (c->p.lattice == 0) ? 4 : 6;
    c->zp = z;
    c->lqp = l1 * l2;
    c->pcord = (int *)malloc(sizeof(int) * (size_t)c->lqp * z);
#define PWR(v, l) ((v) > (l) ? (v) - (l) : ((v) < 1 ? (v) + (l) : (v)))
#define PSQ(x, y) ((PWR(x, l1) - 1) * l2 + PWR(y, l2) - 1)
#define PHC(x, y, o) (2 * ((PWR(x, l1) - 1) * l2 + PWR(y, l2) - 1) + (o) - 1)
    int id = 0;
    for (int cls = 0; cls < (z == 4 ? 4 : 3); cls++)
        for (int i1 = 1; i1 <= l1; i1++)
            for (int i2 = 1; i2 <= l2; i2++) {
                int k;
                if (z == 4) k = (i1 % 2 == 1 && i2 % 2 == 1) ? 0 : (i1 % 2 == 0 && i2 % 2 == 1) ? 1 : (i1 % 2 == 0) ? 2 : 3;
                else k = ((i1 + i2) % 3 == 2) ? 0 : ((i1 + i2) % 3 == 1) ? 1 : 2;
                if (k != cls) continue;
                int *q = c->pcord + (size_t)id * z;
                if (z == 4) {
                    q[0] = PSQ(i1, i2); q[1] = PSQ(i1 + 1, i2); q[2] = PSQ(i1 + 1, i2 + 1); q[3] = PSQ(i1, i2 + 1);
                } else {
                    q[0] = PHC(i1, i2, 2); q[1] = PHC(i1, i2, 1); q[2] = PHC(i1, i2 - 1, 2);
                    q[3] = PHC(i1 - 1, i2 - 1, 1); q[4] = PHC(i1 - 1, i2 - 1, 2); q[5] = PHC(i1 - 1, i2, 1);
                }
                id++;
            }
#undef PSQ
#undef PHC
#undef PWR
    double th = c->p.theta * M_PI, f = (z == 4) ? 0.5 : 1.0;
    zc z0 = 1.0 / c->nn_nf, z1 = c->p.alpha * cexp(CMPLX(0.0, -th)), z1h = c->p.alpha * cexp(CMPLX(0.0, th));
    zc V[36] = {0};
    /* the ring link i -> j = i+1 (mod z) of the reference's column-major initialisers
     * (dqmc_set_plqconf.f90:42-57): V(j, i) = z1, V(i, j) = z1h for even i; V(j, i) = -z1h,
     * V(i, j) = -z1 for odd i */
    for (int i = 0; i < z; i++) V[IDX(i, i, z)] = z0;
    for (int i = 0; i < z; i++) {
        int j = (i + 1) % z;
        V[IDX(j, i, z)] = (i % 2 == 0) ? z1 : -z1h;
        V[IDX(i, j, z)] = (i % 2 == 0) ? z1h : -z1;
    }
    zc iV[36];
    for (int k = 0; k < z * z; k++) iV[k] = I * V[k];
    const double eta[4] = {-sqrt(2.0 * (3.0 + sqrt(6.0))), -sqrt(2.0 * (3.0 - sqrt(6.0))), sqrt(2.0 * (3.0 - sqrt(6.0))),
                           sqrt(2.0 * (3.0 + sqrt(6.0)))};
    const double gam[4] = {1.0 - sqrt(6.0) / 3.0, 1.0 + sqrt(6.0) / 3.0, 1.0 + sqrt(6.0) / 3.0, 1.0 - sqrt(6.0) / 3.0};
    double a = sqrt(c->p.dtau * c->p.rhub_plq);
    for (int is = 0; is < 4; is++) {
        c->pgaml[is] = gam[is];
        expm_herm(z, iV, a * eta[is], c->pb[is]);
        expm_herm(z, iV, -a * eta[is], c->pbi[is]);
        c->pph[is] = cexp(CMPLX(0.0, -a * eta[is] * f * (c->p.nflr + c->p.nu)));
        for (int fl = 1; fl <= 3; fl++) {
            int isp = (is + fl) % 4;
            expm_herm(z, iV, a * (eta[isp] - eta[is]), c->pdl[is][fl - 1]);
            for (int i = 0; i < z; i++) c->pdl[is][fl - 1][IDX(i, i, z)] -= 1.0;
            c->ppr[is][fl - 1] = cexp(CMPLX(0.0, a * (eta[is] - eta[isp]) * f * (c->p.nflr + c->p.nu)));
        }
    }
}
static int fp_is(const orc_ctx *c, int nt, int p) { return c->confp[(size_t)(nt - 1) * c->lqp + p] - 1; }
static zc zpow(zc z, int n);
/* one plaquette: rows P of M <- T M(P, :) (left_forward with bmat, left_backward / _hc with
 * bmat_inv: dqmc_left_forward_prop.f90:22-40) */
static void pl_rows1(orc_ctx *c, int p, const zc *T, zc *M)
{
    int N = c->ndim, z = c->zp;
    const int *P = c->pcord + (size_t)p * z;
    zc x[6];
    for (int j = 0; j < N; j++) {
        for (int m = 0; m < z; m++) x[m] = M[IDX(P[m], j, N)];
        for (int r = 0; r < z; r++) {
            zc v = 0;
            for (int m = 0; m < z; m++) v += T[IDX(r, m, z)] * x[m];
            M[IDX(P[r], j, N)] = v;
        }
    }
}
/* columns P of M <- M(:, P) T (right_backward with bmat_inv, right_forward with bmat:
 * dqmc_right_backward_prop.f90:22-40) */
static void pl_cols1(orc_ctx *c, int p, const zc *T, zc *M)
{
    int N = c->ndim, z = c->zp;
    const int *P = c->pcord + (size_t)p * z;
    zc x[6];
    for (int i = 0; i < N; i++) {
        for (int m = 0; m < z; m++) x[m] = M[IDX(i, P[m], N)];
        for (int r = 0; r < z; r++) {
            zc v = 0;
            for (int m = 0; m < z; m++) v += x[m] * T[IDX(m, r, z)];
            M[IDX(i, P[r], N)] = v;
        }
    }
}
/* dqmc_update_plqu (src/dqmc_config_plqu/dqmc_update_plqu.f90:25-84): M = 1 + (1 - G)_PP
 * delta(iflip, is), ratio1 = det M (here by LU; the reference's s_inv_det_qr_z by QR), ratiotot
 * = ratio1^nflr gaml(isp) / gaml(is) phase_ratio(iflip, is); on acceptance G <- G - G(:, P)
 * delta M^-1 (1 - G)(P, :) */
static void pl_update(orc_ctx *c, int nt, int p)
{
    int N = c->ndim, z = c->zp;
    const int *P = c->pcord + (size_t)p * z;
    zc *G = c->G;
    int is = fp_is(c, nt, p);
    int iflip = (int)ceil(sfmt_stream(&c->rng) * 3.0);
    int isp = (is + iflip) % 4;
    const zc *dl = c->pdl[is][iflip - 1];
    double Mr[72];
    zc Mz[36];
    for (int a = 0; a < z; a++)
        for (int b = 0; b < z; b++) {
            zc v = (a == b) ? 1.0 : 0.0;
            for (int m = 0; m < z; m++) v += ((a == m ? 1.0 : 0.0) - G[IDX(P[a], P[m], N)]) * dl[IDX(m, b, z)];
            Mz[IDX(a, b, z)] = v;
            Mr[2 * IDX(a, b, z)] = creal(v);
            Mr[2 * IDX(a, b, z) + 1] = cimag(v);
        }
    double ld[2];
    int info = orc_inv_logdet(z, Mr, ld);   /* Mr <- M^-1 */
    zc ratio1 = info ? 0.0 : cexp(CMPLX(ld[0], ld[1]));
    zc ratiotot = zpow(ratio1, c->p.nflr) * (c->pgaml[isp] / c->pgaml[is]) * c->ppr[is][iflip - 1];
    double rr = creal(ratiotot * c->phase) / creal(c->phase);
    double random = sfmt_stream(&c->rng);
    c->pacc[1] += 1.0;
    (void)Mz;
    if (fabs(rr) > random) {
        c->pacc[0] += 1.0;
        c->phase *= ratiotot / cabs(ratiotot);
        zc S[36];
        for (int a = 0; a < z; a++)
            for (int b = 0; b < z; b++) {
                zc v = 0;
                for (int m = 0; m < z; m++) v += dl[IDX(a, m, z)] * CMPLX(Mr[2 * IDX(m, b, z)], Mr[2 * IDX(m, b, z) + 1]);
                S[IDX(a, b, z)] = v;
            }
        zc *cp = zalloc((size_t)N * z), *sv = zalloc((size_t)N * z);
        for (int m = 0; m < z; m++)
            for (int i = 0; i < N; i++) cp[IDX(i, m, N)] = G[IDX(i, P[m], N)];
        for (int j = 0; j < N; j++)
            for (int a = 0; a < z; a++) {
                zc v = 0;
                for (int m = 0; m < z; m++) v += S[IDX(a, m, z)] * ((j == P[m] ? 1.0 : 0.0) - G[IDX(P[m], j, N)]);
                sv[IDX(a, j, z)] = -v;
            }
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) {
                zc v = 0;
                for (int m = 0; m < z; m++) v += cp[IDX(i, m, N)] * sv[IDX(m, j, z)];
                G[IDX(i, j, N)] += v;
            }
        free(cp); free(sv);
        c->confp[(size_t)(nt - 1) * c->lqp + p] = isp + 1;
    }
}
/* every plaquette of slice nt on rows (left: bmat 1..lq; left_hc: bmat_inv lq..1) or columns
 * (right_backward: bmat_inv 1..lq) */
static void pl_rows(orc_ctx *c, int nt, int inv, int rev, zc *M)
{
    for (int q = 0; q < c->lqp; q++) {
        int p = rev ? c->lqp - 1 - q : q, is = fp_is(c, nt, p);
        pl_rows1(c, p, inv ? c->pbi[is] : c->pb[is], M);
    }
}
static void pl_cols(orc_ctx *c, int nt, int inv, int rev, zc *M)
{
    for (int q = 0; q < c->lqp; q++) {
        int p = rev ? c->lqp - 1 - q : q, is = fp_is(c, nt, p);
        pl_cols1(c, p, inv ? c->pbi[is] : c->pb[is], M);
    }
}

void orc_get_conf(const orc_ctx *c, int32_t *conf) { memcpy(conf, c->conf, sizeof(int32_t) * (size_t)c->ndim * c->ltrot); }
void orc_set_conf(orc_ctx *c, const int32_t *conf) { memcpy(c->conf, conf, sizeof(int32_t) * (size_t)c->ndim * c->ltrot); }

/* ------------------------------------------------------------------------------------ */
/* operator plugins (h0conf, uconf) acting on an N x N matrix                            */
/* ------------------------------------------------------------------------------------ */
/* kinetic, from the left: rows <- T * rows (dqmc_config_h0/dqmc_left_forward_prop.f90:18-44, :243) */
static void k_left(orc_ctx *c, int nf, const zc *T, zc *M)
{
    int N = c->ndim;
    if (!c->lwrapT || !(c->p.rt > 0)) return;
    if (c->p.kmode == 1) {
        zgemm('N', 'N', N, N, N, T, N, M, N, 0, c->Gtmp, N);
        memcpy(M, c->Gtmp, sizeof(zc) * (size_t)N * N);
        return;
    }
    for (int b = 0; b < c->nn_lf; b++) {
        int ist = b + nf * c->nn_lf;
        int i1 = c->nnlf[b], i2 = c->nnlist[i1 * c->nn_nf + nf];
        const zc *u = T + 4 * ist;
        for (int j = 0; j < N; j++) {
            zc a = M[IDX(i1, j, N)], bb = M[IDX(i2, j, N)];
            M[IDX(i1, j, N)] = u[0] * a + u[2] * bb;
            M[IDX(i2, j, N)] = u[1] * a + u[3] * bb;
        }
    }
}
/* kinetic, from the right: cols <- cols * T (dqmc_right_backward_prop.f90:18-44, :268) */
static void k_right(orc_ctx *c, int nf, const zc *T, zc *M)
{
    int N = c->ndim;
    if (!c->lwrapT || !(c->p.rt > 0)) return;
    if (c->p.kmode == 1) {
        zgemm('N', 'N', N, N, N, M, N, T, N, 0, c->Gtmp, N);
        memcpy(M, c->Gtmp, sizeof(zc) * (size_t)N * N);
        return;
    }
    for (int b = 0; b < c->nn_lf; b++) {
        int ist = b + nf * c->nn_lf;
        int i1 = c->nnlf[b], i2 = c->nnlist[i1 * c->nn_nf + nf];
        const zc *u = T + 4 * ist;
        zc *ca = M + (size_t)i1 * N, *cb = M + (size_t)i2 * N;
        for (int j = 0; j < N; j++) {
            zc a = ca[j], bb = cb[j];
            ca[j] = a * u[0] + bb * u[1];
            cb[j] = a * u[2] + bb * u[3];
        }
    }
}
/* the kinetic half-step of one slice, from the left, families in the given order */
static void khalf_left(orc_ctx *c, const zc *T, zc *M, int reverse)
{
    int nfam = (c->p.kmode == 1) ? 1 : c->nn_nf;
    for (int q = 0; q < nfam; q++) k_left(c, reverse ? nfam - 1 - q : q, T, M);
}

/* uconf row/column scalings (dqmc_config_u/dqmc_left_forward_prop.f90:27-32 and siblings) */
static void v_rows(orc_ctx *c, int nt, const zc *tab, int conjugate, zc *M)
{
    int N = c->ndim;
    if (!c->lwrapu) return;
    const int32_t *cf = c->conf + (size_t)(nt - 1) * N;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            zc b = tab[cf[i] - 1];
            M[IDX(i, j, N)] *= conjugate ? conj(b) : b;
        }
}
static void v_cols(orc_ctx *c, int nt, const zc *tab, zc *M)
{
    int N = c->ndim;
    if (!c->lwrapu) return;
    const int32_t *cf = c->conf + (size_t)(nt - 1) * N;
    for (int j = 0; j < N; j++) {
        zc b = tab[cf[j] - 1];
        for (int i = 0; i < N; i++) M[IDX(i, j, N)] *= b;
    }
}

/* Bmat_left_forward (src/dqmc_ft_core.f90:70-119): M <- B(nt1..nt2) M */
static void bmat_left_forward(orc_ctx *c, int nt1, int nt2, zc *M)
{
    for (int nt = nt2; nt <= nt1; nt++) {
        khalf_left(c, c->urt, M, 0);
        /* plaquettes 1..lq (src/dqmc_ft_core.f90:91-96) */
        if (c->lwrapp) pl_rows(c, nt, 0, 0, M);
        /* bond field, families 1..nn_nf, nflag 2 then 1 (src/dqmc_ft_core.f90:98-104) */
        for (int nf = 0; nf < c->nn_nf && c->lwrapv; nf++) { fv_left(c, nt, nf, 0, M); fv_left(c, nt, nf, 1, M); }
        v_rows(c, nt, c->bmat, 0, M);
        khalf_left(c, c->urt, M, 1);
    }
}
/* Bmat_left_forward_hc (src/dqmc_ft_core.f90:121-171): M <- B(nt1..nt2)^H M */
static void bmat_left_forward_hc(orc_ctx *c, int nt1, int nt2, zc *M)
{
    for (int nt = nt1; nt >= nt2; nt--) {
        khalf_left(c, c->urt, M, 0);
        v_rows(c, nt, c->bmat, 1, M);
        /* then the bond field, families nn_nf..1 (src/dqmc_ft_core.f90:141-156: the on-site field
         * comes first in the conjugated product); the real bond factors are their own conjugates */
        for (int nf = c->nn_nf - 1; nf >= 0 && c->lwrapv; nf--) { fv_left(c, nt, nf, 0, M); fv_left(c, nt, nf, 1, M); }
        /* plaquettes lq..1 with left_forward_prop_hc = bmat_inv (src/dqmc_ft_core.f90:158-163) */
        if (c->lwrapp) pl_rows(c, nt, 1, 1, M);
        khalf_left(c, c->urt, M, 1);
    }
}

/* ------------------------------------------------------------------------------------ */
/* local updates                                                                          */
/* ------------------------------------------------------------------------------------ */
static zc zpow(zc z, int n) { zc r = 1; for (int i = 0; i < n; i++) r *= z; return r; }

/* the Metropolis decision shared by all update flavours (dqmc_update_delay.f90:55-76) */
static int accept_site(orc_ctx *c, zc ratio1, int is, int iflip, int isp, double random, zc *ratiotot_out)
{
    zc ratiotot = zpow(ratio1, c->p.nflr);
    ratiotot = ratiotot * (c->gaml[isp - 1] / c->gaml[is - 1]) * c->phase_ratio[(iflip - 1) + (c->lcomp - 1) * (is - 1)];
    double ratio_re = creal(ratiotot * c->phase) / creal(c->phase);
    double ratio_re_abs = ratio_re < 0 ? -ratio_re : ratio_re;
    *ratiotot_out = ratiotot;
    return ratio_re_abs > random;
}
static void accept_phase(orc_ctx *c, zc ratiotot)
{
    double weight = sqrt(creal(ratiotot * conj(ratiotot)));
    c->phase = c->phase * ratiotot / weight;
}

/* fast rank-1 update: src/dqmc_config_u/dqmc_update.f90:33-86 */
static void update_fast(orc_ctx *c, int nt)
{
    int N = c->ndim, L = c->lcomp;
    zc *G = c->G;
    zc *u = zalloc(N), *v = zalloc(N);
    double accm = 0;
    for (int s = 0; s < N; s++) {
        int32_t *cf = c->conf + (size_t)(nt - 1) * N;
        int is = cf[s];
        int iflip = (int)ceil(sfmt_stream(&c->rng) * (L - 1));
        int isp = (is + iflip - 1) % L + 1;
        for (int j = 0; j < N; j++) v[j] = -G[IDX(s, j, N)];
        v[s] = 1.0 - G[IDX(s, s, N)];
        zc db = c->delta_bmat[(iflip - 1) + (L - 1) * (is - 1)];
        zc ratio1 = v[s] * db + 1.0;
        zc sscl = db / ratio1;
        zc rt;
        double random;
        int acc = accept_site(c, ratio1, is, iflip, isp, (random = sfmt_stream(&c->rng)), &rt);
        if (acc) {
            accm += 1;
            accept_phase(c, rt);
            for (int i = 0; i < N; i++) u[i] = G[IDX(i, s, N)];
            for (int j = 0; j < N; j++) {
                zc f = sscl * v[j];
                for (int i = 0; i < N; i++) G[IDX(i, j, N)] -= u[i] * f;
            }
            cf[s] = isp;
        }
    }
    c->accepted += accm;
    c->tried += N;
    free(u);
    free(v);
}

/* delay update: src/dqmc_config_u/dqmc_update_delay.f90:39-128 */
static void update_delay(orc_ctx *c, int nt)
{
    int N = c->ndim, L = c->lcomp, K = c->nublock;
    zc *G = c->G;
    zc *diagg = zalloc(N), *avec = zalloc((size_t)N * K), *bvec = zalloc((size_t)N * K);
    double accm = 0;
    int ik = 0;
    for (int i = 0; i < N; i++) diagg[i] = G[IDX(i, i, N)];
    int32_t *cf = c->conf + (size_t)(nt - 1) * N;
    for (int s = 0; s < N; s++) {
        int is = cf[s];
        int iflip = (int)ceil(sfmt_stream(&c->rng) * (L - 1));
        int isp = (is + iflip - 1) % L + 1;
        zc db = c->delta_bmat[(iflip - 1) + (L - 1) * (is - 1)];
        zc ratio1 = (1.0 - diagg[s]) * db + 1.0;
        zc sscl = db / ratio1;
        zc rt;
        double random = sfmt_stream(&c->rng);
        if (accept_site(c, ratio1, is, iflip, isp, random, &rt)) {
            accm += 1;
            accept_phase(c, rt);
            zc *a = avec + (size_t)ik * N, *b = bvec + (size_t)ik * N;
            for (int i = 0; i < N; i++) { a[i] = G[IDX(i, s, N)]; b[i] = G[IDX(s, i, N)]; }
            for (int m = 0; m < ik; m++) {
                zc bs = bvec[IDX(s, m, N)], as = avec[IDX(s, m, N)];
                const zc *am = avec + (size_t)m * N, *bm = bvec + (size_t)m * N;
                for (int i = 0; i < N; i++) { a[i] += bs * am[i]; b[i] += as * bm[i]; }
            }
            for (int i = 0; i < N; i++) a[i] *= sscl;
            b[s] -= 1.0;
            for (int i = 0; i < N; i++) diagg[i] += a[i] * b[i];
            ik++;
            cf[s] = isp;
        }
        if (ik == K || s == N - 1) {
            /* G += avec * bvec^T (zgemm 'N','T', always nublock columns, zero padded) */
            zgemm('N', 'T', N, N, K, avec, N, bvec, N, 1.0, G, N);
            ik = 0;
            if (s < N - 1) {
                for (int i = 0; i < N; i++) diagg[i] = G[IDX(i, i, N)];
                memset(avec, 0, sizeof(zc) * (size_t)N * K);
                memset(bvec, 0, sizeof(zc) * (size_t)N * K);
            }
        }
    }
    c->accepted += accm;
    c->tried += N;
    free(diagg); free(avec); free(bvec);
}

/* submatrix update: src/dqmc_config_u/dqmc_update_submatrix.f90:55-162 */
static void update_submatrix(orc_ctx *c, int nt)
{
    int N = c->ndim, L = c->lcomp, K = c->nublock;
    zc *G = c->G;
    int *pvec = (int *)calloc(K, sizeof(int));
    zc *gi = zalloc((size_t)K * K), *Gcut = zalloc((size_t)N * N);
    zc *cvec = zalloc(K), *bv = zalloc(K), *v1 = zalloc(K), *v4 = zalloc(K);
    zc *gm1 = zalloc((size_t)N * K), *gm2 = zalloc((size_t)K * N), *Gmat1 = zalloc((size_t)N * K);
    memcpy(Gcut, G, sizeof(zc) * (size_t)N * N);
    double accm = 0;
    int ik = 0;
    int32_t *cf = c->conf + (size_t)(nt - 1) * N;
#define GI(i, m) gi[IDX(i, m, K)]
    for (int s = 0; s < N; s++) {
        int is = cf[s];
        int iflip = (int)ceil(sfmt_stream(&c->rng) * (L - 1));
        int isp = (is + iflip - 1) % L + 1;
        for (int i = 0; i < ik; i++) { cvec[i] = G[IDX(s, pvec[i], N)]; bv[i] = G[IDX(pvec[i], s, N)]; }
        for (int i = 0; i < ik; i++) {
            zc acc = 0;
            for (int m = 0; m < ik; m++) acc += cvec[m] * GI(m, i);
            v1[i] = acc;
        }
        zc v = 0;
        for (int i = 0; i < ik; i++) v += v1[i] * bv[i];
        zc v2 = v + G[IDX(s, s, N)] - 1.0;
        zc db = c->delta_bmat[(iflip - 1) + (L - 1) * (is - 1)];
        zc ratio1 = -v2 * db + 1.0;
        zc rt;
        double random = sfmt_stream(&c->rng);
        if (accept_site(c, ratio1, is, iflip, isp, random, &rt)) {
            accm += 1;
            accept_phase(c, rt);
            pvec[ik] = s;
            zc v3 = 1.0 / (-v + 1.0 + 1.0 / db - G[IDX(s, s, N)]);
            GI(ik, ik) = v3;
            for (int i = 0; i < ik; i++) {
                zc acc = 0;
                for (int m = 0; m < ik; m++) acc += GI(i, m) * bv[m];
                v4[i] = acc;
            }
            for (int i = 0; i < ik; i++) { GI(ik, i) = v3 * v1[i]; GI(i, ik) = v4[i] * v3; }
            for (int m = 0; m < ik; m++)
                for (int i = 0; i < ik; i++) GI(i, m) += v4[i] * v3 * v1[m];
            zc f = db / (1.0 + db);
            for (int i = 0; i < N; i++) Gcut[IDX(i, s, N)] -= f * Gcut[IDX(i, s, N)];
            for (int i = 0; i < N; i++) gm1[IDX(i, ik, N)] = G[IDX(i, s, N)];
            ik++;
            cf[s] = isp;
        }
        if (ik == K || s == N - 1) {
            for (int i = 0; i < ik; i++)
                for (int j = 0; j < N; j++) gm2[IDX(i, j, K)] = Gcut[IDX(pvec[i], j, N)];
            if (ik > 0) {
                zgemm('N', 'N', N, ik, ik, gm1, N, gi, K, 0, Gmat1, N);
                zgemm('N', 'N', N, N, ik, Gmat1, N, gm2, K, 1.0, Gcut, N);
            }
            memcpy(G, Gcut, sizeof(zc) * (size_t)N * N);
            ik = 0;
        }
    }
#undef GI
    c->accepted += accm;
    c->tried += N;
    free(pvec); free(gi); free(Gcut); free(cvec); free(bv); free(v1); free(v4);
    free(gm1); free(gm2); free(Gmat1);
}

static void do_update(orc_ctx *c, int nt)
{
    if (c->p.update == 0) update_fast(c, nt);
    else if (c->p.update == 1) update_delay(c, nt);
    else update_submatrix(c, nt);
}

/* ------------------------------------------------------------------------------------ */
/* learned block proposals: a restatement of the block move of the HIP library, which has */
/* no counterpart in the reference.  It reuses the local update's Delta, field factor and  */
/* Metropolis test (dqmc_update_delay.f90:45-76) for a whole flipped set F:                 */
/* R = det(1 + Delta_F (1 - G)_FF)^nflr prod qf, accepted with                              */
/* min(1, |Re(R phase)| / |Re(phase)| exp(E_eff(C') - E_eff(C))).  The features are         */
/* recomputed from the whole field at every move (the library updates them per slice).     */
/* ------------------------------------------------------------------------------------ */
#define PROP_H 128
#define PROP_NW (PROP_H * PROP_H + 7 * PROP_H + 1)

static uint64_t smix_next(uint64_t *s)
{
    /* splitmix64 (Steele, Lea, Flood 2014) */
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double smix_unif(uint64_t *s) { return (double)(smix_next(s) >> 11) * (1.0 / 9007199254740992.0); }

void orc_set_proposal(orc_ctx *c, int z, const int32_t *nbr, int nprop, int local, const float *w, const double *norm)
{
    free(c->prop_nbr);
    free(c->prop_w);
    c->prop_nbr = NULL;
    c->prop_w = NULL;
    c->prop_on = (w != NULL);
    c->prop_local = c->prop_on ? local : 1;
    if (!c->prop_on) return;
    c->prop_z = z;
    c->prop_nprop = nprop;
    c->prop_nbr = (int32_t *)malloc(sizeof(int32_t) * (size_t)c->ndim * z);
    memcpy(c->prop_nbr, nbr, sizeof(int32_t) * (size_t)c->ndim * z);
    c->prop_w = (float *)malloc(sizeof(float) * PROP_NW);
    memcpy(c->prop_w, w, sizeof(float) * PROP_NW);
    memcpy(c->prop_norm, norm, sizeof(double) * 10);
    c->prop_rng = 0x9E3779B97F4A7C15ull * (uint64_t)(uint32_t)c->p.seed + 0x632BE59BD9B4E019ull;
    memset(c->prop_stats, 0, sizeof(c->prop_stats));
}

void orc_get_proposal_stats(const orc_ctx *c, double *out) { memcpy(out, c->prop_stats, sizeof(c->prop_stats)); }

/* the four global couplings of the field: sum s, sum_i s_i sum_{j~i} s_j, sum s s(t+1),
 * sum_i s_i sum_{j~i} s_j(t+1) (s = 2 conf - 3, slices periodic) */
static void prop_totals(const orc_ctx *c, double T[4])
{
    int N = c->ndim, lt = c->ltrot, z = c->prop_z;
    T[0] = T[1] = T[2] = T[3] = 0;
    for (int t = 0; t < lt; t++) {
        const int32_t *r0 = c->conf + (size_t)t * N, *r1 = c->conf + (size_t)((t + 1) % lt) * N;
        for (int i = 0; i < N; i++) {
            int s = 2 * r0[i] - 3;
            T[0] += s;
            T[2] += s * (2 * r1[i] - 3);
            for (int k = 0; k < z; k++) {
                int j = c->prop_nbr[(size_t)i * z + k];
                T[1] += s * (2 * r0[j] - 3);
                T[3] += s * (2 * r1[j] - 3);
            }
        }
    }
}

/* g(x) of the 4 -> 128 -> 128 -> 1 tanh network in double from the float weights */
static double prop_mlp(const orc_ctx *c, const double T[4])
{
    const float *w = c->prop_w;
    const float *W1 = w, *b1 = w + 4 * PROP_H, *W2 = w + 5 * PROP_H, *b2 = W2 + PROP_H * PROP_H, *w3 = b2 + PROP_H,
                *b3 = w3 + PROP_H;
    double ns = (double)c->ndim * c->ltrot, nbl = (double)c->ndim * c->prop_z * c->ltrot;
    double f[4] = {T[0] / ns, T[1] / nbl, T[2] / ns, T[3] / nbl}, x[4], h1[PROP_H];
    for (int k = 0; k < 4; k++) x[k] = (f[k] - c->prop_norm[k]) / c->prop_norm[4 + k];
    for (int j = 0; j < PROP_H; j++) {
        double a = b1[j];
        for (int k = 0; k < 4; k++) a += (double)W1[4 * j + k] * x[k];
        h1[j] = tanh(a);
    }
    double out = b3[0];
    for (int j = 0; j < PROP_H; j++) {
        double a = b2[j];
        for (int k = 0; k < PROP_H; k++) a += (double)W2[j * PROP_H + k] * h1[k];
        out += (double)w3[j] * tanh(a);
    }
    return out;
}

static void block_move(orc_ctx *c, int nt)
{
    int N = c->ndim, lt = c->ltrot, z = c->prop_z;
    int32_t *row = c->conf + (size_t)(nt - 1) * N;
    const int32_t *rm = c->conf + (size_t)((nt - 2 + lt) % lt) * N, *rp = c->conf + (size_t)(nt % lt) * N;
    int32_t *orig = (int32_t *)malloc(sizeof(int32_t) * N);
    memcpy(orig, row, sizeof(int32_t) * N);
    double escale = (double)N * lt * c->prop_norm[9];
    double T[4];
    prop_totals(c, T);
    double g0 = prop_mlp(c, T), gc = g0;
    int ninner = 0;
    for (int a = 0; a < c->prop_nprop; a++) {
        double u1 = smix_unif(&c->prop_rng), u2 = smix_unif(&c->prop_rng);
        int i = (int)(u1 * N);
        if (i > N - 1) i = N - 1;
        int s = 2 * row[i] - 3, nn = 0, st = 0;
        for (int k = 0; k < z; k++) {
            int j = c->prop_nbr[(size_t)i * z + k];
            nn += 2 * row[j] - 3;
            st += (2 * rm[j] - 3) + (2 * rp[j] - 3);
        }
        double T2[4] = {T[0] - 2 * s, T[1] - 4 * s * nn, T[2] - 2 * s * ((2 * rm[i] - 3) + (2 * rp[i] - 3)), T[3] - 2 * s * st};
        double gn = prop_mlp(c, T2), de = escale * (gn - gc);
        if (de <= 0.0 || u2 < exp(-de)) {
            memcpy(T, T2, sizeof(T));
            gc = gn;
            ninner++;
            row[i] = 3 - row[i];
        }
    }
    double u3 = smix_unif(&c->prop_rng);
    int F[64], m = 0;
    for (int i = 0; i < N; i++)
        if (row[i] != orig[i] && m < 64) F[m++] = i;
    int acc = 0;
    if (m > 0) {
        zc *M = zalloc((size_t)m * m), dl[64], ld;
        zc qprod = 1;
        for (int a = 0; a < m; a++) {
            int is = orig[F[a]], isp = 3 - is;
            dl[a] = c->delta_bmat[is - 1];
            qprod *= c->phase_ratio[is - 1] * (c->gaml[isp - 1] / c->gaml[is - 1]);
        }
        for (int b = 0; b < m; b++)
            for (int a = 0; a < m; a++)
                M[IDX(a, b, m)] = (a == b ? 1.0 : 0.0) + dl[a] * ((a == b ? 1.0 : 0.0) - c->G[IDX(F[a], F[b], N)]);
        int info = zinv_logdet(m, M, &ld);
        if (!info) {
            zc ratiotot = cexp((double)c->p.nflr * ld) * qprod;
            double ratio_re = creal(ratiotot * c->phase) / creal(c->phase);
            double lacc = log(fabs(ratio_re)) + escale * (gc - g0);
            acc = fabs(ratio_re) > 0.0 && log(u3) < lacc;
            if (acc) {
                /* X = M^-1 Delta_F (E_F - G_F:) (m x N); G <- G - G_:F X */
                zc *R = zalloc((size_t)m * N), *X = zalloc((size_t)m * N);
                for (int col = 0; col < N; col++)
                    for (int a = 0; a < m; a++)
                        R[IDX(a, col, m)] = dl[a] * ((F[a] == col ? 1.0 : 0.0) - c->G[IDX(F[a], col, N)]);
                zgemm('N', 'N', m, N, m, M, m, R, m, 0.0, X, m);
                zc *GF = zalloc((size_t)N * m);
                for (int a = 0; a < m; a++)
                    for (int i = 0; i < N; i++) GF[IDX(i, a, N)] = -c->G[IDX(i, F[a], N)];
                zgemm('N', 'N', N, N, m, GF, N, X, m, 1.0, c->G, N);
                free(R); free(X); free(GF);
                accept_phase(c, ratiotot);
            }
        }
        free(M);
    }
    if (!acc) memcpy(row, orig, sizeof(int32_t) * N);
    c->prop_stats[0] += 1;
    c->prop_stats[1] += acc;
    c->prop_stats[2] += m;
    c->prop_stats[3] += acc ? m : 0;
    c->prop_stats[4] += ninner;
    free(orig);
}

/* ------------------------------------------------------------------------------------ */
/* stabilization and Green's functions                                                    */
/* ------------------------------------------------------------------------------------ */
/* s_zmcpt (lib/s_matrix.f90:785-823): columns sorted by decreasing 2-norm (norms computed
 * once, first maximum wins); jpvt 1-based, A <- A P */
static void zmcpt(int m, int n, zc *A, int32_t *jpvt)
{
    double *vn = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; k++) { vn[k] = dznrm2(m, A + (size_t)k * m, 1); jpvt[k] = k + 1; }
    zc *tmp = zalloc(m);
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(vn[k]);
        for (int j = k + 1; j < n; j++) if (fabs(vn[j]) > best) { best = fabs(vn[j]); p = j; }
        if (p != k) {
            memcpy(tmp, A + (size_t)p * m, sizeof(zc) * m);
            memcpy(A + (size_t)p * m, A + (size_t)k * m, sizeof(zc) * m);
            memcpy(A + (size_t)k * m, tmp, sizeof(zc) * m);
            int32_t it = jpvt[p]; jpvt[p] = jpvt[k]; jpvt[k] = it;
            vn[p] = vn[k];
        }
    }
    free(tmp);
    free(vn);
}

/*
 * first point of the right (0b, n==1) or left (b0, n==nst) UDV chain, column-pivoted QR:
 *   dqmc_ft_stablize_0b_qr.f90:25-32 and dqmc_ft_stablize_b0_qr.f90:25-32.
 * Bd holds B*U_prev; outputs U = Q, D = Re diag(R), V = D^-1 triu(R) P^T, logdetQ.
 */
static void stabilize_first(orc_ctx *c, zc *Bd, zc *Uo, double *Do, zc *Vo, zc *ldq)
{
    int N = c->ndim;
    int32_t *jp = (int32_t *)malloc(sizeof(int32_t) * N);
    zc *R = zalloc((size_t)N * N);
    qr_full(N, N, Bd, Uo, R, jp, ldq, 1);
    for (int i = 0; i < N; i++) Do[i] = creal(R[IDX(i, i, N)]);
    /* s_invdiag_d_x_zr then s_zmp: column k of D^-1 R lands in column jpvt(k) */
    memset(Vo, 0, sizeof(zc) * (size_t)N * N);
    for (int k = 0; k < N; k++)
        for (int i = 0; i <= k; i++) Vo[IDX(i, jp[k] - 1, N)] = R[IDX(i, k, N)] / Do[i];
    free(R);
    free(jp);
}

/* the prepivoted branch needs the pivot order, so it is written out separately */
static void stabilize_prepivot(orc_ctx *c, zc *Bd, const double *Dprev, const zc *Vprev,
                               zc *Uo, double *Do, zc *Vo, zc *ldq)
{
    int N = c->ndim;
    int32_t *jp = (int32_t *)malloc(sizeof(int32_t) * N), *jd = (int32_t *)malloc(sizeof(int32_t) * N);
    zc *R = zalloc((size_t)N * N), *PV = zalloc((size_t)N * N);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) Bd[IDX(i, j, N)] *= Dprev[j];           /* s_z_x_diag_d */
    zmcpt(N, N, Bd, jp);                                                    /* s_zmcpt */
    qr_full(N, N, Bd, Uo, R, jd, ldq, 0);                                   /* s_zgeQR_logdetQ */
    for (int i = 0; i < N; i++) Do[i] = creal(R[IDX(i, i, N)]);
    for (int j = 0; j < N; j++)                                             /* s_invdiag_d_x_zr */
        for (int i = 0; i < N; i++) R[IDX(i, j, N)] = (i <= j) ? R[IDX(i, j, N)] / Do[i] : 0;
    for (int k = 0; k < N; k++)                                             /* s_zpm: P^T V_prev */
        for (int j = 0; j < N; j++) PV[IDX(k, j, N)] = Vprev[IDX(jp[k] - 1, j, N)];
    zgemm('N', 'N', N, N, N, R, N, PV, N, 0, Vo, N);
    free(jp); free(jd); free(R); free(PV);
}

/* s_dvec_min_max (lib/s_matrix.f90:752-783) */
static void dvec_min_max(int n, const double *d, double *dmax, double *dmin)
{
    for (int i = 0; i < n; i++) {
        if (fabs(d[i]) > 1.0) { dmax[i] = fabs(d[i]); dmin[i] = d[i] >= 0 ? 1.0 : -1.0; }
        else { dmax[i] = 1.0; dmin[i] = d[i]; }
    }
}

/* green_equaltime (src/dqmc_ft_core/green_equaltime.f90:1-107) */
static int green_equaltime(orc_ctx *c, const zc *ur, const double *dr, const zc *vr, const zc *vl,
                           const double *dl, const zc *ul, zc *gtt, zc ldqr, zc ldql, zc *logw)
{
    int N = c->ndim;
    double *drmax = malloc(sizeof(double) * N), *drmin = malloc(sizeof(double) * N);
    double *dlmax = malloc(sizeof(double) * N), *dlmin = malloc(sizeof(double) * N);
    dvec_min_max(N, dr, drmax, drmin);
    dvec_min_max(N, dl, dlmax, dlmin);
    zc *uu = zalloc((size_t)N * N), *vv = zalloc((size_t)N * N), *M = zalloc((size_t)N * N), *B = zalloc((size_t)N * N);
    zgemm('C', 'N', N, N, N, ur, N, ul, N, 0, uu, N);
    zgemm('N', 'C', N, N, N, vr, N, vl, N, 0, vv, N);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
            M[IDX(i, j, N)] = uu[IDX(i, j, N)] / (drmax[i] * dlmax[j]) + vv[IDX(i, j, N)] * drmin[i] * dlmin[j];
    zc zld;
    int info = zinv_logdet(N, M, &zld);
    /* s_v_invd_u: B = (ul / dlmax) M^-1 ; s_v_invd_uH: G = (B / drmax) ur^H */
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) uu[IDX(i, j, N)] = ul[IDX(i, j, N)] / dlmax[j];
    zgemm('N', 'N', N, N, N, uu, N, M, N, 0, B, N);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) B[IDX(i, j, N)] /= drmax[j];
    zgemm('N', 'C', N, N, N, B, N, ur, N, 0, gtt, N);
    double r1 = 0, r2 = 0;
    for (int i = 0; i < N; i++) r1 += log(dlmax[i]);
    for (int i = 0; i < N; i++) r2 += log(drmax[i]);
    *logw = -ldql + ldqr + zld + (r1 + r2);
    free(drmax); free(drmin); free(dlmax); free(dlmin); free(uu); free(vv); free(M); free(B);
    return info;
}

/* green_equaltimebb (src/dqmc_ft_core/green_equaltimebb.f90:1-84) */
static int green_equaltimebb(orc_ctx *c, const zc *ur, const double *dr, const zc *vr, zc *gtt, zc ldqr, zc *logw)
{
    int N = c->ndim;
    double *drmax = malloc(sizeof(double) * N), *drmin = malloc(sizeof(double) * N);
    dvec_min_max(N, dr, drmax, drmin);
    zc *M = zalloc((size_t)N * N);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
            M[IDX(i, j, N)] = conj(ur[IDX(j, i, N)]) / drmax[i] + vr[IDX(i, j, N)] * drmin[i];
    zc zld;
    int info = zinv_logdet(N, M, &zld);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) M[IDX(i, j, N)] /= drmax[j];
    zgemm('N', 'C', N, N, N, M, N, ur, N, 0, gtt, N);
    double r = 0;
    for (int i = 0; i < N; i++) r += log(drmax[i]);
    *logw = ldqr + zld + r;
    free(drmax); free(drmin); free(M);
    return info;
}

/* set_phase (src/dqmc_ft_core/set_phase.f90:1-24) */
static zc set_phase(orc_ctx *c, zc logw)
{
    zc w = cexp(CMPLX(0.0, (double)c->p.nflr * cimag(logw)));
    if (c->lwrapp)
        for (size_t k = 0; k < (size_t)c->ltrot * c->lqp; k++) w *= c->pph[c->confp[k] - 1];
    if (c->lwrapu && c->p.rhub > 0.0)
        for (size_t k = 0; k < (size_t)c->ltrot * c->ndim; k++) w *= c->phase_tab[c->conf[k] - 1];
    return w;
}

/* dqmc_ft_stablize_0b_qr(n) */
static void stablize_0b(orc_ctx *c, int n)
{
    int N = c->ndim;
    zc *Bd = zalloc((size_t)N * N);
    memcpy(Bd, c->Ust[n - 1], sizeof(zc) * (size_t)N * N);
    bmat_left_forward(c, c->wrap_step[2 * (n - 1) + 1], c->wrap_step[2 * (n - 1)], Bd);
    if (n == 1) stabilize_first(c, Bd, c->Ust[n], c->Dst[n], c->Vst[n], &c->logdetQst[n]);
    else stabilize_prepivot(c, Bd, c->Dst[n - 1], c->Vst[n - 1], c->Ust[n], c->Dst[n], c->Vst[n], &c->logdetQst[n]);
    free(Bd);
}

/* dqmc_ft_stablize_b0_qr(n): writes point n-1 of the left chain */
static void stablize_b0(orc_ctx *c, int n)
{
    int N = c->ndim;
    zc *Bd = zalloc((size_t)N * N), *U = zalloc((size_t)N * N), *V = zalloc((size_t)N * N);
    double *D = (double *)calloc(N, sizeof(double));
    zc ldq;
    memcpy(Bd, c->Ust[n], sizeof(zc) * (size_t)N * N);
    bmat_left_forward_hc(c, c->wrap_step[2 * (n - 1) + 1], c->wrap_step[2 * (n - 1)], Bd);
    if (n == c->nst) stabilize_first(c, Bd, U, D, V, &ldq);
    else stabilize_prepivot(c, Bd, c->Dst[n], c->Vst[n], U, D, V, &ldq);
    memcpy(c->Ust[n - 1], U, sizeof(zc) * (size_t)N * N);
    memcpy(c->Vst[n - 1], V, sizeof(zc) * (size_t)N * N);
    memcpy(c->Dst[n - 1], D, sizeof(double) * N);
    c->logdetQst[n - 1] = ldq;
    free(Bd); free(U); free(V); free(D);
}

static void set_identity_point(orc_ctx *c, int n)
{
    int N = c->ndim;
    memset(c->Ust[n], 0, sizeof(zc) * (size_t)N * N);
    memset(c->Vst[n], 0, sizeof(zc) * (size_t)N * N);
    for (int i = 0; i < N; i++) { c->Ust[n][IDX(i, i, N)] = 1; c->Vst[n][IDX(i, i, N)] = 1; c->Dst[n][i] = 1; }
    c->logdetQst[n] = 0;
}

int orc_sweep_start(orc_ctx *c)
{
    if (c->nst <= 0) return -1;
    set_identity_point(c, 0);
    for (int n = 1; n <= c->nst; n++) stablize_0b(c, n);
    int info = green_equaltimebb(c, c->Ust[c->nst], c->Dst[c->nst], c->Vst[c->nst], c->G, c->logdetQst[c->nst], &c->logweightf);
    c->phase = set_phase(c, c->logweightf);
    return info;
}

/* ---- correlation bins of equaltime_measure ------------------------------------------------ */
#define NPBC(v, l) ((v) > (l) ? (v) - (l) : ((v) < 1 ? (v) + (l) : (v)))

/* the cell grid (l1, l2, l3) and the site of cell (x, y, z) (1-based, periodic) and orbital o, in
 * the numbering of setup_lattice */
static void cell_dims(const orc_ctx *c, int *l1, int *l2, int *l3)
{
    *l1 = c->p.la;
    *l2 = c->p.lattice == 2 ? 1 : c->p.lb;
    *l3 = c->p.lattice == 3 ? c->p.lc : 1;
}
static int site_at(const orc_ctx *c, int x, int y, int z, int o)
{
    int l1, l2, l3;
    cell_dims(c, &l1, &l2, &l3);
    x = NPBC(x, l1); y = NPBC(y, l2); z = NPBC(z, l3);
    switch (c->p.lattice) {
    case 1: return 2 * ((x - 1) * l2 + y - 1) + o;
    case 2: return x - 1;
    case 3: return (z - 1) * l1 * l2 + (x - 1) * l2 + y - 1;
    default: return (x - 1) * l2 + y - 1;
    }
}
/* latt%imj of the cells of sites i and j, 0-based (src/lattice.f90:471-479, :116-122, :1136-1145) */
static int imj_of(const orc_ctx *c, int i, int j)
{
    int l1, l2, l3;
    cell_dims(c, &l1, &l2, &l3);
    return (NPBC(c->cz[i] - c->cz[j], l3) - 1) * l1 * l2 + NPBC(c->cy[i] - c->cy[j], l2) +
           (NPBC(c->cx[i] - c->cx[j], l1) - 1) * l2 - 1;
}

/* Switches on the correlation bins of equaltime_measure (src/dqmc_measure/equaltime_measure.f90:
 * 44-68): zcpcm, zspsm, znn, pair_onsite as (imj, no_i, no_j), zn(no), the bond correlations zjj,
 * zbb, pair_nn, pair_sn per imj with zj and zb, and the energy terms zqsq, ztsq, ztq (plaquettes,
 * phase theta) and zint_v.  Per unit cell u: A site a (2u on the honeycomb), its neighbour
 * nnlist(a,1), its second neighbour snlist(a,1) (lattice.f90:887, :710, :309, :1370) and the
 * current partner (snlist on the honeycomb, nnlist elsewhere: equaltime_measure.f90:378-394,
 * equaltime_measure1d.f90:241-246); plaquette cycles from lattice.f90:954-957 (square) and
 * :767-772 (honeycomb). */
int orc_set_eqt(orc_ctx *c, double theta)
{
    int lq = c->lq, l1, l2, l3;
    cell_dims(c, &l1, &l2, &l3);
    free(c->eqb); free(c->ca); free(c->cnn); free(c->csn); free(c->cjj); free(c->plq);
    c->eqb = (zc *)calloc((size_t)20 * lq + 8, sizeof(zc));
    c->ca = (int *)calloc(lq, sizeof(int)); c->cnn = (int *)calloc(lq, sizeof(int));
    c->csn = (int *)calloc(lq, sizeof(int)); c->cjj = (int *)calloc(lq, sizeof(int));
    c->zplq = c->p.lattice == 0 ? 4 : (c->p.lattice == 1 ? 6 : 0);
    c->plq = (int *)calloc((size_t)lq * (c->zplq ? c->zplq : 1), sizeof(int));
    for (int u = 0; u < lq; u++) {
        int a = c->p.lattice == 1 ? 2 * u : u;
        int x = c->cx[a], y = c->cy[a], z = c->cz[a];
        c->ca[u] = a;
        c->cnn[u] = c->nnlist[a * c->nn_nf];
        switch (c->p.lattice) {
        case 1: c->csn[u] = site_at(c, x + 1, y + 1, 1, 0); break;
        case 2: c->csn[u] = site_at(c, x + 2, 1, 1, 0); break;
        case 3: c->csn[u] = site_at(c, x + 1, y + 1, z, 0); break;
        default: c->csn[u] = site_at(c, x + 1, y + 1, 1, 0); break;
        }
        c->cjj[u] = c->p.lattice == 1 ? c->csn[u] : c->cnn[u];
        int *q = c->plq + (size_t)u * c->zplq;
        if (c->p.lattice == 0) {
            q[0] = site_at(c, x, y, 1, 0); q[1] = site_at(c, x + 1, y, 1, 0);
            q[2] = site_at(c, x + 1, y + 1, 1, 0); q[3] = site_at(c, x, y + 1, 1, 0);
        } else if (c->p.lattice == 1) {
            q[0] = site_at(c, x, y, 1, 1); q[1] = site_at(c, x, y, 1, 0);
            q[2] = site_at(c, x, y - 1, 1, 1); q[3] = site_at(c, x - 1, y - 1, 1, 0);
            q[4] = site_at(c, x - 1, y - 1, 1, 1); q[5] = site_at(c, x - 1, y, 1, 0);
        }
    }
    c->theta = theta;
    c->eqt = 1;
    return 0;
}

/* eqb (complex): [0, 4lq) zcpcm, [4lq, 8lq) zspsm, [8lq, 12lq) znn, [12lq, 16lq) pair_onsite, each
 * (imj, no_i, no_j) at imj + lq (no_i + 2 no_j); then zjj, zbb, pair_nn, pair_sn (lq each); then
 * zn(1), zn(2), zj, zb, zqsq, ztsq, ztq, zint_v (zqsq without its 1/9, zint_v without rv); all
 * summed with zphi */
void orc_get_eqt(const orc_ctx *c, double *out)
{
    if (c->eqb) memcpy(out, c->eqb, sizeof(zc) * ((size_t)20 * c->lq + 8));
}

static void eqt_bins(orc_ctx *c, const zc *G, zc zphi)
{
    const int N = c->ndim, lq = c->lq;
    const double nf = (double)c->p.nflr;
    zc *b = c->eqb;
#define GF(i, j) G[IDX(i, j, N)]
#define GC(i, j) (((i) == (j) ? 1.0 : 0.0) - G[IDX(j, i, N)])
    /* measure_cpcm, measure_spsm, measure_nn, measure_paircorr's on-site part (:265-318, :484-499) */
    for (int j = 0; j < N; j++) {
        b[20 * lq + c->sub[j]] += GC(j, j) * zphi;
        for (int i = 0; i < N; i++) {
            size_t k = imj_of(c, i, j) + (size_t)lq * (c->sub[i] + 2 * c->sub[j]);
            zc gc = GC(i, j), g = GF(i, j);
            b[k] += gc * zphi;
            b[4 * lq + k] += gc * g * (1.0 - 1.0 / (nf * nf)) * zphi;
            b[8 * lq + k] += (GC(i, i) * GC(j, j) + gc * g / nf) * zphi;
            b[12 * lq + k] += g * g * zphi;
        }
    }
    /* measure_jj, measure_bondcorr, measure_paircorr's bond parts (:367-416, :418-482, :500-543);
     * expar = 1 (no field, flux or dimerisation) */
    for (int v = 0; v < lq; v++) {
        int j0 = c->ca[v], jj = c->cjj[v], jn = c->cnn[v], js = c->csn[v];
        c->eqb[20 * lq + 3] += (GC(j0, jn) + GC(jn, j0)) * zphi;
        for (int u = 0; u < lq; u++) {
            int i0 = c->ca[u], ij = c->cjj[u], in = c->cnn[u], is = c->csn[u];
            int imj = imj_of(c, i0, j0);
            zc z1 = (GC(i0, ij) - GC(ij, i0)) * (GC(j0, jj) - GC(jj, j0));
            zc z3 = GC(i0, jj) * GF(ij, j0) + GC(ij, j0) * GF(i0, jj) - GC(i0, j0) * GF(ij, jj) - GC(ij, jj) * GF(i0, j0);
            b[16 * lq + imj] += (-z1 - z3 / nf) * zphi;
            zc bb = (GC(j0, jn) + GC(jn, j0)) * (GC(i0, in) + GC(in, i0)) +
                    (GC(i0, jn) * GF(in, j0) + GC(in, j0) * GF(i0, jn) + GC(i0, j0) * GF(in, jn) + GC(in, jn) * GF(i0, j0)) / nf;
            b[17 * lq + imj] += bb * zphi;
            b[18 * lq + imj] += GF(i0, j0) * GF(in, jn) * zphi;
            b[19 * lq + imj] += GF(i0, j0) * GF(is, js) * zphi;
        }
    }
    for (int u = 0; u < lq; u++) {
        int i0 = c->ca[u], ij = c->cjj[u];
        b[20 * lq + 2] += I * (GC(i0, ij) - GC(ij, i0)) * zphi;
    }
    /* zqsq, ztsq, ztq (:143-263) over the plaquettes; zint_v (:127-141) over the bonds */
    zc qsq = 0, tsq = 0, tq = 0, iv = 0;
    const int zp = c->zplq;
    for (int u = 0; u < lq && zp; u++) {
        const int *q = c->plq + (size_t)u * zp;
        zc dq = 0, dt = 0;
        for (int s1 = 0; s1 < zp; s1++) {
            int i1 = q[s1], i1a = q[(s1 + 1) % zp];
            zc z1 = ((s1 + 1) % 2 == 0) ? -cexp(I * c->theta) : cexp(-I * c->theta);
            dq += GC(i1, i1);
            dt += 2.0 * creal(z1 * GC(i1, i1a));
        }
        qsq += dq * dq * nf;
        tsq += dt * dt * nf;
        tq += dt * (dq / 3.0) * nf;
        for (int s1 = 0; s1 < zp; s1++) {
            int i1 = q[s1], i1a = q[(s1 + 1) % zp];
            zc z1 = ((s1 + 1) % 2 == 0) ? -cexp(I * c->theta) : cexp(-I * c->theta), z2 = conj(z1);
            for (int s2 = 0; s2 < zp; s2++) {
                int i2 = q[s2], i2a = q[(s2 + 1) % zp];
                zc z3 = ((s2 + 1) % 2 == 0) ? -cexp(I * c->theta) : cexp(-I * c->theta), z4 = conj(z3);
                qsq += GC(i1, i2) * GF(i1, i2);
                tsq += z1 * z3 * GC(i1, i2a) * GF(i1a, i2) + z1 * z4 * GC(i1, i2) * GF(i1a, i2a) +
                       z2 * z3 * GC(i1a, i2a) * GF(i1, i2) + z2 * z4 * GC(i1a, i2) * GF(i1, i2a);
                tq += (z1 / 3.0) * GC(i1, i2) * GF(i1a, i2) + (z2 / 3.0) * GC(i1a, i2) * GF(i1, i2);
            }
        }
    }
    for (int bnd = 0; bnd < c->nn_lf; bnd++)
        for (int f = 0; f < c->nn_nf; f++) {
            int i0 = c->nnlf[bnd], in = c->nnlist[i0 * c->nn_nf + f];
            iv += creal(GC(i0, in));
        }
    b[20 * lq + 4] += qsq * zphi;
    b[20 * lq + 5] += tsq * zphi;
    b[20 * lq + 6] += tq * zphi;
    b[20 * lq + 7] += iv * zphi;
#undef GF
#undef GC
}
#undef NPBC

/* the staggering of site s: + on even x + y (square), x (chain), x + y + z (cubic), on sublattice A
 * (honeycomb) */
static double stag_of(const orc_ctx *c, int s)
{
    int par;
    switch (c->p.lattice) {
    case 1: return c->sub[s] ? -1.0 : 1.0;
    case 2: par = c->cx[s]; break;
    case 3: par = c->cx[s] + c->cy[s] + c->cz[s]; break;
    default: par = c->cx[s] + c->cy[s]; break;
    }
    return (par & 1) ? -1.0 : 1.0;
}

/* equaltime_measure (src/dqmc_measure/equaltime_measure.f90:1-69): sign, density, zkint,
 * zint_u and the staggered spin structure factor built from zspsm (:283-299) */
static void measure(orc_ctx *c)
{
    int N = c->ndim;
    const zc *G = c->G;
    c->nobs += 1;
    double sg = creal(c->phase) < 0 ? -1.0 : 1.0;
    zc zphi = (sg / creal(c->phase)) * c->phase;
    zc zne = 0, zk = 0, zu = 0, ss = 0;
    for (int i = 0; i < N; i++) {
        zc gc = 1.0 - G[IDX(i, i, N)];
        zne += gc;
        zu += gc * gc;
    }
    for (int b = 0; b < c->nn_lf; b++) {
        int i0 = c->nnlf[b];
        for (int nf = 0; nf < c->nn_nf; nf++) {
            int in = c->nnlist[i0 * c->nn_nf + nf];
            zc gc = -G[IDX(in, i0, N)] + (in == i0 ? 1.0 : 0.0);
            zk += -2.0 * creal(gc);
        }
    }
    /* the chain's zkint carries nflr * rt (src/dqmc_measure/equaltime_measure1d.f90:76) */
    if (c->p.lattice == 2) zk *= c->p.nflr * c->p.rt;
    double fac = 1.0 - 1.0 / (double)(c->p.nflr * c->p.nflr);
    for (int j = 0; j < N; j++) {
        double sj = stag_of(c, j);
        for (int i = 0; i < N; i++) {
            double si = stag_of(c, i);
            zc gc = (i == j ? 1.0 : 0.0) - G[IDX(j, i, N)];
            ss += si * sj * gc * G[IDX(i, j, N)];
        }
    }
    ss *= fac / (double)c->lq;
    c->obs[0] += sg;
    c->obs[1] += zne * zphi / (double)c->lq;
    c->obs[2] += zk * zphi;
    c->obs[3] += zu * zphi;
    c->obs[4] += ss * zphi;
    if (c->eqt) eqt_bins(c, G, zphi);
}

static void record_wrap(orc_ctx *c, const zc *gscr, zc logw_tmp, int info)
{
    int N = c->ndim;
    double mx = 0;
    for (size_t k = 0; k < (size_t)N * N; k++) { double d = cabs(gscr[k] - c->G[k]); if (d > mx) mx = d; }
    if (mx > c->max_wrap_error) c->max_wrap_error = mx;
    zc ph = set_phase(c, logw_tmp);
    double pd = cabs(c->phase - ph);
    if (pd > c->max_phasediff) c->max_phasediff = pd;
    c->nchecks += 1;
    c->sumlog10err += log10(mx > 0 ? mx : 1e-300);
    if (info == 0) {
        memcpy(c->G, gscr, sizeof(zc) * (size_t)N * N);
        c->logweightf = logw_tmp;
        c->phase = ph;
    }
    if (c->trace && c->trace_n < c->trace_cap) {
        memcpy(c->trace + (size_t)2 * N * N * c->trace_n, c->G, sizeof(zc) * (size_t)N * N);
        c->trace_n++;
    }
}

/* dqmc_ft_sweep_b0 (src/dqmc_ft_core/dqmc_ft_sweep_b0.f90:1-216) */
int orc_sweep_b0(orc_ctx *c, int lupdate, int lmeasure)
{
    int N = c->ndim, nst = c->nst;
    if (nst <= 0) return -1;
    zc *gs = zalloc((size_t)N * N);
    set_identity_point(c, nst);
    int rc = 0;
    for (int nt = c->ltrot; nt >= 1; nt--) {
        if (lmeasure) measure(c);
        if (c->lwrapT) {
            int nfam = (c->p.kmode == 1) ? 1 : c->nn_nf;
            for (int q = 0; q < nfam; q++) { k_right(c, q, c->urt, c->G); k_left(c, q, c->urtm1, c->G); }
        }
        if (c->lwrapu) {
            if (lupdate && c->prop_local) do_update(c, nt);
            if (lupdate && c->prop_on) block_move(c, nt);
            v_cols(c, nt, c->bmat, c->G);
            v_rows(c, nt, c->bmat_inv, 0, c->G);
        }
        /* bond field, families nn_nf..1 (src/dqmc_ft_core/dqmc_ft_sweep_b0.f90:78-89) */
        for (int nf = c->nn_nf - 1; nf >= 0 && c->lwrapv; nf--) {
            fv_right(c, nt, nf, 1, c->G);
            fv_left(c, nt, nf, 1, c->G);
            if (lupdate) fv_update(c, nt, nf);
            fv_right(c, nt, nf, 2, c->G);
            fv_left(c, nt, nf, 2, c->G);
        }
        /* plaquettes lq..1: update, right_forward (bmat), left_backward (bmat_inv)
         * (src/dqmc_ft_core/dqmc_ft_sweep_b0.f90:91-100) */
        for (int p = c->lqp - 1; p >= 0 && c->lwrapp; p--) {
            if (lupdate) pl_update(c, nt, p);
            int is = fp_is(c, nt, p);
            pl_cols1(c, p, c->pb[is], c->G);
            pl_rows1(c, p, c->pbi[is], c->G);
        }
        if (c->lwrapT) {
            int nfam = (c->p.kmode == 1) ? 1 : c->nn_nf;
            for (int q = nfam - 1; q >= 0; q--) { k_right(c, q, c->urt, c->G); k_left(c, q, c->urtm1, c->G); }
        }
        if ((c->iwrap_nt[nt - 1] > 0 || c->nwrap == 1) || nt == 1) {
            int n = c->iwrap_nt[nt - 1];
            zc *UR = zalloc((size_t)N * N), *VR = zalloc((size_t)N * N);
            double *DR = (double *)malloc(sizeof(double) * N);
            memcpy(UR, c->Ust[n], sizeof(zc) * (size_t)N * N);
            memcpy(VR, c->Vst[n], sizeof(zc) * (size_t)N * N);
            memcpy(DR, c->Dst[n], sizeof(double) * N);
            zc ldqr = c->logdetQst[n];
            stablize_b0(c, n + 1);
            zc lw;
            int info = green_equaltime(c, UR, DR, VR, c->Vst[n], c->Dst[n], c->Ust[n], gs, ldqr, c->logdetQst[n], &lw);
            record_wrap(c, gs, lw, info);
            free(UR); free(VR); free(DR);
            if (info != 0) { rc = info; break; }
        }
    }
    free(gs);
    return rc;
}

/* dqmc_ft_sweep_0b (src/dqmc_ft_core/dqmc_ft_sweep_0b.f90:1-288), without the ltau branch */
int orc_sweep_0b(orc_ctx *c, int lupdate, int lmeasure)
{
    int N = c->ndim;
    if (c->nst <= 0) return -1;
    zc *gs = zalloc((size_t)N * N);
    set_identity_point(c, 0);
    int rc = 0;
    for (int nt = 1; nt <= c->ltrot; nt++) {
        if (c->lwrapT) {
            int nfam = (c->p.kmode == 1) ? 1 : c->nn_nf;
            for (int q = 0; q < nfam; q++) { k_left(c, q, c->urt, c->G); k_right(c, q, c->urtm1, c->G); }
        }
        /* plaquettes 1..lq: left_forward, right_backward, update
         * (src/dqmc_ft_core/dqmc_ft_sweep_0b.f90:82-93) */
        for (int p = 0; p < c->lqp && c->lwrapp; p++) {
            int is = fp_is(c, nt, p);
            pl_rows1(c, p, c->pb[is], c->G);
            pl_cols1(c, p, c->pbi[is], c->G);
            if (lupdate) pl_update(c, nt, p);
        }
        /* bond field, families 1..nn_nf (src/dqmc_ft_core/dqmc_ft_sweep_0b.f90:95-106) */
        for (int nf = 0; nf < c->nn_nf && c->lwrapv; nf++) {
            fv_left(c, nt, nf, 0, c->G);
            fv_right(c, nt, nf, 0, c->G);
            if (lupdate) fv_update(c, nt, nf);
            fv_left(c, nt, nf, 1, c->G);
            fv_right(c, nt, nf, 1, c->G);
        }
        if (c->lwrapu) {
            v_rows(c, nt, c->bmat, 0, c->G);
            v_cols(c, nt, c->bmat_inv, c->G);
            if (lupdate && c->prop_local) do_update(c, nt);
            if (lupdate && c->prop_on) block_move(c, nt);
        }
        if (c->lwrapT) {
            int nfam = (c->p.kmode == 1) ? 1 : c->nn_nf;
            for (int q = nfam - 1; q >= 0; q--) { k_left(c, q, c->urt, c->G); k_right(c, q, c->urtm1, c->G); }
        }
        if (lmeasure) measure(c);
        if (c->iwrap_nt[nt] > 0) {
            int n = c->iwrap_nt[nt];
            zc *UL = zalloc((size_t)N * N), *VL = zalloc((size_t)N * N);
            double *DL = (double *)malloc(sizeof(double) * N);
            memcpy(UL, c->Ust[n], sizeof(zc) * (size_t)N * N);
            memcpy(VL, c->Vst[n], sizeof(zc) * (size_t)N * N);
            memcpy(DL, c->Dst[n], sizeof(double) * N);
            zc ldql = c->logdetQst[n];
            stablize_0b(c, n);
            zc lw;
            int info = green_equaltime(c, c->Ust[n], c->Dst[n], c->Vst[n], VL, DL, UL, gs, c->logdetQst[n], ldql, &lw);
            record_wrap(c, gs, lw, info);
            free(UL); free(VL); free(DL);
            if (info != 0) { rc = info; break; }
        }
    }
    free(gs);
    return rc;
}

void orc_get_g(const orc_ctx *c, double *g) { memcpy(g, c->G, sizeof(zc) * (size_t)c->ndim * c->ndim); }

/* ------------------------------------------------------------------------------------ */
/* time-displaced Green's functions (ltau)                                                */
/* ------------------------------------------------------------------------------------ */
/* ltau with ntdm = nint(dyntau/dtau) (src/model_para.f90:457): gt0, g0t, g00 and the bins
 * znn_tau(lq, 2, 2, ntdm+1) (src/dqmc_measure_2d.f90:30, :61) */
int orc_set_dyn(orc_ctx *c, int ntdm)
{
    if (ntdm < 0 || ntdm > c->ltrot) return -1;
    size_t nn = (size_t)c->ndim * c->ndim;
    free(c->gt0); free(c->g0t); free(c->g00); free(c->znn);
    c->gt0 = zalloc(nn); c->g0t = zalloc(nn); c->g00 = zalloc(nn);
    c->znn = zalloc((size_t)c->lq * 4 * (ntdm + 1));
    c->ltau = 1;
    c->ntdm = ntdm;
    c->xmax_dyn = 0; c->dyncount = 0; c->sumlog10dyn = 0;
    return 0;
}

/* green_tau (src/dqmc_ft_core/green_tau.f90:1-122): G(t,t), G(t,0), G(0,t), G(0,0) from the right
 * chain (ur, dr, vr) and the left chain (vl, dl, ul) at the same point; s_v_invd_u(V, d, U) =
 * V d^-1 U, s_v_d_u(V, d, U) = V d U and the uH forms with U^H (lib/s_matrix.f90:597-720) */
static int green_tau(orc_ctx *c, const zc *ur, const double *dr, const zc *vr_in, const zc *vl, const double *dl,
                     const zc *ul, zc *g00, zc *gt0, zc *g0t, zc *gtt, zc ldqr, zc ldql, zc *logw)
{
    int N = c->ndim;
    size_t nn = (size_t)N * N;
    double *drmax = malloc(sizeof(double) * N), *drmin = malloc(sizeof(double) * N);
    double *dlmax = malloc(sizeof(double) * N), *dlmin = malloc(sizeof(double) * N);
    dvec_min_max(N, dr, drmax, drmin);
    dvec_min_max(N, dl, dlmax, dlmin);
    zc *vlh = zalloc(nn), *vr = zalloc(nn), *uu = zalloc(nn), *vv = zalloc(nn), *M = zalloc(nn), *B = zalloc(nn), *T = zalloc(nn);
    memcpy(vr, vr_in, sizeof(zc) * nn);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) vlh[IDX(i, j, N)] = conj(vl[IDX(j, i, N)]);
    zgemm('C', 'N', N, N, N, ur, N, ul, N, 0, uu, N);
    zgemm('N', 'C', N, N, N, vr, N, vl, N, 0, vv, N);
    /* G(t,t) */
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
            M[IDX(i, j, N)] = uu[IDX(i, j, N)] / (drmax[i] * dlmax[j]) + vv[IDX(i, j, N)] * drmin[i] * dlmin[j];
    zc zld;
    int info = zinv_logdet(N, M, &zld);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) T[IDX(i, j, N)] = ul[IDX(i, j, N)] / dlmax[j];
    zgemm('N', 'N', N, N, N, T, N, M, N, 0, B, N);                      /* B = ul dlmax^-1 M^-1 */
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) T[IDX(i, j, N)] = B[IDX(i, j, N)] / drmax[j];
    zgemm('N', 'C', N, N, N, T, N, ur, N, 0, gtt, N);
    double r1 = 0, r2 = 0;
    for (int i = 0; i < N; i++) r1 += log(dlmax[i]);
    for (int i = 0; i < N; i++) r2 += log(drmax[i]);
    *logw = -ldql + ldqr + zld + (r1 + r2);
    /* G(t,0) = B drmin vr */
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) T[IDX(i, j, N)] = B[IDX(i, j, N)] * drmin[j];
    zgemm('N', 'N', N, N, N, T, N, vr, N, 0, gt0, N);
    /* vr <- vr^-1, vlh <- (vl^H)^-1, vv = vlh vr */
    zc dummy;
    int info2 = zinv_logdet(N, vr, &dummy);
    int info3 = zinv_logdet(N, vlh, &dummy);
    zgemm('N', 'N', N, N, N, vlh, N, vr, N, 0, vv, N);
    /* G(0,0) */
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
            M[IDX(i, j, N)] = vv[IDX(i, j, N)] / (dlmax[i] * drmax[j]) + conj(uu[IDX(j, i, N)]) * dlmin[i] * drmin[j];
    int info4 = zinv_logdet(N, M, &dummy);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) T[IDX(i, j, N)] = vr[IDX(i, j, N)] / drmax[j];
    zgemm('N', 'N', N, N, N, T, N, M, N, 0, B, N);                      /* B = vr^-1 drmax^-1 M^-1 */
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) T[IDX(i, j, N)] = B[IDX(i, j, N)] / dlmax[j];
    zgemm('N', 'N', N, N, N, T, N, vlh, N, 0, g00, N);
    /* G(0,t) = B (-dlmin) ul^H */
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) T[IDX(i, j, N)] = -B[IDX(i, j, N)] * dlmin[j];
    zgemm('N', 'C', N, N, N, T, N, ul, N, 0, g0t, N);
    free(drmax); free(drmin); free(dlmax); free(dlmin);
    free(vlh); free(vr); free(uu); free(vv); free(M); free(B); free(T);
    return info ? info : (info2 ? info2 : (info3 ? info3 : info4));
}

/* dyn_measure (src/dqmc_measure/dyn_measure.f90:1-101), the active znn_tau term: bin nt (1-based)
 * += ((1/2 - Re gtt(i,i)) (1/2 - Re g00(j,j)) - g0t(j,i) gt0(i,j) / nflr) zphi, summed into
 * (imj, orbital of i, orbital of j); imj = npbc(yi - yj) + (npbc(xi - xj) - 1) l2 in 2D
 * (src/lattice.f90:471-479), + (npbc(zi - zj) - 1) l1 l2 in 3D (:1136-1145), npbc(xi - xj) in 1D
 * (:116-123) */
static void dyn_measure(orc_ctx *c, int nt, const zc *gtt)
{
    int N = c->ndim, l1 = c->p.la, l2 = c->p.lb, l3 = c->p.lattice == 3 ? c->p.lc : 1, lq = c->lq;
    if (c->p.lattice == 2) l2 = 1;
    double sg = creal(c->phase) < 0 ? -1.0 : 1.0;
    zc zphi = (sg / creal(c->phase)) * c->phase;
    zc *bin = c->znn + (size_t)lq * 4 * (nt - 1);
#define NPBC(v, l) ((v) > (l) ? (v) - (l) : ((v) < 1 ? (v) + (l) : (v)))
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            int imj = (NPBC(c->cz[i] - c->cz[j], l3) - 1) * l1 * l2 + NPBC(c->cy[i] - c->cy[j], l2) +
                      (NPBC(c->cx[i] - c->cx[j], l1) - 1) * l2;
            zc v = (0.5 - creal(gtt[IDX(i, i, N)])) * (0.5 - creal(c->g00[IDX(j, j, N)]))
                 - c->g0t[IDX(j, i, N)] * c->gt0[IDX(i, j, N)] / (double)c->p.nflr;
            bin[(imj - 1) + (size_t)lq * (c->sub[i] + 2 * c->sub[j])] += v * zphi;
        }
#undef NPBC
}

static void dyn_trace(orc_ctx *c)
{
    int N = c->ndim;
    size_t nn = (size_t)N * N;
    if (!c->dtrace || c->dtrace_n >= c->dtrace_cap) return;
    double *d = c->dtrace + (size_t)2 * 4 * nn * c->dtrace_n;
    memcpy(d, c->gt0, sizeof(zc) * nn);
    memcpy(d + 2 * nn, c->g0t, sizeof(zc) * nn);
    memcpy(d + 4 * nn, c->G, sizeof(zc) * nn);
    memcpy(d + 6 * nn, c->g00, sizeof(zc) * nn);
    c->dtrace_n++;
}

/* M <- M B(nt)^-1, the right-hand half of the 0b wrap (Bmat_right_backward, one slice) */
static void bmat_right_backward_slice(orc_ctx *c, int nt, zc *M)
{
    int nfam = (c->p.kmode == 1) ? 1 : c->nn_nf;
    if (c->lwrapT) for (int q = 0; q < nfam; q++) k_right(c, q, c->urtm1, M);
    /* plaquettes 1..lq (right_backward: bmat_inv, src/dqmc_ft_core.f90:257-262) */
    if (c->lwrapp) pl_cols(c, nt, 1, 0, M);
    /* bond field: right_backward nflag 2 then 1, families 1..nn_nf */
    for (int nf = 0; nf < c->nn_nf && c->lwrapv; nf++) { fv_right(c, nt, nf, 0, M); fv_right(c, nt, nf, 1, M); }
    v_cols(c, nt, c->bmat_inv, M);
    if (c->lwrapT) for (int q = nfam - 1; q >= 0; q--) k_right(c, q, c->urtm1, M);
}

static double max_abs_diff(size_t n, const zc *a, const zc *b)
{
    double mx = 0;
    for (size_t k = 0; k < n; k++) { double d = cabs(a[k] - b[k]); if (d > mx) mx = d; }
    return mx;
}

/* dqmc_ft_sweep_0b(lupdate = .false., lmeasure_equaltime, lmeasure_dyn = .true.) with ltau
 * (src/dqmc_ft_core/dqmc_ft_sweep_0b.f90:36-49, :160-204 with -DDYNERROR, dqmc_ft_core/dyn.f90):
 * the field is not updated; G(t,0), G(0,t) follow B(t) slice by slice and are replaced by
 * green_tau at every stabilization point; bins 1 .. ntdm+1 are measured (the reference writes
 * bin nt+1 for every slice, past the array when ntdm < ltrot) */
int orc_sweep_0b_dyn(orc_ctx *c, int lmeasure)
{
    int N = c->ndim;
    size_t nn = (size_t)N * N;
    if (c->nst <= 0 || !c->ltau) return -1;
    zc *gs = zalloc(nn), *gt0s = zalloc(nn), *g0ts = zalloc(nn);
    set_identity_point(c, 0);
    memcpy(c->g00, c->G, sizeof(zc) * nn);
    memcpy(c->gt0, c->G, sizeof(zc) * nn);
    memcpy(c->g0t, c->G, sizeof(zc) * nn);
    for (int i = 0; i < N; i++) c->g0t[IDX(i, i, N)] -= 1;
    dyn_measure(c, 1, c->G);
    dyn_trace(c);
    int rc = 0;
    for (int nt = 1; nt <= c->ltrot; nt++) {
        if (c->lwrapT) {
            int nfam = (c->p.kmode == 1) ? 1 : c->nn_nf;
            for (int q = 0; q < nfam; q++) { k_left(c, q, c->urt, c->G); k_right(c, q, c->urtm1, c->G); }
        }
        if (c->lwrapp) {
            pl_rows(c, nt, 0, 0, c->G);
            pl_cols(c, nt, 1, 0, c->G);
        }
        for (int nf = 0; nf < c->nn_nf && c->lwrapv; nf++) {
            fv_left(c, nt, nf, 0, c->G);
            fv_right(c, nt, nf, 0, c->G);
            fv_left(c, nt, nf, 1, c->G);
            fv_right(c, nt, nf, 1, c->G);
        }
        if (c->lwrapu) {
            v_rows(c, nt, c->bmat, 0, c->G);
            v_cols(c, nt, c->bmat_inv, c->G);
        }
        if (c->lwrapT) {
            int nfam = (c->p.kmode == 1) ? 1 : c->nn_nf;
            for (int q = nfam - 1; q >= 0; q--) { k_left(c, q, c->urt, c->G); k_right(c, q, c->urtm1, c->G); }
        }
        if (lmeasure) measure(c);
        if (c->iwrap_nt[nt] > 0) {
            int n = c->iwrap_nt[nt];
            zc *UL = zalloc(nn), *VL = zalloc(nn);
            double *DL = (double *)malloc(sizeof(double) * N);
            memcpy(UL, c->Ust[n], sizeof(zc) * nn);
            memcpy(VL, c->Vst[n], sizeof(zc) * nn);
            memcpy(DL, c->Dst[n], sizeof(double) * N);
            zc ldql = c->logdetQst[n];
            stablize_0b(c, n);
            /* -DDYNERROR: the propagated G(t,0), G(0,t) against green_tau's */
            bmat_left_forward(c, nt, nt, c->gt0);
            bmat_right_backward_slice(c, nt, c->g0t);
            zc lw;
            int info = green_tau(c, c->Ust[n], c->Dst[n], c->Vst[n], VL, DL, UL, c->g00, gt0s, g0ts, gs,
                                 c->logdetQst[n], ldql, &lw);
            double e1 = max_abs_diff(nn, c->gt0, gt0s), e2 = max_abs_diff(nn, c->g0t, g0ts);
            if (e1 > c->xmax_dyn) c->xmax_dyn = e1;
            if (e2 > c->xmax_dyn) c->xmax_dyn = e2;
            c->dyncount += 2;
            c->sumlog10dyn += log10(e1 > 0 ? e1 : 1e-300) + log10(e2 > 0 ? e2 : 1e-300);
            memcpy(c->gt0, gt0s, sizeof(zc) * nn);
            memcpy(c->g0t, g0ts, sizeof(zc) * nn);
            record_wrap(c, gs, lw, info);
            free(UL); free(VL); free(DL);
            if (info != 0) { rc = info; break; }
        } else {
            bmat_left_forward(c, nt, nt, c->gt0);
            bmat_right_backward_slice(c, nt, c->g0t);
        }
        if (nt + 1 <= c->ntdm + 1) dyn_measure(c, nt + 1, c->G);
        dyn_trace(c);
    }
    free(gs); free(gt0s); free(g0ts);
    return rc;
}

void orc_get_gtau(const orc_ctx *c, double *gt0, double *g0t, double *g00)
{
    size_t nn = (size_t)c->ndim * c->ndim;
    memcpy(gt0, c->gt0, sizeof(zc) * nn);
    memcpy(g0t, c->g0t, sizeof(zc) * nn);
    memcpy(g00, c->g00, sizeof(zc) * nn);
}
void orc_get_znn(const orc_ctx *c, double *znn) { memcpy(znn, c->znn, sizeof(zc) * (size_t)c->lq * 4 * (c->ntdm + 1)); }
void orc_reset_znn(orc_ctx *c) { memset(c->znn, 0, sizeof(zc) * (size_t)c->lq * 4 * (c->ntdm + 1)); }
void orc_get_dyn_error(const orc_ctx *c, double *out) { out[0] = c->xmax_dyn; out[1] = c->dyncount; out[2] = c->sumlog10dyn; }
/* per slice tau = 0 .. ltrot: gt0, g0t, gtt, g00 (4 N x N complex) into buf (cap entries) */
void orc_set_dyn_trace(orc_ctx *c, double *buf, int cap) { c->dtrace = buf; c->dtrace_cap = cap; c->dtrace_n = 0; }
int orc_dyn_trace_count(const orc_ctx *c) { return c->dtrace_n; }

void orc_get_diag(const orc_ctx *c, double *d)
{
    d[0] = c->max_wrap_error; d[1] = c->max_phasediff;
    d[2] = creal(c->phase); d[3] = cimag(c->phase);
    d[4] = creal(c->logweightf); d[5] = cimag(c->logweightf);
    d[6] = c->accepted; d[7] = c->tried; d[8] = c->nchecks; d[9] = c->sumlog10err;
}
void orc_reset_diag(orc_ctx *c)
{
    c->max_wrap_error = c->max_phasediff = c->nchecks = c->sumlog10err = c->accepted = c->tried = 0;
}
void orc_get_obs(const orc_ctx *c, double *o)
{
    o[0] = c->nobs;
    for (int k = 0; k < 5; k++) { o[1 + 2 * k] = creal(c->obs[k]); o[2 + 2 * k] = cimag(c->obs[k]); }
}
void orc_reset_obs(orc_ctx *c)
{
    c->nobs = 0;
    for (int k = 0; k < 5; k++) c->obs[k] = 0;
    if (c->eqb) memset(c->eqb, 0, sizeof(zc) * ((size_t)20 * c->lq + 8));
}
void orc_set_trace(orc_ctx *c, double *buf, int cap) { c->trace = buf; c->trace_cap = cap; c->trace_n = 0; }
int orc_trace_count(const orc_ctx *c) { return c->trace_n; }

int orc_get_table(const orc_ctx *c, int kind, double *out)
{
    int L = c->lcomp, N = c->ndim;
    zc *z = (zc *)out;
    switch (kind) {
    case 0:
        for (int i = 0; i < L; i++) { z[3 * i] = c->bmat[i]; z[3 * i + 1] = c->bmat_inv[i]; z[3 * i + 2] = c->phase_tab[i]; }
        return 3 * L;
    case 1: memcpy(z, c->delta_bmat, sizeof(zc) * (L - 1) * L); return (L - 1) * L;
    case 2: memcpy(z, c->phase_ratio, sizeof(zc) * (L - 1) * L); return (L - 1) * L;
    case 3: memcpy(out, c->gaml, sizeof(double) * L); return L;
    case 4: case 5: {
        size_t n = (c->p.kmode == 1) ? (size_t)N * N : (size_t)c->nn_nf * c->nn_lf * 4;
        memcpy(z, kind == 4 ? c->urt : c->urtm1, sizeof(zc) * n);
        return (int)n;
    }
    case 6:
        for (int nf = 0; nf < c->nn_nf; nf++)
            for (int b = 0; b < c->nn_lf; b++) {
                int i0 = c->nnlf[b];
                out[2 * (nf * c->nn_lf + b)] = i0;
                out[2 * (nf * c->nn_lf + b) + 1] = c->nnlist[i0 * c->nn_nf + nf];
            }
        return 2 * c->nn_nf * c->nn_lf;
    case 7: for (int i = 0; i <= c->ltrot; i++) out[i] = c->iwrap_nt[i]; return c->ltrot + 1;
    case 8: for (int i = 0; i < 2 * c->nst; i++) out[i] = c->wrap_step[i]; return 2 * c->nst;
    default: return -1;
    }
}

/* ==================================================================================== */
/* Projective DQMC (t-V model, bond HS field, submatrix-LR update)                       */
/* src/dqmc_proj_core.f90 and the files under src/dqmc_proj_core/ and src/dqmc_config_v/          */
/* ==================================================================================== */
struct orc_pctx {
    orc_pparams p;
    orc_ctx *b;            /* lattice, schedule, kinetic tables, measurement sums, RNG */
    int N, ne, ltrot, nst, nn_nf, nn_lf, nustock;
    double alpha;
    zc bp[2], bm[2], bpi[2], bmi[2], dp[2], dm[2];
    int32_t *conf;         /* [ltrot][nn_nf][N]; entries i < nn_lf are the bonds' fields */
    double *wtrial;        /* trial Hamiltonian eigenvalues, ascending */
    zc *projR;             /* N x ne (projL = projR) */
    zc *UR, *UL, *ULR, *ULRINV, *G, *tmp;
    zc **Ust;              /* (nst+1) x (N x ne) */
    double *logwDV;
    zc phase, logweightf;
    double max_wrap_error, max_phasediff, nchecks, sumlog10err, accepted, tried;
    int lwrapv;            /* bond field present (rv != 0); the on-site field lives in b (b->lwrapu) */
};

/* vconf%set_conf, lprojv = .false. (src/dqmc_config_v/dqmc_set_conf.f90:14-77):
 * alpha = acosh(exp(dtau V / 2)), eta = (-1, +1) */
static void pq_setup_v(orc_pctx *c)
{
    double al = acosh(exp(c->p.dtau * c->p.rv * 0.5));
    const double eta[2] = {-1.0, 1.0};
    c->alpha = al;
    for (int is = 0; is < 2; is++) {
        c->bp[is] = exp(al * eta[is]);
        c->bpi[is] = exp(-al * eta[is]);
        c->bm[is] = exp(-al * eta[is]);
        c->bmi[is] = exp(al * eta[is]);
        int isp = (is + 1) % 2;              /* iflip = 1 */
        c->dp[is] = exp(al * (eta[isp] - eta[is])) - 1.0;
        c->dm[is] = exp(-al * (eta[isp] - eta[is])) - 1.0;
    }
}

orc_pctx *orc_pcreate(const orc_pparams *p)
{
    orc_params q;
    memset(&q, 0, sizeof(q));
    q.lattice = p->lattice; q.la = p->la; q.lb = p->lb; q.beta = p->beta; q.dtau = p->dtau;
    q.rt = p->rt; q.mu = p->mu; q.rhub = p->rhub; q.nu = p->nu; q.nflr = p->nflr; q.exact_hs = p->exact_hs;
    q.kmode = p->kmode; q.update = 2; q.nwrap = p->nwrap; q.nublock = 1; q.seed = p->seed; q.lc = 1;
    orc_pctx *c = (orc_pctx *)calloc(1, sizeof(orc_pctx));
    c->p = *p;
    orc_ctx *b = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    b->p = q;
    setup_lattice(b);
    b->ltrot = (int)lround(p->beta / p->dtau);
    setup_schedule(b);
    b->lwrapu = (p->rhub != 0.0);
    if (b->lwrapu) {
        setup_hs(b);      /* u0conf%set_conf (src/dqmc_config_u/dqmc_set_conf.f90), lproju = .false. */
        b->conf = (int32_t *)calloc((size_t)b->ndim * b->ltrot, sizeof(int32_t));
        for (size_t k = 0; k < (size_t)b->ndim * b->ltrot; k++) b->conf[k] = 1;
    }
    setup_kinetic(b);
    sfmt_init(&b->rng, p->seed);
    c->b = b;
    c->lwrapv = (p->rv != 0.0);
    c->N = b->ndim;
    c->ne = c->N / 2 + (int)(p->nu * (double)c->N / 8.0);   /* src/model_para.f90:376 */
    c->ltrot = b->ltrot; c->nst = b->nst;
    c->nn_nf = b->nn_nf; c->nn_lf = b->nn_lf;
    c->nustock = p->nustock;
    pq_setup_v(c);
    int N = c->N, ne = c->ne;
    c->conf = (int32_t *)calloc((size_t)N * c->nn_nf * c->ltrot, sizeof(int32_t));
    for (size_t k = 0; k < (size_t)N * c->nn_nf * c->ltrot; k++) c->conf[k] = 1;
    c->wtrial = (double *)calloc(N, sizeof(double));
    c->projR = zalloc((size_t)N * ne);
    c->UR = zalloc((size_t)N * ne);
    c->UL = zalloc((size_t)N * ne);
    c->ULR = zalloc((size_t)ne * ne);
    c->ULRINV = zalloc((size_t)ne * ne);
    c->G = zalloc((size_t)N * N);
    c->tmp = zalloc((size_t)N * (N > ne ? N : ne));
    c->Ust = (zc **)calloc(c->nst + 1, sizeof(zc *));
    for (int n = 0; n <= c->nst; n++) c->Ust[n] = zalloc((size_t)N * ne);
    c->logwDV = (double *)calloc(c->nst + 2, sizeof(double));
    c->phase = 1.0;
    b->G = c->G;
    return c;
}

void orc_pdestroy(orc_pctx *c)
{
    if (!c) return;
    orc_ctx *b = c->b;
    free(b->urt); free(b->urtm1); free(b->nnlf); free(b->nnlist); free(b->cx); free(b->cy); free(b->sub);
    free(b->iwrap_nt); free(b->wrap_step); free(b->conf); free(b);
    for (int n = 0; n <= c->nst; n++) free(c->Ust[n]);
    free(c->Ust); free(c->logwDV); free(c->conf); free(c->wtrial); free(c->projR);
    free(c->UR); free(c->UL); free(c->ULR); free(c->ULRINV); free(c->G); free(c->tmp);
    free(c);
}

void orc_psizes(const orc_pctx *c, int *out)
{
    out[0] = c->N; out[1] = c->ne; out[2] = c->ltrot; out[3] = c->nst;
    out[4] = c->nn_nf; out[5] = c->nn_lf; out[6] = c->nustock; out[7] = c->b->lq;
}

/* vconf%input_conf from scratch (src/dqmc_config_v/dqmc_input_conf.f90:55-66, one rank):
 * nt outer, site i, family nf innermost; conf(i,nf,nt) = ceiling(u * 2) */
void orc_pinit_conf(orc_pctx *c)
{
    int N = c->N, nf = c->nn_nf;
    /* dqmc_start draws u0conf%input_conf before v0conf%input_conf (src/dqmc_proj_ctrl.f90:27-36) */
    if (c->b->lwrapu) orc_init_conf(c->b);
    if (!c->lwrapv) return;
    for (int nt = 0; nt < c->ltrot; nt++)
        for (int i = 0; i < N; i++)
            for (int f = 0; f < nf; f++)
                c->conf[((size_t)nt * nf + f) * N + i] = (int32_t)ceil(sfmt_stream(&c->b->rng) * 2.0);
}
void orc_pget_conf(const orc_pctx *c, int32_t *conf) { memcpy(conf, c->conf, sizeof(int32_t) * (size_t)c->N * c->nn_nf * c->ltrot); }
void orc_pget_uconf(const orc_pctx *c, int32_t *conf)
{
    if (c->b->lwrapu) memcpy(conf, c->b->conf, sizeof(int32_t) * (size_t)c->N * c->ltrot);
}
void orc_pset_uconf(orc_pctx *c, const int32_t *conf)
{
    if (c->b->lwrapu) memcpy(c->b->conf, conf, sizeof(int32_t) * (size_t)c->N * c->ltrot);
}
void orc_pset_conf(orc_pctx *c, const int32_t *conf) { memcpy(c->conf, conf, sizeof(int32_t) * (size_t)c->N * c->nn_nf * c->ltrot); }

/* cyclic Jacobi diagonalisation of a real symmetric matrix (a destroyed, v = eigenvectors,
 * w ascending); the trial Hamiltonians of this path are real, so zheev's eigenvalues and an
 * orthonormal basis of each eigenspace are what it returns up to rounding and phases */
static void jacobi_sym(int n, double *a, double *v, double *w)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) v[IDX(i, j, n)] = (i == j) ? 1.0 : 0.0;
    double scale = 0;
    for (size_t k = 0; k < (size_t)n * n; k++) scale += a[k] * a[k];
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int q = 0; q < n; q++)
            for (int p = 0; p < q; p++) off += a[IDX(p, q, n)] * a[IDX(p, q, n)];
        if (off <= 1e-34 * scale || off == 0) break;
        for (int q = 1; q < n; q++)
            for (int p = 0; p < q; p++) {
                double apq = a[IDX(p, q, n)];
                if (apq == 0) continue;
                double th = (a[IDX(q, q, n)] - a[IDX(p, p, n)]) / (2.0 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; k++) {   /* columns p, q */
                    double akp = a[IDX(k, p, n)], akq = a[IDX(k, q, n)];
                    a[IDX(k, p, n)] = cs * akp - sn * akq;
                    a[IDX(k, q, n)] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; k++) {   /* rows p, q */
                    double apk = a[IDX(p, k, n)], aqk = a[IDX(q, k, n)];
                    a[IDX(p, k, n)] = cs * apk - sn * aqk;
                    a[IDX(q, k, n)] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = v[IDX(k, p, n)], vkq = v[IDX(k, q, n)];
                    v[IDX(k, p, n)] = cs * vkp - sn * vkq;
                    v[IDX(k, q, n)] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = a[IDX(i, i, n)];
    /* ascending order (selection sort, swapping eigenvector columns) */
    for (int i = 0; i < n; i++) {
        int m = i;
        for (int j = i + 1; j < n; j++) if (w[j] < w[m]) m = j;
        if (m != i) {
            double t = w[i]; w[i] = w[m]; w[m] = t;
            for (int k = 0; k < n; k++) { double s = v[IDX(k, i, n)]; v[IDX(k, i, n)] = v[IDX(k, m, n)]; v[IDX(k, m, n)] = s; }
        }
    }
}

/* set_trial_h0 (src/dqmc_config_h0/set_trial_h0.f90:1-29) with expar = 1, then dqmc_set_proj
 * (src/dqmc_proj_core/dqmc_set_proj.f90:1-78): the lowest ne eigenvectors.  Returns the gap
 * w(ne+1) - w(ne) ("degen"). */
double orc_pset_trial(orc_pctx *c)
{
    int N = c->N, ne = c->ne;
    double *H = (double *)calloc((size_t)N * N, sizeof(double));
    double *V = (double *)calloc((size_t)N * N, sizeof(double));
    orc_ctx *b = c->b;
    for (int i = 0; i < c->nn_lf; i++) {
        int i0 = b->nnlf[i];
        for (int nf = 0; nf < c->nn_nf; nf++) {
            int in = b->nnlist[i0 * c->nn_nf + nf];
            double z1 = -c->p.rt + c->p.rndness * (sfmt_stream(&b->rng) - 1.0);
            H[IDX(i0, in, N)] += z1;
            H[IDX(in, i0, N)] += z1;
        }
    }
    for (int i = 0; i < N; i++) H[IDX(i, i, N)] -= c->p.mu;
    jacobi_sym(N, H, V, c->wtrial);
    for (int j = 0; j < ne; j++)
        for (int i = 0; i < N; i++) c->projR[IDX(i, j, N)] = V[IDX(i, j, N)];
    free(H); free(V);
    return (ne < N) ? c->wtrial[ne] - c->wtrial[ne - 1] : 0.0;
}
void orc_pset_proj(orc_pctx *c, const double *projR) { memcpy(c->projR, projR, sizeof(zc) * (size_t)c->N * c->ne); }
void orc_pget_proj(const orc_pctx *c, double *projR, double *w)
{
    memcpy(projR, c->projR, sizeof(zc) * (size_t)c->N * c->ne);
    if (w) memcpy(w, c->wtrial, sizeof(double) * c->N);
}

/* ---- operators on rectangular matrices ------------------------------------------------ */
/* kinetic family nf from the left on an N x m matrix (rows mixed) */
static void pk_left(orc_pctx *c, int nf, const zc *T, zc *M, int m)
{
    orc_ctx *b = c->b;
    int N = c->N;
    if (!b->lwrapT || !(c->p.rt > 0)) return;
    if (c->p.kmode == 1) {
        zgemm('N', 'N', N, m, N, T, N, M, N, 0, c->tmp, N);
        memcpy(M, c->tmp, sizeof(zc) * (size_t)N * m);
        return;
    }
    for (int q = 0; q < c->nn_lf; q++) {
        int ist = q + nf * c->nn_lf;
        int i1 = b->nnlf[q], i2 = b->nnlist[i1 * c->nn_nf + nf];
        const zc *u = T + 4 * ist;
        for (int j = 0; j < m; j++) {
            zc x = M[IDX(i1, j, N)], y = M[IDX(i2, j, N)];
            M[IDX(i1, j, N)] = u[0] * x + u[2] * y;
            M[IDX(i2, j, N)] = u[1] * x + u[3] * y;
        }
    }
}
/* kinetic family nf from the right on an m x N matrix (columns mixed) */
static void pk_right(orc_pctx *c, int nf, const zc *T, zc *M, int m)
{
    orc_ctx *b = c->b;
    int N = c->N;
    if (!b->lwrapT || !(c->p.rt > 0)) return;
    if (c->p.kmode == 1) {
        zgemm('N', 'N', m, N, N, M, m, T, N, 0, c->tmp, m);
        memcpy(M, c->tmp, sizeof(zc) * (size_t)m * N);
        return;
    }
    for (int q = 0; q < c->nn_lf; q++) {
        int ist = q + nf * c->nn_lf;
        int i1 = b->nnlf[q], i2 = b->nnlist[i1 * c->nn_nf + nf];
        const zc *u = T + 4 * ist;
        zc *ca = M + (size_t)i1 * m, *cb = M + (size_t)i2 * m;
        for (int j = 0; j < m; j++) {
            zc x = ca[j], y = cb[j];
            ca[j] = x * u[0] + y * u[1];
            cb[j] = x * u[2] + y * u[3];
        }
    }
}
static int pq_nfam(const orc_pctx *c) { return (c->p.kmode == 1) ? 1 : c->nn_nf; }

#define HAD 0.70710678118654752440   /* entries of u = ut = [[1,1],[1,-1]]/sqrt(2) (dqmc_config_v.f90:19-20) */
static const double pq_u[4] = {HAD, HAD, HAD, -HAD};   /* column-major u(1,1),u(2,1),u(1,2),u(2,2) */

static int pq_is(const orc_pctx *c, int nt, int nf, int q) { return c->conf[((size_t)(nt - 1) * c->nn_nf + nf) * c->N + q]; }

/* vconf left props on an N x m matrix (rows i1, i2 of each bond of family nf):
 *   mode 0: left_forward nflag 2 (dqmc_config_v/dqmc_left_forward_prop.f90:28-45): rows <- diag(bp,bm) ut rows
 *   mode 1: left_forward nflag 1 (:50-69) and left_backward nflag 2 (dqmc_left_backward_prop.f90): rows <- u rows
 *   mode 2: left_backward nflag 1 (dqmc_left_backward_prop.f90): rows <- u diag(bpi,bmi) rows */
static void pv_left(orc_pctx *c, int nt, int nf, int mode, zc *M, int m)
{
    int N = c->N;
    for (int q = 0; q < c->nn_lf; q++) {
        int i1 = c->b->nnlf[q], i2 = c->b->nnlist[i1 * c->nn_nf + nf];
        int is = pq_is(c, nt, nf, q) - 1;
        for (int j = 0; j < m; j++) {
            zc x = M[IDX(i1, j, N)], y = M[IDX(i2, j, N)];
            if (mode == 2) { x = c->bpi[is] * x; y = c->bmi[is] * y; }
            zc v1 = pq_u[0] * x + pq_u[2] * y;
            zc v2 = pq_u[1] * x + pq_u[3] * y;
            if (mode == 0) { v1 = c->bp[is] * v1; v2 = c->bm[is] * v2; }
            M[IDX(i1, j, N)] = v1;
            M[IDX(i2, j, N)] = v2;
        }
    }
}
/* vconf right props on an m x N matrix (columns i1, i2):
 *   mode 0: right_backward nflag 2 (dqmc_right_backward_prop.f90): cols <- (cols u) diag(bpi,bmi)
 *   mode 1: right_backward nflag 1 and right_forward nflag 2: cols <- cols u (u = ut symmetric)
 *   mode 2: right_forward nflag 1 (dqmc_right_forward_prop.f90): cols <- (cols diag(bp,bm)) ut */
static void pv_right(orc_pctx *c, int nt, int nf, int mode, zc *M, int m)
{
    for (int q = 0; q < c->nn_lf; q++) {
        int i1 = c->b->nnlf[q], i2 = c->b->nnlist[i1 * c->nn_nf + nf];
        int is = pq_is(c, nt, nf, q) - 1;
        zc *ca = M + (size_t)i1 * m, *cb = M + (size_t)i2 * m;
        for (int j = 0; j < m; j++) {
            zc x = ca[j], y = cb[j];
            if (mode == 2) { x = c->bp[is] * x; y = c->bm[is] * y; }
            zc v1 = x * pq_u[0] + y * pq_u[1];
            zc v2 = x * pq_u[2] + y * pq_u[3];
            if (mode == 0) { v1 = c->bpi[is] * v1; v2 = c->bmi[is] * v2; }
            ca[j] = v1;
            cb[j] = v2;
        }
    }
}

/* uconf props on the rows of an N x ne matrix (UR) or the columns of an ne x N matrix (UL),
 * by bmat (inv = 0) or bmat_inv (inv = 1) of the field at slice nt:
 * left_forward / left_backward (src/dqmc_config_u/dqmc_left_forward_prop.f90:23-32,
 * dqmc_left_backward_prop.f90:24-33), right_forward / right_backward
 * (dqmc_right_forward_prop.f90:24-33, dqmc_right_backward_prop.f90:24-33) */
static void pu_left(orc_pctx *c, int nt, int inv, zc *M)
{
    orc_ctx *b = c->b;
    int N = c->N;
    for (int i = 0; i < N; i++) {
        int is = b->conf[(size_t)(nt - 1) * N + i];
        zc f = inv ? b->bmat_inv[is - 1] : b->bmat[is - 1];
        for (int j = 0; j < c->ne; j++) M[IDX(i, j, N)] *= f;
    }
}
static void pu_right(orc_pctx *c, int nt, int inv, zc *M)
{
    orc_ctx *b = c->b;
    int N = c->N, ne = c->ne;
    for (int i = 0; i < N; i++) {
        int is = b->conf[(size_t)(nt - 1) * N + i];
        zc f = inv ? b->bmat_inv[is - 1] : b->bmat[is - 1];
        for (int r = 0; r < ne; r++) M[IDX(r, i, ne)] *= f;
    }
}

/* Bmat_left_forward of the projector core (src/dqmc_proj_core.f90:68-116) */
static void pq_bmat_left_forward(orc_pctx *c, int nt1, int nt2, zc *M)
{
    int nfam = pq_nfam(c);
    for (int nt = nt2; nt <= nt1; nt++) {
        for (int q = 0; q < nfam; q++) pk_left(c, q, c->b->urt, M, c->ne);
        if (c->lwrapv)
            for (int nf = 0; nf < c->nn_nf; nf++) {
                pv_left(c, nt, nf, 0, M, c->ne);
                pv_left(c, nt, nf, 1, M, c->ne);
            }
        if (c->b->lwrapu) pu_left(c, nt, 0, M);
        for (int q = nfam - 1; q >= 0; q--) pk_left(c, q, c->b->urt, M, c->ne);
    }
}

/* s_zgeQR_pqmc (lib/s_matrix.f90:1822-1996): Householder QR (zgeqrf) of the tall orientation,
 * Q formed by zungqr, log|det R| from the real diagonal, and Q's first column negated when an
 * odd number of R's diagonal entries is negative.  trans = 1: LQ of an ne x N matrix done as
 * QR of its plain transpose.  a (tall: N x ne, or wide: ne x N) is replaced by Q. */
static double zgeqr_pqmc(int rdim, int cdim, zc *a)
{
    int trans = rdim < cdim;
    int m = trans ? cdim : rdim, k = trans ? rdim : cdim;
    zc *w = zalloc((size_t)m * k), *tau = zalloc(k);
    if (!trans) memcpy(w, a, sizeof(zc) * (size_t)m * k);
    else
        for (int j = 0; j < m; j++)
            for (int i = 0; i < k; i++) w[IDX(j, i, m)] = a[IDX(i, j, k)];
    zgeqr_core(m, k, w, m, tau, NULL, 0);
    int isgn = 0;
    double rl = 0;
    for (int i = 0; i < k; i++) {
        double r = creal(w[IDX(i, i, m)]);
        if (r < 0) { isgn++; rl += log(-r); } else rl += log(r);
    }
    for (int j = 0; j < k; j++)
        for (int i = 0; i <= j; i++) w[IDX(i, j, m)] = (i == j) ? 1.0 : 0.0;
    zungqr(m, k, w, m, tau);
    if (isgn % 2 == 1)
        for (int i = 0; i < m; i++) w[IDX(i, 0, m)] = -w[IDX(i, 0, m)];
    if (!trans) memcpy(a, w, sizeof(zc) * (size_t)m * k);
    else
        for (int j = 0; j < m; j++)
            for (int i = 0; i < k; i++) a[IDX(i, j, k)] = w[IDX(j, i, m)];
    free(w); free(tau);
    return rl;
}
static double pq_qr(orc_pctx *c, zc *a, int trans)
{
    return trans ? zgeqr_pqmc(c->ne, c->N, a) : zgeqr_pqmc(c->N, c->ne, a);
}
/* exported for the invariant tests: a (rdim x cdim) is replaced by Q, returns log|det R| */
double orc_qr_pqmc(int rdim, int cdim, double *a) { return zgeqr_pqmc(rdim, cdim, (zc *)a); }

/* ULR = UL UR, ULRINV = ULR^-1 (s_inv_logdet_lu_z, or s_invlu_z when logdet is not wanted) */
static int pq_ulrinv(orc_pctx *c, const zc *ul, const zc *ur, zc *ulrinv, zc *logdet)
{
    int N = c->N, ne = c->ne;
    zgemm('N', 'N', ne, ne, N, ul, ne, ur, N, 0, ulrinv, ne);
    zc ld;
    int info = zinv_logdet(ne, ulrinv, &ld);
    if (logdet) *logdet = ld;
    return info;
}

/* green_equaltime (src/dqmc_proj_core/green_equaltime.f90:1-25): G = I - UR ULRINV UL */
static void pq_green(orc_pctx *c, const zc *ul, const zc *ur, const zc *ulrinv, zc *g)
{
    int N = c->N, ne = c->ne;
    zc *t = zalloc((size_t)N * ne);
    zgemm('N', 'N', N, ne, ne, ur, N, ulrinv, ne, 0, t, N);
    zgemm('N', 'N', N, N, ne, t, N, ul, ne, 0, g, N);
    for (size_t k = 0; k < (size_t)N * N; k++) g[k] = -g[k];
    for (int i = 0; i < N; i++) g[IDX(i, i, N)] += 1.0;
    free(t);
}

/* set_phase (src/dqmc_proj_core/set_phase.f90:1-24): the on-site field's phases when U > 0 */
static zc pq_phase(const orc_pctx *c, zc logw)
{
    zc w = cexp(CMPLX(0.0, (double)c->p.nflr * cimag(logw)));
    orc_ctx *b = c->b;
    if (b->lwrapu && c->p.rhub > 0.0)
        for (size_t k = 0; k < (size_t)c->ltrot * c->N; k++) w *= b->phase_tab[b->conf[k] - 1];
    return w;
}

/* submatrix-LR update of family nf at slice nt (src/dqmc_config_v/dqmc_proj_update_submatrix_LR.f90:1-298) */
static void pq_update(orc_pctx *c, int nt, int nf)
{
    int N = c->N, ne = c->ne, ns = c->nustock, nl = c->nn_lf;
    orc_ctx *b = c->b;
    zc *ulP = zalloc((size_t)ne * ns), *urP = zalloc((size_t)ns * ne);
    zc *Rrows = zalloc((size_t)ns * ne), *Felms = zalloc((size_t)ns * ns);
    zc *Ginv = zalloc((size_t)ns * ns), *Vmat = zalloc((size_t)ns * ne), *Umat = zalloc((size_t)ne * ns);
    zc *Vtmp = zalloc((size_t)ns * ne), *Utmp = zalloc((size_t)ne * ns);
    zc *ikk = zalloc((size_t)ns * 2), *kik = zalloc((size_t)2 * ns), *Pikk = zalloc((size_t)ns * 2), *Pkik = zalloc((size_t)2 * ns);
    int *xvec = (int *)calloc(ns, sizeof(int)), *svec = (int *)calloc(ns, sizeof(int));
    zc *Dvec = zalloc(ns);
    int ik = 0, istk_s = 0, istk_e = -1, nstk = 0;
    double accm = 0;
    for (int ib = 0; ib < nl; ib++) {
        int i1 = b->nnlf[ib], i2 = b->nnlist[i1 * c->nn_nf + nf];
        int32_t *cf = &c->conf[((size_t)(nt - 1) * c->nn_nf + nf) * N + ib];
        int is = *cf;
        int iflip = (int)ceil(sfmt_stream(&b->rng) * 1.0);
        int isp = (is + iflip - 1) % 2 + 1;
        if (nstk == 0) {
            int nup = ns < (nl - ib) * 2 ? ns : (nl - ib) * 2;
            memset(ulP, 0, sizeof(zc) * (size_t)ne * ns);
            memset(urP, 0, sizeof(zc) * (size_t)ns * ne);
            for (int q = 0; q < nup / 2; q++) {
                int i3 = b->nnlf[ib + q], i4 = b->nnlist[i3 * c->nn_nf + nf];
                for (int r = 0; r < ne; r++) {
                    ulP[IDX(r, 2 * q, ne)] = c->UL[IDX(r, i3, ne)];
                    ulP[IDX(r, 2 * q + 1, ne)] = c->UL[IDX(r, i4, ne)];
                    urP[IDX(2 * q, r, ns)] = c->UR[IDX(i3, r, N)];
                    urP[IDX(2 * q + 1, r, ns)] = c->UR[IDX(i4, r, N)];
                }
            }
            zgemm('N', 'N', nup, ne, ne, urP, ns, c->ULRINV, ne, 0, Rrows, ns);
            zgemm('N', 'N', nup, nup, ne, Rrows, ns, ulP, ne, 0, Felms, ns);
            istk_s = ib; istk_e = ib + nup / 2 - 1; nstk = nup / 2;
        }
        int s1 = (ib - istk_s) * 2, s2 = s1 + 1;
        zc v[4] = {0, 0, 0, 0};   /* column-major 2x2: v11, v21, v12, v22 */
        if (ik > 0) {
            for (int m = 0; m < ik; m++) {
                Pkik[IDX(0, m, 2)] = Felms[IDX(s1, svec[m], ns)];
                Pkik[IDX(1, m, 2)] = Felms[IDX(s2, svec[m], ns)];
                Pikk[IDX(m, 0, ns)] = Felms[IDX(svec[m], s1, ns)];
                Pikk[IDX(m, 1, ns)] = Felms[IDX(svec[m], s2, ns)];
            }
            zgemm('N', 'N', ik, 2, ik, Ginv, ns, Pikk, ns, 0, ikk, ns);
            zgemm('N', 'N', 2, 2, ik, Pkik, 2, ikk, ns, 1, v, 2);
        }
        v[0] = Felms[IDX(s1, s1, ns)] - v[0];
        v[1] = Felms[IDX(s2, s1, ns)] - v[1];
        v[2] = Felms[IDX(s1, s2, ns)] - v[2];
        v[3] = Felms[IDX(s2, s2, ns)] - v[3];
        zc dp = c->dp[is - 1], dm = c->dm[is - 1];
        v[0] = 1.0 + v[0] * dp;
        v[1] = v[1] * dp;
        v[2] = v[2] * dm;
        v[3] = 1.0 + v[3] * dm;
        zc ratio1 = v[0] * v[3] - v[2] * v[1];
        double ratio_re = creal(ratio1 * c->phase) / creal(c->phase);
        double ratio_abs = ratio_re < 0 ? -ratio_re : ratio_re;
        double random = sfmt_stream(&b->rng);
        if (ratio_abs > random) {
            accm += 1;
            double weight = sqrt(creal(ratio1 * conj(ratio1)));
            c->phase = c->phase * ratio1 / weight;
            zc gk[4];
            gk[0] = v[3] * dp / ratio1;
            gk[2] = -v[2] * dp / ratio1;
            gk[1] = -v[1] * dm / ratio1;
            gk[3] = v[0] * dm / ratio1;
            Ginv[IDX(ik, ik, ns)] = gk[0]; Ginv[IDX(ik + 1, ik, ns)] = gk[1];
            Ginv[IDX(ik, ik + 1, ns)] = gk[2]; Ginv[IDX(ik + 1, ik + 1, ns)] = gk[3];
            if (ik > 0) {
                zgemm('N', 'N', 2, ik, ik, Pkik, 2, Ginv, ns, 0, kik, 2);
                /* 21 block = -gk kik */
                for (int m = 0; m < ik; m++)
                    for (int r = 0; r < 2; r++)
                        Ginv[IDX(ik + r, m, ns)] = -(gk[IDX(r, 0, 2)] * kik[IDX(0, m, 2)] + gk[IDX(r, 1, 2)] * kik[IDX(1, m, 2)]);
                /* 12 block = -ikk gk */
                for (int r = 0; r < 2; r++)
                    for (int m = 0; m < ik; m++)
                        Ginv[IDX(m, ik + r, ns)] = -(ikk[IDX(m, 0, ns)] * gk[IDX(0, r, 2)] + ikk[IDX(m, 1, ns)] * gk[IDX(1, r, 2)]);
                /* 11 block += -ikk * (21 block) */
                for (int q = 0; q < ik; q++)
                    for (int m = 0; m < ik; m++)
                        Ginv[IDX(m, q, ns)] -= ikk[IDX(m, 0, ns)] * Ginv[IDX(ik, q, ns)] + ikk[IDX(m, 1, ns)] * Ginv[IDX(ik + 1, q, ns)];
            }
            ik += 2;
            xvec[ik - 2] = i1; xvec[ik - 1] = i2;
            svec[ik - 2] = s1; svec[ik - 1] = s2;
            Dvec[ik - 2] = dp; Dvec[ik - 1] = dm;
            for (int r = 0; r < ne; r++) {
                Vmat[IDX(ik - 2, r, ns)] = Rrows[IDX(s1, r, ns)];
                Vmat[IDX(ik - 1, r, ns)] = Rrows[IDX(s2, r, ns)];
                Umat[IDX(r, ik - 2, ne)] = ulP[IDX(r, s1, ne)];
                Umat[IDX(r, ik - 1, ne)] = ulP[IDX(r, s2, ne)];
            }
            *cf = isp;
        }
        if (ib == istk_e || ib == nl - 1) {
            if (ik > 0) {
                zgemm('N', 'N', ik, ne, ik, Ginv, ns, Vmat, ns, 0, Vtmp, ns);
                zgemm('N', 'N', ne, ik, ne, c->ULRINV, ne, Umat, ne, 0, Utmp, ne);
                zgemm('N', 'N', ne, ne, ik, Utmp, ne, Vtmp, ns, 0, c->tmp, ne);
                for (size_t k = 0; k < (size_t)ne * ne; k++) c->ULRINV[k] -= c->tmp[k];
                for (int m = 0; m < ik; m++) {
                    int x = xvec[m];
                    zc d = Dvec[m];
                    for (int r = 0; r < ne; r++) c->UR[IDX(x, r, N)] = c->UR[IDX(x, r, N)] + d * c->UR[IDX(x, r, N)];
                }
            }
            ik = 0; istk_s = 0; istk_e = -1; nstk = 0;
            memset(Ginv, 0, sizeof(zc) * (size_t)ns * ns);
            memset(Vmat, 0, sizeof(zc) * (size_t)ns * ne);
            memset(Umat, 0, sizeof(zc) * (size_t)ne * ns);
        }
    }
    c->accepted += accm;
    c->tried += nl;
    free(ulP); free(urP); free(Rrows); free(Felms); free(Ginv); free(Vmat); free(Umat);
    free(Vtmp); free(Utmp); free(ikk); free(kik); free(Pikk); free(Pkik); free(xvec); free(svec); free(Dvec);
}

/* uconf submatrix-LR update at slice nt (src/dqmc_config_u/dqmc_proj_update_submatrix_LR.f90:83-276):
 * stocks of nustock consecutive sites; Rrows = UR(P,:) ULRINV, Felms = Rrows UL(:,P); per site
 * the 1x1 ratio 1 + (F_kk - F_kP Gamma^-1 F_Pk) Delta, Gamma^-1 grown by bordering; at the end
 * of a stock ULRINV -= (ULRINV Umat)(Gamma^-1 Vmat) and the accepted rows of UR scaled by 1 + Delta */
static void pq_update_u(orc_pctx *c, int nt)
{
    int N = c->N, ne = c->ne, ns = c->nustock;
    orc_ctx *b = c->b;
    int lc = b->lcomp;
    zc *ulP = zalloc((size_t)ne * ns), *urP = zalloc((size_t)ns * ne);
    zc *Rrows = zalloc((size_t)ns * ne), *Felms = zalloc((size_t)ns * ns);
    zc *Ginv = zalloc((size_t)ns * ns), *Vmat = zalloc((size_t)ns * ne), *Umat = zalloc((size_t)ne * ns);
    zc *Vtmp = zalloc((size_t)ns * ne), *Utmp = zalloc((size_t)ne * ns);
    zc *ikk = zalloc(ns), *kik = zalloc(ns), *Pikk = zalloc(ns), *Pkik = zalloc(ns), *Dvec = zalloc(ns);
    int *xvec = (int *)calloc(ns, sizeof(int)), *svec = (int *)calloc(ns, sizeof(int));
    int ik = 0, istk_s = 0, istk_e = -1, nstk = 0;
    double accm = 0;
    for (int i = 0; i < N; i++) {
        int32_t *cf = &b->conf[(size_t)(nt - 1) * N + i];
        int is = *cf;
        int iflip = (int)ceil(sfmt_stream(&b->rng) * (lc - 1));
        int isp = (is + iflip - 1) % lc + 1;
        if (nstk == 0) {
            int nup = ns < N - i ? ns : N - i;
            memset(ulP, 0, sizeof(zc) * (size_t)ne * ns);
            memset(urP, 0, sizeof(zc) * (size_t)ns * ne);
            for (int q = 0; q < nup; q++)
                for (int r = 0; r < ne; r++) {
                    ulP[IDX(r, q, ne)] = c->UL[IDX(r, i + q, ne)];
                    urP[IDX(q, r, ns)] = c->UR[IDX(i + q, r, N)];
                }
            zgemm('N', 'N', nup, ne, ne, urP, ns, c->ULRINV, ne, 0, Rrows, ns);
            zgemm('N', 'N', nup, nup, ne, Rrows, ns, ulP, ne, 0, Felms, ns);
            istk_s = i; istk_e = i + nup - 1; nstk = nup;
        }
        int s = i - istk_s;
        zc vuk = 0;
        if (ik > 0) {
            for (int m = 0; m < ik; m++) {
                Pkik[m] = Felms[IDX(s, svec[m], ns)];
                Pikk[m] = Felms[IDX(svec[m], s, ns)];
            }
            for (int m = 0; m < ik; m++) {
                zc t = 0;
                for (int q = 0; q < ik; q++) t += Ginv[IDX(m, q, ns)] * Pikk[q];
                ikk[m] = t;
            }
            for (int m = 0; m < ik; m++) vuk += Pkik[m] * ikk[m];
        }
        vuk = Felms[IDX(s, s, ns)] - vuk;
        zc d = b->delta_bmat[(iflip - 1) + (lc - 1) * (is - 1)];
        zc ratio1 = 1.0 + vuk * d;
        zc ratiotot = zpow(ratio1, c->p.nflr) * (b->gaml[isp - 1] / b->gaml[is - 1]) *
                      b->phase_ratio[(iflip - 1) + (lc - 1) * (is - 1)];
        double ratio_re = creal(ratiotot * c->phase) / creal(c->phase);
        double random = sfmt_stream(&b->rng);
        if ((ratio_re < 0 ? -ratio_re : ratio_re) > random) {
            accm += 1;
            double weight = sqrt(creal(ratiotot * conj(ratiotot)));
            c->phase = c->phase * ratiotot / weight;
            zc gkk = d / ratio1;
            Ginv[IDX(ik, ik, ns)] = gkk;
            if (ik > 0) {
                for (int q = 0; q < ik; q++) {
                    zc t = 0;
                    for (int m = 0; m < ik; m++) t += Pkik[m] * Ginv[IDX(m, q, ns)];
                    kik[q] = t;
                }
                for (int q = 0; q < ik; q++) Ginv[IDX(ik, q, ns)] = -gkk * kik[q];
                for (int m = 0; m < ik; m++) Ginv[IDX(m, ik, ns)] = -ikk[m] * gkk;
                for (int q = 0; q < ik; q++)
                    for (int m = 0; m < ik; m++) Ginv[IDX(m, q, ns)] -= ikk[m] * Ginv[IDX(ik, q, ns)];
            }
            ik += 1;
            xvec[ik - 1] = i; svec[ik - 1] = s; Dvec[ik - 1] = d;
            for (int r = 0; r < ne; r++) {
                Vmat[IDX(ik - 1, r, ns)] = Rrows[IDX(s, r, ns)];
                Umat[IDX(r, ik - 1, ne)] = ulP[IDX(r, s, ne)];
            }
            *cf = isp;
        }
        if (i == istk_e || i == N - 1) {
            if (ik > 0) {
                zgemm('N', 'N', ik, ne, ik, Ginv, ns, Vmat, ns, 0, Vtmp, ns);
                zgemm('N', 'N', ne, ik, ne, c->ULRINV, ne, Umat, ne, 0, Utmp, ne);
                zgemm('N', 'N', ne, ne, ik, Utmp, ne, Vtmp, ns, 0, c->tmp, ne);
                for (size_t k = 0; k < (size_t)ne * ne; k++) c->ULRINV[k] -= c->tmp[k];
                for (int m = 0; m < ik; m++) {
                    int x = xvec[m];
                    zc dd = Dvec[m];
                    for (int r = 0; r < ne; r++) c->UR[IDX(x, r, N)] = c->UR[IDX(x, r, N)] + dd * c->UR[IDX(x, r, N)];
                }
            }
            ik = 0; istk_s = 0; istk_e = -1; nstk = 0;
            memset(Ginv, 0, sizeof(zc) * (size_t)ns * ns);
            memset(Vmat, 0, sizeof(zc) * (size_t)ns * ne);
            memset(Umat, 0, sizeof(zc) * (size_t)ne * ns);
        }
    }
    c->accepted += accm;
    c->tried += N;
    free(ulP); free(urP); free(Rrows); free(Felms); free(Ginv); free(Vmat); free(Umat);
    free(Vtmp); free(Utmp); free(ikk); free(kik); free(Pikk); free(Pkik); free(Dvec); free(xvec); free(svec);
}

/* dqmc_proj_sweep_start_0b (src/dqmc_proj_core/dqmc_proj_sweep_start_0b.f90:1-48) */
int orc_psweep_start(orc_pctx *c)
{
    int N = c->N, ne = c->ne;
    if (c->nst <= 0) return -1;
    memcpy(c->UR, c->projR, sizeof(zc) * (size_t)N * ne);
    c->logweightf = 0;
    c->logwDV[0] = 0;
    for (int n = 1; n <= c->nst; n++) {
        pq_bmat_left_forward(c, c->b->wrap_step[2 * (n - 1) + 1], c->b->wrap_step[2 * (n - 1)], c->UR);
        double lw = pq_qr(c, c->UR, 0);
        memcpy(c->Ust[n], c->UR, sizeof(zc) * (size_t)N * ne);
        c->logwDV[n] = c->logwDV[n - 1] + lw;
    }
    for (int l = 0; l < ne; l++)
        for (int i = 0; i < N; i++) c->UL[IDX(l, i, ne)] = conj(c->projR[IDX(i, l, N)]);
    zc z;
    int info = pq_ulrinv(c, c->UL, c->UR, c->ULRINV, &z);
    c->phase = pq_phase(c, z);
    c->logweightf = c->logwDV[c->nst] + z;
    return info;
}

static int pq_in_window(const orc_pctx *c, int nt)
{
    int lt = c->ltrot, ml = c->p.mid_len;
    return nt > (lt / 2 - (ml + 1) / 2) && nt <= (lt / 2 + ml / 2);
}

static void pq_measure(orc_pctx *c)
{
    pq_ulrinv(c, c->UL, c->UR, c->ULRINV, NULL);
    pq_green(c, c->UL, c->UR, c->ULRINV, c->G);
    c->b->G = c->G;
    c->b->phase = c->phase;
    measure(c->b);
}

/* the scratch recomputation at a stabilization point, and the wrap-error check of
 * dqmc_proj_sweep_0b.f90:121-190 (under WRAPERROR): G from the propagated UL, UR against G
 * from the orthonormalised ones, and the phase difference */
static int pq_scratch(orc_pctx *c, const zc *ul_old, const zc *ur_old, double logw_dv)
{
    int N = c->N, ne = c->ne;
    zc z;
    int info = pq_ulrinv(c, c->UL, c->UR, c->ULRINV, &z);
    zc lw = logw_dv + z;
    zc ph = pq_phase(c, lw);
    pq_green(c, c->UL, c->UR, c->ULRINV, c->G);
    if (c->p.wrap_check) {
        zc *inv = zalloc((size_t)ne * ne), *g2 = zalloc((size_t)N * N);
        pq_ulrinv(c, ul_old, ur_old, inv, NULL);
        pq_green(c, ul_old, ur_old, inv, g2);
        double mx = 0;
        for (size_t k = 0; k < (size_t)N * N; k++) { double d = cabs(g2[k] - c->G[k]); if (d > mx) mx = d; }
        if (mx > c->max_wrap_error) c->max_wrap_error = mx;
        double pd = cabs(c->phase - ph);
        if (pd > c->max_phasediff) c->max_phasediff = pd;
        c->nchecks += 1;
        c->sumlog10err += log10(mx > 0 ? mx : 1e-300);
        free(inv); free(g2);
    }
    c->phase = ph;
    c->logweightf = lw;
    if (c->b->trace && c->b->trace_n < c->b->trace_cap) {
        memcpy(c->b->trace + (size_t)2 * N * N * c->b->trace_n, c->G, sizeof(zc) * (size_t)N * N);
        c->b->trace_n++;
    }
    return info;
}

/* dqmc_proj_sweep_b0 (src/dqmc_proj_core/dqmc_proj_sweep_b0.f90:1-220) */
int orc_psweep_b0(orc_pctx *c, int lupdate, int lmeasure)
{
    int N = c->N, ne = c->ne, nst = c->nst, nfam = pq_nfam(c);
    orc_ctx *b = c->b;
    if (nst <= 0) return -1;
    for (int l = 0; l < ne; l++)
        for (int i = 0; i < N; i++) c->UL[IDX(l, i, ne)] = conj(c->projR[IDX(i, l, N)]);
    memcpy(c->Ust[0], c->projR, sizeof(zc) * (size_t)N * ne);
    zc *ul_old = zalloc((size_t)N * ne), *ur_old = zalloc((size_t)N * ne);
    int rc = 0;
    for (int nt = c->ltrot; nt >= 1; nt--) {
        if (lmeasure && pq_in_window(c, nt)) pq_measure(c);
        for (int q = 0; q < nfam; q++) { pk_right(c, q, b->urt, c->UL, ne); pk_left(c, q, b->urtm1, c->UR, ne); }
        if (b->lwrapu) {
            if (lupdate) pq_update_u(c, nt);
            pu_right(c, nt, 0, c->UL);
            pu_left(c, nt, 1, c->UR);
        }
        for (int nf = c->nn_nf - 1; nf >= 0 && c->lwrapv; nf--) {
            pv_right(c, nt, nf, 1, c->UL, ne);
            pv_left(c, nt, nf, 1, c->UR, ne);
            if (lupdate) pq_update(c, nt, nf);
            pv_right(c, nt, nf, 2, c->UL, ne);
            pv_left(c, nt, nf, 2, c->UR, ne);
        }
        for (int q = nfam - 1; q >= 0; q--) { pk_right(c, q, b->urt, c->UL, ne); pk_left(c, q, b->urtm1, c->UR, ne); }
        if ((b->iwrap_nt[nt - 1] > 0 || b->nwrap == 1) || nt == 1) {
            int n = b->iwrap_nt[nt - 1];
            memcpy(ul_old, c->UL, sizeof(zc) * (size_t)N * ne);
            memcpy(ur_old, c->UR, sizeof(zc) * (size_t)N * ne);
            memcpy(c->UR, c->Ust[n], sizeof(zc) * (size_t)N * ne);
            double lw = pq_qr(c, c->UL, 1);
            if (n + 1 == nst) c->logwDV[n + 1] = lw;
            else c->logwDV[n + 1] = c->logwDV[n + 2] + lw;
            double lsum = (n == 0) ? c->logwDV[n + 1] : c->logwDV[n + 1] + c->logwDV[n];
            int info = pq_scratch(c, ul_old, ur_old, lsum);
            for (int l = 0; l < ne; l++)
                for (int i = 0; i < N; i++) c->Ust[n][IDX(i, l, N)] = c->UL[IDX(l, i, ne)];
            if (info) { rc = info; break; }
        }
    }
    free(ul_old); free(ur_old);
    return rc;
}

/* dqmc_proj_sweep_0b (src/dqmc_proj_core/dqmc_proj_sweep_0b.f90:1-235), without the ltau branch */
int orc_psweep_0b(orc_pctx *c, int lupdate, int lmeasure)
{
    int N = c->N, ne = c->ne, nst = c->nst, nfam = pq_nfam(c);
    orc_ctx *b = c->b;
    if (nst <= 0) return -1;
    memcpy(c->UR, c->projR, sizeof(zc) * (size_t)N * ne);
    for (int l = 0; l < ne; l++)
        for (int i = 0; i < N; i++) c->Ust[nst][IDX(i, l, N)] = conj(c->projR[IDX(i, l, N)]);
    c->logwDV[0] = 0;
    zc *ul_old = zalloc((size_t)N * ne), *ur_old = zalloc((size_t)N * ne);
    int rc = 0;
    for (int nt = 1; nt <= c->ltrot; nt++) {
        for (int q = 0; q < nfam; q++) { pk_left(c, q, b->urt, c->UR, ne); pk_right(c, q, b->urtm1, c->UL, ne); }
        for (int nf = 0; nf < c->nn_nf && c->lwrapv; nf++) {
            pv_left(c, nt, nf, 0, c->UR, ne);
            pv_right(c, nt, nf, 0, c->UL, ne);
            if (lupdate) pq_update(c, nt, nf);
            pv_left(c, nt, nf, 1, c->UR, ne);
            pv_right(c, nt, nf, 1, c->UL, ne);
        }
        if (b->lwrapu) {
            pu_left(c, nt, 0, c->UR);
            pu_right(c, nt, 1, c->UL);
            if (lupdate) pq_update_u(c, nt);
        }
        for (int q = nfam - 1; q >= 0; q--) { pk_left(c, q, b->urt, c->UR, ne); pk_right(c, q, b->urtm1, c->UL, ne); }
        if (lmeasure && pq_in_window(c, nt)) pq_measure(c);
        if (b->iwrap_nt[nt] > 0) {
            int n = b->iwrap_nt[nt];
            memcpy(ul_old, c->UL, sizeof(zc) * (size_t)N * ne);
            memcpy(ur_old, c->UR, sizeof(zc) * (size_t)N * ne);
            for (int l = 0; l < ne; l++)
                for (int i = 0; i < N; i++) c->UL[IDX(l, i, ne)] = c->Ust[n][IDX(i, l, N)];
            double lw = pq_qr(c, c->UR, 0);
            c->logwDV[n] = c->logwDV[n - 1] + lw;
            double lsum = (n == nst) ? c->logwDV[n] : c->logwDV[n] + c->logwDV[n + 1];
            int info = pq_scratch(c, ul_old, ur_old, lsum);
            memcpy(c->Ust[n], c->UR, sizeof(zc) * (size_t)N * ne);
            if (info) { rc = info; break; }
        }
    }
    free(ul_old); free(ur_old);
    return rc;
}

/* G = I - UR (UL UR)^-1 UL from the current UL, UR (fresh inverse) */
void orc_pget_g(orc_pctx *c, double *g)
{
    zc *inv = zalloc((size_t)c->ne * c->ne);
    pq_ulrinv(c, c->UL, c->UR, inv, NULL);
    pq_green(c, c->UL, c->UR, inv, (zc *)g);
    free(inv);
}
void orc_pget_state(const orc_pctx *c, double *ul, double *ur, double *ulrinv)
{
    if (ul) memcpy(ul, c->UL, sizeof(zc) * (size_t)c->N * c->ne);
    if (ur) memcpy(ur, c->UR, sizeof(zc) * (size_t)c->N * c->ne);
    if (ulrinv) memcpy(ulrinv, c->ULRINV, sizeof(zc) * (size_t)c->ne * c->ne);
}
void orc_pget_diag(const orc_pctx *c, double *d)
{
    d[0] = c->max_wrap_error; d[1] = c->max_phasediff;
    d[2] = creal(c->phase); d[3] = cimag(c->phase);
    d[4] = creal(c->logweightf); d[5] = cimag(c->logweightf);
    d[6] = c->accepted; d[7] = c->tried; d[8] = c->nchecks; d[9] = c->sumlog10err;
}
void orc_preset_diag(orc_pctx *c)
{
    c->max_wrap_error = c->max_phasediff = c->nchecks = c->sumlog10err = c->accepted = c->tried = 0;
}
void orc_pget_obs(const orc_pctx *c, double *o) { orc_get_obs(c->b, o); }
int orc_pset_eqt(orc_pctx *c, double theta) { return orc_set_eqt(c->b, theta); }
void orc_pget_eqt(const orc_pctx *c, double *out) { orc_get_eqt(c->b, out); }
void orc_preset_obs(orc_pctx *c) { orc_reset_obs(c->b); }
void orc_pset_trace(orc_pctx *c, double *buf, int cap) { orc_set_trace(c->b, buf, cap); }
int orc_ptrace_count(const orc_pctx *c) { return c->b->trace_n; }
/* kind 0: (bp, bm, bpi, bmi) for is = 1, 2 -> 8 complex; kind 1: (dp, dm) for is = 1, 2 -> 4 complex;
 * other kinds as orc_get_table */
int orc_pget_table(const orc_pctx *c, int kind, double *out)
{
    zc *z = (zc *)out;
    if (kind == 100) {
        for (int is = 0; is < 2; is++) { z[4 * is] = c->bp[is]; z[4 * is + 1] = c->bm[is]; z[4 * is + 2] = c->bpi[is]; z[4 * is + 3] = c->bmi[is]; }
        return 8;
    }
    if (kind == 101) {
        for (int is = 0; is < 2; is++) { z[2 * is] = c->dp[is]; z[2 * is + 1] = c->dm[is]; }
        return 4;
    }
    return orc_get_table(c->b, kind, out);
}

// tv_kernels.hip -- the bond field of the finite-temperature t-V model (vconf, lprojv = .false.):
// the per-bond 2x2 props of one family on rows or columns of a chain matrix
// (src/dqmc_config_v/dqmc_{left,right}_{forward,backward}_prop.f90) and the sequential update
// of one family's bonds on G (src/dqmc_config_v/dqmc_update.f90:35-96).  The field of bond q of
// family nf at slice nt is confv[((nt - 1) nn_nf + nf) n + q] (1 or 2); u = ut = [[1, 1],
// [1, -1]] / sqrt(2); vtb = (bp, bm, bpi, bmi) of component 1 then 2, vdl = (dp, dm) of 1 then 2.
#pragma once

#define V_S 0.70710678118654752440

// rows (side 0) or columns (side 1) i1, i2 of every bond of family nf, one (bond, index) per
// thread:
//   rows: mode 0 diag(bp, bm) ut (left_forward nflag 2), 1 u (left_forward nflag 1,
//         left_backward nflag 2), 2 u diag(bpi, bmi) (left_backward nflag 1)
//   cols: mode 0 (. u) diag(bpi, bmi) (right_backward nflag 2), 1 . u (right_backward nflag 1,
//         right_forward nflag 2), 2 (. diag(bp, bm)) ut (right_forward nflag 1)
extern "C" __global__ void __launch_bounds__(256)
k_vprop(const DChain *chains, int msel, int pt, int nf, int mode, int side, int nt)
{
    const DChain &c = chains[blockIdx.z];
    if (!c.lwrapv) return;
    const int n = c.n, nl = c.nn_lf;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n * nl) return;
    const int q = side ? e / n : e % nl, j = side ? e % n : e / nl;
    const int ist = nf * nl + q;
    const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
    const int is = c.confv[((size_t)(nt - 1) * c.nn_nf + nf) * n + q] - 1;
    GAS cplx *M = gp(chain_mat(c, msel, pt));
    GAS cplx *px = side ? M + (size_t)i1 * n + j : M + (size_t)j * n + i1;
    GAS cplx *py = side ? M + (size_t)i2 * n + j : M + (size_t)j * n + i2;
    cplx x = gld(px), y = gld(py);
    const cplx bp = c.vtb[4 * is], bm = c.vtb[4 * is + 1], bpi = c.vtb[4 * is + 2], bmi = c.vtb[4 * is + 3];
    if (mode == 2) {
        x = cmul(side ? bp : bpi, x);
        y = cmul(side ? bm : bmi, y);
    }
    cplx v1 = cscale(cadd(x, y), V_S), v2 = cscale(csub(x, y), V_S);
    if (mode == 0) {
        v1 = cmul(side ? bpi : bp, v1);
        v2 = cmul(side ? bmi : bm, v2);
    }
    gst(px, v1);
    gst(py, v2);
}

// the bonds of family nf in order: ratio det(1 + (1 - G)_bb diag(dp, dm)) with no flavour power
// and no phase factor, accepted on |Re(ratio phase)| / |Re(phase)| > u2 (u1 picks the flip,
// iflip = ceiling(u1)); on acceptance G <- G - G_{:,b} diag(dp, dm) M^-1 (E_b - G_{b,:}) and
// phase <- phase ratio / |ratio|.  One workgroup per chain; every thread takes the same decisions
// from the same values, so the phase lives in registers.  u = the pass's bond draws of this
// family: slice `slice` of the pass, the p-th family updated in it (2 per bond).  LDS: the two
// rows and two columns of the rank-2 update.
extern "C" __global__ void __launch_bounds__(256)
k_vupdate(const DChain *chains, int nt, int nf, int slice, int p)
{
    extern __shared__ cplx vsh[];
    const DChain &c = chains[blockIdx.x];
    if (!c.lwrapv) return;
    const int n = c.n, nl = c.nn_lf, tid = threadIdx.x;
    cplx *r1 = vsh, *r2 = vsh + n, *c1 = vsh + 2 * n, *c2 = vsh + 3 * n;
    GAS cplx *G = gp(c.G);
    const double *u = c.unifv + (size_t)2 * nl * (c.nn_nf * slice + p);
    int32_t *cf = c.confv + ((size_t)(nt - 1) * c.nn_nf + nf) * n;
    cplx phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    double nacc = 0.0;
    for (int q = 0; q < nl; q++) {
        const int ist = nf * nl + q;
        const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
        const int is = cf[q] - 1;
        const int iflip = (int)ceil(u[2 * q]);
        const int isp = (is + iflip) & 1;
        const cplx dp = c.vdl[2 * is], dm = c.vdl[2 * is + 1];
        const cplx g11 = gld(G + (size_t)i1 * n + i1), g12 = gld(G + (size_t)i2 * n + i1);
        const cplx g21 = gld(G + (size_t)i1 * n + i2), g22 = gld(G + (size_t)i2 * n + i2);
        const cplx m11 = cadd(make_double2(1.0, 0.0), cmul(make_double2(1.0 - g11.x, -g11.y), dp));
        const cplx m12 = cmul(make_double2(-g12.x, -g12.y), dm);
        const cplx m21 = cmul(make_double2(-g21.x, -g21.y), dp);
        const cplx m22 = cadd(make_double2(1.0, 0.0), cmul(make_double2(1.0 - g22.x, -g22.y), dm));
        const cplx ratio = csub(cmul(m11, m22), cmul(m12, m21));
        const double rr = cmul(ratio, phase).x / phase.x;
        if (!(fabs(rr) > u[2 * q + 1])) continue;
        nacc += 1.0;
        phase = cmul(phase, cscale(ratio, 1.0 / hypot(ratio.x, ratio.y)));
        const cplx s11 = cdiv(cmul(dp, m22), ratio), s12 = cdiv(cmul(dp, make_double2(-m12.x, -m12.y)), ratio);
        const cplx s21 = cdiv(cmul(dm, make_double2(-m21.x, -m21.y)), ratio), s22 = cdiv(cmul(dm, m11), ratio);
        __syncthreads();                                   // the previous bond's update has read r, c
        for (int j = tid; j < n; j += 256) {
            const cplx gi1 = gld(G + (size_t)j * n + i1), gi2 = gld(G + (size_t)j * n + i2);
            const cplx a = make_double2((j == i1 ? 1.0 : 0.0) - gi1.x, -gi1.y);
            const cplx b = make_double2((j == i2 ? 1.0 : 0.0) - gi2.x, -gi2.y);
            const cplx t1 = cadd(cmul(s11, a), cmul(s12, b)), t2 = cadd(cmul(s21, a), cmul(s22, b));
            r1[j] = make_double2(-t1.x, -t1.y);
            r2[j] = make_double2(-t2.x, -t2.y);
            c1[j] = gld(G + (size_t)i1 * n + j);
            c2[j] = gld(G + (size_t)i2 * n + j);
        }
        __syncthreads();
        for (size_t k = tid; k < (size_t)n * n; k += 256) {
            const int i = (int)(k % n), j = (int)(k / n);
            gst(G + k, cadd(gld(G + k), cadd(cmul(c1[i], r1[j]), cmul(c2[i], r2[j]))));
        }
        __threadfence_block();
        __syncthreads();
        if (tid == 0) cf[q] = isp + 1;
    }
    if (tid == 0) {
        c.sc[SC_PHASE] = phase.x;
        c.sc[SC_PHASE + 1] = phase.y;
        c.vacc[0] += nacc;
        c.vacc[1] += (double)nl;
    }
}

// The delay form of the same update (src/dqmc_config_v/dqmc_update_delay.f90:28-128): the 2x2
// block of G at (i1, i2) is read as G + sum_m avec(:, m) bvec(:, m)^T over the ik delayed
// vectors; an accepted bond adds the pair avec(:, ik:ik+1) = [G(:, i1) G(:, i2)]' smat,
// bvec(:, ik:ik+1) = [G(i1, :) - e_i1, G(i2, :) - e_i2]' (both with the delayed sums), smat =
// diag(dp, dm) M^-1; every K vectors (K even) and after the family's last bond one MFMA rank-K
// product adds them to G.  Decisions, draws, phase and field bookkeeping as k_vupdate.  One
// workgroup of 256 threads per chain; LDS: the delayed rows of i1 and i2 (4 K) and the decision.
extern "C" __global__ void __launch_bounds__(256)
k_vupdate_delay(const DChain *chains, int nt, int nf, int slice, int p)
{
    extern __shared__ cplx vsh[];
    const DChain &c = chains[blockIdx.x];
    if (!c.lwrapv) return;
    const int n = c.n, nl = c.nn_lf, tid = threadIdx.x, lane = tid & 63;
    const int K = max(2, c.nublock & ~1);
    cplx *a1 = vsh, *a2 = vsh + K, *b1 = vsh + 2 * K, *b2 = vsh + 3 * K;   // avec(i1, :), avec(i2, :), bvec(i1, :), bvec(i2, :)
    __shared__ cplx s_g[4];
    __shared__ int s_acc;
    GAS cplx *G = gp(c.G), *A = gp(c.avec), *B = gp(c.bvec);
    const double *u = c.unifv + (size_t)2 * nl * (c.nn_nf * slice + p);
    int32_t *cf = c.confv + ((size_t)(nt - 1) * c.nn_nf + nf) * n;
    cplx phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    double nacc = 0.0;
    int ik = 0;
    for (int q = 0; q < nl; q++) {
        const int ist = nf * nl + q;
        const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
        // the delayed rows of the bond's two sites, then the 2x2 block summed by wave 0
        for (int m = tid; m < ik; m += blockDim.x) {
            a1[m] = gld(A + (size_t)m * n + i1);
            a2[m] = gld(A + (size_t)m * n + i2);
            b1[m] = gld(B + (size_t)m * n + i1);
            b2[m] = gld(B + (size_t)m * n + i2);
        }
        __syncthreads();
        if (tid < 64) {
            cplx t11 = make_double2(0.0, 0.0), t12 = t11, t21 = t11, t22 = t11;
            if (lane < ik) {
                t11 = cmul(a1[lane], b1[lane]);
                t12 = cmul(a1[lane], b2[lane]);
                t21 = cmul(a2[lane], b1[lane]);
                t22 = cmul(a2[lane], b2[lane]);
            }
            t11.x = wave_sum(t11.x); t11.y = wave_sum(t11.y);
            t12.x = wave_sum(t12.x); t12.y = wave_sum(t12.y);
            t21.x = wave_sum(t21.x); t21.y = wave_sum(t21.y);
            t22.x = wave_sum(t22.x); t22.y = wave_sum(t22.y);
            if (lane == 0) {
                s_g[0] = cadd(gld(G + (size_t)i1 * n + i1), t11);
                s_g[1] = cadd(gld(G + (size_t)i2 * n + i1), t12);
                s_g[2] = cadd(gld(G + (size_t)i1 * n + i2), t21);
                s_g[3] = cadd(gld(G + (size_t)i2 * n + i2), t22);
            }
        }
        __syncthreads();
        const int is = cf[q] - 1;
        const int iflip = (int)ceil(u[2 * q]);
        const int isp = (is + iflip) & 1;
        const cplx dp = c.vdl[2 * is], dm = c.vdl[2 * is + 1];
        const cplx g11 = s_g[0], g12 = s_g[1], g21 = s_g[2], g22 = s_g[3];
        const cplx m11 = cadd(make_double2(1.0, 0.0), cmul(make_double2(1.0 - g11.x, -g11.y), dp));
        const cplx m12 = cmul(make_double2(-g12.x, -g12.y), dm);
        const cplx m21 = cmul(make_double2(-g21.x, -g21.y), dp);
        const cplx m22 = cadd(make_double2(1.0, 0.0), cmul(make_double2(1.0 - g22.x, -g22.y), dm));
        const cplx ratio = csub(cmul(m11, m22), cmul(m12, m21));
        const double rr = cmul(ratio, phase).x / phase.x;
        const bool acc = fabs(rr) > u[2 * q + 1];
        if (acc) {
            nacc += 1.0;
            phase = cmul(phase, cscale(ratio, 1.0 / hypot(ratio.x, ratio.y)));
            const cplx s11 = cdiv(cmul(dp, m22), ratio), s12 = cdiv(cmul(dp, make_double2(-m12.x, -m12.y)), ratio);
            const cplx s21 = cdiv(cmul(dm, make_double2(-m21.x, -m21.y)), ratio), s22 = cdiv(cmul(dm, m11), ratio);
            // thread j builds entry j of the new pair (its delayed sums read column m of avec/bvec)
            for (int j = tid; j < n; j += blockDim.x) {
                cplx g1 = gld(G + (size_t)i1 * n + j), g2 = gld(G + (size_t)i2 * n + j);   // G(j, i1), G(j, i2)
                cplx r1 = gld(G + (size_t)j * n + i1), r2 = gld(G + (size_t)j * n + i2);   // G(i1, j), G(i2, j)
                if (j == i1) r1.x -= 1.0;
                if (j == i2) r2.x -= 1.0;
                for (int m = 0; m < ik; m++) {
                    const cplx am = gld(A + (size_t)m * n + j), bm = gld(B + (size_t)m * n + j);
                    g1 = cadd(g1, cmul(am, b1[m]));
                    g2 = cadd(g2, cmul(am, b2[m]));
                    r1 = cadd(r1, cmul(a1[m], bm));
                    r2 = cadd(r2, cmul(a2[m], bm));
                }
                gst(A + (size_t)ik * n + j, cadd(cmul(g1, s11), cmul(g2, s21)));
                gst(A + (size_t)(ik + 1) * n + j, cadd(cmul(g1, s12), cmul(g2, s22)));
                gst(B + (size_t)ik * n + j, r1);
                gst(B + (size_t)(ik + 1) * n + j, r2);
            }
            if (tid == 0) cf[q] = isp + 1;
            ik += 2;
        }
        __threadfence_block();
        __syncthreads();
        if (ik > 0 && (ik >= K || q == nl - 1)) {
            wg_rank_k(c.G, n, c.avec, c.bvec, ik, nullptr, nullptr);
            __threadfence_block();
            __syncthreads();
            ik = 0;
        }
    }
    if (tid == 0) {
        c.sc[SC_PHASE] = phase.x;
        c.sc[SC_PHASE + 1] = phase.y;
        c.vacc[0] += nacc;
        c.vacc[1] += (double)nl;
    }
}

// The submatrix form of the same update (src/dqmc_config_v/dqmc_update_submatrix.f90:63-250):
// G stays fixed within a block; the 2x2 block of the current G at a bond (i1, i2) is
// G_bb + v with v = G(b, P) Gamma^-1 G(P, b) over the accepted sites P (pairs, in order).  An
// accepted bond borders Gamma^-1 by a 2x2 block: v3 = v5^-1 with v5 = diag(1 + 1/dp, 1 + 1/dm)
// - G_bb - v, the new rows v3 v1 (v1 = G(b, P) Gamma^-1), the new columns v4 v3 (v4 = Gamma^-1
// G(P, b)), the old block + v4 v3 v1; the columns i1, i2 of the updated G take the factors
// 1/(1 + dp), 1/(1 + dm).  Every K entries (K even) and after the family's last bond the block
// is flushed: Gmat1 = G(:, P) Gamma^-1 as an MFMA product, G <- (G + Gmat1 G(P, :)) diag(scale).
// (The reference's v3(2,1) at :153 is -v5(2,1) v5(2,2) / (v5(1,1) det v5), not the inverse's
// -v5(2,1) / det v5; this kernel uses the inverse, with which the submatrix, delay and fast
// forms take the same decisions.)  Draws, decisions, phase and field bookkeeping as k_vupdate.
// One workgroup of 256 threads per chain; LDS: Gamma^-1 (leading dimension K + 1), G(b, P),
// G(P, b), v1, v4 (2 K each).
extern "C" __global__ void __launch_bounds__(256)
k_vupdate_submatrix(const DChain *chains, int nt, int nf, int slice, int p)
{
    extern __shared__ cplx vsh[];
    const DChain &c = chains[blockIdx.x];
    if (!c.lwrapv) return;
    const int n = c.n, nl = c.nn_lf, tid = threadIdx.x, lane = tid & 63;
    const int K = max(2, min(c.nublock, SH_KMAX) & ~1), KL = K + 1;
    cplx *gi = vsh;                     // Gamma^-1, column-major, leading dimension K + 1
    cplx *cv = gi + (size_t)KL * K;     // G(i1, P), G(i2, P): 2 x K, entry (r, m) at r + 2 m
    cplx *bv = cv + 2 * K;              // G(P, i1), G(P, i2): K x 2, entry (m, c) at m + K c
    cplx *v1 = bv + 2 * K;              // G(b, P) Gamma^-1: 2 x K, (r, m) at r + 2 m
    cplx *v4 = v1 + 2 * K;              // Gamma^-1 G(P, b): K x 2, (m, c) at m + K c
    __shared__ int pvec[SH_KMAX];
    __shared__ cplx s_gb[8], s_v[4];   // s_gb: G_bb by the parity of the bond
    GAS cplx *G = gp(c.G);
    cplx *gm1 = c.avec, *gmat1 = c.bvec, *gm2t = c.W1;
    double *csr = c.rwork, *csi = c.rwork + n;
    const double *u = c.unifv + (size_t)2 * nl * (c.nn_nf * slice + p);
    int32_t *cf = c.confv + ((size_t)(nt - 1) * c.nn_nf + nf) * n;
    cplx phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    for (int r = tid; r < n; r += blockDim.x) { csr[r] = 1.0; csi[r] = 0.0; }
    __syncthreads();
    double nacc = 0.0;
    int ik = 0;
    for (int q = 0; q < nl; q++) {
        const int ist = nf * nl + q;
        const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
        for (int m = tid; m < ik; m += blockDim.x) {
            const size_t pm = (size_t)pvec[m];
            cv[2 * m] = gld(G + pm * n + i1);
            cv[2 * m + 1] = gld(G + pm * n + i2);
            bv[m] = gld(G + (size_t)i1 * n + pm);
            bv[m + K] = gld(G + (size_t)i2 * n + pm);
        }
        // G_bb goes to the buffer of this bond's parity: after a rejected bond, a thread may still
        // read the previous bond's values while this one is written
        cplx *s_g = s_gb + 4 * (q & 1);
        if (tid == blockDim.x - 1) {
            s_g[0] = gld(G + (size_t)i1 * n + i1);
            s_g[1] = gld(G + (size_t)i2 * n + i1);
            s_g[2] = gld(G + (size_t)i1 * n + i2);
            s_g[3] = gld(G + (size_t)i2 * n + i2);
        }
        __syncthreads();
        // v1 = G(b, P) Gamma^-1 (lane i: column i) and v = v1 G(P, b), summed by wave 0 (ik <= 64)
        if (tid < 64) {
            cplx t11 = make_double2(0.0, 0.0), t12 = t11, t21 = t11, t22 = t11;
            if (lane < ik) {
                cplx a1 = make_double2(0.0, 0.0), a2 = a1;
                for (int m = 0; m < ik; m++) {
                    const cplx g = gi[m + lane * KL];
                    a1 = cadd(a1, cmul(cv[2 * m], g));
                    a2 = cadd(a2, cmul(cv[2 * m + 1], g));
                }
                v1[2 * lane] = a1;
                v1[2 * lane + 1] = a2;
                t11 = cmul(a1, bv[lane]);
                t12 = cmul(a1, bv[lane + K]);
                t21 = cmul(a2, bv[lane]);
                t22 = cmul(a2, bv[lane + K]);
            }
            t11.x = wave_sum(t11.x); t11.y = wave_sum(t11.y);
            t12.x = wave_sum(t12.x); t12.y = wave_sum(t12.y);
            t21.x = wave_sum(t21.x); t21.y = wave_sum(t21.y);
            t22.x = wave_sum(t22.x); t22.y = wave_sum(t22.y);
            if (lane == 0) {
                s_v[0] = t11; s_v[1] = t12; s_v[2] = t21; s_v[3] = t22;
            }
        }
        __syncthreads();
        const int is = cf[q] - 1;
        const int iflip = (int)ceil(u[2 * q]);
        const int isp = (is + iflip) & 1;
        const cplx dp = c.vdl[2 * is], dm = c.vdl[2 * is + 1];
        // the 2x2 block of the current G
        const cplx g11 = cadd(s_g[0], s_v[0]), g12 = cadd(s_g[1], s_v[1]);
        const cplx g21 = cadd(s_g[2], s_v[2]), g22 = cadd(s_g[3], s_v[3]);
        const cplx m11 = cadd(make_double2(1.0, 0.0), cmul(make_double2(1.0 - g11.x, -g11.y), dp));
        const cplx m12 = cmul(make_double2(-g12.x, -g12.y), dm);
        const cplx m21 = cmul(make_double2(-g21.x, -g21.y), dp);
        const cplx m22 = cadd(make_double2(1.0, 0.0), cmul(make_double2(1.0 - g22.x, -g22.y), dm));
        const cplx ratio = csub(cmul(m11, m22), cmul(m12, m21));
        const double rr = cmul(ratio, phase).x / phase.x;
        if (fabs(rr) > u[2 * q + 1]) {
            nacc += 1.0;
            phase = cmul(phase, cscale(ratio, 1.0 / hypot(ratio.x, ratio.y)));
            // v3 = v5^-1
            const cplx w11 = csub(cadd(make_double2(1.0, 0.0), cinv(dp)), g11);
            const cplx w12 = make_double2(-g12.x, -g12.y), w21 = make_double2(-g21.x, -g21.y);
            const cplx w22 = csub(cadd(make_double2(1.0, 0.0), cinv(dm)), g22);
            const cplx idet = cinv(csub(cmul(w11, w22), cmul(w12, w21)));
            const cplx v3_11 = cmul(w22, idet), v3_12 = cmul(make_double2(-w12.x, -w12.y), idet);
            const cplx v3_21 = cmul(make_double2(-w21.x, -w21.y), idet), v3_22 = cmul(w11, idet);
            for (int i = tid; i < ik; i += blockDim.x) {
                cplx a1 = make_double2(0.0, 0.0), a2 = a1;
                for (int m = 0; m < ik; m++) {
                    const cplx g = gi[i + m * KL];
                    a1 = cadd(a1, cmul(g, bv[m]));
                    a2 = cadd(a2, cmul(g, bv[m + K]));
                }
                v4[i] = a1;
                v4[i + K] = a2;
            }
            __syncthreads();
            // old block += v4 (v3 v1); new rows v3 v1; new columns v4 v3; corner v3
            for (int e = tid; e < ik * ik; e += blockDim.x) {
                const int i = e % ik, m = e / ik;
                const cplx n1 = cadd(cmul(v3_11, v1[2 * m]), cmul(v3_12, v1[2 * m + 1]));
                const cplx n2 = cadd(cmul(v3_21, v1[2 * m]), cmul(v3_22, v1[2 * m + 1]));
                gi[i + m * KL] = cadd(gi[i + m * KL], cadd(cmul(v4[i], n1), cmul(v4[i + K], n2)));
            }
            for (int i = tid; i < ik; i += blockDim.x) {
                gi[ik + i * KL] = cadd(cmul(v3_11, v1[2 * i]), cmul(v3_12, v1[2 * i + 1]));
                gi[ik + 1 + i * KL] = cadd(cmul(v3_21, v1[2 * i]), cmul(v3_22, v1[2 * i + 1]));
                gi[i + ik * KL] = cadd(cmul(v4[i], v3_11), cmul(v4[i + K], v3_21));
                gi[i + (ik + 1) * KL] = cadd(cmul(v4[i], v3_12), cmul(v4[i + K], v3_22));
            }
            if (tid == 0) {
                gi[ik + ik * KL] = v3_11;
                gi[ik + 1 + ik * KL] = v3_21;
                gi[ik + (ik + 1) * KL] = v3_12;
                gi[ik + 1 + (ik + 1) * KL] = v3_22;
                pvec[ik] = i1;
                pvec[ik + 1] = i2;
                cf[q] = isp + 1;
                const cplx f1 = cinv(cadd(make_double2(1.0, 0.0), dp)), f2 = cinv(cadd(make_double2(1.0, 0.0), dm));
                const cplx s1 = cmul(make_double2(csr[i1], csi[i1]), f1), s2 = cmul(make_double2(csr[i2], csi[i2]), f2);
                csr[i1] = s1.x; csi[i1] = s1.y;
                csr[i2] = s2.x; csi[i2] = s2.y;
            }
            for (int r = tid; r < n; r += blockDim.x) {
                gm1[(size_t)r + (size_t)ik * n] = gld(G + (size_t)i1 * n + r);
                gm1[(size_t)r + (size_t)(ik + 1) * n] = gld(G + (size_t)i2 * n + r);
            }
            ik += 2;
            __threadfence_block();
            __syncthreads();
        }
        if (ik > 0 && (ik >= K || q == nl - 1)) {
            // Gmat1 = G(:, P) Gamma^-1 (Gamma^-1 staged through W2 at leading dimension ik: the
            // bonds of a family are disjoint, so ik <= n); gm2t(j, i) = G(p_i, j)
            cplx *gig = c.W2;
            for (int e = tid; e < ik * ik; e += blockDim.x) {
                const int i = e % ik, j = e / ik;
                gig[i + (size_t)j * ik] = gi[i + j * KL];
            }
            for (int r = tid; r < n; r += blockDim.x)
                for (int i = 0; i < ik; i++) gm2t[(size_t)r + (size_t)i * n] = gld(G + (size_t)r * n + pvec[i]);
            __threadfence_block();
            __syncthreads();
            for (int tm = 0; tm < n; tm += GT)
                zgemm_tile<0, 0>(gm1, n, gig, ik, gmat1, n, n, ik, ik, tm, 0, 1.0, 0.0, 0.0, 0.0);
            __threadfence_block();
            __syncthreads();
            wg_rank_k(c.G, n, gmat1, gm2t, ik, csr, csi);
            __threadfence_block();
            __syncthreads();
            for (int r = tid; r < n; r += blockDim.x) { csr[r] = 1.0; csi[r] = 0.0; }
            ik = 0;
            __syncthreads();
        }
    }
    if (tid == 0) {
        c.sc[SC_PHASE] = phase.x;
        c.sc[SC_PHASE + 1] = phase.y;
        c.vacc[0] += nacc;
        c.vacc[1] += (double)nl;
    }
}

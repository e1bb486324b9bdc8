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

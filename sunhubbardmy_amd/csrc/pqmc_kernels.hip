// pqmc_kernels.hip -- device side of the projective sweep (src/dqmc_proj_core, vconf of
// src/dqmc_config_v): rectangular propagators UR (N x ne) and UL (ne x N), their inverse
// overlap ULRINV = (UL UR)^-1 (ne x ne), the bond-field operators, the submatrix-LR update and
// the QR / LU steps of the stabilization.  Every kernel is batched over chains that share N, ne
// and the stock size; blockIdx.z (elementwise, GEMM) or blockIdx.x (one workgroup per chain)
// selects the chain.
#include "sunhip_kernels.h"

// scalar slots of the projective chain beyond the shared layout (SC_PHASE, SC_LOGW, SC_ZLD,
// SC_INFO, SC_OBS ... keep their meaning)
// (the projective sweep has no saved chain point, so SC_LDQ_SAVE's first slot holds the angle
// sum of the on-site field's phases)
enum { PS_RLOG = 30, PS_ISGN = 31, PS_UANG = SC_LDQ_SAVE };

enum {
    P_UR = 0, P_UL, P_ULRINV, P_ULR, P_T1, P_T2, P_ULO, P_URO, P_INV2, P_G, P_G2, P_KT, P_KTM1,
    P_URP, P_ULP, P_RROWS, P_FELMS, P_GINV, P_VMAT, P_UMAT, P_VTMP, P_UTMP, P_UST, P_PROJ
};

struct PChain {
    int n, ne, nn_nf, nn_lf, kmode, ns, ltrot, nst, nflr, lq, pb, trace_cap;
    cplx *UR, *UL, *ULRINV, *ULR, *T1, *T2, *ULo, *URo, *INV2, *G, *G2, *Ust, *proj;
    cplx *urP, *ulP, *Rrows, *Felms, *Ginv, *Vmat, *Umat, *Vtmp, *Utmp, *Dvec;
    cplx *tau, *tblk, *wk, *trace;
    int32_t *piv, *conf, *xvec, *svec, *ikd;
    double *unif, *sc, *logwDV, *rwork;
    uint32_t *rng;
    const int32_t *bonds;
    const cplx *kt, *ktm1;
    const double *stag;
    cplx vt[8];   // (bp, bm, bpi, bmi) for field value 1 then 2
    cplx vd[4];   // (dp, dm) for field value 1 then 2
    // the on-site field (src/dqmc_config_u), when lwrapu: conf_u [ltrot][n], values 1..ulcomp
    int lwrapu, lwrapv, ulcomp, uphase_from_conf;
    int32_t *uconf;
    cplx ub[4], ubi[4], ud[12], uratio[12];   // bmat, bmat_inv, delta_bmat(iflip, is), phase_ratio
    double ug[4], upa[4];                      // gaml, arg of phase(is)
    // the G-based updates of the on-site field: the slice's field before the update
    int32_t *uold;
};

// uniforms of one slice in a pass: 2 per site of the on-site field, 2 per bond of every family
__host__ __device__ __forceinline__ size_t p_slice_draws(const PChain &c)
{
    return (c.lwrapu ? (size_t)2 * c.n : 0) + (c.lwrapv ? (size_t)2 * c.nn_lf * c.nn_nf : 0);
}

__device__ __forceinline__ cplx *pmat(const PChain &c, int sel, int pt)
{
    const size_t nne = (size_t)c.n * c.ne;
    switch (sel) {
    case P_UR: return c.UR;
    case P_UL: return c.UL;
    case P_ULRINV: return c.ULRINV;
    case P_ULR: return c.ULR;
    case P_T1: return c.T1;
    case P_T2: return c.T2;
    case P_ULO: return c.ULo;
    case P_URO: return c.URo;
    case P_INV2: return c.INV2;
    case P_G: return c.G;
    case P_G2: return c.G2;
    case P_KT: return (cplx *)c.kt;
    case P_KTM1: return (cplx *)c.ktm1;
    case P_URP: return c.urP;
    case P_ULP: return c.ulP;
    case P_RROWS: return c.Rrows;
    case P_FELMS: return c.Felms;
    case P_GINV: return c.Ginv;
    case P_VMAT: return c.Vmat;
    case P_UMAT: return c.Umat;
    case P_VTMP: return c.Vtmp;
    case P_UTMP: return c.Utmp;
    case P_UST: return c.Ust + nne * pt;
    case P_PROJ: return c.proj;
    }
    return nullptr;
}

// ---- batched ZGEMM with explicit shapes: C = alpha A B + beta C --------------------------
struct PGemm {
    int sa, pa, sb, pb, sc, pc;
    int m, n, k, lda, ldb, ldc;
    double ar, ai, br, bi;
};

extern "C" __global__ void __launch_bounds__(256)
kp_gemm(const PChain *chains, PGemm g)
{
    const PChain &c = chains[blockIdx.z];
    const int tm = blockIdx.x * GT, tn = blockIdx.y * GT;
    if (tm >= g.m || tn >= g.n) return;
    zgemm_tile<0, 0>(pmat(c, g.sa, g.pa), g.lda, pmat(c, g.sb, g.pb), g.ldb, pmat(c, g.sc, g.pc), g.ldc, g.m, g.n,
                     g.k, tm, tn, g.ar, g.ai, g.br, g.bi);
}

// ---- copies -------------------------------------------------------------------------------
// dst[e] = src[e] (or conj) for e < count
extern "C" __global__ void __launch_bounds__(256)
kp_copy(const PChain *chains, int dsel, int dpt, int ssel, int spt, int count, int conj)
{
    const PChain &c = chains[blockIdx.z];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    cplx v = pmat(c, ssel, spt)[e];
    if (conj) v.y = -v.y;
    pmat(c, dsel, dpt)[e] = v;
}

// dst (cols x rows, ld cols) = src^T (or src^H) of src (rows x cols, ld rows); 32 x 32 tiles
extern "C" __global__ void __launch_bounds__(256)
kp_transpose(const PChain *chains, int dsel, int dpt, int ssel, int spt, int rows, int cols, int conj)
{
    __shared__ cplx t[32][33];
    const PChain &c = chains[blockIdx.z];
    const int tr = (rows + 31) / 32;
    const int r0 = (blockIdx.x % tr) * 32, c0 = (blockIdx.x / tr) * 32;
    if (c0 >= cols) return;
    const cplx *S = pmat(c, ssel, spt);
    cplx *D = pmat(c, dsel, dpt);
    for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
        int i = e % 32, j = e / 32;
        if (r0 + i < rows && c0 + j < cols) t[j][i] = S[(size_t)(r0 + i) + (size_t)(c0 + j) * rows];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
        int j = e % 32, i = e / 32;   // D(c0+j, r0+i) = S(r0+i, c0+j)
        if (r0 + i < rows && c0 + j < cols) {
            cplx v = t[j][i];
            if (conj) v.y = -v.y;
            D[(size_t)(c0 + j) + (size_t)(r0 + i) * cols] = v;
        }
    }
}

// M (n x n) += diag(1)
extern "C" __global__ void __launch_bounds__(256)
kp_add_identity(const PChain *chains, int sel)
{
    const PChain &c = chains[blockIdx.z];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    cplx *M = pmat(c, sel, 0);
    M[(size_t)i + (size_t)i * c.n].x += 1.0;
}

// ---- propagation: kinetic families and bond-field operators -------------------------------
// One launch applies a short program of operators to UR from the left (blockIdx.y = 0, CPB
// columns of UR in LDS) and to UL from the right (blockIdx.y = 1, RPB rows of UL in LDS).
// op kinds: 0 kinetic families 0..nn_nf-1, 1 kinetic families nn_nf-1..0 (checkerboard
// tables kselL / kselR), 2 bond field of family arg with modes modeL / modeR:
//   left  0: rows <- diag(bp,bm) ut rows    (dqmc_config_v/dqmc_left_forward_prop.f90, nflag 2)
//   left  1: rows <- u rows                 (left_forward nflag 1, left_backward nflag 2)
//   left  2: rows <- u diag(bpi,bmi) rows   (dqmc_left_backward_prop.f90, nflag 1)
//   right 0: cols <- cols u diag(bpi,bmi)   (dqmc_right_backward_prop.f90, nflag 2)
//   right 1: cols <- cols u                 (right_backward nflag 1, right_forward nflag 2)
//   right 2: cols <- cols diag(bp,bm) ut    (dqmc_right_forward_prop.f90, nflag 1)
// u = ut = [[1, 1], [1, -1]] / sqrt(2) (src/dqmc_config_v.f90:19-20).
// kind 3: the on-site field, UR rows / UL columns times bmat (mode 0) or bmat_inv (mode 1) of
// the site's value (src/dqmc_config_u/dqmc_left_forward_prop.f90:23-32 and its three siblings).
#define POPS_MAX 12
struct POps {
    int nops, nt, kselL, kselR, sides;   // sides: 1 = UR, 2 = UL
    int kind[POPS_MAX], arg[POPS_MAX], modeL[POPS_MAX], modeR[POPS_MAX];
};

#define HADAMARD 0.70710678118654752440

// the eight bond-field factors in registers (uniform loads); t(is, k) selects without memory access
struct BondTab {
    cplx v[8];
    __device__ BondTab(const PChain &c)
    {
        for (int k = 0; k < 8; k++) v[k] = c.vt[k];
    }
    __device__ __forceinline__ cplx operator()(int is, int k) const { return is == 2 ? v[4 + k] : v[k]; }
};

__device__ __forceinline__ void v_left_pair(const BondTab &tb, int is, int mode, cplx &x, cplx &y)
{
    const cplx s = make_double2(HADAMARD, 0.0), ms = make_double2(-HADAMARD, 0.0);
    const cplx t[4] = {tb(is, 0), tb(is, 1), tb(is, 2), tb(is, 3)};
    if (mode == 2) { x = cmul(t[2], x); y = cmul(t[3], y); }
    cplx v1 = cadd(cmul(s, x), cmul(s, y));
    cplx v2 = cadd(cmul(s, x), cmul(ms, y));
    if (mode == 0) { v1 = cmul(t[0], v1); v2 = cmul(t[1], v2); }
    x = v1;
    y = v2;
}

__device__ __forceinline__ void v_right_pair(const BondTab &tb, int is, int mode, cplx &x, cplx &y)
{
    const cplx s = make_double2(HADAMARD, 0.0), ms = make_double2(-HADAMARD, 0.0);
    const cplx t[4] = {tb(is, 0), tb(is, 1), tb(is, 2), tb(is, 3)};
    if (mode == 2) { x = cmul(x, t[0]); y = cmul(y, t[1]); }
    cplx v1 = cadd(cmul(x, s), cmul(y, s));
    cplx v2 = cadd(cmul(x, s), cmul(y, ms));
    if (mode == 0) { v1 = cmul(t[2], v1); v2 = cmul(t[3], v2); }
    x = v1;
    y = v2;
}

extern "C" __global__ void __launch_bounds__(256)
kp_prop(const PChain *chains, POps ops)
{
    extern __shared__ cplx lds[];
    const PChain &c = chains[blockIdx.z];
    const int n = c.n, ne = c.ne, nl = c.nn_lf;
    const int groups = max(1, (int)blockDim.x / nl);
    const int bt = threadIdx.x % nl, g = threadIdx.x / nl;
    if (!(ops.sides & (1 << blockIdx.y))) return;
    const BondTab tb(c);
    if (blockIdx.y == 0) {
        // UR from the left, CPB columns
        const int c0 = blockIdx.x * CPB;
        if (c0 >= ne) return;
        const int ncol = min(CPB, ne - c0);
        cplx *M = c.UR;
        for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) lds[e] = M[(size_t)(e % n) + (size_t)(c0 + e / n) * n];
        __syncthreads();
        for (int o = 0; o < ops.nops; o++) {
            if (ops.kind[o] < 2) {
                if (c.kmode == 0) cb_left_fams(c, lds, ncol, ops.kselL ? c.ktm1 : c.kt, ops.kind[o] == 1);
                continue;
            }
            if (ops.kind[o] == 3) {
                const int32_t *uc = c.uconf + (size_t)(ops.nt - 1) * n;
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    const int is = uc[i] - 1;
                    const cplx fct = ops.modeL[o] ? c.ubi[is] : c.ub[is];
                    for (int cc = 0; cc < ncol; cc++) lds[cc * n + i] = cmul(fct, lds[cc * n + i]);
                }
                __syncthreads();
                continue;
            }
            const int f = ops.arg[o], mode = ops.modeL[o];
            const int32_t *cf = c.conf + ((size_t)(ops.nt - 1) * c.nn_nf + f) * n;
            for (int bb = bt; bb < nl; bb += blockDim.x) {
                if (g >= groups) break;
                const int ist = f * nl + bb;
                const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
                const int is = cf[bb];
                for (int cc = g; cc < ncol; cc += groups) {
                    cplx x = lds[cc * n + i1], y = lds[cc * n + i2];
                    v_left_pair(tb, is, mode, x, y);
                    lds[cc * n + i1] = x;
                    lds[cc * n + i2] = y;
                }
            }
            __syncthreads();
        }
        for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) M[(size_t)(e % n) + (size_t)(c0 + e / n) * n] = lds[e];
    } else {
        // UL from the right, RPB rows; lds[rr * ldr + col]
        const int r0 = blockIdx.x * RPB;
        if (r0 >= ne) return;
        const int nrow = min(RPB, ne - r0), ldr = cb_ldr(n);
        cplx *M = c.UL;
        for (int e = threadIdx.x; e < n * RPB; e += blockDim.x) {
            int rr = e % RPB, col = e / RPB;
            lds[rr * ldr + col] = (rr < nrow) ? M[(size_t)(r0 + rr) + (size_t)col * ne] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        for (int o = 0; o < ops.nops; o++) {
            if (ops.kind[o] < 2) {
                if (c.kmode == 0) cb_right_fams(c, lds, nrow, ops.kselR ? c.ktm1 : c.kt, ops.kind[o] == 1);
                continue;
            }
            if (ops.kind[o] == 3) {
                const int32_t *uc = c.uconf + (size_t)(ops.nt - 1) * n;
                for (int e = threadIdx.x; e < n * RPB; e += blockDim.x) {
                    const int rr = e / n, col = e % n;
                    const int is = uc[col] - 1;
                    const cplx fct = ops.modeR[o] ? c.ubi[is] : c.ub[is];
                    lds[rr * ldr + col] = cmul(lds[rr * ldr + col], fct);
                }
                __syncthreads();
                continue;
            }
            const int f = ops.arg[o], mode = ops.modeR[o];
            const int32_t *cf = c.conf + ((size_t)(ops.nt - 1) * c.nn_nf + f) * n;
            for (int bb = bt; bb < nl; bb += blockDim.x) {
                if (g >= groups) break;
                const int ist = f * nl + bb;
                const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
                const int is = cf[bb];
                for (int rr = g; rr < nrow; rr += groups) {
                    cplx x = lds[rr * ldr + i1], y = lds[rr * ldr + i2];
                    v_right_pair(tb, is, mode, x, y);
                    lds[rr * ldr + i1] = x;
                    lds[rr * ldr + i2] = y;
                }
            }
            __syncthreads();
        }
        for (int e = threadIdx.x; e < n * RPB; e += blockDim.x) {
            int rr = e % RPB, col = e / RPB;
            if (rr < nrow) M[(size_t)(r0 + rr) + (size_t)col * ne] = lds[rr * ldr + col];
        }
    }
}

// ---- submatrix-LR update (src/dqmc_config_v/dqmc_proj_update_submatrix_LR.f90) -----------
// stock start: the stock's rows of UR (urP, nup x ne, ld ns) and columns of UL (ulP, ne x
// nup, ld ne) in bond order (i1, i2 per bond); Gamma^-1, Vmat, Umat and the count cleared
// nf < 0: the on-site field's stock, the consecutive sites ib0 .. ib0 + nup - 1
extern "C" __global__ void __launch_bounds__(256)
kp_stock_gather(const PChain *chains, int nf, int ib0, int nup)
{
    const PChain &c = chains[blockIdx.z];
    const int ns = c.ns, ne = c.ne, n = c.n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    auto site = [&](int s) {
        if (nf < 0) return ib0 + s;
        const int ist = nf * c.nn_lf + ib0 + s / 2;
        return (int)c.bonds[2 * ist + (s & 1)];
    };
    if (e < ns * ne) {
        const int s = e % ns, l = e / ns;       // urP(s, l)
        cplx v = make_double2(0.0, 0.0);
        if (s < nup) v = c.UR[(size_t)site(s) + (size_t)l * n];
        c.urP[e] = v;
        c.Vmat[e] = make_double2(0.0, 0.0);
    }
    if (e < ne * ns) {
        const int l = e % ne, s = e / ne;       // ulP(l, s)
        cplx v = make_double2(0.0, 0.0);
        if (s < nup) v = c.UL[(size_t)l + (size_t)site(s) * ne];
        c.ulP[e] = v;
        c.Umat[e] = make_double2(0.0, 0.0);
    }
    if (e < ns * ns) c.Ginv[e] = make_double2(0.0, 0.0);
    if (e == 0) c.ikd[0] = 0;
}

// the stock's bonds in order, one workgroup per chain: 2x2 ratio det(I + Vcal D) with
// Vcal = PFP_kk - PFP_kik Gamma^-1 PFP_ikk (:146-175), Metropolis on Re(ratio phase)/Re(phase),
// and on acceptance the bordering of Gamma^-1 by one 2x2 block (:176-222).  Felms and Gamma^-1
// live in LDS (2 ns^2 complex).  The stock's field, acceptance uniforms and bond sites are staged
// in LDS and the field, xvec and Dvec written back at the end, so that the bond loop itself
// touches no global memory (a global access there would put a memory round trip on every
// bond's critical path).
extern "C" __global__ void __launch_bounds__(256)
kp_stock_bonds(const PChain *chains, int nt, int nf, int ib0, int nb, int uoff)
{
    extern __shared__ cplx bsh[];
    const PChain &c = chains[blockIdx.x];
    const int ns = c.ns, ne = c.ne;
    cplx *F = bsh;                    // Felms, ld ns
    cplx *Gi = bsh + ns * ns;         // Gamma^-1, ld ns
    __shared__ cplx ikk[64][2], kik[2][64], n21[2][64], vuk[4], gkk[4];
    __shared__ int svec[64];
    __shared__ int s_acc, s_ik;
    __shared__ double s_accm;
    __shared__ cplx s_phase;
    __shared__ int cfl[32], bsite[64], xv[64];   // field and sites of the stock's nb <= 32 bonds
    __shared__ double ua[32];                    // their acceptance uniforms
    __shared__ cplx vdl[4], dv[64];
    int32_t *cf = c.conf + ((size_t)(nt - 1) * c.nn_nf + nf) * c.n + ib0;
    const double *u = c.unif + uoff;
    for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
        F[e] = c.Felms[e];
        Gi[e] = make_double2(0.0, 0.0);
    }
    if (threadIdx.x < nb) {
        const int q = threadIdx.x, ist = nf * c.nn_lf + ib0 + q;
        cfl[q] = cf[q];
        ua[q] = u[2 * q + 1];
        bsite[2 * q] = c.bonds[2 * ist];
        bsite[2 * q + 1] = c.bonds[2 * ist + 1];
    }
    if (threadIdx.x < 4) vdl[threadIdx.x] = c.vd[threadIdx.x];
    if (threadIdx.x == 0) {
        s_ik = 0;
        s_accm = 0.0;
        s_phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    }
    __syncthreads();
    for (int q = 0; q < nb; q++) {
        const int ik = s_ik;
        const int s1 = 2 * q, s2 = s1 + 1;
        if (ik > 0) {
            // ikk = Gamma^-1 PFP_ikk (ik x 2): two threads per entry, each half of the sum
            {
                const int o = threadIdx.x >> 1, h = threadIdx.x & 1;
                const int m = o % ik, col = o / ik;
                const int sc_ = col ? s2 : s1;
                double ar = 0.0, ai = 0.0;
                if (o < 2 * ik) {
                    // four terms' reads issued together, the sum in the order p = h, h + 2, ...
                    int p = h;
                    for (; p + 6 < ik; p += 8) {
                        const int q0 = svec[p], q1 = svec[p + 2], q2 = svec[p + 4], q3 = svec[p + 6];
                        const cplx g0 = Gi[m + p * ns], g1 = Gi[m + (p + 2) * ns], g2 = Gi[m + (p + 4) * ns],
                                   g3 = Gi[m + (p + 6) * ns];
                        const cplx f0 = F[q0 + sc_ * ns], f1 = F[q1 + sc_ * ns], f2 = F[q2 + sc_ * ns],
                                   f3 = F[q3 + sc_ * ns];
                        ar += g0.x * f0.x - g0.y * f0.y; ai += g0.x * f0.y + g0.y * f0.x;
                        ar += g1.x * f1.x - g1.y * f1.y; ai += g1.x * f1.y + g1.y * f1.x;
                        ar += g2.x * f2.x - g2.y * f2.y; ai += g2.x * f2.y + g2.y * f2.x;
                        ar += g3.x * f3.x - g3.y * f3.y; ai += g3.x * f3.y + g3.y * f3.x;
                    }
                    for (; p < ik; p += 2) {
                        const cplx g = Gi[m + p * ns], f = F[svec[p] + sc_ * ns];
                        ar += g.x * f.x - g.y * f.y;
                        ai += g.x * f.y + g.y * f.x;
                    }
                }
                ar += dpp_d<0xB1>(ar);   // the partner lane (lane ^ 1)
                ai += dpp_d<0xB1>(ai);
                if (o < 2 * ik && h == 0) ikk[m][col] = make_double2(ar, ai);
            }
            __syncthreads();
            // vuk = PFP_kik ikk (2 x 2): one wave per entry, one lane per term
            {
                const int w = threadIdx.x >> 6, m = threadIdx.x & 63;
                if (w < 4) {
                    const int r = w & 1, col = w >> 1;
                    const int sr = r ? s2 : s1;
                    double ar = 0.0, ai = 0.0;
                    if (m < ik) {
                        const cplx f = F[sr + svec[m] * ns], x = ikk[m][col];
                        ar = f.x * x.x - f.y * x.y;
                        ai = f.x * x.y + f.y * x.x;
                    }
                    ar = wave_sum_dpp(ar);
                    ai = wave_sum_dpp(ai);
                    if (m == 0) vuk[r + 2 * col] = make_double2(ar, ai);
                }
            }
        } else if (threadIdx.x < 4) {
            vuk[threadIdx.x] = make_double2(0.0, 0.0);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int is = cfl[q];
            // iflip = ceiling(u * (lcomp - 1)) = 1 for the two-valued bond field
            const cplx dp = vdl[2 * (is - 1)], dm = vdl[2 * (is - 1) + 1];
            cplx v11 = csub(F[s1 + s1 * ns], vuk[0]), v21 = csub(F[s2 + s1 * ns], vuk[1]);
            cplx v12 = csub(F[s1 + s2 * ns], vuk[2]), v22 = csub(F[s2 + s2 * ns], vuk[3]);
            v11 = cadd(make_double2(1.0, 0.0), cmul(v11, dp));
            v21 = cmul(v21, dp);
            v12 = cmul(v12, dm);
            v22 = cadd(make_double2(1.0, 0.0), cmul(v22, dm));
            const cplx ratio1 = csub(cmul(v11, v22), cmul(v12, v21));
            const cplx ph = s_phase;
            const double ratio_re = cmul(ratio1, ph).x / ph.x;
            const double ratio_abs = ratio_re < 0.0 ? -ratio_re : ratio_re;
            const int acc = ratio_abs > ua[q];
            s_acc = acc;
            if (acc) {
                s_accm += 1.0;
                const double wgt = sqrt(cmul(ratio1, cconj(ratio1)).x);
                s_phase = cscale(cmul(ph, ratio1), 1.0 / wgt);
                gkk[0] = cdiv(cmul(v22, dp), ratio1);
                gkk[2] = cdiv(cmul(make_double2(-v12.x, -v12.y), dp), ratio1);
                gkk[1] = cdiv(cmul(make_double2(-v21.x, -v21.y), dm), ratio1);
                gkk[3] = cdiv(cmul(v11, dm), ratio1);
                cfl[q] = 3 - is;
                xv[ik] = bsite[2 * q];
                xv[ik + 1] = bsite[2 * q + 1];
                dv[ik] = dp;
                dv[ik + 1] = dm;
            }
        }
        __syncthreads();
        if (!s_acc) continue;
        if (ik > 0) {
            // kik = PFP_kik Gamma^-1 (2 x ik): two threads per entry
            {
                const int o = threadIdx.x >> 1, h = threadIdx.x & 1;
                const int r = o % 2, p = o / 2;
                const int sr = r ? s2 : s1;
                double ar = 0.0, ai = 0.0;
                if (o < 2 * ik) {
                    int m = h;
                    for (; m + 6 < ik; m += 8) {
                        const int q0 = svec[m], q1 = svec[m + 2], q2 = svec[m + 4], q3 = svec[m + 6];
                        const cplx f0 = F[sr + q0 * ns], f1 = F[sr + q1 * ns], f2 = F[sr + q2 * ns], f3 = F[sr + q3 * ns];
                        const cplx g0 = Gi[m + p * ns], g1 = Gi[(m + 2) + p * ns], g2 = Gi[(m + 4) + p * ns],
                                   g3 = Gi[(m + 6) + p * ns];
                        ar += f0.x * g0.x - f0.y * g0.y; ai += f0.x * g0.y + f0.y * g0.x;
                        ar += f1.x * g1.x - f1.y * g1.y; ai += f1.x * g1.y + f1.y * g1.x;
                        ar += f2.x * g2.x - f2.y * g2.y; ai += f2.x * g2.y + f2.y * g2.x;
                        ar += f3.x * g3.x - f3.y * g3.y; ai += f3.x * g3.y + f3.y * g3.x;
                    }
                    for (; m < ik; m += 2) {
                        const cplx f = F[sr + svec[m] * ns], g = Gi[m + p * ns];
                        ar += f.x * g.x - f.y * g.y;
                        ai += f.x * g.y + f.y * g.x;
                    }
                }
                ar += dpp_d<0xB1>(ar);
                ai += dpp_d<0xB1>(ai);
                if (o < 2 * ik && h == 0) kik[r][p] = make_double2(ar, ai);
            }
            __syncthreads();
            // 21 block = -gkk kik
            for (int e = threadIdx.x; e < 2 * ik; e += blockDim.x) {
                const int r = e % 2, p = e / 2;
                cplx a = cadd(cmul(gkk[r], kik[0][p]), cmul(gkk[r + 2], kik[1][p]));
                n21[r][p] = make_double2(-a.x, -a.y);
            }
            __syncthreads();
            // 11 block -= ikk n21; 12 block = -ikk gkk; 21 block stored
            for (int e = threadIdx.x; e < ik * ik; e += blockDim.x) {
                const int m = e % ik, p = e / ik;
                cplx a = cadd(cmul(ikk[m][0], n21[0][p]), cmul(ikk[m][1], n21[1][p]));
                Gi[m + p * ns] = csub(Gi[m + p * ns], a);
            }
            for (int e = threadIdx.x; e < 2 * ik; e += blockDim.x) {
                const int m = e % ik, col = e / ik;
                cplx a = cadd(cmul(ikk[m][0], gkk[2 * col]), cmul(ikk[m][1], gkk[2 * col + 1]));
                Gi[m + (ik + col) * ns] = make_double2(-a.x, -a.y);
                const int r = e % 2, p = e / 2;
                Gi[(ik + r) + p * ns] = n21[r][p];
            }
        }
        if (threadIdx.x < 4) Gi[(ik + (threadIdx.x & 1)) + (ik + (threadIdx.x >> 1)) * ns] = gkk[threadIdx.x];
        if (threadIdx.x == 0) {
            svec[ik] = s1;
            svec[ik + 1] = s2;
            s_ik = ik + 2;
        }
        __syncthreads();
    }
    // Gamma^-1, the accepted rows of Rrows (Vmat) and columns of ulP (Umat) for the flush; the
    // field, the accepted sites and their deltas
    const int ik = s_ik;
    for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) c.Ginv[e] = Gi[e];
    for (int q = threadIdx.x; q < nb; q += blockDim.x) cf[q] = cfl[q];
    for (int k = threadIdx.x; k < ik; k += blockDim.x) {
        c.xvec[k] = xv[k];
        c.Dvec[k] = dv[k];
    }
    for (int e = threadIdx.x; e < ik * ne; e += blockDim.x) {
        const int m = e % ik, l = e / ik;
        c.Vmat[m + (size_t)l * ns] = c.Rrows[svec[m] + (size_t)l * ns];
        const int l2 = e % ne, m2 = e / ne;
        c.Umat[l2 + (size_t)m2 * ne] = c.ulP[l2 + (size_t)svec[m2] * ne];
    }
    if (threadIdx.x == 0) {
        c.ikd[0] = ik;
        c.sc[SC_PHASE] = s_phase.x;
        c.sc[SC_PHASE + 1] = s_phase.y;
        c.sc[SC_ACC] += s_accm;
        c.sc[SC_TRIED] += nb;
    }
}

// the on-site field's stock of consecutive sites i0 .. i0 + nup - 1, one workgroup per chain
// (src/dqmc_config_u/dqmc_proj_update_submatrix_LR.f90:83-237): per site the 1 x 1 ratio
// 1 + (F_kk - F_kP Gamma^-1 F_Pk) Delta(iflip, is), ratiotot = ratio1^nflr gaml(isp)/gaml(is)
// phase_ratio(iflip, is), the Metropolis test on Re(ratiotot phase)/Re(phase), and on
// acceptance Gamma^-1 bordered by one row and column.  Felms and Gamma^-1 live in LDS, and so do
// the stock's field, draws and HS tables (the field, xvec and Dvec are written back at the end):
// the site loop itself touches no global memory.
extern "C" __global__ void __launch_bounds__(256)
kp_stock_sites(const PChain *chains, int nt, int i0, int nup, size_t uoff)
{
    extern __shared__ cplx bsh[];
    const PChain &c = chains[blockIdx.x];
    const int ns = c.ns, ne = c.ne, lc = c.ulcomp;
    cplx *F = bsh;                    // Felms, ld ns
    cplx *Gi = bsh + ns * ns;         // Gamma^-1, ld ns
    __shared__ cplx ikk[64], kik[64], s_vuk, s_gkk;
    __shared__ int svec[64];
    __shared__ int s_acc, s_ik;
    __shared__ double s_accm;
    __shared__ cplx s_phase;
    __shared__ int cfl[64], xv[64];              // field of the stock's nup <= 64 sites, accepted sites
    __shared__ double u1l[64], u2l[64], ugl[4];  // their two uniforms; gaml
    __shared__ cplx udl[12], url[12], dv[64];    // delta_bmat, phase_ratio; accepted deltas
    int32_t *cf = c.uconf + (size_t)(nt - 1) * c.n + i0;
    const double *u = c.unif + uoff;
    const int nflr = c.nflr;
    for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
        F[e] = c.Felms[e];
        Gi[e] = make_double2(0.0, 0.0);
    }
    if (threadIdx.x < nup) {
        cfl[threadIdx.x] = cf[threadIdx.x];
        u1l[threadIdx.x] = u[2 * threadIdx.x];
        u2l[threadIdx.x] = u[2 * threadIdx.x + 1];
    }
    if (threadIdx.x < 12) {
        udl[threadIdx.x] = c.ud[threadIdx.x];
        url[threadIdx.x] = c.uratio[threadIdx.x];
    }
    if (threadIdx.x < 4) ugl[threadIdx.x] = c.ug[threadIdx.x];
    if (threadIdx.x == 0) {
        s_ik = 0;
        s_accm = 0.0;
        s_phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int q = 0; q < nup; q++) {
        const int ik = s_ik;
        if (ik > 0) {
            // ikk = Gamma^-1 F(P, q): four threads per entry, each a quarter of the sum
            {
                const int m = threadIdx.x >> 2, h = threadIdx.x & 3;
                double ar = 0.0, ai = 0.0;
                if (m < ik) {
                    // four terms' reads issued together, the sum in the order p = h, h + 4, ...
                    int p = h;
                    for (; p + 12 < ik; p += 16) {
                        const int q0 = svec[p], q1 = svec[p + 4], q2 = svec[p + 8], q3 = svec[p + 12];
                        const cplx g0 = Gi[m + p * ns], g1 = Gi[m + (p + 4) * ns], g2 = Gi[m + (p + 8) * ns],
                                   g3 = Gi[m + (p + 12) * ns];
                        const cplx f0 = F[q0 + q * ns], f1 = F[q1 + q * ns], f2 = F[q2 + q * ns], f3 = F[q3 + q * ns];
                        ar += g0.x * f0.x - g0.y * f0.y; ai += g0.x * f0.y + g0.y * f0.x;
                        ar += g1.x * f1.x - g1.y * f1.y; ai += g1.x * f1.y + g1.y * f1.x;
                        ar += g2.x * f2.x - g2.y * f2.y; ai += g2.x * f2.y + g2.y * f2.x;
                        ar += g3.x * f3.x - g3.y * f3.y; ai += g3.x * f3.y + g3.y * f3.x;
                    }
                    for (; p < ik; p += 4) {
                        const cplx g = Gi[m + p * ns], f = F[svec[p] + q * ns];
                        ar += g.x * f.x - g.y * f.y;
                        ai += g.x * f.y + g.y * f.x;
                    }
                }
                ar += dpp_d<0xB1>(ar);   // lane ^ 1, then lane ^ 2 (in-quad DPP)
                ai += dpp_d<0xB1>(ai);
                ar += dpp_d<0x4E>(ar);
                ai += dpp_d<0x4E>(ai);
                if (m < ik && h == 0) ikk[m] = make_double2(ar, ai);
            }
            __syncthreads();
            // vuk = F(q, P) ikk: wave 0, one lane per term
            if (w == 0) {
                double ar = 0.0, ai = 0.0;
                if (lane < ik) {
                    const cplx f = F[q + svec[lane] * ns], x = ikk[lane];
                    ar = f.x * x.x - f.y * x.y;
                    ai = f.x * x.y + f.y * x.x;
                }
                ar = wave_sum_dpp(ar);
                ai = wave_sum_dpp(ai);
                if (lane == 0) s_vuk = make_double2(ar, ai);
            }
        } else if (threadIdx.x == 0) {
            s_vuk = make_double2(0.0, 0.0);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int is = cfl[q];
            const int iflip = (int)ceil(u1l[q] * (lc - 1));
            const int isp = (is + iflip - 1) % lc + 1;
            const int k = (iflip - 1) + (lc - 1) * (is - 1);
            const cplx d = udl[k];
            const cplx ratio1 = cadd(make_double2(1.0, 0.0), cmul(csub(F[q + q * ns], s_vuk), d));
            cplx rt = ratio1;
            for (int f = 1; f < nflr; f++) rt = cmul(rt, ratio1);
            rt = cmul(cscale(rt, ugl[isp - 1] / ugl[is - 1]), url[k]);
            const cplx ph = s_phase;
            const double ratio_re = cmul(rt, ph).x / ph.x;
            const double ratio_abs = ratio_re < 0.0 ? -ratio_re : ratio_re;
            const int acc = ratio_abs > u2l[q];
            s_acc = acc;
            if (acc) {
                s_accm += 1.0;
                const double wgt = sqrt(cmul(rt, cconj(rt)).x);
                s_phase = cscale(cmul(ph, rt), 1.0 / wgt);
                s_gkk = cdiv(d, ratio1);
                cfl[q] = isp;
                xv[ik] = i0 + q;
                dv[ik] = d;
            }
        }
        __syncthreads();
        if (!s_acc) continue;
        const cplx gkk = s_gkk;
        if (ik > 0) {
            // kik = F(q, P) Gamma^-1: four threads per entry
            {
                const int p = threadIdx.x >> 2, h = threadIdx.x & 3;
                double ar = 0.0, ai = 0.0;
                if (p < ik) {
                    int m = h;
                    for (; m + 12 < ik; m += 16) {
                        const int q0 = svec[m], q1 = svec[m + 4], q2 = svec[m + 8], q3 = svec[m + 12];
                        const cplx f0 = F[q + q0 * ns], f1 = F[q + q1 * ns], f2 = F[q + q2 * ns], f3 = F[q + q3 * ns];
                        const cplx g0 = Gi[m + p * ns], g1 = Gi[(m + 4) + p * ns], g2 = Gi[(m + 8) + p * ns],
                                   g3 = Gi[(m + 12) + p * ns];
                        ar += f0.x * g0.x - f0.y * g0.y; ai += f0.x * g0.y + f0.y * g0.x;
                        ar += f1.x * g1.x - f1.y * g1.y; ai += f1.x * g1.y + f1.y * g1.x;
                        ar += f2.x * g2.x - f2.y * g2.y; ai += f2.x * g2.y + f2.y * g2.x;
                        ar += f3.x * g3.x - f3.y * g3.y; ai += f3.x * g3.y + f3.y * g3.x;
                    }
                    for (; m < ik; m += 4) {
                        const cplx f = F[q + svec[m] * ns], g = Gi[m + p * ns];
                        ar += f.x * g.x - f.y * g.y;
                        ai += f.x * g.y + f.y * g.x;
                    }
                }
                ar += dpp_d<0xB1>(ar);
                ai += dpp_d<0xB1>(ai);
                ar += dpp_d<0x4E>(ar);
                ai += dpp_d<0x4E>(ai);
                if (p < ik && h == 0) kik[p] = make_double2(ar, ai);
            }
            __syncthreads();
            // 11 block -= ikk (-gkk kik); 12 block = -ikk gkk; 21 block = -gkk kik
            for (int e = threadIdx.x; e < ik * ik; e += blockDim.x) {
                const int m = e % ik, p = e / ik;
                const cplx n21 = cmul(gkk, kik[p]);          // minus the 21 entry
                Gi[m + p * ns] = cadd(Gi[m + p * ns], cmul(ikk[m], n21));
            }
            for (int e = threadIdx.x; e < ik; e += blockDim.x) {
                const cplx a = cmul(ikk[e], gkk), b = cmul(gkk, kik[e]);
                Gi[e + ik * ns] = make_double2(-a.x, -a.y);
                Gi[ik + e * ns] = make_double2(-b.x, -b.y);
            }
        }
        if (threadIdx.x == 0) {
            Gi[ik + ik * ns] = gkk;
            svec[ik] = q;
            s_ik = ik + 1;
        }
        __syncthreads();
    }
    const int ik = s_ik;
    for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) c.Ginv[e] = Gi[e];
    for (int q = threadIdx.x; q < nup; q += blockDim.x) cf[q] = cfl[q];
    for (int k = threadIdx.x; k < ik; k += blockDim.x) {
        c.xvec[k] = xv[k];
        c.Dvec[k] = dv[k];
    }
    for (int e = threadIdx.x; e < ik * ne; e += blockDim.x) {
        const int m = e % ik, l = e / ik;
        c.Vmat[m + (size_t)l * ns] = c.Rrows[svec[m] + (size_t)l * ns];
        const int l2 = e % ne, m2 = e / ne;
        c.Umat[l2 + (size_t)m2 * ne] = c.ulP[l2 + (size_t)svec[m2] * ne];
    }
    if (threadIdx.x == 0) {
        c.ikd[0] = ik;
        c.sc[SC_PHASE] = s_phase.x;
        c.sc[SC_PHASE + 1] = s_phase.y;
        c.sc[SC_ACC] += s_accm;
        c.sc[SC_TRIED] += nup;
    }
}

// UR(x_m, :) += D_m UR(x_m, :) for the stock's accepted rows (:249-254)
extern "C" __global__ void __launch_bounds__(256)
kp_stock_scale(const PChain *chains)
{
    const PChain &c = chains[blockIdx.z];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int ik = c.ikd[0];
    if (e >= ik * c.ne) return;
    const int m = e % ik, l = e / ik;
    cplx *p = c.UR + (size_t)c.xvec[m] + (size_t)l * c.n;
    *p = cadd(*p, cmul(c.Dvec[m], *p));
}

// ---- stabilization --------------------------------------------------------------------------
// s_zgeQR_pqmc (lib/s_matrix.f90:1822-1996) on the tall matrix asel (N x ne): Householder QR,
// log|det R| and the count of negative diagonal entries of R to PS_RLOG / PS_ISGN, Q formed
// into qsel with its first column negated for an odd count
extern "C" __global__ void __launch_bounds__(BLK_NT)
kp_qr_formq(const PChain *chains, int asel, int apt, int qsel, int qpt)
{
    extern __shared__ cplx bsh[];
    const PChain &c = chains[blockIdx.x];
    const int n = c.n, ne = c.ne;
    cplx *A = pmat(c, asel, apt);
    cplx *Q = pmat(c, qsel, qpt);
    qr_blk_body(A, n, ne, n, c.tau, c.tblk, c.wk, c.pb, bsh);
    __shared__ int s_neg;
    if (threadIdx.x == 0) {
        int neg = 0;
        double rl = 0.0;
        for (int i = 0; i < ne; i++) {
            double r = A[(size_t)i + (size_t)i * n].x;
            if (r < 0.0) { neg++; rl += log(-r); } else rl += log(r);
        }
        s_neg = neg;
        c.sc[PS_RLOG] = rl;
        c.sc[PS_ISGN] = (double)neg;
    }
    __syncthreads();
    formq_blk_body(A, n, ne, n, Q, n, c.tblk, c.wk, c.pb, bsh);
    __syncthreads();
    if (s_neg & 1)
        for (int i = threadIdx.x; i < n; i += blockDim.x) Q[i] = make_double2(-Q[i].x, -Q[i].y);
}

// X = A^-1 for A (ne x ne) by LU with partial pivoting (s_inv_logdet_lu_z); A is overwritten
// by its factors; with slot >= 0 the log det goes to sc[slot], sc[slot+1]; a zero pivot is
// recorded in sc[SC_INFO] (first zero pivot + 1, kept until the host reads it)
extern "C" __global__ void __launch_bounds__(BLK_NT)
kp_lu_inv(const PChain *chains, int asel, int xsel, int slot)
{
    extern __shared__ cplx bsh[];
    const PChain &c = chains[blockIdx.x];
    const int ne = c.ne;
    cplx *A = pmat(c, asel, 0);
    cplx *X = pmat(c, xsel, 0);
    __shared__ cplx s_ld;
    __shared__ double s_inf;
    lu_blk_body(A, ne, ne, c.piv, &s_ld, &s_inf, bsh);
    for (size_t e = threadIdx.x; e < (size_t)ne * ne; e += blockDim.x)
        X[e] = make_double2((e % ne) == (e / ne) ? 1.0 : 0.0, 0.0);
    __syncthreads();
    solve_blk_body(A, ne, ne, c.piv, X, ne, ne, c.T1, bsh);
    if (threadIdx.x == 0) {
        if (slot >= 0) {
            c.sc[slot] = s_ld.x;
            c.sc[slot + 1] = s_ld.y;
        }
        if (s_inf != 0.0 && c.sc[SC_INFO] == 0.0) c.sc[SC_INFO] = s_inf;
    }
}

// weight bookkeeping after a scratch inverse (z = log det ULR in SC_ZLD):
//   op 0: sweep start  logweightf = logwDV(nst) + z, phase = exp(i nflr Im z)
//   op 1: 0b point n   logwDV(n) = logwDV(n-1) + rlog, logw = logwDV(n) [+ logwDV(n+1)] + z
//   op 2: b0 point n   logwDV(n+1) = [logwDV(n+2) +] rlog, logw = logwDV(n+1) [+ logwDV(n)] + z
//   op 3: 0b pass start logwDV(0) = 0
//   op 4: start point n logwDV(n) = logwDV(n-1) + rlog
// (dqmc_proj_sweep_start_0b.f90:16-38, dqmc_proj_sweep_0b.f90:121-148, dqmc_proj_sweep_b0.f90:113-141)
extern "C" __global__ void kp_weights(const PChain *chains, int nchains, int op, int n, int phasediff)
{
    const int ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= nchains) return;
    const PChain &c = chains[ci];
    double *sc = c.sc, *lw = c.logwDV;
    const cplx z = make_double2(sc[SC_ZLD], sc[SC_ZLD + 1]);
    double dv = 0.0;
    if (op == 3) { lw[0] = 0.0; return; }
    if (op == 4) { lw[n] = lw[n - 1] + sc[PS_RLOG]; return; }
    if (op == 0) {
        dv = lw[c.nst];
    } else if (op == 1) {
        lw[n] = lw[n - 1] + sc[PS_RLOG];
        dv = (n == c.nst) ? lw[n] : lw[n] + lw[n + 1];
    } else {
        lw[n + 1] = (n + 1 == c.nst) ? sc[PS_RLOG] : lw[n + 2] + sc[PS_RLOG];
        dv = (n == 0) ? lw[n + 1] : lw[n + 1] + lw[n];
    }
    const cplx logw = make_double2(dv + z.x, z.y);
    // set_phase: exp(i nflr Im logw) times the on-site field's phases (angle sum from kp_uphase)
    const double a = (double)c.nflr * logw.y + ((c.lwrapu && c.uphase_from_conf) ? sc[PS_UANG] : 0.0);
    const cplx ph = make_double2(cos(a), sin(a));
    if (phasediff) {
        const cplx old = make_double2(sc[SC_PHASE], sc[SC_PHASE + 1]);
        const double d = sqrt(cabs2(csub(old, ph)));
        if (d > sc[SC_MAXPH]) sc[SC_MAXPH] = d;
    }
    sc[SC_PHASE] = ph.x;
    sc[SC_PHASE + 1] = ph.y;
    sc[SC_LOGW] = logw.x;
    sc[SC_LOGW + 1] = logw.y;
    if (!(isfinite(logw.x) && isfinite(logw.y))) sc[SC_NONFINITE] = 1.0;
}

// wrap check: max |G2 - G| (s_compare_max_z), one workgroup per chain; the log10 sum and the
// count as the reference's WRAPERROR block; optionally G joins the trace
extern "C" __global__ void __launch_bounds__(256)
kp_wrap_check(const PChain *chains, int compare, int trace)
{
    __shared__ double red[4];
    const PChain &c = chains[blockIdx.x];
    const size_t nn = (size_t)c.n * c.n;
    if (compare) {
        double mx = 0.0;
        for (size_t e = threadIdx.x; e < nn; e += blockDim.x) {
            double d = sqrt(cabs2(csub(c.G2[e], c.G[e])));
            if (!(d <= mx)) mx = d;   // NaN propagates
        }
        for (int o = 32; o > 0; o >>= 1) {
            double v = __shfl_xor(mx, o, 64);
            if (!(v <= mx)) mx = v;
        }
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = red[0];
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) if (!(red[w] <= m)) m = red[w];
            if (!isfinite(m)) c.sc[SC_NONFINITE] = 1.0;
            if (m > c.sc[SC_MAXWRAP]) c.sc[SC_MAXWRAP] = m;
            c.sc[SC_NCHK] += 1.0;
            c.sc[SC_SUMLOG] += log10(m > 0.0 ? m : 1e-300);
        }
    }
    if (trace && c.trace) {
        const int k = (int)c.sc[SC_TRACE_N];
        if (k < c.trace_cap) {
            for (size_t e = threadIdx.x; e < nn; e += blockDim.x) c.trace[(size_t)k * nn + e] = c.G[e];
            __syncthreads();
            if (threadIdx.x == 0) c.sc[SC_TRACE_N] = k + 1;
        }
    }
}

// sum over the on-site field of the arguments of phase(conf) into sc[PS_UANG] (set_phase,
// src/dqmc_proj_core/set_phase.f90:16-22); one workgroup per chain
extern "C" __global__ void __launch_bounds__(256)
kp_uphase(const PChain *chains)
{
    const PChain &c = chains[blockIdx.x];
    __shared__ double red[4];
    double a = 0.0;
    if (c.lwrapu)
        for (size_t k = threadIdx.x; k < (size_t)c.n * c.ltrot; k += blockDim.x) a += c.upa[c.uconf[k] - 1];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) c.sc[PS_UANG] = (red[0] + red[1]) + (red[2] + red[3]);
}

// draws of the projective chain: mode 0 the fresh fields (the on-site field first, N ltrot
// draws conf(i,nt) = ceiling(u lcomp) with i inner, as u0conf%input_conf precedes
// v0conf%input_conf in dqmc_start; then the bond field, N nn_nf ltrot draws, conf(i,nf,nt)
// with nf innermost), mode 1 the pass's uniforms (per slice 2 N for the on-site field and
// 2 nn_lf nn_nf for the bonds), mode 2 count uniforms of chain `which` into out
extern "C" __global__ void __launch_bounds__(64)
kp_draw(const PChain *chains, int mode, int which, int count, double *out)
{
    const int ci = (mode == 2) ? which : (int)blockIdx.x;
    const PChain &c = chains[ci];
    if (mode == 0) {
        const int n = c.n, nf = c.nn_nf;
        if (c.lwrapu) {
            int32_t *uc = c.uconf;
            const int lc = c.ulcomp;
            sfmt_draw_block(c.rng, (size_t)n * c.ltrot, [=](size_t k, double u) { uc[k] = (int32_t)ceil(u * lc); });
        }
        if (c.lwrapv) {
            int32_t *conf = c.conf;
            sfmt_draw_block(c.rng, (size_t)n * nf * c.ltrot, [=](size_t k, double u) {
                const size_t f = k % nf, i = (k / nf) % n, t = k / ((size_t)nf * n);
                conf[(t * nf + f) * n + i] = (int32_t)ceil(u * 2.0);
            });
        }
    } else if (mode == 1) {
        double *un = c.unif;
        sfmt_draw_block(c.rng, p_slice_draws(c) * c.ltrot, [=](size_t k, double u) { un[k] = u; });
    } else {
        sfmt_draw_block(c.rng, (size_t)count, [=](size_t k, double u) { out[k] = u; });
    }
}

extern "C" __global__ void kp_seed(const PChain *chains, int nchains, const int32_t *seeds)
{
    const int ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= nchains) return;
    sfmt_seed_state(chains[ci].rng, seeds[ci]);
}


// ------------------------------------------------------------------------------------------
// The G-based updates of the on-site field (src/dqmc_config_u/dqmc_proj_update_delay.f90,
// dqmc_proj_update_submatrix_G.f90): the finite-temperature delay / submatrix kernels run on
// G = 1 - UR (UL UR)^-1 UL through the chain's view; then the rows of UR whose site flipped are
// multiplied by bmat(new) / bmat(old) = 1 + delta_bmat(iflip, old) (urrecord of the reference).
// ------------------------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(256)
kp_u_save(const PChain *chains, int nt)
{
    const PChain &c = chains[blockIdx.z];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c.n) c.uold[i] = c.uconf[(size_t)(nt - 1) * c.n + i];
}

extern "C" __global__ void __launch_bounds__(256)
kp_u_rescale(const PChain *chains, int nt)
{
    const PChain &c = chains[blockIdx.z];
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)c.n * c.ne) return;
    const int i = (int)(e % c.n), lc = c.ulcomp;
    const int old = c.uold[i], nw = c.uconf[(size_t)(nt - 1) * c.n + i];
    if (old == nw) return;
    const int iflip = ((nw - old) % lc + lc) % lc;           // nw = mod(old + iflip - 1, lc) + 1
    const cplx d = c.ud[(iflip - 1) + (lc - 1) * (old - 1)];
    c.UR[e] = cmul(c.UR[e], make_double2(1.0 + d.x, d.y));
}

// the bond field of family nf at slice nt before a G-based update, and UR's rows i1, i2 of every
// bond it flipped times 1 + delta_bmat_p, 1 + delta_bmat_m of the old value
// (src/dqmc_config_v/dqmc_proj_update_delay.f90:124-125; UR is in the family's rotated basis)
extern "C" __global__ void __launch_bounds__(256)
kp_v_save(const PChain *chains, int nt, int nf)
{
    const PChain &c = chains[blockIdx.z];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < c.nn_lf) c.uold[q] = c.conf[((size_t)(nt - 1) * c.nn_nf + nf) * c.n + q];
}

extern "C" __global__ void __launch_bounds__(256)
kp_v_rescale(const PChain *chains, int nt, int nf)
{
    const PChain &c = chains[blockIdx.z];
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)c.nn_lf * c.ne) return;
    const int q = (int)(e % c.nn_lf), j = (int)(e / c.nn_lf);
    const int old = c.uold[q], nw = c.conf[((size_t)(nt - 1) * c.nn_nf + nf) * c.n + q];
    if (old == nw) return;
    const int ist = nf * c.nn_lf + q;
    const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
    const cplx dp = c.vd[2 * (old - 1)], dm = c.vd[2 * (old - 1) + 1];
    cplx *a = c.UR + (size_t)j * c.n;
    a[i1] = cmul(a[i1], make_double2(1.0 + dp.x, dp.y));
    a[i2] = cmul(a[i2], make_double2(1.0 + dm.x, dm.y));
}

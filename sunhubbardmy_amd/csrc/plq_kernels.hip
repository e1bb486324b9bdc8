// plq_kernels.hip -- the plaquette field of the finite-temperature sweep (hconf, lprojplqu =
// .false.): the z x z factor of every plaquette applied to rows or columns of a chain matrix
// (src/dqmc_config_plqu/dqmc_{left,right}_{forward,backward}_prop.f90, _hc) and the sequential
// per-plaquette props and updates of one slice on G (dqmc_update_plqu.f90:1-86 with its place in
// dqmc_ft_sweep_0b.f90:82-93 and dqmc_ft_sweep_b0.f90:91-100).  Plaquettes overlap, so their
// factors are applied one after the other in the reference's order; a factor touches the zp rows
// (or columns) of its sites only.
#pragma once

#define PLQ_CB 8      // columns (rows) of a chain matrix per workgroup in k_plq_prop
#define PLQ_ZMAX 6

// the zp x zp factor of component `is` (0-based): pb (inv = 0) or pbi (inv = 1), column-major
__device__ __forceinline__ const cplx *plq_factor(const DChain &c, int inv, int is)
{
    return (inv ? c.pbi : c.pb) + (size_t)is * c.zp * c.zp;
}

// All plaquettes of slice nt on a chain matrix, plaquette 1..lqp (rev = 0) or lqp..1 (rev = 1):
//   side 0: rows   M(P, :) <- T M(P, :)   (left_forward: T = bmat; left_backward and
//                                          left_forward_hc: T = bmat_inv = bmat^H)
//   side 1: columns M(:, P) <- M(:, P) T (right_backward: T = bmat_inv; right_forward: bmat)
// A workgroup owns PLQ_CB whole columns (side 0) or rows (side 1) in LDS (leading dimension
// n + 1); thread (c, r) with r < zp forms entry r of its line's plaquette block.  Grid:
// (ceil(n / PLQ_CB), 1, chains).
extern "C" __global__ void __launch_bounds__(256)
k_plq_prop(const DChain *chains, int msel, int pt, int side, int inv, int rev, int nt)
{
    extern __shared__ cplx plqsh[];
    const DChain &c = chains[blockIdx.z];
    if (!c.lwrapp) return;
    const int n = c.n, z = c.zp, lqp = c.lqp, ld = n + 1;
    const int l0 = blockIdx.x * PLQ_CB;
    if (l0 >= n) return;
    const int nl = min(PLQ_CB, n - l0);
    cplx *line = plqsh;                                 // PLQ_CB x (n + 1)
    cplx *tab = line + (size_t)PLQ_CB * ld;             // 4 z^2: the factors of the four components
    int *cyc = (int *)(tab + 4 * z * z);                // lqp x z plaquette sites
    int *cf = cyc + lqp * z;                            // lqp: the field of slice nt
    GAS cplx *M = gp(chain_mat(c, msel, pt));
    for (int e = threadIdx.x; e < 4 * z * z; e += blockDim.x) tab[e] = (inv ? c.pbi : c.pb)[e];
    for (int e = threadIdx.x; e < lqp * z; e += blockDim.x) cyc[e] = c.plqc[e];
    for (int e = threadIdx.x; e < lqp; e += blockDim.x) cf[e] = c.confp[(size_t)(nt - 1) * lqp + e] - 1;
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
        const int l = e / n, i = e % n;
        line[l * ld + i] = side ? gld(M + (size_t)i * n + l0 + l) : gld(M + (size_t)(l0 + l) * n + i);
    }
    __syncthreads();
    const int l = threadIdx.x >> 5, r = threadIdx.x & 31;
    const bool act = l < nl && r < z;
    for (int q = 0; q < lqp; q++) {
        const int p = rev ? lqp - 1 - q : q;
        const int *P = cyc + p * z;
        const cplx *T = tab + (size_t)cf[p] * z * z;
        cplx v = make_double2(0.0, 0.0);
        if (act) {
            const cplx *x = line + l * ld;
            for (int m = 0; m < z; m++)
                v = cadd(v, side ? cmul(x[P[m]], T[m + r * z]) : cmul(T[r + m * z], x[P[m]]));
        }
        __syncthreads();
        if (act) line[l * ld + P[r]] = v;
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
        const int l2 = e / n, i = e % n;
        if (side)
            gst(M + (size_t)i * n + l0 + l2, line[l2 * ld + i]);
        else
            gst(M + (size_t)(l0 + l2) * n + i, line[l2 * ld + i]);
    }
}

// rows P of G <- T rows (thread j: column j) or columns P <- columns T (thread i: row i)
__device__ static void plq_apply_g(GAS cplx *G, int n, int z, const int *P, const cplx *T, int side)
{
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        cplx x[PLQ_ZMAX];
        for (int m = 0; m < z; m++) x[m] = side ? gld(G + (size_t)P[m] * n + j) : gld(G + (size_t)j * n + P[m]);
        for (int r = 0; r < z; r++) {
            cplx v = make_double2(0.0, 0.0);
            for (int m = 0; m < z; m++) v = cadd(v, side ? cmul(x[m], T[m + r * z]) : cmul(T[r + m * z], x[m]));
            if (side)
                gst(G + (size_t)P[r] * n + j, v);
            else
                gst(G + (size_t)j * n + P[r], v);
        }
    }
    __threadfence_block();
    __syncthreads();
}

// One slice of plaquettes on G with their updates, one workgroup of 256 threads per chain:
//   dir 0 (0b): plaquette 1..lqp: G <- B_p G B_p^-1 (left_forward, right_backward), update p
//   dir 1 (b0): plaquette lqp..1: update p, G <- B_p^-1 G B_p (right_forward, left_backward)
// The update (dqmc_update_plqu.f90:25-84): iflip = ceiling(u1 * 3), isp = mod(is + iflip - 1, 4)
// + 1, M = 1 + (1 - G)_PP delta(iflip, is), ratio1 = det M, ratiotot = ratio1^nflr gaml(isp) /
// gaml(is) phase_ratio(iflip, is), accepted on |Re(ratiotot phase)| / |Re(phase)| > u2; then
// G <- G - G(:, P) delta M^-1 (1 - G)(P, :), phase <- phase ratiotot / |ratiotot|.  M^-1 and
// det M by Gauss-Jordan with row pivoting on one thread (z <= 6).  slice = the slice's index in
// the pass (its 2 lqp draws, in update order).  LDS: G(:, P) and (1 - G)(P, :) (z n each).
extern "C" __global__ void __launch_bounds__(256)
k_plq_slice(const DChain *chains, int nt, int dir, int slice, int lupdate)
{
    extern __shared__ cplx plqsh[];
    const DChain &c = chains[blockIdx.x];
    if (!c.lwrapp) return;
    const int n = c.n, z = c.zp, lqp = c.lqp, tid = threadIdx.x;
    cplx *colP = plqsh;                     // n x z: G(:, P), column m at m n
    cplx *rowP = colP + (size_t)z * n;      // z x n: (1 - G)(P, :), then -S (1 - G)(P, :); (r, j) at r + z j
    __shared__ cplx s_gpp[PLQ_ZMAX * PLQ_ZMAX], s_S[PLQ_ZMAX * PLQ_ZMAX];
    __shared__ int s_acc;
    GAS cplx *G = gp(c.G);
    int32_t *cf = c.confp + (size_t)(nt - 1) * lqp;
    const double *u = lupdate ? c.unifp + (size_t)2 * lqp * slice : nullptr;
    cplx phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    double nacc = 0.0;
    int P[PLQ_ZMAX];
    for (int q = 0; q < lqp; q++) {
        const int p = dir ? lqp - 1 - q : q;
        for (int m = 0; m < z; m++) P[m] = c.plqc[p * z + m];
        const int is = cf[p] - 1;
        if (dir == 0) {
            plq_apply_g(G, n, z, P, plq_factor(c, 0, is), 0);
            plq_apply_g(G, n, z, P, plq_factor(c, 1, is), 1);
        }
        int isnow = is;
        if (lupdate) {
            for (int e = tid; e < z * z; e += blockDim.x)
                s_gpp[e] = gld(G + (size_t)P[e / z] * n + P[e % z]);   // G(P_a, P_b) at a + z b
            __syncthreads();
            if (tid == 0) {
                const double u1 = u[2 * q], u2 = u[2 * q + 1];
                const int iflip = (int)ceil(u1 * 3.0);
                const int isp = (is + iflip) % 4;
                const cplx *dl = c.pdl + (size_t)(3 * is + iflip - 1) * z * z;
                // A = M = 1 + (1 - G_PP) delta, B = identity; Gauss-Jordan to A^-1 with det
                cplx A[PLQ_ZMAX * PLQ_ZMAX], B[PLQ_ZMAX * PLQ_ZMAX];
                for (int a = 0; a < z; a++)
                    for (int b = 0; b < z; b++) {
                        cplx v = make_double2(a == b ? 1.0 : 0.0, 0.0);
                        for (int m = 0; m < z; m++) {
                            const cplx g = s_gpp[a + z * m];
                            const cplx om = make_double2((a == m ? 1.0 : 0.0) - g.x, -g.y);
                            v = cadd(v, cmul(om, dl[m + z * b]));
                        }
                        A[a + z * b] = v;
                        B[a + z * b] = make_double2(a == b ? 1.0 : 0.0, 0.0);
                    }
                cplx det = make_double2(1.0, 0.0);
                for (int k = 0; k < z; k++) {
                    int pv = k;
                    double best = cabs2(A[k + z * k]);
                    for (int a = k + 1; a < z; a++)
                        if (cabs2(A[a + z * k]) > best) { best = cabs2(A[a + z * k]); pv = a; }
                    if (pv != k) {
                        det = make_double2(-det.x, -det.y);
                        for (int b = 0; b < z; b++) {
                            cplx t = A[k + z * b]; A[k + z * b] = A[pv + z * b]; A[pv + z * b] = t;
                            t = B[k + z * b]; B[k + z * b] = B[pv + z * b]; B[pv + z * b] = t;
                        }
                    }
                    const cplx piv = A[k + z * k];
                    det = cmul(det, piv);
                    const cplx ip = cinv(piv);
                    for (int b = 0; b < z; b++) { A[k + z * b] = cmul(A[k + z * b], ip); B[k + z * b] = cmul(B[k + z * b], ip); }
                    for (int a = 0; a < z; a++) {
                        if (a == k) continue;
                        const cplx f = A[a + z * k];
                        for (int b = 0; b < z; b++) {
                            A[a + z * b] = csub(A[a + z * b], cmul(f, A[k + z * b]));
                            B[a + z * b] = csub(B[a + z * b], cmul(f, B[k + z * b]));
                        }
                    }
                }
                cplx rt = make_double2(1.0, 0.0);
                for (int f = 0; f < c.nflr; f++) rt = cmul(rt, det);
                rt = cscale(rt, c.pgaml[isp] / c.pgaml[is]);
                rt = cmul(rt, c.ppr[(iflip - 1) + 3 * is]);
                const double rr = cmul(rt, phase).x / phase.x;
                const int acc = fabs(rr) > u2;
                if (acc) {
                    // thread 0 takes every decision, so it alone carries the phase
                    phase = cmul(phase, cscale(rt, 1.0 / hypot(rt.x, rt.y)));
                    // S = delta M^-1
                    for (int a = 0; a < z; a++)
                        for (int b = 0; b < z; b++) {
                            cplx v = make_double2(0.0, 0.0);
                            for (int m = 0; m < z; m++) v = cadd(v, cmul(dl[a + z * m], B[m + z * b]));
                            s_S[a + z * b] = v;
                        }
                    cf[p] = isp + 1;
                }
                s_acc = acc ? isp + 1 : 0;
            }
            __syncthreads();
            const int acc = s_acc;
            if (acc) {
                isnow = acc - 1;
                nacc += 1.0;
                for (int j = tid; j < n; j += blockDim.x) {
                    for (int m = 0; m < z; m++) {
                        colP[j + (size_t)m * n] = gld(G + (size_t)P[m] * n + j);
                        const cplx g = gld(G + (size_t)j * n + P[m]);
                        rowP[m + (size_t)z * j] = make_double2((j == P[m] ? 1.0 : 0.0) - g.x, -g.y);
                    }
                    // -S (1 - G)(P, j) in place: thread j owns column j of rowP
                    cplx t[PLQ_ZMAX];
                    for (int a = 0; a < z; a++) {
                        cplx v = make_double2(0.0, 0.0);
                        for (int m = 0; m < z; m++) v = cadd(v, cmul(s_S[a + z * m], rowP[m + (size_t)z * j]));
                        t[a] = make_double2(-v.x, -v.y);
                    }
                    for (int a = 0; a < z; a++) rowP[a + (size_t)z * j] = t[a];
                }
                __syncthreads();
                for (size_t e = tid; e < (size_t)n * n; e += blockDim.x) {
                    const int i = (int)(e % n), j = (int)(e / n);
                    cplx v = gld(G + e);
                    for (int m = 0; m < z; m++) v = cadd(v, cmul(colP[i + (size_t)m * n], rowP[m + (size_t)z * j]));
                    gst(G + e, v);
                }
                __threadfence_block();
                __syncthreads();
            }
        }
        if (dir == 1) {
            plq_apply_g(G, n, z, P, plq_factor(c, 0, isnow), 1);
            plq_apply_g(G, n, z, P, plq_factor(c, 1, isnow), 0);
        }
    }
    if (tid == 0) {
        c.sc[SC_PHASE] = phase.x;
        c.sc[SC_PHASE + 1] = phase.y;
        if (lupdate) {
            c.pacc[0] += nacc;
            c.pacc[1] += (double)lqp;
        }
    }
}

// blocked.hip -- blocked factorizations of one chain's matrix inside one workgroup:
// panels of up to 16 columns are factored in LDS, and every trailing update is a rank-16
// product on FP64 MFMA (v_mfma_f64_16x16x4_f64), so each panel streams the trailing matrix
// once instead of once per column.  The bodies take explicit sizes and leading dimensions so
// that the square N x N matrices of the finite-temperature sweep and the tall N x ne
// matrices of the projective sweep share them.
//   qr_blk_body     Householder QR without pivoting (zgeqrf + zlarft), T factors kept per panel
//   formq_blk_body  Q = H_1 ... H_n from the reflectors and the T factors (zungqr)
//   lu_blk_body     LU with partial pivoting, pivot = largest |Re|+|Im| (zgetrf)
//   solve_blk_body  (P L U) Y = X for nrhs right-hand sides
#include "sunhip_kernels.h"

#define PB 16          // largest panel width (T factors are stored PB x PB)
#define BLK_NT 512     // threads per workgroup (8 waves)
#define BLK_NMAX 512   // LDS holds a 512 x 16 complex panel (128 KiB)

// MFMA operand layout (gfx950): A(16x4): lane l holds A[l%16][l/16]; B(4x16): lane l holds
// B[l/16][l%16]; D(16x16): lane l, element r holds D[l/16 + 4r][l%16].
__device__ __forceinline__ double4_t mfma64(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ static double blk_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += red[i];
    return s;
}

__device__ static void blk_argmax(double v, int idx, double *rv, int *ri, int *outi)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    wave_argmax_dpp(v, idx);
    __syncthreads();
    if (lane == 0) { rv[w] = v; ri[w] = idx; }
    __syncthreads();
    double bv = rv[0];
    int bi = ri[0];
    for (int i = 1; i < nw; i++)
        if (rv[i] > bv || (rv[i] == bv && ri[i] < bi)) { bv = rv[i]; bi = ri[i]; }
    *outi = bi;
}

// the m x b panel A(k0 + r, k0 + p) into P (leading dimension ldp), eight elements per thread in
// flight (clamped addresses)
__device__ __forceinline__ void panel_stage(cplx *P, int ldp, const cplx *A_, int lda, int k0, int m, int b)
{
    GAS const cplx *A = gp(A_);
    for (int base = threadIdx.x; base < m * b; base += 8 * (int)blockDim.x) {
        cplx v8[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = min(base + u * (int)blockDim.x, m * b - 1), r = e % m, p = e / m;
            v8[u] = gld(A + ((size_t)(k0 + r) + (size_t)(k0 + p) * lda));
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = base + u * (int)blockDim.x;
            if (e < m * b) P[(e / m) * ldp + e % m] = v8[u];
        }
    }
}

// V[r][p] of a Householder panel stored LAPACK-style (unit diagonal implicit, zeros above)
__device__ __forceinline__ cplx vpanel(const cplx *P, int ldp, int r, int p)
{
    if (r > p) return P[p * ldp + r];
    return make_double2(r == p ? 1.0 : 0.0, 0.0);
}

// The products below read their global operands through clamped, in-bounds addresses with no
// branch around the load (rows and columns past the edge compute values that are never stored),
// so a tile's loads are all in flight together and the next tile's are issued before the
// current tile's MFMAs.

// W (16 x nc, global, ld 16) = V^H C, V = m x 16 Householder panel in LDS, C global (m x nc, ldc).
// The rows are walked in blocks of 16 whose C values are loaded two blocks ahead into one of
// two alternating register sets.
__device__ __forceinline__ void vhc_block(double4_t &ar, double4_t &ai, const cplx *P, int ldp, int m, int b,
                                          GAS cplx *C, size_t ccol, int bk, int kq, int l16, cplx (&y)[4])
{
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int r = 16 * bk + 4 * u + kq;
        cplx x = make_double2(0.0, 0.0);
        const cplx yy = r < m ? y[u] : make_double2(0.0, 0.0);
        if (r < m && l16 < b) x = cconj(vpanel(P, ldp, r, l16));         // X[p][k] = conj V[k][p]
        ar = mfma64(x.x, yy.x, ar);
        ar = mfma64(-x.y, yy.y, ar);
        ai = mfma64(x.x, yy.y, ai);
        ai = mfma64(x.y, yy.x, ai);
        y[u] = gld(C + ((size_t)min(16 * (bk + 2) + 4 * u + kq, m - 1) + ccol));   // Y[k][q] = C[k][c0+q]
    }
}

__device__ static void wg_vhc(cplx *W_, const cplx *P, int ldp, int m, int b, const cplx *C_, int ldc, int nc)
{
    GAS cplx *W = gp(W_), *C = gp(C_);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    const int ntile = (nc + 15) >> 4, nb = (m + 15) >> 4;
    for (int t = w; t < ntile; t += nw) {
        const int c0 = t * 16;
        const size_t ccol = (size_t)min(c0 + l16, nc - 1) * ldc;
        double4_t ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
        cplx ya[4], yb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            ya[u] = gld(C + ((size_t)min(4 * u + kq, m - 1) + ccol));
            yb[u] = gld(C + ((size_t)min(16 + 4 * u + kq, m - 1) + ccol));
        }
        for (int bk = 0; bk < nb; bk += 2) {
            vhc_block(ar, ai, P, ldp, m, b, C, ccol, bk, kq, l16, ya);
            if (bk + 1 < nb) vhc_block(ar, ai, P, ldp, m, b, C, ccol, bk + 1, kq, l16, yb);
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const int p = kq + 4 * rr, q = c0 + l16;
            if (q < nc) gst(W + (p + 16 * (size_t)q), make_double2(ar[rr], ai[rr]));
        }
    }
}

// W <- op(T) W, T 16 x 16 upper triangular in LDS (col-major), op = T^H (herm) or T
__device__ static void wg_tmul(cplx *W_, int nc, const cplx *T, int b, bool herm)
{
    GAS cplx *W = gp(W_);
    for (int q = threadIdx.x; q < nc; q += blockDim.x) {
        cplx x[PB], y[PB];
#pragma unroll
        for (int p = 0; p < PB; p++) {
            const cplx v = gld(W + (min(p, b - 1) + 16 * (size_t)q));
            x[p] = (p < b) ? v : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int p = 0; p < PB; p++) {
            cplx s = make_double2(0.0, 0.0);
#pragma unroll
            for (int pp = 0; pp < PB; pp++) {
                if (herm) {            // (T^H)[p][pp] = conj T[pp][p], nonzero for pp <= p
                    if (pp <= p) s = cadd(s, cmul(cconj(T[p * PB + pp]), x[pp]));
                } else {               // T[p][pp], nonzero for pp >= p
                    if (pp >= p) s = cadd(s, cmul(T[pp * PB + p], x[pp]));
                }
            }
            y[p] = s;
        }
#pragma unroll
        for (int p = 0; p < PB; p++)
            if (p < b) gst(W + (p + 16 * (size_t)q), y[p]);
    }
}

// C (mrows x ncols, global, ldc) -= A B with A = 16 columns in LDS (A[p*lda + r], unit-panel
// accessor when vp = 1) and B = 16 x ncols in global memory (B[p + c*ldb]).  The accumulator
// holds C^T tiles so that the global reads and writes of C are 16-element column runs.  A
// wave's tiles alternate between two register sets; a set is refilled with the C tile and B
// fragment of the wave's tile after next as soon as they are consumed.
struct R16Args {
    GAS cplx *C;
    GAS cplx *B;
    const cplx *A;
    int ldc, mrows, ncols, lda, roff, b, ldb, mt, ntiles;
    bool vp;
};

__device__ __forceinline__ void r16_load(const R16Args &g, int t, int kq, int l16, cplx (&cv)[4], cplx (&xv)[4])
{
    t = min(t, g.ntiles - 1);
    const int r0 = (t % g.mt) * 16, c0 = (t / g.mt) * 16;
    const size_t rowc = (size_t)min(r0 + l16, g.mrows - 1);
#pragma unroll
    for (int j = 0; j < 4; j++) cv[j] = gld(g.C + (rowc + (size_t)min(c0 + kq + 4 * j, g.ncols - 1) * g.ldc));
    const size_t xc = (size_t)min(c0 + l16, g.ncols - 1) * g.ldb;
#pragma unroll
    for (int s = 0; s < 4; s++) xv[s] = gld(g.B + (min(4 * s + kq, g.b - 1) + xc));
}

__device__ __forceinline__ void r16_tile(const R16Args &g, int t, int step, int kq, int l16, cplx (&cv)[4], cplx (&xv)[4])
{
    const int r0 = (t % g.mt) * 16, c0 = (t / g.mt) * 16;
    const int row = r0 + l16;
    double4_t ar, ai;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        ar[j] = cv[j].x;
        ai[j] = cv[j].y;
    }
    // the C tile of the wave's tile after next into the registers just read
    const int tn = min(t + step, g.ntiles - 1);
    const int nr0 = (tn % g.mt) * 16, nc0 = (tn / g.mt) * 16;
    const size_t nrowc = (size_t)min(nr0 + l16, g.mrows - 1);
#pragma unroll
    for (int j = 0; j < 4; j++) cv[j] = gld(g.C + (nrowc + (size_t)min(nc0 + kq + 4 * j, g.ncols - 1) * g.ldc));
    const size_t nxc = (size_t)min(nc0 + l16, g.ncols - 1) * g.ldb;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int p = 4 * s + kq;
        cplx x = make_double2(0.0, 0.0), y = make_double2(0.0, 0.0);
        if (p < g.b) {
            x = xv[s];                                                          // X[c][p] = B[p][c]
            if (row < g.mrows) y = g.vp ? vpanel(g.A, g.lda, g.roff + row, p) : g.A[p * g.lda + g.roff + row];  // Y[p][r]
        }
        ar = mfma64(-x.x, y.x, ar);
        ar = mfma64(x.y, y.y, ar);
        ai = mfma64(-x.x, y.y, ai);
        ai = mfma64(-x.y, y.x, ai);
        // and its B fragment, entry by entry as the current one is consumed
        xv[s] = gld(g.B + (min(4 * s + kq, g.b - 1) + nxc));
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int col = c0 + kq + 4 * j;
        if (row < g.mrows && col < g.ncols) gst(g.C + ((size_t)row + (size_t)col * g.ldc), make_double2(ar[j], ai[j]));
    }
}

__device__ static void wg_rank16(cplx *C, int ldc, int mrows, int ncols, const cplx *A, int lda, int roff, int b,
                                 bool vp, const cplx *B, int ldb)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int l16 = lane & 15, kq = lane >> 4;
    R16Args g;
    g.C = gp(C);
    g.B = gp(B);
    g.A = A;
    g.ldc = ldc; g.mrows = mrows; g.ncols = ncols; g.lda = lda; g.roff = roff; g.b = b; g.ldb = ldb;
    g.mt = (mrows + 15) >> 4;
    g.ntiles = g.mt * ((ncols + 15) >> 4);
    g.vp = vp;
    if (w >= g.ntiles) return;
    cplx ca[4], xa[4], cb[4], xb[4];
    r16_load(g, w, kq, l16, ca, xa);
    r16_load(g, w + nw, kq, l16, cb, xb);
    for (int t = w; t < g.ntiles; t += 2 * nw) {
        r16_tile(g, t, 2 * nw, kq, l16, ca, xa);
        if (t + nw < g.ntiles) r16_tile(g, t + nw, 2 * nw, kq, l16, cb, xb);
    }
}

// T factor of one panel (zlarft, forward, columnwise): S = V^H V over the panel's m rows,
// T(:,j) = -tau_j T(:,0:j) S(0:j, j); V is the m x b panel in LDS (ld M, unit diagonal implicit).
// Row i of T only needs row i itself and S (T(i, j) = -tau_j sum_{i <= q < j} T(i, q) S(q, j)),
// so lane i builds its row in registers, all rows at once.
__device__ static void panel_tfactor(const cplx *P, int M, int m, int b, const cplx *tau, cplx *T, cplx *S)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int pr = w; pr < b * b; pr += nw) {
        int i = pr % b, j = pr / b;
        if (i >= j) continue;
        double sr = 0.0, si = 0.0;
        for (int r = j + lane; r < m; r += 64) {
            cplx vi = vpanel(P, M, r, i), vj = vpanel(P, M, r, j);
            sr += vi.x * vj.x + vi.y * vj.y;
            si += vi.x * vj.y - vi.y * vj.x;
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) S[j * PB + i] = make_double2(sr, si);
    }
    __syncthreads();
    if (threadIdx.x < PB) {
        const int i = threadIdx.x;
        GAS cplx *tg = gp(tau);
        cplx tv[PB], trow[PB];
#pragma unroll
        for (int j = 0; j < PB; j++) {
            tv[j] = gld(tg + min(j, b - 1));
            trow[j] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int j = 0; j < PB; j++) {
            if (j < b && j > i) {
                cplx acc = make_double2(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < PB; q++)
                    if (q >= i && q < j) acc = cadd(acc, cmul(trow[q], S[j * PB + q]));
                trow[j] = cmul(make_double2(-tv[j].x, -tv[j].y), acc);
            } else if (j < b && j == i) {
                trow[j] = tv[j];
            }
        }
#pragma unroll
        for (int j = 0; j < PB; j++) T[j * PB + i] = trow[j];
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------
// QR without pivoting, blocked: A (M x Nc, M >= Nc, lda) -> reflectors below the diagonal and
// R on and above it; tau[Nc]; the panel's T factor to tblk[panel]; pb = panel width <= PB;
// lds holds M x pb panel + 2 PB x PB; wk holds PB x Nc.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void qr_blk_body(cplx *A, int M, int Nc, int lda, cplx *tau_out, cplx *tblk, cplx *wk, int pb,
                                   cplx *lds)
{
    cplx *P = lds;                      // M x pb panel, ld M
    cplx *T = lds + (size_t)M * pb;     // PB x PB
    cplx *S = T + PB * PB;              // PB x PB Gram matrix
    __shared__ double red[16];
    __shared__ cplx s_par[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int k0 = 0; k0 < Nc; k0 += pb) {
        const int m = M - k0, b = min(pb, Nc - k0);
        panel_stage(P, M, A, lda, k0, m, b);
        __syncthreads();
        for (int j = 0; j < b; j++) {
            double s = 0.0;
            for (int r = j + 1 + threadIdx.x; r < m; r += blockDim.x) s += cabs2(P[j * M + r]);
            s = blk_sum(s, red);
            if (threadIdx.x == 0) {
                cplx alpha = P[j * M + j];
                cplx tau = make_double2(0.0, 0.0), scal = make_double2(1.0, 0.0);
                double beta = alpha.x;
                if (!(s == 0.0 && alpha.y == 0.0)) {
                    double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + s);
                    beta = alpha.x >= 0 ? -nrm : nrm;
                    tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
                    scal = cinv(make_double2(alpha.x - beta, alpha.y));
                    P[j * M + j] = make_double2(beta, 0.0);
                }
                s_par[0] = tau;
                s_par[1] = scal;
                tau_out[k0 + j] = tau;
            }
            __syncthreads();
            const cplx tau = s_par[0], scal = s_par[1];
            for (int r = j + 1 + threadIdx.x; r < m; r += blockDim.x) P[j * M + r] = cmul(P[j * M + r], scal);
            __syncthreads();
            // H_j^H on the remaining panel columns, one wave per column
            const cplx tc = cconj(tau);
            for (int q = j + 1 + w; q < b; q += nw) {
                double sr = 0.0, si = 0.0;
                for (int r = j + lane; r < m; r += 64) {
                    cplx v = (r == j) ? make_double2(1.0, 0.0) : P[j * M + r], a = P[q * M + r];
                    sr += v.x * a.x + v.y * a.y;
                    si += v.x * a.y - v.y * a.x;
                }
                sr = wave_sum(sr);
                si = wave_sum(si);
                cplx f = cmul(tc, make_double2(sr, si));
                for (int r = j + lane; r < m; r += 64) {
                    cplx v = (r == j) ? make_double2(1.0, 0.0) : P[j * M + r];
                    P[q * M + r] = csub(P[q * M + r], cmul(v, f));
                }
            }
            __syncthreads();
        }
        panel_tfactor(P, M, m, b, tau_out + k0, T, S);
        for (int e = threadIdx.x; e < PB * PB; e += blockDim.x) tblk[(size_t)(k0 / pb) * PB * PB + e] = T[e];
        for (int e = threadIdx.x; e < m * b; e += blockDim.x) {
            int r = e % m, p = e / m;
            A[(size_t)(k0 + r) + (size_t)(k0 + p) * lda] = P[p * M + r];
        }
        const int nc = Nc - k0 - b;
        if (nc > 0) {
            cplx *Ct = A + (size_t)k0 + (size_t)(k0 + b) * lda;
            wg_vhc(wk, P, M, m, b, Ct, lda, nc);
            __syncthreads();
            wg_tmul(wk, nc, T, b, true);
            __syncthreads();
            wg_rank16(Ct, lda, m, nc, P, M, 0, b, true, wk, PB);
        }
        __syncthreads();
    }
}

// Q (M x Nc, ldq) = H_1 ... H_Nc applied to the first Nc columns of the identity: backward
// accumulation by panels, Q(k0:, k0:) <- (I - V T V^H) Q(k0:, k0:)
__device__ __forceinline__ void formq_blk_body(const cplx *A, int M, int Nc, int lda, cplx *Q, int ldq, const cplx *tblk,
                                      cplx *wk, int pb, cplx *lds)
{
    cplx *P = lds;
    cplx *T = lds + (size_t)M * pb;
    for (size_t e = threadIdx.x; e < (size_t)M * Nc; e += blockDim.x) {
        int i = (int)(e % M), j = (int)(e / M);
        Q[(size_t)i + (size_t)j * ldq] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    const int np = (Nc + pb - 1) / pb;
    for (int pi = np - 1; pi >= 0; pi--) {
        const int k0 = pi * pb, m = M - k0, b = min(pb, Nc - k0), nc = Nc - k0;
        panel_stage(P, M, A, lda, k0, m, b);
        for (int e = threadIdx.x; e < PB * PB; e += blockDim.x) T[e] = tblk[(size_t)pi * PB * PB + e];
        __syncthreads();
        cplx *Ct = Q + (size_t)k0 + (size_t)k0 * ldq;
        wg_vhc(wk, P, M, m, b, Ct, ldq, nc);
        __syncthreads();
        wg_tmul(wk, nc, T, b, false);
        __syncthreads();
        wg_rank16(Ct, ldq, m, nc, P, M, 0, b, true, wk, PB);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// LU with partial pivoting, blocked (right-looking), in place on A (n x n, lda); piv[k] = row
// interchanged with k (0-based); log det as s_inv_logdet_lu_z (Im folded into (-pi, pi]);
// thread 0 writes *ldet and *info (first zero pivot + 1, or 0)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void lu_blk_body(cplx *A, int n, int lda, int32_t *piv, cplx *ldet, double *info, cplx *lds)
{
    cplx *P = lds;                       // panel, ld n
    __shared__ double rv[16];
    __shared__ int ri[16];
    __shared__ int s_info;
    __shared__ int s_pv[PB];
    GAS cplx *Ag = gp(A);
    if (threadIdx.x == 0) s_info = 0;
    for (int k0 = 0; k0 < n; k0 += PB) {
        const int m = n - k0, b = min(PB, m);
        panel_stage(P, n, A, lda, k0, m, b);
        __syncthreads();
        for (int j = 0; j < b; j++) {
            double bv = -1.0;
            int bi = j;
            for (int r = j + threadIdx.x; r < m; r += blockDim.x) {
                cplx a = P[j * n + r];
                double v = fabs(a.x) + fabs(a.y);
                if (v > bv) { bv = v; bi = r; }
            }
            int p;
            blk_argmax(bv, bi, rv, ri, &p);
            if (threadIdx.x == 0) { s_pv[j] = p; piv[k0 + j] = k0 + p; }
            if (p != j)
                for (int q = threadIdx.x; q < b; q += blockDim.x) {
                    cplx t = P[q * n + p];
                    P[q * n + p] = P[q * n + j];
                    P[q * n + j] = t;
                }
            __syncthreads();
            cplx pv = P[j * n + j];
            bool zero = (pv.x == 0.0 && pv.y == 0.0);
            if (zero && threadIdx.x == 0 && s_info == 0) s_info = k0 + j + 1;
            if (!zero) {
                cplx pinv = cinv(pv);
                for (int r = j + 1 + threadIdx.x; r < m; r += blockDim.x) P[j * n + r] = cmul(P[j * n + r], pinv);
            }
            __syncthreads();
            for (int e = threadIdx.x; e < (m - j - 1) * (b - j - 1); e += blockDim.x) {
                int r = j + 1 + e % (m - j - 1), q = j + 1 + e / (m - j - 1);
                P[q * n + r] = csub(P[q * n + r], cmul(P[j * n + r], P[q * n + j]));
            }
            __syncthreads();
        }
        for (int e = threadIdx.x; e < m * b; e += blockDim.x) {
            int r = e % m, p = e / m;
            A[(size_t)(k0 + r) + (size_t)(k0 + p) * lda] = P[p * n + r];
        }
        // the panel's interchanges on the columns left and right of it, in order.  When every
        // swap pairs a panel row with a row below the panel that no other swap names (the usual
        // case) the swaps touch distinct rows, so they run eight at a time with their loads
        // issued before their stores; otherwise one after another.
        bool dep = false;
        if ((int)threadIdx.x < b) {
            const int t = threadIdx.x, pt = s_pv[t];
            if (pt != t) {
                dep = pt < b;
                for (int jj = 0; jj < t; jj++) dep = dep || s_pv[jj] == pt;
            }
        }
        const bool indep = !__syncthreads_or(dep ? 1 : 0);
        for (int q = threadIdx.x; q < n; q += blockDim.x) {
            if (q >= k0 && q < k0 + b) continue;
            cplx *col = A + (size_t)q * lda;
            if (indep) {
                GAS cplx *cg = gp(col) + k0;
                for (int j0 = 0; j0 < b; j0 += 8) {
                    cplx xa[8], xb[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int j = min(j0 + u, b - 1);
                        xa[u] = gld(cg + j);
                        xb[u] = gld(cg + s_pv[j]);
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int j = j0 + u;
                        if (j < b && s_pv[j] != j) {
                            gst(cg + j, xb[u]);
                            gst(cg + s_pv[j], xa[u]);
                        }
                    }
                }
            } else {
                for (int j = 0; j < b; j++) {
                    int p = s_pv[j];
                    if (p != j) {
                        cplx t = col[k0 + p];
                        col[k0 + p] = col[k0 + j];
                        col[k0 + j] = t;
                    }
                }
            }
        }
        __syncthreads();
        const int nc = n - k0 - b;
        if (nc > 0) {
            // U12 = L11^-1 A12 (unit lower 16 x 16 from the panel), one thread per column
            for (int q = threadIdx.x; q < nc; q += blockDim.x) {
                cplx *col = A + (size_t)k0 + (size_t)(k0 + b + q) * lda;
                cplx x[PB];
#pragma unroll
                for (int i = 0; i < PB; i++) x[i] = (i < b) ? col[i] : make_double2(0.0, 0.0);
#pragma unroll
                for (int i = 0; i < PB; i++)
#pragma unroll
                    for (int ii = 0; ii < PB; ii++)
                        if (ii < i && i < b) x[i] = csub(x[i], cmul(P[ii * n + i], x[ii]));
#pragma unroll
                for (int i = 0; i < PB; i++)
                    if (i < b) col[i] = x[i];
            }
            __syncthreads();
            // A22 -= L21 U12
            cplx *C = A + (size_t)(k0 + b) + (size_t)(k0 + b) * lda;
            wg_rank16(C, lda, m - b, nc, P, n, b, b, false, A + (size_t)k0 + (size_t)(k0 + b) * lda, lda);
        }
        __syncthreads();
    }
    // log det = sum_i log(+-U_ii), summed by the whole workgroup
    __syncthreads();
    double lr = 0.0, li = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        cplx d = gld(gp(A) + ((size_t)i + (size_t)i * lda));
        if (piv[i] != i) d = make_double2(-d.x, -d.y);
        const cplx l = clog_(d);
        lr += l.x;
        li += l.y;
    }
    lr = blk_sum(lr, rv);
    li = blk_sum(li, rv);
    if (threadIdx.x == 0) {
        const double k2 = round(li / 2.0 / M_PI);
        *ldet = make_double2(lr, li - 2.0 * M_PI * k2);
        *info = (double)s_info;
    }
}

// (P L U) Y = X in place of X (n x nrhs, ldx), with the LU factors of lu_blk_body in A: row
// interchanges through Y (n x nrhs work, ld n), then block forward and back substitution with
// rank-16 updates on MFMA; lds holds n x PB + n ints
__device__ __forceinline__ void solve_blk_body(const cplx *A, int n, int lda, const int32_t *piv, cplx *X, int ldx, int nrhs,
                                      cplx *Y, cplx *lds)
{
    cplx *P = lds;                            // 16 columns of L or U, ld n
    int *perm = (int *)(lds + (size_t)n * PB);
    // final row order of the interchanges
    if (threadIdx.x == 0) {
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int k = 0; k < n; k++) {
            int p = piv[k];
            int t = perm[p]; perm[p] = perm[k]; perm[k] = t;
        }
    }
    __syncthreads();
    for (size_t e = threadIdx.x; e < (size_t)n * nrhs; e += blockDim.x) {
        int i = (int)(e % n), j = (int)(e / n);
        Y[e] = X[(size_t)perm[i] + (size_t)j * ldx];
    }
    __syncthreads();
    for (size_t e = threadIdx.x; e < (size_t)n * nrhs; e += blockDim.x) {
        int i = (int)(e % n), j = (int)(e / n);
        X[(size_t)i + (size_t)j * ldx] = Y[e];
    }
    __syncthreads();
    // forward: unit lower L
    for (int k0 = 0; k0 < n; k0 += PB) {
        const int b = min(PB, n - k0), mb = n - k0 - b;
        for (int e = threadIdx.x; e < (n - k0) * b; e += blockDim.x) {
            int r = e % (n - k0), p = e / (n - k0);
            P[p * n + r] = A[(size_t)(k0 + r) + (size_t)(k0 + p) * lda];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nrhs; q += blockDim.x) {
            cplx *col = X + (size_t)k0 + (size_t)q * ldx;
            cplx x[PB];
#pragma unroll
            for (int i = 0; i < PB; i++) x[i] = (i < b) ? col[i] : make_double2(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < PB; i++)
#pragma unroll
                for (int ii = 0; ii < PB; ii++)
                    if (ii < i && i < b) x[i] = csub(x[i], cmul(P[ii * n + i], x[ii]));
#pragma unroll
            for (int i = 0; i < PB; i++)
                if (i < b) col[i] = x[i];
        }
        __syncthreads();
        if (mb > 0) wg_rank16(X + (size_t)(k0 + b), ldx, mb, nrhs, P, n, b, b, false, X + (size_t)k0, ldx);
        __syncthreads();
    }
    // backward: upper U
    const int np = (n + PB - 1) / PB;
    for (int pi = np - 1; pi >= 0; pi--) {
        const int k0 = pi * PB, b = min(PB, n - k0);
        // rows 0 .. k0+b of the 16 columns
        for (int e = threadIdx.x; e < (k0 + b) * b; e += blockDim.x) {
            int r = e % (k0 + b), p = e / (k0 + b);
            P[p * n + r] = A[(size_t)r + (size_t)(k0 + p) * lda];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nrhs; q += blockDim.x) {
            cplx *col = X + (size_t)k0 + (size_t)q * ldx;
            cplx x[PB];
#pragma unroll
            for (int i = 0; i < PB; i++) x[i] = (i < b) ? col[i] : make_double2(0.0, 0.0);
#pragma unroll
            for (int i = PB - 1; i >= 0; i--) {
                if (i < b) {
#pragma unroll
                    for (int ii = 0; ii < PB; ii++)
                        if (ii > i && ii < b) x[i] = csub(x[i], cmul(P[ii * n + k0 + i], x[ii]));
                    x[i] = cdiv(x[i], P[i * n + k0 + i]);
                }
            }
#pragma unroll
            for (int i = 0; i < PB; i++)
                if (i < b) col[i] = x[i];
        }
        __syncthreads();
        if (k0 > 0) wg_rank16(X, ldx, k0, nrhs, P, n, 0, b, false, X + (size_t)k0, ldx);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// The finite-temperature chain's N x N factorizations (one workgroup per chain)
// ------------------------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(BLK_NT)
k_qr_blk(const DChain *chains, int sel, int out_pt)
{
    extern __shared__ cplx bsh[];
    __shared__ double red[16];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    qr_blk_body(chain_mat(c, sel, 0), n, n, n, c.tau, c.tblk, c.wk, PB, bsh);
    // log det Q = sum_i log(1 - tau_i 2 Re(tau_i) / |tau_i|^2), summed by the whole workgroup
    double sr = 0.0, si = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const cplx t = gld(gp(c.tau) + i);
        if (t.x != 0.0 || t.y != 0.0) {
            const cplx q = cdiv(cscale(t, 2.0 * t.x), cmul(t, cconj(t)));
            const cplx l = clog_(make_double2(1.0 - q.x, -q.y));
            sr += l.x;
            si += l.y;
        }
    }
    sr = blk_sum(sr, red);
    si = blk_sum(si, red);
    if (threadIdx.x == 0) c.ldq[out_pt] = make_double2(sr, si);
}

extern "C" __global__ void __launch_bounds__(BLK_NT)
k_formq_blk(const DChain *chains, int sel, int out_pt)
{
    extern __shared__ cplx bsh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    formq_blk_body(chain_mat(c, sel, 0), n, n, n, c.U + (size_t)n * n * out_pt, n, c.tblk, c.wk, PB, bsh);
}

extern "C" __global__ void __launch_bounds__(BLK_NT)
k_lu_blk(const DChain *chains)
{
    extern __shared__ cplx bsh[];
    const DChain &c = chains[blockIdx.x];
    __shared__ cplx s_ld;
    __shared__ double s_inf;
    lu_blk_body(c.W3, c.n, c.n, c.piv, &s_ld, &s_inf, bsh);
    // the row order of the interchanges, for k_solve_cb: the pivots into LDS, the swaps in order
    // by one thread, the result out
    const int n = c.n;
    int *perm = (int *)bsh, *pv = perm + n;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        perm[i] = i;
        pv[i] = c.piv[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        c.sc[SC_ZLD] = s_ld.x;
        c.sc[SC_ZLD + 1] = s_ld.y;
        c.sc[SC_INFO] = s_inf;
        for (int k = 0; k < n; k++) {
            const int p = pv[k], t = perm[p];
            perm[p] = perm[k];
            perm[k] = t;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) c.lperm[i] = perm[i];
}

// ------------------------------------------------------------------------------------------
// (P L U) Y = X in place of X for N <= 256, 16 right-hand sides per workgroup (zgetrs with the
// factors of k_lu_blk).  The workgroup's 16 columns stay in registers through the forward and
// the backward substitution, in the MFMA accumulator layout: wave w holds the 16-row tiles
// t = w, w + 4, w + 8, w + 12; lane l, element r of a tile is row 16 t + 4 r + l / 16, column
// l % 16.  Panel p (rows 16 p .. 16 p + 15) is solved by the wave that holds tile p against the
// 16 x 16 diagonal block (staged in LDS), published to LDS, and every wave then updates its
// tiles below (forward, L) or above (backward, U) by one rank-16 MFMA product per tile whose
// factor block is read from global memory.  A chain's workgroups run on one XCD
// (blockIdx.x % 8 picks it) so that they share the factor's panels in that XCD's L2.
// ------------------------------------------------------------------------------------------
#define SCB_T 4   // tiles per wave: N <= 256

// forward (unit lower, lower = 1) or backward (upper with diagonal reciprocals dinv) solve of
// one 16 x 16 tile held in accumulator layout, against D11[a * 16 + b] in LDS
template <bool LOWER>
__device__ __forceinline__ void cb_tri(double4_t &xr, double4_t &xi, const cplx *D11, const cplx *dinv, int lq, int l16)
{
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int i = LOWER ? q : 15 - q;
        const int ri = i >> 2, qi = i & 3;
        if (!LOWER && lq == qi) {
            const cplx d = dinv[i];
            const double a = xr[ri], b = xi[ri];
            xr[ri] = a * d.x - b * d.y;
            xi[ri] = a * d.y + b * d.x;
        }
        const int src = l16 + 16 * qi;
        const double vr = __shfl(xr[ri], src, 64), vi = __shfl(xi[ri], src, 64);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int a = 4 * r + lq;
            if (LOWER ? a > i : a < i) {
                const cplx m = D11[a * 16 + i];
                xr[r] -= m.x * vr - m.y * vi;
                xi[r] -= m.x * vi + m.y * vr;
            }
        }
    }
}

// tiles of this wave past (forward) or before (backward) panel p -= F(tile rows, panel cols) Xp
template <bool FWD>
__device__ __forceinline__ void cb_update(double4_t (&xr)[SCB_T], double4_t (&xi)[SCB_T], GAS cplx *A, int n, int p,
                                          int w, int lq, int l16, const cplx *Xp)
{
    const int k0 = 16 * p;
    cplx bv[4];
#pragma unroll
    for (int s = 0; s < 4; s++) bv[s] = Xp[(4 * s + lq) * 16 + l16];
#pragma unroll
    for (int tt = 0; tt < SCB_T; tt++) {
        const int tile = w + 4 * tt;
        if (FWD ? tile > p : tile < p) {
            const int row = 16 * tile + l16;
            const size_t rowc = (size_t)min(row, n - 1);
            cplx av[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int kk = k0 + 4 * s + lq;
                const cplx a = gld(A + (rowc + (size_t)min(kk, n - 1) * n));
                av[s] = (row < n && kk < n) ? a : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const cplx a = av[s], b = bv[s];
                xr[tt] = mfma64(-a.x, b.x, xr[tt]);
                xr[tt] = mfma64(a.y, b.y, xr[tt]);
                xi[tt] = mfma64(-a.x, b.y, xi[tt]);
                xi[tt] = mfma64(-a.y, b.x, xi[tt]);
            }
        }
    }
}

// the diagonal block of panel p (identity past n), four entries per lane, into registers
__device__ __forceinline__ void cb_diag_load(GAS cplx *A, int n, int p, int lane, cplx (&d)[4])
{
    const int k0 = 16 * p;
#pragma unroll
    for (int e4 = 0; e4 < 4; e4++) {
        const int e = lane + 64 * e4, a = e >> 4, b = e & 15;
        const cplx v = gld(A + ((size_t)min(k0 + a, n - 1) + (size_t)min(k0 + b, n - 1) * n));
        d[e4] = (k0 + a < n && k0 + b < n) ? v : make_double2(a == b ? 1.0 : 0.0, 0.0);
    }
}

// the owner of panel p stages the diagonal block (loaded ahead into d), solves its tile,
// publishes it
template <bool FWD>
__device__ __forceinline__ void cb_panel(double4_t (&xr)[SCB_T], double4_t (&xi)[SCB_T], const cplx (&d)[4], int p,
                                         int lane, int lq, int l16, cplx *D11, cplx *dinv, cplx *Xp)
{
#pragma unroll
    for (int e4 = 0; e4 < 4; e4++) D11[lane + 64 * e4] = d[e4];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (!FWD && lane < 16) dinv[lane] = cinv(D11[lane * 16 + lane]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int tp = p >> 2;
#pragma unroll
    for (int tt = 0; tt < SCB_T; tt++) {
        if (tt == tp) {
            cb_tri<FWD>(xr[tt], xi[tt], D11, dinv, lq, l16);
#pragma unroll
            for (int r = 0; r < 4; r++) Xp[(4 * r + lq) * 16 + l16] = make_double2(xr[tt][r], xi[tt][r]);
        }
    }
}

extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_solve_cb(const DChain *chains, int nch, int ncb, int xsel)
{
    const int bid = blockIdx.x, xcd = bid & 7, kb = bid >> 3;
    const int chain = xcd + 8 * (kb / ncb), cb = kb % ncb;
    if (chain >= nch) return;
    const DChain &c = chains[chain];
    const int n = c.n, j0 = 16 * cb;
    if (j0 >= n) return;
    GAS cplx *A = gp(c.W3), *X = gp(chain_mat(c, xsel, 0));
    GAS int32_t *perm = gp(c.lperm);
    __shared__ cplx D11[256], Xp[2][256], dinv[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int nt = (n + 15) >> 4;
    const int col = j0 + l16;
    const size_t colc = (size_t)min(col, n - 1) * n;
    double4_t xr[SCB_T], xi[SCB_T];
    // the rows of P X
#pragma unroll
    for (int tt = 0; tt < SCB_T; tt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 16 * (w + 4 * tt) + 4 * r + lq;
            const int pr = perm[min(row, n - 1)];
            const cplx v = gld(X + ((size_t)pr + colc));
            xr[tt][r] = row < n ? v.x : 0.0;
            xi[tt][r] = row < n ? v.y : 0.0;
        }
    // the owner of the next panel loads its diagonal block while the current panel is solved
    // and applied
    cplx d[4];
    if (w == 0) cb_diag_load(A, n, 0, lane, d);
    for (int p = 0; p < nt; p++) {
        if (w == (p & 3)) cb_panel<true>(xr, xi, d, p, lane, lq, l16, D11, dinv, Xp[p & 1]);
        if (w == ((p + 1) & 3)) cb_diag_load(A, n, p + 1 < nt ? p + 1 : nt - 1, lane, d);
        __syncthreads();
        cb_update<true>(xr, xi, A, n, p, w, lq, l16, Xp[p & 1]);
    }
    __syncthreads();
    for (int p = nt - 1; p >= 0; p--) {
        if (w == (p & 3)) cb_panel<false>(xr, xi, d, p, lane, lq, l16, D11, dinv, Xp[p & 1]);
        if (w == ((p + 3) & 3)) cb_diag_load(A, n, p > 0 ? p - 1 : 0, lane, d);
        __syncthreads();
        cb_update<false>(xr, xi, A, n, p, w, lq, l16, Xp[p & 1]);
    }
#pragma unroll
    for (int tt = 0; tt < SCB_T; tt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 16 * (w + 4 * tt) + 4 * r + lq;
            if (row < n && col < n) gst(X + ((size_t)row + (size_t)col * n), make_double2(xr[tt][r], xi[tt][r]));
        }
}

extern "C" __global__ void __launch_bounds__(BLK_NT)
k_solve_blk(const DChain *chains, int xsel)
{
    extern __shared__ cplx bsh[];
    const DChain &c = chains[blockIdx.x];
    solve_blk_body(c.W3, c.n, c.n, c.piv, chain_mat(c, xsel, 0), c.n, c.n, c.W2, bsh);
}

// T factors of the reflectors that the column-at-a-time (pivoted) QR left in matrix sel, panel
// by panel, so that k_formq_blk forms Q by rank-16 MFMA products
extern "C" __global__ void __launch_bounds__(BLK_NT)
k_tfactors(const DChain *chains, int sel)
{
    extern __shared__ cplx bsh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    cplx *P = bsh;
    cplx *T = bsh + (size_t)n * PB;
    cplx *S = T + PB * PB;
    const cplx *A = chain_mat(c, sel, 0);
    for (int k0 = 0; k0 < n; k0 += PB) {
        const int m = n - k0, b = min(PB, m);
        for (int e = threadIdx.x; e < m * b; e += blockDim.x) {
            int r = e % m, p = e / m;
            P[p * n + r] = A[(size_t)(k0 + r) + (size_t)(k0 + p) * n];
        }
        __syncthreads();
        panel_tfactor(P, n, m, b, c.tau + k0, T, S);
        for (int e = threadIdx.x; e < PB * PB; e += blockDim.x) c.tblk[(size_t)(k0 / PB) * PB * PB + e] = T[e];
        __syncthreads();
    }
}

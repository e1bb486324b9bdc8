// kernels.hip -- CDNA4 (gfx950) kernels of the DQMC sweep.  Every kernel is batched over
// independent Markov chains: blockIdx.z (GEMM, elementwise) or blockIdx.x (one workgroup per
// chain) selects the chain, whose matrix dimension may differ from its neighbours'.
#include "sunhip_kernels.h"

#define FULL_DBL 1

// ------------------------------------------------------------------------------------------
// SFMT19937 on the device (lib/m_spring.f90:187-379): every chain owns its generator; state =
// 624 uint32 lanes + index, kept in global memory between calls.
// ------------------------------------------------------------------------------------------
__device__ static void sfmt_refill(uint32_t *w)
{
    const uint32_t msk[4] = {0xdfffffefu, 0xddfecb7fu, 0xbffaffffu, 0xbffffff6u};
    uint32_t r1[4], r2[4];
    for (int l = 0; l < 4; l++) { r1[l] = w[4 * 154 + l]; r2[l] = w[4 * 155 + l]; }
    for (int i = 0; i < 156; i++) {
        uint32_t a[4], b[4];
        int ib = i + 122; if (ib >= 156) ib -= 156;
        for (int l = 0; l < 4; l++) { a[l] = w[4 * i + l]; b[l] = w[4 * ib + l]; }
        uint64_t ahi = ((uint64_t)a[3] << 32) | a[2], alo = ((uint64_t)a[1] << 32) | a[0];
        uint64_t xhi = (ahi << 8) | (alo >> 56), xlo = alo << 8;
        uint64_t chi = ((uint64_t)r1[3] << 32) | r1[2], clo = ((uint64_t)r1[1] << 32) | r1[0];
        uint64_t ylo = (clo >> 8) | (chi << 56), yhi = chi >> 8;
        uint32_t x[4] = {(uint32_t)xlo, (uint32_t)(xlo >> 32), (uint32_t)xhi, (uint32_t)(xhi >> 32)};
        uint32_t y[4] = {(uint32_t)ylo, (uint32_t)(ylo >> 32), (uint32_t)yhi, (uint32_t)(yhi >> 32)};
        uint32_t o[4];
        for (int l = 0; l < 4; l++) o[l] = a[l] ^ x[l] ^ ((b[l] >> 11) & msk[l]) ^ y[l] ^ (r2[l] << 18);
        for (int l = 0; l < 4; l++) { w[4 * i + l] = o[l]; r1[l] = r2[l]; r2[l] = o[l]; }
    }
}

__device__ static double sfmt_next(uint32_t *w, int &idx)
{
    if (idx >= 624) { sfmt_refill(w); idx = 0; }
    return ((double)w[idx++] + 0.5) * (1.0 / 4294967296.0);
}

// init_gen_rand(seed) and the period certification (lib/m_spring.f90:187-243)
__device__ static void sfmt_seed_state(uint32_t *w, int32_t seed)
{
    const uint32_t parity[4] = {0x00000001u, 0x00000000u, 0x00000000u, 0x13c9e684u};
    w[0] = (uint32_t)seed;
    for (int i = 1; i < 624; i++) w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + (uint32_t)i;
    uint32_t inner = 0;
    for (int i = 0; i < 4; i++) inner ^= w[i] & parity[i];
    for (int sh = 16; sh > 0; sh >>= 1) inner ^= inner >> sh;
    if ((inner & 1u) != 1u) {
        bool done = false;
        for (int i = 0; i < 4 && !done; i++) {
            uint32_t work = 1;
            for (int j = 0; j < 32; j++) {
                if (work & parity[i]) { w[i] ^= work; done = true; break; }
                work <<= 1;
            }
        }
    }
    w[624] = 624;
}

extern "C" __global__ void k_sfmt_seed(const DChain *chains, int nchains, const int32_t *seeds)
{
    int ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= nchains) return;
    sfmt_seed_state(chains[ci].rng, seeds[ci]);
}

// Draws tot numbers from the generator whose state (624 words + index) is at g, one workgroup
// of 64 threads: the state is kept in LDS, lane 0 runs each refill (the recursion is sequential
// over the 156 words), and the whole wave converts and stores the block; emit(k, u) receives
// draw k (0-based) of this call.
template <typename Emit>
__device__ static void sfmt_draw_block(uint32_t *g, size_t tot, Emit emit)
{
    __shared__ uint32_t w[624];
    for (int i = threadIdx.x; i < 624; i += blockDim.x) w[i] = g[i];
    int idx = (int)g[624];
    __syncthreads();
    size_t k = 0;
    while (k < tot) {
        if (idx >= 624) {
            if (threadIdx.x == 0) sfmt_refill(w);
            __syncthreads();
            idx = 0;
        }
        const int cnt = (int)min((size_t)(624 - idx), tot - k);
        for (int j = threadIdx.x; j < cnt; j += blockDim.x)
            emit(k + j, ((double)w[idx + j] + 0.5) * (1.0 / 4294967296.0));
        idx += cnt;
        k += cnt;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < 624; i += blockDim.x) g[i] = w[i];
    if (threadIdx.x == 0) g[624] = (uint32_t)idx;
    __syncthreads();   // a following call in the same workgroup reads the state back
}

// one workgroup per chain; mode 0: fresh HS field conf(i,nt) = ceiling(u*lcomp) (nt outer);
// mode 1: the pass's 2*N*ltrot uniforms
extern "C" __global__ void __launch_bounds__(64)
k_sfmt_draw(const DChain *chains, int nchains, int mode)
{
    const DChain &c = chains[blockIdx.x];
    if (mode == 0) {
        // hconf%input_plqconf, u0conf%input_conf, then v0conf%input_conf (src/dqmc_ft_ctrl.f90:19-35)
        if (c.lwrapp) {
            int32_t *cp = c.confp;
            sfmt_draw_block(c.rng, (size_t)c.lqp * c.ltrot, [=](size_t k, double u) { cp[k] = (int32_t)ceil(u * 4); });
        }
        if (c.lwrapu) {
            const int lc = c.lcomp;
            int32_t *conf = c.conf;
            sfmt_draw_block(c.rng, (size_t)c.n * c.ltrot, [=](size_t k, double u) { conf[k] = (int32_t)ceil(u * lc); });
        }
        if (c.lwrapv) {
            int32_t *cv = c.confv;
            sfmt_draw_block(c.rng, (size_t)c.n * c.nn_nf * c.ltrot, [=](size_t k, double u) { cv[k] = (int32_t)ceil(u * 2); });
        }
    } else if (!c.lwrapv && !c.lwrapp) {
        double *un = c.unif;
        sfmt_draw_block(c.rng, (size_t)2 * c.n * c.ltrot, [=](size_t k, double u) { un[k] = u; });
    } else {
        // per slice the on-site draws (2 n, when the Hubbard field is on), the bond draws
        // (2 nn_nf nn_lf) and the plaquette draws (2 lqp) in the order the pass consumes them:
        // b0 (mode 1) on-site, bonds, plaquettes; 0b (mode 2) plaquettes, bonds, on-site
        // (dqmc_ft_sweep_b0.f90:70-100, dqmc_ft_sweep_0b.f90:82-112)
        double *un = c.unif, *uv = c.unifv, *up = c.unifp;
        const size_t nu = c.lwrapu ? (size_t)2 * c.n : 0, nv = c.lwrapv ? (size_t)2 * c.nn_nf * c.nn_lf : 0;
        const size_t np = c.lwrapp ? (size_t)2 * c.lqp : 0, per = nu + nv + np;
        const bool ufirst = mode == 1;
        sfmt_draw_block(c.rng, per * c.ltrot, [=](size_t k, double u) {
            const size_t s = k / per, r = k % per;
            if (ufirst) {
                if (r < nu) un[s * nu + r] = u;
                else if (r < nu + nv) uv[s * nv + r - nu] = u;
                else up[s * np + r - nu - nv] = u;
            } else {
                if (r < np) up[s * np + r] = u;
                else if (r < np + nv) uv[s * nv + r - np] = u;
                else un[s * nu + r - np - nv] = u;
            }
        });
    }
}

// ------------------------------------------------------------------------------------------
// Kinetic and interaction propagation, checkerboard form (BREAKUP_T).
//   column pass (left products, rows mixed): a workgroup owns CPB full columns in LDS.
//   row pass (right products, columns mixed): a workgroup owns RPB full rows in LDS (ld n + 1).
// Program bits: 1 = families 0..nn_nf-1, 2 = V scaling, 4 = families nn_nf-1..0,
//               8 = V scaling before everything.  Several slices can be chained (B products).
// Bond 2x2 blocks are column-major [T11, T21, T12, T22] (dqmc_left_forward_prop.f90:18-44).
// ------------------------------------------------------------------------------------------
#define CPB 8
#define RPB 8

// Global <-> LDS staging with ST_BATCH independent 16-byte accesses in flight per thread: a
// batch's loads are all issued before its first LDS store, so a workgroup waits for one memory
// latency per batch instead of one per element.  ld(e) returns element e of the tile, st(e, v)
// writes it back.
#define ST_BATCH 8
template <int B, class Ld, class Ix>
__device__ __forceinline__ void stage_in_map(cplx *lds, int total, Ld ld, Ix ix)
{
    for (int base = threadIdx.x; base < total; base += B * blockDim.x) {
        cplx v[B];
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int e = base + u * (int)blockDim.x;
            v[u] = e < total ? ld(e) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int e = base + u * (int)blockDim.x;
            if (e < total) lds[ix(e)] = v[u];
        }
    }
}

template <class Ld>
__device__ __forceinline__ void stage_in(cplx *lds, int total, Ld ld)
{
    stage_in_map<ST_BATCH>(lds, total, ld, [](int e) { return e; });
}

template <class St>
__device__ __forceinline__ void stage_out(const cplx *lds, int total, St st)
{
    for (int base = threadIdx.x; base < total; base += ST_BATCH * blockDim.x) {
        cplx v[ST_BATCH];
#pragma unroll
        for (int u = 0; u < ST_BATCH; u++) {
            const int e = base + u * (int)blockDim.x;
            v[u] = e < total ? lds[e] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < ST_BATCH; u++) {
            const int e = base + u * (int)blockDim.x;
            if (e < total) st(e, v[u]);
        }
    }
}

__device__ __forceinline__ cplx vtab(const DChain &c, int vsel, int is)
{
    // vsel 0: bmat, 1: bmat_inv, 2: conj(bmat)
    cplx b = (vsel == 1) ? c.bmat_inv[is - 1] : c.bmat[is - 1];
    if (vsel == 2) b.y = -b.y;
    return b;
}

// one bond per thread (and a stride of columns): the bond's endpoints and 2x2 block are read
// once per family and applied to every column the thread covers
template <class Chain>
__device__ __forceinline__ void cb_left_fams(const Chain &c, cplx *lds, int ncol, const cplx *T, bool rev)
{
    const int nf = c.nn_nf, nl = c.nn_lf, n = c.n;
    const int groups = max(1, (int)blockDim.x / nl);
    const int b = threadIdx.x % nl, g = threadIdx.x / nl;
    for (int q = 0; q < nf; q++) {
        int f = rev ? nf - 1 - q : q;
        for (int bb = b; bb < nl; bb += blockDim.x) {
            if (g >= groups) break;
            const int ist = f * nl + bb;
            const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
            const cplx t0 = T[4 * ist], t1 = T[4 * ist + 1], t2 = T[4 * ist + 2], t3 = T[4 * ist + 3];
            for (int cc = g; cc < ncol; cc += groups) {
                cplx x = lds[cc * n + i1], y = lds[cc * n + i2];
                lds[cc * n + i1] = cadd(cmul(t0, x), cmul(t2, y));
                lds[cc * n + i2] = cadd(cmul(t1, x), cmul(t3, y));
            }
        }
        __syncthreads();
    }
}

// the four factors of the HS components, staged in LDS once per pass (a register table indexed by
// the component is lowered to a private-memory lookup)
__device__ __forceinline__ const cplx *vtab_lds(const DChain &c, int vsel)
{
    __shared__ cplx vt[4];
    if (threadIdx.x < 4) vt[threadIdx.x] = vtab(c, vsel, threadIdx.x + 1);
    __syncthreads();
    return vt;
}

__device__ static void cb_left_v(const DChain &c, cplx *lds, int ncol, int nt, int vsel)
{
    const int n = c.n;
    GAS const int32_t *cf = gp(c.conf + (size_t)(nt - 1) * n);
    const cplx *vt = vtab_lds(c, vsel);
#pragma unroll 8
    for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
        int r = e % n;
        lds[e] = cmul(vt[cf[r] - 1], lds[e]);
    }
    __syncthreads();
}

extern "C" __global__ void __launch_bounds__(256)
k_cb_cols(const DChain *chains, int msel, int pt, int ksel, int prog, int vsel, int nt0, int nt1, int ntstep)
{
    extern __shared__ cplx lds[];
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const int c0 = blockIdx.x * CPB;
    if (c0 >= n) return;
    const int ncol = min(CPB, n - c0);
    cplx *M = chain_mat(c, msel, pt);
    const cplx *T = ksel ? c.ktm1 : c.kt;
    // the CPB columns are one contiguous run of n * ncol elements
    GAS cplx *Mc = gp(M) + (size_t)c0 * n;
    stage_in(lds, n * ncol, [=](int e) { return gld(Mc + (e)); });
    __syncthreads();
    // slices nt0, nt0+ntstep, ..., nt1 (ntstep = +-1); the count bounds the loop
    const int nslice = (ntstep > 0 ? nt1 - nt0 : nt0 - nt1) + 1;
    for (int q = 0, nt = nt0; q < nslice; q++, nt += ntstep) {
        if (nt < 1 || nt > c.ltrot) break;
        if ((prog & 8) && c.lwrapu) cb_left_v(c, lds, ncol, nt, vsel);
        if ((prog & 1) && c.lwrapT) cb_left_fams(c, lds, ncol, T, false);
        if ((prog & 2) && c.lwrapu) cb_left_v(c, lds, ncol, nt, vsel);
        if ((prog & 4) && c.lwrapT) cb_left_fams(c, lds, ncol, T, true);
    }
    stage_out(lds, n * ncol, [=](int e, cplx v) { gst(Mc + (e), v); });
}

// rows in LDS are row-major with the padded leading dimension ldr = n + 1 (lds[rr * ldr + col]):
// the lanes of a family pass touch consecutive sites of one row, and the staging writes of
// eight consecutive rows fall on different banks
__host__ __device__ __forceinline__ int cb_ldr(int n) { return n + 1; }

template <class Chain>
__device__ __forceinline__ void cb_right_fams(const Chain &c, cplx *lds, int nrow, const cplx *T, bool rev)
{
    const int ldr = cb_ldr(c.n);
    const int nf = c.nn_nf, nl = c.nn_lf;
    const int groups = max(1, (int)blockDim.x / nl);
    const int b = threadIdx.x % nl, g = threadIdx.x / nl;
    for (int q = 0; q < nf; q++) {
        int f = rev ? nf - 1 - q : q;
        for (int bb = b; bb < nl; bb += blockDim.x) {
            if (g >= groups) break;
            const int ist = f * nl + bb;
            const int i1 = c.bonds[2 * ist], i2 = c.bonds[2 * ist + 1];
            const cplx t0 = T[4 * ist], t1 = T[4 * ist + 1], t2 = T[4 * ist + 2], t3 = T[4 * ist + 3];
            for (int rr = g; rr < nrow; rr += groups) {
                cplx x = lds[rr * ldr + i1], y = lds[rr * ldr + i2];
                lds[rr * ldr + i1] = cadd(cmul(x, t0), cmul(y, t1));
                lds[rr * ldr + i2] = cadd(cmul(x, t2), cmul(y, t3));
            }
        }
        __syncthreads();
    }
}

__device__ static void cb_right_v(const DChain &c, cplx *lds, int nt, int vsel)
{
    const int n = c.n;
    GAS const int32_t *cf = gp(c.conf + (size_t)(nt - 1) * n);
    const cplx *vt = vtab_lds(c, vsel);
    const int ldr = cb_ldr(n);
#pragma unroll 8
    for (int e = threadIdx.x; e < n * RPB; e += blockDim.x) {
        const int rr = e / n, col = e % n;
        lds[rr * ldr + col] = cmul(lds[rr * ldr + col], vt[cf[col] - 1]);
    }
    __syncthreads();
}

extern "C" __global__ void __launch_bounds__(256)
k_cb_rows(const DChain *chains, int msel, int pt, int ksel, int prog, int vsel, int nt)
{
    extern __shared__ cplx lds[];
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const int r0 = blockIdx.x * RPB;
    if (r0 >= n) return;
    const int nrow = min(RPB, n - r0);
    cplx *M = chain_mat(c, msel, pt);
    const cplx *T = ksel ? c.ktm1 : c.kt;
    // lds[rr * ldr + col] = M(r0 + rr, col): eight consecutive threads read one 128-byte run
    GAS cplx *Mg = gp(M);
    const int ldr = cb_ldr(n);
    stage_in_map<ST_BATCH>(lds, n * RPB, [=](int e) {
        const int rr = e % RPB, col = e / RPB;
        return rr < nrow ? gld(Mg + ((size_t)(r0 + rr) + (size_t)col * n)) : make_double2(0.0, 0.0);
    }, [=](int e) { return (e % RPB) * ldr + e / RPB; });
    __syncthreads();
    if ((prog & 8) && c.lwrapu) cb_right_v(c, lds, nt, vsel);
    if ((prog & 1) && c.lwrapT) cb_right_fams(c, lds, nrow, T, false);
    if ((prog & 2) && c.lwrapu) cb_right_v(c, lds, nt, vsel);
    if ((prog & 4) && c.lwrapT) cb_right_fams(c, lds, nrow, T, true);
    for (int base = threadIdx.x; base < n * RPB; base += ST_BATCH * blockDim.x) {
        cplx v[ST_BATCH];
#pragma unroll
        for (int u = 0; u < ST_BATCH; u++) {
            const int e = base + u * (int)blockDim.x;
            v[u] = e < n * RPB ? lds[(e % RPB) * ldr + e / RPB] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < ST_BATCH; u++) {
            const int e = base + u * (int)blockDim.x, rr = e % RPB, col = e / RPB;
            if (e < n * RPB && rr < nrow) gst(Mg + ((size_t)(r0 + rr) + (size_t)col * n), v[u]);
        }
    }
}

// exp(V) scalings for the dense-K form: rows (side 0) or columns (side 1), or both (side 2:
// rows by vsel, columns by vsel2)
extern "C" __global__ void __launch_bounds__(256)
k_vscale(const DChain *chains, int msel, int pt, int side, int vsel, int vsel2, int nt)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    if (!c.lwrapu) return;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    int r = (int)(e % n), col = (int)(e / n);
    const int32_t *cf = c.conf + (size_t)(nt - 1) * n;
    cplx *M = chain_mat(c, msel, pt);
    cplx v = M[e];
    if (side == 0 || side == 2) v = cmul(vtab(c, vsel, cf[r]), v);
    if (side == 1) v = cmul(v, vtab(c, vsel, cf[col]));
    if (side == 2) v = cmul(v, vtab(c, vsel2, cf[col]));
    M[e] = v;
}

// ------------------------------------------------------------------------------------------
// Batched ZGEMM on FP64 MFMA (v_mfma_f64_16x16x4_f64), C = alpha op(A) op(B) + beta C,
// all N x N of the chain.  Workgroup tile 64x64, 4 waves of 32x32, K chunk 16 staged in LDS.
// The accumulator holds the transposed tile so that stores are 256-byte column runs.
// MFMA operand layout (probed on gfx950): A(16x4): lane l holds A[l%16][l/16]; B(4x16): lane l
// holds B[l/16][l%16]; D(16x16): lane l, element r holds D[l/16 + 4r][l%16].
// ------------------------------------------------------------------------------------------
#define GT 64
#define GK 16

__device__ __forceinline__ double4_t mfma(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// one 64 x 64 tile (tm, tn) of C (m x n, ldc) = alpha op(A) op(B) + beta C, op(A) m x kk,
// op(B) kk x n; op = 0: as stored, 1: conjugate transpose
template <int OPA, int OPB>
__device__ __forceinline__ void zgemm_tile(const cplx *A_, int lda, const cplx *B_, int ldb, cplx *C_, int ldc, int m, int n,
                                  int kk, int tm, int tn, double ar, double ai, double br, double bi)
{
    GAS cplx *A = gp(A_), *B = gp(B_), *C = gp(C_);
    __shared__ double sAr[GT][GK + 1], sAi[GT][GK + 1];
    __shared__ double sBr[GK][GT + 1], sBi[GK][GT + 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    double4_t accr[2][2], acci[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) { accr[i][j][r] = 0.0; acci[i][j][r] = 0.0; }
    // the next k block's operands are loaded into registers (clamped, branch-free addresses)
    // while the current block's MFMAs run from LDS
    cplx va[4], vb[4];
    auto load_k = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int idx = t + 256 * q;
            int i, k, j;
            if (OPA == 0) { i = idx % GT; k = idx / GT; } else { k = idx % GK; i = idx / GK; }
            const int ic = min(tm + i, m - 1), kc = min(k0 + k, kk - 1);
            const cplx a = OPA == 0 ? gld(A + ((size_t)ic + (size_t)kc * lda)) : cconj(gld(A + ((size_t)kc + (size_t)ic * lda)));
            va[q] = (tm + i < m && k0 + k < kk) ? a : make_double2(0.0, 0.0);
            if (OPB == 0) { k = idx % GK; j = idx / GK; } else { j = idx % GT; k = idx / GT; }
            const int jc = min(tn + j, n - 1), kc2 = min(k0 + k, kk - 1);
            const cplx u = OPB == 0 ? gld(B + ((size_t)kc2 + (size_t)jc * ldb)) : cconj(gld(B + ((size_t)jc + (size_t)kc2 * ldb)));
            vb[q] = (tn + j < n && k0 + k < kk) ? u : make_double2(0.0, 0.0);
        }
    };
    load_k(0);
    for (int k0 = 0; k0 < kk; k0 += GK) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int idx = t + 256 * q;
            int i, k, j;
            if (OPA == 0) { i = idx % GT; k = idx / GT; } else { k = idx % GK; i = idx / GK; }
            sAr[i][k] = va[q].x; sAi[i][k] = va[q].y;
            if (OPB == 0) { k = idx % GK; j = idx / GK; } else { j = idx % GT; k = idx / GT; }
            sBr[k][j] = vb[q].x; sBi[k][j] = vb[q].y;
        }
        __syncthreads();
        if (k0 + GK < kk) load_k(k0 + GK);
#pragma unroll
        for (int ks = 0; ks < GK / 4; ks++) {
            const int k = 4 * ks + (lane >> 4), l16 = lane & 15;
            double xr[2], xi[2], yr[2], yi[2];
            for (int s = 0; s < 2; s++) {
                // transposed product: X = op(B)^T rows (output columns), Y = op(A)^T cols (output rows)
                xr[s] = sBr[k][32 * wn + 16 * s + l16]; xi[s] = sBi[k][32 * wn + 16 * s + l16];
                yr[s] = sAr[32 * wm + 16 * s + l16][k]; yi[s] = sAi[32 * wm + 16 * s + l16][k];
            }
            for (int sn = 0; sn < 2; sn++)
                for (int sm = 0; sm < 2; sm++) {
                    accr[sm][sn] = mfma(xr[sn], yr[sm], accr[sm][sn]);
                    accr[sm][sn] = mfma(-xi[sn], yi[sm], accr[sm][sn]);
                    acci[sm][sn] = mfma(xr[sn], yi[sm], acci[sm][sn]);
                    acci[sm][sn] = mfma(xi[sn], yr[sm], acci[sm][sn]);
                }
        }
        __syncthreads();
    }
    const bool zb = (br == 0.0 && bi == 0.0);
    for (int sm = 0; sm < 2; sm++)
        for (int sn = 0; sn < 2; sn++)
            for (int r = 0; r < 4; r++) {
                int row = tm + 32 * wm + 16 * sm + (lane & 15);
                int col = tn + 32 * wn + 16 * sn + (lane >> 4) + 4 * r;
                if (row < m && col < n) {
                    size_t o = (size_t)row + (size_t)col * ldc;
                    double vr = accr[sm][sn][r], vi = acci[sm][sn][r];
                    double outr = ar * vr - ai * vi, outi = ar * vi + ai * vr;
                    if (!zb) { cplx cv = gld(C + (o)); outr += br * cv.x - bi * cv.y; outi += br * cv.y + bi * cv.x; }
                    gst(C + (o), make_double2(outr, outi));
                }
            }
}

template <int OPA, int OPB>
__global__ void __launch_bounds__(256)
k_zgemm(const DChain *chains, int selA, int ptA, int selB, int ptB, int selC, int ptC,
        double ar, double ai, double br, double bi)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const int tm = blockIdx.x * GT, tn = blockIdx.y * GT;
    if (tm >= n || tn >= n) return;
    zgemm_tile<OPA, OPB>(chain_mat(c, selA, ptA), n, chain_mat(c, selB, ptB), n, chain_mat(c, selC, ptC), n, n, n, n,
                         tm, tn, ar, ai, br, bi);
}

template __global__ void k_zgemm<0, 0>(const DChain *, int, int, int, int, int, int, double, double, double, double);
template __global__ void k_zgemm<1, 0>(const DChain *, int, int, int, int, int, int, double, double, double, double);
template __global__ void k_zgemm<0, 1>(const DChain *, int, int, int, int, int, int, double, double, double, double);

// ------------------------------------------------------------------------------------------
// Workgroup-level rank-k update on MFMA: G(n x n) = (G + A B^T) diag(cs) with A, B n x k
// column-major (leading dimension n).  Used by the delay flush (cs = 1) and the submatrix flush.
// Computed as G^T tiles so that every global access is a 16-element column run.
// ------------------------------------------------------------------------------------------
__device__ static void wg_rank_k(cplx *G, int n, const cplx *A, const cplx *B, int k, const double *csr,
                                 const double *csi)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nt = (n + 15) >> 4;
    const int l16 = lane & 15, kq = lane >> 4;
    for (int tile = w; tile < nt * nt; tile += nw) {
        const int i0 = (tile % nt) * 16, j0 = (tile / nt) * 16;
        // accumulator = G^T tile: element (p, q) = G[i0+q][j0+p], lane holds p = l/16 + 4r, q = l%16
        double4_t accr, acci;
        const int row = i0 + l16;
        for (int r = 0; r < 4; r++) {
            int col = j0 + kq + 4 * r;
            cplx g = (row < n && col < n) ? G[(size_t)row + (size_t)col * n] : make_double2(0.0, 0.0);
            accr[r] = g.x; acci[r] = g.y;
        }
        for (int m0 = 0; m0 < k; m0 += 4) {
            const int m = m0 + kq;
            cplx x = make_double2(0.0, 0.0), y = make_double2(0.0, 0.0);
            if (m < k) {
                if (j0 + l16 < n) x = B[(size_t)(j0 + l16) + (size_t)m * n];   // X[p][m] = B[j0+p][m]
                if (i0 + l16 < n) y = A[(size_t)(i0 + l16) + (size_t)m * n];   // Y[m][q] = A[i0+q][m]
            }
            accr = mfma(x.x, y.x, accr);
            accr = mfma(-x.y, y.y, accr);
            acci = mfma(x.x, y.y, acci);
            acci = mfma(x.y, y.x, acci);
        }
        for (int r = 0; r < 4; r++) {
            int col = j0 + kq + 4 * r;
            if (row < n && col < n) {
                cplx v = make_double2(accr[r], acci[r]);
                if (csr) v = cmul(v, make_double2(csr[col], csi[col]));
                G[(size_t)row + (size_t)col * n] = v;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// The same rank-k update G += A B^T for k <= K known at compile time, organised for latency:
// a wave owns 16-row strips of G; the strip's A fragment (K/4 complex per lane) stays in
// registers while the wave walks the 16-column tiles.  Every load has a clamped, in-bounds
// address and no branch around it (rows and columns past n compute values that are never
// stored; entries past k are zero in the A fragment), so nothing waits on a load before its
// use: G tiles are loaded two tiles ahead into one of two alternating register sets (the flush
// alone takes 210 VGPRs so; three sets take 251, four spill), the B fragment one tile ahead (a
// second B set would spill in the delay kernel).  At k = 32 the
// update reads and writes G for 8 flops per byte, so it is bound by HBM, not by the MFMAs; G
// streams through with nontemporal accesses so that the A and B fragments, re-read for every
// tile, stay in L2.
// ------------------------------------------------------------------------------------------
template <int KS>
__device__ __forceinline__ void rank_k_tile(GAS cplx *G, GAS cplx *B, int n, int k, int row, bool rok, int kq, int l16,
                                            int tj, const cplx (&y)[KS], cplx (&x)[KS], cplx (&g)[4])
{
    const int j0 = tj * 16;
    double4_t accr, acci;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        accr[r] = g[r].x;
        acci[r] = g[r].y;
    }
    // this register set now takes tile tj + 2 (clamped: past the last tile it reloads the last
    // columns, which are never used); the B fragment takes tile tj + 1
    const size_t rowc = (size_t)min(row, n - 1);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int col = min(j0 + 32 + kq + 4 * r, n - 1);
        g[r] = gld_nt(G + (rowc + (size_t)col * n));
    }
    const size_t xcol = (size_t)min(j0 + 16 + l16, n - 1);
#pragma unroll
    for (int s = 0; s < KS; s++) {
        accr = mfma(x[s].x, y[s].x, accr);
        accr = mfma(-x[s].y, y[s].y, accr);
        acci = mfma(x[s].x, y[s].y, acci);
        acci = mfma(x[s].y, y[s].x, acci);
        const int mc = min(4 * s + kq, k - 1);
        x[s] = gld(B + (xcol + (size_t)mc * n));
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int col = j0 + kq + 4 * r;
        if (rok && col < n) gst_nt(G + ((size_t)row + (size_t)col * n), make_double2(accr[r], acci[r]));
    }
}

template <int K>
__device__ static void wg_rank_k_strips(cplx *G_, int n, const cplx *A_, const cplx *B_, int k)
{
    constexpr int KS = K / 4;
    GAS cplx *G = gp(G_), *A = gp(A_), *B = gp(B_);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nt = (n + 15) >> 4;
    const int l16 = lane & 15, kq = lane >> 4;
    for (int si = w; si < nt; si += nw) {
        const int row = si * 16 + l16;
        const bool rok = row < n;
        const size_t rowc = (size_t)min(row, n - 1);
        const size_t xcol = (size_t)min(l16, n - 1);
        cplx y[KS], x[KS], ga[4], gb[4];
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int m = 4 * s + kq, mc = min(m, k - 1);
            const cplx a = gld(A + (rowc + (size_t)mc * n));
            y[s] = m < k ? a : make_double2(0.0, 0.0);
            x[s] = gld(B + (xcol + (size_t)mc * n));
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            ga[r] = gld_nt(G + (rowc + (size_t)min(kq + 4 * r, n - 1) * n));
            gb[r] = gld_nt(G + (rowc + (size_t)min(16 + kq + 4 * r, n - 1) * n));
        }
        for (int tj = 0; tj < nt; tj += 2) {
            rank_k_tile<KS>(G, B, n, k, row, rok, kq, l16, tj, y, x, ga);
            if (tj + 1 < nt) rank_k_tile<KS>(G, B, n, k, row, rok, kq, l16, tj + 1, y, x, gb);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Metropolis decision of one site (dqmc_update_delay.f90:45-76), evaluated identically by
// every thread of the chain's workgroup.
// ------------------------------------------------------------------------------------------
struct SiteDraw {
    int is, isp, fl;
    cplx db;
};

// the HS tables a site decision reads, copied to LDS once per launch (the same field names as
// DChain so that the decision helpers take either)
struct HsTab {
    int lcomp, nflr;
    double gaml[4];
    cplx delta_bmat[12], phase_ratio[12];
    cplx qf[12];   // gaml(isp)/gaml(is) phase_ratio(iflip, is): the field-only factor of ratiotot
};

__device__ static void hs_tab_load(HsTab &t, const DChain &c)
{
    if (threadIdx.x == 0) {
        t.lcomp = c.lcomp;
        t.nflr = c.nflr;
        for (int k = 0; k < 4; k++) t.gaml[k] = c.gaml[k];
        for (int k = 0; k < 12; k++) { t.delta_bmat[k] = c.delta_bmat[k]; t.phase_ratio[k] = c.phase_ratio[k]; }
        for (int is = 1; is <= c.lcomp; is++)
            for (int fl = 1; fl < c.lcomp; fl++) {
                const int k = (fl - 1) + (c.lcomp - 1) * (is - 1), isp = (is + fl - 1) % c.lcomp + 1;
                t.qf[k] = cscale(c.phase_ratio[k], c.gaml[isp - 1] / c.gaml[is - 1]);
            }
    }
}

// The decision of dqmc_update_delay.f90:55-76 without divisions: ratiotot = ratio1^nflr qf and
// accept when |Re(ratiotot phase)| > u2 |Re(phase)|, which is |ratio_re| > u2 for the reference's
// ratio_re = Re(ratiotot phase) / Re(phase).  Only the direction of phase matters here, so the
// caller may carry it unnormalised.
__device__ __forceinline__ bool site_accept_fast(const HsTab &t, cplx ratio1, const SiteDraw &d, double u2,
                                                 cplx phase, cplx *ratiotot)
{
    cplx rt = ratio1;
    for (int f = 1; f < t.nflr; f++) rt = cmul(rt, ratio1);
    rt = cmul(rt, t.qf[(d.fl - 1) + (t.lcomp - 1) * (d.is - 1)]);
    *ratiotot = rt;
    return fabs(cmul(rt, phase).x) > u2 * fabs(phase.x);
}

template <class Tab>
__device__ __forceinline__ bool site_accept(const Tab &c, cplx ratio1, const SiteDraw &d, double u2,
                                            cplx phase, cplx *ratiotot)
{
    cplx rt = make_double2(1.0, 0.0);
    for (int f = 0; f < c.nflr; f++) rt = cmul(rt, ratio1);
    rt = cscale(rt, c.gaml[d.isp - 1] / c.gaml[d.is - 1]);
    rt = cmul(rt, c.phase_ratio[(d.fl - 1) + (c.lcomp - 1) * (d.is - 1)]);
    double ratio_re = cmul(rt, phase).x / phase.x;
    double ratio_abs = ratio_re < 0.0 ? -ratio_re : ratio_re;
    *ratiotot = rt;
    return ratio_abs > u2;
}

template <class Tab>
__device__ __forceinline__ SiteDraw site_draw(const Tab &c, int is, double u1)
{
    SiteDraw d;
    d.is = is;
    d.fl = (int)ceil(u1 * (c.lcomp - 1));
    d.isp = (is + d.fl - 1) % c.lcomp + 1;
    d.db = c.delta_bmat[(d.fl - 1) + (c.lcomp - 1) * (is - 1)];
    return d;
}

__device__ __forceinline__ cplx phase_after(cplx phase, cplx rt)
{
    double wgt = sqrt(cmul(rt, cconj(rt)).x);
    return cscale(cmul(phase, rt), 1.0 / wgt);
}

// delay update of one slice (src/dqmc_config_u/dqmc_update_delay.f90:39-128); one workgroup per
// chain, the delayed vectors avec / bvec in global memory (any N and depth; the register kernel
// below covers N <= 256).  The slice's HS field and uniforms and the HS tables are staged in LDS
// (the field goes back at the end), so that the site loop reads global memory only for G's row
// and column and the delayed vectors of an accepted site, four steps of the products in flight.
// LDS: diag(G) (n), the accepted site's rows of avec / bvec (2 x 64), uniforms (2n), field (n).
extern "C" __global__ void __launch_bounds__(256)
k_update_delay(const DChain *chains, int nt, int slice_in_pass)
{
    extern __shared__ cplx sh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n, K = c.nublock;
    cplx *diagg = sh;                 // n
    cplx *rowa = sh + n;              // 64: avec(s, 0:ik)
    cplx *rowb = rowa + 64;           // 64: bvec(s, 0:ik)
    double *us = (double *)(rowb + 64);   // 2n
    int *cfs = (int *)(us + 2 * n);   // n
    __shared__ HsTab tab;
    cplx *G = c.G, *avec = c.avec, *bvec = c.bvec;
    GAS cplx *Gg = gp(c.G), *ag = gp(c.avec), *bg = gp(c.bvec);
    int32_t *cf = c.conf + (size_t)(nt - 1) * n;
    const double *u = c.unif + (size_t)2 * n * slice_in_pass;
    cplx phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        diagg[r] = G[(size_t)r + (size_t)r * n];
        cfs[r] = cf[r];
    }
    for (int q = threadIdx.x; q < 2 * n; q += blockDim.x) us[q] = u[q];
    hs_tab_load(tab, c);
    __syncthreads();
    // every load issued so far has landed (the phase above would otherwise keep the compiler
    // waiting on all outstanding memory operations inside the site loop)
    __builtin_amdgcn_s_waitcnt(0);
    int ik = 0;
    double accm = 0.0;
    for (int s = 0; s < n; s++) {
        const SiteDraw d = site_draw(tab, cfs[s], us[2 * s]);
        cplx dg = diagg[s];
        cplx ratio1 = cadd(cmul(make_double2(1.0 - dg.x, -dg.y), d.db), make_double2(1.0, 0.0));
        cplx rt;
        const bool acc = __builtin_amdgcn_readfirstlane((int)site_accept(tab, ratio1, d, us[2 * s + 1], phase, &rt)) != 0;
        if (acc) {
            accm += 1.0;
            phase = phase_after(phase, rt);
            cplx sscl = cdiv(d.db, ratio1);
            for (int m = threadIdx.x; m < ik; m += blockDim.x) {
                rowa[m] = gld(ag + ((size_t)s + (size_t)m * n));
                rowb[m] = gld(bg + ((size_t)s + (size_t)m * n));
            }
            __syncthreads();
            // every thread has read this site's field above: thread 0 records the flip
            if (threadIdx.x == 0) cfs[s] = d.isp;
            for (int r = threadIdx.x; r < n; r += blockDim.x) {
                cplx a = gld(Gg + ((size_t)r + (size_t)s * n));
                cplx b = gld(Gg + ((size_t)s + (size_t)r * n));
                // the sums in the order m = 0, 1, ..., four steps' loads issued together
                int m = 0;
                for (; m + 4 <= ik; m += 4) {
                    cplx av[4], bv[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        av[j] = gld(ag + ((size_t)r + (size_t)(m + j) * n));
                        bv[j] = gld(bg + ((size_t)r + (size_t)(m + j) * n));
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        a = cadd(a, cmul(rowb[m + j], av[j]));
                        b = cadd(b, cmul(rowa[m + j], bv[j]));
                    }
                }
                for (; m < ik; m++) {
                    a = cadd(a, cmul(rowb[m], gld(ag + ((size_t)r + (size_t)m * n))));
                    b = cadd(b, cmul(rowa[m], gld(bg + ((size_t)r + (size_t)m * n))));
                }
                a = cmul(a, sscl);
                if (r == s) b.x -= 1.0;
                diagg[r] = cadd(diagg[r], cmul(a, b));
                avec[(size_t)r + (size_t)ik * n] = a;
                bvec[(size_t)r + (size_t)ik * n] = b;
            }
            ik++;
            __syncthreads();
        }
        if (ik == K || s == n - 1) {
            // G += avec bvec^T on the prefetching MFMA strips at the depth of the accepted count
            if (ik > 0) {
                if (ik <= 16) {
                    wg_rank_k_strips<16>(G, n, avec, bvec, ik);
                } else if (ik <= 32) {
                    wg_rank_k_strips<32>(G, n, avec, bvec, ik);
                } else if (ik <= 48) {
                    wg_rank_k_strips<48>(G, n, avec, bvec, ik);
                } else {
                    wg_rank_k_strips<32>(G, n, avec, bvec, 32);
                    __syncthreads();
                    wg_rank_k_strips<32>(G, n, avec + (size_t)32 * n, bvec + (size_t)32 * n, ik - 32);
                }
            }
            __syncthreads();
            ik = 0;
            if (s < n - 1)
                for (int r = threadIdx.x; r < n; r += blockDim.x) diagg[r] = gld(Gg + ((size_t)r + (size_t)r * n));
            __syncthreads();
            __builtin_amdgcn_s_waitcnt(0);
        }
    }
    for (int r = threadIdx.x; r < n; r += blockDim.x) cf[r] = cfs[r];
    if (threadIdx.x == 0) {
        c.sc[SC_PHASE] = phase.x;
        c.sc[SC_PHASE + 1] = phase.y;
        c.sc[SC_ACC] += accm;
        c.sc[SC_TRIED] += n;
    }
}

// the factors of one site's draw that do not depend on G (the submatrix kernel evaluates them
// for a staged block of sites at once, off the per-site critical path)
struct SubSite {
    cplx db, idb, f, pr;   // delta, 1 / delta, 1 / (1 + delta), phase_ratio(fl, is)
    double gr, u2;         // gaml(isp) / gaml(is), the acceptance uniform
    int isp;
};

// lane l's complex value as a wave-uniform (scalar) operand; l must be wave-uniform
__device__ __forceinline__ cplx readlane_c(cplx v, int l)
{
    auto rd = [l](double x) {
        const long long b = __double_as_longlong(x);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    };
    return make_double2(rd(v.x), rd(v.y));
}

// submatrix update of one slice (src/dqmc_config_u/dqmc_update_submatrix.f90:55-162).
// Gcuta is kept implicitly as G diag(cs); the flush writes (G + (G(:,p) Gamma^-1) G(p,:)) diag(cs),
// where cs differs from 1 only at the accepted sites p.
// G does not change between flushes, so the columns and rows of G at the next dss sites are
// staged in LDS (as in the delay kernel): G(s, p), G(p, s), G(s, s) and the column and row of G
// an accepted site records come from LDS.  Per site, v1 = G(s, p) Gamma^-1 and
// v4 = Gamma^-1 G(p, s) are formed before the decision, each by two waves over halves of p
// (partial sums in LDS, double-buffered by site parity so that no barrier separates one site's
// reads from the next site's writes); the decision needs v = v1 . G(p, s) only.
// The site loop itself reads and writes no global memory (its draws are staged with the block
// and the accepted sites' fields, columns and rows are taken at the flush).
// Scratch: avec = G(:,p) columns, W1 = G(p,:)^T, bvec = Gmat1, W2 = Gamma^-1 (ld ik).
// LDS: Gamma^-1 (K (K+1)), the K column scales, dss columns and dss rows of G.
extern "C" __global__ void __launch_bounds__(256)
k_update_submatrix(const DChain *chains, int nt, int slice_in_pass, int dss)
{
    extern __shared__ cplx sh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n, K = c.nublock, KL = K + 1;
    dss = max(1, min(dss, 8));        // the staged draws hold 8 sites
    cplx *gi = sh;                    // Gamma^-1, column-major, leading dimension K + 1 (lanes reading
                                      // one entry of K columns each fall on different banks)
    cplx *csk = gi + KL * K;          // K: cs of the accepted sites
    cplx *gcol = csk + K;             // dss x n: G(:, s0 + j)
    cplx *grow = gcol + (size_t)dss * n;  // dss x n: G(s0 + j, :)
    __shared__ int pvec[256];
    __shared__ cplx part[2][256];     // [parity][wave * 64 + i]: v1 halves (waves 0, 1), v4 halves (2, 3)
    __shared__ cplx vsum[2][2];       // [parity][wave]: v halves
    __shared__ HsTab tab;
    __shared__ SubSite sfac[8];       // the draws of the staged sites
    __shared__ int ispk[64];          // the new HS field of the accepted sites, written at the flush
    cplx *G = c.G, *gm1 = c.avec, *gmat1 = c.bvec, *gm2t = c.W1;
    GAS cplx *Gg = gp(c.G);
    int32_t *cf = c.conf + (size_t)(nt - 1) * n;
    GAS const int32_t *cfg = gp(cf);
    GAS const double *ug = gp(c.unif + (size_t)2 * n * slice_in_pass);
    cplx phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    hs_tab_load(tab, c);
    __syncthreads();
    // every load issued so far has landed: without this the compiler cannot tell, inside the
    // site loop, that the phase loaded above is no longer in flight, and waits for all
    // outstanding memory operations (the accepted sites' stores) at each use
    __builtin_amdgcn_s_waitcnt(0);
    int ik = 0, s0 = -1;
    double accm = 0.0;
    for (int s = 0; s < n; s++) {
        if (s0 < 0 || s >= s0 + dss) {
            // stage G(:, s..s+nb) (one contiguous run) and G(s..s+nb, :) (nb-element runs);
            // valid until the next flush changes G.  A lane may still read the previous block.
            if (s0 >= 0) __syncthreads();
            s0 = s;
            const int nb = min(dss, n - s), tot = nb * n;
            GAS cplx *Gc = Gg + (size_t)s * n;
            // the sites' draws come with the block (thread t: site s + t), so that the site loop
            // itself waits on no global load (a wait there would also wait for stores in flight)
            int cfv = 1;
            double u1v = 0.5, u2v = 0.0;
            if (threadIdx.x < nb) {
                cfv = cfg[s + threadIdx.x];
                u1v = ug[2 * (s + threadIdx.x)];
                u2v = ug[2 * (s + threadIdx.x) + 1];
            }
            // 8 column and 8 row elements per thread in flight (clamped addresses)
            for (int base = threadIdx.x; base < tot; base += 8 * (int)blockDim.x) {
                cplx a[8], b[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int e = min(base + u * (int)blockDim.x, tot - 1);
                    a[u] = gld(Gc + e);
                    b[u] = gld(Gg + ((size_t)(s + e % nb) + (size_t)(e / nb) * n));
                }
                if (base == (int)threadIdx.x && threadIdx.x < nb) {
                    const SiteDraw dd = site_draw(tab, cfv, u1v);
                    SubSite q;
                    q.db = dd.db;
                    q.idb = cinv(dd.db);
                    q.f = cdiv(make_double2(1.0, 0.0), cadd(make_double2(1.0, 0.0), dd.db));
                    q.pr = tab.phase_ratio[(dd.fl - 1) + (tab.lcomp - 1) * (dd.is - 1)];
                    q.gr = tab.gaml[dd.isp - 1] / tab.gaml[dd.is - 1];
                    q.u2 = u2v;
                    q.isp = dd.isp;
                    sfac[threadIdx.x] = q;
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int e = base + u * (int)blockDim.x;
                    if (e < tot) {
                        gcol[e] = a[u];
                        grow[(e % nb) * n + e / nb] = b[u];
                    }
                }
            }
            __syncthreads();
        }
        const cplx *gcs = gcol + (size_t)(s - s0) * n, *grs = grow + (size_t)(s - s0) * n;
        const int par = s & 1;
        const SubSite q = sfac[s - s0];
        {
            // waves 0, 1: v1(i) = sum_m G(s, p_m) Gamma^-1(m, i); waves 2, 3: v4(i) =
            // sum_m Gamma^-1(i, m) G(p_m, s); m over [0, h) (even waves) or [h, ik) (odd)
            // lane j holds G(s, p_j) (v1 waves) or G(p_j, s), broadcast to the wave by readlane;
            // lanes past ik compute on clamped indices and store nothing
            const int i = lane, h = (ik + 1) >> 1;
            const int mb = (w & 1) ? h : 0, me = (w & 1) ? ik : h;
            const int pj = i < ik ? pvec[i] : s, ic = min(i, K - 1);
            const cplx bvi = gcs[pj], own = w < 2 ? grs[pj] : bvi;
            cplx a = make_double2(0.0, 0.0), pv = make_double2(0.0, 0.0);
            // four LDS reads in flight per step; the sum keeps the order m = mb, mb + 1, ...
            const int gs = w < 2 ? 1 : KL, go = w < 2 ? ic * KL : ic;   // Gamma^-1 entry (m, ic) or (ic, m)
            int m = mb;
            for (; m + 4 <= me; m += 4) {
                const cplx g0 = gi[go + m * gs], g1 = gi[go + (m + 1) * gs], g2 = gi[go + (m + 2) * gs],
                           g3 = gi[go + (m + 3) * gs];
                if (w < 2) {
                    a = cadd(a, cmul(readlane_c(own, m), g0));
                    a = cadd(a, cmul(readlane_c(own, m + 1), g1));
                    a = cadd(a, cmul(readlane_c(own, m + 2), g2));
                    a = cadd(a, cmul(readlane_c(own, m + 3), g3));
                } else {
                    a = cadd(a, cmul(g0, readlane_c(own, m)));
                    a = cadd(a, cmul(g1, readlane_c(own, m + 1)));
                    a = cadd(a, cmul(g2, readlane_c(own, m + 2)));
                    a = cadd(a, cmul(g3, readlane_c(own, m + 3)));
                }
            }
            for (; m < me; m++) {
                const cplx g = gi[go + m * gs];
                a = w < 2 ? cadd(a, cmul(readlane_c(own, m), g)) : cadd(a, cmul(g, readlane_c(own, m)));
            }
            if (w < 2 && i < ik) pv = cmul(a, bvi);
            if (i < ik) part[par][w * 64 + i] = a;
            if (w < 2) {
                pv.x = wave_sum_dpp(pv.x);
                pv.y = wave_sum_dpp(pv.y);
                if (i == 0) vsum[par][w] = pv;
            }
        }
        // G(s, s) is in the staged column; read before the barrier, after which the next
        // block may be staged
        const cplx gss = gcs[s];
        __syncthreads();
        const cplx v = cadd(vsum[par][0], vsum[par][1]);
        cplx v2 = make_double2(v.x + gss.x - 1.0, v.y + gss.y);
        cplx ratio1 = csub(make_double2(1.0, 0.0), cmul(v2, q.db));
        // the decision of site_accept with its draw-only factors precomputed (the same
        // operations in the same order); every thread evaluates it, and reading it from lane 0
        // makes the branch scalar
        cplx rt = make_double2(1.0, 0.0);
        for (int q = 0; q < tab.nflr; q++) rt = cmul(rt, ratio1);
        rt = cmul(cscale(rt, q.gr), q.pr);
        const double ratio_re = cmul(rt, phase).x / phase.x;
        const bool acc = __builtin_amdgcn_readfirstlane((int)(fabs(ratio_re) > q.u2)) != 0;
        if (acc) {
            accm += 1.0;
            phase = phase_after(phase, rt);
            cplx den = make_double2(-v.x + 1.0 - gss.x, -v.y - gss.y);
            den = cadd(den, q.idb);
            cplx v3 = cinv(den);
            // Gamma^-1 += (v4 v3) v1 and its new border: lane i of every wave holds v4(i) v3 and
            // v1(i); wave w updates the columns m = w, w + 4, ... with v1(m) broadcast by readlane
            const cplx *pp = part[par];
            const int i = lane, ic = min(i, K - 1);
            const cplx v1own = cadd(pp[ic], pp[64 + ic]);
            const cplx v4v3 = cmul(cadd(pp[128 + ic], pp[192 + ic]), v3);
#pragma unroll 4
            for (int m = w; m < ik; m += 4) {
                const cplx v1m = readlane_c(v1own, m);
                if (i < ik) gi[i + m * KL] = cadd(gi[i + m * KL], cmul(v4v3, v1m));
            }
            if (w == 0 && i < ik) {
                gi[ik + i * KL] = cmul(v3, v1own);
                gi[i + ik * KL] = v4v3;
            }
            if (threadIdx.x == 0) {
                gi[ik + ik * KL] = v3;
                pvec[ik] = s;
                ispk[ik] = q.isp;
                // cs(s) is 1 until its site is accepted (reset at every flush)
                csk[ik] = cmul(make_double2(1.0, 0.0), q.f);
            }
            ik++;
            __syncthreads();
        }
        if (ik == K || s == n - 1) {
            if (ik > 0) {
                // Gmat1 = G(:,p) Gamma^-1 as an MFMA product (Gamma^-1 staged through W2, which
                // the update does not use)
                cplx *gig = c.W2;
                for (int e = threadIdx.x; e < ik * ik; e += blockDim.x) {
                    const int i = e % ik, j = e / ik;
                    gig[i + (size_t)j * ik] = gi[i + j * KL];  // ld ik: ik^2 entries
                }
                if (threadIdx.x < ik) cf[pvec[threadIdx.x]] = ispk[threadIdx.x];
                // the accepted sites' columns G(:, p) and rows G(p, :) (the site loop itself
                // touches no global memory), 8 + 8 elements per thread in flight
                {
                    const int tot = ik * n;
                    for (int base = threadIdx.x; base < tot; base += 8 * (int)blockDim.x) {
                        cplx a[8], b[8];
#pragma unroll
                        for (int u = 0; u < 8; u++) {
                            const int e = min(base + u * (int)blockDim.x, tot - 1), i = e / n, r = e - i * n;
                            a[u] = gld(Gg + ((size_t)r + (size_t)pvec[i] * n));
                            b[u] = gld(Gg + ((size_t)pvec[i] + (size_t)r * n));
                        }
#pragma unroll
                        for (int u = 0; u < 8; u++) {
                            const int e = base + u * (int)blockDim.x;
                            if (e < tot) {
                                gm1[e] = a[u];      // gm1(r, i) = G(r, p_i), ld n
                                gm2t[e] = b[u];     // gm2t(r, i) = G(p_i, r), ld n
                            }
                        }
                    }
                }
                __syncthreads();
                for (int tm = 0; tm < n; tm += GT)
                    zgemm_tile<0, 0>(gm1, n, gig, ik, gmat1, n, n, ik, ik, tm, 0, 1.0, 0.0, 0.0, 0.0);
                __syncthreads();
                // G += Gmat1 G(p,:): one pass over G up to 48 sites (fragment depth the next
                // multiple of 16), two passes of up to 32 beyond (a depth-64 fragment would not
                // fit in registers)
                if (ik <= 16) {
                    wg_rank_k_strips<16>(G, n, gmat1, gm2t, ik);
                } else if (ik <= 32) {
                    wg_rank_k_strips<32>(G, n, gmat1, gm2t, ik);
                } else if (ik <= 48) {
                    wg_rank_k_strips<48>(G, n, gmat1, gm2t, ik);
                } else {
                    wg_rank_k_strips<32>(G, n, gmat1, gm2t, 32);
                    __syncthreads();
                    wg_rank_k_strips<32>(G, n, gmat1 + (size_t)32 * n, gm2t + (size_t)32 * n, ik - 32);
                }
                __syncthreads();
                // the columns of the accepted sites times their cs, 8 elements per thread in flight
                const int tot = ik * n;
                for (int base = threadIdx.x; base < tot; base += 8 * (int)blockDim.x) {
                    cplx v[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int e = min(base + u * (int)blockDim.x, tot - 1), i = e / n;
                        v[u] = gld(Gg + ((size_t)(e - i * n) + (size_t)pvec[i] * n));
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int e = base + u * (int)blockDim.x, i = e / n;
                        if (e < tot) gst(Gg + ((size_t)(e - i * n) + (size_t)pvec[i] * n), cmul(v[u], csk[i]));
                    }
                }
            }
            __syncthreads();
            // the flush's last prefetches are never consumed; retiring them here keeps the
            // compiler from waiting on them (and on every store in flight) inside the site loop
            __builtin_amdgcn_s_waitcnt(0);
            ik = 0;
            s0 = -1;                  // G changed: stage again
        }
    }
    if (threadIdx.x == 0) {
        c.sc[SC_PHASE] = phase.x;
        c.sc[SC_PHASE + 1] = phase.y;
        c.sc[SC_ACC] += accm;
        c.sc[SC_TRIED] += n;
    }
}

// ------------------------------------------------------------------------------------------
// Equal-time measurement (src/dqmc_measure/equaltime_measure.f90:1-69, zkint :71-114,
// zint_u :116-125, zspsm :283-299 -> Fourier component at the staggering vector).
// One workgroup per chain; the G(i,j) G(j,i) sum is taken over 32x32 tile pairs staged in LDS.
// ------------------------------------------------------------------------------------------
__device__ static double block_sum(double v, double *red)
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

// partial sums of S = sum_ij s_i s_j G(i,j) G(j,i) over one 32x32 tile pair (bi <= bj) per
// workgroup, written to rwork[2 tp]
extern "C" __global__ void __launch_bounds__(256)
k_measure_tiles(const DChain *chains)
{
    __shared__ cplx ta[32][33], tb[32][33];
    __shared__ double red[16];
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const int nb = (n + 31) / 32;
    int tp = blockIdx.x;
    if (tp >= nb * (nb + 1) / 2) return;
    int bi = 0;
    while (tp >= nb - bi) { tp -= nb - bi; bi++; }
    const int bj = bi + tp;
    // the eight G entries of a thread (four per tile, 256 threads) and its stag values are loaded
    // in one batch through global-space pointers (clamped addresses), ahead of the LDS stores
    GAS cplx *G = gp(c.G);
    GAS const double *stg = gp(c.stag);
    constexpr int MT_U = 1024 / 256;
    cplx va[MT_U], vb[MT_U];
    double si_[MT_U], sj_[MT_U];
#pragma unroll
    for (int u = 0; u < MT_U; u++) {
        const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
        const int r1 = bi * 32 + q, c1 = bj * 32 + p, r2 = bj * 32 + q, c2 = bi * 32 + p;
        va[u] = gld(G + ((size_t)min(r1, n - 1) + (size_t)min(c1, n - 1) * n));
        vb[u] = gld(G + ((size_t)min(r2, n - 1) + (size_t)min(c2, n - 1) * n));
        si_[u] = stg[min(r1, n - 1)];
        sj_[u] = stg[min(c1, n - 1)];
    }
#pragma unroll
    for (int u = 0; u < MT_U; u++) {
        const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
        const int r1 = bi * 32 + q, c1 = bj * 32 + p, r2 = bj * 32 + q, c2 = bi * 32 + p;
        ta[q][p] = (r1 < n && c1 < n) ? va[u] : make_double2(0.0, 0.0);
        tb[q][p] = (r2 < n && c2 < n) ? vb[u] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const double wgt = (bi == bj) ? 1.0 : 2.0;
    double ssr = 0.0, ssi = 0.0;
#pragma unroll
    for (int u = 0; u < MT_U; u++) {
        const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
        const int i = bi * 32 + q, j = bj * 32 + p;
        if (i < n && j < n) {
            const cplx pr = cmul(ta[q][p], tb[p][q]);   // G(i,j) G(j,i)
            const double sg = si_[u] * sj_[u] * wgt;
            ssr += sg * pr.x;
            ssi += sg * pr.y;
        }
    }
    ssr = block_sum(ssr, red);
    ssi = block_sum(ssi, red);
    if (threadIdx.x == 0) {
        c.rwork[2 * blockIdx.x] = ssr;
        c.rwork[2 * blockIdx.x + 1] = ssi;
    }
}

// diagonal and bond terms, the tile partials, and the zphi-weighted bin sums; one workgroup per chain
extern "C" __global__ void __launch_bounds__(256)
k_measure(const DChain *chains)
{
    __shared__ double red[16];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    const cplx *G = c.G;
    double zne_r = 0, zne_i = 0, zu_r = 0, zu_i = 0, zk = 0, trr = 0, tri = 0, ssr = 0, ssi = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        cplx g = G[(size_t)i + (size_t)i * n];
        cplx gc = make_double2(1.0 - g.x, -g.y);
        zne_r += gc.x; zne_i += gc.y;
        cplx g2 = cmul(gc, gc);
        zu_r += g2.x; zu_i += g2.y;
        trr += g.x; tri += g.y;
    }
    const int nbond = c.nn_nf * c.nn_lf;
    for (int b = threadIdx.x; b < nbond; b += blockDim.x) {
        int i0 = c.bonds[2 * b], i1 = c.bonds[2 * b + 1];
        // gfc(i0,i1) = delta - gf(i1,i0); kinetic term -2 Re(gfc(i0,i1))
        double gcr = (i0 == i1 ? 1.0 : 0.0) - G[(size_t)i1 + (size_t)i0 * n].x;
        zk += -2.0 * gcr;
    }
    const int nb = (n + 31) / 32, ntp = nb * (nb + 1) / 2;
    for (int t = threadIdx.x; t < ntp; t += blockDim.x) { ssr += c.rwork[2 * t]; ssi += c.rwork[2 * t + 1]; }
    zne_r = block_sum(zne_r, red); zne_i = block_sum(zne_i, red);
    zu_r = block_sum(zu_r, red); zu_i = block_sum(zu_i, red);
    zk = block_sum(zk, red);
    trr = block_sum(trr, red); tri = block_sum(tri, red);
    ssr = block_sum(ssr, red); ssi = block_sum(ssi, red);
    if (threadIdx.x == 0) {
        double *sc = c.sc;
        cplx phase = make_double2(sc[SC_PHASE], sc[SC_PHASE + 1]);
        double sg = phase.x < 0 ? -1.0 : 1.0;
        cplx zphi = cscale(phase, sg / phase.x);
        double fac = 1.0 - 1.0 / (double)(c.nflr * c.nflr);
        cplx ss = cscale(make_double2(trr - ssr, tri - ssi), fac / (double)c.lq);
        cplx v[5];
        v[0] = make_double2(sg, 0.0);
        v[1] = cmul(cscale(make_double2(zne_r, zne_i), 1.0 / (double)c.lq), zphi);
        v[2] = cmul(make_double2(zk * c.kin_weight, 0.0), zphi);
        v[3] = cmul(make_double2(zu_r, zu_i), zphi);
        v[4] = cmul(ss, zphi);
        sc[SC_NOBS] += 1.0;
        for (int k = 0; k < 5; k++) { sc[SC_OBS + 2 * k] += v[k].x; sc[SC_OBS + 2 * k + 1] += v[k].y; }
    }
}

// ------------------------------------------------------------------------------------------
// Householder QR (zgeqp3 / zgeqrf semantics, LAPACK zlarfg reflectors), one workgroup per
// chain, in place on matrix `sel`; tau -> c.tau, jpvt -> c.jpvt (1-based), log det Q -> ldq.
// pivot = 1: column pivoting on downdated partial norms with the tol3z recompute rule.
// ------------------------------------------------------------------------------------------
__device__ static void block_argmax(double v, int idx, double *rv, int *ri, double *outv, int *outi)
{
    // first index of the largest value
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    wave_argmax_dpp(v, idx);
    __syncthreads();
    if (lane == 0) { rv[w] = v; ri[w] = idx; }
    __syncthreads();
    double bv = rv[0];
    int bi = ri[0];
    for (int i = 1; i < nw; i++)
        if (rv[i] > bv || (rv[i] == bv && ri[i] < bi)) { bv = rv[i]; bi = ri[i]; }
    *outv = bv;
    *outi = bi;
}

// One wave's share of rows r0 + lane + 64 q (q < NQ) of a column, loaded with clamped addresses
// in one batch (entries past n are never used).
template <int NQ>
__device__ __forceinline__ void qr_col_load(GAS cplx *col, int r0, int n, int lane, cplx (&x)[NQ])
{
#pragma unroll
    for (int q = 0; q < NQ; q++) x[q] = gld(col + min(r0 + lane + 64 * q, n - 1));
}

// H^H = I - conj(tau) v v^H on one trailing column held in registers (rows i + lane + 64 q), its
// norm downdate as zlaqp2 (the vn1/vn2 recomputation below tol3z)
template <int NQ>
__device__ __forceinline__ void qr_apply_col(GAS cplx *a, const cplx (&x)[NQ], int i, int j, int n, int lane, cplx tc,
                                             const cplx *v, bool pivot, double *vn1, double *vn2, double tol3z)
{
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int r = i + lane + 64 * q;
        if (r < n) {
            const cplx vv = v[r], aa = x[q];
            sr += vv.x * aa.x + vv.y * aa.y;     // conj(v) a
            si += vv.x * aa.y - vv.y * aa.x;
        }
    }
    sr = wave_sum_dpp(sr);
    si = wave_sum_dpp(si);
    const cplx f = cmul(tc, make_double2(sr, si));
    double nrm2 = 0.0;
    cplx ai_new = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int r = i + lane + 64 * q;
        if (r < n) {
            const cplx aa = csub(x[q], cmul(v[r], f));
            gst(a + r, aa);
            if (r > i) nrm2 += cabs2(aa);
            else ai_new = aa;   // r == i happens on lane 0 only
        }
    }
    if (pivot) {
        nrm2 = wave_sum_dpp(nrm2);
        if (lane == 0 && vn1[j] != 0.0) {
            double t = sqrt(cabs2(ai_new)) / vn1[j];
            t = 1.0 - t * t;
            if (t < 0.0) t = 0.0;
            double t2 = t * (vn1[j] / vn2[j]) * (vn1[j] / vn2[j]);
            if (t2 <= tol3z) {
                vn1[j] = (i < n - 1) ? sqrt(nrm2) : 0.0;
                vn2[j] = vn1[j];
            } else {
                vn1[j] *= sqrt(t);
            }
        }
    }
}

// Householder QR, column-at-a-time, with column pivoting (pivot = 1, zgeqp3's choice by the
// downdated partial norms) or in the given order (pivot = 0); R above the diagonal, the
// reflectors below, tau to c.tau, log det Q to ldq[out_pt].  Every column a wave works on is
// read once into registers (NQ = rows / 64), and the wave's next trailing column is loaded while
// the current one is reduced.
template <int NQ>
__global__ void __launch_bounds__(1024)
k_qr(const DChain *chains, int sel, int pivot, int out_pt)
{
    extern __shared__ double qrsh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    double *vn1 = qrsh, *vn2 = qrsh + n;
    cplx *v = (cplx *)(qrsh + 2 * n);          // n
    __shared__ double rv[16];
    __shared__ int ri[16];
    __shared__ cplx s_tau;
    GAS cplx *A = gp(chain_mat(c, sel, 0));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // the unpivoted call keeps the prepivot order that k_prepivot left in jpvt
    if (pivot)
        for (int j = threadIdx.x; j < n; j += blockDim.x) c.jpvt[j] = j + 1;
    if (pivot) {
        for (int j = w; j < n; j += nw) {
            cplx x[NQ];
            qr_col_load<NQ>(A + (size_t)j * n, 0, n, lane, x);
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (lane + 64 * q < n) s += cabs2(x[q]);
            s = wave_sum_dpp(s);
            if (lane == 0) { vn1[j] = sqrt(s); vn2[j] = vn1[j]; }
        }
    }
    __syncthreads();
    const double tol3z = sqrt(2.220446049250313e-16);
    for (int i = 0; i < n; i++) {
        if (pivot) {
            double bv = -1.0;
            int bi = i;
            for (int j = i + threadIdx.x; j < n; j += blockDim.x) {
                double a = fabs(vn1[j]);
                if (a > bv) { bv = a; bi = j; }
            }
            double mv;
            int p;
            block_argmax(bv, bi, rv, ri, &mv, &p);
            if (p != i) {
                for (int r = threadIdx.x; r < n; r += blockDim.x) {
                    const cplx t = gld(A + ((size_t)r + (size_t)p * n)), u = gld(A + ((size_t)r + (size_t)i * n));
                    gst(A + ((size_t)r + (size_t)p * n), u);
                    gst(A + ((size_t)r + (size_t)i * n), t);
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    int32_t it = c.jpvt[p]; c.jpvt[p] = c.jpvt[i]; c.jpvt[i] = it;
                    vn1[p] = vn1[i];
                    vn2[p] = vn2[i];
                }
            }
            __syncthreads();
        }
        // reflector of column i (zlarfg), wave 0
        if (w == 0) {
            GAS cplx *col = A + (size_t)i * n;
            cplx x[NQ];
            qr_col_load<NQ>(col, i + 1, n, lane, x);
            const cplx alpha = gld(col + i);
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (i + 1 + lane + 64 * q < n) s += cabs2(x[q]);
            s = wave_sum_dpp(s);
            double xnorm = sqrt(s);
            cplx tau = make_double2(0.0, 0.0);
            cplx scal = make_double2(0.0, 0.0);
            double beta = 0.0;
            bool triv = (xnorm == 0.0 && alpha.y == 0.0);
            if (!triv) {
                double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + s);
                beta = alpha.x >= 0 ? -nrm : nrm;
                tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
                scal = cinv(make_double2(alpha.x - beta, alpha.y));
            }
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int r = i + 1 + lane + 64 * q;
                if (r < n) {
                    const cplx xs = triv ? x[q] : cmul(x[q], scal);
                    gst(col + r, xs);
                    v[r] = xs;
                }
            }
            if (lane == 0) {
                if (!triv) gst(col + i, make_double2(beta, 0.0));
                v[i] = make_double2(1.0, 0.0);
                s_tau = tau;
                c.tau[i] = tau;
            }
        }
        __syncthreads();
        const cplx tc = cconj(s_tau);
        // apply H^H to the trailing columns, one wave per column; the wave's next column is
        // in flight while the current one is reduced
        int j = i + 1 + w;
        if (NQ > 4) {
            // (two register sets of more than four slots would spill at 1024 threads)
            for (; j < n; j += nw) {
                cplx xa[NQ];
                qr_col_load<NQ>(A + (size_t)j * n, i, n, lane, xa);
                qr_apply_col<NQ>(A + (size_t)j * n, xa, i, j, n, lane, tc, v, pivot, vn1, vn2, tol3z);
            }
        } else if (j < n) {
            cplx xa[NQ], xb[NQ];
            qr_col_load<NQ>(A + (size_t)j * n, i, n, lane, xa);
            for (; j < n; j += 2 * nw) {
                const int j2 = j + nw;
                if (j2 < n) qr_col_load<NQ>(A + (size_t)j2 * n, i, n, lane, xb);
                qr_apply_col<NQ>(A + (size_t)j * n, xa, i, j, n, lane, tc, v, pivot, vn1, vn2, tol3z);
                if (j2 < n) {
                    if (j2 + nw < n) qr_col_load<NQ>(A + (size_t)(j2 + nw) * n, i, n, lane, xa);
                    qr_apply_col<NQ>(A + (size_t)j2 * n, xb, i, j2, n, lane, tc, v, pivot, vn1, vn2, tol3z);
                }
            }
        }
        __syncthreads();
    }
    // log det Q, summed by the whole workgroup
    double sr = 0.0, si = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const cplx t = gld(gp(c.tau) + q);
        if (t.x != 0.0 || t.y != 0.0) {
            const cplx qq = cdiv(cscale(t, 2.0 * t.x), cmul(t, cconj(t)));
            const cplx l = clog_(make_double2(1.0 - qq.x, -qq.y));
            sr += l.x;
            si += l.y;
        }
    }
    __shared__ double red[16];
    sr = block_sum(sr, red);
    si = block_sum(si, red);
    if (threadIdx.x == 0) c.ldq[out_pt] = make_double2(sr, si);
}

template __global__ void k_qr<4>(const DChain *, int, int, int);
template __global__ void k_qr<8>(const DChain *, int, int, int);
template __global__ void k_qr<16>(const DChain *, int, int, int);

// Q = H_1 ... H_n from the reflectors in matrix `sel` (zung2r), written to chain point U[pt]
// (dest 0) or Us (dest 1)
extern "C" __global__ void __launch_bounds__(1024)
k_formq(const DChain *chains, int sel, int out_pt)
{
    extern __shared__ double qsh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    cplx *v = (cplx *)qsh;
    const cplx *A = chain_mat(c, sel, 0);
    cplx *Q = c.U + (size_t)n * n * out_pt;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (size_t e = threadIdx.x; e < (size_t)n * n; e += blockDim.x) Q[e] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int i = n - 1; i >= 0; i--) {
        const cplx tau = c.tau[i];
        for (int r = i + threadIdx.x; r < n; r += blockDim.x)
            v[r] = (r == i) ? make_double2(1.0, 0.0) : A[(size_t)r + (size_t)i * n];
        __syncthreads();
        for (int j = i + 1 + w; j < n; j += nw) {
            cplx *q = Q + (size_t)j * n;
            double sr = 0.0, si = 0.0;
            for (int r = i + lane; r < n; r += 64) {
                cplx vv = v[r], qq = q[r];
                sr += vv.x * qq.x + vv.y * qq.y;
                si += vv.x * qq.y - vv.y * qq.x;
            }
            sr = wave_sum_dpp(sr);
            si = wave_sum_dpp(si);
            cplx f = cmul(tau, make_double2(sr, si));
            for (int r = i + lane; r < n; r += 64) q[r] = csub(q[r], cmul(v[r], f));
        }
        __syncthreads();
        for (int r = i + threadIdx.x; r < n; r += blockDim.x) {
            cplx val = (r == i) ? make_double2(1.0 - tau.x, -tau.y)
                                : make_double2(-(tau.x * v[r].x - tau.y * v[r].y), -(tau.x * v[r].y + tau.y * v[r].x));
            Q[(size_t)r + (size_t)i * n] = val;
        }
        __syncthreads();
    }
}

// D = Re diag(R), V = D^-1 triu(R) P^T (first point, column-pivoted QR) -> D[pt], V[pt]
extern "C" __global__ void __launch_bounds__(256)
k_v_from_qrcp(const DChain *chains, int sel, int pt)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    int i = (int)(e % n), k = (int)(e / n);
    const cplx *R = chain_mat(c, sel, 0);
    double di = R[(size_t)i + (size_t)i * n].x;
    cplx val = (i <= k) ? cscale(R[(size_t)i + (size_t)k * n], 1.0 / di) : make_double2(0.0, 0.0);
    c.V[(size_t)n * n * pt + (size_t)i + (size_t)(c.jpvt[k] - 1) * n] = val;
    if (k == 0) c.D[(size_t)n * pt + i] = di;
}

// prepivot: Btmp = Bd diag(Dprev); columns sorted by decreasing norm (s_zmcpt); the sorted
// matrix goes to W2, the order to jpvt.  One workgroup per chain; the selection sort (first
// maximum wins, the displaced entry takes the maximum's slot) runs in wave 0 without barriers.
// A wave scales two columns at a time and the copy moves eight entries per thread at a time,
// every batch's loads issued before its first store (clamped addresses).
extern "C" __global__ void __launch_bounds__(1024)
k_prepivot(const DChain *chains, int dpt)
{
    extern __shared__ double psh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    volatile double *vn = psh;
    volatile int *jp = (volatile int *)(psh + n);
    const double *Dp = c.D + (size_t)n * dpt;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    GAS cplx *B = gp(c.W1);
    for (int j = w; j < n; j += 2 * nw) {
        const int j2 = j + nw;
        const bool two = j2 < n;
        const double d1 = Dp[j], d2 = Dp[two ? j2 : j];
        GAS cplx *B1 = B + (size_t)j * n, *B2 = B + (size_t)(two ? j2 : j) * n;
        double s1 = 0.0, s2 = 0.0;
        for (int r0 = lane; r0 < n; r0 += 4 * 64) {
            cplx a1[4], a2[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = min(r0 + 64 * u, n - 1);
                a1[u] = gld(B1 + r);
                a2[u] = gld(B2 + r);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = r0 + 64 * u;
                if (r < n) {
                    const cplx b1 = cscale(a1[u], d1);
                    gst(B1 + r, b1);
                    s1 += cabs2(b1);
                    if (two) {
                        const cplx b2 = cscale(a2[u], d2);
                        gst(B2 + r, b2);
                        s2 += cabs2(b2);
                    }
                }
            }
        }
        s1 = wave_sum_dpp(s1);
        s2 = wave_sum_dpp(s2);
        if (lane == 0) {
            vn[j] = sqrt(s1);
            jp[j] = j + 1;
            if (two) { vn[j2] = sqrt(s2); jp[j2] = j2 + 1; }
        }
    }
    __syncthreads();
    if (w == 0) {
        for (int k = 0; k < n; k++) {
            double bv = -1.0;
            int bi = k;
            for (int j = k + lane; j < n; j += 64) {
                double a = fabs(vn[j]);
                if (a > bv) { bv = a; bi = j; }
            }
            wave_argmax_dpp(bv, bi);
            if (lane == 0 && bi != k) {
                int it = jp[bi]; jp[bi] = jp[k]; jp[k] = it;
                vn[bi] = vn[k];
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) c.jpvt[j] = jp[j];
    GAS cplx *W2 = gp(c.W2);
    const int nn = n * n;
    for (int base = threadIdx.x; base < nn; base += 8 * (int)blockDim.x) {
        cplx v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = min(base + u * (int)blockDim.x, nn - 1), r = e % n, k = e / n;
            v[u] = gld(B + ((size_t)r + (size_t)(jp[k] - 1) * n));
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = base + u * (int)blockDim.x;
            if (e < nn) gst(W2 + e, v[u]);
        }
    }
}

// after the prepivoted QR of W2: D[pt] = Re diag R, W3 = D^-1 triu(R), W4 = P^T V[vpt]
extern "C" __global__ void __launch_bounds__(256)
k_after_qr(const DChain *chains, int pt, int vsel, int vpt)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    int i = (int)(e % n), j = (int)(e / n);
    // every load (the pivot index first, then the row of V it names) is issued before the
    // first store, through global-space pointers
    GAS cplx *R = gp(c.W2), *Vp = gp(chain_mat(c, vsel, vpt));
    const int jp = gp(c.jpvt)[i];
    const double di = gld(R + ((size_t)i + (size_t)i * n)).x;
    const cplx r = gld(R + e);
    const cplx vrow = gld(Vp + ((size_t)(jp - 1) + (size_t)j * n));
    gst(gp(c.W3) + e, (i <= j) ? cscale(r, 1.0 / di) : make_double2(0.0, 0.0));
    gst(gp(c.W4) + e, vrow);
    if (j == 0) gp(c.D)[(size_t)n * pt + i] = di;
}

// ------------------------------------------------------------------------------------------
// green_equaltime / green_equaltimebb pieces (src/dqmc_ft_core/green_equaltime.f90:1-107,
// green_equaltimebb.f90:1-84).  Right point R = (U[ptR], D[ptR], V[ptR]) or the saved point;
// left point likewise.  rsel/lsel: 0 = chain point, 1 = saved copy (Us, Ds, Vs).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double dmax_of(double d) { return fabs(d) > 1.0 ? fabs(d) : 1.0; }
__device__ __forceinline__ double dmin_of(double d) { return fabs(d) > 1.0 ? (d >= 0 ? 1.0 : -1.0) : d; }

__device__ __forceinline__ const double *dvec_of(const DChain &c, int saved, int pt)
{
    return saved ? c.Ds : c.D + (size_t)c.n * pt;
}

// M = uu/(drmax dlmax) + vv drmin dlmin into W3 (uu in W1, vv in W2); bb = 1: M = UR^H/drmax + VR drmin
extern "C" __global__ void __launch_bounds__(256)
k_green_m(const DChain *chains, int bb, int rsaved, int rpt, int lsaved, int lpt)
{
    __shared__ cplx tile[32][33];
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const int bx = blockIdx.x % ((n + 31) / 32), by = blockIdx.x / ((n + 31) / 32);
    if (by * 32 >= n) return;
    // every thread's four entries (256 threads, 32 x 32 tile) are loaded in one batch through
    // global-space pointers with clamped addresses, ahead of the stores
    GAS const double *dr = gp(dvec_of(c, rsaved, rpt));
    constexpr int GU = 1024 / 256;
    if (!bb) {
        GAS const double *dl = gp(dvec_of(c, lsaved, lpt));
        GAS cplx *W1 = gp(c.W1), *W2 = gp(c.W2), *W3 = gp(c.W3);
        cplx x1[GU], x2[GU];
        double di[GU], dj[GU];
#pragma unroll
        for (int u = 0; u < GU; u++) {
            const int e = threadIdx.x + 256 * u;
            const int i = min(bx * 32 + (e % 32), n - 1), j = min(by * 32 + (e / 32), n - 1);
            const size_t o = (size_t)i + (size_t)j * n;
            x1[u] = gld(W1 + o);
            x2[u] = gld(W2 + o);
            di[u] = dr[i];
            dj[u] = dl[j];
        }
#pragma unroll
        for (int u = 0; u < GU; u++) {
            const int e = threadIdx.x + 256 * u;
            const int i = bx * 32 + (e % 32), j = by * 32 + (e / 32);
            if (i < n && j < n) {
                const double a = 1.0 / (dmax_of(di[u]) * dmax_of(dj[u]));
                const double b = dmin_of(di[u]) * dmin_of(dj[u]);
                gst(W3 + ((size_t)i + (size_t)j * n), cadd(cscale(x1[u], a), cscale(x2[u], b)));
            }
        }
    } else {
        GAS cplx *ur = gp(rsaved ? c.Us : c.U + (size_t)n * n * rpt);
        GAS cplx *vr = gp(rsaved ? c.Vs : c.V + (size_t)n * n * rpt);
        GAS cplx *W3 = gp(c.W3);
        // tile of UR^H: element (i, j) = conj(UR(j, i)); read UR block (by, bx) coalesced
        cplx xu[GU], xv[GU];
        double di[GU];
#pragma unroll
        for (int u = 0; u < GU; u++) {
            const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
            const int r = min(by * 32 + q, n - 1), cc = min(bx * 32 + p, n - 1);
            xu[u] = gld(ur + ((size_t)r + (size_t)cc * n));
            const int i = min(bx * 32 + q, n - 1), j = min(by * 32 + p, n - 1);
            xv[u] = gld(vr + ((size_t)i + (size_t)j * n));
            di[u] = dr[i];
        }
#pragma unroll
        for (int u = 0; u < GU; u++) {
            const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
            const int r = by * 32 + q, cc = bx * 32 + p;
            tile[q][p] = (r < n && cc < n) ? xu[u] : make_double2(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < GU; u++) {
            const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
            const int i = bx * 32 + q, j = by * 32 + p;
            if (i < n && j < n) {
                const cplx uh = cconj(tile[p][q]);
                gst(W3 + ((size_t)i + (size_t)j * n), cadd(cscale(uh, 1.0 / dmax_of(di[u])), cscale(xv[u], dmin_of(di[u]))));
            }
        }
    }
}

// right-hand side X = DRmax^-1 UR^H into W4 (tiled transpose)
extern "C" __global__ void __launch_bounds__(256)
k_green_rhs(const DChain *chains, int rsaved, int rpt)
{
    __shared__ cplx tile[32][33];
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const int nb = (n + 31) / 32;
    const int bx = blockIdx.x % nb, by = blockIdx.x / nb;
    if (by >= nb) return;
    // the thread's four entries and scale factors in one batch (global-space pointers, clamped
    // addresses), ahead of the LDS stores
    GAS const double *dr = gp(dvec_of(c, rsaved, rpt));
    GAS cplx *ur = gp(rsaved ? c.Us : c.U + (size_t)n * n * rpt);
    GAS cplx *W4 = gp(c.W4);
    constexpr int GU = 1024 / 256;
    cplx xu[GU];
    double di[GU];
#pragma unroll
    for (int u = 0; u < GU; u++) {
        const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
        xu[u] = gld(ur + ((size_t)min(by * 32 + q, n - 1) + (size_t)min(bx * 32 + p, n - 1) * n));
        di[u] = dr[min(bx * 32 + q, n - 1)];
    }
#pragma unroll
    for (int u = 0; u < GU; u++) {
        const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
        const int r = by * 32 + q, cc = bx * 32 + p;
        tile[q][p] = (r < n && cc < n) ? xu[u] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < GU; u++) {
        const int e = threadIdx.x + 256 * u, q = e % 32, p = e / 32;
        const int i = bx * 32 + q, j = by * 32 + p;
        if (i < n && j < n) gst(W4 + ((size_t)i + (size_t)j * n), cscale(cconj(tile[p][q]), 1.0 / dmax_of(di[u])));
    }
}

// W1 = UL diag(1/dlmax)
extern "C" __global__ void __launch_bounds__(256)
k_green_ulscale(const DChain *chains, int lsaved, int lpt)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    int j = (int)(e / n);
    const double *dl = dvec_of(c, lsaved, lpt);
    const cplx *ul = lsaved ? c.Us : c.U + (size_t)n * n * lpt;
    c.W1[e] = cscale(ul[e], 1.0 / dmax_of(dl[j]));
}

// LU with partial pivoting (zgetrf, pivot = largest |Re|+|Im|) of W3 in place, one workgroup per
// chain; log det with the imaginary part folded into (-pi, pi] (s_inv_logdet_lu_z) -> SC_ZLD
extern "C" __global__ void __launch_bounds__(1024)
k_lu(const DChain *chains)
{
    extern __shared__ cplx lsh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    cplx *lcol = lsh;           // n
    cplx *lrow = lsh + n;       // n
    __shared__ double rv[16];
    __shared__ int ri[16];
    __shared__ int s_piv;
    __shared__ int s_info;
    cplx *A = c.W3;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (threadIdx.x == 0) s_info = 0;
    {
        double bv = -1.0;
        int bi = 0;
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            cplx a = A[r];
            double v = fabs(a.x) + fabs(a.y);
            if (v > bv) { bv = v; bi = r; }
        }
        double mv;
        int p;
        block_argmax(bv, bi, rv, ri, &mv, &p);
        if (threadIdx.x == 0) s_piv = p;
        __syncthreads();
    }
    for (int k = 0; k < n; k++) {
        const int p = s_piv;
        if (threadIdx.x == 0) c.piv[k] = p;
        if (p != k)
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                cplx t = A[(size_t)p + (size_t)j * n];
                A[(size_t)p + (size_t)j * n] = A[(size_t)k + (size_t)j * n];
                A[(size_t)k + (size_t)j * n] = t;
            }
        __syncthreads();
        cplx akk = A[(size_t)k + (size_t)k * n];
        bool zero = (akk.x == 0.0 && akk.y == 0.0);
        if (zero && threadIdx.x == 0 && s_info == 0) s_info = k + 1;
        cplx pinv = zero ? make_double2(0.0, 0.0) : cinv(akk);
        for (int r = k + 1 + threadIdx.x; r < n; r += blockDim.x) {
            cplx l = zero ? A[(size_t)r + (size_t)k * n] : cmul(A[(size_t)r + (size_t)k * n], pinv);
            A[(size_t)r + (size_t)k * n] = l;
            lcol[r] = l;
        }
        for (int j = k + 1 + threadIdx.x; j < n; j += blockDim.x) lrow[j] = A[(size_t)k + (size_t)j * n];
        __syncthreads();
        double bv = -1.0;
        int bi = k + 1;
        for (int j = k + 1 + w; j < n; j += nw) {
            cplx *a = A + (size_t)j * n;
            cplx f = lrow[j];
            bool track = (j == k + 1);
            for (int r = k + 1 + lane; r < n; r += 64) {
                cplx v = csub(a[r], cmul(lcol[r], f));
                a[r] = v;
                if (track) {
                    double m = fabs(v.x) + fabs(v.y);
                    if (m > bv) { bv = m; bi = r; }
                }
            }
        }
        double mv;
        int np;
        block_argmax(bv, bi, rv, ri, &mv, &np);
        if (threadIdx.x == 0) s_piv = (k + 1 < n) ? np : 0;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cplx ld = make_double2(0.0, 0.0);
        for (int i = 0; i < n; i++) {
            cplx d = A[(size_t)i + (size_t)i * n];
            if (c.piv[i] != i) d = make_double2(-d.x, -d.y);
            ld = cadd(ld, clog_(d));
        }
        double im = ld.y;
        double k2 = round(im / 2.0 / M_PI);
        ld.y = im - 2.0 * M_PI * k2;
        c.sc[SC_ZLD] = ld.x;
        c.sc[SC_ZLD + 1] = ld.y;
        c.sc[SC_INFO] = (double)s_info;
    }
}

// solve (P L U) Y = X for X = matrix `xsel`, in place; one wave per right-hand-side column,
// the column held in registers (n <= 1024)
#define SOLVE_SLOTS 16
extern "C" __global__ void __launch_bounds__(512)
k_lu_solve(const DChain *chains, int xsel)
{
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    const cplx *A = c.W3;
    cplx *X = chain_mat(c, xsel, 0);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int ns = (n + 63) / 64;
    for (int j = w; j < n; j += nw) {
        cplx *x = X + (size_t)j * n;
        cplx reg[SOLVE_SLOTS];
#pragma unroll
        for (int s = 0; s < SOLVE_SLOTS; s++) {
            int r = lane + 64 * s;
            reg[s] = (s < ns && r < n) ? x[r] : make_double2(0.0, 0.0);
        }
        // row interchanges, in order
        for (int k = 0; k < n; k++) {
            int p = c.piv[k];
            if (p == k) continue;
            int sk = k >> 6, lk = k & 63, sp = p >> 6, lp = p & 63;
            double kr = 0, ki = 0, pr = 0, pi = 0;
#pragma unroll
            for (int s = 0; s < SOLVE_SLOTS; s++) {
                if (s == sk) { kr = __shfl(reg[s].x, lk, 64); ki = __shfl(reg[s].y, lk, 64); }
                if (s == sp) { pr = __shfl(reg[s].x, lp, 64); pi = __shfl(reg[s].y, lp, 64); }
            }
#pragma unroll
            for (int s = 0; s < SOLVE_SLOTS; s++) {
                if (s == sk && lane == lk) reg[s] = make_double2(pr, pi);
                if (s == sp && lane == lp) reg[s] = make_double2(kr, ki);
            }
        }
        // L y = b (unit lower)
        for (int k = 0; k < n - 1; k++) {
            int sk = k >> 6, lk = k & 63;
            double yr = 0, yi = 0;
#pragma unroll
            for (int s = 0; s < SOLVE_SLOTS; s++)
                if (s == sk) { yr = __shfl(reg[s].x, lk, 64); yi = __shfl(reg[s].y, lk, 64); }
            const cplx yk = make_double2(yr, yi);
            const cplx *lk_col = A + (size_t)k * n;
#pragma unroll
            for (int s = 0; s < SOLVE_SLOTS; s++) {
                int r = lane + 64 * s;
                if (s >= sk && s < ns && r > k && r < n) reg[s] = csub(reg[s], cmul(lk_col[r], yk));
            }
        }
        // U x = y
        for (int k = n - 1; k >= 0; k--) {
            int sk = k >> 6, lk = k & 63;
            cplx ukk = A[(size_t)k + (size_t)k * n];
#pragma unroll
            for (int s = 0; s < SOLVE_SLOTS; s++)
                if (s == sk && lane == lk) reg[s] = cdiv(reg[s], ukk);
            double yr = 0, yi = 0;
#pragma unroll
            for (int s = 0; s < SOLVE_SLOTS; s++)
                if (s == sk) { yr = __shfl(reg[s].x, lk, 64); yi = __shfl(reg[s].y, lk, 64); }
            const cplx yk = make_double2(yr, yi);
            const cplx *uk_col = A + (size_t)k * n;
#pragma unroll
            for (int s = 0; s < SOLVE_SLOTS; s++) {
                int r = lane + 64 * s;
                if (s <= sk && r < k) reg[s] = csub(reg[s], cmul(uk_col[r], yk));
            }
        }
#pragma unroll
        for (int s = 0; s < SOLVE_SLOTS; s++) {
            int r = lane + 64 * s;
            if (s < ns && r < n) x[r] = reg[s];
        }
    }
}

// logweightf scratch: general (bb = 0): -ldql + ldqr + zld + sum log dlmax + sum log drmax;
// bb = 1: ldqr + zld + sum log drmax
extern "C" __global__ void __launch_bounds__(256)
k_green_logw(const DChain *chains, int bb, int rsaved, int rpt, int lsaved, int lpt, int dst_logw)
{
    __shared__ double red[16];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    const double *dr = dvec_of(c, rsaved, rpt);
    const double *dl = bb ? nullptr : dvec_of(c, lsaved, lpt);
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s += log(dmax_of(dr[i]));
        if (!bb) s += log(dmax_of(dl[i]));
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        cplx ldqr = rsaved ? make_double2(c.sc[SC_LDQ_SAVE], c.sc[SC_LDQ_SAVE + 1]) : c.ldq[rpt];
        cplx lw = cadd(ldqr, make_double2(c.sc[SC_ZLD] + s, c.sc[SC_ZLD + 1]));
        if (!bb) {
            cplx ldql = lsaved ? make_double2(c.sc[SC_LDQ_SAVE], c.sc[SC_LDQ_SAVE + 1]) : c.ldq[lpt];
            lw = csub(lw, ldql);
        }
        c.sc[dst_logw] = lw.x;
        c.sc[dst_logw + 1] = lw.y;
    }
}

// wrap check and replacement (dqmc_ft_sweep_0b.f90:212-259): max|Gs - G|, set_phase of the
// scratch weight, G <- Gs when the LU succeeded; compare = 0 at sweep start (plain set)
extern "C" __global__ void __launch_bounds__(1024)
k_wrap(const DChain *chains, int compare)
{
    __shared__ double red[16];
    __shared__ double cnt_sh[8][16];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n;
    double *sc = c.sc;
    double mx = 0.0;
    bool nonfinite = false;
    // eight entries of Gs (and G) per thread in flight at a time, clamped addresses; the maximum
    // and the flag do not depend on the order
    GAS cplx *Gsg = gp(c.Gs), *Gg = gp(c.G);
    const int nn = n * n;
    for (int base = threadIdx.x; base < nn; base += 8 * (int)blockDim.x) {
        cplx a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = min(base + u * (int)blockDim.x, nn - 1);
            a[u] = gld(Gsg + e);
            if (compare) b[u] = gld(Gg + e);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (base + u * (int)blockDim.x >= nn) continue;
            if (!isfinite(a[u].x) || !isfinite(a[u].y)) nonfinite = true;
            if (compare) {
                const double d = hypot(a[u].x - b[u].x, a[u].y - b[u].y);
                if (d > mx) mx = d;
            }
        }
    }
    // max reduction
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) red[w] = mx;
    __syncthreads();
    mx = 0.0;
    for (int i = 0; i < nw; i++) mx = fmax(mx, red[i]);
    // set_phase: exp(i nflr Im logw) * prod_{i,nt} phase(conf) = exp(i (nflr Im logw + sum cnt_l arg_l))
    // counts 0-3: the on-site field's components; 4-7: the plaquette field's (hconf%phase of
    // every plaquette and slice, set_phase.f90:9-15)
    double cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (c.phase_from_conf && c.lwrapu) {
        const int tot = n * c.ltrot;
        GAS const int32_t *cg = gp(c.conf);
        for (int base = threadIdx.x; base < tot; base += 8 * (int)blockDim.x) {
            int v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = cg[min(base + u * (int)blockDim.x, tot - 1)];
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (base + u * (int)blockDim.x < tot) {
                    cnt[0] += (v[u] == 1); cnt[1] += (v[u] == 2); cnt[2] += (v[u] == 3); cnt[3] += (v[u] == 4);
                }
        }
    }
    if (c.lwrapp) {
        size_t tot = (size_t)c.lqp * c.ltrot;
        for (size_t e = threadIdx.x; e < tot; e += blockDim.x) {
            int v = c.confp[e];
            cnt[4] += (v == 1); cnt[5] += (v == 2); cnt[6] += (v == 3); cnt[7] += (v == 4);
        }
    }
    for (int l = 0; l < 8; l++) {
        double v = cnt[l];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) cnt_sh[l][w] = v;
    }
    int nf_flag = __syncthreads_or(nonfinite ? 1 : 0);
    if (threadIdx.x == 0) {
        double ang = (double)c.nflr * sc[SC_LOGW_TMP + 1];
        for (int l = 0; l < c.lcomp && l < 4; l++) {
            double s = 0.0;
            for (int i = 0; i < nw; i++) s += cnt_sh[l][i];
            ang += s * c.parg[l];
        }
        if (c.lwrapp)
            for (int l = 0; l < 4; l++) {
                double s = 0.0;
                for (int i = 0; i < nw; i++) s += cnt_sh[4 + l][i];
                ang += s * c.pparg[l];
            }
        cplx ph = make_double2(cos(ang), sin(ang));
        int info = (int)sc[SC_INFO];
        if (nf_flag) sc[SC_NONFINITE] = 1.0;
        if (compare) {
            if (mx > sc[SC_MAXWRAP]) sc[SC_MAXWRAP] = mx;
            cplx ph0 = make_double2(sc[SC_PHASE], sc[SC_PHASE + 1]);
            double pd = hypot(ph0.x - ph.x, ph0.y - ph.y);
            if (pd > sc[SC_MAXPH]) sc[SC_MAXPH] = pd;
            sc[SC_NCHK] += 1.0;
            sc[SC_SUMLOG] += log10(mx > 0 ? mx : 1e-300);
        }
        if (info == 0) {
            sc[SC_LOGW] = sc[SC_LOGW_TMP];
            sc[SC_LOGW + 1] = sc[SC_LOGW_TMP + 1];
            sc[SC_PHASE] = ph.x;
            sc[SC_PHASE + 1] = ph.y;
        }
        red[0] = (double)info;
    }
    __syncthreads();
    if ((int)red[0] == 0) {
        for (int base = threadIdx.x; base < nn; base += 8 * (int)blockDim.x) {
            cplx a[8];
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] = gld(Gsg + min(base + u * (int)blockDim.x, nn - 1));
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (base + u * (int)blockDim.x < nn) gst(Gg + base + u * (int)blockDim.x, a[u]);
        }
    }
    __syncthreads();
    int tn = (int)sc[SC_TRACE_N];
    if (c.trace && tn < c.trace_cap) {
        for (size_t e = threadIdx.x; e < (size_t)n * n; e += blockDim.x) c.trace[(size_t)tn * n * n + e] = c.G[e];
        __syncthreads();
        if (threadIdx.x == 0) sc[SC_TRACE_N] = tn + 1;
    }
}

// ------------------------------------------------------------------------------------------
// small utility kernels
// ------------------------------------------------------------------------------------------
// set chain point pt to (I, 1, I, 0)
extern "C" __global__ void __launch_bounds__(256)
k_set_identity(const DChain *chains, int pt)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    int i = (int)(e % n), j = (int)(e / n);
    cplx v = make_double2(i == j ? 1.0 : 0.0, 0.0);
    c.U[(size_t)n * n * pt + e] = v;
    c.V[(size_t)n * n * pt + e] = v;
    if (j == 0) c.D[(size_t)n * pt + i] = 1.0;
    if (e == 0) c.ldq[pt] = make_double2(0.0, 0.0);
}

// copy between matrices; mode 1 additionally saves D[pt] and ldq[pt] of the chain point
extern "C" __global__ void __launch_bounds__(256)
k_copy(const DChain *chains, int dsel, int dpt, int ssel, int spt)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    chain_mat(c, dsel, dpt)[e] = chain_mat(c, ssel, spt)[e];
}

// save chain point pt (U, D, V, logdetQ) into (Us, Ds, Vs, SC_LDQ_SAVE)
extern "C" __global__ void __launch_bounds__(256)
k_save_point(const DChain *chains, int pt)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    c.Us[e] = c.U[(size_t)n * n * pt + e];
    c.Vs[e] = c.V[(size_t)n * n * pt + e];
    if (e < (size_t)n) c.Ds[e] = c.D[(size_t)n * pt + e];
    if (e == 0) { c.sc[SC_LDQ_SAVE] = c.ldq[pt].x; c.sc[SC_LDQ_SAVE + 1] = c.ldq[pt].y; }
}

// ------------------------------------------------------------------------------------------
// Delay update with the delayed vectors held in registers (src/dqmc_config_u/dqmc_update_delay.f90:39-128).
// Wave w covers rows 32w .. 32w+31: lane l < 32 keeps avec(32w + l, 0:K), lane l + 32 keeps
// bvec of the same row, so a row's two new entries meet through a lane shuffle.  Each site
// costs one workgroup barrier: the rows avec(s, :) and bvec(s, :) that site s needs are
// published by row s's lanes at the end of site s - 1 into one of two LDS buffers (by the
// parity of s), and diagg(s) is only read at site s, so no lane rewrites a value another lane
// may still be reading.  avec/bvec reach memory only for the MFMA flush.
// ------------------------------------------------------------------------------------------
#define DSB 8   // sites per staged block of G columns and rows

// sum_{m < ik} other[m] reg[m] in chunks of four entries: ik is uniform over the workgroup,
// so each chunk is a scalar branch and only the filled chunks are read (entries of a partly
// filled chunk beyond ik are zero in reg).  Two accumulators alternate over m.
template <int K>
__device__ __forceinline__ cplx delay_dot(int ik, const cplx (&reg)[K], const cplx *other)
{
    double vr = 0.0, vi = 0.0, wr = 0.0, wi = 0.0;
#pragma unroll
    for (int c = 0; c < K / 4; c++) {
        if (4 * c < ik) {
            const cplx o0 = other[4 * c], o1 = other[4 * c + 1], o2 = other[4 * c + 2], o3 = other[4 * c + 3];
            vr += o0.x * reg[4 * c].x - o0.y * reg[4 * c].y;
            vi += o0.x * reg[4 * c].y + o0.y * reg[4 * c].x;
            wr += o1.x * reg[4 * c + 1].x - o1.y * reg[4 * c + 1].y;
            wi += o1.x * reg[4 * c + 1].y + o1.y * reg[4 * c + 1].x;
            vr += o2.x * reg[4 * c + 2].x - o2.y * reg[4 * c + 2].y;
            vi += o2.x * reg[4 * c + 2].y + o2.y * reg[4 * c + 2].x;
            wr += o3.x * reg[4 * c + 3].x - o3.y * reg[4 * c + 3].y;
            wi += o3.x * reg[4 * c + 3].y + o3.y * reg[4 * c + 3].x;
        }
    }
    return make_double2(vr + wr, vi + wi);
}

// the value of lane l ^ 32 (v_permlane32_swap exchanges the two 32-lane halves)
__device__ __forceinline__ double partner32(double v)
{
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    const bool upper = (threadIdx.x & 32) != 0;
    return __hiloint2double((int)(upper ? b[0] : b[1]), (int)(upper ? a[0] : a[1]));
}

// reg[ik] = v by a binary tree of uniform branches; the empty volatile asm keeps each leaf a
// real branch target instead of a select over all K registers
template <int K, int LO, int HI>
__device__ __forceinline__ void delay_put(int ik, cplx (&reg)[K], cplx v)
{
    if constexpr (HI - LO == 1) {
        reg[LO] = v;
        asm volatile("" ::: "memory");
    } else {
        constexpr int MID = (LO + HI) / 2;
        if (ik < MID)
            delay_put<K, LO, MID>(ik, reg, v);
        else
            delay_put<K, MID, HI>(ik, reg, v);
    }
}

template <int K>
__global__ void __launch_bounds__(K >= 8 ? 512 : 1024)
k_update_delay_reg(const DChain *chains, int nt, int slice_in_pass)
{
    extern __shared__ cplx sh[];
    const DChain &c = chains[blockIdx.x];
    const int n = c.n, kb = c.nublock;
    const int lane = threadIdx.x & 63;
    const bool isA = lane < 32;
    const int r = 32 * (threadIdx.x >> 6) + (lane & 31);
    cplx *diagg = sh;                   // n
    cplx *gcol = diagg + n;             // DSB x n: G(:, s0+j)
    cplx *grow = gcol + DSB * n;        // DSB x n: G(s0+j, :)
    cplx *rows = grow + DSB * n;        // 2 buffers x (avec row K, bvec row K)
    double *us = (double *)(rows + 4 * K);  // 2n uniforms of this slice
    int *cfs = (int *)(us + 2 * n);     // n: HS field of this slice (read-only here)
    __shared__ HsTab tab;
    GAS cplx *G = gp(c.G);
    int32_t *cf = c.conf + (size_t)(nt - 1) * n;
    const double *u = c.unif + (size_t)2 * n * slice_in_pass;
    cplx phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    cplx reg[K];
#pragma unroll
    for (int m = 0; m < K; m++) reg[m] = make_double2(0.0, 0.0);
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        diagg[q] = gld(G + ((size_t)q + (size_t)q * n));
        cfs[q] = cf[q];
    }
    for (int q = threadIdx.x; q < 2 * n; q += blockDim.x) us[q] = u[q];
    for (int q = threadIdx.x; q < 4 * K; q += blockDim.x) rows[q] = make_double2(0.0, 0.0);
    hs_tab_load(tab, c);
    int ik = 0, s0 = -1;
    double accm = 0.0;
    for (int s = 0; s < n; s++) {
        if (s0 < 0 || s >= s0 + DSB) {
            // stage G(:, s..s+DSB) and G(s..s+DSB, :); valid until the next flush changes G.
            // The previous site's lanes may still read the old block.
            if (s0 >= 0) __syncthreads();
            s0 = s;
            const int nb = min(DSB, n - s);
            // columns s..s+nb are one contiguous run; a row block is read as nb-element runs
            GAS cplx *Gc = G + (size_t)s * n;
            stage_in_map<4>(gcol, nb * n, [=](int e) { return gld_nt(Gc + (e)); }, [](int e) { return e; });
            stage_in_map<4>(grow, nb * n, [=](int e) { return gld_nt(G + ((size_t)(s + e % nb) + (size_t)(e / nb) * n)); },
                            [=](int e) { return (e % nb) * n + e / nb; });
        }
        __syncthreads();
        const SiteDraw d = site_draw(tab, cfs[s], us[2 * s]);
        const cplx dg = diagg[s];
        const cplx ratio1 = cadd(cmul(make_double2(1.0 - dg.x, -dg.y), d.db), make_double2(1.0, 0.0));
        // delta / ratio1, evaluated next to the decision (independent of it) so that the two
        // dependency chains overlap
        const cplx sscl = cdiv(d.db, ratio1);
        cplx rt;
        // every thread evaluates the same decision; reading it from lane 0 makes the branch
        // and the delay count scalar
        const bool acc = __builtin_amdgcn_readfirstlane((int)site_accept_fast(tab, ratio1, d, us[2 * s + 1], phase, &rt)) != 0;
        if (acc) {
            accm += 1.0;
            phase = cmul(phase, rt);   // normalised at the next flush
            const cplx *rw = rows + (s & 1) * 2 * K;      // avec(s, :), then bvec(s, :)
            // A lanes: sum_m bvec(s,m) avec(r,m); B lanes: sum_m avec(s,m) bvec(r,m)
            const cplx *other = isA ? rw + K : rw;
            const int iku = __builtin_amdgcn_readfirstlane(ik);
            const cplx dot = delay_dot<K>(iku, reg, other);
            cplx v = make_double2(0.0, 0.0);
            if (r < n) {
                if (isA) {
                    const cplx gc = gcol[(s - s0) * n + r];
                    v = cmul(make_double2(gc.x + dot.x, gc.y + dot.y), sscl);
                } else {
                    const cplx gr = grow[(s - s0) * n + r];
                    v = make_double2(gr.x + dot.x, gr.y + dot.y);
                    if (r == s) v.x -= 1.0;
                }
            }
            delay_put<K, 0, K>(iku, reg, v);
            // the row's other entry from the partner lane (A <-> B)
            const cplx w = make_double2(partner32(v.x), partner32(v.y));
            if (isA && r > s && r < n) diagg[r] = cadd(diagg[r], cmul(v, w));
            if (threadIdx.x == 0) cf[s] = d.isp;
            ik++;
        }
        // row s+1 publishes its filled vector entries for the next site (the reader's products
        // cover the same chunks, and its entries past ik are zero)
        if (r == s + 1 && r < n) {
            cplx *pw = rows + ((s + 1) & 1) * 2 * K + (isA ? 0 : K);
            const int iku = __builtin_amdgcn_readfirstlane(ik);
#pragma unroll
            for (int c4 = 0; c4 < K / 4; c4++)
                if (4 * c4 < iku) {
                    pw[4 * c4] = reg[4 * c4];
                    pw[4 * c4 + 1] = reg[4 * c4 + 1];
                    pw[4 * c4 + 2] = reg[4 * c4 + 2];
                    pw[4 * c4 + 3] = reg[4 * c4 + 3];
                }
        }
        if (ik == kb || s == n - 1) {
            __syncthreads();
            if (ik > 0) {
                if (r < n) {
                    GAS cplx *dst = gp(isA ? c.avec : c.bvec);
#pragma unroll
                    for (int m = 0; m < K; m++)
                        if (m < ik) gst(dst + ((size_t)r + (size_t)m * n), reg[m]);
                }
                // the vectors restart from zero after the flush; clearing them here frees
                // their registers for the flush
#pragma unroll
                for (int m = 0; m < K; m++) reg[m] = make_double2(0.0, 0.0);
                __syncthreads();
                wg_rank_k_strips<K>(c.G, n, c.avec, c.bvec, ik);
            }
            __syncthreads();
            phase = cscale(phase, 1.0 / sqrt(cabs2(phase)));
            ik = 0;
            s0 = -1;
            if (s < n - 1)
                for (int q = threadIdx.x; q < n; q += blockDim.x) diagg[q] = gld(G + ((size_t)q + (size_t)q * n));
        }
    }
    if (threadIdx.x == 0) {
        c.sc[SC_PHASE] = phase.x;
        c.sc[SC_PHASE + 1] = phase.y;
        c.sc[SC_ACC] += accm;
        c.sc[SC_TRIED] += n;
    }
}

template __global__ void k_update_delay_reg<4>(const DChain *, int, int);
template __global__ void k_update_delay_reg<8>(const DChain *, int, int);
template __global__ void k_update_delay_reg<16>(const DChain *, int, int);
template __global__ void k_update_delay_reg<32>(const DChain *, int, int);

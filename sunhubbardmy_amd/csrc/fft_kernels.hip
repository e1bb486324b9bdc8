// fft_kernels.hip -- kinetic propagation by the discrete Fourier transform of the lattice (the
// FFT build of the reference: src/dqmc_config_h0/dqmc_left_forward_prop.f90:95-134,
// dqmc_right_backward_prop.f90:100-144, tables from dqmc_set_h0conf.f90:225-291).
//
// On a Bravais lattice with one site per cell and translation-invariant hopping,
// exp(-+dtau K/2) = F^H diag(e_k) F with F the unitary DFT over the cells and
// e_k = exp(-+dtau eps_k / 2) (exph0k / exph0kinv).  Sites are numbered
// s = i0 + n0 (i1 + n1 i2) (n0 the fastest of the up to three lattice directions), and e_k is
// stored in the same order of the wave-vector index.
//
// A workgroup owns `fcb` whole columns (left products: rows mixed) or rows (right products:
// columns mixed) of one chain's matrix in LDS, the HS-field scalings of the slice are fused as in
// the checkerboard kernels, and each kinetic half-step is a forward transform, the diagonal, and
// an inverse transform.  A transform is one pass per lattice direction; in a pass every thread
// holds one line (length L <= 24) in registers and transforms it by a mixed-radix Cooley-Tukey
// recursion unrolled at compile time, so the matrix is read and written once per pass from LDS
// and once per launch from HBM.  Lines of a column are padded (one slot per n0-run, an odd
// leading dimension per column) so that neither the lines along n0 nor the staging accesses of
// consecutive rows fall on one LDS bank.
#include "sunhip_kernels.h"

#define FFT_LMAX 24

__host__ __device__ constexpr int fft_spf(int l)
{
    for (int p = 2; p * p <= l; p++)
        if (l % p == 0) return p;
    return l;
}

// omega_L^m = exp(-+ 2 pi i m / L) from the table of the full line length LT (tw[j] =
// exp(-2 pi i j / LT)); inv selects the conjugate (the inverse transform)
template <int L, int LT>
__device__ __forceinline__ cplx fft_w(const cplx *tw, int m, bool inv)
{
    const cplx w = tw[(m % L) * (LT / L)];
    return make_double2(w.x, inv ? -w.y : w.y);
}

// in-place unnormalised DFT of x[0..L): X[k] = sum_n x[n] omega_L^(n k).  L = P Q with P the
// smallest prime factor: P-point transforms over n = Q n1 + n2, twiddles omega_L^(n2 k1), Q-point
// transforms, output in natural order (k = k1 + P k2).  Every index is a compile-time constant
// after unrolling, so x stays in registers.
template <int L, int LT>
__device__ __forceinline__ void fft_reg(cplx (&x)[L], const cplx *tw, bool inv)
{
    if constexpr (L == 1) {
        return;
    } else if constexpr (L == 2) {
        const cplx a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (fft_spf(L) == L) {
        cplx y[L];
#pragma unroll
        for (int k = 0; k < L; k++) {
            cplx s = x[0];
#pragma unroll
            for (int m = 1; m < L; m++) {
                const int e = (m * k) % L;
                s = e == 0 ? cadd(s, x[m]) : cadd(s, cmul(x[m], fft_w<L, LT>(tw, e, inv)));
            }
            y[k] = s;
        }
#pragma unroll
        for (int k = 0; k < L; k++) x[k] = y[k];
    } else {
        constexpr int P = fft_spf(L), Q = L / P;
        cplx y[L];
#pragma unroll
        for (int n2 = 0; n2 < Q; n2++) {
            cplx a[P];
#pragma unroll
            for (int n1 = 0; n1 < P; n1++) a[n1] = x[Q * n1 + n2];
            fft_reg<P, LT>(a, tw, inv);
#pragma unroll
            for (int k1 = 0; k1 < P; k1++) {
                const int e = (n2 * k1) % L;
                y[n2 + Q * k1] = e == 0 ? a[k1] : cmul(a[k1], fft_w<L, LT>(tw, e, inv));
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < P; k1++) {
            cplx b[Q];
#pragma unroll
            for (int n2 = 0; n2 < Q; n2++) b[n2] = y[n2 + Q * k1];
            fft_reg<Q, LT>(b, tw, inv);
#pragma unroll
            for (int k2 = 0; k2 < Q; k2++) x[k1 + P * k2] = b[k2];
        }
    }
}

// geometry of one chain's transform in LDS: directions n[0..2] (n[0] fastest), padded index
// of site s is s + s / n0, the leading dimension of a column (or row) is ld (odd)
struct FftGeo {
    int n[3];
    int ld;
};

__host__ __device__ __forceinline__ int fft_ld(int n, int n0)
{
    const int ld = n + n / n0;
    return ld | 1;
}

__device__ __forceinline__ int fft_pad(int s, int n0) { return s + s / n0; }

// one pass along direction d over `nvec` vectors (columns or rows) of the LDS block: every
// thread transforms whole lines
template <int L>
__device__ __forceinline__ void fft_lines(cplx *lds, const FftGeo &g, int d, int nvec, const cplx *tw, bool inv)
{
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    // lines along d: the other two directions and the vector index
    const int o1 = d == 0 ? n1 : n0, o2 = d == 2 ? n1 : n2;
    const int nlines = o1 * o2 * nvec;
    const int stride = d == 0 ? 1 : d == 1 ? n0 + 1 : (n0 + 1) * n1;
    for (int line = threadIdx.x; line < nlines; line += blockDim.x) {
        const int a = line % o1, rest = line / o1, b = rest % o2, v = rest / o2;
        int base;
        if (d == 0) base = (n0 + 1) * (a + n1 * b);
        else if (d == 1) base = a + (n0 + 1) * n1 * b;
        else base = a + (n0 + 1) * b;
        base += v * g.ld;
        cplx x[L];
#pragma unroll
        for (int j = 0; j < L; j++) x[j] = lds[base + j * stride];
        fft_reg<L, L>(x, tw, inv);
#pragma unroll
        for (int j = 0; j < L; j++) lds[base + j * stride] = x[j];
    }
}

__device__ static void fft_pass(cplx *lds, const FftGeo &g, int d, int nvec, const cplx *tw, bool inv)
{
    switch (d == 0 ? g.n[0] : d == 1 ? g.n[1] : g.n[2]) {
#define FFT_CASE(L) case L: fft_lines<L>(lds, g, d, nvec, tw, inv); break;
    FFT_CASE(2) FFT_CASE(3) FFT_CASE(4) FFT_CASE(5) FFT_CASE(6) FFT_CASE(7) FFT_CASE(8) FFT_CASE(9)
    FFT_CASE(10) FFT_CASE(11) FFT_CASE(12) FFT_CASE(13) FFT_CASE(14) FFT_CASE(15) FFT_CASE(16)
    FFT_CASE(18) FFT_CASE(20) FFT_CASE(24)
#undef FFT_CASE
    default: break;   // length 1: nothing to do (other lengths are refused by sh_create)
    }
    __syncthreads();
}

// the twiddle tables of the three directions, tw[d][j] = exp(-2 pi i j / n_d)
__device__ __forceinline__ void fft_twiddles(cplx (*tw)[FFT_LMAX], const FftGeo &g)
{
    for (int e = threadIdx.x; e < 3 * FFT_LMAX; e += blockDim.x) {
        const int d = e / FFT_LMAX, j = e % FFT_LMAX;
        const int l = d == 0 ? g.n[0] : d == 1 ? g.n[1] : g.n[2];
        if (j < l) {
            double s, c;
            sincospi(2.0 * (double)j / (double)l, &s, &c);
            tw[d][j] = make_double2(c, -s);
        }
    }
}

// kinetic half-step on nvec vectors: transform with sign `first` (true: the conjugate exponent),
// scale by e_k / N, transform back
__device__ static void fft_khalf(cplx *lds, const FftGeo &g, int nvec, int n, const cplx (*tw)[FFT_LMAX],
                                 const cplx *ek, bool first_inv)
{
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (g.n[d] > 1) fft_pass(lds, g, d, nvec, tw[d], first_inv);
    const double sc = 1.0 / (double)n;
    GAS const cplx *E = gp(ek);
    for (int e = threadIdx.x; e < n * nvec; e += blockDim.x) {
        const int k = e % n, v = e / n;
        const int p = v * g.ld + fft_pad(k, g.n[0]);
        lds[p] = cmul(lds[p], cscale(gld(E + k), sc));
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (g.n[d] > 1) fft_pass(lds, g, d, nvec, tw[d], !first_inv);
}

// exp(V) on the site index of every vector (rows of the column block, columns of the row block)
__device__ static void fft_v(const DChain &c, cplx *lds, const FftGeo &g, int nvec, int nt, int vsel)
{
    const int n = c.n;
    GAS const int32_t *cf = gp(c.conf + (size_t)(nt - 1) * n);
    const cplx *vt = vtab_lds(c, vsel);
    for (int e = threadIdx.x; e < n * nvec; e += blockDim.x) {
        const int s = e % n, v = e / n;
        const int p = v * g.ld + fft_pad(s, g.n[0]);
        lds[p] = cmul(vt[cf[s] - 1], lds[p]);
    }
    __syncthreads();
}

__device__ __forceinline__ FftGeo fft_geo(const DChain &c)
{
    FftGeo g;
    g.n[0] = c.fd0;
    g.n[1] = c.fd1;
    g.n[2] = c.fd2;
    g.ld = fft_ld(c.n, c.fd0);
    return g;
}

// left products (rows mixed) on fcb columns per workgroup; program bits as in k_cb_cols
extern "C" __global__ void __launch_bounds__(256)
k_fft_cols(const DChain *chains, int msel, int pt, int ksel, int prog, int vsel, int nt0, int nt1, int ntstep,
           int fcb)
{
    extern __shared__ cplx lds[];
    __shared__ cplx tw[3][FFT_LMAX];
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const int c0 = blockIdx.x * fcb;
    if (c0 >= n) return;
    const int ncol = min(fcb, n - c0);
    const FftGeo g = fft_geo(c);
    fft_twiddles(tw, g);
    cplx *M = chain_mat(c, msel, pt);
    const cplx *E = ksel ? c.ktm1 : c.kt;
    GAS cplx *Mc = gp(M) + (size_t)c0 * n;
    const int n0 = g.n[0], ld = g.ld;
    // the columns are one contiguous run of n * ncol elements
    stage_in_map<ST_BATCH>(lds, n * ncol, [=](int e) { return gld(Mc + (e)); },
                           [=](int e) { return (e / n) * ld + fft_pad(e % n, n0); });
    __syncthreads();
    const int nslice = (ntstep > 0 ? nt1 - nt0 : nt0 - nt1) + 1;
    for (int q = 0, nt = nt0; q < nslice; q++, nt += ntstep) {
        if (nt < 1 || nt > c.ltrot) break;
        if ((prog & 8) && c.lwrapu) fft_v(c, lds, g, ncol, nt, vsel);
        if ((prog & 1) && c.lwrapT) fft_khalf(lds, g, ncol, n, tw, E, false);
        if ((prog & 2) && c.lwrapu) fft_v(c, lds, g, ncol, nt, vsel);
        if ((prog & 4) && c.lwrapT) fft_khalf(lds, g, ncol, n, tw, E, false);
    }
    for (int base = threadIdx.x; base < n * ncol; base += ST_BATCH * blockDim.x) {
        cplx v[ST_BATCH];
#pragma unroll
        for (int u = 0; u < ST_BATCH; u++) {
            const int e = base + u * (int)blockDim.x;
            v[u] = e < n * ncol ? lds[(e / n) * ld + fft_pad(e % n, n0)] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < ST_BATCH; u++) {
            const int e = base + u * (int)blockDim.x;
            if (e < n * ncol) gst(Mc + (e), v[u]);
        }
    }
}

// right products (columns mixed) on fcb rows per workgroup: M <- M F^H diag(e_k) F, i.e. per row
// the conjugate transform, the diagonal, the forward transform
extern "C" __global__ void __launch_bounds__(256)
k_fft_rows(const DChain *chains, int msel, int pt, int ksel, int prog, int vsel, int nt, int fcb)
{
    extern __shared__ cplx lds[];
    __shared__ cplx tw[3][FFT_LMAX];
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const int r0 = blockIdx.x * fcb;
    if (r0 >= n) return;
    const int nrow = min(fcb, n - r0);
    const FftGeo g = fft_geo(c);
    fft_twiddles(tw, g);
    cplx *M = chain_mat(c, msel, pt);
    const cplx *E = ksel ? c.ktm1 : c.kt;
    GAS cplx *Mg = gp(M);
    const int n0 = g.n[0], ld = g.ld;
    // consecutive threads read consecutive rows of one column
    stage_in_map<ST_BATCH>(lds, n * nrow, [=](int e) { return gld(Mg + ((size_t)(r0 + e % nrow) + (size_t)(e / nrow) * n)); },
                           [=](int e) { return (e % nrow) * ld + fft_pad(e / nrow, n0); });
    __syncthreads();
    if ((prog & 8) && c.lwrapu) fft_v(c, lds, g, nrow, nt, vsel);
    if ((prog & 1) && c.lwrapT) fft_khalf(lds, g, nrow, n, tw, E, true);
    if ((prog & 2) && c.lwrapu) fft_v(c, lds, g, nrow, nt, vsel);
    if ((prog & 4) && c.lwrapT) fft_khalf(lds, g, nrow, n, tw, E, true);
    for (int base = threadIdx.x; base < n * nrow; base += ST_BATCH * blockDim.x) {
        cplx v[ST_BATCH];
#pragma unroll
        for (int u = 0; u < ST_BATCH; u++) {
            const int e = base + u * (int)blockDim.x;
            v[u] = e < n * nrow ? lds[(e % nrow) * ld + fft_pad(e / nrow, n0)] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < ST_BATCH; u++) {
            const int e = base + u * (int)blockDim.x;
            if (e < n * nrow) gst(Mg + ((size_t)(r0 + e % nrow) + (size_t)(e / nrow) * n), v[u]);
        }
    }
}

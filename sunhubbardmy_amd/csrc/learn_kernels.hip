// learn.hip -- learned block proposals for the finite-temperature sweep: the effective-action
// network (training kernels in FP32; inference accumulates in FP64 from the FP32 weights so
// that the accept/reject sequence is reproducible off the GPU), the HS-field features, and the
// block move with the exact determinant ratio of the flipped set.  Device side; the host side
// is learn.hip.
//
// The reference has no global or multi-site move (its lstglobal/lworm flags,
// src/model_para.f90:84-85, are never read); the block move below reuses its local-update
// quantities: Delta = delta_bmat(1, is), the field factor gaml(isp)/gaml(is) phase_ratio(1, is)
// and the Metropolis test on |Re(ratiotot phase)| / |Re(phase)|
// (src/dqmc_config_u/dqmc_update_delay.f90:45-76).

#define SHL_H 128
#define SHL_NWT (SHL_H * SHL_H + 7 * SHL_H + 1)
#define SHL_MB 64
#define SHL_TRAIN_WG 64
enum {
    WO_W1 = 0,
    WO_B1 = 4 * SHL_H,
    WO_W2 = 5 * SHL_H,
    WO_B2 = 5 * SHL_H + SHL_H * SHL_H,
    WO_W3 = 6 * SHL_H + SHL_H * SHL_H,
    WO_B3 = 7 * SHL_H + SHL_H * SHL_H
};

// per-chain device state of the proposals
struct PropChain {
    const int32_t *nbr;   // n * z neighbour sites
    int32_t *psum;        // ltrot * 4 per-slice partial sums of the features
    uint64_t *rng;        // splitmix64 state
    cplx *pa, *pb;        // n * 64: the two factors of the rank-|F| update of G (pb conjugated)
    int32_t *mapply;      // |F| of the block move just accepted, 0 if none (k_block_apply)
    double *stats;        // 5 counters (sh_get_proposal_stats)
    int z, nb2;           // list length, sum of list lengths (twice the bond count)
};

// what a block move reads of the network: weights and the input/output normalisation
struct PropNet {
    double xmu[4], xsd[4], ymu, ysd;
    const float *w;
    int nprop, pad;
};

// splitmix64 (Steele, Lea, Flood, OOPSLA 2014): one 64-bit state word per chain
__host__ __device__ __forceinline__ uint64_t smix_next(uint64_t &s)
{
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ double smix_unif(uint64_t &s) { return (double)(smix_next(s) >> 11) * 0x1.0p-53; }
static inline uint64_t smix_seed(int32_t seed) { return 0x9E3779B97F4A7C15ull * (uint64_t)(uint32_t)seed + 0x632BE59BD9B4E019ull; }

__device__ __forceinline__ int wave_sum_i(int v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_sum_f(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum of NV ints over the workgroup, returned to every thread (scr: 16 * NV ints of LDS)
template <int NV>
__device__ __forceinline__ void wg_sum_i(int (&v)[NV], int *scr)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum_i(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) scr[w * NV + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        int s = 0;
        for (int q = 0; q < nw; q++) s += scr[q * NV + k];
        v[k] = s;
    }
    __syncthreads();
}

// normalised network inputs x_k = (T_k / norm_k - xmu_k) / xsd_k of the totals T (sums over
// slices of the per-slice sums), as x_k = T_k fa_k + fb_k with the factors formed once
__device__ __forceinline__ void feat_of(const int (&T)[4], const double (&fa)[4], const double (&fb)[4], double (&x)[4])
{
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = fma((double)T[k], fa[k], fb[k]);
}

// tanh(a) = 1 - 2 / (e^{2a} + 1): one exponential and one division; exact limits +-1, absolute
// error a few ulp of 1 (the network only needs tanh to that absolute accuracy)
__device__ __forceinline__ double tanh_exp(double a) { return 1.0 - 2.0 / (exp(2.0 * a) + 1.0); }

// g(x) of the 4 -> 128 -> 128 -> 1 tanh network, evaluated by a workgroup of exactly 256
// threads; thread t owns half (t & 1) of row t >> 1 of W2 in w2r.  Every thread returns the
// same value.  h1: 128 doubles, red: 4 doubles of LDS.
__device__ __forceinline__ double mlp_eval64(const double (&x)[4], const float *W1s, const float *b1s, const float *b2s,
                                             const float *w3s, float b3, const float (&w2r)[64], double *h1, double *red)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t < SHL_H) {
        double a = (double)b1s[t];
        a += (double)W1s[4 * t] * x[0] + (double)W1s[4 * t + 1] * x[1] + (double)W1s[4 * t + 2] * x[2] +
             (double)W1s[4 * t + 3] * x[3];
        h1[t] = tanh_exp(a);
    }
    __syncthreads();
    const int j = t >> 1, half = t & 1;
    const double *hh = h1 + half * 64;
    double p0 = 0.0, p1 = 0.0;
#pragma unroll
    for (int k = 0; k < 64; k += 2) {
        p0 += (double)w2r[k] * hh[k];
        p1 += (double)w2r[k + 1] * hh[k + 1];
    }
    double p = p0 + p1;
    p += __shfl_xor(p, 1, 64);
    const double h2 = tanh_exp(p + (double)b2s[j]);
    double q = half ? 0.0 : (double)w3s[j] * h2;
    q = wave_sum(q);
    if (lane == 0) red[w] = q;
    __syncthreads();
    return (double)b3 + ((red[0] + red[1]) + (red[2] + red[3]));
}

// ------------------------------------------------------------------------------------------
// per-slice partial sums of the features: psum[t] = (sum s, sum s nn(s), sum s s(t+1),
// sum s nn(s(t+1))) over the sites of slice t; grid (ltrot, nchains)
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_prop_psum(const DChain *chains, const PropChain *props)
{
    const DChain &c = chains[blockIdx.y];
    const PropChain &p = props[blockIdx.y];
    const int t = blockIdx.x, n = c.n, lt = c.ltrot;
    if (t >= lt) return;
    __shared__ int scr[16 * 4];
    const int32_t *r0 = c.conf + (size_t)t * n, *r1 = c.conf + (size_t)((t + 1) % lt) * n;
    int v[4] = {0, 0, 0, 0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int s = 2 * r0[i] - 3, s1 = 2 * r1[i] - 3;
        int nn = 0, nt = 0;
        for (int k = 0; k < p.z; k++) {
            const int j = p.nbr[(size_t)i * p.z + k];
            nn += 2 * r0[j] - 3;
            nt += 2 * r1[j] - 3;
        }
        v[0] += s;
        v[1] += s * nn;
        v[2] += s * s1;
        v[3] += s * nt;
    }
    wg_sum_i<4>(v, scr);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) p.psum[(size_t)t * 4 + k] = v[k];
}

// features and energy per site of every chain: out[c] = (f0, f1, f2, f3, E/Ns) with
// E = -(nflr Re logweightf + sum over the field of log gaml(conf))
__global__ __launch_bounds__(64) void k_prop_features(const DChain *chains, const PropChain *props, double *out)
{
    const DChain &c = chains[blockIdx.x];
    const PropChain &p = props[blockIdx.x];
    double T[4] = {0.0, 0.0, 0.0, 0.0};    // integer sums, exact in double
    for (int t = threadIdx.x; t < c.ltrot; t += 64)
        for (int k = 0; k < 4; k++) T[k] += (double)p.psum[(size_t)t * 4 + k];
    for (int k = 0; k < 4; k++) T[k] = wave_sum(T[k]);
    if (threadIdx.x == 0) {
        const double ns = (double)c.n * c.ltrot, nbl = (double)p.nb2 * c.ltrot;
        const double n2 = 0.5 * (ns + T[0]), n1 = ns - n2;    // spins at conf 2 and conf 1
        const double e = -(c.nflr * c.sc[SC_LOGW] + n1 * log(fabs(c.gaml[0])) + n2 * log(fabs(c.gaml[1])));
        double *o = out + (size_t)blockIdx.x * 5;
        o[0] = T[0] / ns;
        o[1] = T[1] / nbl;
        o[2] = T[2] / ns;
        o[3] = T[3] / nbl;
        o[4] = e / ns;
    }
}

// ------------------------------------------------------------------------------------------
// One block move of slice nt per chain (one workgroup of 256 threads), right after the local
// update of that slice, with G the equal-time Green's function the local update used.
//  1. stage the spins of slices nt-1, nt, nt+1; refresh the partial sums the local update
//     changed; totals -> features -> g0
//  2. nprop Metropolis attempts on E_eff at uniform sites of slice nt (spins in LDS only)
//  3. F = flipped sites; M = 1 + Delta_F (1 - G)_FF; Gauss-Jordan inverse with partial
//     pivoting gives det M and M^-1
//  4. accept with min(1, |Re(R phase)| / |Re(phase)| exp(E_eff(C') - E_eff(C))),
//     R = det(M)^nflr prod qf
//  5. on accept G <- G - G_:F M^-1 Delta_F (E_F - G_F:) as a rank-|F| MFMA update, flip the
//     field, rotate the phase
// LDS: 3n spins, the neighbour lists, M (64 x 64), pivot rows/columns, network vectors and
// small weights.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_block_move(const DChain *chains, const PropChain *props, PropNet pn, int nt)
{
    extern __shared__ cplx lsh[];
    const DChain &c = chains[blockIdx.x];
    const PropChain &p = props[blockIdx.x];
    const int n = c.n, lt = c.ltrot, z = p.z, t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int tau = nt - 1, tm = (tau + lt - 1) % lt, tp = (tau + 1) % lt;
    cplx *Mm = (cplx *)lsh;                         // 64 x 64, leading dimension 64
    cplx *rowk = Mm + SHL_MB * SHL_MB;              // 64
    cplx *colk = rowk + SHL_MB;                     // 64
    cplx *dl = colk + SHL_MB;                       // 64: Delta of each flipped site
    double *h1 = (double *)(dl + SHL_MB);           // 128
    double *red = h1 + SHL_H;                       // 8
    float *W1s = (float *)(red + 8);                // 4H
    float *b1s = W1s + 4 * SHL_H, *b2s = b1s + SHL_H, *w3s = b2s + SHL_H;
    int *F = (int *)(w3s + SHL_H);                  // 64
    int *isF = F + SHL_MB;                          // 64: field value before the move
    int *piv = isF + SHL_MB;                        // 64
    int *cols = piv + SHL_MB;                       // 64
    int *iscr = cols + SHL_MB;                      // 16 * 12
    int *R = iscr + 16 * 12;                        // 3n spins (+-1): slices tm, tau, tp
    int *nb = R + 3 * n;                            // n * z neighbour lists
    int32_t *cf = c.conf + (size_t)tau * n;
    const int32_t *cm = c.conf + (size_t)tm * n, *cp = c.conf + (size_t)tp * n;
    cplx *G = c.G;

    for (int k = t; k < 4 * SHL_H; k += 256) W1s[k] = pn.w[WO_W1 + k];
    for (int k = t; k < SHL_H; k += 256) {
        b1s[k] = pn.w[WO_B1 + k];
        b2s[k] = pn.w[WO_B2 + k];
        w3s[k] = pn.w[WO_W3 + k];
    }
    const float b3 = pn.w[WO_B3];
    float w2r[64];
    {
        const float *src = pn.w + WO_W2 + (size_t)(t >> 1) * SHL_H + (t & 1) * 64;
#pragma unroll
        for (int k = 0; k < 64; k++) w2r[k] = src[k];
    }
    for (int i = t; i < n; i += 256) {
        R[i] = 2 * cm[i] - 3;
        R[n + i] = 2 * cf[i] - 3;
        R[2 * n + i] = 2 * cp[i] - 3;
    }
    for (int e = t; e < n * z; e += 256) nb[e] = p.nbr[e];
    __syncthreads();

    // partial sums of slices tm (rows tm, tau) and tau (rows tau, tp), and the other slices
    auto slice_sums = [&](int (&v)[12]) {
        for (int k = 0; k < 12; k++) v[k] = 0;
        for (int i = t; i < n; i += 256) {
            const int s0 = R[i], s1 = R[n + i], s2 = R[2 * n + i];
            int n0 = 0, n1 = 0, n2 = 0;
            for (int k = 0; k < z; k++) {
                const int j = nb[i * z + k];
                n0 += R[j];
                n1 += R[n + j];
                n2 += R[2 * n + j];
            }
            v[0] += s0; v[1] += s0 * n0; v[2] += s0 * s1; v[3] += s0 * n1;
            v[4] += s1; v[5] += s1 * n1; v[6] += s1 * s2; v[7] += s1 * n2;
        }
    };
    int v[12];
    slice_sums(v);
    for (int q = t; q < lt; q += 256)
        if (q != tm && q != tau)
            for (int k = 0; k < 4; k++) v[8 + k] += p.psum[(size_t)q * 4 + k];
    wg_sum_i<12>(v, iscr);
    int T[4];
    for (int k = 0; k < 4; k++) T[k] = v[k] + v[4 + k] + v[8 + k];

    const double ns = (double)n * lt, nbl = (double)p.nb2 * lt;
    const double escale = ns * pn.ysd;
    double fa[4], fb[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        fa[k] = 1.0 / (((k & 1) ? nbl : ns) * pn.xsd[k]);
        fb[k] = -pn.xmu[k] / pn.xsd[k];
    }
    uint64_t rs = p.rng[0];
    double x[4];
    feat_of(T, fa, fb, x);
    const double g0 = mlp_eval64(x, W1s, b1s, b2s, w3s, b3, w2r, h1, red);
    double gc = g0;
    int ninner = 0;
    for (int a = 0; a < pn.nprop; a++) {
        const double u1 = smix_unif(rs), u2 = smix_unif(rs);
        const int i = min((int)(u1 * n), n - 1);
        const int s = R[n + i];
        int nn = 0, st = 0;
        for (int k = 0; k < z; k++) {
            const int j = nb[i * z + k];
            nn += R[n + j];
            st += R[j] + R[2 * n + j];
        }
        const int T2[4] = {T[0] - 2 * s, T[1] - 4 * s * nn, T[2] - 2 * s * (R[i] + R[2 * n + i]), T[3] - 2 * s * st};
        feat_of(T2, fa, fb, x);
        const double gn = mlp_eval64(x, W1s, b1s, b2s, w3s, b3, w2r, h1, red);
        const double de = escale * (gn - gc);
        // every thread holds the same de and u2, so the decision is uniform; the spins of this
        // attempt were read before the evaluation's barriers
        if (de <= 0.0 || u2 < exp(-de)) {
            for (int k = 0; k < 4; k++) T[k] = T2[k];
            gc = gn;
            ninner++;
            if (t == 0) R[n + i] = -s;
        }
        __syncthreads();
    }
    const double u3 = smix_unif(rs);

    // the flipped set, in increasing site order (at most nprop <= 64 sites)
    if (w == 0) {
        int cnt = 0;
        for (int base = 0; base < n; base += 64) {
            const int i = base + lane;
            const int old = i < n ? cf[i] : 1;
            const bool ch = i < n && R[n + i] != 2 * old - 3;
            const unsigned long long mk = __ballot(ch);
            if (ch) {
                const int pos = cnt + __popcll(mk & ((1ull << lane) - 1ull));
                if (pos < SHL_MB) {
                    F[pos] = i;
                    isF[pos] = old;
                }
            }
            cnt += __popcll(mk);
        }
        if (lane == 0) iscr[0] = cnt;
    }
    __syncthreads();
    const int m = min(iscr[0], SHL_MB);
    bool acc = false;
    cplx unit = make_double2(1.0, 0.0);
    if (m > 0) {
        for (int a = t; a < m; a += 256) dl[a] = c.delta_bmat[isF[a] - 1];
        __syncthreads();
        for (int e = t; e < m * m; e += 256) {
            const int a = e % m, b = e / m;
            const cplx g = G[(size_t)F[a] + (size_t)F[b] * n];
            cplx val = cmul(dl[a], make_double2((a == b ? 1.0 : 0.0) - g.x, -g.y));
            if (a == b) val.x += 1.0;
            Mm[a + b * SHL_MB] = val;
        }
        __syncthreads();
        // in-place Gauss-Jordan with row interchanges
        double ldr = 0.0, ldi = 0.0;
        int nswap = 0;
        bool sing = false;
        for (int k = 0; k < m; k++) {
            if (w == 0) {
                double best = -1.0;
                int bi = k;
                if (k + lane < m) {
                    best = cabs2(Mm[(k + lane) + k * SHL_MB]);
                    bi = k + lane;
                }
                for (int o = 32; o > 0; o >>= 1) {
                    const double ob = __shfl_xor(best, o, 64);
                    const int oi = __shfl_xor(bi, o, 64);
                    if (ob > best || (ob == best && oi < bi)) {
                        best = ob;
                        bi = oi;
                    }
                }
                if (bi != k)
                    for (int q = lane; q < m; q += 64) {
                        const cplx tmp = Mm[k + q * SHL_MB];
                        Mm[k + q * SHL_MB] = Mm[bi + q * SHL_MB];
                        Mm[bi + q * SHL_MB] = tmp;
                    }
                if (lane == 0) piv[k] = bi;
            }
            __syncthreads();
            if (piv[k] != k) nswap++;
            const cplx pv = Mm[k + k * SHL_MB];
            const bool zero = (pv.x == 0.0 && pv.y == 0.0);
            sing = sing || zero;
            if (!zero) {
                ldr += 0.5 * log(cabs2(pv));
                ldi += atan2(pv.y, pv.x);
            }
            const cplx pinv = zero ? make_double2(0.0, 0.0) : cinv(pv);
            for (int q = t; q < m; q += 256) {
                rowk[q] = (q == k) ? pinv : cmul(Mm[k + q * SHL_MB], pinv);
                colk[q] = Mm[q + k * SHL_MB];
            }
            __syncthreads();
            for (int e = t; e < m * m; e += 256) {
                const int i = e % m, q = e / m;
                if (i == k) {
                    Mm[k + q * SHL_MB] = rowk[q];
                } else {
                    const cplx base = (q == k) ? make_double2(0.0, 0.0) : Mm[i + q * SHL_MB];
                    Mm[i + q * SHL_MB] = csub(base, cmul(colk[i], rowk[q]));
                }
            }
            __syncthreads();
        }
        // ratiotot = det(M)^nflr prod qf; det(M) = (-1)^nswap exp(ldr + i ldi)
        cplx qprod = make_double2(1.0, 0.0);
        for (int a = 0; a < m; a++) {
            const int is = isF[a], isp = 3 - is;
            qprod = cmul(qprod, cscale(c.phase_ratio[is - 1], c.gaml[isp - 1] / c.gaml[is - 1]));
        }
        const double th = c.nflr * (ldi + ((nswap & 1) ? M_PI : 0.0)) + atan2(qprod.y, qprod.x);
        const double lmag = c.nflr * ldr + 0.5 * log(cabs2(qprod));
        unit = make_double2(cos(th), sin(th));
        const cplx ph = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
        const double pr = fabs(cmul(unit, ph).x) / fabs(ph.x);
        const double lacc = lmag + log(pr) + escale * (gc - g0);
        acc = !sing && pr > 0.0 && log(u3) < lacc;
    }
    if (acc) {
        // column order of M^-1: the row interchanges undone on the columns, last first
        if (t == 0) {
            for (int j = 0; j < m; j++) cols[j] = j;
            for (int l = m - 1; l >= 0; l--) {
                const int q = piv[l], tmp = cols[l];
                cols[l] = cols[q];
                cols[q] = tmp;
            }
        }
        __syncthreads();
        // pa = -G_:F M^-1 (n x m), 16 columns per item; the item's m entries of G are loaded
        // eight at a time ahead of their use (clamped addresses), the sums in the order b = 0, 1, ...
        const int ngrp = (m + 15) / 16;
        GAS cplx *Gg = gp(G);
        for (int it = t; it < n * ngrp; it += 256) {
            const int row = it % n, g0i = (it / n) * 16;
            double ar[16], ai[16];
#pragma unroll
            for (int q = 0; q < 16; q++) { ar[q] = 0.0; ai[q] = 0.0; }
            for (int b0 = 0; b0 < m; b0 += 8) {
                cplx gbv[8];
#pragma unroll
                for (int u = 0; u < 8; u++) gbv[u] = gld(Gg + ((size_t)row + (size_t)F[min(b0 + u, m - 1)] * n));
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int b = b0 + u;
                    if (b < m) {
                        const cplx gb = gbv[u];
#pragma unroll
                        for (int q = 0; q < 16; q++) {
                            const int col = min(g0i + q, m - 1);
                            const cplx mi = Mm[b + cols[col] * SHL_MB];
                            ar[q] += gb.x * mi.x - gb.y * mi.y;
                            ai[q] += gb.x * mi.y + gb.y * mi.x;
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 16; q++)
                if (g0i + q < m) p.pa[(size_t)row + (size_t)(g0i + q) * n] = make_double2(-ar[q], -ai[q]);
        }
        // pb(col, a) = conj(Delta_a (delta(F_a, col) - G(F_a, col))), eight rows' entries in
        // flight per thread; k_block_apply then adds pa pb^T to G over tiles of the whole chip
        for (int base = t; base < n * m; base += 8 * 256) {
            cplx gv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int e = min(base + u * 256, n * m - 1), a = e / n, col = e - a * n;
                gv[u] = gld(Gg + ((size_t)F[a] + (size_t)col * n));
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int e = base + u * 256, a = e / n, col = e - a * n;
                if (e < n * m)
                    p.pb[(size_t)col + (size_t)a * n] =
                        cconj(cmul(dl[a], make_double2((F[a] == col ? 1.0 : 0.0) - gv[u].x, -gv[u].y)));
            }
        }
        for (int a = t; a < m; a += 256) cf[F[a]] = 3 - isF[a];
    } else {
        for (int a = t; a < m; a += 256) R[n + F[a]] = 2 * isF[a] - 3;
    }
    __syncthreads();
    // the partial sums of slices tm and tau for the field as it now stands
    slice_sums(v);
    wg_sum_i<12>(v, iscr);
    if (t == 0) {
        for (int k = 0; k < 4; k++) {
            p.psum[(size_t)tm * 4 + k] = v[k];
            p.psum[(size_t)tau * 4 + k] = v[4 + k];
        }
        if (acc) {
            cplx ph = cmul(make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]), unit);
            ph = cscale(ph, 1.0 / sqrt(cabs2(ph)));
            c.sc[SC_PHASE] = ph.x;
            c.sc[SC_PHASE + 1] = ph.y;
        }
        p.rng[0] = rs;
        p.stats[0] += 1.0;
        p.stats[1] += acc ? 1.0 : 0.0;
        p.stats[2] += m;
        p.stats[3] += acc ? m : 0;
        p.stats[4] += ninner;
        p.mapply[0] = acc ? m : 0;
    }
}

// The rank-|F| update G <- G + pa pb^T of the chains whose block move of this slice was
// accepted, as 64 x 64 MFMA tiles over the whole chip (grid: the largest chain's tiles x
// chains); every other chain's workgroups return at once.  pb is stored conjugated, so the
// conjugate transpose of zgemm_tile gives pb^T.
__global__ __launch_bounds__(256) void k_block_apply(const DChain *chains, const PropChain *props)
{
    const DChain &c = chains[blockIdx.z];
    const PropChain &p = props[blockIdx.z];
    const int m = p.mapply[0], n = c.n, ntl = (n + GT - 1) / GT;
    if (m <= 0 || (int)blockIdx.x >= ntl * ntl) return;
    const int tm = ((int)blockIdx.x % ntl) * GT, tn = ((int)blockIdx.x / ntl) * GT;
    zgemm_tile<0, 1>(p.pa, n, p.pb, n, c.G, n, n, n, m, tm, tn, 1.0, 0.0, 1.0, 0.0);
}

// ------------------------------------------------------------------------------------------
// Training: forward and backward of the network over a minibatch in FP32.  Workgroup b
// handles samples [b*per, min((b+1)*per, count)) of the index list and writes its gradient
// sum (weights layout) to gpart[b] and its sum of squared errors to lpart[b].  mode 1 only
// evaluates: pred[k] = ymu + ysd g(x_k).
// ------------------------------------------------------------------------------------------
struct NormF {
    double xmu[4], xsd[4], ymu, ysd;
};

__global__ __launch_bounds__(256) void k_mlp_batch(const float *wt, const double *samp, const int32_t *idx, int count,
                                                   int per, NormF nm, float scale, int mode, float *gpart,
                                                   double *lpart, double *pred)
{
    extern __shared__ cplx lsh[];
    float *W2s = (float *)lsh;                 // H x H
    float *W1s = W2s + SHL_H * SHL_H;          // 4H
    float *b1s = W1s + 4 * SHL_H, *b2s = b1s + SHL_H, *w3s = b2s + SHL_H;
    float *h1 = w3s + SHL_H;                   // H
    float *da2 = h1 + SHL_H;                   // H
    float *red = da2 + SHL_H;                  // 8
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, j = t >> 1, half = t & 1;
    for (int k = t; k < SHL_H * SHL_H; k += 256) W2s[k] = wt[WO_W2 + k];
    for (int k = t; k < 4 * SHL_H; k += 256) W1s[k] = wt[WO_W1 + k];
    for (int k = t; k < SHL_H; k += 256) {
        b1s[k] = wt[WO_B1 + k];
        b2s[k] = wt[WO_B2 + k];
        w3s[k] = wt[WO_W3 + k];
    }
    const float b3 = wt[WO_B3];
    __syncthreads();
    float w2r[64], g2[64];
#pragma unroll
    for (int k = 0; k < 64; k++) {
        w2r[k] = W2s[j * SHL_H + half * 64 + k];
        g2[k] = 0.0f;
    }
    float gw1[4] = {0.f, 0.f, 0.f, 0.f}, gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f;
    double loss = 0.0;
    const int s0 = blockIdx.x * per, s1 = min(count, s0 + per);
    for (int sidx = s0; sidx < s1; sidx++) {
        const int k0 = idx ? idx[sidx] : sidx;
        const double *sp = samp + (size_t)k0 * 5;
        float x[4];
#pragma unroll
        for (int k = 0; k < 4; k++) x[k] = (float)((sp[k] - nm.xmu[k]) / nm.xsd[k]);
        const float y = (float)((sp[4] - nm.ymu) / nm.ysd);
        if (t < SHL_H) {
            const float a = b1s[t] + W1s[4 * t] * x[0] + W1s[4 * t + 1] * x[1] + W1s[4 * t + 2] * x[2] + W1s[4 * t + 3] * x[3];
            h1[t] = tanhf(a);
        }
        __syncthreads();
        const float *hh = h1 + half * 64;
        float p0 = 0.f, p1 = 0.f;
#pragma unroll
        for (int k = 0; k < 64; k += 2) {
            p0 += w2r[k] * hh[k];
            p1 += w2r[k + 1] * hh[k + 1];
        }
        float pp = p0 + p1;
        pp += __shfl_xor(pp, 1, 64);
        const float h2 = tanhf(pp + b2s[j]);
        float q = half ? 0.f : w3s[j] * h2;
        q = wave_sum_f(q);
        if (lane == 0) red[w] = q;
        __syncthreads();
        const float out = b3 + ((red[0] + red[1]) + (red[2] + red[3]));
        if (mode == 1) {
            if (t == 0) pred[k0] = nm.ymu + nm.ysd * (double)out;
            __syncthreads();
            continue;
        }
        const float e = out - y;
        if (t == 0) loss += (double)e * e;
        const float dout = 2.0f * e * scale;
        const float a2 = dout * w3s[j] * (1.0f - h2 * h2);
        if (!half) {
            gw3 += dout * h2;
            gb2 += a2;
            da2[j] = a2;
        }
        if (t == 0) gb3 += dout;
#pragma unroll
        for (int k = 0; k < 64; k++) g2[k] += a2 * hh[k];
        __syncthreads();
        // dh1(k) = sum_j da2(j) W2(j, k) for k = t >> 1, the two halves of j summed by lane pairs
        float sacc = 0.f;
        const int kk = t >> 1;
        for (int jj = 0; jj < 64; jj++) sacc += da2[half * 64 + jj] * W2s[(half * 64 + jj) * SHL_H + kk];
        sacc += __shfl_xor(sacc, 1, 64);
        const float a1 = sacc * (1.0f - h1[kk] * h1[kk]);
        if (!half) {
            gb1 += a1;
#pragma unroll
            for (int k = 0; k < 4; k++) gw1[k] += a1 * x[k];
        }
        __syncthreads();
    }
    if (mode == 1) return;
    float *gp = gpart + (size_t)blockIdx.x * SHL_NWT;
#pragma unroll
    for (int k = 0; k < 64; k++) gp[WO_W2 + j * SHL_H + half * 64 + k] = g2[k];
    if (!half) {
        gp[WO_B2 + j] = gb2;
        gp[WO_W3 + j] = gw3;
        gp[WO_B1 + j] = gb1;
#pragma unroll
        for (int k = 0; k < 4; k++) gp[WO_W1 + 4 * j + k] = gw1[k];
    }
    if (t == 0) {
        gp[WO_B3] = gb3;
        lpart[blockIdx.x] = loss;
    }
}

// sum of the workgroup gradients in a fixed order, then one Adam step (lr > 0) or the plain
// gradient into gout; lout = (sum of the squared errors) / ldiv
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(const float *gpart, int nwg, const double *lpart, float *wt,
                                                         float *am, float *av, float lr, float c1, float c2,
                                                         float *gout, double *lout, double ldiv)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nwg; b++) s += lpart[b];
        *lout = s / ldiv;
    }
    if (i >= SHL_NWT) return;
    float g = 0.f;
    for (int b = 0; b < nwg; b++) g += gpart[(size_t)b * SHL_NWT + i];
    if (gout) gout[i] = g;
    if (lr > 0.f) {
        const float m = 0.9f * am[i] + 0.1f * g;
        const float v = 0.999f * av[i] + 0.001f * g * g;
        am[i] = m;
        av[i] = v;
        wt[i] -= lr * (m * c1) / (sqrtf(v * c2) + 1e-8f);
    }
}


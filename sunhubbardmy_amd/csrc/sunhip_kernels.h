// sunhip_kernels.h -- device-side data layout and small helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double2 cplx;                                   // (re, im), the complex(dp) layout
typedef double double4_t __attribute__((ext_vector_type(4)));

#define SH_KMAX 64   // largest delay depth / submatrix block

// Global-memory view of a device pointer.  The chain descriptors hold generic pointers, whose
// accesses compile to FLAT instructions; the result of a FLAT load can only be awaited together
// with every other outstanding memory and LDS access, so loads issued ahead of use stall the
// first consumer anyway.  Through an address-space-1 pointer the same accesses compile to
// GLOBAL instructions with per-load wait counts.  Only for pointers into device allocations.
// HIP's vector class has no members for address-space-1 objects, so complex elements are read
// and written through gld / gst (one 16-byte access each).
#define GAS __attribute__((address_space(1)))
typedef double gdbl2 __attribute__((ext_vector_type(2)));
template <class T>
__device__ __forceinline__ GAS T *gp(const T *p)
{
    return (GAS T *)p;
}
__device__ __forceinline__ double2 gld(const GAS double2 *p)
{
    const gdbl2 v = *(const GAS gdbl2 *)p;
    return make_double2(v.x, v.y);
}
// streaming variants (nontemporal): for data read or written once per pass, so that it does not
// displace operands that are re-read from L2
__device__ __forceinline__ double2 gld_nt(const GAS double2 *p)
{
    const gdbl2 v = __builtin_nontemporal_load((const GAS gdbl2 *)p);
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void gst_nt(GAS double2 *p, double2 v)
{
    gdbl2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (GAS gdbl2 *)p);
}
__device__ __forceinline__ void gst(GAS double2 *p, double2 v)
{
    gdbl2 w;
    w.x = v.x;
    w.y = v.y;
    *(GAS gdbl2 *)p = w;
}

// per-chain scalar slots (double array `sc`)
enum {
    SC_PHASE = 0,      // 2: phase (re, im)
    SC_LOGW = 2,       // 2: logweightf
    SC_LOGW_TMP = 4,   // 2: scratch logweightf from green_equaltime
    SC_ZLD = 6,        // 2: log det of the LU-factored matrix
    SC_MAXWRAP = 8,
    SC_MAXPH = 9,
    SC_NCHK = 10,
    SC_SUMLOG = 11,
    SC_ACC = 12,
    SC_TRIED = 13,
    SC_INFO = 14,      // first zero pivot (+1) of the last LU, 0 if none
    SC_NOBS = 15,
    SC_OBS = 16,       // 10: complex sums of sign, density/lq, zkint, zint_u, S_stag
    SC_TRACE_N = 26,
    SC_LDQ_SAVE = 27,  // 2: logdetQ of the saved chain point
    SC_NONFINITE = 29,
    SC_SIZE = 32       // 30, 31: per-sweep-type extras
};

// matrix selectors resolved per chain on the device
enum { M_G = 0, M_GS, M_W1, M_W2, M_W3, M_W4, M_US, M_VS, M_U, M_V, M_KT, M_KTM1, M_GT0, M_G0T, M_G00, M_GT0S, M_G0TS };

struct DChain {
    int n, lq, nn_nf, nn_lf, lcomp, nflr, kmode, nublock, update, lwrapT, lwrapu, phase_from_conf;
    int ltrot, nst, pad0, pad1;
    cplx *G, *Gs, *W1, *W2, *W3, *W4, *Us, *Vs;
    cplx *U, *V;          // (nst+1) points, n*n each
    double *D, *Ds;       // (nst+1)*n, n
    cplx *ldq;            // nst+1
    cplx *avec, *bvec;    // n*nublock
    cplx *tau;            // n
    cplx *wk;             // 16 x n work block of the blocked factorizations
    cplx *tblk;           // (n/16 + 1) triangular 16 x 16 factors of the block reflectors
    int32_t *jpvt, *piv;  // n
    int32_t *lperm;       // n: row order of the LU's interchanges (row i of P X is row lperm[i] of X)
    double *rwork;        // 4n
    int32_t *conf;        // ltrot*n
    double *unif;         // 2*n*ltrot
    uint32_t *rng;        // 624 words + index
    const int32_t *bonds; // nn_nf*nn_lf*2
    const cplx *kt, *ktm1;
    const double *stag;
    double *sc;           // SC_SIZE
    cplx *trace;
    int trace_cap, pad2;
    cplx bmat[4], bmat_inv[4], delta_bmat[12], phase_ratio[12];
    double gaml[4], parg[4];
    // time-displaced measurement (ltau): G(t,0), G(0,t), G(0,0), green_tau's scratch G(t,0) and
    // G(0,t), the znn_tau bins [ntdm+1][2][2][lq], cell x, y, z (1-based) and orbital of every site
    cplx *gt0, *g0t, *g00, *gt0s, *g0ts, *znn;
    const int32_t *sxy;
    double *dynerr;       // largest |propagated - green_tau| of G(t,0), G(0,t)
    int dl1, dl2, dl3, ntdm, ltau;
    double kin_weight;    // factor of the zkint bin (nflr * rt for the chain, else 1)
    // correlation bins of the equal-time measurement (null: off): 20 lq + 8 complex; per unit
    // cell the A site, nnlist(., 1), snlist(., 1) and the current partner; plaquette site cycles
    cplx *eqb;
    const int32_t *eqc, *plq;
    int zplq, pad3;
    double theta;
    // bond field of the t-V model (lwrapv): field [ltrot][nn_nf][n], the pass's bond draws
    // (2 nn_nf nn_lf per slice, in the order the families are updated), acceptance sums;
    // vtb = (bp, bm, bpi, bmi) of component 1 then 2, vdl = (dp, dm) of 1 then 2
    int lwrapv, pad5;
    int32_t *confv;
    double *unifv, *vacc;
    cplx vtb[8], vdl[4];
    // the DFT form of the kinetic term: lattice directions in site order (fd0 fastest); kt / ktm1
    // then hold exph0k / exph0kinv [n]
    int fd0, fd1, fd2, pad6;
    // plaquette field (lwrapp): field [ltrot][lqp] (1..4), the pass's plaquette draws (2 lqp per
    // slice, in the order the plaquettes are updated), acceptance sums; plaquette cycles
    // [lqp][zp]; per component the zp x zp column-major factors exp(i a eta V) (pb) and
    // exp(-i a eta V) (pbi), per (is, iflip) exp(i a (eta_isp - eta_is) V) - 1 (pdl, block
    // 3 is + iflip, 0-based); phase_ratio at iflip + 3 is, gaml, arg of the phase per component
    int lwrapp, zp, lqp, pad7;
    int32_t *confp;
    double *unifp, *pacc;
    const int32_t *plqc;
    const cplx *pb, *pbi, *pdl;
    cplx ppr[12];
    double pgaml[4], pparg[4];
};

__device__ __forceinline__ cplx cmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cplx cconj(cplx a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double cabs2(cplx a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b)
{
    // Smith's algorithm, as compilers lower Fortran complex division
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, d = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    } else {
        double r = b.x / b.y, d = b.x * r + b.y;
        return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
    }
}
__device__ __forceinline__ cplx cinv(cplx b) { return cdiv(make_double2(1.0, 0.0), b); }
// complex log with the principal branch (Im in (-pi, pi])
__device__ __forceinline__ cplx clog_(cplx z) { return make_double2(log(hypot(z.x, z.y)), atan2(z.y, z.x)); }
__device__ __forceinline__ cplx cexp_(cplx z)
{
    double e = exp(z.x);
    return make_double2(e * cos(z.y), e * sin(z.y));
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// the same sum without the LDS crossbar (needs all 64 lanes active): in-quad DPP exchanges
// (lane ^ 1, lane ^ 2), the half-row and row mirrors, then the 16- and 32-lane swaps of gfx950.
// Each step adds two partial sums of the same shape, so every lane ends with the same value.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false),
                            __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false));
}
// the lane's partner value at the same DPP / lane-swap steps as wave_sum_dpp, for one dword
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ int swap16_i(int v)
{
    const auto a = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    return (int)((threadIdx.x & 16) ? a[0] : a[1]);
}
__device__ __forceinline__ int swap32_i(int v)
{
    const auto a = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return (int)((threadIdx.x & 32) ? a[0] : a[1]);
}
__device__ __forceinline__ double swap16_d(double v)
{
    return __hiloint2double(swap16_i(__double2hiint(v)), swap16_i(__double2loint(v)));
}
__device__ __forceinline__ double swap32_d(double v)
{
    return __hiloint2double(swap32_i(__double2hiint(v)), swap32_i(__double2loint(v)));
}

__device__ __forceinline__ double wave_sum_dpp(double v)
{
    v += dpp_d<0xB1>(v);     // quad_perm [1, 0, 3, 2]
    v += dpp_d<0x4E>(v);     // quad_perm [2, 3, 0, 1]
    v += dpp_d<0x141>(v);    // row_half_mirror
    v += dpp_d<0x140>(v);    // row_mirror
    const bool odd16 = (threadIdx.x & 16) != 0, odd32 = (threadIdx.x & 32) != 0;
    {
        const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
        const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        v += __hiloint2double((int)(odd16 ? b[0] : b[1]), (int)(odd16 ? a[0] : a[1]));
    }
    {
        const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
        const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        v += __hiloint2double((int)(odd32 ? b[0] : b[1]), (int)(odd32 ? a[0] : a[1]));
    }
    return v;
}

// (largest value, smallest index among equal values) over the 64 lanes (all active), by the
// steps of wave_sum_dpp: the choice is associative and commutative, so every lane ends with the
// same pair whatever the order of the steps
__device__ __forceinline__ void wave_argmax_dpp(double &v, int &idx)
{
    auto take = [&](double v2, int i2) {
        if (v2 > v || (v2 == v && i2 < idx)) { v = v2; idx = i2; }
    };
    take(dpp_d<0xB1>(v), dpp_i<0xB1>(idx));
    take(dpp_d<0x4E>(v), dpp_i<0x4E>(idx));
    take(dpp_d<0x141>(v), dpp_i<0x141>(idx));
    take(dpp_d<0x140>(v), dpp_i<0x140>(idx));
    take(swap16_d(v), swap16_i(idx));
    take(swap32_d(v), swap32_i(idx));
}

__device__ __forceinline__ cplx *chain_mat(const DChain &c, int sel, int pt)
{
    size_t nn = (size_t)c.n * c.n;
    switch (sel) {
    case M_G: return c.G;
    case M_GS: return c.Gs;
    case M_W1: return c.W1;
    case M_W2: return c.W2;
    case M_W3: return c.W3;
    case M_W4: return c.W4;
    case M_US: return c.Us;
    case M_VS: return c.Vs;
    case M_U: return c.U + nn * pt;
    case M_V: return c.V + nn * pt;
    case M_KT: return (cplx *)c.kt;
    case M_KTM1: return (cplx *)c.ktm1;
    case M_GT0: return c.gt0;
    case M_G0T: return c.g0t;
    case M_G00: return c.g00;
    case M_GT0S: return c.gt0s;
    case M_G0TS: return c.g0ts;
    }
    return nullptr;
}

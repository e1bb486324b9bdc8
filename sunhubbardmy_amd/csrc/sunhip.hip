// sunhip.hip -- host side of libsunhip.so: the C ABI of include/sunhip.h.
// The slice loop mirrors src/dqmc_ft_core/dqmc_ft_sweep_b0.f90 and dqmc_ft_sweep_0b.f90; every
// launch advances all chains of the batch at once.
#include "kernels.hip"
#include "blocked.hip"
#include "pqmc_kernels.hip"
#include "learn_kernels.hip"
#include "dyn_kernels.hip"
#include "eqt_kernels.hip"
#include "tv_kernels.hip"
#include "plq_kernels.hip"
#include "fft_kernels.hip"
#include "../../include/sunhip.h"
#include "../../include/sunhip_learn.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

// timer classes: one per kernel group, in the order reported by sh_timers
enum {
    T_WRAP = 0,   // k_cb_cols / k_cb_rows / dense-K GEMMs on G, per slice
    T_UPDATE,     // k_update_delay / k_update_submatrix
    T_BPROD,      // B products of the stabilization (k_cb_cols over nwrap slices)
    T_QR,         // k_prepivot + k_qr
    T_FORMQ,      // k_formq
    T_GEMM,       // ZGEMMs of the stabilization and green_equaltime
    T_LU,         // k_lu
    T_SOLVE,      // k_lu_solve
    T_MEASURE,    // k_measure
    T_OTHER,      // elementwise helpers, wrap check
    T_PROPOSAL,   // k_block_move
    T_NCLS
};

struct EvRec {
    int cls;
    hipEvent_t a, b;
};

}  // namespace

// what every context of the library holds: the stream, its allocations, the last error and
// the per-kernel-group timers
struct ctx_base {
    int dev = 0;
    hipStream_t st = nullptr;
    std::vector<void *> allocs;
    std::string err;
    bool timing = false;
    std::vector<EvRec> evs;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    double tms[T_NCLS] = {0};
    double tcnt[T_NCLS] = {0};
};

struct sh_ctx : ctx_base {
    int nch = 0, nmax = 0, ltrot = 0, nwrap = 0, nst = 0, kmode = 0, update = 0;
    int fft_fcb = 0, fft_ld = 0;   // DFT kinetic form: vectors per workgroup, largest padded length
    std::vector<int> iwrap_nt, wrap_step;
    std::vector<DChain> hc;
    DChain *dc = nullptr;
    int32_t *d_seeds = nullptr;
    bool started = false;
    bool legacy_update = false;   // test hook: the L2-resident delay update kernel
    int ntdm = -1;                // time-displaced measurement: bins 0 .. ntdm, -1 = off
    size_t eqt_lds = 0;           // LDS of k_eqt, 0 while no chain has correlation bins
    int lwrapv = 0;               // the batch carries the t-V bond field
    int lwrapp = 0;               // the batch carries the plaquette field
    // push/pop_stage backups
    std::vector<void *> stage;
    std::vector<cplx *> trace_bufs;
    std::vector<int32_t> seeds;
    // learned block proposals (learn.hip)
    int prop_z = 0;
    bool prop_on = false, prop_local = true;
    int32_t *d_nbr = nullptr;
    double *d_feat = nullptr;
    std::vector<PropChain> hprop;
    PropChain *d_prop = nullptr;
    float *pw = nullptr;
    PropNet pnet = {};
    // hipGraph-captured passes (sh_set_graphs): one executable graph per pass kind and flags,
    // captured again whenever a setter changes what a pass launches (graph_epoch)
    bool use_graphs = false;
    int graph_epoch = 1;
    hipGraphExec_t gexec[8] = {};
    int gepoch[8] = {};
    // autotuner (sh_autotune): saved chain state per chain (G, HS field, scalars)
    std::vector<void *> tune_save;
};

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);            \
            return SH_E_HIP;                                                       \
        }                                                                          \
    } while (0)

static int fail(ctx_base *ctx, int code, const std::string &m)
{
    if (ctx) ctx->err = m;
    return code;
}

template <typename T>
static T *dalloc(ctx_base *ctx, size_t n)
{
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(T) + 256) != hipSuccess) return nullptr;
    hipMemset(p, 0, n * sizeof(T) + 256);
    ctx->allocs.push_back(p);
    return (T *)p;
}

template <typename T>
static T *dupload(ctx_base *ctx, const T *src, size_t n)
{
    T *p = dalloc<T>(ctx, n);
    if (p && src && n) hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice);
    return p;
}

// ---- timing ------------------------------------------------------------------------------
static hipEvent_t ev_get(ctx_base *ctx)
{
    if (ctx->pool_used == ctx->pool.size()) {
        hipEvent_t e;
        hipEventCreate(&e);
        ctx->pool.push_back(e);
    }
    return ctx->pool[ctx->pool_used++];
}

struct Tic {
    ctx_base *ctx;
    int cls;
    hipEvent_t a = nullptr;
    Tic(ctx_base *c, int k) : ctx(c), cls(k)
    {
        if (ctx->timing) {
            a = ev_get(ctx);
            hipEventRecord(a, ctx->st);
        }
    }
    ~Tic()
    {
        if (ctx->timing) {
            hipEvent_t b = ev_get(ctx);
            hipEventRecord(b, ctx->st);
            ctx->evs.push_back({cls, a, b});
        }
    }
};

static void timers_collect(ctx_base *ctx)
{
    if (ctx->evs.empty()) return;
    hipStreamSynchronize(ctx->st);
    for (auto &r : ctx->evs) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.a, r.b);
        ctx->tms[r.cls] += ms;
        ctx->tcnt[r.cls] += 1.0;
    }
    ctx->evs.clear();
    ctx->pool_used = 0;
}

// ---- launch helpers ----------------------------------------------------------------------
static dim3 grid_elem(sh_ctx *ctx) { return dim3((unsigned)(((size_t)ctx->nmax * ctx->nmax + 255) / 256), 1, ctx->nch); }
static dim3 grid_tiles32(sh_ctx *ctx)
{
    unsigned nb = (ctx->nmax + 31) / 32;
    return dim3(nb * nb, 1, ctx->nch);
}

template <int OPA, int OPB>
static void gemm(sh_ctx *ctx, int selA, int ptA, int selB, int ptB, int selC, int ptC, double ar = 1.0,
                 double br = 0.0)
{
    unsigned nb = (ctx->nmax + GT - 1) / GT;
    hipLaunchKernelGGL((k_zgemm<OPA, OPB>), dim3(nb, nb, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, selA, ptA, selB,
                       ptB, selC, ptC, ar, 0.0, br, 0.0);
}

static void copy_mat(sh_ctx *ctx, int dsel, int dpt, int ssel, int spt)
{
    hipLaunchKernelGGL(k_copy, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, dsel, dpt, ssel, spt);
}

static size_t lds_cols(sh_ctx *ctx) { return (size_t)ctx->nmax * CPB * sizeof(cplx); }
static size_t lds_fft(sh_ctx *ctx) { return (size_t)ctx->fft_fcb * ctx->fft_ld * sizeof(cplx); }
static size_t lds_rows(sh_ctx *ctx) { return (size_t)cb_ldr(ctx->nmax) * RPB * sizeof(cplx); }
// k_plq_prop: PLQ_CB lines of n + 1, the four factors, the plaquette cycles and the slice's field;
// k_plq_slice: one plaquette's G(:, P) and (1 - G)(P, :)
static size_t lds_plq_prop(sh_ctx *ctx)
{
    size_t b = 0;
    for (const auto &c : ctx->hc)
        b = std::max(b, (size_t)PLQ_CB * (c.n + 1) * sizeof(cplx) + (size_t)4 * c.zp * c.zp * sizeof(cplx) +
                            (size_t)c.lqp * (c.zp + 1) * sizeof(int));
    return b;
}
static size_t lds_plq_slice(sh_ctx *ctx)
{
    size_t b = 0;
    for (const auto &c : ctx->hc) b = std::max(b, (size_t)2 * c.zp * c.n * sizeof(cplx));
    return b;
}

// left product (rows mixed) on matrix sel: program bits as in k_cb_cols
static void left_pass(sh_ctx *ctx, int sel, int pt, int ksel, int prog, int vsel, int nt0, int nt1, int step)
{
    if (ctx->kmode == SH_K_CHECKERBOARD) {
        hipLaunchKernelGGL(k_cb_cols, dim3((ctx->nmax + CPB - 1) / CPB, 1, ctx->nch), dim3(256), lds_cols(ctx),
                           ctx->st, ctx->dc, sel, pt, ksel, prog, vsel, nt0, nt1, step);
        return;
    }
    if (ctx->kmode == SH_K_FFT) {
        const int f = ctx->fft_fcb;
        hipLaunchKernelGGL(k_fft_cols, dim3((ctx->nmax + f - 1) / f, 1, ctx->nch), dim3(256), lds_fft(ctx), ctx->st,
                           ctx->dc, sel, pt, ksel, prog, vsel, nt0, nt1, step, f);
        return;
    }
    // dense kinetic term: matrices K (ksel 0 -> urt, 1 -> urtm1) as GEMMs through W2
    const int kmat = ksel ? M_KTM1 : M_KT;
    const int nslice = (step > 0 ? nt1 - nt0 : nt0 - nt1) + 1;
    for (int q = 0, nt = nt0; q < nslice; q++, nt += step) {
        if (prog & 8) hipLaunchKernelGGL(k_vscale, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, sel, pt, 0, vsel, 0, nt);
        if (prog & 1) { gemm<0, 0>(ctx, kmat, 0, sel, pt, M_W2, 0); copy_mat(ctx, sel, pt, M_W2, 0); }
        if (prog & 2) hipLaunchKernelGGL(k_vscale, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, sel, pt, 0, vsel, 0, nt);
        if (prog & 4) { gemm<0, 0>(ctx, kmat, 0, sel, pt, M_W2, 0); copy_mat(ctx, sel, pt, M_W2, 0); }
    }
}

// right product (columns mixed) on matrix sel (G, or G(0,t) of the time-displaced pass)
static void right_pass(sh_ctx *ctx, int ksel, int prog, int vsel, int nt, int sel = M_G)
{
    if (ctx->kmode == SH_K_CHECKERBOARD) {
        hipLaunchKernelGGL(k_cb_rows, dim3((ctx->nmax + RPB - 1) / RPB, 1, ctx->nch), dim3(256), lds_rows(ctx),
                           ctx->st, ctx->dc, sel, 0, ksel, prog, vsel, nt);
        return;
    }
    if (ctx->kmode == SH_K_FFT) {
        const int f = ctx->fft_fcb;
        hipLaunchKernelGGL(k_fft_rows, dim3((ctx->nmax + f - 1) / f, 1, ctx->nch), dim3(256), lds_fft(ctx), ctx->st,
                           ctx->dc, sel, 0, ksel, prog, vsel, nt, f);
        return;
    }
    const int kmat = ksel ? M_KTM1 : M_KT;
    if (prog & 8) hipLaunchKernelGGL(k_vscale, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, sel, 0, 1, vsel, 0, nt);
    if (prog & 1) { gemm<0, 0>(ctx, sel, 0, kmat, 0, M_W2, 0); copy_mat(ctx, sel, 0, M_W2, 0); }
    if (prog & 2) hipLaunchKernelGGL(k_vscale, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, sel, 0, 1, vsel, 0, nt);
    if (prog & 4) { gemm<0, 0>(ctx, sel, 0, kmat, 0, M_W2, 0); copy_mat(ctx, sel, 0, M_W2, 0); }
}

// ---- stabilization -------------------------------------------------------------------------
static size_t lds_qr(sh_ctx *ctx) { return (size_t)ctx->nmax * 32; }
static bool blocked(sh_ctx *ctx) { return ctx->nmax <= BLK_NMAX; }
static size_t lds_blk(sh_ctx *ctx) { return ((size_t)ctx->nmax * PB + 2 * PB * PB) * sizeof(cplx) + ctx->nmax * 4; }

static void lu_factor(sh_ctx *ctx)
{
    Tic t(ctx, T_LU);
    if (blocked(ctx))
        hipLaunchKernelGGL(k_lu_blk, dim3(ctx->nch), dim3(BLK_NT), lds_blk(ctx), ctx->st, ctx->dc);
    else
        hipLaunchKernelGGL(k_lu, dim3(ctx->nch), dim3(1024), (size_t)ctx->nmax * 2 * sizeof(cplx), ctx->st, ctx->dc);
}

static void lu_solve(sh_ctx *ctx, int xsel)
{
    Tic t(ctx, T_SOLVE);
    if (ctx->nmax <= 16 * 4 * SCB_T) {
        // 16 right-hand sides per workgroup, a chain's workgroups on one XCD
        const int ncb = (ctx->nmax + 15) / 16, nch8 = (ctx->nch + 7) / 8 * 8;
        hipLaunchKernelGGL(k_solve_cb, dim3(nch8 * ncb), dim3(256), 0, ctx->st, ctx->dc, ctx->nch, ncb, xsel);
    } else if (blocked(ctx))
        hipLaunchKernelGGL(k_solve_blk, dim3(ctx->nch), dim3(BLK_NT), lds_blk(ctx), ctx->st, ctx->dc, xsel);
    else
        hipLaunchKernelGGL(k_lu_solve, dim3(ctx->nch), dim3(512), 0, ctx->st, ctx->dc, xsel);
}

// column-at-a-time QR of matrix sel (pivot: zgeqp3's column pivoting), rows held in registers
// in NQ = ceil(N / 64) slots
static void launch_qr(sh_ctx *ctx, int sel, int pivot, int pt)
{
    auto k = ctx->nmax <= 256 ? k_qr<4> : ctx->nmax <= 512 ? k_qr<8> : k_qr<16>;
    hipLaunchKernelGGL(k, dim3(ctx->nch), dim3(1024), lds_qr(ctx), ctx->st, ctx->dc, sel, pivot, pt);
}

// QR of W1 (first point: column pivoting) or of the prepivoted W1 diag(D[dpt]) -> point `pt`
static void udv_step(sh_ctx *ctx, bool first, int pt, int dpt)
{
    if (first) {
        {
            Tic t(ctx, T_QR);
            launch_qr(ctx, M_W1, 1, pt);
        }
        {
            Tic t(ctx, T_FORMQ);
            if (blocked(ctx)) {
                hipLaunchKernelGGL(k_tfactors, dim3(ctx->nch), dim3(BLK_NT), lds_blk(ctx), ctx->st, ctx->dc, M_W1);
                hipLaunchKernelGGL(k_formq_blk, dim3(ctx->nch), dim3(BLK_NT), lds_blk(ctx), ctx->st, ctx->dc, M_W1, pt);
            } else {
                hipLaunchKernelGGL(k_formq, dim3(ctx->nch), dim3(1024), (size_t)ctx->nmax * sizeof(cplx), ctx->st,
                                   ctx->dc, M_W1, pt);
            }
        }
        Tic t(ctx, T_OTHER);
        hipLaunchKernelGGL(k_v_from_qrcp, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, M_W1, pt);
        return;
    }
    {
        Tic t(ctx, T_QR);
        hipLaunchKernelGGL(k_prepivot, dim3(ctx->nch), dim3(1024), (size_t)ctx->nmax * 12, ctx->st, ctx->dc, dpt);
        if (blocked(ctx))
            hipLaunchKernelGGL(k_qr_blk, dim3(ctx->nch), dim3(BLK_NT), lds_blk(ctx), ctx->st, ctx->dc, M_W2, pt);
        else
            launch_qr(ctx, M_W2, 0, pt);
    }
    {
        Tic t(ctx, T_FORMQ);
        if (blocked(ctx))
            hipLaunchKernelGGL(k_formq_blk, dim3(ctx->nch), dim3(BLK_NT), lds_blk(ctx), ctx->st, ctx->dc, M_W2, pt);
        else
            hipLaunchKernelGGL(k_formq, dim3(ctx->nch), dim3(1024), (size_t)ctx->nmax * sizeof(cplx), ctx->st,
                               ctx->dc, M_W2, pt);
    }
    {
        Tic t(ctx, T_OTHER);
        // V_new = D^-1 R P^T V_prev: W3 = D^-1 triu R, W4 = P^T V[dpt]
        hipLaunchKernelGGL(k_after_qr, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, pt, (int)M_V, dpt);
    }
    Tic t(ctx, T_GEMM);
    gemm<0, 0>(ctx, M_W3, 0, M_W4, 0, M_V, pt);
}

// bond-field props of family nf on rows (side 0) or columns (side 1) of matrix sel
static void vprop(sh_ctx *ctx, int sel, int nf, int mode, int side, int nt)
{
    const unsigned nb = (unsigned)(((size_t)ctx->nmax * ctx->nmax / 2 + 255) / 256);   // n nn_lf <= n^2 / 2
    hipLaunchKernelGGL(k_vprop, dim3(nb, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, sel, 0, nf, mode, side, nt);
}
// the bond update follows the build's algorithm as the on-site one does: the delay form
// (dqmc_config_v/dqmc_update_delay.f90) or the submatrix form (dqmc_update_submatrix.f90) with
// the build's depth, else the fast form (dqmc_update.f90)
static void vupdate(sh_ctx *ctx, int nt, int nf, int slice, int p)
{
    int kmax = 0;
    for (auto &c : ctx->hc) kmax = std::max(kmax, c.nublock);
    if (ctx->update == SH_U_DELAY && kmax >= 2 && !ctx->legacy_update) {
        hipLaunchKernelGGL(k_vupdate_delay, dim3(ctx->nch), dim3(256), (size_t)4 * SH_KMAX * sizeof(cplx), ctx->st,
                           ctx->dc, nt, nf, slice, p);
        return;
    }
    if (ctx->update == SH_U_SUBMATRIX && !ctx->legacy_update) {
        const int K = std::max(2, std::min(kmax, SH_KMAX) & ~1);
        hipLaunchKernelGGL(k_vupdate_submatrix, dim3(ctx->nch), dim3(256), (size_t)(K * (K + 1) + 8 * K) * sizeof(cplx),
                           ctx->st, ctx->dc, nt, nf, slice, p);
        return;
    }
    hipLaunchKernelGGL(k_vupdate, dim3(ctx->nch), dim3(256), (size_t)4 * ctx->nmax * sizeof(cplx), ctx->st, ctx->dc,
                       nt, nf, slice, p);
}
// every plaquette of slice nt on a chain matrix: rows (side 0) or columns (side 1), the factors
// exp(i a eta V) (inv 0) or their inverses (inv 1), plaquettes 1..lqp (rev 0) or lqp..1 (rev 1)
static void pprop(sh_ctx *ctx, int sel, int pt, int side, int inv, int rev, int nt)
{
    hipLaunchKernelGGL(k_plq_prop, dim3((ctx->nmax + PLQ_CB - 1) / PLQ_CB, 1, ctx->nch), dim3(256),
                       lds_plq_prop(ctx), ctx->st, ctx->dc, sel, pt, side, inv, rev, nt);
}
// the plaquettes of slice nt on G in pass order: 0b (dir 0) B_p G B_p^-1 for p = 1..lqp, b0
// (dir 1) B_p^-1 G B_p for p = lqp..1 (dqmc_ft_sweep_0b.f90:82-93, dqmc_ft_sweep_b0.f90:91-100);
// with updates one sequential kernel per chain, else all rows and all columns at once
static void plq_slice(sh_ctx *ctx, int nt, int dir, int slice, bool upd)
{
    if (upd) {
        hipLaunchKernelGGL(k_plq_slice, dim3(ctx->nch), dim3(256), lds_plq_slice(ctx), ctx->st, ctx->dc, nt, dir,
                           slice, 1);
    } else if (dir == 0) {
        pprop(ctx, M_G, 0, 0, 0, 0, nt);
        pprop(ctx, M_G, 0, 1, 1, 0, nt);
    } else {
        pprop(ctx, M_G, 0, 1, 0, 1, nt);
        pprop(ctx, M_G, 0, 0, 1, 1, nt);
    }
}
// one slice of Bmat_left_forward with the bond or the plaquette field: K-half, the plaquettes
// 1..lq, the bond families 1..nn_nf (nflag 2 then 1), V, K-half (src/dqmc_ft_core.f90:80-119); of
// Bmat_left_forward_hc: K-half, V^H, the families nn_nf..1, the plaquettes lq..1 with bmat^H =
// bmat_inv, K-half (:128-171: the conjugated product takes the fields in reverse order)
static void bprod_slice_v(sh_ctx *ctx, int nt, int hc)
{
    const int nfam = ctx->hc[0].nn_nf;
    left_pass(ctx, M_W1, 0, 0, hc ? 1 | 2 : 1, hc ? 2 : 0, nt, nt, 1);
    if (ctx->lwrapp && !hc) pprop(ctx, M_W1, 0, 0, 0, 0, nt);
    for (int q = 0; q < nfam && ctx->lwrapv; q++) {
        const int nf = hc ? nfam - 1 - q : q;
        vprop(ctx, M_W1, nf, 0, 0, nt);
        vprop(ctx, M_W1, nf, 1, 0, nt);
    }
    if (ctx->lwrapp && hc) pprop(ctx, M_W1, 0, 0, 1, 1, nt);
    left_pass(ctx, M_W1, 0, 0, hc ? 4 : 2 | 4, 0, nt, nt, 1);
}

// dqmc_ft_stablize_0b_qr(n): point n of the right chain from point n-1
static void stablize_0b(sh_ctx *ctx, int n)
{
    {
        Tic t(ctx, T_BPROD);
        copy_mat(ctx, M_W1, 0, M_U, n - 1);
        int ws1 = ctx->wrap_step[2 * (n - 1)], ws2 = ctx->wrap_step[2 * (n - 1) + 1];
        // Bmat_left_forward: slices ws1..ws2, each K-half, V, K-half (src/dqmc_ft_core.f90:70-119)
        if (ctx->lwrapv || ctx->lwrapp)
            for (int nt = ws1; nt <= ws2; nt++) bprod_slice_v(ctx, nt, 0);
        else
            left_pass(ctx, M_W1, 0, 0, 1 | 2 | 4, 0, ws1, ws2, +1);
    }
    udv_step(ctx, n == 1, n, n - 1);
}

// dqmc_ft_stablize_b0_qr(n): point n-1 of the left chain from point n
static void stablize_b0(sh_ctx *ctx, int n)
{
    {
        Tic t(ctx, T_BPROD);
        copy_mat(ctx, M_W1, 0, M_U, n);
        int ws1 = ctx->wrap_step[2 * (n - 1)], ws2 = ctx->wrap_step[2 * (n - 1) + 1];
        // Bmat_left_forward_hc: slices ws2 down to ws1 with conj(bmat) (src/dqmc_ft_core.f90:121-171)
        if (ctx->lwrapv || ctx->lwrapp)
            for (int nt = ws2; nt >= ws1; nt--) bprod_slice_v(ctx, nt, 1);
        else
            left_pass(ctx, M_W1, 0, 0, 1 | 2 | 4, 2, ws2, ws1, -1);
    }
    udv_step(ctx, n == ctx->nst, n - 1, n);
}

// green_equaltime into Gs + scratch logweight (rsaved/lsaved select the saved chain point)
static void green(sh_ctx *ctx, int rsaved, int rpt, int lsaved, int lpt)
{
    const int URs = rsaved ? M_US : M_U, VRs = rsaved ? M_VS : M_V;
    const int ULs = lsaved ? M_US : M_U, VLs = lsaved ? M_VS : M_V;
    {
        Tic t(ctx, T_GEMM);
        gemm<1, 0>(ctx, URs, rsaved ? 0 : rpt, ULs, lsaved ? 0 : lpt, M_W1, 0);   // uu = UR^H UL
        gemm<0, 1>(ctx, VRs, rsaved ? 0 : rpt, VLs, lsaved ? 0 : lpt, M_W2, 0);   // vv = VR VL^H
    }
    {
        Tic t(ctx, T_OTHER);
        hipLaunchKernelGGL(k_green_m, grid_tiles32(ctx), dim3(256), 0, ctx->st, ctx->dc, 0, rsaved, rpt, lsaved, lpt);
    }
    lu_factor(ctx);
    {
        Tic t(ctx, T_OTHER);
        hipLaunchKernelGGL(k_green_rhs, grid_tiles32(ctx), dim3(256), 0, ctx->st, ctx->dc, rsaved, rpt);
    }
    lu_solve(ctx, M_W4);
    {
        Tic t(ctx, T_OTHER);
        hipLaunchKernelGGL(k_green_ulscale, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, lsaved, lpt);
    }
    {
        Tic t(ctx, T_GEMM);
        gemm<0, 0>(ctx, M_W1, 0, M_W4, 0, M_GS, 0);
    }
    Tic t(ctx, T_OTHER);
    hipLaunchKernelGGL(k_green_logw, dim3(ctx->nch), dim3(256), 0, ctx->st, ctx->dc, 0, rsaved, rpt, lsaved, lpt,
                       (int)SC_LOGW_TMP);
}

// green_tau (src/dqmc_ft_core/green_tau.f90:1-122) at right point rpt against the saved left
// point: G(t,t) into Gs with the scratch weight (as green), G(t,0) and G(0,t) into the scratch
// gt0s / g0ts, G(0,0) into g00.  The explicit inverses of the reference become LU solves:
// G(t,0) = (UL dlmax^-1) M^-1 (drmin VR); with A = VR^-1 drmax^-1 and M2 the G(0,0) matrix,
// G(0,0) = A M2^-1 (dlmax^-1 (VL^H)^-1) and G(0,t) = A M2^-1 (-dlmin UL^H).
static void green_tau(sh_ctx *ctx, int rpt)
{
    auto elem = [&](int op, int dsel) {
        hipLaunchKernelGGL(k_dyn_elem, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, op, dsel, rpt);
    };
    green(ctx, 0, rpt, 1, 0);          // Gs = G(t,t), W1 = UL dlmax^-1, W3 = LU of M, SC_LOGW_TMP
    {
        Tic t(ctx, T_OTHER);
        hipLaunchKernelGGL(k_dyn_info, dim3((ctx->nch + 63) / 64), dim3(64), 0, ctx->st, ctx->dc, ctx->nch, 0);
        elem(2, M_W4);                 // W4 = drmin VR
    }
    lu_solve(ctx, M_W4);
    {
        Tic t(ctx, T_GEMM);
        gemm<0, 0>(ctx, M_W1, 0, M_W4, 0, M_GT0S, 0);
    }
    // VR^-1 into W4, (VL^H)^-1 into W2
    {
        Tic t(ctx, T_OTHER);
        copy_mat(ctx, M_W3, 0, M_V, rpt);
    }
    lu_factor(ctx);
    {
        Tic t(ctx, T_OTHER);
        elem(1, M_W4);
    }
    lu_solve(ctx, M_W4);
    {
        Tic t(ctx, T_OTHER);
        elem(3, M_W3);
    }
    lu_factor(ctx);
    {
        Tic t(ctx, T_OTHER);
        elem(1, M_W2);
    }
    lu_solve(ctx, M_W2);
    {
        Tic t(ctx, T_GEMM);
        gemm<0, 0>(ctx, M_W2, 0, M_W4, 0, M_W1, 0);     // (VL^H)^-1 VR^-1
        gemm<1, 0>(ctx, M_US, 0, M_U, rpt, M_W3, 0);     // UL^H UR
    }
    {
        Tic t(ctx, T_OTHER);
        elem(4, M_W3);                                   // M2
    }
    lu_factor(ctx);
    {
        Tic t(ctx, T_OTHER);
        elem(5, M_W2);                                   // dlmax^-1 (VL^H)^-1
        elem(6, M_W4);                                   // A = VR^-1 drmax^-1
    }
    lu_solve(ctx, M_W2);
    {
        Tic t(ctx, T_GEMM);
        gemm<0, 0>(ctx, M_W4, 0, M_W2, 0, M_G00, 0);
    }
    {
        Tic t(ctx, T_OTHER);
        elem(7, M_W2);                                   // -dlmin UL^H
    }
    lu_solve(ctx, M_W2);
    {
        Tic t(ctx, T_GEMM);
        gemm<0, 0>(ctx, M_W4, 0, M_W2, 0, M_G0TS, 0);
    }
    Tic t(ctx, T_OTHER);
    hipLaunchKernelGGL(k_dyn_info, dim3((ctx->nch + 63) / 64), dim3(64), 0, ctx->st, ctx->dc, ctx->nch, 1);
}

// green_equaltimebb: G(beta,beta) from the right chain point nst
static void green_bb(sh_ctx *ctx, int rpt)
{
    hipLaunchKernelGGL(k_green_m, grid_tiles32(ctx), dim3(256), 0, ctx->st, ctx->dc, 1, 0, rpt, 0, 0);
    lu_factor(ctx);
    hipLaunchKernelGGL(k_green_rhs, grid_tiles32(ctx), dim3(256), 0, ctx->st, ctx->dc, 0, rpt);
    lu_solve(ctx, M_W4);
    Tic t(ctx, T_OTHER);
    copy_mat(ctx, M_GS, 0, M_W4, 0);
    hipLaunchKernelGGL(k_green_logw, dim3(ctx->nch), dim3(256), 0, ctx->st, ctx->dc, 1, 0, rpt, 0, 0, (int)SC_LOGW_TMP);
}

static void wrap_replace(sh_ctx *ctx, int compare)
{
    Tic t(ctx, T_OTHER);
    hipLaunchKernelGGL(k_wrap, dim3(ctx->nch), dim3(1024), 0, ctx->st, ctx->dc, compare);
}

static void set_identity(sh_ctx *ctx, int pt)
{
    hipLaunchKernelGGL(k_set_identity, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, pt);
}

// register-resident delay update: the template depth covering every chain's nublock, and the
// thread count 2 * rows; 0 when the register budget does not allow it
// the submatrix kernel's LDS: Gamma^-1, the K column scales and dss staged columns and rows of G;
// dss = the largest staging depth <= 8 within SUB_LDS_CAP (>= 1 up to N = 1024 at K = 64)
// the kernel also holds 44 KiB of static LDS (the MFMA tiles of the flush, the per-site partial
// sums); static plus dynamic stays within the 160 KiB of a CU
#define SUB_LDS_CAP ((size_t)112 * 1024)
static size_t lds_sub(int n, int K, int dss) { return (size_t)(K * (K + 1) + K + 2 * (size_t)dss * n) * sizeof(cplx); }
static int sub_stage(int n, int K)
{
    int dss = 8;
    while (dss > 1 && lds_sub(n, K, dss) > SUB_LDS_CAP) dss--;
    return dss;
}
static int delay_reg_depth(sh_ctx *ctx, int *threads)
{
    int kmax = 0;
    for (auto &c : ctx->hc) kmax = std::max(kmax, c.nublock);
    int kt = kmax <= 4 ? 4 : kmax <= 8 ? 8 : kmax <= 16 ? 16 : kmax <= 32 ? 32 : 0;
    int rows = (ctx->nmax + 31) / 32 * 32;
    int nt = 2 * rows;
    if (kt == 0 || nt > (kt >= 8 ? 512 : 1024)) return 0;
    // LDS: diag, DSB staged columns and rows, two row buffers, uniforms and HS field
    if ((size_t)(1 + 2 * DSB) * ctx->nmax * sizeof(cplx) + 4 * kt * sizeof(cplx) + (size_t)ctx->nmax * 20 > 150 * 1024)
        return 0;
    *threads = nt;
    return kt;
}

static void update_slice(sh_ctx *ctx, int nt, int slice_in_pass)
{
    Tic t(ctx, T_UPDATE);
    int nthr = 0;
    int kt = (ctx->update == SH_U_DELAY && !ctx->legacy_update) ? delay_reg_depth(ctx, &nthr) : 0;
    if (kt) {
        size_t lds = (size_t)(ctx->nmax + 2 * DSB * ctx->nmax + 4 * kt) * sizeof(cplx) + (size_t)ctx->nmax * 20;
        switch (kt) {
        case 4: hipLaunchKernelGGL(k_update_delay_reg<4>, dim3(ctx->nch), dim3(nthr), lds, ctx->st, ctx->dc, nt, slice_in_pass); break;
        case 8: hipLaunchKernelGGL(k_update_delay_reg<8>, dim3(ctx->nch), dim3(nthr), lds, ctx->st, ctx->dc, nt, slice_in_pass); break;
        case 16: hipLaunchKernelGGL(k_update_delay_reg<16>, dim3(ctx->nch), dim3(nthr), lds, ctx->st, ctx->dc, nt, slice_in_pass); break;
        default: hipLaunchKernelGGL(k_update_delay_reg<32>, dim3(ctx->nch), dim3(nthr), lds, ctx->st, ctx->dc, nt, slice_in_pass); break;
        }
    } else if (ctx->update == SH_U_DELAY) {
        // diag(G), two rows of the delayed vectors, the slice's uniforms and HS field
        size_t lds = (size_t)(ctx->nmax + 2 * 64) * sizeof(cplx) + (size_t)ctx->nmax * 20;
        hipLaunchKernelGGL(k_update_delay, dim3(ctx->nch), dim3(256), lds, ctx->st, ctx->dc, nt, slice_in_pass);
    } else {
        int K = 0;
        for (auto &c : ctx->hc) K = std::max(K, c.nublock);
        const int dss = sub_stage(ctx->nmax, K);
        hipLaunchKernelGGL(k_update_submatrix, dim3(ctx->nch), dim3(256), lds_sub(ctx->nmax, K, dss), ctx->st, ctx->dc, nt,
                           slice_in_pass, dss);
    }
}

static void measure(sh_ctx *ctx)
{
    Tic t(ctx, T_MEASURE);
    unsigned nb = (ctx->nmax + 31) / 32;
    hipLaunchKernelGGL(k_measure_tiles, dim3(nb * (nb + 1) / 2, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc);
    hipLaunchKernelGGL(k_measure, dim3(ctx->nch), dim3(256), 0, ctx->st, ctx->dc);
    if (ctx->eqt_lds) hipLaunchKernelGGL(k_eqt, dim3(ctx->nch), dim3(256), ctx->eqt_lds, ctx->st, ctx->dc);
}

static int check_launch(sh_ctx *ctx)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, SH_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
}

// A pass: its launches issued directly, or (sh_set_graphs, timers off) captured once into an
// executable hipGraph per pass kind and flags and replayed, so that a pass costs one host launch.
// Kernel arguments are captured by value; the chain descriptors are read from device memory at
// run time.  Every setter that changes which kernels a pass launches, their shapes or by-value
// arguments (proposals, correlation bins, the time-displaced pass, the autotuned depth and tiles,
// the delay-kernel switch) advances graph_epoch, which retires the captured graphs.
template <class Body>
static int run_pass(sh_ctx *ctx, int key, Body body)
{
    if (!ctx->use_graphs || ctx->timing) {
        body();
        return check_launch(ctx);
    }
    if (!ctx->gexec[key] || ctx->gepoch[key] != ctx->graph_epoch) {
        if (ctx->gexec[key]) {
            hipGraphExecDestroy(ctx->gexec[key]);
            ctx->gexec[key] = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
        body();
        hipError_t e = hipStreamEndCapture(ctx->st, &g);
        if (e != hipSuccess || !g) return fail(ctx, SH_E_HIP, std::string("pass capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&ctx->gexec[key], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            ctx->gexec[key] = nullptr;
            return fail(ctx, SH_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        ctx->gepoch[key] = ctx->graph_epoch;
    }
    HIPCHK(hipGraphLaunch(ctx->gexec[key], ctx->st));
    return check_launch(ctx);
}

// learned block proposals (learn.hip): the per-pass refresh of the feature sums and one block
// move of a slice, both no-ops while proposals are off
static void prop_pass_start(sh_ctx *ctx);
static void block_move(sh_ctx *ctx, int nt);

// ---- C ABI ---------------------------------------------------------------------------------
extern "C" {

int sh_create(sh_ctx **out, int nchains, const sh_model *m, int device)
{
    if (!out || nchains <= 0 || !m) return SH_E_ARG;
    *out = nullptr;
    sh_ctx *ctx = new sh_ctx();
    ctx->dev = device;
    ctx->nch = nchains;
    ctx->ltrot = m[0].ltrot;
    ctx->nwrap = m[0].nwrap;
    ctx->nst = m[0].nst;
    ctx->kmode = m[0].kmode;
    ctx->update = m[0].update;
    for (int c = 0; c < nchains; c++) {
        const sh_model &q = m[c];
        if (q.ltrot != ctx->ltrot || q.nwrap != ctx->nwrap || q.nst != ctx->nst || q.kmode != ctx->kmode ||
            q.update != ctx->update) {
            int rc = fail(ctx, SH_E_ARG, "all chains of a batch must share ltrot, nwrap, nst, kmode and update");
            delete ctx;
            return rc;
        }
        if (!q.iwrap_nt || !q.wrap_step || !q.bonds || !q.urt || !q.urtm1 || !q.bmat || !q.bmat_inv ||
            !q.phase_arg || !q.delta_bmat || !q.phase_ratio || !q.gaml || !q.stag || q.ltrot <= 0) {
            delete ctx;
            return SH_E_ARG;
        }
        if (q.ndim <= 0 || q.ndim > 1024 || q.nst <= 0 || q.lcomp < 2 || q.lcomp > 4 || q.nublock < 1 ||
            q.nublock > 64 || (q.update == SH_U_SUBMATRIX && q.nublock > 64) || q.nflr < 1) {
            delete ctx;
            return SH_E_ARG;
        }
        if ((q.kmode == SH_K_CHECKERBOARD || q.vtab) && (q.nn_nf <= 0 || q.nn_lf <= 0)) { delete ctx; return SH_E_ARG; }
        if (q.kmode != SH_K_CHECKERBOARD && q.kmode != SH_K_DENSE && q.kmode != SH_K_FFT) { delete ctx; return SH_E_ARG; }
        if (q.kmode == SH_K_FFT) {
            auto ok = [](int l) { return (l >= 1 && l <= 16) || l == 18 || l == 20 || l == 24; };
            const int *d = q.fft_dims;
            if (!ok(d[0]) || !ok(d[1]) || !ok(d[2]) || d[0] * d[1] * d[2] != q.ndim) {
                int rc = fail(ctx, SH_E_ARG, "FFT kinetic term: fft_dims must multiply to ndim, each 1..16, 18, 20 or 24");
                delete ctx;
                return rc;
            }
            ctx->fft_ld = std::max(ctx->fft_ld, fft_ld(q.ndim, d[0]));
        }
        if ((q.vtab != nullptr) != (m[0].vtab != nullptr) || (q.vtab && !q.vdelta)) {
            int rc = fail(ctx, SH_E_ARG, "all chains of a batch must have the bond field or none");
            delete ctx;
            return rc;
        }
        // the pass loops take the family count and the on-site switch of the batch, so chains
        // with a bond field share nn_nf, and every chain shares lwrapu (U != 0)
        if ((q.vtab && q.nn_nf != m[0].nn_nf) || q.lwrapu != m[0].lwrapu) {
            int rc = fail(ctx, SH_E_ARG, "chains of a batch must share lwrapu, and nn_nf when they carry the bond field");
            delete ctx;
            return rc;
        }
        if ((q.zplq != 0) != (m[0].zplq != 0)) {
            int rc = fail(ctx, SH_E_ARG, "all chains of a batch must have the plaquette field or none");
            delete ctx;
            return rc;
        }
        if (q.zplq) {
            // square (4) or honeycomb (6) plaquettes of distinct sites; G(:, P) and (1 - G)(P, :)
            // of one plaquette stay in LDS
            bool ok = (q.zplq == 4 || q.zplq == 6) && q.lqp >= 1 && q.plq_cord && q.plq_b && q.plq_bi &&
                      q.plq_delta && q.plq_pratio && q.plq_parg && q.plq_gaml &&
                      (size_t)2 * q.zplq * q.ndim * sizeof(cplx) <= 128 * 1024;
            for (int p = 0; ok && p < q.lqp; p++)
                for (int a = 0; a < q.zplq && ok; a++) {
                    const int sa = q.plq_cord[p * q.zplq + a];
                    ok = sa >= 0 && sa < q.ndim;
                    for (int b = 0; b < a && ok; b++) ok = sa != q.plq_cord[p * q.zplq + b];
                }
            if (!ok) {
                int rc = fail(ctx, SH_E_ARG, "plaquette field: zplq 4 or 6, plaquettes of distinct sites in 0..ndim-1, "
                                             "every table given, 2 zplq ndim complex within 128 KiB");
                delete ctx;
                return rc;
            }
        }
        for (int b = 0; b < q.nn_nf * q.nn_lf * 2; b++)
            if (q.bonds[b] < 0 || q.bonds[b] >= q.ndim) { delete ctx; return SH_E_ARG; }
        for (int i = 0; i <= q.ltrot; i++)
            if (q.iwrap_nt[i] < 0 || q.iwrap_nt[i] > q.nst) { delete ctx; return SH_E_ARG; }
        for (int i = 0; i < 2 * q.nst; i++)
            if (q.wrap_step[i] < 1 || q.wrap_step[i] > q.ltrot) { delete ctx; return SH_E_ARG; }
        ctx->nmax = std::max(ctx->nmax, q.ndim);
    }
    // DFT kinetic form: as many vectors per workgroup (<= 16) as fit in 96 KiB of LDS
    if (ctx->kmode == SH_K_FFT) {
        ctx->fft_fcb = 16;
        while (ctx->fft_fcb > 1 && (size_t)ctx->fft_fcb * ctx->fft_ld * sizeof(cplx) > 96 * 1024) ctx->fft_fcb /= 2;
    }
    // arguments are validated before the first HIP call
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return SH_E_HIP; }
    ctx->iwrap_nt.assign(m[0].iwrap_nt, m[0].iwrap_nt + m[0].ltrot + 1);
    ctx->wrap_step.assign(m[0].wrap_step, m[0].wrap_step + 2 * m[0].nst);
    if (hipStreamCreate(&ctx->st) != hipSuccess) { delete ctx; return SH_E_HIP; }
    // kernels whose dynamic LDS grows with N get exactly what the largest chain needs
    {
        struct { const void *f; size_t bytes; } big[] = {
            {(const void *)k_cb_rows, lds_rows(ctx)}, {(const void *)k_cb_cols, lds_cols(ctx)},
            {(const void *)k_fft_rows, lds_fft(ctx)}, {(const void *)k_fft_cols, lds_fft(ctx)},
            // the cap sub_stage works to: the attribute is per function, shared with the
            // projective library's launches of the same kernel
            {(const void *)k_update_submatrix, SUB_LDS_CAP},
            {(const void *)k_vupdate_submatrix, (size_t)(SH_KMAX * (SH_KMAX + 1) + 8 * SH_KMAX) * sizeof(cplx)},
            {(const void *)k_update_delay_reg<4>, (size_t)(1 + 2 * DSB) * ctx->nmax * sizeof(cplx) + ctx->nmax * 20 + 16 * sizeof(cplx)},
            {(const void *)k_update_delay_reg<8>, (size_t)(1 + 2 * DSB) * ctx->nmax * sizeof(cplx) + ctx->nmax * 20 + 32 * sizeof(cplx)},
            {(const void *)k_update_delay_reg<16>, (size_t)(1 + 2 * DSB) * ctx->nmax * sizeof(cplx) + ctx->nmax * 20 + 64 * sizeof(cplx)},
            {(const void *)k_update_delay_reg<32>, (size_t)(1 + 2 * DSB) * ctx->nmax * sizeof(cplx) + ctx->nmax * 20 + 128 * sizeof(cplx)},
            {(const void *)k_qr_blk, blocked(ctx) ? lds_blk(ctx) : 0},
            {(const void *)k_formq_blk, blocked(ctx) ? lds_blk(ctx) : 0},
            {(const void *)k_tfactors, blocked(ctx) ? lds_blk(ctx) : 0},
            {(const void *)k_lu_blk, blocked(ctx) ? lds_blk(ctx) : 0},
            {(const void *)k_solve_blk, blocked(ctx) ? lds_blk(ctx) : 0}};
        for (auto &b : big) {
            if (b.bytes <= 64 * 1024 || b.bytes > 150 * 1024) continue;
            hipFuncAttributes fa;
            if (hipFuncGetAttributes(&fa, b.f) != hipSuccess || fa.sharedSizeBytes + b.bytes > 160 * 1024) {
                (void)hipGetLastError();
                int rc = fail(ctx, SH_E_HIP, "static plus dynamic LDS of a kernel above the 160 KiB of a CU");
                sh_destroy(ctx);
                return rc;
            }
            if (hipFuncSetAttribute(b.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b.bytes) != hipSuccess) {
                (void)hipGetLastError();
                int rc = fail(ctx, SH_E_HIP, "dynamic LDS request above the device limit");
                sh_destroy(ctx);
                return rc;
            }
        }
    }
    std::vector<int32_t> seeds(nchains);
    ctx->hc.resize(nchains);
    for (int ci = 0; ci < nchains; ci++) {
        const sh_model &q = m[ci];
        DChain &c = ctx->hc[ci];
        memset(&c, 0, sizeof(c));
        const int n = q.ndim;
        const size_t nn = (size_t)n * n;
        c.n = n; c.lq = q.lq; c.nn_nf = q.nn_nf; c.nn_lf = q.nn_lf; c.lcomp = q.lcomp; c.nflr = q.nflr;
        c.kmode = q.kmode; c.nublock = q.nublock; c.update = q.update; c.lwrapT = q.lwrapT; c.lwrapu = q.lwrapu;
        c.phase_from_conf = q.phase_from_conf; c.ltrot = q.ltrot; c.nst = q.nst;
        c.kin_weight = q.kin_weight != 0.0 ? q.kin_weight : 1.0;
        if (q.vtab) {
            c.lwrapv = 1;
            for (int k = 0; k < 8; k++) c.vtb[k] = make_double2(q.vtab[2 * k], q.vtab[2 * k + 1]);
            for (int k = 0; k < 4; k++) c.vdl[k] = make_double2(q.vdelta[2 * k], q.vdelta[2 * k + 1]);
            c.confv = dalloc<int32_t>(ctx, (size_t)n * q.nn_nf * q.ltrot);
            c.unifv = dalloc<double>(ctx, (size_t)2 * q.nn_nf * q.nn_lf * q.ltrot);
            c.vacc = dalloc<double>(ctx, 2);
            if (!c.confv || !c.unifv || !c.vacc) {
                int rc = fail(ctx, SH_E_HIP, "device allocation failed");
                sh_destroy(ctx);
                return rc;
            }
            std::vector<int32_t> ones((size_t)n * q.nn_nf * q.ltrot, 1);
            hipMemcpy(c.confv, ones.data(), ones.size() * sizeof(int32_t), hipMemcpyHostToDevice);
            hipMemset(c.vacc, 0, 2 * sizeof(double));
        }
        if (q.zplq) {
            const int z = q.zplq;
            c.lwrapp = 1; c.zp = z; c.lqp = q.lqp;
            c.confp = dalloc<int32_t>(ctx, (size_t)q.lqp * q.ltrot);
            c.unifp = dalloc<double>(ctx, (size_t)2 * q.lqp * q.ltrot);
            c.pacc = dalloc<double>(ctx, 2);
            c.plqc = dupload<int32_t>(ctx, q.plq_cord, (size_t)q.lqp * z);
            c.pb = dupload<cplx>(ctx, (const cplx *)q.plq_b, (size_t)4 * z * z);
            c.pbi = dupload<cplx>(ctx, (const cplx *)q.plq_bi, (size_t)4 * z * z);
            c.pdl = dupload<cplx>(ctx, (const cplx *)q.plq_delta, (size_t)12 * z * z);
            if (!c.confp || !c.unifp || !c.pacc || !c.plqc || !c.pb || !c.pbi || !c.pdl) {
                int rc = fail(ctx, SH_E_HIP, "device allocation failed");
                sh_destroy(ctx);
                return rc;
            }
            for (int k = 0; k < 12; k++) c.ppr[k] = make_double2(q.plq_pratio[2 * k], q.plq_pratio[2 * k + 1]);
            for (int k = 0; k < 4; k++) { c.pgaml[k] = q.plq_gaml[k]; c.pparg[k] = q.plq_parg[k]; }
            std::vector<int32_t> ones((size_t)q.lqp * q.ltrot, 1);
            hipMemcpy(c.confp, ones.data(), ones.size() * sizeof(int32_t), hipMemcpyHostToDevice);
            hipMemset(c.pacc, 0, 2 * sizeof(double));
        }
        c.G = dalloc<cplx>(ctx, nn); c.Gs = dalloc<cplx>(ctx, nn);
        c.W1 = dalloc<cplx>(ctx, nn); c.W2 = dalloc<cplx>(ctx, nn); c.W3 = dalloc<cplx>(ctx, nn);
        c.W4 = dalloc<cplx>(ctx, nn); c.Us = dalloc<cplx>(ctx, nn); c.Vs = dalloc<cplx>(ctx, nn);
        c.U = dalloc<cplx>(ctx, nn * (q.nst + 1)); c.V = dalloc<cplx>(ctx, nn * (q.nst + 1));
        c.D = dalloc<double>(ctx, (size_t)n * (q.nst + 1)); c.Ds = dalloc<double>(ctx, n);
        c.ldq = dalloc<cplx>(ctx, q.nst + 1);
        // room for every depth the autotuner may pick
        const int kb = std::max(q.nublock, SH_KMAX);
        c.avec = dalloc<cplx>(ctx, (size_t)n * kb); c.bvec = dalloc<cplx>(ctx, (size_t)n * kb);
        c.tau = dalloc<cplx>(ctx, n); c.jpvt = dalloc<int32_t>(ctx, n); c.piv = dalloc<int32_t>(ctx, n);
        c.lperm = dalloc<int32_t>(ctx, n);
        c.wk = dalloc<cplx>(ctx, (size_t)PB * n);
        c.tblk = dalloc<cplx>(ctx, (size_t)(n / PB + 1) * PB * PB);
        c.rwork = dalloc<double>(ctx, 4 * (size_t)n);
        c.conf = dalloc<int32_t>(ctx, (size_t)n * q.ltrot);
        c.unif = dalloc<double>(ctx, (size_t)2 * n * q.ltrot);
        c.rng = dalloc<uint32_t>(ctx, 625);
        c.bonds = dupload<int32_t>(ctx, q.bonds, (size_t)2 * q.nn_nf * q.nn_lf);
        size_t kl = (q.kmode == SH_K_CHECKERBOARD) ? (size_t)4 * q.nn_nf * q.nn_lf : q.kmode == SH_K_FFT ? (size_t)n : nn;
        if (q.kmode == SH_K_FFT) { c.fd0 = q.fft_dims[0]; c.fd1 = q.fft_dims[1]; c.fd2 = q.fft_dims[2]; }
        c.kt = dupload<cplx>(ctx, (const cplx *)q.urt, kl);
        c.ktm1 = dupload<cplx>(ctx, (const cplx *)q.urtm1, kl);
        c.stag = dupload<double>(ctx, q.stag, n);
        c.sc = dalloc<double>(ctx, SC_SIZE);
        if (!c.G || !c.Gs || !c.W1 || !c.W2 || !c.W3 || !c.W4 || !c.Us || !c.Vs || !c.U || !c.V || !c.D || !c.Ds ||
            !c.ldq || !c.avec || !c.bvec || !c.tau || !c.jpvt || !c.piv || !c.rwork || !c.conf || !c.unif || !c.wk || !c.tblk ||
            !c.rng || !c.bonds || !c.kt || !c.ktm1 || !c.stag || !c.sc) {
            int rc = fail(ctx, SH_E_HIP, "device allocation failed");
            sh_destroy(ctx);
            return rc;
        }
        for (int l = 0; l < q.lcomp; l++) {
            c.bmat[l] = make_double2(q.bmat[2 * l], q.bmat[2 * l + 1]);
            c.bmat_inv[l] = make_double2(q.bmat_inv[2 * l], q.bmat_inv[2 * l + 1]);
            c.gaml[l] = q.gaml[l];
            c.parg[l] = q.phase_arg[l];
        }
        for (int k = 0; k < (q.lcomp - 1) * q.lcomp; k++) {
            c.delta_bmat[k] = make_double2(q.delta_bmat[2 * k], q.delta_bmat[2 * k + 1]);
            c.phase_ratio[k] = make_double2(q.phase_ratio[2 * k], q.phase_ratio[2 * k + 1]);
        }
        // every conf entry starts at 1 (valid index) until init_conf / set_conf
        std::vector<int32_t> ones((size_t)n * q.ltrot, 1);
        hipMemcpy(c.conf, ones.data(), ones.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        double sc0[SC_SIZE] = {0};
        sc0[SC_PHASE] = 1.0;
        hipMemcpy(c.sc, sc0, sizeof(sc0), hipMemcpyHostToDevice);
        seeds[ci] = q.seed;
    }
    ctx->lwrapv = ctx->hc[0].lwrapv;
    ctx->lwrapp = ctx->hc[0].lwrapp;
    if (ctx->lwrapp) {
        // the plaquette kernels' LDS: whole lines of a chain matrix plus the tables, and one
        // plaquette's rows and columns of G
        const size_t a = lds_plq_prop(ctx), b = lds_plq_slice(ctx);
        if ((a > 64 * 1024 && hipFuncSetAttribute((const void *)k_plq_prop, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)a) != hipSuccess) ||
            (b > 64 * 1024 && hipFuncSetAttribute((const void *)k_plq_slice, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)b) != hipSuccess)) {
            (void)hipGetLastError();
            int rc = fail(ctx, SH_E_HIP, "dynamic LDS request above the device limit");
            sh_destroy(ctx);
            return rc;
        }
    }
    ctx->dc = dupload<DChain>(ctx, ctx->hc.data(), nchains);
    ctx->d_seeds = dupload<int32_t>(ctx, seeds.data(), nchains);
    ctx->seeds = seeds;
    if (!ctx->dc || !ctx->d_seeds) { sh_destroy(ctx); return SH_E_HIP; }
    hipLaunchKernelGGL(k_sfmt_seed, dim3((nchains + 63) / 64), dim3(64), 0, ctx->st, ctx->dc, nchains, ctx->d_seeds);
    if (hipStreamSynchronize(ctx->st) != hipSuccess || hipGetLastError() != hipSuccess) {
        sh_destroy(ctx);
        return SH_E_HIP;
    }
    *out = ctx;
    return 0;
}

int sh_destroy(sh_ctx *ctx)
{
    if (!ctx) return 0;
    if (ctx->st) hipStreamSynchronize(ctx->st);
    for (void *p : ctx->allocs) hipFree(p);
    for (void *p : ctx->stage) hipFree(p);
    for (cplx *p : ctx->trace_bufs) if (p) hipFree(p);
    for (auto e : ctx->pool) hipEventDestroy(e);
    for (auto g : ctx->gexec) if (g) hipGraphExecDestroy(g);
    for (void *p : ctx->tune_save) hipFree(p);
    if (ctx->st) hipStreamDestroy(ctx->st);
    delete ctx;
    return 0;
}

const char *sh_last_error(const sh_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sh_init_conf(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    hipLaunchKernelGGL(k_sfmt_draw, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, ctx->nch, 0);
    return check_launch(ctx);
}

int sh_set_conf(sh_ctx *ctx, int chain, const int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    size_t n = (size_t)c.n * c.ltrot;
    for (size_t k = 0; k < n; k++)
        if (conf[k] < 1 || conf[k] > c.lcomp) return fail(ctx, SH_E_ARG, "conf value out of 1..lcomp");
    HIPCHK(hipMemcpyAsync(c.conf, conf, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_conf(sh_ctx *ctx, int chain, int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    HIPCHK(hipMemcpyAsync(conf, c.conf, (size_t)c.n * c.ltrot * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_set_confv(sh_ctx *ctx, int chain, const int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.lwrapv) return fail(ctx, SH_E_STATE, "the chain has no bond field");
    const size_t cnt = (size_t)c.n * c.nn_nf * c.ltrot;
    for (size_t k = 0; k < cnt; k++)
        if (conf[k] < 1 || conf[k] > 2) return fail(ctx, SH_E_ARG, "bond field values must be 1 or 2");
    HIPCHK(hipMemcpy(c.confv, conf, cnt * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

int sh_get_confv(sh_ctx *ctx, int chain, int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.lwrapv) return fail(ctx, SH_E_STATE, "the chain has no bond field");
    HIPCHK(hipMemcpyAsync(conf, c.confv, (size_t)c.n * c.nn_nf * c.ltrot * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_vacc(sh_ctx *ctx, int chain, double *acc)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !acc) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.lwrapv) return fail(ctx, SH_E_STATE, "the chain has no bond field");
    HIPCHK(hipMemcpyAsync(acc, c.vacc, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_set_confp(sh_ctx *ctx, int chain, const int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.lwrapp) return fail(ctx, SH_E_STATE, "the chain has no plaquette field");
    const size_t cnt = (size_t)c.lqp * c.ltrot;
    for (size_t k = 0; k < cnt; k++)
        if (conf[k] < 1 || conf[k] > 4) return fail(ctx, SH_E_ARG, "plaquette field values are 1..4");
    HIPCHK(hipMemcpyAsync(c.confp, conf, cnt * sizeof(int32_t), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_confp(sh_ctx *ctx, int chain, int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.lwrapp) return fail(ctx, SH_E_STATE, "the chain has no plaquette field");
    HIPCHK(hipMemcpyAsync(conf, c.confp, (size_t)c.lqp * c.ltrot * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_pacc(sh_ctx *ctx, int chain, double *acc)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !acc) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.lwrapp) return fail(ctx, SH_E_STATE, "the chain has no plaquette field");
    HIPCHK(hipMemcpyAsync(acc, c.pacc, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_set_dyn(sh_ctx *ctx, int chain, int ntdm, int l1, int l2, int l3, const int32_t *sites)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !sites || ntdm < 0 || ntdm > ctx->ltrot || l1 < 1 || l2 < 1 ||
        l3 < 1)
        return SH_E_ARG;
    ctx->graph_epoch++;
    if (ctx->ntdm >= 0 && ntdm != ctx->ntdm) return fail(ctx, SH_E_ARG, "all chains of a batch share ntdm");
    DChain &c = ctx->hc[chain];
    if (l1 * l2 * l3 != c.lq) return fail(ctx, SH_E_ARG, "l1 * l2 * l3 must be the number of unit cells");
    for (int i = 0; i < c.n; i++)
        if (sites[4 * i] < 1 || sites[4 * i] > l1 || sites[4 * i + 1] < 1 || sites[4 * i + 1] > l2 ||
            sites[4 * i + 2] < 1 || sites[4 * i + 2] > l3 || sites[4 * i + 3] < 0 || sites[4 * i + 3] > 1)
            return fail(ctx, SH_E_ARG, "site cell coordinates out of 1..l1, 1..l2, 1..l3 or orbital out of 0..1");
    if (c.lq > 512) return fail(ctx, SH_E_ARG, "more unit cells than the znn_tau bins in LDS hold (512)");
    const size_t nn = (size_t)c.n * c.n;
    if (!c.ltau) {
        c.gt0 = dalloc<cplx>(ctx, nn); c.g0t = dalloc<cplx>(ctx, nn); c.g00 = dalloc<cplx>(ctx, nn);
        c.gt0s = dalloc<cplx>(ctx, nn); c.g0ts = dalloc<cplx>(ctx, nn);
        c.znn = dalloc<cplx>(ctx, (size_t)4 * c.lq * (ctx->ltrot + 1));
        c.dynerr = dalloc<double>(ctx, 2);
        if (!c.sxy) c.sxy = dalloc<int32_t>(ctx, (size_t)4 * c.n);
        if (!c.gt0 || !c.g0t || !c.g00 || !c.gt0s || !c.g0ts || !c.znn || !c.dynerr || !c.sxy)
            return fail(ctx, SH_E_HIP, "device allocation failed");
    }
    HIPCHK(hipMemcpy((void *)c.sxy, sites, sizeof(int32_t) * 4 * c.n, hipMemcpyHostToDevice));
    c.dl1 = l1;
    c.dl2 = l2;
    c.dl3 = l3;
    c.ntdm = ntdm;
    c.ltau = 1;
    ctx->ntdm = ntdm;
    HIPCHK(hipMemcpy(ctx->dc + chain, &c, sizeof(DChain), hipMemcpyHostToDevice));
    return 0;
}

// LDS bytes of k_eqt for a chain
static size_t eqt_lds(const DChain &c) { return sizeof(cplx) * ((size_t)16 * c.lq + 2 * EQT_T * (EQT_T + 1) + c.n); }

// the correlation bins of one chain (shared by the finite-temperature and projective batches):
// validation, device tables, the LDS size of k_eqt over every chain of the batch that has bins
static int eqt_setup(ctx_base *ctx, DChain &c, const std::vector<DChain> &all, int l1, int l2, int l3,
                     const int32_t *sites, const int32_t *cells, int zplq, const int32_t *plq, double theta,
                     size_t &lds_out)
{
    if (!sites || !cells || l1 < 1 || l2 < 1 || l3 < 1 || zplq < 0 || (zplq > 0 && !plq)) return SH_E_ARG;
    if (l1 * l2 * l3 != c.lq) return fail(ctx, SH_E_ARG, "l1 * l2 * l3 must be the number of unit cells");
    if (eqt_lds(c) > 160 * 1024) return fail(ctx, SH_E_ARG, "the correlation bins need 16 lq + n + 544 <= 10240");
    for (int i = 0; i < c.n; i++)
        if (sites[4 * i] < 1 || sites[4 * i] > l1 || sites[4 * i + 1] < 1 || sites[4 * i + 1] > l2 ||
            sites[4 * i + 2] < 1 || sites[4 * i + 2] > l3 || sites[4 * i + 3] < 0 || sites[4 * i + 3] > 1)
            return fail(ctx, SH_E_ARG, "site cell coordinates out of 1..l1, 1..l2, 1..l3 or orbital out of 0..1");
    for (int k = 0; k < 4 * c.lq; k++)
        if (cells[k] < 0 || cells[k] >= c.n) return fail(ctx, SH_E_ARG, "cell partner site out of range");
    for (int k = 0; k < zplq * c.lq; k++)
        if (plq[k] < 0 || plq[k] >= c.n) return fail(ctx, SH_E_ARG, "plaquette site out of range");
    if (!c.sxy) c.sxy = dalloc<int32_t>(ctx, (size_t)4 * c.n);
    if (!c.eqb) c.eqb = dalloc<cplx>(ctx, (size_t)20 * c.lq + 8);
    if (!c.eqc) c.eqc = dalloc<int32_t>(ctx, (size_t)4 * c.lq);
    if (zplq > 0 && c.zplq != zplq) c.plq = dalloc<int32_t>(ctx, (size_t)zplq * c.lq);
    if (!c.sxy || !c.eqb || !c.eqc || (zplq > 0 && !c.plq)) return fail(ctx, SH_E_HIP, "device allocation failed");
    HIPCHK(hipMemcpy((void *)c.sxy, sites, sizeof(int32_t) * 4 * c.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void *)c.eqc, cells, sizeof(int32_t) * 4 * c.lq, hipMemcpyHostToDevice));
    if (zplq > 0) HIPCHK(hipMemcpy((void *)c.plq, plq, sizeof(int32_t) * zplq * c.lq, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c.eqb, 0, sizeof(cplx) * (20 * (size_t)c.lq + 8)));
    c.dl1 = l1;
    c.dl2 = l2;
    c.dl3 = l3;
    c.zplq = zplq;
    c.theta = theta;
    size_t lds = 0;
    for (const DChain &d : all)
        if (d.eqb || &d == &c) lds = std::max(lds, eqt_lds(d));
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)k_eqt, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return fail(ctx, SH_E_HIP, "k_eqt: LDS request refused");
    lds_out = lds;
    return 0;
}

int sh_set_eqt(sh_ctx *ctx, int chain, int l1, int l2, int l3, const int32_t *sites, const int32_t *cells, int zplq,
               const int32_t *plq, double theta)
{
    if (!ctx || chain < 0 || chain >= ctx->nch) return SH_E_ARG;
    ctx->graph_epoch++;
    DChain &c = ctx->hc[chain];
    size_t lds = 0;
    int rc = eqt_setup(ctx, c, ctx->hc, l1, l2, l3, sites, cells, zplq, plq, theta, lds);
    if (rc) return rc;
    ctx->eqt_lds = lds;
    HIPCHK(hipMemcpy(ctx->dc + chain, &c, sizeof(DChain), hipMemcpyHostToDevice));
    return 0;
}

int sh_get_eqt(sh_ctx *ctx, int chain, double *bins)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !bins) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.eqb) return fail(ctx, SH_E_STATE, "correlation bins not set up (sh_set_eqt)");
    HIPCHK(hipMemcpyAsync(bins, c.eqb, sizeof(cplx) * (20 * (size_t)c.lq + 8), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_dyn_obs(sh_ctx *ctx, int chain, double *znn, double *dynerr)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !znn) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.ltau) return fail(ctx, SH_E_STATE, "time-displaced measurement not set up (sh_set_dyn)");
    HIPCHK(hipMemcpyAsync(znn, c.znn, sizeof(cplx) * 4 * c.lq * (c.ntdm + 1), hipMemcpyDeviceToHost, ctx->st));
    if (dynerr) HIPCHK(hipMemcpyAsync(dynerr, c.dynerr, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_reset_dyn_obs(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    for (auto &c : ctx->hc)
        if (c.ltau) {
            HIPCHK(hipMemsetAsync(c.znn, 0, sizeof(cplx) * 4 * c.lq * (ctx->ltrot + 1), ctx->st));
            HIPCHK(hipMemsetAsync(c.dynerr, 0, 2 * sizeof(double), ctx->st));
        }
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_gtau(sh_ctx *ctx, int chain, double *gt0, double *g0t, double *g00)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !gt0 || !g0t || !g00) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    if (!c.ltau) return fail(ctx, SH_E_STATE, "time-displaced measurement not set up (sh_set_dyn)");
    const size_t nn = (size_t)c.n * c.n * sizeof(cplx);
    HIPCHK(hipMemcpyAsync(gt0, c.gt0, nn, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(g0t, c.g0t, nn, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(g00, c.g00, nn, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_rng_state(sh_ctx *ctx, int chain, uint32_t *state)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !state) return SH_E_ARG;
    HIPCHK(hipMemcpyAsync(state, ctx->hc[chain].rng, 625 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->st));
    uint64_t ps = 0;
    if (!ctx->hprop.empty()) HIPCHK(hipMemcpyAsync(&ps, ctx->hprop[chain].rng, sizeof(ps), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    state[625] = (uint32_t)ps;
    state[626] = (uint32_t)(ps >> 32);
    return 0;
}

int sh_set_rng_state(sh_ctx *ctx, int chain, const uint32_t *state)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !state) return SH_E_ARG;
    if (state[624] > 624) return fail(ctx, SH_E_ARG, "SFMT read index out of 0..624");
    HIPCHK(hipMemcpyAsync(ctx->hc[chain].rng, state, 625 * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->st));
    const uint64_t ps = (uint64_t)state[625] | ((uint64_t)state[626] << 32);
    if (!ctx->hprop.empty()) HIPCHK(hipMemcpyAsync(ctx->hprop[chain].rng, &ps, sizeof(ps), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

static int first_info(sh_ctx *ctx)
{
    // the reference stops on a failed LU in b0; report the first chain's pivot index
    for (int ci = 0; ci < ctx->nch; ci++) {
        double v[2];
        if (hipMemcpy(v, ctx->hc[ci].sc + SC_INFO, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return SH_E_HIP;
        if (v[0] != 0.0) return (int)v[0];
    }
    return 0;
}

int sh_sweep_start_0b(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    set_identity(ctx, 0);
    for (int n = 1; n <= ctx->nst; n++) stablize_0b(ctx, n);
    green_bb(ctx, ctx->nst);
    wrap_replace(ctx, 0);
    int rc = check_launch(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    int info = first_info(ctx);
    if (info > 0) return fail(ctx, info, "LU breakdown in sweep_start");
    ctx->started = true;
    return 0;
}

int sh_sweep_b0(sh_ctx *ctx, int lupdate, int lmeasure)
{
    if (!ctx) return SH_E_ARG;
    if (!ctx->started) return fail(ctx, SH_E_STATE, "sweep_b0 before sweep_start_0b");
    const bool upd = lupdate && ctx->hc[0].lwrapu, updv = lupdate && ctx->lwrapv, updp = lupdate && ctx->lwrapp;
    const bool split = ctx->lwrapv || ctx->lwrapp;
    auto body = [&]() {
    set_identity(ctx, ctx->nst);
    if (upd || updv || updp) hipLaunchKernelGGL(k_sfmt_draw, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, ctx->nch, 1);
    if (upd) prop_pass_start(ctx);
    int slice = 0;
    for (int nt = ctx->ltrot; nt >= 1; nt--, slice++) {
        if (lmeasure) measure(ctx);
        {
            Tic t(ctx, T_WRAP);
            // K-half: G B^{-1}: right_forward (urt), left_backward (urtm1), families 1..nn_nf
            right_pass(ctx, 0, 1, 0, nt);
            left_pass(ctx, M_G, 0, 1, 1, 0, nt, nt, 1);
        }
        if (upd && ctx->prop_local) update_slice(ctx, nt, slice);
        if (upd) block_move(ctx, nt);
        if (!split) {
            Tic t(ctx, T_WRAP);
            // V (right_forward bmat, left_backward bmat_inv) then the K-half nn_nf..1
            right_pass(ctx, 0, 8 | 4, 0, nt);
            left_pass(ctx, M_G, 0, 1, 8 | 4, 1, nt, nt, 1);
        } else {
            // V, the bond field by families nn_nf..1 (nflag 2, update, nflag 1:
            // src/dqmc_ft_core/dqmc_ft_sweep_b0.f90:70-89), the plaquettes lq..1 (update, then
            // right_forward / left_backward: :91-100), then the K-half nn_nf..1
            Tic t(ctx, T_WRAP);
            if (ctx->hc[0].lwrapu) {
                right_pass(ctx, 0, 8, 0, nt);
                left_pass(ctx, M_G, 0, 1, 8, 1, nt, nt, 1);
            }
            const int nfam = ctx->lwrapv ? ctx->hc[0].nn_nf : 0;
            for (int p = 0; p < nfam; p++) {
                const int nf = nfam - 1 - p;
                vprop(ctx, M_G, nf, 1, 1, nt);
                vprop(ctx, M_G, nf, 1, 0, nt);
                if (updv) vupdate(ctx, nt, nf, slice, p);
                vprop(ctx, M_G, nf, 2, 1, nt);
                vprop(ctx, M_G, nf, 2, 0, nt);
            }
            if (ctx->lwrapp) plq_slice(ctx, nt, 1, slice, updp);
            right_pass(ctx, 0, 4, 0, nt);
            left_pass(ctx, M_G, 0, 1, 4, 1, nt, nt, 1);
        }
        if ((ctx->iwrap_nt[nt - 1] > 0 || ctx->nwrap == 1) || nt == 1) {
            int n = ctx->iwrap_nt[nt - 1];
            {
                Tic t(ctx, T_OTHER);
                hipLaunchKernelGGL(k_save_point, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, n);
            }
            stablize_b0(ctx, n + 1);
            green(ctx, 1, 0, 0, n);
            wrap_replace(ctx, 1);
        }
    }
    };
    int rc = run_pass(ctx, 0 + 2 * (lupdate != 0) + (lmeasure != 0), body);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    int info = first_info(ctx);
    if (info > 0) return fail(ctx, info, "LU breakdown in green_equaltime (sweep_b0)");
    return 0;
}

// dqmc_ft_sweep_0b(lupdate = .false., lmeasure_equaltime, lmeasure_dyn = .true.) with ltau
// (src/dqmc_ft_core/dqmc_ft_sweep_0b.f90:36-49, :160-204, dqmc_ft_core/dyn.f90): no updates;
// G(t,0) <- B(t) G(t,0) and G(0,t) <- G(0,t) B(t)^-1 slice by slice, replaced by green_tau at
// every stabilization point after the -DDYNERROR comparison; bins 0 .. ntdm measured
static int sweep_0b_dyn(sh_ctx *ctx, int lmeasure)
{
    set_identity(ctx, 0);
    // the bins of up to 512 cells and two diagonals (<= 48 KiB)
    const size_t dyn_lds = (size_t)(8 * 512 + 2 * ctx->nmax) * sizeof(double);
    auto dyn_measure = [&](int b) {
        Tic t(ctx, T_MEASURE);
        hipLaunchKernelGGL(k_dyn_measure, dim3(ctx->nch), dim3(256), dyn_lds, ctx->st, ctx->dc, b);
    };
    {
        Tic t(ctx, T_OTHER);
        hipLaunchKernelGGL(k_dyn_elem, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, 0, (int)M_G, 0);
    }
    dyn_measure(0);
    const bool split = ctx->lwrapv || ctx->lwrapp;
    for (int nt = 1; nt <= ctx->ltrot; nt++) {
        if (!split) {
            Tic t(ctx, T_WRAP);
            left_pass(ctx, M_G, 0, 0, 1 | 2, 0, nt, nt, 1);
            right_pass(ctx, 1, 1 | 2, 1, nt);
            left_pass(ctx, M_G, 0, 0, 4, 0, nt, nt, 1);
            right_pass(ctx, 1, 4, 1, nt);
        } else {
            // K-half, the plaquettes 1..lq, the bond field by families 1..nn_nf (nflag 2 then 1),
            // V, K-half
            Tic t(ctx, T_WRAP);
            const int nfam = ctx->lwrapv ? ctx->hc[0].nn_nf : 0;
            left_pass(ctx, M_G, 0, 0, 1, 0, nt, nt, 1);
            right_pass(ctx, 1, 1, 1, nt);
            if (ctx->lwrapp) plq_slice(ctx, nt, 0, nt - 1, false);
            for (int nf = 0; nf < nfam; nf++) {
                vprop(ctx, M_G, nf, 0, 0, nt);
                vprop(ctx, M_G, nf, 0, 1, nt);
                vprop(ctx, M_G, nf, 1, 0, nt);
                vprop(ctx, M_G, nf, 1, 1, nt);
            }
            left_pass(ctx, M_G, 0, 0, 2 | 4, 0, nt, nt, 1);
            right_pass(ctx, 1, 2 | 4, 1, nt);
        }
        if (lmeasure) measure(ctx);
        if (!split) {
            // G(t,0) <- B(t) G(t,0), G(0,t) <- G(0,t) B(t)^-1 (Bmat_left_forward / _right_backward)
            Tic t(ctx, T_WRAP);
            left_pass(ctx, M_GT0, 0, 0, 1 | 2 | 4, 0, nt, nt, 1);
            right_pass(ctx, 1, 1 | 2 | 4, 1, nt, M_G0T);
        } else {
            Tic t(ctx, T_WRAP);
            const int nfam = ctx->lwrapv ? ctx->hc[0].nn_nf : 0;
            left_pass(ctx, M_GT0, 0, 0, 1, 0, nt, nt, 1);
            right_pass(ctx, 1, 1, 1, nt, M_G0T);
            if (ctx->lwrapp) {
                pprop(ctx, M_GT0, 0, 0, 0, 0, nt);
                pprop(ctx, M_G0T, 0, 1, 1, 0, nt);
            }
            for (int nf = 0; nf < nfam; nf++) {
                vprop(ctx, M_GT0, nf, 0, 0, nt);
                vprop(ctx, M_GT0, nf, 1, 0, nt);
                vprop(ctx, M_G0T, nf, 0, 1, nt);
                vprop(ctx, M_G0T, nf, 1, 1, nt);
            }
            left_pass(ctx, M_GT0, 0, 0, 2 | 4, 0, nt, nt, 1);
            right_pass(ctx, 1, 2 | 4, 1, nt, M_G0T);
        }
        if (ctx->iwrap_nt[nt] > 0) {
            int n = ctx->iwrap_nt[nt];
            {
                Tic t(ctx, T_OTHER);
                hipLaunchKernelGGL(k_save_point, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, n);
            }
            stablize_0b(ctx, n);
            green_tau(ctx, n);
            {
                Tic t(ctx, T_OTHER);
                hipLaunchKernelGGL(k_dyn_err, dim3(ctx->nch), dim3(256), 0, ctx->st, ctx->dc);
                hipLaunchKernelGGL(k_dyn_elem, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, 8, (int)M_GT0, 0);
                hipLaunchKernelGGL(k_dyn_elem, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, 8, (int)M_G0T, 0);
            }
            wrap_replace(ctx, 1);
        }
        if (nt <= ctx->ntdm) dyn_measure(nt);
    }
    int rc = check_launch(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_sweep_0b(sh_ctx *ctx, int lupdate, int lmeasure, int lmeasure_dyn)
{
    if (!ctx) return SH_E_ARG;
    if (!ctx->started) return fail(ctx, SH_E_STATE, "sweep_0b before sweep_start_0b");
    if (lmeasure_dyn) {
        if (lupdate) return fail(ctx, SH_E_ARG, "time-displaced measurement runs without updates (lupdate = 0)");
        for (const auto &c : ctx->hc)
            if (!c.ltau) return fail(ctx, SH_E_STATE, "time-displaced measurement not set up on every chain (sh_set_dyn)");
        return sweep_0b_dyn(ctx, lmeasure);
    }
    const bool upd = lupdate && ctx->hc[0].lwrapu, updv = lupdate && ctx->lwrapv, updp = lupdate && ctx->lwrapp;
    const bool split = ctx->lwrapv || ctx->lwrapp;
    auto body = [&]() {
    set_identity(ctx, 0);
    if (upd || updv || updp) hipLaunchKernelGGL(k_sfmt_draw, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, ctx->nch, 2);
    if (upd) prop_pass_start(ctx);
    for (int nt = 1; nt <= ctx->ltrot; nt++) {
        if (!split) {
            Tic t(ctx, T_WRAP);
            // K-half B G B^{-1}: left_forward (urt), right_backward (urtm1); then V rows/cols
            left_pass(ctx, M_G, 0, 0, 1 | 2, 0, nt, nt, 1);
            right_pass(ctx, 1, 1 | 2, 1, nt);
        } else {
            // K-half, the plaquettes 1..lq (left_forward, right_backward, update:
            // src/dqmc_ft_core/dqmc_ft_sweep_0b.f90:82-93), the bond field by families 1..nn_nf
            // (nflag 2, update, nflag 1: :95-106), then V
            Tic t(ctx, T_WRAP);
            left_pass(ctx, M_G, 0, 0, 1, 0, nt, nt, 1);
            right_pass(ctx, 1, 1, 1, nt);
            if (ctx->lwrapp) plq_slice(ctx, nt, 0, nt - 1, updp);
            const int nfam = ctx->lwrapv ? ctx->hc[0].nn_nf : 0;
            for (int nf = 0; nf < nfam; nf++) {
                vprop(ctx, M_G, nf, 0, 0, nt);
                vprop(ctx, M_G, nf, 0, 1, nt);
                if (updv) vupdate(ctx, nt, nf, nt - 1, nf);
                vprop(ctx, M_G, nf, 1, 0, nt);
                vprop(ctx, M_G, nf, 1, 1, nt);
            }
            if (ctx->hc[0].lwrapu) {
                left_pass(ctx, M_G, 0, 0, 2, 0, nt, nt, 1);
                right_pass(ctx, 1, 2, 1, nt);
            }
        }
        if (upd && ctx->prop_local) update_slice(ctx, nt, nt - 1);
        if (upd) block_move(ctx, nt);
        {
            Tic t(ctx, T_WRAP);
            left_pass(ctx, M_G, 0, 0, 4, 0, nt, nt, 1);
            right_pass(ctx, 1, 4, 1, nt);
        }
        if (lmeasure) measure(ctx);
        if (ctx->iwrap_nt[nt] > 0) {
            int n = ctx->iwrap_nt[nt];
            {
                Tic t(ctx, T_OTHER);
                hipLaunchKernelGGL(k_save_point, grid_elem(ctx), dim3(256), 0, ctx->st, ctx->dc, n);
            }
            stablize_0b(ctx, n);
            green(ctx, 0, n, 1, 0);
            wrap_replace(ctx, 1);
        }
    }
    };
    int rc = run_pass(ctx, 4 + 2 * (lupdate != 0) + (lmeasure != 0), body);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_push_stage(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    if (ctx->stage.empty()) {
        for (auto &c : ctx->hc) {
            size_t nn = (size_t)c.n * c.n;
            size_t bytes = (2 * nn * (c.nst + 1) + nn) * sizeof(cplx) + (size_t)c.n * (c.nst + 1) * sizeof(double) +
                           (c.nst + 1) * sizeof(cplx) + SC_SIZE * sizeof(double);
            void *p = nullptr;
            HIPCHK(hipMalloc(&p, bytes));
            ctx->stage.push_back(p);
        }
    }
    for (int ci = 0; ci < ctx->nch; ci++) {
        DChain &c = ctx->hc[ci];
        size_t nn = (size_t)c.n * c.n;
        char *p = (char *)ctx->stage[ci];
        HIPCHK(hipMemcpyAsync(p, c.U, nn * (c.nst + 1) * sizeof(cplx), hipMemcpyDeviceToDevice, ctx->st)); p += nn * (c.nst + 1) * sizeof(cplx);
        HIPCHK(hipMemcpyAsync(p, c.V, nn * (c.nst + 1) * sizeof(cplx), hipMemcpyDeviceToDevice, ctx->st)); p += nn * (c.nst + 1) * sizeof(cplx);
        HIPCHK(hipMemcpyAsync(p, c.G, nn * sizeof(cplx), hipMemcpyDeviceToDevice, ctx->st)); p += nn * sizeof(cplx);
        HIPCHK(hipMemcpyAsync(p, c.D, (size_t)c.n * (c.nst + 1) * sizeof(double), hipMemcpyDeviceToDevice, ctx->st)); p += (size_t)c.n * (c.nst + 1) * sizeof(double);
        HIPCHK(hipMemcpyAsync(p, c.ldq, (c.nst + 1) * sizeof(cplx), hipMemcpyDeviceToDevice, ctx->st)); p += (c.nst + 1) * sizeof(cplx);
        HIPCHK(hipMemcpyAsync(p, c.sc, SC_SIZE * sizeof(double), hipMemcpyDeviceToDevice, ctx->st));
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_pop_stage(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    if (ctx->stage.empty()) return fail(ctx, SH_E_STATE, "pop_stage without push_stage");
    for (int ci = 0; ci < ctx->nch; ci++) {
        DChain &c = ctx->hc[ci];
        size_t nn = (size_t)c.n * c.n;
        char *p = (char *)ctx->stage[ci];
        HIPCHK(hipMemcpyAsync(c.U, p, nn * (c.nst + 1) * sizeof(cplx), hipMemcpyDeviceToDevice, ctx->st)); p += nn * (c.nst + 1) * sizeof(cplx);
        HIPCHK(hipMemcpyAsync(c.V, p, nn * (c.nst + 1) * sizeof(cplx), hipMemcpyDeviceToDevice, ctx->st)); p += nn * (c.nst + 1) * sizeof(cplx);
        HIPCHK(hipMemcpyAsync(c.G, p, nn * sizeof(cplx), hipMemcpyDeviceToDevice, ctx->st)); p += nn * sizeof(cplx);
        HIPCHK(hipMemcpyAsync(c.D, p, (size_t)c.n * (c.nst + 1) * sizeof(double), hipMemcpyDeviceToDevice, ctx->st)); p += (size_t)c.n * (c.nst + 1) * sizeof(double);
        HIPCHK(hipMemcpyAsync(c.ldq, p, (c.nst + 1) * sizeof(cplx), hipMemcpyDeviceToDevice, ctx->st));
        // gfc = gf and the UDV chain come back; phase and logweightf are module state that
        // pop_stage does not restore in the reference, so sc stays as is
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_g(sh_ctx *ctx, int chain, double *g)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !g) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    HIPCHK(hipMemcpyAsync(g, c.G, (size_t)c.n * c.n * sizeof(cplx), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_diag(sh_ctx *ctx, int chain, double *d)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !d) return SH_E_ARG;
    double sc[SC_SIZE];
    HIPCHK(hipMemcpyAsync(sc, ctx->hc[chain].sc, sizeof(sc), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    d[0] = sc[SC_MAXWRAP]; d[1] = sc[SC_MAXPH];
    d[2] = sc[SC_PHASE]; d[3] = sc[SC_PHASE + 1];
    d[4] = sc[SC_LOGW]; d[5] = sc[SC_LOGW + 1];
    d[6] = sc[SC_ACC]; d[7] = sc[SC_TRIED]; d[8] = sc[SC_NCHK]; d[9] = sc[SC_SUMLOG];
    if (sc[SC_NONFINITE] != 0.0) return fail(ctx, SH_E_NONFINITE, "non-finite scratch Green's function");
    return 0;
}

int sh_reset_diag(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    for (auto &c : ctx->hc) {
        HIPCHK(hipMemsetAsync(c.sc + SC_MAXWRAP, 0, 6 * sizeof(double), ctx->st));
        HIPCHK(hipMemsetAsync(c.sc + SC_NONFINITE, 0, sizeof(double), ctx->st));
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_get_obs(sh_ctx *ctx, int chain, double *o)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !o) return SH_E_ARG;
    double sc[SC_SIZE];
    HIPCHK(hipMemcpyAsync(sc, ctx->hc[chain].sc, sizeof(sc), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    o[0] = sc[SC_NOBS];
    for (int k = 0; k < 10; k++) o[1 + k] = sc[SC_OBS + k];
    return 0;
}

int sh_get_obs_all(sh_ctx *ctx, double *o, int reset)
{
    if (!ctx || !o) return SH_E_ARG;
    std::vector<double> sc((size_t)SC_SIZE * ctx->nch);
    for (int ci = 0; ci < ctx->nch; ci++)
        HIPCHK(hipMemcpyAsync(sc.data() + (size_t)SC_SIZE * ci, ctx->hc[ci].sc, SC_SIZE * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->st));
    if (reset)
        for (auto &c : ctx->hc) HIPCHK(hipMemsetAsync(c.sc + SC_NOBS, 0, 11 * sizeof(double), ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int ci = 0; ci < ctx->nch; ci++) {
        const double *q = sc.data() + (size_t)SC_SIZE * ci;
        o[11 * ci] = q[SC_NOBS];
        for (int k = 0; k < 10; k++) o[11 * ci + 1 + k] = q[SC_OBS + k];
    }
    return 0;
}

int sh_reset_obs(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    for (auto &c : ctx->hc) {
        HIPCHK(hipMemsetAsync(c.sc + SC_NOBS, 0, 11 * sizeof(double), ctx->st));
        if (c.eqb) HIPCHK(hipMemsetAsync(c.eqb, 0, sizeof(cplx) * (20 * (size_t)c.lq + 8), ctx->st));
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_set_trace(sh_ctx *ctx, int chain, int cap)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || cap < 0) return SH_E_ARG;
    ctx->graph_epoch++;
    if ((int)ctx->trace_bufs.size() < ctx->nch) ctx->trace_bufs.resize(ctx->nch, nullptr);
    DChain &c = ctx->hc[chain];
    if (ctx->trace_bufs[chain]) { HIPCHK(hipFree(ctx->trace_bufs[chain])); ctx->trace_bufs[chain] = nullptr; }
    c.trace = nullptr;
    c.trace_cap = 0;
    if (cap > 0) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, (size_t)cap * c.n * c.n * sizeof(cplx)));
        ctx->trace_bufs[chain] = (cplx *)p;
        c.trace = (cplx *)p;
        c.trace_cap = cap;
    }
    double zero = 0.0;
    HIPCHK(hipMemcpy(c.sc + SC_TRACE_N, &zero, sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->dc + chain, &c, sizeof(DChain), hipMemcpyHostToDevice));
    return 0;
}

int sh_get_trace(sh_ctx *ctx, int chain, double *buf, int *count)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !count) return SH_E_ARG;
    const DChain &c = ctx->hc[chain];
    double tn = 0.0;
    HIPCHK(hipMemcpy(&tn, c.sc + SC_TRACE_N, sizeof(double), hipMemcpyDeviceToHost));
    *count = (int)tn;
    if (buf && c.trace && *count > 0)
        HIPCHK(hipMemcpy(buf, c.trace, (size_t)(*count) * c.n * c.n * sizeof(cplx), hipMemcpyDeviceToHost));
    return 0;
}

int sh_timers(sh_ctx *ctx, int enable, double *t)
{
    if (!ctx) return SH_E_ARG;
    ctx->graph_epoch++;
    timers_collect(ctx);
    if (t) {
        for (int k = 0; k < T_NCLS; k++) { t[k] = ctx->tms[k]; t[T_NCLS + k] = ctx->tcnt[k]; }
    }
    for (int k = 0; k < T_NCLS; k++) { ctx->tms[k] = 0; ctx->tcnt[k] = 0; }
    ctx->timing = enable != 0;
    return 0;
}

int sh_synchronize(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return check_launch(ctx);
}

int sh_set_graphs(sh_ctx *ctx, int enable)
{
    if (!ctx) return SH_E_ARG;
    ctx->use_graphs = enable != 0;
    ctx->graph_epoch++;
    return 0;
}

// ---- autotuner -------------------------------------------------------------------------------
// The update of one slice (slice ltrot, the first of a b0 pass) is timed on the chains' current
// G at every candidate depth, with the state every trial changes (G, the slice's HS field, the
// scalar slots) restored before each trial and at the end, and the slice's uniforms replaced
// by a fixed low-discrepancy sequence (the next updating pass draws its own).  Then the DFT
// kinetic kernels at every vectors-per-workgroup that fits, on the scratch matrix W1.
}  // extern "C"

__global__ void k_tune_unif(const DChain *chains)
{
    const DChain &c = chains[blockIdx.x];
    for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
        const double a = 0.5 + 0.6180339887498949 * i, b = 0.25 + 0.7548776662466927 * i;
        c.unif[2 * i] = a - floor(a);
        c.unif[2 * i + 1] = b - floor(b);
    }
}

static int tune_state(sh_ctx *ctx, bool save)
{
    const int nt = ctx->ltrot;
    for (int ci = 0; ci < ctx->nch; ci++) {
        const DChain &c = ctx->hc[ci];
        const size_t gb = (size_t)c.n * c.n * sizeof(cplx), cb = (size_t)c.n * sizeof(int32_t);
        char *p = (char *)ctx->tune_save[ci];
        int32_t *cf = c.conf + (size_t)(nt - 1) * c.n;
        if (save) {
            HIPCHK(hipMemcpyAsync(p, c.G, gb, hipMemcpyDeviceToDevice, ctx->st));
            HIPCHK(hipMemcpyAsync(p + gb, cf, cb, hipMemcpyDeviceToDevice, ctx->st));
            HIPCHK(hipMemcpyAsync(p + gb + cb, c.sc, SC_SIZE * sizeof(double), hipMemcpyDeviceToDevice, ctx->st));
        } else {
            HIPCHK(hipMemcpyAsync(c.G, p, gb, hipMemcpyDeviceToDevice, ctx->st));
            HIPCHK(hipMemcpyAsync(cf, p + gb, cb, hipMemcpyDeviceToDevice, ctx->st));
            HIPCHK(hipMemcpyAsync(c.sc, p + gb + cb, SC_SIZE * sizeof(double), hipMemcpyDeviceToDevice, ctx->st));
        }
    }
    return 0;
}

static int set_nublock(sh_ctx *ctx, int k)
{
    for (auto &c : ctx->hc) c.nublock = k;
    HIPCHK(hipMemcpy(ctx->dc, ctx->hc.data(), sizeof(DChain) * ctx->nch, hipMemcpyHostToDevice));
    return 0;
}

// best of `reps` timed runs of f (after one untimed run), in ms; before() runs outside the timing
template <class Before, class F>
static double tune_time(sh_ctx *ctx, int reps, Before before, F f)
{
    hipEvent_t a = ev_get(ctx), b = ev_get(ctx);
    double best = 1e30;
    for (int r = 0; r <= reps; r++) {
        before();
        hipEventRecord(a, ctx->st);
        f();
        hipEventRecord(b, ctx->st);
        hipEventSynchronize(b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, a, b);
        if (r > 0) best = std::min(best, (double)ms);
    }
    ctx->pool_used -= 2;
    return best;
}

extern "C" {

int sh_autotune(sh_ctx *ctx, int *tuned)
{
    if (!ctx || !tuned) return SH_E_ARG;
    if (!ctx->started) return fail(ctx, SH_E_STATE, "sh_autotune before sweep_start_0b");
    if (ctx->timing) return fail(ctx, SH_E_STATE, "sh_autotune with the timers on");
    HIPCHK(hipStreamSynchronize(ctx->st));
    const int k0 = ctx->hc[0].nublock;
    int best_k = k0;
    if (ctx->hc[0].lwrapu) {
        if (ctx->tune_save.empty()) {
            for (auto &c : ctx->hc) {
                void *p = nullptr;
                const size_t bytes = (size_t)c.n * c.n * sizeof(cplx) + (size_t)c.n * sizeof(int32_t) + SC_SIZE * sizeof(double);
                HIPCHK(hipMalloc(&p, bytes));
                ctx->tune_save.push_back(p);
            }
        }
        int rc = tune_state(ctx, true);
        if (rc) return rc;
        hipLaunchKernelGGL(k_tune_unif, dim3(ctx->nch), dim3(256), 0, ctx->st, ctx->dc);
        std::vector<int> cand;
        if (ctx->update == SH_U_DELAY) {
            if (k0 > 1) {
                // register-resident depths where they fit, the memory-resident kernel above
                for (int k : {4, 8, 16, 32}) cand.push_back(k);
                int thr = 0;
                for (auto &c : ctx->hc) c.nublock = 32;
                const bool reg32 = delay_reg_depth(ctx, &thr) != 0;
                for (auto &c : ctx->hc) c.nublock = k0;
                if (!reg32) cand.push_back(64);
            }
        } else {
            for (int k : {8, 16, 24, 32, 48, 64}) cand.push_back(k);
        }
        double best = 1e30;
        for (int k : cand) {
            if (k > ctx->nmax) continue;
            if ((rc = set_nublock(ctx, k))) return rc;
            const double ms = tune_time(ctx, 3, [&]() { tune_state(ctx, false); }, [&]() { update_slice(ctx, ctx->ltrot, 0); });
            if (hipGetLastError() != hipSuccess) return fail(ctx, SH_E_HIP, "autotune launch failed");
            if (ms < best) { best = ms; best_k = k; }
        }
        if ((rc = set_nublock(ctx, best_k))) return rc;
        if ((rc = tune_state(ctx, false))) return rc;
    }
    int best_f = 0;
    if (ctx->kmode == SH_K_FFT) {
        double best = 1e30;
        const int f0 = ctx->fft_fcb;
        best_f = f0;
        for (int f = 2; f <= 16; f *= 2) {
            if ((size_t)f * ctx->fft_ld * sizeof(cplx) > 96 * 1024) continue;
            ctx->fft_fcb = f;
            const double ms = tune_time(ctx, 3, []() {}, [&]() {
                left_pass(ctx, M_W1, 0, 0, 1 | 2 | 4, 0, ctx->ltrot, ctx->ltrot, 1);
                right_pass(ctx, 1, 1 | 2 | 4, 1, ctx->ltrot, M_W1);
            });
            if (hipGetLastError() != hipSuccess) return fail(ctx, SH_E_HIP, "autotune launch failed");
            if (ms < best) { best = ms; best_f = f; }
        }
        ctx->fft_fcb = best_f;
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->graph_epoch++;
    tuned[0] = best_k;
    tuned[1] = best_f;
    return 0;
}

}  // extern "C"

// ---- test hooks (include/sunhip_testing.h) --------------------------------------------------
#include "../../include/sunhip_testing.h"

static void *dbg_ptr(sh_ctx *ctx, int chain, int sel, int pt, size_t *bytes)
{
    const DChain &c = ctx->hc[chain];
    size_t n = c.n, nn = n * n;
    *bytes = nn * sizeof(cplx);
    switch (sel) {
    case 0: return c.G;
    case 1: return c.Gs;
    case 2: return c.W1;
    case 3: return c.W2;
    case 4: return c.W3;
    case 5: return c.W4;
    case 6: return c.Us;
    case 7: return c.Vs;
    case 8: return c.U + nn * pt;
    case 9: return c.V + nn * pt;
    case 100: *bytes = n * sizeof(double); return c.D + n * pt;
    case 102: *bytes = n * sizeof(cplx); return c.tau;
    case 103: *bytes = sizeof(cplx); return c.ldq + pt;
    case 104: *bytes = SC_SIZE * sizeof(double); return c.sc;
    }
    return nullptr;
}

extern "C" int sh_dbg_put(sh_ctx *ctx, int chain, int sel, int pt, const double *src)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !src) return SH_E_ARG;
    if (sel == 8 || sel == 9 || sel == 100 || sel == 103)
        if (pt < 0 || pt > ctx->hc[chain].nst) return SH_E_ARG;
    size_t bytes = 0;
    void *p = dbg_ptr(ctx, chain, sel, pt, &bytes);
    if (!p) return SH_E_ARG;
    HIPCHK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int sh_dbg_get(sh_ctx *ctx, int chain, int sel, int pt, double *dst)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !dst) return SH_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->st));
    const DChain &c = ctx->hc[chain];
    if (sel == 101 || sel == 105) {
        std::vector<int32_t> v(c.n);
        HIPCHK(hipMemcpy(v.data(), sel == 101 ? c.jpvt : c.piv, c.n * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int i = 0; i < c.n; i++) dst[i] = v[i];
        return 0;
    }
    if (sel == 8 || sel == 9 || sel == 100 || sel == 103)
        if (pt < 0 || pt > c.nst) return SH_E_ARG;
    size_t bytes = 0;
    void *p = dbg_ptr(ctx, chain, sel, pt, &bytes);
    if (!p) return SH_E_ARG;
    HIPCHK(hipMemcpy(dst, p, bytes, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int sh_dbg_op(sh_ctx *ctx, int op, const int *a)
{
    if (!ctx || !a) return SH_E_ARG;
    auto okpt = [&](int p) { return p >= 0 && p <= ctx->nst; };
    auto okslice = [&](int s) { return s >= 1 && s <= ctx->ltrot; };
    switch (op) {
    case 0: gemm<0, 0>(ctx, M_W1, 0, M_W2, 0, M_W3, 0); break;
    case 1: gemm<1, 0>(ctx, M_W1, 0, M_W2, 0, M_W3, 0); break;
    case 2: gemm<0, 1>(ctx, M_W1, 0, M_W2, 0, M_W3, 0); break;
    case 3:
        if (!okpt(a[0])) return SH_E_ARG;
        udv_step(ctx, true, a[0], 0);
        break;
    case 4:
        if (!okpt(a[0]) || !okpt(a[1]) || a[0] == a[1]) return SH_E_ARG;
        udv_step(ctx, false, a[0], a[1]);
        break;
    case 5:
        lu_factor(ctx);
        lu_solve(ctx, M_W4);
        break;
    case 6:
        if (!okslice(a[3]) || !okslice(a[4]) || (a[5] != 1 && a[5] != -1)) return SH_E_ARG;
        if ((a[5] > 0) != (a[4] >= a[3]) && a[3] != a[4]) return SH_E_ARG;
        left_pass(ctx, M_W1, 0, a[1], a[0], a[2], a[3], a[4], a[5]);
        break;
    case 7:
        if (!okslice(a[3])) return SH_E_ARG;
        right_pass(ctx, a[1], a[0], a[2], a[3]);
        break;
    case 8:
        if ((!a[0] && !okpt(a[1])) || (!a[2] && !okpt(a[3]))) return SH_E_ARG;
        green(ctx, a[0], a[1], a[2], a[3]);
        break;
    case 9:
        if (!okpt(a[0])) return SH_E_ARG;
        green_bb(ctx, a[0]);
        break;
    case 10:
        if (!okslice(a[0]) || a[1] < 0 || a[1] >= ctx->ltrot) return SH_E_ARG;
        update_slice(ctx, a[0], a[1]);
        break;
    case 11:
        hipLaunchKernelGGL(k_sfmt_draw, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, ctx->nch, 1);
        break;
    case 12: measure(ctx); break;
    case 13: wrap_replace(ctx, a[0]); break;
    case 14: ctx->legacy_update = a[0] != 0; ctx->graph_epoch++; break;
    default: return SH_E_ARG;
    }
    int rc = check_launch(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

// ---- projective sweep (include/sunhip_proj.h) -------------------------------------------------
#include "pqmc.hip"
#include "learn.hip"

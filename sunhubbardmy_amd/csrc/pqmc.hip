// pqmc.hip -- host side of the projective sweep: the C ABI of include/sunhip_proj.h.
// The slice loops mirror src/dqmc_proj_core/dqmc_proj_sweep_b0.f90 and dqmc_proj_sweep_0b.f90
// and the updates mirror src/dqmc_config_v/dqmc_proj_update_submatrix_LR.f90 (bond field) and
// src/dqmc_config_u/dqmc_proj_update_submatrix_LR.f90 (on-site field); every launch advances
// all chains of the batch.  Included by sunhip.hip after the finite-temperature code.
#include "../../include/sunhip_proj.h"

struct shp_ctx : ctx_base {
    int nch = 0, n = 0, ne = 0, ltrot = 0, nwrap = 0, nst = 0, kmode = 0, ns = 0, mid_len = 1;
    int nn_nf = 0, nn_lf = 0, wrap_check = 1, pb = PB;
    int update = SHP_U_LR, gblock = 0;   // update algorithm of the on-site field and its depth
    bool lwrapu = false, lwrapv = true;
    size_t slice_draws = 0;       // uniforms per slice of a pass
    bool tracing = false, started = false;
    std::vector<int> iwrap_nt, wrap_step;
    std::vector<PChain> hc;
    PChain *dc = nullptr;
    DChain *dview = nullptr;      // measurement view of each chain (G, bonds, staggering, bins)
    std::vector<DChain> hview;    // host copy of the views
    size_t eqt_lds = 0;           // LDS of k_eqt, 0 while no chain has correlation bins
    int32_t *d_seeds = nullptr;
    std::vector<cplx *> trace_bufs;
};

namespace {

size_t p_lds_prop(const shp_ctx *c) { return (size_t)cb_ldr(c->n) * std::max(CPB, RPB) * sizeof(cplx); }
size_t p_lds_bonds(const shp_ctx *c) { return (size_t)2 * c->ns * c->ns * sizeof(cplx); }
size_t p_lds_qr(const shp_ctx *c) { return ((size_t)c->n * c->pb + 2 * PB * PB) * sizeof(cplx); }
size_t p_lds_lu(const shp_ctx *c) { return (size_t)c->ne * PB * sizeof(cplx) + (size_t)c->ne * sizeof(int); }

void pgemm(shp_ctx *ctx, int sa, int pa, int sb, int pb_, int sc, int pc, int m, int n, int k, int lda, int ldb,
           int ldc, double ar = 1.0, double br = 0.0)
{
    PGemm g{sa, pa, sb, pb_, sc, pc, m, n, k, lda, ldb, ldc, ar, 0.0, br, 0.0};
    dim3 grid((m + GT - 1) / GT, (n + GT - 1) / GT, ctx->nch);
    hipLaunchKernelGGL(kp_gemm, grid, dim3(256), 0, ctx->st, ctx->dc, g);
}

void pcopy(shp_ctx *ctx, int dsel, int dpt, int ssel, int spt, int count, int conj = 0)
{
    hipLaunchKernelGGL(kp_copy, dim3((count + 255) / 256, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, dsel, dpt,
                       ssel, spt, count, conj);
}

void ptranspose(shp_ctx *ctx, int dsel, int dpt, int ssel, int spt, int rows, int cols, int conj = 0)
{
    unsigned nb = (unsigned)(((rows + 31) / 32) * ((cols + 31) / 32));
    hipLaunchKernelGGL(kp_transpose, dim3(nb, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, dsel, dpt, ssel, spt,
                       rows, cols, conj);
}

// ---- operator programs ------------------------------------------------------------------------
POps ops_begin(int nt, int sides, int kselL, int kselR)
{
    POps o;
    memset(&o, 0, sizeof(o));
    o.nt = nt;
    o.sides = sides;
    o.kselL = kselL;
    o.kselR = kselR;
    return o;
}
void op_kin(POps &o, bool rev) { o.kind[o.nops++] = rev ? 1 : 0; }
void op_v(POps &o, int f, int modeL, int modeR)
{
    o.kind[o.nops] = 2; o.arg[o.nops] = f; o.modeL[o.nops] = modeL; o.modeR[o.nops] = modeR;
    o.nops++;
}
// the on-site field: UR rows / UL columns times bmat (0) or bmat_inv (1)
void op_u(POps &o, int modeL, int modeR)
{
    o.kind[o.nops] = 3; o.arg[o.nops] = 0; o.modeL[o.nops] = modeL; o.modeR[o.nops] = modeR;
    o.nops++;
}

void launch_prop(shp_ctx *ctx, const POps &o)
{
    unsigned nb = (unsigned)((ctx->ne + std::min(CPB, RPB) - 1) / std::min(CPB, RPB));
    hipLaunchKernelGGL(kp_prop, dim3(nb, 2, ctx->nch), dim3(256), p_lds_prop(ctx), ctx->st, ctx->dc, o);
}

// dense kinetic half-step on the selected sides: UR <- T_L UR, UL <- UL T_R
void kin_dense(shp_ctx *ctx, int sides, int kselL, int kselR)
{
    const int n = ctx->n, ne = ctx->ne;
    if (sides & 1) {
        pgemm(ctx, kselL ? P_KTM1 : P_KT, 0, P_UR, 0, P_T1, 0, n, ne, n, n, n, n);
        pcopy(ctx, P_UR, 0, P_T1, 0, n * ne);
    }
    if (sides & 2) {
        pgemm(ctx, P_UL, 0, kselR ? P_KTM1 : P_KT, 0, P_T1, 0, ne, n, n, ne, n, ne);
        pcopy(ctx, P_UL, 0, P_T1, 0, n * ne);
    }
}

// checkerboard: one launch runs the whole program; dense: the kinetic steps are GEMM pairs and
// the field operators between them are grouped into launches
void run_ops(shp_ctx *ctx, const POps &o)
{
    Tic t(ctx, T_WRAP);
    if (ctx->kmode != SH_K_DENSE) {
        launch_prop(ctx, o);
        return;
    }
    POps part = ops_begin(o.nt, o.sides, o.kselL, o.kselR);
    for (int k = 0; k < o.nops; k++) {
        if (o.kind[k] < 2) {
            if (part.nops) { launch_prop(ctx, part); part.nops = 0; }
            kin_dense(ctx, o.sides, o.kselL, o.kselR);
        } else {
            part.kind[part.nops] = o.kind[k]; part.arg[part.nops] = o.arg[k];
            part.modeL[part.nops] = o.modeL[k]; part.modeR[part.nops] = o.modeR[k];
            part.nops++;
        }
    }
    if (part.nops) launch_prop(ctx, part);
}

void green(shp_ctx *ctx, int ul, int ur, int inv, int g);
void inverse(shp_ctx *ctx, int ul, int ur, int inv, int slot);

// the G-based forms for the bond field (src/dqmc_config_v/dqmc_proj_update_delay.f90,
// dqmc_proj_update_submatrix_G.f90): G from the current propagators; the finite-temperature bond
// update on it through the chain's view, the delay form or every accepted bond applied to G at
// once (its draws: the pass buffer at 2 nn_lf nn_nf per slice, the bond field alone); UR's rows of
// the flipped bonds; a fresh (UL UR)^-1
void update_family_g(shp_ctx *ctx, int nt, int nf, size_t uoff0)
{
    const int nl = ctx->nn_lf;
    green(ctx, P_UL, P_UR, P_ULRINV, P_G);
    Tic t(ctx, T_UPDATE);
    hipLaunchKernelGGL(kp_v_save, dim3((nl + 255) / 256, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, nt, nf);
    const int slice = (int)(uoff0 / ctx->slice_draws);
    const int p = (int)((uoff0 % ctx->slice_draws) / (2 * (size_t)nl));
    if (ctx->update == SHP_U_DELAY) {
        hipLaunchKernelGGL(k_vupdate_delay, dim3(ctx->nch), dim3(256), (size_t)4 * 64 * sizeof(cplx), ctx->st,
                           ctx->dview, nt, nf, slice, p);
    } else if (ctx->update == SHP_U_SUBMATRIX_G) {
        const int K = std::max(2, ctx->gblock & ~1);
        hipLaunchKernelGGL(k_vupdate_submatrix, dim3(ctx->nch), dim3(256), (size_t)(K * (K + 1) + 8 * K) * sizeof(cplx),
                           ctx->st, ctx->dview, nt, nf, slice, p);
    } else
        hipLaunchKernelGGL(k_vupdate, dim3(ctx->nch), dim3(256), (size_t)4 * ctx->n * sizeof(cplx), ctx->st,
                           ctx->dview, nt, nf, slice, p);
    hipLaunchKernelGGL(kp_v_rescale, dim3((unsigned)(((size_t)nl * ctx->ne + 255) / 256), 1, ctx->nch), dim3(256), 0,
                       ctx->st, ctx->dc, nt, nf);
    inverse(ctx, P_UL, P_UR, P_ULRINV, -1);
}

// ---- submatrix-LR update of family nf at slice nt (the fast form: stocks of one bond), or a
// G-based form; uoff = its first uniform in the pass ----
void update_family(shp_ctx *ctx, int nt, int nf, size_t uoff0)
{
    if (ctx->update == SHP_U_DELAY || ctx->update == SHP_U_SUBMATRIX_G) {
        update_family_g(ctx, nt, nf, uoff0);
        return;
    }
    Tic t(ctx, T_UPDATE);
    const int ns = ctx->ns, ne = ctx->ne, nl = ctx->nn_lf;
    const int per = ctx->update == SHP_U_FAST ? 1 : ns / 2;
    for (int ib0 = 0; ib0 < nl; ib0 += per) {
        const int nb = std::min(per, nl - ib0), nup = 2 * nb;
        const int cnt = std::max(ns * ne, ns * ns);
        hipLaunchKernelGGL(kp_stock_gather, dim3((cnt + 255) / 256, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc,
                           nf, ib0, nup);
        // Rrows = urP ULRINV, Felms = Rrows ulP (:121-133)
        pgemm(ctx, P_URP, 0, P_ULRINV, 0, P_RROWS, 0, nup, ne, ne, ns, ne, ns);
        pgemm(ctx, P_RROWS, 0, P_ULP, 0, P_FELMS, 0, nup, nup, ne, ns, ne, ns);
        const int uoff = (int)(uoff0 + 2 * (size_t)ib0);
        hipLaunchKernelGGL(kp_stock_bonds, dim3(ctx->nch), dim3(256), p_lds_bonds(ctx), ctx->st, ctx->dc, nt, nf, ib0,
                           nb, uoff);
        // Woodbury flush ULRINV -= (ULRINV Umat)(Gamma^-1 Vmat) over the whole stock width; rows
        // and columns past the accepted count are zero (:240-266)
        pgemm(ctx, P_GINV, 0, P_VMAT, 0, P_VTMP, 0, ns, ne, ns, ns, ns, ns);
        pgemm(ctx, P_ULRINV, 0, P_UMAT, 0, P_UTMP, 0, ne, ns, ne, ne, ne, ne);
        pgemm(ctx, P_UTMP, 0, P_VTMP, 0, P_ULRINV, 0, ne, ne, ns, ne, ns, ne, -1.0, 1.0);
        hipLaunchKernelGGL(kp_stock_scale, dim3((ns * ne + 255) / 256, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc);
    }
}

// the G-based forms (dqmc_proj_update_delay.f90, dqmc_proj_update_submatrix_G.f90): G from the
// current UR, (UL UR)^-1, UL; the finite-temperature delay or submatrix kernel on it through the
// chain's view (its uniforms are the pass buffer at 2 N per slice: the on-site field alone);
// UR's flipped rows; a fresh (UL UR)^-1

void update_u_g(shp_ctx *ctx, int nt, size_t uoff0)
{
    const int n = ctx->n;
    green(ctx, P_UL, P_UR, P_ULRINV, P_G);
    Tic t(ctx, T_UPDATE);
    hipLaunchKernelGGL(kp_u_save, dim3((n + 255) / 256, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, nt);
    const int slice = (int)(uoff0 / (2 * (size_t)n));
    if (ctx->update == SHP_U_DELAY) {
        const size_t lds = (size_t)(n + 2 * 64) * sizeof(cplx) + (size_t)n * 20;
        hipLaunchKernelGGL(k_update_delay, dim3(ctx->nch), dim3(256), lds, ctx->st, ctx->dview, nt, slice);
    } else {
        const int K = ctx->gblock, dss = sub_stage(n, K);
        hipLaunchKernelGGL(k_update_submatrix, dim3(ctx->nch), dim3(256), lds_sub(n, K, dss), ctx->st, ctx->dview, nt,
                           slice, dss);
    }
    hipLaunchKernelGGL(kp_u_rescale, dim3((unsigned)(((size_t)n * ctx->ne + 255) / 256), 1, ctx->nch), dim3(256), 0,
                       ctx->st, ctx->dc, nt);
    inverse(ctx, P_UL, P_UR, P_ULRINV, -1);
}

// ---- update of the on-site field at slice nt: submatrix-LR in stocks of ns consecutive sites
// (the fast form: stocks of one site), or a G-based form ----
void update_u(shp_ctx *ctx, int nt, size_t uoff0)
{
    if (ctx->update == SHP_U_DELAY || ctx->update == SHP_U_SUBMATRIX_G) {
        update_u_g(ctx, nt, uoff0);
        return;
    }
    Tic t(ctx, T_UPDATE);
    const int ns = ctx->ns, ne = ctx->ne, n = ctx->n;
    // the fast update is the same algebra one site at a time: stocks of one site
    const int step = ctx->update == SHP_U_FAST ? 1 : ns;
    for (int i0 = 0; i0 < n; i0 += step) {
        const int nup = std::min(step, n - i0);
        const int cnt = std::max(ns * ne, ns * ns);
        hipLaunchKernelGGL(kp_stock_gather, dim3((cnt + 255) / 256, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc,
                           -1, i0, nup);
        pgemm(ctx, P_URP, 0, P_ULRINV, 0, P_RROWS, 0, nup, ne, ne, ns, ne, ns);
        pgemm(ctx, P_RROWS, 0, P_ULP, 0, P_FELMS, 0, nup, nup, ne, ns, ne, ns);
        hipLaunchKernelGGL(kp_stock_sites, dim3(ctx->nch), dim3(256), p_lds_bonds(ctx), ctx->st, ctx->dc, nt, i0, nup,
                           uoff0 + 2 * (size_t)i0);
        pgemm(ctx, P_GINV, 0, P_VMAT, 0, P_VTMP, 0, ns, ne, ns, ns, ns, ns);
        pgemm(ctx, P_ULRINV, 0, P_UMAT, 0, P_UTMP, 0, ne, ns, ne, ne, ne, ne);
        pgemm(ctx, P_UTMP, 0, P_VTMP, 0, P_ULRINV, 0, ne, ne, ns, ne, ns, ne, -1.0, 1.0);
        hipLaunchKernelGGL(kp_stock_scale, dim3((ns * ne + 255) / 256, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc);
    }
}

// ---- scratch quantities ---------------------------------------------------------------------
// inv = (ul ur)^-1; with slot >= 0 its log det goes to sc[slot]
void inverse(shp_ctx *ctx, int ul, int ur, int inv, int slot)
{
    {
        Tic t(ctx, T_GEMM);
        pgemm(ctx, ul, 0, ur, 0, P_ULR, 0, ctx->ne, ctx->ne, ctx->n, ctx->ne, ctx->n, ctx->ne);
    }
    Tic t(ctx, T_LU);
    hipLaunchKernelGGL(kp_lu_inv, dim3(ctx->nch), dim3(BLK_NT), p_lds_lu(ctx), ctx->st, ctx->dc, P_ULR, inv, slot);
}

// g = 1 - ur inv ul (src/dqmc_proj_core/green_equaltime.f90:14-21)
void green(shp_ctx *ctx, int ul, int ur, int inv, int g)
{
    Tic t(ctx, T_GEMM);
    const int n = ctx->n, ne = ctx->ne;
    pgemm(ctx, ur, 0, inv, 0, P_T1, 0, n, ne, ne, n, ne, n);
    pgemm(ctx, P_T1, 0, ul, 0, g, 0, n, n, ne, n, ne, n, -1.0, 0.0);
    hipLaunchKernelGGL(kp_add_identity, dim3((n + 255) / 256, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, g);
}

// equal-time measurement in the middle window: ULRINV is replaced by a scratch inverse
// (s_invlu_z), as the reference does (dqmc_proj_sweep_0b.f90:108-113)
void measure(shp_ctx *ctx)
{
    inverse(ctx, P_UL, P_UR, P_ULRINV, -1);
    green(ctx, P_UL, P_UR, P_ULRINV, P_G);
    Tic t(ctx, T_MEASURE);
    const int nb = (ctx->n + 31) / 32;
    hipLaunchKernelGGL(k_measure_tiles, dim3(nb * (nb + 1) / 2, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dview);
    hipLaunchKernelGGL(k_measure, dim3(ctx->nch), dim3(256), 0, ctx->st, ctx->dview);
    if (ctx->eqt_lds) hipLaunchKernelGGL(k_eqt, dim3(ctx->nch), dim3(256), ctx->eqt_lds, ctx->st, ctx->dview);
}

bool in_window(const shp_ctx *ctx, int nt)
{
    const int lt = ctx->ltrot, ml = ctx->mid_len;
    return nt > (lt / 2 - (ml + 1) / 2) && nt <= (lt / 2 + ml / 2);
}

void weights(shp_ctx *ctx, int op, int n, int phasediff)
{
    if (ctx->lwrapu && op <= 2) hipLaunchKernelGGL(kp_uphase, dim3(ctx->nch), dim3(256), 0, ctx->st, ctx->dc);
    hipLaunchKernelGGL(kp_weights, dim3((ctx->nch + 63) / 64), dim3(64), 0, ctx->st, ctx->dc, ctx->nch, op, n,
                       phasediff);
}

// after the QR of a stabilization point: the scratch inverse and weight and, when checked or
// traced, the scratch G compared with G from the propagators saved in ULo / URo
void scratch(shp_ctx *ctx, int op, int n)
{
    inverse(ctx, P_UL, P_UR, P_ULRINV, SC_ZLD);
    {
        Tic t(ctx, T_OTHER);
        weights(ctx, op, n, ctx->wrap_check);
    }
    if (ctx->wrap_check || ctx->tracing) green(ctx, P_UL, P_UR, P_ULRINV, P_G);
    if (ctx->wrap_check) {
        inverse(ctx, P_ULO, P_URO, P_INV2, -1);
        green(ctx, P_ULO, P_URO, P_INV2, P_G2);
    }
    if (ctx->wrap_check || ctx->tracing) {
        Tic t(ctx, T_OTHER);
        hipLaunchKernelGGL(kp_wrap_check, dim3(ctx->nch), dim3(256), 0, ctx->st, ctx->dc, ctx->wrap_check,
                           ctx->tracing ? 1 : 0);
    }
}

void save_old(shp_ctx *ctx)
{
    if (!ctx->wrap_check) return;
    Tic t(ctx, T_OTHER);
    pcopy(ctx, P_ULO, 0, P_UL, 0, ctx->n * ctx->ne);
    pcopy(ctx, P_URO, 0, P_UR, 0, ctx->n * ctx->ne);
}

// s_zgeQR_pqmc of asel (N x ne) with Q into T1
void qr_to_t1(shp_ctx *ctx, int asel, int apt)
{
    Tic t(ctx, T_QR);
    hipLaunchKernelGGL(kp_qr_formq, dim3(ctx->nch), dim3(BLK_NT), p_lds_qr(ctx), ctx->st, ctx->dc, asel, apt, P_T1, 0);
}

// Bmat_left_forward on UR (src/dqmc_proj_core.f90:68-116): K half, the bond fields of every
// family (nflag 2 then 1), K half reversed
void bmat_left_forward(shp_ctx *ctx, int nt1, int nt2)
{
    for (int nt = nt2; nt <= nt1; nt++) {
        POps o = ops_begin(nt, 1, 0, 0);
        op_kin(o, false);
        for (int f = 0; f < ctx->nn_nf && ctx->lwrapv; f++) {
            op_v(o, f, 0, 0);
            op_v(o, f, 1, 0);
        }
        if (ctx->lwrapu) op_u(o, 0, 0);
        op_kin(o, true);
        run_ops(ctx, o);
    }
}

int first_info_p(shp_ctx *ctx)
{
    int info = 0;
    for (auto &c : ctx->hc) {
        double v = 0.0;
        if (hipMemcpy(&v, c.sc + SC_INFO, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
        if (v != 0.0 && info == 0) info = (int)v;
    }
    return info;
}

int check_launch_p(shp_ctx *ctx)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, SH_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
}

int finish_pass(shp_ctx *ctx, const char *what)
{
    int rc = check_launch_p(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    int info = first_info_p(ctx);
    if (info > 0) return fail(ctx, info, std::string("singular UL UR overlap in ") + what);
    if (info < 0) return fail(ctx, SH_E_HIP, "reading the chain state failed");
    return 0;
}

}  // namespace

// ---- C ABI ---------------------------------------------------------------------------------
extern "C" {

int shp_destroy(shp_ctx *ctx)
{
    if (!ctx) return 0;
    if (ctx->st) hipStreamSynchronize(ctx->st);
    for (void *p : ctx->allocs) hipFree(p);
    for (cplx *p : ctx->trace_bufs) if (p) hipFree(p);
    for (auto e : ctx->pool) hipEventDestroy(e);
    if (ctx->st) hipStreamDestroy(ctx->st);
    delete ctx;
    return 0;
}

const char *shp_last_error(const shp_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int shp_create(shp_ctx **out, int nchains, const shp_model *m, int device)
{
    if (!out || nchains <= 0 || !m) return SH_E_ARG;
    *out = nullptr;
    shp_ctx *ctx = new shp_ctx();
    const shp_model &q0 = m[0];
    ctx->dev = device;
    ctx->nch = nchains;
    ctx->n = q0.ndim; ctx->ne = q0.ne; ctx->ltrot = q0.ltrot; ctx->nwrap = q0.nwrap; ctx->nst = q0.nst;
    ctx->kmode = q0.kmode; ctx->ns = q0.nustock; ctx->mid_len = q0.mid_len; ctx->nn_nf = q0.nn_nf;
    ctx->nn_lf = q0.nn_lf; ctx->wrap_check = q0.wrap_check ? 1 : 0;
    ctx->lwrapv = q0.vtab != nullptr;
    ctx->lwrapu = q0.ubmat != nullptr;
    ctx->update = q0.update;
    ctx->gblock = q0.gblock;
    ctx->slice_draws = (ctx->lwrapu ? (size_t)2 * q0.ndim : 0) + (ctx->lwrapv ? (size_t)2 * q0.nn_lf * q0.nn_nf : 0);
    for (int c = 0; c < nchains; c++) {
        const shp_model &q = m[c];
        if (q.ndim != ctx->n || q.ne != ctx->ne || q.ltrot != ctx->ltrot || q.nwrap != ctx->nwrap ||
            q.nst != ctx->nst || q.kmode != ctx->kmode || q.nustock != ctx->ns || q.nn_nf != ctx->nn_nf ||
            q.nn_lf != ctx->nn_lf || q.mid_len != ctx->mid_len || (q.wrap_check != 0) != (ctx->wrap_check != 0)) {
            int rc = fail(ctx, SH_E_ARG, "all chains of a projective batch must share their sizes and schedule");
            delete ctx;
            return rc;
        }
        if (!q.iwrap_nt || !q.wrap_step || !q.bonds || !q.urt || !q.urtm1 || (q.vtab && !q.vdelta) || !q.stag ||
            (q.vtab != nullptr) != (m[0].vtab != nullptr) || (q.ubmat != nullptr) != (m[0].ubmat != nullptr)) {
            delete ctx;
            return SH_E_ARG;
        }
        if (q.update != q0.update || q.gblock != q0.gblock || q.update < SHP_U_LR ||
            q.update > SHP_U_SUBMATRIX_G ||
            (q.update >= SHP_U_DELAY && ((q.ubmat != nullptr) == (q.vtab != nullptr) || q.gblock < 1 || q.gblock > 64))) {
            int rc = fail(ctx, SH_E_ARG, "update: one algorithm for the batch; the G-based forms need one field alone "
                                         "(on-site or bond) and a depth 1..64");
            delete ctx;
            return rc;
        }
        if (q.ubmat && (q.ulcomp < 2 || q.ulcomp > 4 || !q.ubmat_inv || !q.udelta || !q.uratio || !q.ugaml ||
                        !q.uphase_arg || q.nn_nf * 2 + 3 > POPS_MAX)) {
            int rc = fail(ctx, SH_E_ARG, "incomplete on-site field tables");
            delete ctx;
            return rc;
        }
        if (q.ndim < 2 || q.ndim > 1024 || q.ne < 1 || q.ne > q.ndim || q.ne > 512 || q.ltrot < 1 || q.nst < 1 ||
            q.nustock < 2 || q.nustock > 64 || (q.nustock & 1) || q.nn_nf < 1 || q.nn_lf < 1 ||
            q.nn_nf * 2 + 2 > POPS_MAX || q.nflr < 1 || q.mid_len < 0 || q.lq < 1 ||
            (q.kmode != SH_K_CHECKERBOARD && q.kmode != SH_K_DENSE)) {
            delete ctx;
            return SH_E_ARG;
        }
        for (int b = 0; b < q.nn_nf * q.nn_lf * 2; b++)
            if (q.bonds[b] < 0 || q.bonds[b] >= q.ndim) { delete ctx; return SH_E_ARG; }
        for (int i = 0; i <= q.ltrot; i++)
            if (q.iwrap_nt[i] < 0 || q.iwrap_nt[i] > q.nst) { delete ctx; return SH_E_ARG; }
        for (int i = 0; i < 2 * q.nst; i++)
            if (q.wrap_step[i] < 1 || q.wrap_step[i] > q.ltrot) { delete ctx; return SH_E_ARG; }
    }
    // panel width of the QR: 16 columns while an N x 16 panel fits in LDS, else 8
    ctx->pb = ((size_t)ctx->n * PB + 2 * PB * PB) * sizeof(cplx) <= 150 * 1024 ? PB : PB / 2;
    if (p_lds_qr(ctx) > 150 * 1024 || p_lds_lu(ctx) > 150 * 1024 || p_lds_prop(ctx) > 150 * 1024) {
        int rc = fail(ctx, SH_E_ARG, "matrix sizes beyond the LDS budget of the projective kernels");
        delete ctx;
        return rc;
    }
    // arguments are validated before the first HIP call
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return SH_E_HIP; }
    ctx->iwrap_nt.assign(q0.iwrap_nt, q0.iwrap_nt + q0.ltrot + 1);
    ctx->wrap_step.assign(q0.wrap_step, q0.wrap_step + 2 * q0.nst);
    if (hipStreamCreate(&ctx->st) != hipSuccess) { delete ctx; return SH_E_HIP; }
    {
        struct { const void *f; size_t bytes; } big[] = {
            {(const void *)kp_prop, p_lds_prop(ctx)}, {(const void *)kp_stock_bonds, p_lds_bonds(ctx)},
            {(const void *)kp_stock_sites, p_lds_bonds(ctx)},
            {(const void *)kp_qr_formq, p_lds_qr(ctx)}, {(const void *)kp_lu_inv, p_lds_lu(ctx)}};
        for (auto &b : big) {
            if (b.bytes <= 64 * 1024) continue;
            if (hipFuncSetAttribute(b.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b.bytes) != hipSuccess) {
                (void)hipGetLastError();
                int rc = fail(ctx, SH_E_HIP, "dynamic LDS request above the device limit");
                shp_destroy(ctx);
                return rc;
            }
        }
    }
    const int n = ctx->n, ne = ctx->ne, ns = ctx->ns;
    const size_t nne = (size_t)n * ne, nn = (size_t)n * n;
    std::vector<int32_t> seeds(nchains);
    std::vector<DChain> views(nchains);
    ctx->hc.resize(nchains);
    for (int ci = 0; ci < nchains; ci++) {
        const shp_model &q = m[ci];
        PChain &c = ctx->hc[ci];
        memset(&c, 0, sizeof(c));
        c.n = n; c.ne = ne; c.nn_nf = q.nn_nf; c.nn_lf = q.nn_lf; c.kmode = q.kmode; c.ns = ns; c.ltrot = q.ltrot;
        c.nst = q.nst; c.nflr = q.nflr; c.lq = q.lq; c.pb = ctx->pb;
        c.UR = dalloc<cplx>(ctx, nne); c.UL = dalloc<cplx>(ctx, nne); c.T1 = dalloc<cplx>(ctx, nne);
        c.T2 = dalloc<cplx>(ctx, nne); c.ULo = dalloc<cplx>(ctx, nne); c.URo = dalloc<cplx>(ctx, nne);
        c.ULRINV = dalloc<cplx>(ctx, (size_t)ne * ne); c.ULR = dalloc<cplx>(ctx, (size_t)ne * ne);
        c.INV2 = dalloc<cplx>(ctx, (size_t)ne * ne);
        c.G = dalloc<cplx>(ctx, nn); c.G2 = dalloc<cplx>(ctx, nn);
        c.Ust = dalloc<cplx>(ctx, nne * (q.nst + 1)); c.proj = dalloc<cplx>(ctx, nne);
        c.urP = dalloc<cplx>(ctx, (size_t)ns * ne); c.ulP = dalloc<cplx>(ctx, (size_t)ns * ne);
        c.Rrows = dalloc<cplx>(ctx, (size_t)ns * ne); c.Felms = dalloc<cplx>(ctx, (size_t)ns * ns);
        c.Ginv = dalloc<cplx>(ctx, (size_t)ns * ns); c.Vmat = dalloc<cplx>(ctx, (size_t)ns * ne);
        c.Umat = dalloc<cplx>(ctx, (size_t)ns * ne); c.Vtmp = dalloc<cplx>(ctx, (size_t)ns * ne);
        c.Utmp = dalloc<cplx>(ctx, (size_t)ns * ne); c.Dvec = dalloc<cplx>(ctx, ns);
        c.tau = dalloc<cplx>(ctx, ne); c.tblk = dalloc<cplx>(ctx, (size_t)(ne / (PB / 2) + 2) * PB * PB);
        c.wk = dalloc<cplx>(ctx, (size_t)PB * std::max(n, ne));
        c.piv = dalloc<int32_t>(ctx, ne);
        c.conf = dalloc<int32_t>(ctx, (size_t)n * q.nn_nf * q.ltrot);
        c.xvec = dalloc<int32_t>(ctx, ns); c.svec = dalloc<int32_t>(ctx, ns); c.ikd = dalloc<int32_t>(ctx, 4);
        c.unif = dalloc<double>(ctx, std::max(ctx->slice_draws, (size_t)1) * q.ltrot);
        c.lwrapu = ctx->lwrapu ? 1 : 0;
        c.lwrapv = ctx->lwrapv ? 1 : 0;
        if (ctx->lwrapu) {
            c.ulcomp = q.ulcomp;
            c.uphase_from_conf = q.uphase_from_conf ? 1 : 0;
            c.uconf = dalloc<int32_t>(ctx, (size_t)n * q.ltrot);
            for (int l = 0; l < q.ulcomp; l++) {
                c.ub[l] = make_double2(q.ubmat[2 * l], q.ubmat[2 * l + 1]);
                c.ubi[l] = make_double2(q.ubmat_inv[2 * l], q.ubmat_inv[2 * l + 1]);
                c.ug[l] = q.ugaml[l];
                c.upa[l] = q.uphase_arg[l];
            }
            for (int k = 0; k < (q.ulcomp - 1) * q.ulcomp; k++) {
                c.ud[k] = make_double2(q.udelta[2 * k], q.udelta[2 * k + 1]);
                c.uratio[k] = make_double2(q.uratio[2 * k], q.uratio[2 * k + 1]);
            }
        }
        c.sc = dalloc<double>(ctx, SC_SIZE);
        c.logwDV = dalloc<double>(ctx, q.nst + 2);
        c.rwork = dalloc<double>(ctx, 4 * (size_t)n);
        c.rng = dalloc<uint32_t>(ctx, 625);
        c.bonds = dupload<int32_t>(ctx, q.bonds, (size_t)2 * q.nn_nf * q.nn_lf);
        size_t kl = (q.kmode == SH_K_CHECKERBOARD) ? (size_t)4 * q.nn_nf * q.nn_lf : nn;
        c.kt = dupload<cplx>(ctx, (const cplx *)q.urt, kl);
        c.ktm1 = dupload<cplx>(ctx, (const cplx *)q.urtm1, kl);
        c.stag = dupload<double>(ctx, q.stag, n);
        if (!c.UR || !c.UL || !c.T1 || !c.T2 || !c.ULo || !c.URo || !c.ULRINV || !c.ULR || !c.INV2 || !c.G || !c.G2 ||
            !c.Ust || !c.proj || !c.urP || !c.ulP || !c.Rrows || !c.Felms || !c.Ginv || !c.Vmat || !c.Umat ||
            !c.Vtmp || !c.Utmp || !c.Dvec || !c.tau || !c.tblk || !c.wk || !c.piv || !c.conf || !c.xvec ||
            !c.svec || !c.ikd || !c.unif || !c.sc || !c.logwDV || !c.rwork || !c.rng || !c.bonds || !c.kt ||
            !c.ktm1 || !c.stag || (ctx->lwrapu && !c.uconf)) {
            int rc = fail(ctx, SH_E_HIP, "device allocation failed");
            shp_destroy(ctx);
            return rc;
        }
        if (q.vtab) {
            for (int k = 0; k < 8; k++) c.vt[k] = make_double2(q.vtab[2 * k], q.vtab[2 * k + 1]);
            for (int k = 0; k < 4; k++) c.vd[k] = make_double2(q.vdelta[2 * k], q.vdelta[2 * k + 1]);
        }
        std::vector<int32_t> ones((size_t)n * q.nn_nf * q.ltrot, 1);
        hipMemcpy(c.conf, ones.data(), ones.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        if (ctx->lwrapu) hipMemcpy(c.uconf, ones.data(), (size_t)n * q.ltrot * sizeof(int32_t), hipMemcpyHostToDevice);
        double sc0[SC_SIZE] = {0};
        sc0[SC_PHASE] = 1.0;
        hipMemcpy(c.sc, sc0, sizeof(sc0), hipMemcpyHostToDevice);
        seeds[ci] = q.seed;
        // the measurement kernels read G, the bonds, the staggering and the bins through a view
        DChain &v = views[ci];
        memset(&v, 0, sizeof(v));
        v.n = n; v.lq = q.lq; v.nn_nf = q.nn_nf; v.nn_lf = q.nn_lf; v.nflr = q.nflr; v.lcomp = 2;
        v.G = c.G; v.bonds = c.bonds; v.stag = c.stag; v.sc = c.sc; v.rwork = c.rwork;
        v.kin_weight = q.kin_weight != 0.0 ? q.kin_weight : 1.0;
        if (ctx->update >= SHP_U_DELAY) {
            // ... and the G-based updates of the on-site field the HS field, its draws and tables,
            // the delayed vectors and the submatrix scratch
            const int K = ctx->gblock;
            c.uold = dalloc<int32_t>(ctx, n);
            v.ltrot = q.ltrot; v.nublock = K;
            v.update = ctx->update == SHP_U_DELAY ? SH_U_DELAY : SH_U_SUBMATRIX;
            if (ctx->lwrapu) {
                v.conf = c.uconf; v.unif = c.unif;
                v.lcomp = q.ulcomp;
                v.lwrapu = 1;
                for (int l = 0; l < q.ulcomp; l++) v.gaml[l] = q.ugaml[l];
                for (int k = 0; k < (q.ulcomp - 1) * q.ulcomp; k++) {
                    v.delta_bmat[k] = c.ud[k];
                    v.phase_ratio[k] = c.uratio[k];
                }
            } else {
                // the bond field: its draws, tables and field; acceptance into the chain's
                // (accepted, tried) slots
                v.confv = c.conf; v.unifv = c.unif; v.lwrapv = 1;
                for (int k = 0; k < 8; k++) v.vtb[k] = c.vt[k];
                for (int k = 0; k < 4; k++) v.vdl[k] = c.vd[k];
                v.vacc = c.sc + SC_ACC;
            }
            // (the bond field's submatrix form borders by 2 x 2 blocks: at least two columns)
            const int Ka = std::max(K, 2);
            v.avec = dalloc<cplx>(ctx, (size_t)n * Ka);
            v.bvec = dalloc<cplx>(ctx, (size_t)n * Ka);
            v.W1 = dalloc<cplx>(ctx, (size_t)n * Ka);
            v.W2 = dalloc<cplx>(ctx, (size_t)Ka * Ka);
            if (!c.uold || !v.avec || !v.bvec || !v.W1 || !v.W2) {
                int rc = fail(ctx, SH_E_HIP, "device allocation failed");
                shp_destroy(ctx);
                return rc;
            }
        }
    }
    if (ctx->update == SHP_U_SUBMATRIX_G) {
        const int Kb = std::max(2, ctx->gblock & ~1);
        const size_t lds = SUB_LDS_CAP;   // sub_stage's cap; the attribute is shared per function
        const size_t ldsv = (size_t)(Kb * (Kb + 1) + 8 * Kb) * sizeof(cplx);
        if ((lds > 64 * 1024 &&
             hipFuncSetAttribute((const void *)k_update_submatrix, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
                 hipSuccess) ||
            (ldsv > 64 * 1024 &&
             hipFuncSetAttribute((const void *)k_vupdate_submatrix, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsv) !=
                 hipSuccess)) {
            (void)hipGetLastError();
            int rc = fail(ctx, SH_E_HIP, "dynamic LDS request above the device limit");
            shp_destroy(ctx);
            return rc;
        }
    }
    ctx->dc = dupload<PChain>(ctx, ctx->hc.data(), nchains);
    ctx->dview = dupload<DChain>(ctx, views.data(), nchains);
    ctx->hview = views;
    ctx->d_seeds = dupload<int32_t>(ctx, seeds.data(), nchains);
    if (!ctx->dc || !ctx->dview || !ctx->d_seeds) { shp_destroy(ctx); return SH_E_HIP; }
    hipLaunchKernelGGL(kp_seed, dim3((nchains + 63) / 64), dim3(64), 0, ctx->st, ctx->dc, nchains, ctx->d_seeds);
    if (hipStreamSynchronize(ctx->st) != hipSuccess || hipGetLastError() != hipSuccess) {
        shp_destroy(ctx);
        return SH_E_HIP;
    }
    *out = ctx;
    return 0;
}

int shp_init_conf(shp_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    hipLaunchKernelGGL(kp_draw, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, 0, 0, 0, (double *)nullptr);
    int rc = check_launch_p(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_set_conf(shp_ctx *ctx, int chain, const int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    const size_t n = (size_t)ctx->n * ctx->nn_nf * ctx->ltrot;
    for (size_t k = 0; k < n; k++)
        if (conf[k] < 1 || conf[k] > 2) return fail(ctx, SH_E_ARG, "bond field values must be 1 or 2");
    HIPCHK(hipMemcpyAsync(ctx->hc[chain].conf, conf, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_get_conf(shp_ctx *ctx, int chain, int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    const size_t n = (size_t)ctx->n * ctx->nn_nf * ctx->ltrot;
    HIPCHK(hipMemcpyAsync(conf, ctx->hc[chain].conf, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_set_uconf(shp_ctx *ctx, int chain, const int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    if (!ctx->lwrapu) return fail(ctx, SH_E_STATE, "the model has no on-site field");
    const PChain &c = ctx->hc[chain];
    const size_t n = (size_t)ctx->n * ctx->ltrot;
    for (size_t k = 0; k < n; k++)
        if (conf[k] < 1 || conf[k] > c.ulcomp) return fail(ctx, SH_E_ARG, "on-site field values out of 1..ulcomp");
    HIPCHK(hipMemcpyAsync(c.uconf, conf, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_get_uconf(shp_ctx *ctx, int chain, int32_t *conf)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !conf) return SH_E_ARG;
    if (!ctx->lwrapu) return fail(ctx, SH_E_STATE, "the model has no on-site field");
    const size_t n = (size_t)ctx->n * ctx->ltrot;
    HIPCHK(hipMemcpyAsync(conf, ctx->hc[chain].uconf, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_get_rng_state(shp_ctx *ctx, int chain, uint32_t *state)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !state) return SH_E_ARG;
    HIPCHK(hipMemcpyAsync(state, ctx->hc[chain].rng, 625 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_set_rng_state(shp_ctx *ctx, int chain, const uint32_t *state)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !state) return SH_E_ARG;
    if (state[624] > 624) return fail(ctx, SH_E_ARG, "SFMT read index out of 0..624");
    HIPCHK(hipMemcpyAsync(ctx->hc[chain].rng, state, 625 * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_draw_uniforms(shp_ctx *ctx, int chain, int count, double *out)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || count < 0 || (count > 0 && !out)) return SH_E_ARG;
    if (count == 0) return 0;
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, (size_t)count * sizeof(double)));
    hipLaunchKernelGGL(kp_draw, dim3(1), dim3(64), 0, ctx->st, ctx->dc, 2, chain, count, d);
    int rc = check_launch_p(ctx);
    if (!rc && hipMemcpyAsync(out, d, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, ctx->st) != hipSuccess)
        rc = fail(ctx, SH_E_HIP, "copy of the drawn numbers failed");
    if (hipStreamSynchronize(ctx->st) != hipSuccess && !rc) rc = fail(ctx, SH_E_HIP, "stream synchronisation failed");
    hipFree(d);
    return rc;
}

int shp_set_proj(shp_ctx *ctx, int chain, const double *proj)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !proj) return SH_E_ARG;
    HIPCHK(hipMemcpyAsync(ctx->hc[chain].proj, proj, (size_t)ctx->n * ctx->ne * sizeof(cplx), hipMemcpyHostToDevice,
                          ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_sweep_start_0b(shp_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    const int n = ctx->n, ne = ctx->ne;
    pcopy(ctx, P_UR, 0, P_PROJ, 0, n * ne);
    weights(ctx, 3, 0, 0);
    for (int p = 1; p <= ctx->nst; p++) {
        bmat_left_forward(ctx, ctx->wrap_step[2 * (p - 1) + 1], ctx->wrap_step[2 * (p - 1)]);
        qr_to_t1(ctx, P_UR, 0);
        pcopy(ctx, P_UR, 0, P_T1, 0, n * ne);
        pcopy(ctx, P_UST, p, P_T1, 0, n * ne);
        weights(ctx, 4, p, 0);
    }
    ptranspose(ctx, P_UL, 0, P_PROJ, 0, n, ne, 1);
    inverse(ctx, P_UL, P_UR, P_ULRINV, SC_ZLD);
    weights(ctx, 0, 0, 0);
    int rc = finish_pass(ctx, "sweep_start_0b");
    if (rc) return rc;
    ctx->started = true;
    return 0;
}

int shp_sweep_b0(shp_ctx *ctx, int lupdate, int lmeasure)
{
    if (!ctx) return SH_E_ARG;
    if (!ctx->started) return fail(ctx, SH_E_STATE, "sweep_b0 before sweep_start_0b");
    const int n = ctx->n, ne = ctx->ne, nf = ctx->nn_nf;
    ptranspose(ctx, P_UL, 0, P_PROJ, 0, n, ne, 1);
    pcopy(ctx, P_UST, 0, P_PROJ, 0, n * ne);
    if (lupdate) hipLaunchKernelGGL(kp_draw, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, 1, 0, 0, (double *)nullptr);
    for (int nt = ctx->ltrot, s = 0; nt >= 1; nt--, s++) {
        if (lmeasure && in_window(ctx, nt)) measure(ctx);
        // K half (right_forward on UL with urt, left_backward on UR with urtm1), then the
        // rotation of the last family (nflag 2)
        // the on-site field first: its update, then right_forward on UL (bmat) and
        // left_backward on UR (bmat_inv) (dqmc_proj_sweep_b0.f90:72-77)
        const size_t sb = (size_t)s * ctx->slice_draws;
        const size_t ub = ctx->lwrapu ? (size_t)2 * n : 0;
        POps o = ops_begin(nt, 3, 1, 0);
        op_kin(o, false);
        if (ctx->lwrapu) {
            run_ops(ctx, o);
            if (lupdate) update_u(ctx, nt, sb);
            o = ops_begin(nt, 3, 1, 0);
            op_u(o, 1, 0);
        }
        if (ctx->lwrapv) op_v(o, nf - 1, 1, 1);
        else op_kin(o, true);
        run_ops(ctx, o);
        for (int f = nf - 1; f >= 0 && ctx->lwrapv; f--) {
            if (lupdate) update_family(ctx, nt, f, sb + ub + (size_t)2 * ctx->nn_lf * (nf - 1 - f));
            POps p = ops_begin(nt, 3, 1, 0);
            op_v(p, f, 2, 2);                       // nflag 1 of this family
            if (f > 0) op_v(p, f - 1, 1, 1);        // nflag 2 of the next one
            else op_kin(p, true);
            run_ops(ctx, p);
        }
        if ((ctx->iwrap_nt[nt - 1] > 0 || ctx->nwrap == 1) || nt == 1) {
            const int pt = ctx->iwrap_nt[nt - 1];
            save_old(ctx);
            pcopy(ctx, P_UR, 0, P_UST, pt, n * ne);
            {
                Tic t(ctx, T_QR);
                ptranspose(ctx, P_T2, 0, P_UL, 0, ne, n);
            }
            qr_to_t1(ctx, P_T2, 0);
            ptranspose(ctx, P_UL, 0, P_T1, 0, n, ne);
            pcopy(ctx, P_UST, pt, P_T1, 0, n * ne);
            scratch(ctx, 2, pt);
        }
    }
    return finish_pass(ctx, "sweep_b0");
}

int shp_sweep_0b(shp_ctx *ctx, int lupdate, int lmeasure, int ldyn)
{
    if (!ctx) return SH_E_ARG;
    if (ldyn) return fail(ctx, SH_E_ARG, "time-displaced measurement is not provided");
    if (!ctx->started) return fail(ctx, SH_E_STATE, "sweep_0b before sweep_start_0b");
    const int n = ctx->n, ne = ctx->ne, nf = ctx->nn_nf;
    pcopy(ctx, P_UR, 0, P_PROJ, 0, n * ne);
    pcopy(ctx, P_UST, ctx->nst, P_PROJ, 0, n * ne, 1);
    weights(ctx, 3, 0, 0);
    if (lupdate) hipLaunchKernelGGL(kp_draw, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, 1, 0, 0, (double *)nullptr);
    for (int nt = 1; nt <= ctx->ltrot; nt++) {
        // K half (left_forward on UR with urt, right_backward on UL with urtm1), then nflag 2
        // of the first family
        // the bond families, then the on-site field: left_forward on UR (bmat), right_backward
        // on UL (bmat_inv), its update (dqmc_proj_sweep_0b.f90:70-86), then the K half
        const size_t sb = (size_t)(nt - 1) * ctx->slice_draws;
        const size_t vb = ctx->lwrapv ? (size_t)2 * ctx->nn_lf * nf : 0;
        POps o = ops_begin(nt, 3, 0, 1);
        op_kin(o, false);
        if (ctx->lwrapv) op_v(o, 0, 0, 0);
        else if (ctx->lwrapu) op_u(o, 0, 1);
        else op_kin(o, true);
        run_ops(ctx, o);
        for (int f = 0; f < nf && ctx->lwrapv; f++) {
            if (lupdate) update_family(ctx, nt, f, sb + (size_t)2 * ctx->nn_lf * f);
            POps p = ops_begin(nt, 3, 0, 1);
            op_v(p, f, 1, 1);                       // nflag 1 of this family
            if (f + 1 < nf) op_v(p, f + 1, 0, 0);   // nflag 2 of the next one
            else if (ctx->lwrapu) op_u(p, 0, 1);
            else op_kin(p, true);
            run_ops(ctx, p);
        }
        if (ctx->lwrapu) {
            if (lupdate) update_u(ctx, nt, sb + vb);
            POps p = ops_begin(nt, 3, 0, 1);
            op_kin(p, true);
            run_ops(ctx, p);
        }
        if (lmeasure && in_window(ctx, nt)) measure(ctx);
        if (ctx->iwrap_nt[nt] > 0) {
            const int pt = ctx->iwrap_nt[nt];
            save_old(ctx);
            ptranspose(ctx, P_UL, 0, P_UST, pt, n, ne);
            qr_to_t1(ctx, P_UR, 0);
            pcopy(ctx, P_UR, 0, P_T1, 0, n * ne);
            scratch(ctx, 1, pt);
            pcopy(ctx, P_UST, pt, P_UR, 0, n * ne);
        }
    }
    return finish_pass(ctx, "sweep_0b");
}

int shp_get_g(shp_ctx *ctx, int chain, double *g)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !g) return SH_E_ARG;
    inverse(ctx, P_UL, P_UR, P_INV2, -1);
    green(ctx, P_UL, P_UR, P_INV2, P_G2);
    int rc = check_launch_p(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(g, ctx->hc[chain].G2, (size_t)ctx->n * ctx->n * sizeof(cplx), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_get_state(shp_ctx *ctx, int chain, double *ul, double *ur, double *ulrinv)
{
    if (!ctx || chain < 0 || chain >= ctx->nch) return SH_E_ARG;
    const PChain &c = ctx->hc[chain];
    const size_t nne = (size_t)ctx->n * ctx->ne * sizeof(cplx);
    if (ul) HIPCHK(hipMemcpyAsync(ul, c.UL, nne, hipMemcpyDeviceToHost, ctx->st));
    if (ur) HIPCHK(hipMemcpyAsync(ur, c.UR, nne, hipMemcpyDeviceToHost, ctx->st));
    if (ulrinv) HIPCHK(hipMemcpyAsync(ulrinv, c.ULRINV, (size_t)ctx->ne * ctx->ne * sizeof(cplx), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_get_diag(shp_ctx *ctx, int chain, double *d)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !d) return SH_E_ARG;
    double sc[SC_SIZE];
    HIPCHK(hipMemcpyAsync(sc, ctx->hc[chain].sc, sizeof(sc), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    d[0] = sc[SC_MAXWRAP]; d[1] = sc[SC_MAXPH];
    d[2] = sc[SC_PHASE]; d[3] = sc[SC_PHASE + 1];
    d[4] = sc[SC_LOGW]; d[5] = sc[SC_LOGW + 1];
    d[6] = sc[SC_ACC]; d[7] = sc[SC_TRIED]; d[8] = sc[SC_NCHK]; d[9] = sc[SC_SUMLOG];
    if (sc[SC_NONFINITE] != 0.0) return fail(ctx, SH_E_NONFINITE, "non-finite weight or Green's function");
    return 0;
}

int shp_reset_diag(shp_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    for (auto &c : ctx->hc) HIPCHK(hipMemsetAsync(c.sc + SC_MAXWRAP, 0, 6 * sizeof(double), ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_get_obs(shp_ctx *ctx, int chain, double *o)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !o) return SH_E_ARG;
    double sc[SC_SIZE];
    HIPCHK(hipMemcpyAsync(sc, ctx->hc[chain].sc, sizeof(sc), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    o[0] = sc[SC_NOBS];
    for (int k = 0; k < 10; k++) o[1 + k] = sc[SC_OBS + k];
    return 0;
}

int shp_set_eqt(shp_ctx *ctx, int chain, int l1, int l2, int l3, const int32_t *sites, const int32_t *cells,
                int zplq, const int32_t *plq, double theta)
{
    if (!ctx || chain < 0 || chain >= ctx->nch) return SH_E_ARG;
    DChain &v = ctx->hview[chain];
    size_t lds = 0;
    int rc = eqt_setup(ctx, v, ctx->hview, l1, l2, l3, sites, cells, zplq, plq, theta, lds);
    if (rc) return rc;
    ctx->eqt_lds = lds;
    HIPCHK(hipMemcpy(ctx->dview + chain, &v, sizeof(DChain), hipMemcpyHostToDevice));
    return 0;
}

int shp_get_eqt(shp_ctx *ctx, int chain, double *bins)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !bins) return SH_E_ARG;
    const DChain &v = ctx->hview[chain];
    if (!v.eqb) return fail(ctx, SH_E_STATE, "correlation bins not set up (shp_set_eqt)");
    HIPCHK(hipMemcpyAsync(bins, v.eqb, sizeof(cplx) * (20 * (size_t)v.lq + 8), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_reset_obs(shp_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    for (auto &c : ctx->hc) HIPCHK(hipMemsetAsync(c.sc + SC_NOBS, 0, 11 * sizeof(double), ctx->st));
    for (auto &v : ctx->hview)
        if (v.eqb) HIPCHK(hipMemsetAsync(v.eqb, 0, sizeof(cplx) * (20 * (size_t)v.lq + 8), ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shp_set_trace(shp_ctx *ctx, int enable, int cap)
{
    if (!ctx || cap < 0) return SH_E_ARG;
    for (cplx *p : ctx->trace_bufs) if (p) HIPCHK(hipFree(p));
    ctx->trace_bufs.assign(ctx->nch, nullptr);
    ctx->tracing = enable && cap > 0;
    for (int ci = 0; ci < ctx->nch; ci++) {
        PChain &c = ctx->hc[ci];
        c.trace = nullptr;
        c.trace_cap = 0;
        if (ctx->tracing) {
            void *p = nullptr;
            HIPCHK(hipMalloc(&p, (size_t)cap * ctx->n * ctx->n * sizeof(cplx)));
            ctx->trace_bufs[ci] = (cplx *)p;
            c.trace = (cplx *)p;
            c.trace_cap = cap;
        }
        double zero = 0.0;
        HIPCHK(hipMemcpy(c.sc + SC_TRACE_N, &zero, sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(ctx->dc, ctx->hc.data(), ctx->nch * sizeof(PChain), hipMemcpyHostToDevice));
    return 0;
}

int shp_get_trace(shp_ctx *ctx, int chain, double *buf, int32_t *count)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !count) return SH_E_ARG;
    const PChain &c = ctx->hc[chain];
    double tn = 0.0;
    HIPCHK(hipMemcpy(&tn, c.sc + SC_TRACE_N, sizeof(double), hipMemcpyDeviceToHost));
    *count = (int32_t)tn;
    if (buf && c.trace && *count > 0)
        HIPCHK(hipMemcpy(buf, c.trace, (size_t)(*count) * ctx->n * ctx->n * sizeof(cplx), hipMemcpyDeviceToHost));
    return 0;
}

int shp_timers(shp_ctx *ctx, int enable, double *t)
{
    if (!ctx) return SH_E_ARG;
    timers_collect(ctx);
    if (t)
        for (int k = 0; k < T_NCLS; k++) { t[k] = ctx->tms[k]; t[T_NCLS + k] = ctx->tcnt[k]; }
    for (int k = 0; k < T_NCLS; k++) { ctx->tms[k] = 0; ctx->tcnt[k] = 0; }
    ctx->timing = enable != 0;
    return 0;
}

int shp_synchronize(shp_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return check_launch_p(ctx);
}

}  // extern "C"

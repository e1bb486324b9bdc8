// learn.hip -- host side of the learned block proposals: the C ABI of include/sunhip_learn.h.
// Included at the end of sunhip.hip; the kernels are in learn_kernels.hip.

struct shl_net : ctx_base {
    void *na_scratch = nullptr;   // this device's no-allocation scratch (shl_loss_grad_noalloc)
    int cap = 0, count = 0;
    float *w = nullptr, *am = nullptr, *av = nullptr, *gpart = nullptr, *gout = nullptr;
    double *samp = nullptr, *lpart = nullptr, *dloss = nullptr, *pred = nullptr;
    int32_t *perm = nullptr;
    int pred_cap = 0;
    double norm[10] = {0};
    bool norm_set = false;
    long adam_t = 0;
    uint64_t rng = 0;
};

// ---- scratch of the no-allocation loss/gradient entry point ---------------------------------
// The partial sums of shl_loss_grad_noalloc: SHL_TRAIN_WG gradient vectors (float) and their
// loss partials (double), 4.4 MB in one file-scope device array that every GPU the library
// runs on holds until the process ends.  Host code never passes the array's name to a kernel:
// its device address is taken once per device by hipGetSymbolAddress (in shl_create, with the
// device current) and kept in noalloc_ptr[device].  Calls that overlap on one GPU share the
// array and race.
#define SHL_NA_GFLOATS ((size_t)SHL_TRAIN_WG * SHL_NWT)
#define SHL_NA_DOUBLES ((SHL_NA_GFLOATS + 1) / 2 + SHL_TRAIN_WG)
#define SHL_NA_MAX_SAMPLES (1 << 24)
static __device__ __align__(256) double shl_noalloc_scratch[SHL_NA_DOUBLES];
static void *noalloc_ptr[64];
static std::mutex noalloc_mu;

static int noalloc_setup(int device, void **out)
{
    if (device < 0 || device >= 64) return SH_E_ARG;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(shl_noalloc_scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return SH_E_HIP;
        }
        noalloc_ptr[device] = p;
    }
    *out = noalloc_ptr[device];
    return 0;
}

static size_t lds_block_move(int nmax, int z)
{
    return (size_t)SHL_MB * SHL_MB * sizeof(cplx) + 3 * SHL_MB * sizeof(cplx) + (SHL_H + 8) * sizeof(double) +
           7 * SHL_H * sizeof(float) + (4 * SHL_MB + 16 * 12) * sizeof(int) + (size_t)(3 + z) * nmax * sizeof(int);
}
static size_t lds_mlp_batch() { return (size_t)(SHL_H * SHL_H + 7 * SHL_H + 8) * sizeof(float) + 2 * SHL_H * sizeof(float); }

static NormF normf_of(const double *nm)
{
    NormF f;
    for (int k = 0; k < 4; k++) { f.xmu[k] = nm[k]; f.xsd[k] = nm[4 + k]; }
    f.ymu = nm[8];
    f.ysd = nm[9];
    return f;
}

// device buffers of the proposals of a context
static int prop_alloc(sh_ctx *ctx)
{
    if (ctx->d_prop) return 0;
    ctx->hprop.assign(ctx->nch, PropChain{});
    for (int ci = 0; ci < ctx->nch; ci++) {
        const DChain &c = ctx->hc[ci];
        PropChain &p = ctx->hprop[ci];
        p.psum = dalloc<int32_t>(ctx, (size_t)4 * c.ltrot);
        p.rng = dalloc<uint64_t>(ctx, 1);
        p.pa = dalloc<cplx>(ctx, (size_t)c.n * SHL_MB);
        p.pb = dalloc<cplx>(ctx, (size_t)c.n * SHL_MB);
        p.stats = dalloc<double>(ctx, 8);
        p.mapply = dalloc<int32_t>(ctx, 1);
        if (!p.psum || !p.rng || !p.pa || !p.pb || !p.stats || !p.mapply)
            return fail(ctx, SH_E_HIP, "device allocation failed");
        HIPCHK(hipMemset(p.mapply, 0, sizeof(int32_t)));
        uint64_t s = smix_seed(ctx->seeds[ci]);
        HIPCHK(hipMemcpy(p.rng, &s, sizeof(s), hipMemcpyHostToDevice));
    }
    ctx->pw = dalloc<float>(ctx, SHL_NWT);
    ctx->d_feat = dalloc<double>(ctx, (size_t)5 * ctx->nch);
    if (!ctx->pw || !ctx->d_feat) return fail(ctx, SH_E_HIP, "device allocation failed");
    ctx->d_prop = dupload<PropChain>(ctx, ctx->hprop.data(), ctx->nch);
    if (!ctx->d_prop) return fail(ctx, SH_E_HIP, "device allocation failed");
    return 0;
}

static void prop_pass_start(sh_ctx *ctx)
{
    if (!ctx->prop_on) return;
    Tic t(ctx, T_OTHER);
    hipLaunchKernelGGL(k_prop_psum, dim3(ctx->ltrot, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, ctx->d_prop);
}

static void block_move(sh_ctx *ctx, int nt)
{
    if (!ctx->prop_on) return;
    Tic t(ctx, T_PROPOSAL);
    hipLaunchKernelGGL(k_block_move, dim3(ctx->nch), dim3(256), lds_block_move(ctx->nmax, ctx->prop_z), ctx->st, ctx->dc,
                       ctx->d_prop, ctx->pnet, nt);
    const int ntl = (ctx->nmax + GT - 1) / GT;
    hipLaunchKernelGGL(k_block_apply, dim3(ntl * ntl, 1, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, ctx->d_prop);
}

extern "C" {

int sh_set_neighbours(sh_ctx *ctx, int z, const int32_t *nbr)
{
    if (ctx) ctx->graph_epoch++;
    if (!ctx || z < 1 || z > 16 || !nbr) return SH_E_ARG;
    size_t tot = 0;
    for (auto &c : ctx->hc) tot += (size_t)c.n * z;
    size_t off = 0;
    for (auto &c : ctx->hc) {
        // symmetric lists: j appears in list(i) as often as i in list(j)
        std::vector<int> cnt;
        for (int i = 0; i < c.n; i++)
            for (int k = 0; k < z; k++) {
                const int j = nbr[off + (size_t)i * z + k];
                if (j < 0 || j >= c.n || j == i) return fail(ctx, SH_E_ARG, "neighbour index out of range or self");
            }
        for (int i = 0; i < c.n; i++)
            for (int k = 0; k < z; k++) {
                const int j = nbr[off + (size_t)i * z + k];
                int a = 0, b = 0;
                for (int q = 0; q < z; q++) {
                    a += nbr[off + (size_t)i * z + q] == j;
                    b += nbr[off + (size_t)j * z + q] == i;
                }
                if (a != b) return fail(ctx, SH_E_ARG, "neighbour lists are not symmetric");
            }
        off += (size_t)c.n * z;
    }
    int rc = prop_alloc(ctx);
    if (rc) return rc;
    if (!ctx->d_nbr || ctx->prop_z != z) {
        ctx->d_nbr = dalloc<int32_t>(ctx, tot);
        if (!ctx->d_nbr) return fail(ctx, SH_E_HIP, "device allocation failed");
    }
    HIPCHK(hipMemcpy(ctx->d_nbr, nbr, tot * sizeof(int32_t), hipMemcpyHostToDevice));
    ctx->prop_z = z;
    ctx->prop_on = false;      // the proposal's LDS size depends on z: sh_set_proposal again
    ctx->prop_local = true;
    off = 0;
    for (int ci = 0; ci < ctx->nch; ci++) {
        PropChain &p = ctx->hprop[ci];
        p.nbr = ctx->d_nbr + off;
        p.z = z;
        p.nb2 = ctx->hc[ci].n * z;
        off += (size_t)ctx->hc[ci].n * z;
    }
    HIPCHK(hipMemcpy(ctx->d_prop, ctx->hprop.data(), sizeof(PropChain) * ctx->nch, hipMemcpyHostToDevice));
    return 0;
}

int sh_features(sh_ctx *ctx, double *out)
{
    if (!ctx || !out) return SH_E_ARG;
    if (!ctx->prop_z) return fail(ctx, SH_E_STATE, "sh_features before sh_set_neighbours");
    hipLaunchKernelGGL(k_prop_psum, dim3(ctx->ltrot, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, ctx->d_prop);
    hipLaunchKernelGGL(k_prop_features, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, ctx->d_prop, ctx->d_feat);
    int rc = check_launch(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, ctx->d_feat, sizeof(double) * 5 * ctx->nch, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int sh_set_proposal(sh_ctx *ctx, const shl_net *net, int nprop, int local_updates)
{
    if (ctx) ctx->graph_epoch++;
    if (!ctx) return SH_E_ARG;
    if (!net) {
        ctx->prop_on = false;
        ctx->prop_local = true;
        return 0;
    }
    if (nprop < 1 || nprop > SHL_MB) return fail(ctx, SH_E_ARG, "nprop must be 1..64");
    if (!ctx->prop_z) return fail(ctx, SH_E_STATE, "sh_set_proposal before sh_set_neighbours");
    if (!net->norm_set) return fail(ctx, SH_E_STATE, "the network has no normalisation (train it or set it)");
    for (auto &c : ctx->hc)
        if (c.lcomp != 2 || c.ltrot < 2 || !c.lwrapu)
            return fail(ctx, SH_E_ARG, "block proposals need lcomp = 2, ltrot >= 2 and an interaction term");
    if (net->dev != ctx->dev) return fail(ctx, SH_E_ARG, "network and context are on different devices");
    const size_t lds = lds_block_move(ctx->nmax, ctx->prop_z);
    if (lds > 150 * 1024) return fail(ctx, SH_E_ARG, "block proposals: ndim * (3 + z) too large for LDS");
    if (lds > 64 * 1024 && hipFuncSetAttribute((const void *)k_block_move, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, SH_E_HIP, "dynamic LDS request above the device limit");
    }
    // the weights as they are now (the network's stream is idle after every shl_ call)
    HIPCHK(hipMemcpy(ctx->pw, net->w, sizeof(float) * SHL_NWT, hipMemcpyDeviceToDevice));
    PropNet &pn = ctx->pnet;
    for (int k = 0; k < 4; k++) { pn.xmu[k] = net->norm[k]; pn.xsd[k] = net->norm[4 + k]; }
    pn.ymu = net->norm[8];
    pn.ysd = net->norm[9];
    pn.w = ctx->pw;
    pn.nprop = nprop;
    ctx->prop_on = true;
    ctx->prop_local = local_updates != 0;
    return 0;
}

int sh_get_proposal_stats(sh_ctx *ctx, int chain, double *stats)
{
    if (!ctx || chain < 0 || chain >= ctx->nch || !stats) return SH_E_ARG;
    if (!ctx->d_prop) {
        for (int k = 0; k < 5; k++) stats[k] = 0.0;
        return 0;
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(hipMemcpy(stats, ctx->hprop[chain].stats, 5 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int sh_reset_proposal_stats(sh_ctx *ctx)
{
    if (!ctx) return SH_E_ARG;
    if (!ctx->d_prop) return 0;
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (auto &p : ctx->hprop) HIPCHK(hipMemset(p.stats, 0, 8 * sizeof(double)));
    return 0;
}

// ---- the network -----------------------------------------------------------------------------

int shl_create(shl_net **out, int capacity, int32_t seed, int device)
{
    if (!out || capacity < 1 || capacity > (1 << 26)) return SH_E_ARG;
    *out = nullptr;
    shl_net *ctx = new shl_net();
    ctx->dev = device;
    ctx->cap = capacity;
    ctx->rng = smix_seed(seed);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->st) != hipSuccess) {
        delete ctx;
        return SH_E_HIP;
    }
    if (hipFuncSetAttribute((const void *)k_mlp_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_mlp_batch()) != hipSuccess) {
        (void)hipGetLastError();
        shl_destroy(ctx);
        return SH_E_HIP;
    }
    // the device address of this GPU's no-allocation scratch, with the device current
    if (noalloc_setup(device, &ctx->na_scratch) != 0) {
        int rc = fail(ctx, SH_E_HIP, "hipGetSymbolAddress of the no-allocation scratch failed");
        shl_destroy(ctx);
        return rc;
    }
    ctx->w = dalloc<float>(ctx, SHL_NWT);
    ctx->am = dalloc<float>(ctx, SHL_NWT);
    ctx->av = dalloc<float>(ctx, SHL_NWT);
    ctx->gout = dalloc<float>(ctx, SHL_NWT);
    ctx->gpart = dalloc<float>(ctx, (size_t)SHL_TRAIN_WG * SHL_NWT);
    ctx->lpart = dalloc<double>(ctx, SHL_TRAIN_WG);
    ctx->dloss = dalloc<double>(ctx, 1);
    ctx->samp = dalloc<double>(ctx, (size_t)5 * capacity);
    ctx->perm = dalloc<int32_t>(ctx, capacity);
    if (!ctx->w || !ctx->am || !ctx->av || !ctx->gout || !ctx->gpart || !ctx->lpart || !ctx->dloss || !ctx->samp ||
        !ctx->perm) {
        shl_destroy(ctx);
        return SH_E_HIP;
    }
    // Glorot-uniform weights, zero biases
    std::vector<float> w(SHL_NWT, 0.0f);
    uint64_t s = ctx->rng ^ 0xD1B54A32D192ED03ull;
    auto fill = [&](int off, int cnt, int fin, int fout) {
        const double a = sqrt(6.0 / (fin + fout));
        for (int k = 0; k < cnt; k++) w[off + k] = (float)((2.0 * smix_unif(s) - 1.0) * a);
    };
    fill(WO_W1, 4 * SHL_H, 4, SHL_H);
    fill(WO_W2, SHL_H * SHL_H, SHL_H, SHL_H);
    fill(WO_W3, SHL_H, SHL_H, 1);
    if (hipMemcpy(ctx->w, w.data(), sizeof(float) * SHL_NWT, hipMemcpyHostToDevice) != hipSuccess) {
        shl_destroy(ctx);
        return SH_E_HIP;
    }
    *out = ctx;
    return 0;
}

int shl_destroy(shl_net *ctx)
{
    if (!ctx) return 0;
    if (ctx->st) hipStreamSynchronize(ctx->st);
    for (void *p : ctx->allocs) hipFree(p);
    if (ctx->st) hipStreamDestroy(ctx->st);
    delete ctx;
    return 0;
}

const char *shl_last_error(const shl_net *ctx) { return ctx ? ctx->err.c_str() : "null network"; }

int shl_get_weights(shl_net *ctx, float *w)
{
    if (!ctx || !w) return SH_E_ARG;
    HIPCHK(hipMemcpy(w, ctx->w, sizeof(float) * SHL_NWT, hipMemcpyDeviceToHost));
    return 0;
}

int shl_set_weights(shl_net *ctx, const float *w)
{
    if (!ctx || !w) return SH_E_ARG;
    for (int k = 0; k < SHL_NWT; k++)
        if (!std::isfinite(w[k])) return fail(ctx, SH_E_ARG, "non-finite weight");
    HIPCHK(hipMemcpy(ctx->w, w, sizeof(float) * SHL_NWT, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ctx->am, 0, sizeof(float) * SHL_NWT));
    HIPCHK(hipMemset(ctx->av, 0, sizeof(float) * SHL_NWT));
    ctx->adam_t = 0;
    return 0;
}

int shl_set_norm(shl_net *ctx, const double *norm)
{
    if (!ctx || !norm) return SH_E_ARG;
    for (int k = 0; k < 10; k++)
        if (!std::isfinite(norm[k])) return fail(ctx, SH_E_ARG, "non-finite normalisation");
    for (int k = 4; k < 8; k++)
        if (!(norm[k] > 0.0)) return fail(ctx, SH_E_ARG, "feature scales must be positive");
    if (!(norm[9] > 0.0)) return fail(ctx, SH_E_ARG, "energy scale must be positive");
    for (int k = 0; k < 10; k++) ctx->norm[k] = norm[k];
    ctx->norm_set = true;
    return 0;
}

int shl_get_norm(shl_net *ctx, double *norm)
{
    if (!ctx || !norm) return SH_E_ARG;
    if (!ctx->norm_set) return fail(ctx, SH_E_STATE, "normalisation not set");
    for (int k = 0; k < 10; k++) norm[k] = ctx->norm[k];
    return 0;
}

int shl_add_samples(shl_net *net, sh_ctx *ctx)
{
    if (!net || !ctx) return SH_E_ARG;
    if (!ctx->prop_z) return fail(net, SH_E_STATE, "shl_add_samples before sh_set_neighbours");
    if (net->dev != ctx->dev) return fail(net, SH_E_ARG, "network and context are on different devices");
    if (net->count + ctx->nch > net->cap) return fail(net, SH_E_ARG, "sample capacity exceeded");
    hipLaunchKernelGGL(k_prop_psum, dim3(ctx->ltrot, ctx->nch), dim3(256), 0, ctx->st, ctx->dc, ctx->d_prop);
    hipLaunchKernelGGL(k_prop_features, dim3(ctx->nch), dim3(64), 0, ctx->st, ctx->dc, ctx->d_prop,
                       net->samp + (size_t)5 * net->count);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->st);
    if (e != hipSuccess) return fail(net, SH_E_HIP, std::string("shl_add_samples: ") + hipGetErrorString(e));
    net->count += ctx->nch;
    return 0;
}

int shl_push_samples(shl_net *ctx, const double *x, int n)
{
    if (!ctx || (!x && n > 0) || n < 0) return SH_E_ARG;
    if (ctx->count + n > ctx->cap) return fail(ctx, SH_E_ARG, "sample capacity exceeded");
    for (size_t k = 0; k < (size_t)5 * n; k++)
        if (!std::isfinite(x[k])) return fail(ctx, SH_E_ARG, "non-finite sample");
    HIPCHK(hipMemcpy(ctx->samp + (size_t)5 * ctx->count, x, sizeof(double) * 5 * n, hipMemcpyHostToDevice));
    ctx->count += n;
    return 0;
}

int shl_num_samples(const shl_net *ctx) { return ctx ? ctx->count : SH_E_ARG; }

int shl_get_samples(shl_net *ctx, double *x, int first, int n)
{
    if (!ctx || !x || first < 0 || n < 0 || first + n > ctx->count) return SH_E_ARG;
    HIPCHK(hipMemcpy(x, ctx->samp + (size_t)5 * first, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
    return 0;
}

int shl_clear_samples(shl_net *ctx)
{
    if (!ctx) return SH_E_ARG;
    ctx->count = 0;
    return 0;
}

// one launch pair over `count` sample indices (idx may be null: samples 0..count-1)
static int mlp_step(shl_net *ctx, const int32_t *idx, int count, float lr, float *gout)
{
    const int nwg = std::min(SHL_TRAIN_WG, std::max(1, (count + 3) / 4));
    const int per = (count + nwg - 1) / nwg;
    NormF nm = normf_of(ctx->norm);
    hipLaunchKernelGGL(k_mlp_batch, dim3(nwg), dim3(256), lds_mlp_batch(), ctx->st, ctx->w, ctx->samp, idx, count, per,
                       nm, 1.0f / (float)count, 0, ctx->gpart, ctx->lpart, (double *)nullptr);
    float c1 = 1.0f, c2 = 1.0f;
    if (lr > 0.f) {
        ctx->adam_t++;
        c1 = (float)(1.0 / (1.0 - pow(0.9, (double)ctx->adam_t)));
        c2 = (float)(1.0 / (1.0 - pow(0.999, (double)ctx->adam_t)));
    }
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((SHL_NWT + 255) / 256), dim3(256), 0, ctx->st, ctx->gpart, nwg,
                       ctx->lpart, ctx->w, ctx->am, ctx->av, lr, c1, c2, gout, ctx->dloss, 1.0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, SH_E_HIP, std::string("training launch: ") + hipGetErrorString(e));
    return 0;
}

static int norm_from_samples(shl_net *ctx)
{
    std::vector<double> x((size_t)5 * ctx->count);
    HIPCHK(hipMemcpy(x.data(), ctx->samp, sizeof(double) * x.size(), hipMemcpyDeviceToHost));
    double nm[10];
    for (int k = 0; k < 5; k++) {
        double s = 0.0, s2 = 0.0;
        for (int q = 0; q < ctx->count; q++) s += x[(size_t)q * 5 + k];
        const double mu = s / ctx->count;
        for (int q = 0; q < ctx->count; q++) s2 += (x[(size_t)q * 5 + k] - mu) * (x[(size_t)q * 5 + k] - mu);
        double sd = sqrt(s2 / ctx->count);
        if (!(sd > 1e-12)) sd = 1.0;
        if (k < 4) { nm[k] = mu; nm[4 + k] = sd; } else { nm[8] = mu; nm[9] = sd; }
    }
    return shl_set_norm(ctx, nm);
}

int shl_train(shl_net *ctx, int epochs, int batch, double lr, double *loss)
{
    if (!ctx || epochs < 0 || batch < 1 || !(lr > 0.0)) return SH_E_ARG;
    if (ctx->count < 1) return fail(ctx, SH_E_STATE, "no samples");
    if (!ctx->norm_set) {
        int rc = norm_from_samples(ctx);
        if (rc) return rc;
    }
    std::vector<int32_t> perm(ctx->count);
    for (int ep = 0; ep < epochs; ep++) {
        for (int k = 0; k < ctx->count; k++) perm[k] = k;
        for (int k = ctx->count - 1; k > 0; k--) {
            const int q = (int)(smix_unif(ctx->rng) * (k + 1));
            std::swap(perm[k], perm[std::min(q, k)]);
        }
        HIPCHK(hipMemcpyAsync(ctx->perm, perm.data(), sizeof(int32_t) * ctx->count, hipMemcpyHostToDevice, ctx->st));
        double tot = 0.0;
        for (int b0 = 0; b0 < ctx->count; b0 += batch) {
            const int nb = std::min(batch, ctx->count - b0);
            int rc = mlp_step(ctx, ctx->perm + b0, nb, (float)lr, nullptr);
            if (rc) return rc;
            double l = 0.0;
            HIPCHK(hipMemcpyAsync(&l, ctx->dloss, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
            HIPCHK(hipStreamSynchronize(ctx->st));
            tot += l;
        }
        if (loss) loss[ep] = tot / ctx->count;
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
}

int shl_loss_grad(shl_net *ctx, const double *x, int n, double *loss, float *grad)
{
    if (!ctx || !x || n < 1 || !loss || !grad) return SH_E_ARG;
    if (!ctx->norm_set) return fail(ctx, SH_E_STATE, "normalisation not set");
    if (n > ctx->cap) return fail(ctx, SH_E_ARG, "more samples than the capacity");
    // the host samples go to the end of the sample buffer's free space or replace it
    // temporarily: evaluate them from a scratch copy
    double *d = dupload<double>(ctx, x, (size_t)5 * n);
    if (!d) return fail(ctx, SH_E_HIP, "device allocation failed");
    double *keep = ctx->samp;
    ctx->samp = d;
    int rc = mlp_step(ctx, nullptr, n, 0.0f, ctx->gout);
    ctx->samp = keep;
    if (rc == 0) {
        double l = 0.0;
        HIPCHK(hipMemcpyAsync(&l, ctx->dloss, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipMemcpyAsync(grad, ctx->gout, sizeof(float) * SHL_NWT, hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        *loss = l / n;
    }
    hipStreamSynchronize(ctx->st);
    hipFree(d);
    ctx->allocs.pop_back();
    return rc;
}

int shl_loss_grad_noalloc(shl_net *ctx, const double *x, int n, double *loss, float *grad)
{
    if (!ctx || !x || !loss || !grad || n < 1) return SH_E_ARG;
    if (n > SHL_NA_MAX_SAMPLES)
        return fail(ctx, SH_E_ARG, "shl_loss_grad_noalloc: at most 16777216 samples per call");
    if (!ctx->norm_set) return fail(ctx, SH_E_STATE, "normalisation not set");
    if (!ctx->na_scratch) return fail(ctx, SH_E_STATE, "no-allocation scratch not set up");
    // the partitions and kernels of shl_loss_grad, with the partial sums in this device's
    // scratch: gradient partials first, the loss partials after them (8-byte aligned)
    float *gpart = (float *)ctx->na_scratch;
    double *lpart = (double *)ctx->na_scratch + (SHL_NA_GFLOATS + 1) / 2;
    const int nwg = std::min(SHL_TRAIN_WG, std::max(1, (n + 3) / 4));
    const int per = (n + nwg - 1) / nwg;
    hipLaunchKernelGGL(k_mlp_batch, dim3(nwg), dim3(256), lds_mlp_batch(), ctx->st, ctx->w, x, (const int32_t *)nullptr,
                       n, per, normf_of(ctx->norm), 1.0f / (float)n, 0, gpart, lpart, (double *)nullptr);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((SHL_NWT + 255) / 256), dim3(256), 0, ctx->st, gpart, nwg, lpart,
                       ctx->w, ctx->am, ctx->av, 0.0f, 1.0f, 1.0f, grad, loss, (double)n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->st);
    if (e != hipSuccess) return fail(ctx, SH_E_HIP, std::string("shl_loss_grad_noalloc: ") + hipGetErrorString(e));
    return 0;
}

int shl_predict(shl_net *ctx, const double *f, int n, double *e)
{
    if (!ctx || !f || n < 1 || !e) return SH_E_ARG;
    if (!ctx->norm_set) return fail(ctx, SH_E_STATE, "normalisation not set");
    std::vector<double> x((size_t)5 * n, 0.0);
    for (int q = 0; q < n; q++)
        for (int k = 0; k < 4; k++) x[(size_t)q * 5 + k] = f[(size_t)q * 4 + k];
    double *d = dupload<double>(ctx, x.data(), x.size());
    double *o = dalloc<double>(ctx, n);
    if (!d || !o) return fail(ctx, SH_E_HIP, "device allocation failed");
    const int nwg = std::min(SHL_TRAIN_WG, n);
    const int per = (n + nwg - 1) / nwg;
    hipLaunchKernelGGL(k_mlp_batch, dim3(nwg), dim3(256), lds_mlp_batch(), ctx->st, ctx->w, d, (const int32_t *)nullptr,
                       n, per, normf_of(ctx->norm), 1.0f, 1, ctx->gpart, ctx->lpart, o);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpyAsync(e, o, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->st);
    hipStreamSynchronize(ctx->st);
    hipFree(o);
    hipFree(d);
    ctx->allocs.pop_back();
    ctx->allocs.pop_back();
    if (err != hipSuccess) return fail(ctx, SH_E_HIP, std::string("shl_predict: ") + hipGetErrorString(err));
    return 0;
}

}  // extern "C"

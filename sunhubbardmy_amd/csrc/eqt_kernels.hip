// eqt_kernels.hip -- the correlation bins of equaltime_measure on the device
// (src/dqmc_measure/equaltime_measure.f90:44-68): zcpcm, zspsm, znn, pair_onsite per (imj, no_i,
// no_j) (:265-318, :484-499), zn (:301-318), the bond correlations zjj, zbb, pair_nn, pair_sn with
// zj, zb (:367-543), and the energy terms zqsq, ztsq, ztq (:143-263) and zint_v (:127-141).
// One workgroup per chain and measured slice, after k_measure.  Bins (complex, summed with
// zphi): [0, 16 lq) the four (imj, no_i, no_j) arrays at q 4 lq + imj + lq (no_i + 2 no_j),
// [16 lq, 20 lq) zjj, zbb, pair_nn, pair_sn, then zn(1), zn(2), zj, zb, zqsq, ztsq, ztq, zint_v.
#pragma once

#define EQT_T 16

// imj (0-based) of sites i and j from their cell coordinates (latt%imj)
__device__ __forceinline__ int eqt_imj(const int32_t *xy, int i, int j, int l1, int l2, int l3)
{
    int dx = xy[4 * i] - xy[4 * j], dy = xy[4 * i + 1] - xy[4 * j + 1], dz = xy[4 * i + 2] - xy[4 * j + 2];
    dx = dx < 1 ? dx + l1 : dx;
    dy = dy < 1 ? dy + l2 : dy;
    dz = dz < 1 ? dz + l3 : dz;
    return (dz - 1) * l1 * l2 + dy + (dx - 1) * l2 - 1;
}

__device__ __forceinline__ void lds_cadd(double *a, cplx v)
{
    atomicAdd(a, v.x);
    atomicAdd(a + 1, v.y);
}

// LDS: bins 16 lq complex, G(I, J) and G(J, I) tiles (16 x 17 each), 1 - G(i, i) for every site
extern "C" __global__ void __launch_bounds__(256)
k_eqt(const DChain *chains)
{
    extern __shared__ double eqsh[];
    __shared__ double red[16];
    const DChain &c = chains[blockIdx.x];
    if (!c.eqb) return;
    const int n = c.n, lq = c.lq, l1 = c.dl1, l2 = c.dl2, l3 = c.dl3, tid = threadIdx.x;
    double *acc = eqsh;                               // 2 * 16 lq
    cplx *ta = (cplx *)(eqsh + 32 * lq);              // [16][17]
    cplx *tb = ta + EQT_T * (EQT_T + 1);
    cplx *gd = tb + EQT_T * (EQT_T + 1);              // [n] 1 - G(i, i)
    GAS const cplx *G = gp(c.G);
    const int32_t *xy = c.sxy;
    const double nf = (double)c.nflr, fs = 1.0 - 1.0 / (nf * nf);
    for (int k = tid; k < 32 * lq; k += 256) acc[k] = 0.0;
    for (int i = tid; i < n; i += 256) {
        cplx g = gld(G + (size_t)i * n + i);
        gd[i] = make_double2(1.0 - g.x, -g.y);
    }
    // the (imj, no_i, no_j) bins over all pairs, 16 x 16 tiles of G and of its transpose
    const int r = tid % EQT_T, s = tid / EQT_T;
    for (int J = 0; J < n; J += EQT_T)
        for (int I = 0; I < n; I += EQT_T) {
            __syncthreads();
            const int i = I + r, j = J + s;
            ta[s * (EQT_T + 1) + r] = (i < n && j < n) ? gld(G + (size_t)j * n + i) : make_double2(0.0, 0.0);
            // tb[s][r] = G(J + r, I + s)
            tb[s * (EQT_T + 1) + r] = (J + r < n && I + s < n) ? gld(G + (size_t)(I + s) * n + J + r) : make_double2(0.0, 0.0);
            __syncthreads();
            if (i < n && j < n) {
                const cplx g = ta[s * (EQT_T + 1) + r];                  // G(i, j)
                const cplx gt = tb[r * (EQT_T + 1) + s];                 // G(j, i)
                const cplx gc = make_double2((i == j ? 1.0 : 0.0) - gt.x, -gt.y);
                const cplx gg = cmul(gc, g);
                const int k = eqt_imj(xy, i, j, l1, l2, l3) + lq * (xy[4 * i + 3] + 2 * xy[4 * j + 3]);
                lds_cadd(acc + 2 * k, gc);
                lds_cadd(acc + 2 * (4 * lq + k), cscale(gg, fs));
                lds_cadd(acc + 2 * (8 * lq + k), cadd(cmul(gd[i], gd[j]), cscale(gg, 1.0 / nf)));
                lds_cadd(acc + 2 * (12 * lq + k), cmul(g, g));
            }
        }
    __syncthreads();
    double *sc = c.sc;
    const cplx phase = make_double2(sc[SC_PHASE], sc[SC_PHASE + 1]);
    const double sg = phase.x < 0 ? -1.0 : 1.0;
    const cplx zphi = cscale(phase, sg / phase.x);
    GAS cplx *B = gp(c.eqb);
    for (int k = tid; k < 16 * lq; k += 256)
        gst(B + k, cadd(gld(B + k), cmul(zphi, make_double2(acc[2 * k], acc[2 * k + 1]))));
    __syncthreads();
    for (int k = tid; k < 8 * lq; k += 256) acc[k] = 0.0;
    __syncthreads();
    // bond correlations over cell pairs: per cell A site, nnlist(., 1), snlist(., 1), current partner
    const int32_t *cs = c.eqc;
#define GG(a, b) gld(G + (size_t)(b) * n + (a))
#define GCC(a, b) csub(make_double2((a) == (b) ? 1.0 : 0.0, 0.0), GG(b, a))
    for (int e = tid; e < lq * lq; e += 256) {
        const int u = e % lq, v = e / lq;
        const int i0 = cs[4 * u], in = cs[4 * u + 1], is = cs[4 * u + 2], ij = cs[4 * u + 3];
        const int j0 = cs[4 * v], jn = cs[4 * v + 1], js = cs[4 * v + 2], jj = cs[4 * v + 3];
        const int imj = eqt_imj(xy, i0, j0, l1, l2, l3);
        const cplx z1 = cmul(csub(GCC(i0, ij), GCC(ij, i0)), csub(GCC(j0, jj), GCC(jj, j0)));
        const cplx z3 = csub(cadd(cmul(GCC(i0, jj), GG(ij, j0)), cmul(GCC(ij, j0), GG(i0, jj))),
                             cadd(cmul(GCC(i0, j0), GG(ij, jj)), cmul(GCC(ij, jj), GG(i0, j0))));
        lds_cadd(acc + 2 * imj, csub(cscale(z1, -1.0), cscale(z3, 1.0 / nf)));
        const cplx bj = cadd(GCC(j0, jn), GCC(jn, j0)), bi = cadd(GCC(i0, in), GCC(in, i0));
        const cplx bc = cadd(cadd(cmul(GCC(i0, jn), GG(in, j0)), cmul(GCC(in, j0), GG(i0, jn))),
                             cadd(cmul(GCC(i0, j0), GG(in, jn)), cmul(GCC(in, jn), GG(i0, j0))));
        lds_cadd(acc + 2 * (lq + imj), cadd(cmul(bj, bi), cscale(bc, 1.0 / nf)));
        const cplx g00 = GG(i0, j0);
        lds_cadd(acc + 2 * (2 * lq + imj), cmul(g00, GG(in, jn)));
        lds_cadd(acc + 2 * (3 * lq + imj), cmul(g00, GG(is, js)));
    }
    // scalars: zn(1), zn(2), zj, zb, zqsq, ztsq, ztq, zint_v
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = 0.0;
    for (int i = tid; i < n; i += 256) {
        // (a register array indexed by the orbital would be lowered to private memory)
        if (xy[4 * i + 3]) {
            v[2] += gd[i].x;
            v[3] += gd[i].y;
        } else {
            v[0] += gd[i].x;
            v[1] += gd[i].y;
        }
    }
    for (int u = tid; u < lq; u += 256) {
        const int i0 = cs[4 * u], in = cs[4 * u + 1], ij = cs[4 * u + 3];
        const cplx dj = csub(GCC(i0, ij), GCC(ij, i0));          // zj += i dj
        v[4] -= dj.y;
        v[5] += dj.x;
        const cplx b = cadd(GCC(i0, in), GCC(in, i0));
        v[6] += b.x;
        v[7] += b.y;
    }
    const int zp = c.zplq;
    const cplx ep = make_double2(cos(c.theta), sin(c.theta));
    const cplx zodd = cconj(ep), zeven = make_double2(-ep.x, -ep.y);   // 1-based odd / even sites
    for (int u = tid; u < lq && zp; u += 256) {
        const int32_t *q = c.plq + (size_t)u * zp;
        cplx dq = make_double2(0.0, 0.0);
        double dt = 0.0;
        for (int s1 = 0; s1 < zp; s1++) {
            const int i1 = q[s1], i1a = q[(s1 + 1) % zp];
            const cplx z1 = (s1 & 1) ? zeven : zodd;
            dq = cadd(dq, gd[i1]);
            dt += 2.0 * cmul(z1, GCC(i1, i1a)).x;
        }
        cplx qsq = cscale(cmul(dq, dq), nf), tsq = make_double2(dt * dt * nf, 0.0), tq = cscale(dq, dt * nf / 3.0);
        for (int s1 = 0; s1 < zp; s1++) {
            const int i1 = q[s1], i1a = q[(s1 + 1) % zp];
            const cplx z1 = (s1 & 1) ? zeven : zodd, z2 = cconj(z1);
            for (int s2 = 0; s2 < zp; s2++) {
                const int i2 = q[s2], i2a = q[(s2 + 1) % zp];
                const cplx z3 = (s2 & 1) ? zeven : zodd, z4 = cconj(z3);
                qsq = cadd(qsq, cmul(GCC(i1, i2), GG(i1, i2)));
                tsq = cadd(tsq, cadd(cadd(cmul(cmul(z1, z3), cmul(GCC(i1, i2a), GG(i1a, i2))),
                                          cmul(cmul(z1, z4), cmul(GCC(i1, i2), GG(i1a, i2a)))),
                                     cadd(cmul(cmul(z2, z3), cmul(GCC(i1a, i2a), GG(i1, i2))),
                                          cmul(cmul(z2, z4), cmul(GCC(i1a, i2), GG(i1, i2a))))));
                tq = cadd(tq, cadd(cmul(cscale(z1, 1.0 / 3.0), cmul(GCC(i1, i2), GG(i1a, i2))),
                                   cmul(cscale(z2, 1.0 / 3.0), cmul(GCC(i1a, i2), GG(i1, i2)))));
            }
        }
        v[8] += qsq.x; v[9] += qsq.y; v[10] += tsq.x; v[11] += tsq.y; v[12] += tq.x; v[13] += tq.y;
    }
    for (int b = tid; b < c.nn_nf * c.nn_lf; b += 256) {
        const int i0 = c.bonds[2 * b], i1 = c.bonds[2 * b + 1];
        v[14] += GCC(i0, i1).x;
    }
#undef GG
#undef GCC
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = block_sum(v[q], red);
    __syncthreads();
    for (int k = tid; k < 4 * lq; k += 256)
        gst(B + 16 * lq + k, cadd(gld(B + 16 * lq + k), cmul(zphi, make_double2(acc[2 * k], acc[2 * k + 1]))));
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < 8; q++)
            gst(B + 20 * lq + q, cadd(gld(B + 20 * lq + q), cmul(zphi, make_double2(v[2 * q], v[2 * q + 1]))));
    }
}

// dyn_kernels.hip -- the time-displaced path (ltau) of the finite-temperature 0b pass:
// the element-wise steps of green_tau (src/dqmc_ft_core/green_tau.f90:1-122) around the
// library's ZGEMMs and LU solves, the propagated-versus-recomputed check (-DDYNERROR) and the
// znn_tau measurement (src/dqmc_measure/dyn_measure.f90:1-101).
#include "sunhip_kernels.h"

// Element-wise steps on one chain's N x N matrices; rpt is the right chain point, the left point
// is the saved one (Us, Vs, Ds).  op:
//   0  start of the pass: G(0,0) = G(t,0) = G, G(0,t) = G - 1
//   1  dst = 1
//   2  dst(i,j) = drmin_i VR(i,j)
//   3  dst = VL^H
//   4  W3(i,j) = W1(i,j) / (dlmax_i drmax_j) + W3(i,j) dlmin_i drmin_j       (the G(0,0) matrix)
//   5  dst(i,j) /= dlmax_i
//   6  dst(i,j) /= drmax_j
//   7  dst(i,j) = -dlmin_i conj(UL(j,i))
//   8  dst = src (G(t,0) <- green_tau's, G(0,t) likewise: dsel M_GT0 / M_G0T)
extern "C" __global__ void __launch_bounds__(256)
k_dyn_elem(const DChain *chains, int op, int dsel, int rpt)
{
    const DChain &c = chains[blockIdx.z];
    const int n = c.n;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n || !c.ltau) return;
    const int i = (int)(e % n), j = (int)(e / n);
    const size_t et = (size_t)j + (size_t)i * n;      // (j, i)
    GAS cplx *D = gp(chain_mat(c, dsel, 0));
    const double *dr = c.D + (size_t)n * rpt, *dl = c.Ds;
    switch (op) {
    case 0: {
        const cplx g = gld(gp(c.G) + e);
        gst(gp(c.g00) + e, g);
        gst(gp(c.gt0) + e, g);
        gst(gp(c.g0t) + e, make_double2(g.x - (i == j ? 1.0 : 0.0), g.y));
        break;
    }
    case 1: gst(D + e, make_double2(i == j ? 1.0 : 0.0, 0.0)); break;
    case 2: gst(D + e, cscale(gld(gp(c.V + (size_t)n * n * rpt) + e), dmin_of(dr[i]))); break;
    case 3: gst(D + e, cconj(gld(gp(c.Vs) + et))); break;
    case 4: {
        const cplx a = gld(gp(c.W1) + e), b = gld(gp(c.W3) + e);
        gst(gp(c.W3) + e, cadd(cscale(a, 1.0 / (dmax_of(dl[i]) * dmax_of(dr[j]))), cscale(b, dmin_of(dl[i]) * dmin_of(dr[j]))));
        break;
    }
    case 5: gst(D + e, cscale(gld(D + e), 1.0 / dmax_of(dl[i]))); break;
    case 6: gst(D + e, cscale(gld(D + e), 1.0 / dmax_of(dr[j]))); break;
    case 7: gst(D + e, cscale(cconj(gld(gp(c.Us) + et)), -dmin_of(dl[i]))); break;
    case 8: gst(D + e, gld(gp(dsel == M_GT0 ? c.gt0s : c.g0ts) + e)); break;
    }
}

// -DDYNERROR: the largest |propagated - green_tau| of G(t,0) and G(0,t) into c.dynerr[0], the
// count of compared matrices into c.dynerr[1]; one workgroup per chain
extern "C" __global__ void __launch_bounds__(256)
k_dyn_err(const DChain *chains)
{
    __shared__ double red[4];
    const DChain &c = chains[blockIdx.x];
    if (!c.ltau) return;
    const size_t nn = (size_t)c.n * c.n;
    double m = 0.0;
    for (size_t e = threadIdx.x; e < nn; e += blockDim.x) {
        const cplx a = gld(gp(c.gt0) + e), b = gld(gp(c.gt0s) + e);
        const cplx u = gld(gp(c.g0t) + e), v = gld(gp(c.g0ts) + e);
        m = fmax(m, fmax(hypot(a.x - b.x, a.y - b.y), hypot(u.x - v.x, u.y - v.y)));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
        c.dynerr[0] = fmax(c.dynerr[0], m);
        c.dynerr[1] += 2.0;
    }
}

// SC_INFO of the G(t,t) LU kept aside while the G(0,0) LUs run (mode 0), put back (mode 1)
extern "C" __global__ void k_dyn_info(const DChain *chains, int nch, int mode)
{
    const int ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= nch) return;
    double *sc = chains[ci].sc;
    if (mode == 0) sc[31] = sc[SC_INFO];
    else sc[SC_INFO] = sc[31];
}

// dyn_measure: bin b (0-based slice) of znn_tau += ((1/2 - Re G(t,t)_ii) (1/2 - Re G(0,0)_jj)
// - G(0,t)_ji G(t,0)_ij / nflr) zphi at (imj, orbital i, orbital j), imj = (npbc(zi - zj) - 1)
// l1 l2 + npbc(yi - yj) + (npbc(xi - xj) - 1) l2 (src/lattice.f90:471-479, :1136-1145; the 1D
// and 2D maps are the cases l2 = l3 = 1 and l3 = 1).  One workgroup per chain; the bins are summed
// in LDS (lq x 4 complex) and added to the chain's znn once.
extern "C" __global__ void __launch_bounds__(256)
k_dyn_measure(const DChain *chains, int b)
{
    extern __shared__ double dbin[];        // 2 * 4 * lq, then 2 * n diagonals
    const DChain &c = chains[blockIdx.x];
    const int n = c.n, lq = c.lq, l1 = c.dl1, l2 = c.dl2, l3 = c.dl3;
    if (!c.ltau || b > c.ntdm) return;
    double *dtt = dbin + 8 * lq, *d00 = dtt + n;
    for (int k = threadIdx.x; k < 8 * lq; k += blockDim.x) dbin[k] = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        dtt[k] = 0.5 - gld(gp(c.G) + ((size_t)k + (size_t)k * n)).x;
        d00[k] = 0.5 - gld(gp(c.g00) + ((size_t)k + (size_t)k * n)).x;
    }
    __syncthreads();
    const cplx phase = make_double2(c.sc[SC_PHASE], c.sc[SC_PHASE + 1]);
    const double sg = phase.x < 0 ? -1.0 : 1.0;
    const cplx zphi = cscale(phase, sg / phase.x);
    const double rf = 1.0 / (double)c.nflr;
    GAS const int32_t *xy = gp(c.sxy);
    for (size_t e = threadIdx.x; e < (size_t)n * n; e += blockDim.x) {
        const int i = (int)(e % n), j = (int)(e / n);
        const cplx gt = gld(gp(c.gt0) + e);                                    // G(t,0)_ij
        const cplx g0 = gld(gp(c.g0t) + ((size_t)j + (size_t)i * n));           // G(0,t)_ji
        const cplx p = cmul(g0, gt);
        const cplx v = make_double2(dtt[i] * d00[j] - rf * p.x, -rf * p.y);
        const cplx w = cmul(v, zphi);
        int dx = xy[4 * i] - xy[4 * j], dy = xy[4 * i + 1] - xy[4 * j + 1], dz = xy[4 * i + 2] - xy[4 * j + 2];
        dx = dx < 1 ? dx + l1 : dx;
        dy = dy < 1 ? dy + l2 : dy;
        dz = dz < 1 ? dz + l3 : dz;
        const int imj = (dz - 1) * l1 * l2 + dy + (dx - 1) * l2 - 1;
        const int k = imj + lq * (xy[4 * i + 3] + 2 * xy[4 * j + 3]);
        atomicAdd(&dbin[2 * k], w.x);
        atomicAdd(&dbin[2 * k + 1], w.y);
    }
    __syncthreads();
    double *zb = (double *)(c.znn + (size_t)4 * lq * b);
    for (int k = threadIdx.x; k < 8 * lq; k += blockDim.x) zb[k] += dbin[k];
}

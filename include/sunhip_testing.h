/*
 * sunhip_testing.h -- test hooks of libsunhip.so (not part of the drop-in boundary).
 *
 * They upload/download one chain's work matrices and run single kernel groups so that the
 * unit tests can check every kernel against a float64 numpy reference.  No counterpart in
 * the reference; callers are the kernel tests only.
 */
#ifndef SUNHIP_TESTING_H
#define SUNHIP_TESTING_H
#include "sunhip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* matrix selectors (N x N complex): 0 G, 1 Gs, 2 W1, 3 W2, 4 W3, 5 W4, 6 Us, 7 Vs, 8 U[pt], 9 V[pt]
 * vector selectors: 100 D[pt] (N doubles), 101 jpvt (N int32 as doubles), 102 tau (N complex),
 *                   103 ldq[pt] (1 complex), 104 scalar slots (32 doubles), 105 piv (N as doubles) */
int sh_dbg_put(sh_ctx *ctx, int chain, int sel, int pt, const double *src);
int sh_dbg_get(sh_ctx *ctx, int chain, int sel, int pt, double *dst);

/* op codes (args in a[]):
 *  0: W3 = W1 W2        1: W3 = W1^H W2        2: W3 = W1 W2^H
 *  3: column-pivoted QR of W1 -> U[a0], D[a0], V[a0], ldq[a0]
 *  4: prepivoted QR of W1 diag(D[a1]) with V[a0] = D^-1 R P^T V[a1] -> U[a0], D[a0], V[a0], ldq[a0]
 *  5: LU of W3 (log det in scalar slot 6,7; info in slot 14), then solve W3 Y = W4 in place
 *  6: left product on W1: k_cb_cols program a0, kinetic table a1, V table a2, slices a3..a4 step a5
 *  7: right product on G: k_cb_rows program a0, kinetic table a1, V table a2, slice a3
 *  8: green_equaltime: right point (saved a0, pt a1), left point (saved a2, pt a3) -> Gs, slot 4,5
 *  9: green_equaltimebb from point a0 -> Gs
 * 10: delay (or submatrix) update of slice a0 using the pass uniforms of slice index a1
 * 11: draw the pass uniforms (2 N ltrot) from each chain's generator
 * 12: measurement of G into the observable sums
 * 13: wrap check/replace (compare = a0)
 * 14: a0 != 0 selects the delay update that keeps its vectors in L2 instead of registers */
int sh_dbg_op(sh_ctx *ctx, int op, const int *a);

#ifdef __cplusplus
}
#endif
#endif

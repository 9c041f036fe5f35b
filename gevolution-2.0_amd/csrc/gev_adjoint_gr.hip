// gev_adjoint_gr.hip -- reverse-mode kernels of the weak-field GR cycle
// (main.cu:608-1317 with the elliptic B, as gev_step runs it):
//
//   T00, Tij, L = deposits(x, q; phi_old)           projection_T00 / Tij / T0i functors
//                                                   (gevolution.hpp:1729-1854, 2582-2965, 2031-2201;
//                                                   L = T0i link sums without 1 + phi_a + phi_b)
//   src  = prepareFTsource T00 form(T00, phi_old, chi_old)      (gevolution.hpp:160-218)
//   phi  = P src                                    (solveModifiedPoissonFT, :806-852)
//   Sij  = prepareFTsource Tij form(phi, Tij)       (:64-138)
//   chi  = C Sij                                    (projectFTscalar, :251-340)
//   B    = V (L (1 + phi_a + phi_b))                (projectFTvector, :449-515, elliptic)
//   q'   = update_q(x, q; phi, chi, B)              (:888-1011)
//   x'   = x + dtau update_pos rate(x, q'; phi, chi, B)   (:1162-1232)
//
// P, C and V are FFT-diagonal: P and V are real symmetric (self-adjoint) as
// real-space maps, and C's adjoint is the same transform pair with the
// conjugate per-mode weights (k_adj_project_ft_scalar).  The per-particle
// reverses of the kick + drift (adj_gr_particle) and of the deposits
// (adj_gr_deposit_particle) are written out term by term, one pass over the
// stencil per particle.  The field adjoints are scattered with atomics (their
// order follows atomic arbitration).  Per-particle adjoints lx, lq are [3][n]
// arrays indexed by particle ID, as in the Newtonian adjoint; positions in box
// units (d/dx = N d/dr).
#include <type_traits>

#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_kops.hpp"
#include "gev_pcommon.hpp"

// FP32 tiles of the reverse kick/drift and deposit: 1 the wide ones
// (16 x 8 x 8), 0 the 8 x 8 x 4 ones (more workgroups per CU, more halo)
#ifndef ADJ_KD_WIDE
#define ADJ_KD_WIDE 1
#endif
#ifndef ADJ_DEP_WIDE
#define ADJ_DEP_WIDE 1
#endif
#ifndef ADJ_CHUNK_X4
#define ADJ_CHUNK_X4 1  // four times the mean particles per tile cell from which the reverse kick/drift walks per-lane chunks
#endif

namespace gev {

template <typename T> __device__ __forceinline__ void gadd(T *p, double v) { atomicAdd(p, (T)v); }

// CIC weight of corner b and its derivatives in r
template <typename R> __device__ __forceinline__ R cw(int b, const R r[3])
{
    const R o = (R)1;
    return (((b >> 2) & 1) ? r[0] : o - r[0]) * (((b >> 1) & 1) ? r[1] : o - r[1]) * ((b & 1) ? r[2] : o - r[2]);
}
template <typename R> __device__ __forceinline__ void cw_grad(int b, const R r[3], R g[3])
{
    const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
    const R o = (R)1;
    const R wx = bx ? r[0] : o - r[0], wy = by ? r[1] : o - r[1], wz = bz ? r[2] : o - r[2];
    g[0] = (bx ? o : -o) * wy * wz;
    g[1] = wx * (by ? o : -o) * wz;
    g[2] = wx * wy * (bz ? o : -o);
}

// field accessor around the particle's cell, in the field's precision
template <typename T> struct FAt {
    const T *f;
    int N, x, y, z;
    __device__ __forceinline__ T operator()(int dx, int dy, int dz) const { return f[site(N, x + dx, y + dy, z + dz)]; }
};

// corner values c[4 dx + 2 dy + dz] of a cube as a field accessor (grad1 of
// it is the r-gradient of their trilinear interpolation)
template <typename R> struct CubeAt {
    const R *c;
    __device__ __forceinline__ R operator()(int dx, int dy, int dz) const { return c[4 * dx + 2 * dy + dz]; }
};

// pgradB of update_q (gevolution.hpp:968-994) is linear in the momenta,
// pg_l = sum_m P[l][m] q_m; the matrix P at offsets r and, for the reverse,
// its r-derivatives: dPd[l][j] = dP[l][l]/dr_j (the second-order stencils of
// B_l along l times the trilinear weights across) and the three mixed
// differences M[l] = d^2 B_l / dr_a dr_b over the cube, a, b the other two
// axes (dP[0][1]/dr_2 = dP[2][1]/dr_0 = M[1], and so on)
template <typename R, class FA>
__device__ __forceinline__ void pgrad_matrix(const FA &B0, const FA &B1, const FA &B2, const R r[3], R P[3][3],
                                             R dPd[3][3], R M[3])
{
    const R o = (R)1, s0 = o - r[0], s1 = o - r[1], s2 = o - r[2];
    // staggered links across each axis (the drift's B interpolation uses the same ones)
    P[0][1] = s2 * (B1(1, 0, 0) - B1(0, 0, 0)) + r[2] * (B1(1, 0, 1) - B1(0, 0, 1));
    P[0][2] = s1 * (B2(1, 0, 0) - B2(0, 0, 0)) + r[1] * (B2(1, 1, 0) - B2(0, 1, 0));
    P[1][0] = s2 * (B0(0, 1, 0) - B0(0, 0, 0)) + r[2] * (B0(0, 1, 1) - B0(0, 0, 1));
    P[1][2] = s0 * (B2(0, 1, 0) - B2(0, 0, 0)) + r[0] * (B2(1, 1, 0) - B2(1, 0, 0));
    P[2][0] = s1 * (B0(0, 0, 1) - B0(0, 0, 0)) + r[1] * (B0(0, 1, 1) - B0(0, 1, 0));
    P[2][1] = s0 * (B1(0, 0, 1) - B1(0, 0, 0)) + r[0] * (B1(1, 0, 1) - B1(1, 0, 0));
    M[0] = B0(0, 1, 1) - B0(0, 0, 1) - B0(0, 1, 0) + B0(0, 0, 0);
    M[1] = B1(1, 0, 1) - B1(0, 0, 1) - B1(1, 0, 0) + B1(0, 0, 0);
    M[2] = B2(1, 1, 0) - B2(0, 1, 0) - B2(1, 0, 0) + B2(0, 0, 0);
    // the diagonal: B_l's three-point stencil along l at the four lines across
    // (u, v the offsets along the two other axes in increasing axis order)
    auto diag = [&](int l, auto at) {
        const R rl = r[l];
        const int a = l == 0 ? 1 : 0, b = l == 2 ? 1 : 2;
        const R ra = r[a], rb = r[b], sa = o - ra, sb = o - rb;
        R X[2][2], L[2][2];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int v = 0; v < 2; v++) {
                const R m = at(-1, u, v), c = at(0, u, v), p = at(1, u, v);
                X[u][v] = (rl - o) * m + (o - (R)2 * rl) * c + rl * p;
                L[u][v] = m - (R)2 * c + p;
            }
        P[l][l] = sa * sb * X[0][0] + ra * sb * X[1][0] + sa * rb * X[0][1] + ra * rb * X[1][1];
        dPd[l][l] = sa * sb * L[0][0] + ra * sb * L[1][0] + sa * rb * L[0][1] + ra * rb * L[1][1];
        dPd[l][a] = sb * (X[1][0] - X[0][0]) + rb * (X[1][1] - X[0][1]);
        dPd[l][b] = sa * (X[0][1] - X[0][0]) + ra * (X[1][1] - X[1][0]);
    };
    diag(0, [&](int t, int u, int v) { return B0(t, u, v); });
    diag(1, [&](int t, int u, int v) { return B1(u, t, v); });
    diag(2, [&](int t, int u, int v) { return B2(u, v, t); });
}

struct AdjKD {
    int nf_k, nf_d;
    double a_k, a_d, dtau_k, dtau_d, bc_k, bc_d;
};

// Reverse of the fused kick + drift of one particle (state before the kick),
// at cell offsets r and momenta q, with the adjoints LX, LQ after the drift:
// nx, nq the adjoints before the kick; the fields' adjoints go to
// sc(f, dx, dy, dz, value) with f = 0 phi, 1 chi, 2..4 B_x..B_z and the
// site's offset from the particle's cell.  Fields are read through the
// accessors Fp, Fc, FB0..FB2 (offsets from the particle's cell).
//
// The forward (update_q, gevolution.hpp:888-1011, then the update_pos rate,
// :1162-1232):
//   e = sqrt(|q|^2 + a_k^2),  Phibar = 4 Phi v2 / e^2 + (1 + a_k^2 Phi / e^2)(v2 + e^2) / e^2
//   q' = q - K (e Phibar G2 - e G1 + P q / bcoeff_k),  K = dtau_k N
//   x' = x + dtau_d (q' fac + b / bcoeff_d),  fac = (1 + (3 - w2 / f^2) Phi - Chi) / f
// with v2 = |q|^2, w2 = |q'|^2, f^2 = w2 + a_d^2, Phi, Chi the CIC values,
// G2 = grad2(phi), G1 = grad1(chi) (the r-gradient of Chi), P the pgradB
// matrix and b the staggered B interpolation, whose r-gradient is P's
// off-diagonal (d(mu.b)/dr_j = sum_{l != j} mu_l P[j][l]).  Its reverse is
// written out term by term: one pass over the stencil per particle in place
// of six forward tangent directions.
template <typename T, class FA, class SC>
__device__ __forceinline__ void adj_gr_particle(const T rt[3], const T qt[3], const T LX[3], const T LQ[3], const FA &Fp,
                                                const FA &Fc, const FA &FB0, const FA &FB1, const FA &FB2, int N,
                                                const AdjKD &s, T nx[3], T nq[3], SC &sc)
{
    // in the container's precision (FP32 for the mixed form)
    using R = T;
    const R r[3] = {rt[0], rt[1], rt[2]};
    const R q[3] = {qt[0], qt[1], qt[2]};
    const R Nd = (R)N;
    const R o = (R)1;
    const int nfm = s.nf_k > s.nf_d ? s.nf_k : s.nf_d;
    // ---- the kick, keeping what its reverse needs
    const R K = (R)(s.dtau_k * N);
    const R a2k = (R)(s.a_k * s.a_k);
    const R v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    const R e2 = v2 + a2k, ie2 = o / e2, e = sqrt(e2);
    const R Phi = cic<R>(Fp, r);
    R gP[3], G2[3], G1[3] = {(R)0, (R)0, (R)0};
    grad1<R>(Fp, r, gP);  // dPhi/dr
    grad2<R>(Fp, r, G2);
    if (nfm >= 2) grad1<R>(Fc, r, G1);
    const R Chi = s.nf_d >= 2 ? cic<R>(Fc, r) : (R)0;
    R P[3][3] = {}, dPd[3][3] = {}, Mx[3] = {};
    if (nfm >= 3) pgrad_matrix<R>(FB0, FB1, FB2, r, P, dPd, Mx);
    const R p1 = Phi * ie2;
    const R Phibar = (R)4 * p1 * v2 + (o + a2k * p1) * (v2 + e2) * ie2;
    const R dPhibar = ((R)4 * v2 + a2k * (v2 + e2) * ie2) * ie2;  // dPhibar/dPhi
    const R dPbv = a2k * ie2 * ie2 * (((R)4 * Phi + o) - (R)2 * Phi * v2 * ie2);  // dPhibar/dv2 (e^2 = v2 + a_k^2)
    const R k2 = s.nf_k >= 2 ? o : (R)0;
    const R ibk = s.nf_k >= 3 ? (R)(1.0 / s.bc_k) : (R)0;
    R qq[3];
#pragma unroll
    for (int l = 0; l < 3; l++)
        qq[l] = q[l] - K * (e * (Phibar * G2[l] - k2 * G1[l]) + ibk * (P[l][0] * q[0] + P[l][1] * q[1] + P[l][2] * q[2]));
    // ---- the drift's reverse at (r, q'), mu = dtau_d LX
    const R a2d = (R)(s.a_d * s.a_d);
    const R w2 = qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2];
    const R f2 = w2 + a2d, ie = o / sqrt(f2), ie_2 = ie * ie, ie3 = ie * ie_2;
    const R ph = s.nf_d >= 1 ? Phi : (R)0;
    const R dtd = (R)s.dtau_d;
    const R mu[3] = {dtd * LX[0], dtd * LX[1], dtd * LX[2]};
    const R mq = mu[0] * qq[0] + mu[1] * qq[1] + mu[2] * qq[2];
    const R qL = qq[0] * LX[0] + qq[1] * LX[1] + qq[2] * LX[2];
    const R t3 = (R)3 - w2 * ie_2;
    const R fac = (o + t3 * ph - Chi) * ie;
    // dfac/dw2
    const R dfac = -(R)0.5 * ie3 * (o + (R)3 * ph - Chi) - ph * ie3 + (R)1.5 * ph * w2 * ie3 * ie_2;
    R lqt[3];
#pragma unroll
    for (int j = 0; j < 3; j++) lqt[j] = LQ[j] + mu[j] * fac + (R)2 * mq * dfac * qq[j];
    const R ibd = s.nf_d >= 3 ? (R)(1.0 / s.bc_d) : (R)0;
    const R cph = s.nf_d >= 1 ? mq * ie * t3 : (R)0, cch = s.nf_d >= 2 ? mq * ie : (R)0;
    R ar[3];
    ar[0] = cph * gP[0] - cch * G1[0] + ibd * (mu[1] * P[0][1] + mu[2] * P[0][2]);
    ar[1] = cph * gP[1] - cch * G1[1] + ibd * (mu[0] * P[1][0] + mu[2] * P[1][2]);
    ar[2] = cph * gP[2] - cch * G1[2] + ibd * (mu[0] * P[2][0] + mu[1] * P[2][1]);
    // ---- the kick's reverse, lambda = lqt
    const R lg2 = lqt[0] * G2[0] + lqt[1] * G2[1] + lqt[2] * G2[2];
    const R lg1 = k2 * (lqt[0] * G1[0] + lqt[1] * G1[1] + lqt[2] * G1[2]);
    // momenta: d(e)/dq_j = q_j / e, dPhibar/dq_j = 2 q_j dPhibar/dv2, P linear
    const R cq = (Phibar * lg2 - lg1) / e + (R)2 * e * lg2 * dPbv;
#pragma unroll
    for (int j = 0; j < 3; j++)
        nq[j] = lqt[j] - K * (q[j] * cq + ibk * (lqt[0] * P[0][j] + lqt[1] * P[1][j] + lqt[2] * P[2][j]));
    // offsets: lambda.G2 is the trilinear interpolation of the corner values
    // c_b = lambda . (phi(b + e_l) - phi(b - e_l)) / 2
    R cb[8];
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
        cb[b] = (R)0.5 * (lqt[0] * (Fp(bx + 1, by, bz) - Fp(bx - 1, by, bz)) +
                          lqt[1] * (Fp(bx, by + 1, bz) - Fp(bx, by - 1, bz)) +
                          lqt[2] * (Fp(bx, by, bz + 1) - Fp(bx, by, bz - 1)));
    }
    R dlg2[3];
    grad1<R>(CubeAt<R>{cb}, r, dlg2);
    R kr[3];
#pragma unroll
    for (int j = 0; j < 3; j++) kr[j] = e * (lg2 * dPhibar * gP[j] + Phibar * dlg2[j]);
    if (s.nf_k >= 2) {
        // lambda.G1 = sum_l lambda_l dChi/dr_l: its r-gradient takes the mixed
        // second differences of chi over the cube
        const R s0 = o - r[0], s1 = o - r[1], s2 = o - r[2];
        const R H01 = s2 * (Fc(1, 1, 0) - Fc(1, 0, 0) - Fc(0, 1, 0) + Fc(0, 0, 0)) +
                      r[2] * (Fc(1, 1, 1) - Fc(1, 0, 1) - Fc(0, 1, 1) + Fc(0, 0, 1));
        const R H02 = s1 * (Fc(1, 0, 1) - Fc(1, 0, 0) - Fc(0, 0, 1) + Fc(0, 0, 0)) +
                      r[1] * (Fc(1, 1, 1) - Fc(1, 1, 0) - Fc(0, 1, 1) + Fc(0, 1, 0));
        const R H12 = s0 * (Fc(0, 1, 1) - Fc(0, 1, 0) - Fc(0, 0, 1) + Fc(0, 0, 0)) +
                      r[0] * (Fc(1, 1, 1) - Fc(1, 1, 0) - Fc(1, 0, 1) + Fc(1, 0, 0));
        kr[0] -= e * (lqt[1] * H01 + lqt[2] * H02);
        kr[1] -= e * (lqt[0] * H01 + lqt[2] * H12);
        kr[2] -= e * (lqt[0] * H02 + lqt[1] * H12);
    }
    if (s.nf_k >= 3) {
        // lambda . P q: the diagonal's own r-derivatives and the mixed ones
        const R l0 = lqt[0], l1 = lqt[1], l2 = lqt[2];
        kr[0] += ibk * (l0 * q[0] * dPd[0][0] + l1 * (q[1] * dPd[1][0] + q[2] * Mx[2]) +
                        l2 * (q[1] * Mx[1] + q[2] * dPd[2][0]));
        kr[1] += ibk * (l0 * (q[0] * dPd[0][1] + q[2] * Mx[2]) + l1 * q[1] * dPd[1][1] +
                        l2 * (q[0] * Mx[0] + q[2] * dPd[2][1]));
        kr[2] += ibk * (l0 * (q[0] * dPd[0][2] + q[1] * Mx[1]) + l1 * (q[0] * Mx[0] + q[1] * dPd[1][2]) +
                        l2 * q[2] * dPd[2][2]);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) nx[j] = LX[j] + Nd * (ar[j] - K * kr[j]);

    // ---- field adjoints of the kick (pre-kick q, lambda = lqt) and of the
    // drift (kicked q', lambda = LX), summed per field over the stencil
    // offsets in registers first: 76 distinct sites take one atomic each
    // instead of 152 scattered adds
    // st[dx + 1][dy + 1][dz + 1]: one field's adjoint at stencil offset (dx, dy, dz)
    R st[4][4][4];
    auto clear = [&]() {
#pragma unroll
        for (int i = 0; i < 64; i++) st[i >> 4][(i >> 2) & 3][i & 3] = (R)0;
    };
    auto add = [&](int dx, int dy, int dz, R val) { st[dx + 1][dy + 1][dz + 1] += val; };
    auto put = [&](int f, int dx, int dy, int dz) { sc(f, dx, dy, dz, st[dx + 1][dy + 1][dz + 1]); };
    // phi: g = Phibar grad2(phi) - grad1(chi) + pgradB / (bcoeff e); q' = q - K e g
    {
        clear();
        const R lG2[3] = {-K * e * Phibar * lqt[0], -K * e * Phibar * lqt[1], -K * e * Phibar * lqt[2]};
        const R lPhi = -K * e * dPhibar * lg2;
        const R lph = s.nf_d >= 1 ? dtd * qL * t3 * ie : (R)0;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
            const R w = cw(b, r);
            add(bx, by, bz, w * (lPhi + lph));
            const R hw = (R)0.5 * w;
            add(bx + 1, by, bz, hw * lG2[0]);
            add(bx - 1, by, bz, -hw * lG2[0]);
            add(bx, by + 1, bz, hw * lG2[1]);
            add(bx, by - 1, bz, -hw * lG2[1]);
            add(bx, by, bz + 1, hw * lG2[2]);
            add(bx, by, bz - 1, -hw * lG2[2]);
        }
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int u = (b >> 2) & 1, v_ = (b >> 1) & 1, w_ = b & 1;
            put(0, u, v_, w_);                              // the cube
        }
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int u = (b >> 1) & 1, v_ = b & 1;
            put(0, -1, u, v_); put(0, 2, u, v_);            // x faces
            put(0, u, -1, v_); put(0, u, 2, v_);            // y faces
            put(0, u, v_, -1); put(0, u, v_, 2);            // z faces
        }
    }
    // chi: the kick's -grad1(chi) and the drift's -chi
    if (s.nf_k >= 2 || s.nf_d >= 2) {
        clear();
        const R lG1[3] = {K * e * lqt[0], K * e * lqt[1], K * e * lqt[2]};
        const R lch = (R)(-s.dtau_d * qL * ie);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
            R val = (R)0;
            if (s.nf_k >= 2) {
                R gw[3];
                cw_grad(b, r, gw);
                val += gw[0] * lG1[0] + gw[1] * lG1[1] + gw[2] * lG1[2];
            }
            if (s.nf_d >= 2) val += cw(b, r) * lch;
            add(bx, by, bz, val);
        }
#pragma unroll
        for (int b = 0; b < 8; b++) put(1, (b >> 2) & 1, (b >> 1) & 1, b & 1);
    }
    // B: pgradB (gevolution.hpp:968-994) is linear in B, its transpose term by
    // term; the drift's staggered B interpolation (:1213-1226)
    if (s.nf_k >= 3 || s.nf_d >= 3) {
        const R c = s.nf_k >= 3 ? -K / s.bc_k : (R)0;
        const R l0 = c * lqt[0], l1 = c * lqt[1], l2 = c * lqt[2];
        const R cd = s.nf_d >= 3 ? (R)(s.dtau_d / s.bc_d) : (R)0;
        const R b0 = cd * LX[0], b1 = cd * LX[1], b2 = cd * LX[2];
        const R s0 = o - r[0], s1 = o - r[1], s2 = o - r[2];
        // B0: x stencil of pg[0], the y and z links of pg[1] and pg[2], the drift
        clear();
        {
            const R t = l0 * q[0];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int by = (b >> 1) & 1, bz = b & 1;
                const R w = (by ? r[1] : s1) * (bz ? r[2] : s2) * t;
                add(-1, by, bz, w * (r[0] - o)); add(0, by, bz, w * (o - (R)2 * r[0])); add(1, by, bz, w * r[0]);
            }
            const R t1 = l1 * q[0], t2 = l2 * q[0];
            add(0, 1, 0, t1 * s2 - t2 * r[1] + b0 * r[1] * s2);
            add(0, 0, 0, -t1 * s2 - t2 * s1 + b0 * s1 * s2);
            add(0, 1, 1, t1 * r[2] + t2 * r[1] + b0 * r[1] * r[2]);
            add(0, 0, 1, -t1 * r[2] + t2 * s1 + b0 * s1 * r[2]);
        }
#pragma unroll
        for (int b = 0; b < 12; b++) put(2, b / 4 - 1, (b >> 1) & 1, b & 1);
        // B1: y stencil of pg[1], the x and z links of pg[0] and pg[2], the drift
        clear();
        {
            const R t = l1 * q[1];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int bx = (b >> 1) & 1, bz = b & 1;
                const R w = (bx ? r[0] : s0) * (bz ? r[2] : s2) * t;
                add(bx, -1, bz, w * (r[1] - o)); add(bx, 0, bz, w * (o - (R)2 * r[1])); add(bx, 1, bz, w * r[1]);
            }
            const R t0 = l0 * q[1], t2 = l2 * q[1];
            add(1, 0, 0, t0 * s2 - t2 * r[0] + b1 * r[0] * s2);
            add(0, 0, 0, -t0 * s2 - t2 * s0 + b1 * s0 * s2);
            add(1, 0, 1, t0 * r[2] + t2 * r[0] + b1 * r[0] * r[2]);
            add(0, 0, 1, -t0 * r[2] + t2 * s0 + b1 * s0 * r[2]);
        }
#pragma unroll
        for (int b = 0; b < 12; b++) put(3, (b >> 1) & 1, b / 4 - 1, b & 1);
        // B2: z stencil of pg[2], the x and y links of pg[0] and pg[1], the drift
        clear();
        {
            const R t = l2 * q[2];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int bx = (b >> 1) & 1, by = b & 1;
                const R w = (bx ? r[0] : s0) * (by ? r[1] : s1) * t;
                add(bx, by, -1, w * (r[2] - o)); add(bx, by, 0, w * (o - (R)2 * r[2])); add(bx, by, 1, w * r[2]);
            }
            const R t0 = l0 * q[2], t1 = l1 * q[2];
            add(1, 0, 0, t0 * s1 - t1 * r[0] + b2 * r[0] * s1);
            add(0, 0, 0, -t0 * s1 - t1 * s0 + b2 * s0 * s1);
            add(1, 1, 0, t0 * r[1] + t1 * r[0] + b2 * r[0] * r[1]);
            add(0, 1, 0, -t0 * r[1] + t1 * s0 + b2 * s0 * r[1]);
        }
#pragma unroll
        for (int b = 0; b < 12; b++) put(4, (b >> 1) & 1, b & 1, b / 4 - 1);
    }
}

// The tiled reverse kick + drift: one workgroup per tile of TX x TY x TZ cells
// reads the fields of the stencil box [x0-1, x0+TX+1] x ... from LDS, as the
// forward update does, and sums the field adjoints of its particles in FP64
// LDS accumulators over the same box (the scatter stencils stay inside it),
// which go to the global arrays with one atomic per site and field at the
// end.  A crowded cell then costs LDS atomics, not global ones.
template <typename T, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) void k_adj_gr_kick_drift(PV<T> v, const T *__restrict__ phi, const T *__restrict__ chi,
                                                          const T *__restrict__ Bi, T *lx, T *lq, T *lphi, T *lchi,
                                                          T *lB, int64_t M, AdjKD s)
{
    constexpr int SX = TX + 3, SY = TY + 3, SZ = TZ + 3, NS = SX * SY * SZ, NROW = TY * TZ;
    __shared__ T fld[5][NS];
    __shared__ double acc[5][NS];
    const int N = v.N;
    const uint32_t msk = (uint32_t)N - 1;
    const int tid = threadIdx.x, lane = tid & 63;
    int x0, y0, z0;
    tile_origin<TX, TY, TZ>(N, x0, y0, z0);
    const int nfmax = s.nf_k > s.nf_d ? s.nf_k : s.nf_d;
    for (int e = tid; e < NS; e += NT) {
        const int lx0 = e % SX, ly = (e / SX) % SY, lz = e / (SX * SY);
        const uint32_t g = ((uint32_t)(x0 - 1 + lx0) & msk) +
                           (uint32_t)N * (((uint32_t)(y0 - 1 + ly) & msk) + (uint32_t)N * ((uint32_t)(z0 - 1 + lz) & msk));
        fld[0][e] = phi[g];
        fld[1][e] = nfmax >= 2 ? chi[g] : (T)0;
        fld[2][e] = nfmax >= 3 ? Bi[g] : (T)0;
        fld[3][e] = nfmax >= 3 ? Bi[M + g] : (T)0;
        fld[4][e] = nfmax >= 3 ? Bi[2 * M + g] : (T)0;
#pragma unroll
        for (int f = 0; f < 5; f++) acc[f][e] = 0.0;
    }
    int32_t rb0, incl;
    tile_rows_load<TX, TY, NROW>(v.start, N, x0, y0, z0, lane, rb0, incl);
    incl = wave_row_scan(incl, lane);
    const int total = __shfl(incl, NROW - 1, 64);
    // A crowded tile gives every lane its own contiguous chunk of Q particles,
    // so the lanes of one instruction sit Q particles apart, mostly in
    // different cells, and their field-adjoint atomics rarely meet on one LDS
    // address; other tiles keep consecutive particles in consecutive lanes
    // (coalesced loads)
    constexpr int NW = NT / 64;
    const bool chunk = ADJ_CHUNK_X4 > 0 && 4 * total >= ADJ_CHUNK_X4 * (TX * TY * TZ);
    const int Q = (total + 63) / 64;
    // the same for every lane of the workgroup
    const int niter = chunk ? (Q + NW - 1) / NW : (total + NT - 1) / NT;
    __syncthreads();
    const int64_t n = v.n;
    for (int it = 0; it < niter; it++) {
        const int sq = it * NW + (tid >> 6);  // position inside the lane's chunk
        const int f = chunk ? (sq < Q ? lane * Q + sq : total) : tid + it * NT;
        const int64_t p = tile_locate<NROW>(rb0, incl, f);
        if (p < 0) continue;
        const int64_t id = v.id[p];
        int x, y, z;
        cell_coords(v.cell[p], N, v.lg, x, y, z);
        T rt[3];
        ref_dist(v, p, x, y, z, rt);
        const T qt[3] = {v.q0[p], v.q1[p], v.q2[p]};
        T LX[3], LQ[3], nx[3], nq[3];
#pragma unroll
        for (int j = 0; j < 3; j++) { LX[j] = lx[j * n + id]; LQ[j] = lq[j * n + id]; }
        const int o = (x - x0 + 1) + SX * ((y - y0 + 1) + SY * (z - z0 + 1));
        const FieldLds<T> Fp{fld[0], o, SX, SX * SY}, Fc{fld[1], o, SX, SX * SY}, FB0{fld[2], o, SX, SX * SY},
            FB1{fld[3], o, SX, SX * SY}, FB2{fld[4], o, SX, SX * SY};
        auto sc = [&](int f, int dx, int dy, int dz, T val) {
            __hip_atomic_fetch_add(&acc[f][o + dx + SX * dy + SX * SY * dz], (double)val, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        adj_gr_particle<T>(rt, qt, LX, LQ, Fp, Fc, FB0, FB1, FB2, N, s, nx, nq, sc);
#pragma unroll
        for (int j = 0; j < 3; j++) { lx[j * n + id] = nx[j]; lq[j * n + id] = nq[j]; }
    }
    __syncthreads();
    T *dst[5] = {lphi, lchi, lB, lB + M, lB + 2 * M};
    for (int e = tid; e < NS; e += NT) {
        const int lx0 = e % SX, ly = (e / SX) % SY, lz = e / (SX * SY);
        const uint32_t g = ((uint32_t)(x0 - 1 + lx0) & msk) +
                           (uint32_t)N * (((uint32_t)(y0 - 1 + ly) & msk) + (uint32_t)N * ((uint32_t)(z0 - 1 + lz) & msk));
#pragma unroll
        for (int f = 0; f < 5; f++) {
            const double a = acc[f][e];
            if (a != 0.0) atomicAdd(dst[f] + g, (T)a);
        }
    }
}

// Reverse of the T00 + Tij + T0i-link deposits of one particle (phi_old the
// deposits' phi; masses per unit of the reference's coeff, as
// gev_project_T00_Tij_T0i_links scales them), at cell offsets r and momenta
// q.  F(f, dx, dy, dz) reads field f at an offset from the particle's cell:
// f = 0 lT00, 1..6 lTij, 7..9 lL (the field adjoints), 10 phi_old.  Out:
// gx, gq the adjoint contributions to the particle (d/dx in box units), and
// sc(b, value) phi_old's adjoint at corner b (when has_phi).
//
// The particle's share of <adjoints, deposits> is one scalar
//   S = sum_b W_b (e + f phi_b + f2 phi_b^2) lT00_b               (gevolution.hpp:1729-1854)
//     + (1 + cf pint) mij / e (sum_d A_d q_d^2 + q0 q1 X01 + q0 q2 X02 + q1 q2 X12)   (:2582-2965)
//     + m0i (q0 Lx + q1 Ly + q2 Lz)                                (:2031-2201, no link factor)
// with W_b the CIC weights, e = sqrt(q^2 + a^2), f = 3 e + q^2 / e,
// f2 = (f^2 / 2 + q^2 / (1 + q^2 / a^2)) / e, cf = 4 + a^2 / e^2, pint the
// CIC of phi, A_d the CIC of lT_dd, X the edge interpolations of the
// off-diagonal lT and L the face interpolations of the link adjoints; its
// r- and q-gradients are written out term by term.
template <typename T, class FA, class SC>
__device__ __forceinline__ void adj_gr_deposit_particle(const T rt[3], const T qt[3], const FA &F, bool has_phi,
                                                        bool has_links, int N, double m00, double mij, double m0i,
                                                        double a_d, T gx[3], T gq[3], SC &sc)
{
    using R = T;  // the container's precision
    const R o = (R)1;
    const R r[3] = {rt[0], rt[1], rt[2]};
    const R q[3] = {qt[0], qt[1], qt[2]};
    const R s0 = o - r[0], s1 = o - r[1], s2 = o - r[2];
    const R a2 = (R)(a_d * a_d);
    R W[8], ph[8], lt[8];
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
        W[b] = cw(b, r);
        ph[b] = has_phi ? F(10, bx, by, bz) : (R)0;
        lt[b] = (R)m00 * F(0, bx, by, bz);
    }
    // T00
    const R q2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    const R e = fsqrt(q2 + a2), ie = frcp(e), ie2 = ie * ie, ie3 = ie2 * ie;
    const R f = (R)3 * e + q2 * ie;
    const R iden = frcp(o + q2 * (R)(1.0 / (a_d * a_d))), h = q2 * iden;
    const R hf = (R)0.5 * f * f + h;
    const R f2 = hf * ie;
    R c0[8], A0 = (R)0, A1 = (R)0, A2 = (R)0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        c0[b] = lt[b] * (e + f * ph[b] + f2 * (ph[b] * ph[b]));
        const R wl = W[b] * lt[b];
        A0 += wl;
        A1 += wl * ph[b];
        A2 += wl * ph[b] * ph[b];
    }
    R gr[3];
    grad1<R>(CubeAt<R>{c0}, r, gr);
    // d/d(q^2) of e, f, f2
    const R df = (R)2.5 * ie - (R)0.5 * q2 * ie3;
    const R df2 = (f * df + iden * iden) * ie - (R)0.5 * hf * ie3;
    const R dS0 = (R)0.5 * ie * A0 + df * A1 + df2 * A2;
    // Tij without and with its factor fT = 1 + cf pint
    R pint = (R)0, cd[8], Ad[3] = {(R)0, (R)0, (R)0};
    const int dc[3] = {1, 4, 6};
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
        pint += W[b] * ph[b];
        R c = (R)0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const R v = F(dc[d], bx, by, bz);
            Ad[d] += W[b] * v;
            c += q[d] * q[d] * v;
        }
        cd[b] = c;
    }
    R gpint[3], gd[3];
    grad1<R>(CubeAt<R>{ph}, r, gpint);
    grad1<R>(CubeAt<R>{cd}, r, gd);
    const R X01 = s2 * F(2, 0, 0, 0) + r[2] * F(2, 0, 0, 1);
    const R X02 = s1 * F(3, 0, 0, 0) + r[1] * F(3, 0, 1, 0);
    const R X12 = s0 * F(5, 0, 0, 0) + r[0] * F(5, 1, 0, 0);
    gd[0] += q[1] * q[2] * (F(5, 1, 0, 0) - F(5, 0, 0, 0));
    gd[1] += q[0] * q[2] * (F(3, 0, 1, 0) - F(3, 0, 0, 0));
    gd[2] += q[0] * q[1] * (F(2, 0, 0, 1) - F(2, 0, 0, 0));
    const R Q = Ad[0] * q[0] * q[0] + Ad[1] * q[1] * q[1] + Ad[2] * q[2] * q[2] + q[0] * q[1] * X01 + q[0] * q[2] * X02 +
                q[1] * q[2] * X12;
    const R mi = (R)mij;
    const R Sij = mi * ie * Q;
    const R cf = (R)4 + a2 * ie2, fT = o + cf * pint;
    const R dQ[3] = {(R)2 * Ad[0] * q[0] + q[1] * X01 + q[2] * X02, (R)2 * Ad[1] * q[1] + q[0] * X01 + q[2] * X12,
                     (R)2 * Ad[2] * q[2] + q[0] * X02 + q[1] * X12};
    // the terms along q: T00's through q^2, cf's (dcf/dq_j = -2 a^2 q_j / e^4), 1 / e's
    const R cq = (R)2 * dS0 - (R)2 * Sij * pint * a2 * ie2 * ie2 - fT * mi * ie3 * Q;
    const R fm = fT * mi * ie;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        gr[j] += Sij * cf * gpint[j] + fm * gd[j];
        gq[j] = cq * q[j] + fm * dQ[j];
    }
    // T0i links: linear in q over the face interpolations of the link adjoints
    if (has_links) {
        const R m = (R)m0i;
        const R Lx = s1 * s2 * F(7, 0, 0, 0) + r[1] * s2 * F(7, 0, 1, 0) + s1 * r[2] * F(7, 0, 0, 1) +
                     r[1] * r[2] * F(7, 0, 1, 1);
        const R Ly = s0 * s2 * F(8, 0, 0, 0) + r[0] * s2 * F(8, 1, 0, 0) + s0 * r[2] * F(8, 0, 0, 1) +
                     r[0] * r[2] * F(8, 1, 0, 1);
        const R Lz = s0 * s1 * F(9, 0, 0, 0) + r[0] * s1 * F(9, 1, 0, 0) + s0 * r[1] * F(9, 0, 1, 0) +
                     r[0] * r[1] * F(9, 1, 1, 0);
        gq[0] += m * Lx;
        gq[1] += m * Ly;
        gq[2] += m * Lz;
        const R dLy0 = s2 * (F(8, 1, 0, 0) - F(8, 0, 0, 0)) + r[2] * (F(8, 1, 0, 1) - F(8, 0, 0, 1));
        const R dLz0 = s1 * (F(9, 1, 0, 0) - F(9, 0, 0, 0)) + r[1] * (F(9, 1, 1, 0) - F(9, 0, 1, 0));
        const R dLx1 = s2 * (F(7, 0, 1, 0) - F(7, 0, 0, 0)) + r[2] * (F(7, 0, 1, 1) - F(7, 0, 0, 1));
        const R dLz1 = s0 * (F(9, 0, 1, 0) - F(9, 0, 0, 0)) + r[0] * (F(9, 1, 1, 0) - F(9, 1, 0, 0));
        const R dLx2 = s1 * (F(7, 0, 0, 1) - F(7, 0, 0, 0)) + r[1] * (F(7, 0, 1, 1) - F(7, 0, 1, 0));
        const R dLy2 = s0 * (F(8, 0, 0, 1) - F(8, 0, 0, 0)) + r[0] * (F(8, 1, 0, 1) - F(8, 1, 0, 0));
        gr[0] += m * (q[1] * dLy0 + q[2] * dLz0);
        gr[1] += m * (q[0] * dLx1 + q[2] * dLz1);
        gr[2] += m * (q[0] * dLx2 + q[1] * dLy2);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) gx[j] = (R)N * gr[j];
    if (has_phi) {
        // d/dphi_b: T00's e + f phi + f2 phi^2 per corner, Tij's fT through the CIC of phi
        const R tt = Sij * cf;
#pragma unroll
        for (int b = 0; b < 8; b++) sc(b, W[b] * (lt[b] * (f + (R)2 * f2 * ph[b]) + tt));
    }
}

// The tiled reverse deposit: one workgroup per tile of TX x TY x TZ cells
// stages the eleven fields its particles read (the sites [x0, x0+TX] x ...)
// in LDS and sums phi_old's adjoint there in FP64, added to lphi once per
// site at the end.
template <typename T, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) void k_adj_gr_deposit(PV<T> v, const T *__restrict__ phi, const T *__restrict__ lT00,
                                                       const T *__restrict__ lTij, const T *__restrict__ lL, T *lx,
                                                       T *lq, T *lphi, int64_t M, double m00, double mij, double m0i,
                                                       double a)
{
    constexpr int SX = TX + 1, SY = TY + 1, SZ = TZ + 1, NS = SX * SY * SZ, NROW = TY * TZ;
    __shared__ T fld[11][NS];
    __shared__ double acc[NS];
    const int N = v.N;
    const uint32_t msk = (uint32_t)N - 1;
    const int tid = threadIdx.x, lane = tid & 63;
    int x0, y0, z0;
    tile_origin<TX, TY, TZ>(N, x0, y0, z0);
    const bool has_phi = phi != nullptr, has_links = lL != nullptr, out_phi = has_phi && lphi != nullptr;
    for (int e = tid; e < NS; e += NT) {
        const int ex = e % SX, ey = (e / SX) % SY, ez = e / (SX * SY);
        const uint32_t g = ((uint32_t)(x0 + ex) & msk) +
                           (uint32_t)N * (((uint32_t)(y0 + ey) & msk) + (uint32_t)N * ((uint32_t)(z0 + ez) & msk));
        fld[0][e] = lT00[g];
#pragma unroll
        for (int c = 0; c < 6; c++) fld[1 + c][e] = lTij[c * M + g];
#pragma unroll
        for (int c = 0; c < 3; c++) fld[7 + c][e] = has_links ? lL[c * M + g] : (T)0;
        fld[10][e] = has_phi ? phi[g] : (T)0;
        acc[e] = 0.0;
    }
    int32_t rb0, incl;
    tile_rows_load<TX, TY, NROW>(v.start, N, x0, y0, z0, lane, rb0, incl);
    incl = wave_row_scan(incl, lane);
    const int total = __shfl(incl, NROW - 1, 64);
    const int niter = (total + NT - 1) / NT;  // the same for every lane of the workgroup
    __syncthreads();
    const int64_t n = v.n;
    for (int it = 0; it < niter; it++) {
        const int64_t p = tile_locate<NROW>(rb0, incl, tid + it * NT);
        if (p < 0) continue;
        const int64_t id = v.id[p];
        int x, y, z;
        cell_coords(v.cell[p], N, v.lg, x, y, z);
        T rt[3];
        ref_dist(v, p, x, y, z, rt);
        const T qt[3] = {v.q0[p], v.q1[p], v.q2[p]};
        const int o = (x - x0) + SX * ((y - y0) + SY * (z - z0));
        auto F = [&](int f, int dx, int dy, int dz) -> T { return fld[f][o + dx + SX * dy + SX * SY * dz]; };
        auto sc = [&](int b, T val) {
            __hip_atomic_fetch_add(&acc[o + ((b >> 2) & 1) + SX * ((b >> 1) & 1) + SX * SY * (b & 1)], (double)val,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        T gx[3], gq[3];
        adj_gr_deposit_particle<T>(rt, qt, F, has_phi, has_links, N, m00, mij, m0i, a, gx, gq, sc);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            lx[j * n + id] += gx[j];
            lq[j * n + id] += gq[j];
        }
    }
    if (out_phi) {
        __syncthreads();
        for (int e = tid; e < NS; e += NT) {
            const int ex = e % SX, ey = (e / SX) % SY, ez = e / (SX * SY);
            const uint32_t g = ((uint32_t)(x0 + ex) & msk) +
                               (uint32_t)N * (((uint32_t)(y0 + ey) & msk) + (uint32_t)N * ((uint32_t)(z0 + ez) & msk));
            if (acc[e] != 0.0) atomicAdd(lphi + g, (T)acc[e]);
        }
    }
}

// Reverse of prepareFTsource, T00 form (gevolution.hpp:160-218):
//   src = c2 (T00 - bg)(1 - 2 phi (1 - phi)) + 1/8 sum_i (phi(s-e_i) - phi(s+e_i))^2
//         + (c3 (1 - 3 phi) - c1) phi - c3 chi
// lT00 = d src / dT00 lsrc; lphi += ...; lchi += -c3 lsrc
template <typename T>
__global__ __launch_bounds__(256) void k_adj_source_T00(const T *__restrict__ phi, const T *__restrict__ T00,
                                                        const T *__restrict__ lsrc, T *lT00, T *lphi, T *lchi, int N,
                                                        int64_t M, double bg, double c1, double c2, double c3)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const int x = (int)(s % N), y = (int)((s / N) % N), z = (int)(s / ((int64_t)N * N));
    auto P = [&](int dx, int dy, int dz) { return (double)phi[site(N, x + dx, y + dy, z + dz)]; };
    auto Ls = [&](int dx, int dy, int dz) { return (double)lsrc[site(N, x + dx, y + dy, z + dz)]; };
    const double p = P(0, 0, 0), l = Ls(0, 0, 0), t = (double)T00[s];
    double acc = l * (c2 * (t - bg) * (-2.0 + 4.0 * p) + c3 * (1.0 - 6.0 * p) - c1);
    // 1/8 (phi(s'-e) - phi(s'+e))^2 at s' = s + e and s' = s - e
    acc += 0.25 * (Ls(1, 0, 0) * (p - P(2, 0, 0)) - Ls(-1, 0, 0) * (P(-2, 0, 0) - p));
    acc += 0.25 * (Ls(0, 1, 0) * (p - P(0, 2, 0)) - Ls(0, -1, 0) * (P(0, -2, 0) - p));
    acc += 0.25 * (Ls(0, 0, 1) * (p - P(0, 0, 2)) - Ls(0, 0, -1) * (P(0, 0, -2) - p));
    lphi[s] = (T)((double)lphi[s] + acc);
    lT00[s] = (T)(c2 * (1.0 - 2.0 * p * (1.0 - p)) * l);
    lchi[s] = (T)((double)lchi[s] - c3 * l);
}

// Reverse of prepareFTsource, Tij form (gevolution.hpp:64-138):
//   S_ii = 1/2 (phi(s+e_i) - phi(s-e_i))^2 + c T_ii
//   S_ij = phi_i phi_j - phi phi_ij + 1/2 (phi^2 - phi_i^2 - phi_j^2 + phi_ij^2) + c T_ij
// (phi_i = phi(s+e_i), phi_ij = phi(s+e_i+e_j)); lphi += the stencil's
// transpose applied to lS, gathered per site
template <typename T>
__global__ __launch_bounds__(256) void k_adj_source_Tij_phi(const T *__restrict__ phi, const T *__restrict__ lS, T *lphi,
                                                            int N, int64_t M)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const int x = (int)(s % N), y = (int)((s / N) % N), z = (int)(s / ((int64_t)N * N));
    auto P = [&](int dx, int dy, int dz) { return (double)phi[site(N, x + dx, y + dy, z + dz)]; };
    auto L = [&](int c, int dx, int dy, int dz) { return (double)lS[c * M + site(N, x + dx, y + dy, z + dz)]; };
    const int dcomp[3] = {0, 3, 5};
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int ex = i == 0, ey = i == 1, ez = i == 2;
        // S_ii at s - e_i uses phi(s) as its +e_i point, at s + e_i as its -e_i point
        acc += L(dcomp[i], -ex, -ey, -ez) * (P(0, 0, 0) - P(-2 * ex, -2 * ey, -2 * ez));
        acc -= L(dcomp[i], ex, ey, ez) * (P(2 * ex, 2 * ey, 2 * ez) - P(0, 0, 0));
    }
    const int oc[3] = {1, 2, 4};           // (01, 02, 12)
    const int oi[3] = {0, 0, 1}, oj[3] = {1, 2, 2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int ix = oi[c] == 0, iy = oi[c] == 1, iz = oi[c] == 2;
        const int jx = oj[c] == 0, jy = oj[c] == 1, jz = oj[c] == 2;
        const int k = oc[c];
        const double p = P(0, 0, 0);
        // s' = s - e_i: d/dphi(s'+e_i) = phi(s'+e_j) - phi(s'+e_i)
        acc += L(k, -ix, -iy, -iz) * (P(jx - ix, jy - iy, jz - iz) - p);
        // s' = s - e_j: d/dphi(s'+e_j) = phi(s'+e_i) - phi(s'+e_j)
        acc += L(k, -jx, -jy, -jz) * (P(ix - jx, iy - jy, iz - jz) - p);
        // s' = s: d/dphi(s') = phi(s') - phi(s'+e_i+e_j)
        acc += L(k, 0, 0, 0) * (p - P(ix + jx, iy + jy, iz + jz));
        // s' = s - e_i - e_j: d/dphi(s'+e_i+e_j) = phi(s'+e_i+e_j) - phi(s')
        acc += L(k, -ix - jx, -iy - jy, -iz - jz) * (p - P(-ix - jx, -iy - jy, -iz - jz));
    }
    lphi[s] = (T)((double)lphi[s] + acc);
}

// Reverse of the link factor of the B solve's input, L_i (1 + phi(s) + phi(s+e_i)):
// u = V^T lB (given); lL_i(s) = (1 + phi(s) + phi(s+e_i)) u_i(s) (written over u);
// lphi(s) += sum_i L_i(s) u_i(s) + L_i(s-e_i) u_i(s-e_i)
template <typename T>
__global__ __launch_bounds__(256) void k_adj_link_factor_phi(const T *__restrict__ phi, const T *__restrict__ L,
                                                             const T *__restrict__ u, T *lphi, int N, int64_t M)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const int x = (int)(s % N), y = (int)((s / N) % N), z = (int)(s / ((int64_t)N * N));
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int64_t sm = site(N, x - (i == 0), y - (i == 1), z - (i == 2));
        acc += (double)L[i * M + s] * (double)u[i * M + s] + (double)L[i * M + sm] * (double)u[i * M + sm];
    }
    lphi[s] = (T)((double)lphi[s] + acc);
}
template <typename T>
__global__ __launch_bounds__(256) void k_adj_link_factor_L(const T *__restrict__ phi, T *u, int N, int64_t M)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const int x = (int)(s % N), y = (int)((s / N) % N), z = (int)(s / ((int64_t)N * N));
    const double p = (double)phi[s];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double pe = (double)phi[site(N, x + (i == 0), y + (i == 1), z + (i == 2))];
        u[i * M + s] = (T)((1.0 + p + pe) * (double)u[i * M + s]);
    }
}

// adjoint of projectFTscalar's per-mode map chi = sum_c w_c S_c / (2 k^4 N):
// lS_c = conj(w_c) lchi / (2 k^4 N), k = 0 -> 0 (6 components out)
template <typename T>
__global__ __launch_bounds__(256) void k_adj_project_ft_scalar(const cx<T> *__restrict__ lchi, cx<T> *__restrict__ lS,
                                                               int N, int P, int64_t NK, const double *__restrict__ gk2,
                                                               const cx<double> *__restrict__ ks)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= NK) return;
    const int kx = (int)(e % P);
    const int64_t rr = e / P;
    const int ky = (int)(rr % N), kz = (int)(rr / N);
    if (kx > N / 2) return;
    const cx<T> lv = lchi[e];
    const cx<double> l = mk((double)lv.x, (double)lv.y);
    const double g0 = gk2[kx], g1 = gk2[ky], g2 = gk2[kz];
    const bool zero = kx == 0 && ky == 0 && kz == 0;
    const double nrm = zero ? 0.0 : kop_scalar_norm(g0 + g1 + g2, N);
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const cx<double> w = conj(kop_scalar_weight(c, g0, g1, g2, ks[kx], ks[ky], ks[kz]));
        const cx<double> o = (w * l) * nrm;
        lS[c * NK + e] = mk((T)o.x, (T)o.y);
    }
}

static inline unsigned nbl(int64_t n) { return (unsigned)((n + 255) / 256); }

template <typename T>
int adj_gr_kick_drift_impl(gev_pcls *p, const T *phi, const T *chi, const T *Bi, T *lx, T *lq, T *lphi, T *lchi,
                           T *lB, int nf_k, int nf_d, double a_k, double a_d, double dtau_k, double dtau_d)
{
    gev_ctx *c = p->ctx;
    AdjKD s{nf_k, nf_d, a_k, a_d, dtau_k, dtau_d, a_k * a_k * (double)c->N, a_d * a_d * (double)c->N};
    const int N = c->N;
    // tiles as the forward update's where the FP64 accumulators fit the LDS
    auto launch = [&](auto tx, auto ty, auto tz, auto nt) {
        constexpr int TX = decltype(tx)::value, TY = decltype(ty)::value, TZ = decltype(tz)::value;
        constexpr int NT = decltype(nt)::value;
        const unsigned nb = (unsigned)((int64_t)(N / TX) * (N / TY) * (N / TZ));
        hipLaunchKernelGGL((k_adj_gr_kick_drift<T, TX, TY, TZ, NT>), dim3(nb), dim3(NT), 0, c->stream, make_pv<T>(p),
                           phi, chi, Bi, lx, lq, lphi, lchi, lB, c->M, s);
    };
    using I = std::integral_constant<int, 4>;
    if (p->n > 0) {
        bool done = false;
        if constexpr (sizeof(T) == 4) {
            if (ADJ_KD_WIDE && N >= 16) {
                launch(std::integral_constant<int, 16>{}, std::integral_constant<int, 8>{},
                       std::integral_constant<int, 8>{}, std::integral_constant<int, 512>{});
                done = true;
            }
        }
        if (done) {
        } else if (N >= 8)
            launch(std::integral_constant<int, 8>{}, std::integral_constant<int, 8>{}, I{},
                   std::integral_constant<int, 256>{});
        else
            launch(I{}, I{}, I{}, std::integral_constant<int, 256>{});
    }
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T>
int adj_gr_deposit_impl(gev_pcls *p, const T *phi, const T *lT00, const T *lTij, const T *lL, T *lx, T *lq, T *lphi,
                        double a, double tij_scale)
{
    gev_ctx *c = p->ctx;
    const double dx = 1.0 / (double)c->N, m = p->mass / (dx * dx * dx);
    const int N = c->N;
    auto launch = [&](auto tx, auto ty, auto tz, auto nt) {
        constexpr int TX = decltype(tx)::value, TY = decltype(ty)::value, TZ = decltype(tz)::value;
        constexpr int NT = decltype(nt)::value;
        const unsigned nb = (unsigned)((int64_t)(N / TX) * (N / TY) * (N / TZ));
        hipLaunchKernelGGL((k_adj_gr_deposit<T, TX, TY, TZ, NT>), dim3(nb), dim3(NT), 0, c->stream, make_pv<T>(p), phi,
                           lT00, lTij, lL, lx, lq, lphi, c->M, phi ? m / a : m, tij_scale * m / a, m, a);
    };
    using I4 = std::integral_constant<int, 4>;
    using I8 = std::integral_constant<int, 8>;
    if (p->n > 0) {
        bool done = false;
        if constexpr (sizeof(T) == 4) {  // FP64 fields take twice the LDS: the smaller tile
            if (ADJ_DEP_WIDE && N >= 16) {
                launch(std::integral_constant<int, 16>{}, I8{}, I8{}, std::integral_constant<int, 256>{});
                done = true;
            }
        }
        if (done) {
        } else if (N >= 8) launch(I8{}, I8{}, I4{}, std::integral_constant<int, 256>{});
        else launch(I4{}, I4{}, I4{}, std::integral_constant<int, 256>{});
    }
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T>
int adj_source_T00_impl(gev_ctx *c, const T *phi, const T *T00, const T *lsrc, T *lT00, T *lphi, T *lchi, double bg,
                        double c1, double c2, double c3)
{
    hipLaunchKernelGGL((k_adj_source_T00<T>), dim3(nbl(c->M)), dim3(256), 0, c->stream, phi, T00, lsrc, lT00, lphi,
                       lchi, c->N, c->M, bg, c1, c2, c3);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int adj_source_Tij_phi_impl(gev_ctx *c, const T *phi, const T *lS, T *lphi)
{
    hipLaunchKernelGGL((k_adj_source_Tij_phi<T>), dim3(nbl(c->M)), dim3(256), 0, c->stream, phi, lS, lphi, c->N,
                       c->M);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int adj_link_factor_impl(gev_ctx *c, const T *phi, const T *L, T *u, T *lphi)
{
    hipLaunchKernelGGL((k_adj_link_factor_phi<T>), dim3(nbl(c->M)), dim3(256), 0, c->stream, phi, L, u, lphi, c->N,
                       c->M);
    hipLaunchKernelGGL((k_adj_link_factor_L<T>), dim3(nbl(c->M)), dim3(256), 0, c->stream, phi, u, c->N, c->M);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int adj_project_ft_scalar_impl(gev_ctx *c, const cx<T> *lchi, cx<T> *lS)
{
    hipLaunchKernelGGL((k_adj_project_ft_scalar<T>), dim3(nbl(c->NK)), dim3(256), 0, c->stream, lchi, lS, c->N, c->P,
                       c->NK, c->gk2, c->kshift);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

#define INST(T)                                                                                                     \
    template int adj_gr_kick_drift_impl<T>(gev_pcls *, const T *, const T *, const T *, T *, T *, T *, T *, T *, int, \
                                           int, double, double, double, double);                                 \
    template int adj_gr_deposit_impl<T>(gev_pcls *, const T *, const T *, const T *, const T *, T *, T *, T *,      \
                                        double, double);                                                           \
    template int adj_source_T00_impl<T>(gev_ctx *, const T *, const T *, const T *, T *, T *, T *, double, double, \
                                        double, double);                                                           \
    template int adj_source_Tij_phi_impl<T>(gev_ctx *, const T *, const T *, T *);                                 \
    template int adj_link_factor_impl<T>(gev_ctx *, const T *, const T *, T *, T *);                               \
    template int adj_project_ft_scalar_impl<T>(gev_ctx *, const cx<T> *, cx<T> *);
INST(double)
INST(float)
#undef INST

}  // namespace gev

/*
 * gev_oracle_tlm.c -- TEST INFRASTRUCTURE, NOT THE PRODUCT.
 *
 * Forward-mode (tangent-linear) derivative of the Newtonian particle-mesh run
 * that the library's adjoint differentiates in reverse:
 *
 *   IC     particle ID = cell c on the cell-centred lattice,
 *          q0 = cq N grad1(psi)(c, r = 1/2),  x0 = x_lat + cx N grad1(psi)(c, r = 1/2)
 *          (initialize_q_ic_basic / displace_pcls_ic_basic, ic_basic.hpp:65-96, 133-168)
 *   cycle  rho = om sum_p W(x_p)                       (meshprojection_project,
 *                                                       gevolution.hpp:1902-1905)
 *          phi = Poisson(rho; 4 pi G / a, 0)           (solveModifiedPoissonFT, :806-852)
 *          q  -= dtau_k a N g(x; phi)                  (update_q_Newton, :1051-1118)
 *          x  += (dtau / a_d) q                        (update_pos_Newton, :1270-1273)
 *   loss   L = sum_s (rho_final(s) - target(s))^2
 *
 * with W the 8 CIC corner weights and g_i = 1/2 sum_c w_c (phi(c + e_i) -
 * phi(c - e_i)) (GRADIENT_ORDER 2).  The tangent of a direction v of psi is
 * carried alongside: the derivative of each corner weight in the particle's
 * cell (W is piecewise multilinear, so this is exact away from cell faces),
 * the same linear Poisson solve, and the derivative of g in the particle's
 * cell.  The run returns L and dL = <grad L, v>, which a reverse-mode gradient
 * must reproduce to rounding (the dot-product test of an adjoint).
 *
 * Pinned by: central differences of L computed by this same deterministic
 * forward (tests/test_oracle_tlm.py; a dropped term of the tangent, a wrong
 * sign of a corner derivative or a missing K or C factor fails them), and by
 * single-particle finite differences of the tangent deposit and kick.
 * All loops are sequential or per-particle, so results do not depend on the
 * thread count.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_API __attribute__((visibility("default")))

/* functions from gev_oracle.c */
void or_fft_r2c(int N, const double *in, double *out);
void or_fft_c2r(int N, const double *in, double *out);
void or_poisson(int N, const double *src, double *pot, double coeff, double modif);
void or_project_T00(int N, long np, const double *pos, const double *q, double pmass, double *T00, double a,
                    const double *phi, double coeff);
double or_update_q_newton(int N, long np, const double *pos, double *q, double dtau, const double *psi,
                          const double *chi, int nfield, double a);
void or_update_pos_newton(int N, long np, double *pos, const double *q, double dtau, double a);
double or_displace_particles(int N, long np, double *pos, const double *xi, double coeff);
double or_initialize_q(int N, long np, const double *pos, double *q, const double *pot, double coeff);

static inline long tsidx(int N, long x, long y, long z)
{
    const long m = N - 1;
    return (x & m) + (long)N * ((y & m) + (long)N * (z & m));
}

static inline void tcell(int N, const double *pos, long np, long i, long c[3], double r[3])
{
    const double dx = 1.0 / (double)N;
    for (int d = 0; d < 3; d++) {
        double p = pos[d * np + i];
        long ci = (long)floor(p * (double)N);
        r[d] = (p - (double)ci * dx) / dx;
        c[d] = ci;
    }
}

/* CIC weight of corner b = (bx, by, bz) and its three partial derivatives in r */
static inline void corner_weight(int b, const double r[3], double *w, double dw[3])
{
    const int bb[3] = {(b >> 2) & 1, (b >> 1) & 1, b & 1};
    double f[3], s[3];
    for (int d = 0; d < 3; d++) {
        f[d] = bb[d] ? r[d] : 1.0 - r[d];
        s[d] = bb[d] ? 1.0 : -1.0;
    }
    *w = f[0] * f[1] * f[2];
    dw[0] = s[0] * f[1] * f[2];
    dw[1] = f[0] * s[1] * f[2];
    dw[2] = f[0] * f[1] * s[2];
}

/* drho += m' sum_b (dW_b/dr . N dpos) at the particle's 8 corners */
OR_API void or_tlm_deposit_mass(int N, long np, const double *pos, const double *dpos, double mprime, double *drho)
{
    for (long i = 0; i < np; i++) {
        long c[3]; double r[3];
        tcell(N, pos, np, i, c, r);
        const double dr[3] = {N * dpos[i], N * dpos[np + i], N * dpos[2 * np + i]};
        for (int b = 0; b < 8; b++) {
            double w, dw[3];
            corner_weight(b, r, &w, dw);
            drho[tsidx(N, c[0] + ((b >> 2) & 1), c[1] + ((b >> 1) & 1), c[2] + (b & 1))] +=
                mprime * (dw[0] * dr[0] + dw[1] * dr[1] + dw[2] * dr[2]);
        }
    }
}

/* dq -= K (dg/dr . N dpos + g(x; dphi)), K = dtau a N, g the CIC-weighted
 * half central difference at the particle */
OR_API void or_tlm_kick_newton(int N, long np, const double *pos, const double *dpos, double *dq, const double *phi,
                               const double *dphi, double K)
{
    for (long i = 0; i < np; i++) {
        long c[3]; double r[3];
        tcell(N, pos, np, i, c, r);
        const double dr[3] = {N * dpos[i], N * dpos[np + i], N * dpos[2 * np + i]};
        double dg[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < 8; b++) {
            const long x = c[0] + ((b >> 2) & 1), y = c[1] + ((b >> 1) & 1), z = c[2] + (b & 1);
            double w, dw[3];
            corner_weight(b, r, &w, dw);
            const double wd = dw[0] * dr[0] + dw[1] * dr[1] + dw[2] * dr[2];
            const double D[3] = {0.5 * (phi[tsidx(N, x + 1, y, z)] - phi[tsidx(N, x - 1, y, z)]),
                                 0.5 * (phi[tsidx(N, x, y + 1, z)] - phi[tsidx(N, x, y - 1, z)]),
                                 0.5 * (phi[tsidx(N, x, y, z + 1)] - phi[tsidx(N, x, y, z - 1)])};
            const double dD[3] = {0.5 * (dphi[tsidx(N, x + 1, y, z)] - dphi[tsidx(N, x - 1, y, z)]),
                                  0.5 * (dphi[tsidx(N, x, y + 1, z)] - dphi[tsidx(N, x, y - 1, z)]),
                                  0.5 * (dphi[tsidx(N, x, y, z + 1)] - dphi[tsidx(N, x, y, z - 1)])};
            for (int d = 0; d < 3; d++) dg[d] += wd * D[d] + w * dD[d];
        }
        for (int d = 0; d < 3; d++) dq[d * np + i] -= K * dg[d];
    }
}

static void poisson_real(int N, const double *src, double *pot, double coeff, double *F)
{
    or_fft_r2c(N, src, F);
    or_poisson(N, F, F, coeff, 0.0);
    or_fft_c2r(N, F, pot);
}

/* The run described at the top.  sched: steps x {a, a_d, dtau_k, dtau}.
 * v may be NULL (forward only; *dL untouched).  rho_out (N^3) may be NULL.
 * pos_out / q_out (3 x N^3, by particle ID) may be NULL.  Returns L. */
OR_API double or_newton_run_tlm(int N, const double *psi, const double *v, const double *target, int steps,
                                const double *sched, double fourpiG, double omega_m, double cx, double cq,
                                double *rho_out, double *pos_out, double *q_out, double *dL)
{
    const long M = (long)N * N * N, np = M;
    const long nk = 2L * (N / 2 + 1) * N * N;
    double *pos = (double *)calloc(3 * np, sizeof(double));
    double *q = (double *)calloc(3 * np, sizeof(double));
    double *dpos = (double *)calloc(3 * np, sizeof(double));
    double *dq = (double *)calloc(3 * np, sizeof(double));
    double *rho = (double *)calloc(M, sizeof(double));
    double *drho = (double *)calloc(M, sizeof(double));
    double *phi = (double *)calloc(M, sizeof(double));
    double *dphi = (double *)calloc(M, sizeof(double));
    double *F = (double *)calloc(nk, sizeof(double));
    const double pmass = omega_m / (double)M;   /* m' = pmass / dx^3 = omega_m */
    /* cell-centred lattice, ID = cell */
    for (long s = 0; s < M; s++) {
        const long x = s % N, y = (s / N) % N, z = s / ((long)N * N);
        pos[s] = (x + 0.5) / N;
        pos[np + s] = (y + 0.5) / N;
        pos[2 * np + s] = (z + 0.5) / N;
    }
    if (v) {  /* both maps are linear in psi and read it at the lattice point */
        or_initialize_q(N, np, pos, dq, v, -cq);
        or_initialize_q(N, np, pos, dpos, v, -cx);
    }
    or_initialize_q(N, np, pos, q, psi, -cq);
    or_displace_particles(N, np, pos, psi, cx);
    for (int k = 0; k < steps; k++) {
        const double a = sched[4 * k], a_d = sched[4 * k + 1], dtk = sched[4 * k + 2], dt = sched[4 * k + 3];
        memset(rho, 0, sizeof(double) * M);
        or_project_T00(N, np, pos, q, pmass, rho, 1.0, NULL, 1.0);
        poisson_real(N, rho, phi, fourpiG / a, F);
        if (v) {
            memset(drho, 0, sizeof(double) * M);
            or_tlm_deposit_mass(N, np, pos, dpos, omega_m, drho);
            poisson_real(N, drho, dphi, fourpiG / a, F);
            or_tlm_kick_newton(N, np, pos, dpos, dq, phi, dphi, dtk * a * N);
            const double C = dt / a_d;
            for (long j = 0; j < 3 * np; j++) dpos[j] += C * dq[j];
        }
        or_update_q_newton(N, np, pos, q, dtk, phi, NULL, 1, a);
        or_update_pos_newton(N, np, pos, q, dt, a_d);
    }
    memset(rho, 0, sizeof(double) * M);
    or_project_T00(N, np, pos, q, pmass, rho, 1.0, NULL, 1.0);
    double L = 0.0;
    for (long s = 0; s < M; s++) L += (rho[s] - target[s]) * (rho[s] - target[s]);
    if (v) {
        memset(drho, 0, sizeof(double) * M);
        or_tlm_deposit_mass(N, np, pos, dpos, omega_m, drho);
        double d = 0.0;
        for (long s = 0; s < M; s++) d += 2.0 * (rho[s] - target[s]) * drho[s];
        *dL = d;
    }
    if (rho_out) memcpy(rho_out, rho, sizeof(double) * M);
    if (pos_out) memcpy(pos_out, pos, sizeof(double) * 3 * np);
    if (q_out) memcpy(q_out, q, sizeof(double) * 3 * np);
    free(pos); free(q); free(dpos); free(dq); free(rho); free(drho); free(phi); free(dphi); free(F);
    return L;
}

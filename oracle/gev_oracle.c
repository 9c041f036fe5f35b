/*
 * gev_oracle.c -- TEST INFRASTRUCTURE, NOT THE PRODUCT.
 *
 * CPU FP64 restatement of gevolution-2.0's particle-mesh (PM) step, used only
 * as the correctness checker (tests/, __graft_entry__.smoke()) and as the
 * timed CPU baseline (bench.py cpu_baseline leg).  Nothing in the shipped
 * library links, loads or calls this file.
 *
 * Conventions (shared with the HIP library so results can be compared):
 *   - real field: double[N^3], site (x,y,z) at x + N*(y + N*z), periodic, no halo;
 *     multi-component fields are component-major (component c at offset c*N^3);
 *     symmetric tensors store (00,01,02,11,12,22) in that order.
 *   - Fourier field: complex (re,im interleaved) [(N/2+1) * N * N], mode
 *     (kx,ky,kz) at kx + (N/2+1)*(ky + N*kz); kx in [0,N/2], ky,kz in [0,N).
 *     Forward transform uses exp(-2 pi i k.x/N), both directions unnormalised
 *     (the reference divides by N^3 explicitly, gevolution.hpp:817).
 *   - particles: SoA, pos[c*np+i] in box units [0,1), q[c*np+i] = q/m.
 *   - N must be a power of two.
 *
 * What pins each function (tests/test_oracle_*.py, tests/test_abi.py); every
 * check is fixed by mathematics or data independent of this file, and every
 * function below is pinned (a swapped weight, corner, factor or conjugate in
 * any of them fails at least one check):
 *   or_fft_r2c / or_fft_c2r   brute-force DFT (N=4,8), numpy.fft fixtures
 *                             (N=8,16), Parseval + round trip, c2r ignoring
 *                             Im of the kx=0 / N/2 planes
 *   or_poisson                7-point Laplacian of the result = source (+modif)
 *   or_project_ft_scalar      longitudinal / pure-trace tensors, closed forms;
 *                             on a random tensor the real-space identity
 *                             Lap^2 chi = (3 DD S - Lap tr S) / (2 N^3) with the
 *                             staggered double divergence DD
 *   or_project_ft_vector      inverse of (-Lap + modif) on transverse input,
 *                             D- transversality of the output
 *   or_evolve_ft_vector       on a random tensor -Lap dB = -(2 a^2 dtau / N)
 *                             T(div S), with the staggered divergence onto the
 *                             links and the discrete Helmholtz projection T;
 *                             trace/longitudinal inputs leave B unchanged
 *   or_prepare_source_T00/Tij hand-evaluated stencils along every axis and axis
 *                             pair, phi = 0 limits
 *   or_project_T00/T0i/Tij    sum rules (energy, momentum, stress), single
 *                             particle weights per site and per link, the
 *                             per-corner / per-link phi factors on random phi,
 *                             linear-phi closed forms (incl. the quadratic T00
 *                             term), and the point-particle stress tensor of
 *                             the weak-field metric expanded to O(phi^2) (T00)
 *                             and O(phi) (Tij)
 *   or_update_q[_newton]      random potentials against tensor products of 1-D
 *                             lattice stencils (central and link gradients, all
 *                             nine B_c d_d terms), separable closed forms, the
 *                             weak-field geodesic equation to O(phi grad phi)
 *   or_update_pos[_newton]    free streaming, random staggered B, and the
 *                             geodesic velocity to O(phi, chi)
 *   or_displace_particles /
 *   or_initialize_q           closed forms on single modes
 *   or_hconf / or_rk4bg /
 *   or_particle_horizon       Einstein-de Sitter closed forms
 *   gev_oracle_step.c         Philox known-answer vectors, natural spline vs
 *                             scipy, P(k) single-mode closed form, IC P(k) vs
 *                             input table, linear growth D(a) (Newton and GR),
 *                             momentum conservation of the step
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_API __attribute__((visibility("default")))

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) { cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cplx cconj(cplx a) { cplx r = {a.re, -a.im}; return r; }
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx cscale(cplx a, double s) { cplx r = {a.re * s, a.im * s}; return r; }

static inline long wrapi(long i, int N) { return i & (long)(N - 1); }
static inline long sidx(int N, long x, long y, long z)
{
    return wrapi(x, N) + (long)N * (wrapi(y, N) + (long)N * wrapi(z, N));
}

OR_API int or_version(void) { return 1; }

OR_API void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

OR_API int or_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * FFT (stands in for LATfield2 PlanFFT / cuFFT, main.cu:343-357,742-948).
 * Iterative radix-2 Cooley-Tukey on one line; 3D transforms line by line.
 * Pinned by: brute-force DFT at N=4,8 and numpy.fft.rfftn fixtures at N=8,16
 * (tests/golden), Parseval, c2r(r2c(x)) = N^3 x.
 * ------------------------------------------------------------------------ */
/* tw[k] = exp(-2 pi i k / n), k < n/2 */
static cplx *twiddles(int n)
{
    cplx *tw = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
    for (int k = 0; k < n / 2; k++) {
        double ang = 2.0 * M_PI * (double)k / (double)n;
        tw[k].re = cos(ang);
        tw[k].im = -sin(ang);
    }
    return tw;
}

static void fft_line(int n, cplx *v, int sign, const cplx *tw)
{
    /* bit reversal */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = v[i]; v[i] = v[j]; v[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int k = 0; k < half; k++) {
            cplx w = tw[k * step];
            if (sign > 0) w.im = -w.im;
            for (int i = 0; i < n; i += len) {
                cplx u = v[i + k];
                cplx t = cmul(w, v[i + k + half]);
                v[i + k] = cadd(u, t);
                v[i + k + half] = csub(u, t);
            }
        }
    }
}

/* complex passes along y (axis 1) and z (axis 2) of a (P x N x N) array */
static void fft_pass_yz(int N, cplx *a, int axis, int sign)
{
    const long P = N / 2 + 1;
    const long stride = (axis == 1) ? P : P * N;
    const long nouter = P * N; /* number of lines */
    cplx *tw = twiddles(N);
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
        for (long l = 0; l < nouter; l++) {
            long kx = l % P, o = l / P; /* o = the other axis index */
            long base = (axis == 1) ? kx + P * N * o : kx + P * o;
            for (int j = 0; j < N; j++) buf[j] = a[base + j * stride];
            fft_line(N, buf, sign, tw);
            for (int j = 0; j < N; j++) a[base + j * stride] = buf[j];
        }
        free(buf);
    }
    free(tw);
}

OR_API void or_fft_r2c(int N, const double *in, double *outd)
{
    cplx *out = (cplx *)outd;
    const long P = N / 2 + 1;
    cplx *tw = twiddles(N);
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)N * N; l++) {
            for (int x = 0; x < N; x++) { buf[x].re = in[l * N + x]; buf[x].im = 0.0; }
            fft_line(N, buf, -1, tw);
            for (int k = 0; k < P; k++) out[l * P + k] = buf[k];
        }
        free(buf);
    }
    free(tw);
    fft_pass_yz(N, out, 1, -1);
    fft_pass_yz(N, out, 2, -1);
}

/* Unnormalised inverse.  The imaginary parts of the kx=0 and kx=N/2 entries
 * are ignored on the final real axis, as every c2r (FFTW, cuFFT, numpy) does. */
OR_API void or_fft_c2r(int N, const double *ind, double *out)
{
    const long P = N / 2 + 1;
    const long nk = P * N * N;
    cplx *t = (cplx *)malloc(sizeof(cplx) * nk);
    memcpy(t, ind, sizeof(cplx) * nk);
    fft_pass_yz(N, t, 2, +1);
    fft_pass_yz(N, t, 1, +1);
    cplx *tw = twiddles(N);
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)N * N; l++) {
            const cplx *h = t + l * P;
            buf[0].re = h[0].re; buf[0].im = 0.0;
            buf[N / 2].re = h[N / 2].re; buf[N / 2].im = 0.0;
            for (int k = 1; k < N / 2; k++) { buf[k] = h[k]; buf[N - k] = cconj(h[k]); }
            fft_line(N, buf, +1, tw);
            for (int x = 0; x < N; x++) out[l * N + x] = buf[x].re;
        }
        free(buf);
    }
    free(tw);
    free(t);
}

/* ------------------------------------------------------------------------
 * k-space helpers (gevolution.hpp:269-274, 827-831)
 * gridk2[i] = (2 N sin(pi i/N))^2 ; kshift[i] = 2 N sin(pi i/N) e^{-i pi i/N}
 * ------------------------------------------------------------------------ */
static void make_k_tables(int N, double *gridk2, cplx *kshift)
{
    for (int i = 0; i < N; i++) {
        double s = 2.0 * (double)N * sin(M_PI * (double)i / (double)N);
        if (kshift) {
            kshift[i].re = s * cos(M_PI * (double)i / (double)N);
            kshift[i].im = -s * sin(M_PI * (double)i / (double)N);
        }
        gridk2[i] = s * s;
    }
}

/* solveModifiedPoissonFT, gevolution.hpp:806-852.
 * pot = src * (-coeff/N^3) / (k2 + modif); k=0 -> 0 when modif == 0.
 * Pinned by: single-mode eigenvalue test against the real-space 7-point
 * Laplacian (tests), linearity, k=0 handling. */
OR_API void or_poisson(int N, const double *srcd, double *potd, double coeff, double modif)
{
    const cplx *src = (const cplx *)srcd;
    cplx *pot = (cplx *)potd;
    const long P = N / 2 + 1;
    double *gk2 = (double *)malloc(sizeof(double) * N);
    make_k_tables(N, gk2, NULL);
    coeff /= -((double)N * (double)N * (double)N);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < (long)N * N; l++) {
        long ky = l % N, kz = l / N;
        for (long kx = 0; kx < P; kx++) {
            long i = kx + P * l;
            double d = gk2[kx] + gk2[ky] + gk2[kz] + modif;
            if (modif == 0.0 && kx == 0 && ky == 0 && kz == 0) { pot[i].re = 0.0; pot[i].im = 0.0; continue; }
            pot[i] = cscale(src[i], coeff / d);
        }
    }
    free(gk2);
}

/* projectFTscalar, gevolution.hpp:251-340.  Sij holds 6 complex components
 * (00,01,02,11,12,22) at stride (N/2+1)N^2.  Pinned by: pure-trace source
 * gives zero, closed-form single longitudinal mode (tests). */
OR_API void or_project_ft_scalar(int N, const double *Sd, double *chid, int add)
{
    const cplx *S = (const cplx *)Sd;
    cplx *chi = (cplx *)chid;
    const long P = N / 2 + 1, nk = P * N * N;
    double *gk2 = (double *)malloc(sizeof(double) * N);
    cplx *ks = (cplx *)malloc(sizeof(cplx) * N);
    make_k_tables(N, gk2, ks);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < (long)N * N; l++) {
        long ky = l % N, kz = l / N;
        for (long kx = 0; kx < P; kx++) {
            long i = kx + P * l;
            double k2x = gk2[kx], k2y = gk2[ky], k2z = gk2[kz];
            double k2 = k2x + k2y + k2z;
            cplx acc = cadd(cadd(cscale(S[i + 0 * nk], k2y + k2z - 2.0 * k2x),
                                 cscale(S[i + 3 * nk], k2x + k2z - 2.0 * k2y)),
                            cscale(S[i + 5 * nk], k2x + k2y - 2.0 * k2z));
            acc = csub(acc, cscale(cmul(cmul(ks[kx], ks[ky]), S[i + 1 * nk]), 6.0));
            acc = csub(acc, cscale(cmul(cmul(ks[kx], ks[kz]), S[i + 2 * nk]), 6.0));
            acc = csub(acc, cscale(cmul(cmul(ks[ky], ks[kz]), S[i + 4 * nk]), 6.0));
            cplx r = cscale(acc, 1.0 / (2.0 * k2 * k2 * (double)N));
            if (kx == 0 && ky == 0 && kz == 0) continue; /* set to zero below */
            if (add) chi[i] = cadd(chi[i], r); else chi[i] = r;
        }
    }
    chi[0].re = 0.0; chi[0].im = 0.0; /* gevolution.hpp:330-333 */
    free(gk2); free(ks);
}

/* projectFTvector, gevolution.hpp:449-515.  Pinned by: transversality
 * ks^*.B = 0 ... (conj convention as in the reference), longitudinal input
 * gives zero, transverse input scaled by 4 coeff/(k2+modif). */
OR_API void or_project_ft_vector(int N, const double *Sd, double *Bd, double coeff, double modif)
{
    const cplx *S = (const cplx *)Sd;
    cplx *B = (cplx *)Bd;
    const long P = N / 2 + 1, nk = P * N * N;
    double *gk2 = (double *)malloc(sizeof(double) * N);
    cplx *ks = (cplx *)malloc(sizeof(cplx) * N);
    make_k_tables(N, gk2, ks);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < (long)N * N; l++) {
        long ky = l % N, kz = l / N;
        for (long kx = 0; kx < P; kx++) {
            long i = kx + P * l;
            double k2 = gk2[kx] + gk2[ky] + gk2[kz];
            cplx s0 = S[i], s1 = S[i + nk], s2 = S[i + 2 * nk];
            cplx tmp = cscale(cadd(cadd(cmul(ks[kx], s0), cmul(ks[ky], s1)), cmul(ks[kz], s2)), 1.0 / k2);
            double f = 4.0 * coeff / (k2 + modif);
            cplx b0 = cscale(csub(s0, cmul(cconj(ks[kx]), tmp)), f);
            cplx b1 = cscale(csub(s1, cmul(cconj(ks[ky]), tmp)), f);
            cplx b2 = cscale(csub(s2, cmul(cconj(ks[kz]), tmp)), f);
            if (kx == 0 && ky == 0 && kz == 0) { b0.re = b0.im = b1.re = b1.im = b2.re = b2.im = 0.0; }
            B[i] = b0; B[i + nk] = b1; B[i + 2 * nk] = b2;
        }
    }
    free(gk2); free(ks);
}

/* evolveFTvector, gevolution.hpp:359-429 (parabolic vector method).
 * Pinned by: transversality of the increment, pure-trace source gives zero,
 * and -Lap dB = -(2 a2dtau / N) T(div S) on a random tensor (tests). */
OR_API void or_evolve_ft_vector(int N, const double *Sd, double *Bd, double a2dtau)
{
    const cplx *S = (const cplx *)Sd;
    cplx *B = (cplx *)Bd;
    const long P = N / 2 + 1, nk = P * N * N;
    double *gk2 = (double *)malloc(sizeof(double) * N);
    cplx *ks = (cplx *)malloc(sizeof(cplx) * N);
    make_k_tables(N, gk2, ks);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < (long)N * N; l++) {
        long ky = l % N, kz = l / N;
        for (long kx = 0; kx < P; kx++) {
            long i = kx + P * l;
            if (kx == 0 && ky == 0 && kz == 0) {
                for (int c = 0; c < 3; c++) { B[i + c * nk].re = 0.0; B[i + c * nk].im = 0.0; }
                continue;
            }
            double g0 = gk2[kx], g1 = gk2[ky], g2 = gk2[kz];
            cplx k0 = ks[kx], k1 = ks[ky], k2c = ks[kz];
            double k4 = (g0 + g1 + g2); k4 *= k4;
            cplx pre = {0.0, -2.0 * a2dtau / k4};
            cplx S00 = S[i], S01 = S[i + nk], S02 = S[i + 2 * nk], S11 = S[i + 3 * nk], S12 = S[i + 4 * nk], S22 = S[i + 5 * nk];
            cplx t0 = cadd(cmul(cconj(k0), csub(csub(csub(cscale(S00, g1 + g2), cscale(S11, g1)), cscale(S22, g2)), cscale(cmul(cmul(k1, k2c), S12), 2.0))),
                           cscale(cadd(cmul(k1, S01), cmul(k2c, S02)), g1 + g2 - g0));
            cplx t1 = cadd(cmul(cconj(k1), csub(csub(csub(cscale(S11, g0 + g2), cscale(S00, g0)), cscale(S22, g2)), cscale(cmul(cmul(k0, k2c), S02), 2.0))),
                           cscale(cadd(cmul(k0, S01), cmul(k2c, S12)), g0 + g2 - g1));
            cplx t2 = cadd(cmul(cconj(k2c), csub(csub(csub(cscale(S22, g0 + g1), cscale(S00, g0)), cscale(S11, g1)), cscale(cmul(cmul(k0, k1), S01), 2.0))),
                           cscale(cadd(cmul(k0, S02), cmul(k1, S12)), g0 + g1 - g2));
            B[i] = cadd(B[i], cmul(pre, t0));
            B[i + nk] = cadd(B[i + nk], cmul(pre, t1));
            B[i + 2 * nk] = cadd(B[i + 2 * nk], cmul(pre, t2));
        }
    }
    free(gk2); free(ks);
}

/* ------------------------------------------------------------------------
 * Real-space source construction
 * ------------------------------------------------------------------------ */

/* prepareFTsource (T00 form) with PHINONLINEAR, gevolution.hpp:160-218.
 * result may alias source.  Returns the sum of the source before the update.
 * Pinned by: hand-evaluated stencil on random 4^3 fields (tests), phi=chi=0
 * reduces to c2 (s - bg). */
OR_API double or_prepare_source_T00(int N, const double *phi, const double *chi, const double *source, double bg,
                                    double *result, double c1, double c2, double c3)
{
    const long M = (long)N * N * N;
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
    for (long l = 0; l < (long)N * N; l++) {
        long y = l % N, z = l / N;
        for (long x = 0; x < N; x++) {
            long i = x + N * l;
            double s = source[i];
            double p = phi[i];
            sum += s;
            double r = c2 * (s - bg);
            r *= 1.0 - 2.0 * p * (1.0 - p);
            double d0 = phi[sidx(N, x - 1, y, z)] - phi[sidx(N, x + 1, y, z)];
            double d1 = phi[sidx(N, x, y - 1, z)] - phi[sidx(N, x, y + 1, z)];
            double d2 = phi[sidx(N, x, y, z - 1)] - phi[sidx(N, x, y, z + 1)];
            r += 0.125 * d0 * d0;
            r += 0.125 * d1 * d1;
            r += 0.125 * d2 * d2;
            r += (c3 * (1.0 - 3.0 * p) - c1) * p - c3 * chi[i];
            result[i] = r;
        }
    }
    (void)M;
    return sum;
}

/* prepareFTsource (Tij form) with PHINONLINEAR, gevolution.hpp:64-138.
 * Sij may alias Tij (6 components each).  Pinned by: hand-evaluated stencil
 * (tests); phi=0 reduces to coeff*Tij. */
OR_API void or_prepare_source_Tij(int N, const double *phi, const double *Tij, double *Sij, double coeff)
{
    const long M = (long)N * N * N;
#pragma omp parallel for schedule(static)
    for (long l = 0; l < (long)N * N; l++) {
        long y = l % N, z = l / N;
        for (long x = 0; x < N; x++) {
            long i = x + N * l;
            double p = phi[i];
            double pp[3] = {phi[sidx(N, x + 1, y, z)], phi[sidx(N, x, y + 1, z)], phi[sidx(N, x, y, z + 1)]};
            double pm[3] = {phi[sidx(N, x - 1, y, z)], phi[sidx(N, x, y - 1, z)], phi[sidx(N, x, y, z - 1)]};
            double p01 = phi[sidx(N, x + 1, y + 1, z)];
            double p02 = phi[sidx(N, x + 1, y, z + 1)];
            double p12 = phi[sidx(N, x, y + 1, z + 1)];
            double T00 = Tij[i], T01 = Tij[i + M], T02 = Tij[i + 2 * M], T11 = Tij[i + 3 * M], T12 = Tij[i + 4 * M], T22 = Tij[i + 5 * M];
            Sij[i] = 0.5 * (pp[0] - pm[0]) * (pp[0] - pm[0]) + coeff * T00;
            Sij[i + 3 * M] = 0.5 * (pp[1] - pm[1]) * (pp[1] - pm[1]) + coeff * T11;
            Sij[i + 5 * M] = 0.5 * (pp[2] - pm[2]) * (pp[2] - pm[2]) + coeff * T22;
            Sij[i + M] = pp[0] * pp[1] - p * p01 + 0.5 * (p * p - pp[0] * pp[0] - pp[1] * pp[1] + p01 * p01) + coeff * T01;
            Sij[i + 2 * M] = pp[0] * pp[2] - p * p02 + 0.5 * (p * p - pp[0] * pp[0] - pp[2] * pp[2] + p02 * p02) + coeff * T02;
            Sij[i + 4 * M] = pp[1] * pp[2] - p * p12 + 0.5 * (p * p - pp[1] * pp[1] - pp[2] * pp[2] + p12 * p12) + coeff * T12;
        }
    }
}

/* ------------------------------------------------------------------------
 * Particle helpers.  LATfield2's dispatch, as its call sites use it
 * (gevolution.hpp:1339,1358 and Particles_gevolution.hpp:517-528): the particle
 * sits in cell c = floor(pos N) and ref_dist = (pos - c dx)/dx in [0,1).
 * ------------------------------------------------------------------------ */
static inline void cell_of(int N, const double *pos, long np, long i, long c[3], double r[3])
{
    const double dx = 1.0 / (double)N;
    for (int d = 0; d < 3; d++) {
        double p = pos[d * np + i];
        long ci = (long)floor(p * (double)N);
        r[d] = (p - (double)ci * dx) / dx;
        c[d] = ci;
    }
}

/* Deterministic parallel deposit: particles are bucketed by z-plane (stable)
 * and planes of equal parity are processed concurrently; a particle in plane
 * cz writes planes cz and cz+1 only, so same-parity planes never collide. */
typedef void (*deposit_fn)(int N, long np, long i, const double *pos, const double *q, void *ctx);

static void deposit_planes(int N, long np, const double *pos, const double *q, deposit_fn fn, void *ctx)
{
    long *cnt = (long *)calloc(N + 1, sizeof(long));
    long *order = (long *)malloc(sizeof(long) * (np > 0 ? np : 1));
    int *plane = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
    for (long i = 0; i < np; i++) {
        long c = (long)floor(pos[2 * np + i] * (double)N);
        if (c < 0) c = 0;
        if (c >= N) c = N - 1;
        plane[i] = (int)c;
        cnt[c + 1]++;
    }
    for (int z = 0; z < N; z++) cnt[z + 1] += cnt[z];
    long *cur = (long *)malloc(sizeof(long) * N);
    memcpy(cur, cnt, sizeof(long) * N);
    for (long i = 0; i < np; i++) order[cur[plane[i]]++] = i;
    for (int parity = 0; parity < 2; parity++) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int z = parity; z < N; z += 2) {
            for (long j = cnt[z]; j < cnt[z + 1]; j++) fn(N, np, order[j], pos, q, ctx);
        }
    }
    free(cnt); free(order); free(plane); free(cur);
}

typedef struct { double *out; const double *phi; double a, mass; } dep_ctx;

/* particle_T00_project host branch, gevolution.hpp:1729-1790 (GR, nfield=2)
 * and the nfield=1 branch (:1819-1829), which is also the Newtonian
 * meshprojection_project mass density (gevolution.hpp:1902-1905).
 * Pinned by: one-particle exact weights and per-corner phi polynomial, sum
 * rule sum(T00) = sum m'(e+...), linear-phi closed form of the phi^2 term, the
 * point-particle a^3 T00 of the weak-field metric to O(phi^2),
 * uniform lattice at rest -> uniform field (tests). */
static void dep_T00(int N, long np, long i, const double *pos, const double *q, void *vctx)
{
    dep_ctx *c = (dep_ctx *)vctx;
    long cc[3]; double r[3];
    cell_of(N, pos, np, i, cc, r);
    double w[8];
    for (int b = 0; b < 8; b++) {
        int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
        w[b] = (bx ? r[0] : 1.0 - r[0]) * (by ? r[1] : 1.0 - r[1]) * (bz ? r[2] : 1.0 - r[2]);
    }
    double mass = c->mass;
    if (c->phi) {
        double a = c->a;
        double q0 = q[i], q1 = q[np + i], q2 = q[2 * np + i];
        double f2 = q0 * q0 + q1 * q1 + q2 * q2;
        double e = sqrt(f2 + a * a);
        double f = 3.0 * e + f2 / e;
        f2 = (0.5 * f * f + f2 / (1.0 + f2 / a / a)) / e;
        mass /= a;
        for (int b = 0; b < 8; b++) {
            long s = sidx(N, cc[0] + ((b >> 2) & 1), cc[1] + ((b >> 1) & 1), cc[2] + (b & 1));
            double p = c->phi[s];
            c->out[s] += w[b] * (e + f * p + f2 * p * p) * mass;
        }
    } else {
        for (int b = 0; b < 8; b++) {
            long s = sidx(N, cc[0] + ((b >> 2) & 1), cc[1] + ((b >> 1) & 1), cc[2] + (b & 1));
            c->out[s] += w[b] * mass;
        }
    }
}

/* out += T00 (caller zeroes).  mass = particle mass; coeff as in the reference. */
OR_API void or_project_T00(int N, long np, const double *pos, const double *q, double pmass, double *T00,
                           double a, const double *phi, double coeff)
{
    double dx = 1.0 / (double)N;
    dep_ctx c = {T00, phi, a, pmass * coeff / (dx * dx * dx)};
    deposit_planes(N, np, pos, q, dep_T00, &c);
}

/* particle_T0i_project host branch, gevolution.hpp:2031-2105 (12 staggered
 * deposits, component i lives on the link x + e_i/2).  Pinned by: total
 * momentum sum rule, single-particle exact weights and per-link phi factors,
 * the linear-phi sum rule at the link midpoints (tests). */
static void dep_T0i(int N, long np, long i, const double *pos, const double *q, void *vctx)
{
    dep_ctx *c = (dep_ctx *)vctx;
    long cc[3]; double r[3];
    cell_of(N, pos, np, i, cc, r);
    const long M = (long)N * N * N;
    double *o = c->out;
    double ph[8];
    for (int b = 0; b < 8; b++) {
        long s = sidx(N, cc[0] + ((b >> 2) & 1), cc[1] + ((b >> 1) & 1), cc[2] + (b & 1));
        ph[b] = c->phi ? c->phi[s] : 0.0;
    }
    double m = c->mass;
    double v0 = q[i], v1 = q[np + i], v2 = q[2 * np + i];
    long x = cc[0], y = cc[1], z = cc[2];
    double g[12];
    if (c->phi) {
        g[0] = (1.0 - r[1]) * (1.0 - r[2]) * (1.0 + ph[0] + ph[4]) * m * v0;
        g[1] = (1.0 - r[0]) * (1.0 - r[2]) * (1.0 + ph[0] + ph[2]) * m * v1;
        g[2] = (1.0 - r[0]) * (1.0 - r[1]) * (1.0 + ph[0] + ph[1]) * m * v2;
        g[3] = r[0] * (1.0 - r[2]) * (1.0 + ph[4] + ph[6]) * m * v1;
        g[4] = r[0] * (1.0 - r[1]) * (1.0 + ph[4] + ph[5]) * m * v2;
        g[5] = r[1] * (1.0 - r[2]) * (1.0 + ph[2] + ph[6]) * m * v0;
        g[6] = r[1] * (1.0 - r[0]) * (1.0 + ph[2] + ph[3]) * m * v2;
        g[7] = r[2] * (1.0 - r[1]) * (1.0 + ph[1] + ph[5]) * m * v0;
        g[8] = r[2] * (1.0 - r[0]) * (1.0 + ph[1] + ph[3]) * m * v1;
        g[9] = r[1] * r[2] * (1.0 + ph[3] + ph[7]) * m * v0;
        g[10] = r[0] * r[2] * (1.0 + ph[5] + ph[7]) * m * v1;
        g[11] = r[0] * r[1] * (1.0 + ph[6] + ph[7]) * m * v2;
    } else {
        g[0] = (1.0 - r[1]) * (1.0 - r[2]) * m * v0;
        g[1] = (1.0 - r[0]) * (1.0 - r[2]) * m * v1;
        g[2] = (1.0 - r[0]) * (1.0 - r[1]) * m * v2;
        g[3] = r[0] * (1.0 - r[2]) * m * v1;
        g[4] = r[0] * (1.0 - r[1]) * m * v2;
        g[5] = r[1] * (1.0 - r[2]) * m * v0;
        g[6] = r[1] * (1.0 - r[0]) * m * v2;
        g[7] = r[2] * (1.0 - r[1]) * m * v0;
        g[8] = r[2] * (1.0 - r[0]) * m * v1;
        g[9] = r[1] * r[2] * m * v0;
        g[10] = r[0] * r[2] * m * v1;
        g[11] = r[0] * r[1] * m * v2;
    }
    o[sidx(N, x, y, z)] += g[0];
    o[M + sidx(N, x, y, z)] += g[1];
    o[2 * M + sidx(N, x, y, z)] += g[2];
    o[M + sidx(N, x + 1, y, z)] += g[3];
    o[2 * M + sidx(N, x + 1, y, z)] += g[4];
    o[sidx(N, x, y + 1, z)] += g[5];
    o[2 * M + sidx(N, x, y + 1, z)] += g[6];
    o[sidx(N, x, y, z + 1)] += g[7];
    o[M + sidx(N, x, y, z + 1)] += g[8];
    o[sidx(N, x, y + 1, z + 1)] += g[9];
    o[M + sidx(N, x + 1, y, z + 1)] += g[10];
    o[2 * M + sidx(N, x + 1, y + 1, z)] += g[11];
}

OR_API void or_project_T0i(int N, long np, const double *pos, const double *q, double pmass, double *T0i,
                           const double *phi, double coeff)
{
    double dx = 1.0 / (double)N;
    dep_ctx c = {T0i, phi, 1.0, pmass * coeff / (dx * dx * dx)};
    deposit_planes(N, np, pos, q, dep_T0i, &c);
}

/* particle_Tij_project host branch without CIC_PROJECT_TIJ,
 * gevolution.hpp:2582-2604, 2681-2775, 2778-2812, 2957-2962: diagonal parts
 * by full CIC (24 deposits), off-diagonal ij on the edge x + (e_i+e_j)/2
 * split over the remaining direction (6 deposits).  Pinned by: trace sum rule
 * sum(Tii) = sum f m' q^2/e, single-particle exact weights with and without
 * phi, the linear-phi sum rule, the point-particle a^3 Tij of the weak-field
 * metric to O(phi) (tests). */
static void dep_Tij(int N, long np, long i, const double *pos, const double *q, void *vctx)
{
    dep_ctx *c = (dep_ctx *)vctx;
    long cc[3]; double r[3];
    cell_of(N, pos, np, i, cc, r);
    const long M = (long)N * N * N;
    double *o = c->out;
    double a = c->a;
    double v[3] = {q[i], q[np + i], q[2 * np + i]};
    double f = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double e = sqrt(f + a * a);
    long x = cc[0], y = cc[1], z = cc[2];
    if (c->phi) {
        const double *P = c->phi;
        f = 4.0 + a * a / (f + a * a);
        f *= (1.0 - r[0]) * ((1.0 - r[1]) * ((1.0 - r[2]) * P[sidx(N, x, y, z)] + r[2] * P[sidx(N, x, y, z + 1)]) +
                             r[1] * ((1.0 - r[2]) * P[sidx(N, x, y + 1, z)] + r[2] * P[sidx(N, x, y + 1, z + 1)])) +
             r[0] * ((1.0 - r[1]) * ((1.0 - r[2]) * P[sidx(N, x + 1, y, z)] + r[2] * P[sidx(N, x + 1, y, z + 1)]) +
                     r[1] * ((1.0 - r[2]) * P[sidx(N, x + 1, y + 1, z)] + r[2] * P[sidx(N, x + 1, y + 1, z + 1)]));
        f += 1.0;
    } else {
        f = 1.0;
    }
    double mass = c->mass;
    static const int diag_comp[3] = {0, 3, 5};
    for (int d = 0; d < 3; d++) {
        double w = f * mass * v[d] * v[d] / e;
        for (int b = 0; b < 8; b++) {
            int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
            double wt = w * (bx ? r[0] : 1.0 - r[0]) * (by ? r[1] : 1.0 - r[1]) * (bz ? r[2] : 1.0 - r[2]);
            o[diag_comp[d] * M + sidx(N, x + bx, y + by, z + bz)] += wt;
        }
    }
    double w = f * mass * v[0] * v[1] / e;
    o[1 * M + sidx(N, x, y, z)] += w * (1.0 - r[2]);
    o[1 * M + sidx(N, x, y, z + 1)] += w * r[2];
    w = f * mass * v[0] * v[2] / e;
    o[2 * M + sidx(N, x, y, z)] += w * (1.0 - r[1]);
    o[2 * M + sidx(N, x, y + 1, z)] += w * r[1];
    w = f * mass * v[1] * v[2] / e;
    o[4 * M + sidx(N, x, y, z)] += w * (1.0 - r[0]);
    o[4 * M + sidx(N, x + 1, y, z)] += w * r[0];
}

OR_API void or_project_Tij(int N, long np, const double *pos, const double *q, double pmass, double *Tij,
                           double a, const double *phi, double coeff)
{
    double dx = 1.0 / (double)N;
    dep_ctx c = {Tij, phi, a, pmass * coeff / (dx * dx * dx) / a};
    deposit_planes(N, np, pos, q, dep_Tij, &c);
}

/* ------------------------------------------------------------------------
 * Particle updates
 * ------------------------------------------------------------------------ */
#define F(fld, dx_, dy_, dz_) (fld)[sidx(N, x + (dx_), y + (dy_), z + (dz_))]

/* CIC-weighted GRADIENT_ORDER=2 central difference of a scalar at the
 * particle, the loop of gevolution.hpp:930-941. */
static void grad2(int N, const double *f, long x, long y, long z, const double r[3], double g[3])
{
    for (int i = 0; i < 3; i++) {
        int ex = (i == 0), ey = (i == 1), ez = (i == 2);
        double s;
        s = 0.5 * (1. - r[0]) * (1. - r[1]) * (1. - r[2]) * (F(f, ex, ey, ez) - F(f, -ex, -ey, -ez));
        s += 0.5 * r[0] * (1. - r[1]) * (1. - r[2]) * (F(f, ex + 1, ey, ez) - F(f, -ex + 1, -ey, -ez));
        s += 0.5 * (1. - r[0]) * r[1] * (1. - r[2]) * (F(f, ex, ey + 1, ez) - F(f, -ex, -ey + 1, -ez));
        s += 0.5 * r[0] * r[1] * (1. - r[2]) * (F(f, ex + 1, ey + 1, ez) - F(f, -ex + 1, -ey + 1, -ez));
        s += 0.5 * (1. - r[0]) * (1. - r[1]) * r[2] * (F(f, ex, ey, ez + 1) - F(f, -ex, -ey, -ez + 1));
        s += 0.5 * r[0] * (1. - r[1]) * r[2] * (F(f, ex + 1, ey, ez + 1) - F(f, -ex + 1, -ey, -ez + 1));
        s += 0.5 * (1. - r[0]) * r[1] * r[2] * (F(f, ex, ey + 1, ez + 1) - F(f, -ex, -ey + 1, -ez + 1));
        s += 0.5 * r[0] * r[1] * r[2] * (F(f, ex + 1, ey + 1, ez + 1) - F(f, -ex + 1, -ey + 1, -ez + 1));
        g[i] = s;
    }
}

/* CIC edge gradient (GRADIENT_ORDER=1 form), used for chi in update_q
 * (gevolution.hpp:950-964) and for the IC displacement/velocity
 * (ic_basic.hpp:73-84, 141-152). */
static void grad1(int N, const double *f, long x, long y, long z, const double r[3], double g[3])
{
    g[0] = (1. - r[1]) * (1. - r[2]) * (F(f, 1, 0, 0) - F(f, 0, 0, 0));
    g[1] = (1. - r[0]) * (1. - r[2]) * (F(f, 0, 1, 0) - F(f, 0, 0, 0));
    g[2] = (1. - r[0]) * (1. - r[1]) * (F(f, 0, 0, 1) - F(f, 0, 0, 0));
    g[0] += r[1] * (1. - r[2]) * (F(f, 1, 1, 0) - F(f, 0, 1, 0));
    g[1] += r[0] * (1. - r[2]) * (F(f, 1, 1, 0) - F(f, 1, 0, 0));
    g[2] += r[0] * (1. - r[1]) * (F(f, 1, 0, 1) - F(f, 1, 0, 0));
    g[0] += (1. - r[1]) * r[2] * (F(f, 1, 0, 1) - F(f, 0, 0, 1));
    g[1] += (1. - r[0]) * r[2] * (F(f, 0, 1, 1) - F(f, 0, 0, 1));
    g[2] += (1. - r[0]) * r[1] * (F(f, 0, 1, 1) - F(f, 0, 1, 0));
    g[0] += r[1] * r[2] * (F(f, 1, 1, 1) - F(f, 0, 1, 1));
    g[1] += r[0] * r[2] * (F(f, 1, 1, 1) - F(f, 1, 0, 1));
    g[2] += r[0] * r[1] * (F(f, 1, 1, 1) - F(f, 1, 1, 0));
}

static double cic_interp(int N, const double *f, long x, long y, long z, const double r[3])
{
    double s = F(f, 0, 0, 0) * (1. - r[0]) * (1. - r[1]) * (1. - r[2]);
    s += F(f, 1, 0, 0) * r[0] * (1. - r[1]) * (1. - r[2]);
    s += F(f, 0, 1, 0) * (1. - r[0]) * r[1] * (1. - r[2]);
    s += F(f, 1, 1, 0) * r[0] * r[1] * (1. - r[2]);
    s += F(f, 0, 0, 1) * (1. - r[0]) * (1. - r[1]) * r[2];
    s += F(f, 1, 0, 1) * r[0] * (1. - r[1]) * r[2];
    s += F(f, 0, 1, 1) * (1. - r[0]) * r[1] * r[2];
    s += F(f, 1, 1, 1) * r[0] * r[1] * r[2];
    return s;
}

/* update_q (GR kick), gevolution.hpp:888-1011, GRADIENT_ORDER=2, no
 * ANISOTROPIC_EXPANSION.  params[0]=a, params[1]=bcoeff (= a^2 N, main.cu:1227).
 * Returns the maximum over particles of the functor's return value v2/a^2.
 * Pinned by: uniform fields -> free streaming, linear potential -> exact
 * constant force, Newtonian limit, random B against the staggered stencils,
 * the weak-field geodesic equation to O(phi grad phi) (tests). */
OR_API double or_update_q(int N, long np, const double *pos, double *q, double dtau, const double *phi,
                          const double *chi, const double *Bi, int nfield, double a, double bcoeff)
{
    const long M = (long)N * N * N;
    const double dx = 1.0 / (double)N;
    double vmax = 0.0;
#pragma omp parallel for reduction(max : vmax) schedule(static)
    for (long i = 0; i < np; i++) {
        long cc[3]; double r[3];
        cell_of(N, pos, np, i, cc, r);
        long x = cc[0], y = cc[1], z = cc[2];
        double v[3] = {q[i], q[np + i], q[2 * np + i]};
        double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        double e2 = v2 + a * a;
        double phiint = cic_interp(N, phi, x, y, z, r);
        phiint /= e2;
        phiint = 4. * phiint * v2 + (1. + a * a * phiint) * (v2 + e2) / e2;
        double g[3];
        grad2(N, phi, x, y, z, r, g);
        g[0] *= phiint; g[1] *= phiint; g[2] *= phiint;
        if (nfield >= 2 && chi) {
            double gc[3];
            grad1(N, chi, x, y, z, r, gc);
            g[0] -= gc[0]; g[1] -= gc[1]; g[2] -= gc[2];
        }
        e2 = sqrt(e2);
        if (nfield >= 3 && Bi) {
            const double *B0 = Bi, *B1 = Bi + M, *B2 = Bi + 2 * M;
            double pg[3];
            pg[0] = ((1. - r[2]) * (F(B1, 1, 0, 0) - F(B1, 0, 0, 0)) + r[2] * (F(B1, 1, 0, 1) - F(B1, 0, 0, 1))) * v[1];
            pg[0] += ((1. - r[1]) * (F(B2, 1, 0, 0) - F(B2, 0, 0, 0)) + r[1] * (F(B2, 1, 1, 0) - F(B2, 0, 1, 0))) * v[2];
            pg[0] += (1. - r[1]) * (1. - r[2]) * ((r[0] - 1.) * F(B0, -1, 0, 0) + (1. - 2. * r[0]) * F(B0, 0, 0, 0) + r[0] * F(B0, 1, 0, 0)) * v[0];
            pg[0] += r[1] * (1. - r[2]) * ((r[0] - 1.) * F(B0, -1, 1, 0) + (1. - 2. * r[0]) * F(B0, 0, 1, 0) + r[0] * F(B0, 1, 1, 0)) * v[0];
            pg[0] += (1. - r[1]) * r[2] * ((r[0] - 1.) * F(B0, -1, 0, 1) + (1. - 2. * r[0]) * F(B0, 0, 0, 1) + r[0] * F(B0, 1, 0, 1)) * v[0];
            pg[0] += r[1] * r[2] * ((r[0] - 1.) * F(B0, -1, 1, 1) + (1. - 2. * r[0]) * F(B0, 0, 1, 1) + r[0] * F(B0, 1, 1, 1)) * v[0];

            pg[1] = ((1. - r[0]) * (F(B2, 0, 1, 0) - F(B2, 0, 0, 0)) + r[0] * (F(B2, 1, 1, 0) - F(B2, 1, 0, 0))) * v[2];
            pg[1] += ((1. - r[2]) * (F(B0, 0, 1, 0) - F(B0, 0, 0, 0)) + r[2] * (F(B0, 0, 1, 1) - F(B0, 0, 0, 1))) * v[0];
            pg[1] += (1. - r[0]) * (1. - r[2]) * ((r[1] - 1.) * F(B1, 0, -1, 0) + (1. - 2. * r[1]) * F(B1, 0, 0, 0) + r[1] * F(B1, 0, 1, 0)) * v[1];
            pg[1] += r[0] * (1. - r[2]) * ((r[1] - 1.) * F(B1, 1, -1, 0) + (1. - 2. * r[1]) * F(B1, 1, 0, 0) + r[1] * F(B1, 1, 1, 0)) * v[1];
            pg[1] += (1. - r[0]) * r[2] * ((r[1] - 1.) * F(B1, 0, -1, 1) + (1. - 2. * r[1]) * F(B1, 0, 0, 1) + r[1] * F(B1, 0, 1, 1)) * v[1];
            pg[1] += r[0] * r[2] * ((r[1] - 1.) * F(B1, 1, -1, 1) + (1. - 2. * r[1]) * F(B1, 1, 0, 1) + r[1] * F(B1, 1, 1, 1)) * v[1];

            pg[2] = ((1. - r[1]) * (F(B0, 0, 0, 1) - F(B0, 0, 0, 0)) + r[1] * (F(B0, 0, 1, 1) - F(B0, 0, 1, 0))) * v[0];
            pg[2] += ((1. - r[0]) * (F(B1, 0, 0, 1) - F(B1, 0, 0, 0)) + r[0] * (F(B1, 1, 0, 1) - F(B1, 1, 0, 0))) * v[1];
            pg[2] += (1. - r[0]) * (1. - r[1]) * ((r[2] - 1.) * F(B2, 0, 0, -1) + (1. - 2. * r[2]) * F(B2, 0, 0, 0) + r[2] * F(B2, 0, 0, 1)) * v[2];
            pg[2] += r[0] * (1. - r[1]) * ((r[2] - 1.) * F(B2, 1, 0, -1) + (1. - 2. * r[2]) * F(B2, 1, 0, 0) + r[2] * F(B2, 1, 0, 1)) * v[2];
            pg[2] += (1. - r[0]) * r[1] * ((r[2] - 1.) * F(B2, 0, 1, -1) + (1. - 2. * r[2]) * F(B2, 0, 1, 0) + r[2] * F(B2, 0, 1, 1)) * v[2];
            pg[2] += r[0] * r[1] * ((r[2] - 1.) * F(B2, 1, 1, -1) + (1. - 2. * r[2]) * F(B2, 1, 1, 0) + r[2] * F(B2, 1, 1, 1)) * v[2];

            g[0] += pg[0] / bcoeff / e2;
            g[1] += pg[1] / bcoeff / e2;
            g[2] += pg[2] / bcoeff / e2;
        }
        v2 = 0.;
        for (int d = 0; d < 3; d++) {
            v[d] -= dtau * e2 * g[d] / dx;
            v2 += v[d] * v[d];
            q[d * np + i] = v[d];
        }
        double ret = v2 / a / a;
        if (ret > vmax) vmax = ret;
    }
    return vmax;
}

/* update_q_Newton, gevolution.hpp:1051-1118 (GRADIENT_ORDER=2).
 * Pinned by: linear potential -> exact constant force (tests). */
OR_API double or_update_q_newton(int N, long np, const double *pos, double *q, double dtau, const double *psi,
                                 const double *chi, int nfield, double a)
{
    const double dx = 1.0 / (double)N;
    double vmax = 0.0;
#pragma omp parallel for reduction(max : vmax) schedule(static)
    for (long i = 0; i < np; i++) {
        long cc[3]; double r[3];
        cell_of(N, pos, np, i, cc, r);
        double g[3];
        grad2(N, psi, cc[0], cc[1], cc[2], r, g);
        if (nfield >= 2 && chi) {
            double gc[3];
            grad1(N, chi, cc[0], cc[1], cc[2], r, gc);
            g[0] -= gc[0]; g[1] -= gc[1]; g[2] -= gc[2];
        }
        double v2 = 0.;
        for (int d = 0; d < 3; d++) {
            double v = q[d * np + i] - dtau * a * g[d] / dx;
            q[d * np + i] = v;
            v2 += v * v;
        }
        double ret = v2 / a / a;
        if (ret > vmax) vmax = ret;
    }
    return vmax;
}

/* Periodic wrap applied when LATfield2's moveParticles re-bins a particle
 * across the global boundary (behaviour inferred from the single-domain
 * periodic lattice). */
static inline double wrap_pos(double p)
{
    if (p >= 1.0) p -= 1.0;
    else if (p < 0.0) p += 1.0;
    if (p >= 1.0) p = 0.0;
    if (p < 0.0) p = 0.0;
    return p;
}

/* update_pos (GR drift), gevolution.hpp:1162-1232.  params {a, a^2 N}.
 * Pinned by: free streaming v = q/sqrt(q^2+a^2) with zero fields, random B
 * against the staggered interpolation, the geodesic velocity to O(phi, chi)
 * (tests). */
OR_API void or_update_pos(int N, long np, double *pos, const double *q, double dtau, const double *phi,
                          const double *chi, const double *Bi, int nfield, double a, double bcoeff)
{
    const long M = (long)N * N * N;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < np; i++) {
        long cc[3]; double r[3];
        cell_of(N, pos, np, i, cc, r);
        long x = cc[0], y = cc[1], z = cc[2];
        double v0 = q[i], v1 = q[np + i], v2q = q[2 * np + i];
        double v2 = v0 * v0 + v1 * v1 + v2q * v2q;
        double e2 = v2 + a * a;
        double ph = 0., ch = 0.;
        if (nfield >= 1) ph = cic_interp(N, phi, x, y, z, r);
        if (nfield >= 2) ch = cic_interp(N, chi, x, y, z, r);
        v2 = (1. + (3. - v2 / e2) * ph - ch) / sqrt(e2);
        double v[3] = {v0 * v2, v1 * v2, v2q * v2};
        if (nfield >= 3) {
            const double *B0 = Bi, *B1 = Bi + M, *B2 = Bi + 2 * M;
            double b[3];
            b[0] = F(B0, 0, 0, 0) * (1. - r[1]) * (1. - r[2]);
            b[1] = F(B1, 0, 0, 0) * (1. - r[0]) * (1. - r[2]);
            b[2] = F(B2, 0, 0, 0) * (1. - r[0]) * (1. - r[1]);
            b[1] += F(B1, 1, 0, 0) * r[0] * (1. - r[2]);
            b[2] += F(B2, 1, 0, 0) * r[0] * (1. - r[1]);
            b[0] += F(B0, 0, 1, 0) * r[1] * (1. - r[2]);
            b[2] += F(B2, 0, 1, 0) * (1. - r[0]) * r[1];
            b[0] += F(B0, 0, 0, 1) * (1. - r[1]) * r[2];
            b[1] += F(B1, 0, 0, 1) * (1. - r[0]) * r[2];
            b[1] += F(B1, 1, 0, 1) * r[0] * r[2];
            b[0] += F(B0, 0, 1, 1) * r[1] * r[2];
            b[2] += F(B2, 1, 1, 0) * r[0] * r[1];
            for (int l = 0; l < 3; l++) pos[l * np + i] = wrap_pos(pos[l * np + i] + dtau * (v[l] + b[l] / bcoeff));
        } else {
            for (int l = 0; l < 3; l++) pos[l * np + i] = wrap_pos(pos[l * np + i] + dtau * v[l]);
        }
    }
}

/* update_pos_Newton, gevolution.hpp:1270-1273. */
OR_API void or_update_pos_newton(int N, long np, double *pos, const double *q, double dtau, double a)
{
    (void)N;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < np; i++)
        for (int l = 0; l < 3; l++) pos[l * np + i] = wrap_pos(pos[l * np + i] + dtau * q[l * np + i] / a);
}

/* displace_pcls_ic_basic (ic_basic.hpp:65-96) with one field: pos += coeff grad(xi)/dx.
 * Returns the maximum displacement |coeff grad xi| (the MAX reduction output). */
OR_API double or_displace_particles(int N, long np, double *pos, const double *xi, double coeff)
{
    const double dx = 1.0 / (double)N;
    double dmax = 0.0;
    double *newpos = (double *)malloc(sizeof(double) * 3 * (np > 0 ? np : 1));
#pragma omp parallel for reduction(max : dmax) schedule(static)
    for (long i = 0; i < np; i++) {
        long cc[3]; double r[3];
        cell_of(N, pos, np, i, cc, r);
        double g[3];
        grad1(N, xi, cc[0], cc[1], cc[2], r, g);
        g[0] /= dx; g[1] /= dx; g[2] /= dx;
        double d = coeff * sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        if (d > dmax) dmax = d;
        for (int l = 0; l < 3; l++) newpos[l * np + i] = wrap_pos(pos[l * np + i] + coeff * g[l]);
    }
    memcpy(pos, newpos, sizeof(double) * 3 * np);
    free(newpos);
    return dmax;
}

/* initialize_q_ic_basic (ic_basic.hpp:133-168): q = -coeff grad(Phi)/dx.
 * Returns max q^2 (the functor's return value). */
OR_API double or_initialize_q(int N, long np, const double *pos, double *q, const double *pot, double coeff)
{
    const double dx = 1.0 / (double)N;
    double vmax = 0.0;
#pragma omp parallel for reduction(max : vmax) schedule(static)
    for (long i = 0; i < np; i++) {
        long cc[3]; double r[3];
        cell_of(N, pos, np, i, cc, r);
        double g[3];
        grad1(N, pot, cc[0], cc[1], cc[2], r, g);
        double v2 = 0.;
        for (int l = 0; l < 3; l++) {
            double v = -(g[l] / dx) * coeff;
            q[l * np + i] = v;
            v2 += v * v;
        }
        if (v2 > vmax) vmax = v2;
    }
    return vmax;
}
#undef F

/* ------------------------------------------------------------------------
 * Background (background.hpp:200-245, 268-290), Omega_rad = 0, no ncdm/fld.
 * Pinned by: Einstein-de Sitter closed forms (tests).
 * ------------------------------------------------------------------------ */
OR_API double or_hconf(double a, double fourpiG, double om, double ol, double orad)
{
    return sqrt((2. * fourpiG / 3.) * ((om / a) + (ol * a * a) + (orad / a / a)));
}

OR_API double or_rk4bg(double a, double fourpiG, double om, double ol, double orad, double dtau)
{
    double k1a = a * or_hconf(a, fourpiG, om, ol, orad);
    double k2a = (a + k1a * dtau / 2.) * or_hconf(a + k1a * dtau / 2., fourpiG, om, ol, orad);
    double k3a = (a + k2a * dtau / 2.) * or_hconf(a + k2a * dtau / 2., fourpiG, om, ol, orad);
    double k4a = (a + k3a * dtau) * or_hconf(a + k3a * dtau, fourpiG, om, ol, orad);
    return a + dtau * (k1a + 2. * k2a + 2. * k3a + k4a) / 6.;
}

/* particleHorizon: tau(a) = int_0^a da'/(a' Hconf(a')), evaluated in the
 * reference's variable s = sqrt(a') (background.hpp:261-290) with composite
 * Gauss-Legendre (8 points x 2048 panels) instead of GSL qng. */
OR_API double or_particle_horizon(double a, double fourpiG, double om, double ol, double orad)
{
    static const double xg[8] = {-0.96028985649753618, -0.79666647741362673, -0.52553240991632899,
                                 -0.18343464249564978, 0.18343464249564978,  0.52553240991632899,
                                 0.79666647741362673,  0.96028985649753618};
    static const double wg[8] = {0.10122853629037669, 0.22238103445337434, 0.31370664587788705, 0.36268378337836177,
                                 0.36268378337836177, 0.31370664587788705, 0.22238103445337434, 0.10122853629037669};
    double lo = sqrt(a) * 1.0e-7, hi = sqrt(a);
    const int panels = 2048;
    double h = (hi - lo) / panels, sum = 0.0;
    for (int p = 0; p < panels; p++) {
        double c = lo + (p + 0.5) * h;
        for (int j = 0; j < 8; j++) {
            double s = c + 0.5 * h * xg[j];
            sum += wg[j] * 0.5 * h * 2. / (s * or_hconf(s * s, 1., om, ol, orad));
        }
    }
    return sum / sqrt(fourpiG);
}

/* ------------------------------------------------------------------------
 * Lightcone particle output (single rank).
 * ------------------------------------------------------------------------ */

#define OR_IDCHECK_ZONE 0.05        /* LIGHTCONE_IDCHECK_ZONE, metadata.hpp:38-40 */
#define OR_GADGET_VEL 3.335640952e-6 /* GADGET_VELOCITY_CONVERSION, metadata.hpp:117 */

static double dist3(const double a[3], const double b[3])
{
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

/* findIntersectingLightcones (tools.hpp:619-862): the periodic copies
 * lcvertex + (u,v,w) of the observer, |u|,|v|,|w| <= ceil(outer)+1, visited
 * u-major, whose look-back shell [inner, outer] and cone may meet the box
 * domain = {lo[0..2], hi[0..2]} (box units).  Rejected: copies farther than
 * outer from the nearest point of the box, and copies with every box corner
 * nearer than inner.  Accepted: copies inside the box's enclosing sphere,
 * copies that see the box centre inside the opening, and copies whose
 * opening meets the enclosing sphere (two angular tests, outside and inside
 * the outer radius).  Returns the number found; the first cap are stored
 * (the reference stores MAX_INTERSECTS = 12). */
OR_API int or_lightcone_vertices(const double lcvertex[3], const double dir[3], double opening, double outer,
                                 double inner, const double domain[6], double *vertex, int cap)
{
    const int range = (int)ceil(outer) + 1;
    const double *lo = domain, *hi = domain + 3;
    double corner[8][3];
    for (int c = 0; c < 8; c++) {
        corner[c][0] = (c & 1) ? hi[0] : lo[0];
        corner[c][1] = (c & 2) ? hi[1] : lo[1];
        corner[c][2] = (c & 4) ? hi[2] : lo[2];
    }
    const double centre[3] = {0.5 * lo[0] + 0.5 * hi[0], 0.5 * lo[1] + 0.5 * hi[1], 0.5 * lo[2] + 0.5 * hi[2]};
    const double rdom = 0.5 * sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                   (hi[2] - lo[2]) * (hi[2] - lo[2]));
    int n = 0;
    for (int u = -range; u <= range; u++)
        for (int v = -range; v <= range; v++)
            for (int w = -range; w <= range; w++) {
                const double vx[3] = {lcvertex[0] + u, lcvertex[1] + v, lcvertex[2] + w};
                /* distance to the nearest point of the box */
                double s2 = 0.0;
                for (int d = 0; d < 3; d++) {
                    const double e = vx[d] < lo[d] ? lo[d] - vx[d] : (vx[d] > hi[d] ? vx[d] - hi[d] : 0.0);
                    s2 += e * e;
                }
                if (sqrt(s2) > outer) continue;
                int all_in = 1;
                for (int c = 0; c < 8 && all_in; c++)
                    if (!(dist3(corner[c], vx) < inner)) all_in = 0;
                if (all_in) continue;
                const double dist = dist3(centre, vx);
                int take = dist <= rdom;
                if (!take) {
                    const double cosang = ((centre[0] - vx[0]) * dir[0] + (centre[1] - vx[1]) * dir[1] +
                                           (centre[2] - vx[2]) * dir[2]) / dist;
                    if (cosang >= opening)
                        take = 1;
                    else if (dist > outer &&
                             acos(cosang) - acos(opening) <=
                                 acos((outer * outer + dist * dist - rdom * rdom) / (2. * outer * dist)))
                        take = 1;
                    else if (dist <= outer && acos(cosang) - acos(opening) <= asin(rdom / dist))
                        take = 1;
                }
                if (take) {
                    if (n < cap)
                        for (int d = 0; d < 3; d++) vertex[3 * n + d] = vx[d];
                    n++;
                }
            }
    return n;
}

static int backlog_has(const long *b, long nb, long id)
{
    long lo = 0, hi = nb;
    while (lo < hi) {
        long mid = lo + (hi - lo) / 2;
        if (b[mid] < id) lo = mid + 1;
        else hi = mid;
    }
    return lo < nb && b[lo] == id;
}

/* saveGadget2, lightcone form, device path (Particles_gevolution.hpp:1375-1470
 * buffer_tracer_particles + bufferTracerParticle :500-592 with
 * EXACT_OUTPUT_REDSHIFTS, the backlog check :2050-2115 and the prelog bins
 * :2120-2170 for one rank).  Walks the particles in order and writes the
 * selected ones in that order:
 *   - ID % tracer_factor == 0;
 *   - the first vertex v with inner <= d < outer (d = |pos - v|,
 *     inner = dist - dtau/2, outer = dist + (1/2 + ZONE) dtau_old) and, unless
 *     opening == -1, (pos - v).direction / d > opening;
 *   - not in the backlog when outer - d <= 2 ZONE dtau_old.
 * Each selected particle is moved to the light-cone crossing,
 * dtau_pos = (dist - d) / (1 + vlos), and kicked by dtau_pos + dtau_old/2
 * with the CIC edge gradient of phi at a + dtau_pos dadtau; positions are
 * shifted by (lightcone vertex - v), times boxsize, not wrapped.
 * prelog[j] = 1 when d - inner < 2 ZONE dtau (the ID goes to the next
 * backlog).  margin[i] (every particle) is the smallest distance of any
 * selection quantity to its threshold, over all vertices, for parity tests.
 * Returns the number written. */
OR_API long or_lightcone_select(int N, long np, const double *pos, const double *q, const long *id,
                                const double *phi, int tracer_factor, const double lcvertex[3], const double dir[3],
                                double opening, double dist, double dtau, double dtau_old, double dadtau, double a,
                                double boxsize, const double *vertex, int nv, const long *backlog, long nbacklog,
                                float *posout, float *velout, long *idout, unsigned char *prelog, double *margin)
{
    const double inner = dist - 0.5 * dtau, outer = dist + (0.5 + OR_IDCHECK_ZONE) * dtau_old;
    const double zc = 2. * OR_IDCHECK_ZONE * dtau_old, zp = 2. * OR_IDCHECK_ZONE * dtau;
    long nw = 0;
    for (long i = 0; i < np; i++) {
        double m = INFINITY;
        const double P[3] = {pos[i], pos[np + i], pos[2 * np + i]};
        int hit = -1;
        double dh = 0.0;
        for (int v = 0; v < nv; v++) {
            const double *vx = vertex + 3 * v;
            const double d = dist3(P, vx);
            const double cosang = ((P[0] - vx[0]) * dir[0] + (P[1] - vx[1]) * dir[1] + (P[2] - vx[2]) * dir[2]) / d;
            double t[5] = {fabs(d - inner), fabs(d - outer), opening == -1. ? INFINITY : fabs(cosang - opening),
                           fabs(outer - d - zc), fabs(d - inner - zp)};
            for (int k = 0; k < 5; k++) m = t[k] < m ? t[k] : m;
            if (hit < 0 && d >= inner && d < outer && (opening == -1. || cosang > opening)) {
                hit = v;
                dh = d;
            }
        }
        if (margin) margin[i] = m;
        if (tracer_factor <= 0 || id[i] % tracer_factor != 0 || hit < 0) continue;
        if (outer - dh <= zc && nbacklog > 0 && backlog_has(backlog, nbacklog, id[i])) continue;
        const double *vx = vertex + 3 * hit;
        const double Q[3] = {q[i], q[np + i], q[2 * np + i]};
        const double q2 = Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2];
        const double vlos = (Q[0] * (P[0] - vx[0]) + Q[1] * (P[1] - vx[1]) + Q[2] * (P[2] - vx[2])) / dh /
                            sqrt(q2 + a * (a + dtau_old * dadtau));
        const double dtp = (dist - dh) / (1. + vlos), dtv = dtp + 0.5 * dtau_old, ap = a + dtp * dadtau;
        double phip = 0.0, g[3] = {0.0, 0.0, 0.0};
        if (phi) {
            long c[3];
            double r[3];
            cell_of(N, pos, np, i, c, r);
            phip = cic_interp(N, phi, c[0], c[1], c[2], r);
            grad1(N, phi, c[0], c[1], c[2], r, g);
        }
        const double e2 = q2 + ap * ap, e = sqrt(e2), w = q2 + e2;
        const double vfac = 1. + (4. - w / e2) * phip;
        const double rescale = 1. / sqrt(ap) / OR_GADGET_VEL;
        for (int d = 0; d < 3; d++) {
            posout[3 * nw + d] = (float)((P[d] + dtp * Q[d] * vfac / e + (lcvertex[d] - vx[d])) * boxsize);
            velout[3 * nw + d] = (float)((Q[d] - dtv * w * g[d] * (double)N / e) * rescale / ap);
        }
        idout[nw] = id[i];
        prelog[nw] = (dh - inner < zp) ? 1 : 0;
        nw++;
    }
    return nw;
}

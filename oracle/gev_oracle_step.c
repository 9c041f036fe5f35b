/*
 * gev_oracle_step.c -- TEST INFRASTRUCTURE, NOT THE PRODUCT.
 *
 * CPU FP64 restatement of the parts of gevolution-2.0 that drive the PM step:
 * the counter-based RNG and the basic IC generator (mPk path), one main-loop
 * cycle in main.cu order, and the power-spectrum estimator.  Same conventions
 * as gev_oracle.c.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_API __attribute__((visibility("default")))

/* functions from gev_oracle.c */
void or_fft_r2c(int N, const double *in, double *out);
void or_fft_c2r(int N, const double *in, double *out);
void or_poisson(int N, const double *src, double *pot, double coeff, double modif);
void or_project_ft_scalar(int N, const double *S, double *chi, int add);
void or_project_ft_vector(int N, const double *S, double *B, double coeff, double modif);
double or_prepare_source_T00(int N, const double *phi, const double *chi, const double *source, double bg, double *result,
                             double c1, double c2, double c3);
void or_prepare_source_Tij(int N, const double *phi, const double *Tij, double *Sij, double coeff);
void or_project_T00(int N, long np, const double *pos, const double *q, double pmass, double *T00, double a,
                    const double *phi, double coeff);
void or_project_T0i(int N, long np, const double *pos, const double *q, double pmass, double *T0i, const double *phi,
                    double coeff);
void or_project_Tij(int N, long np, const double *pos, const double *q, double pmass, double *Tij, double a,
                    const double *phi, double coeff);
double or_update_q(int N, long np, const double *pos, double *q, double dtau, const double *phi, const double *chi,
                   const double *Bi, int nfield, double a, double bcoeff);
double or_update_q_newton(int N, long np, const double *pos, double *q, double dtau, const double *psi,
                          const double *chi, int nfield, double a);
void or_update_pos(int N, long np, double *pos, const double *q, double dtau, const double *phi, const double *chi,
                   const double *Bi, int nfield, double a, double bcoeff);
void or_update_pos_newton(int N, long np, double *pos, const double *q, double dtau, double a);
double or_displace_particles(int N, long np, double *pos, const double *xi, double coeff);
double or_initialize_q(int N, long np, const double *pos, double *q, const double *pot, double coeff);
double or_hconf(double a, double fourpiG, double om, double ol, double orad);
double or_rk4bg(double a, double fourpiG, double om, double ol, double orad, double dtau);

/* ------------------------------------------------------------------------
 * Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based RNG
 * that replaces gevolution's sitmo engine (prng_engine.hpp).  Pinned by the
 * published Random123 known-answer vectors (tests).
 * ------------------------------------------------------------------------ */
OR_API void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u53(uint32_t hi, uint32_t lo)
{
    uint64_t v = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------
 * Natural cubic spline (the gsl_interp_cspline that loadPowerSpectrum builds,
 * ic_basic.hpp:488-492).  Pinned by: exact reproduction of cubic data with
 * zero end curvature and of linear data (tests).
 * ------------------------------------------------------------------------ */
typedef struct { int n; double *x, *y, *m; } spline_t;

static void spline_init(spline_t *s, int n, const double *x, const double *y)
{
    s->n = n;
    s->x = (double *)malloc(sizeof(double) * n);
    s->y = (double *)malloc(sizeof(double) * n);
    s->m = (double *)calloc(n, sizeof(double));
    memcpy(s->x, x, sizeof(double) * n);
    memcpy(s->y, y, sizeof(double) * n);
    if (n < 3) return;
    int k = n - 2;
    double *a = (double *)malloc(sizeof(double) * k), *b = (double *)malloc(sizeof(double) * k),
           *c = (double *)malloc(sizeof(double) * k), *d = (double *)malloc(sizeof(double) * k);
    for (int i = 1; i <= k; i++) {
        double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
        a[i - 1] = h0; b[i - 1] = 2.0 * (h0 + h1); c[i - 1] = h1;
        d[i - 1] = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    }
    for (int i = 1; i < k; i++) { /* Thomas algorithm */
        double w = a[i] / b[i - 1];
        b[i] -= w * c[i - 1];
        d[i] -= w * d[i - 1];
    }
    s->m[k] = d[k - 1] / b[k - 1];
    for (int i = k - 2; i >= 0; i--) s->m[i + 1] = (d[i] - c[i] * s->m[i + 2]) / b[i];
    free(a); free(b); free(c); free(d);
}

static double spline_eval(const spline_t *s, double xv)
{
    int lo = 0, hi = s->n - 1;
    if (xv <= s->x[0]) hi = 1;
    else if (xv >= s->x[s->n - 1]) lo = s->n - 2;
    else {
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (s->x[mid] > xv) hi = mid; else lo = mid;
        }
    }
    lo = (hi - 1 < lo) ? lo : hi - 1;
    double h = s->x[lo + 1] - s->x[lo], t = xv - s->x[lo];
    double m0 = s->m[lo], m1 = s->m[lo + 1];
    double slope = (s->y[lo + 1] - s->y[lo]) / h - h * (2.0 * m0 + m1) / 6.0;
    return s->y[lo] + t * (slope + t * (0.5 * m0 + t * (m1 - m0) / (6.0 * h)));
}

static void spline_free(spline_t *s) { free(s->x); free(s->y); free(s->m); }

OR_API double or_spline_eval(int n, const double *x, const double *y, double xv)
{
    spline_t s;
    spline_init(&s, n, x, y);
    double r = spline_eval(&s, xv);
    spline_free(&s);
    return r;
}

/* ------------------------------------------------------------------------
 * Basic IC generator, mPk-file path with the standard CIC kernel
 * (generateIC_basic, ic_basic.hpp:1834-2505 with "correct displacement = no";
 * generateCICKernel default branch :810-838; generateDisplacementField
 * :1164-1270; displace_pcls_ic_basic :65-96; initialize_q_ic_basic :133-168).
 * Differences by design: Philox instead of sitmo (the realisation cannot be
 * matched), particles on the cell-centred N^3 lattice with ID = x+N(y+Nz),
 * sinc deconvolution by signed frequency, Hermitian pairs on the kx=0 plane
 * drawn once.
 * ------------------------------------------------------------------------ */

/* the random complex amplitude of mode (kx, ky, kz) in array indices;
 * kx = 0 pairs are folded to a canonical member and conjugated. */
static void ic_mode_random(int N, uint32_t seed, long kx, long ky, long kz, double *re, double *im)
{
    long ny = (ky <= N / 2) ? ky : ky - N;
    long nz = (kz <= N / 2) ? kz : kz - N;
    int conj = 0;
    if (kx == 0 && (ny < 0 || (ny == 0 && nz < 0))) { ny = -ny; nz = -nz; conj = 1; }
    uint32_t ctr[4] = {(uint32_t)kx, (uint32_t)(int32_t)ny, (uint32_t)(int32_t)nz, 0u};
    uint32_t key[2] = {seed, 0x67657631u};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    double u1 = u53(o[0], o[1]), u2 = u53(o[2], o[3]);
    double r = sqrt(-2.0 * log(u1));
    *re = r * cos(2.0 * M_PI * u2);
    *im = (conj ? -1.0 : 1.0) * r * sin(2.0 * M_PI * u2);
}

static double sinc_n(int N, long n)
{
    if (n == 0) return 1.0;
    double x = M_PI * (double)n / (double)N;
    return sin(x) / x;
}

/* Fills potFT (complex, unpadded) with the displacement potential of
 * generateDisplacementField; kerFT is the Fourier image of the CIC kernel. */
OR_API void or_ic_displacement_ft(int N, uint32_t seed, int ksphere, double coeff, int npk, const double *kL,
                                  const double *amp, const double *kerFTd, double *potFTd)
{
    const long P = N / 2 + 1;
    const long kmax = N / 2 - 1, kmax2 = kmax * kmax;
    const double *ker = kerFTd;
    double *pot = potFTd;
    spline_t s;
    spline_init(&s, npk, kL, amp);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < (long)N * N; l++) {
        long ky = l % N, kz = l / N;
        long ny = (ky <= N / 2) ? ky : ky - N;
        long nz = (kz <= N / 2) ? kz : kz - N;
        for (long kx = 0; kx < P; kx++) {
            long i = kx + P * l;
            long k2i = kx * kx + ny * ny + nz * nz;
            if (k2i == 0 || kx >= kmax || labs(ny) >= kmax || labs(nz) >= kmax || (ksphere && k2i >= kmax2)) {
                pot[2 * i] = 0.0; pot[2 * i + 1] = 0.0;
                continue;
            }
            double sc = sinc_n(N, kx) * sinc_n(N, ny) * sinc_n(N, nz);
            double k2 = 4.0 * M_PI * M_PI * (double)k2i;
            double re, im;
            ic_mode_random(N, seed, kx, ky, kz, &re, &im);
            double mag = spline_eval(&s, sqrt(k2)) * sc * (1.0 + 7.5 * coeff / k2);
            re *= mag; im *= mag;
            /* divide by the kernel (complex) */
            double kr = ker[2 * i], ki = ker[2 * i + 1];
            double den = kr * kr + ki * ki;
            pot[2 * i] = (re * kr + im * ki) / den;
            pot[2 * i + 1] = (im * kr - re * ki) / den;
        }
    }
    spline_free(&s);
}

/* loadPowerSpectrum conversion (ic_basic.hpp:445-447, 1910-1914):
 * k -> k L, P -> sqrt(P L / 2) / L^2. */
OR_API void or_pk_table(int n, const double *k_hmpc, const double *P_mpch3, double boxsize, double *kL, double *amp)
{
    for (int i = 0; i < n; i++) {
        kL[i] = k_hmpc[i] * boxsize;
        amp[i] = sqrt(0.5 * P_mpch3[i] * boxsize) / boxsize / boxsize;
    }
}

/* Complete IC generation.  Buffers: pos,q [3np], id [np], phi, chi [M],
 * Bi [3M], source [M], Sij [6M].  Returns maxvel (after the GR rescaling
 * of main.cu:462-467 when gr != 0); *max_disp gets the max displacement. */
OR_API double or_ic_generate(int N, uint32_t seed, int ksphere, int gr, double boxsize, double fourpiG, double om,
                             double ol, double a, int npk, const double *k_hmpc, const double *P_mpch3, double *pos,
                             double *q, int64_t *id, double *phi, double *chi, double *Bi, double *source, double *Sij,
                             double *max_disp)
{
    const long M = (long)N * N * N, np = M;
    const long P = N / 2 + 1, nk = P * N * N;
    const double dx = 1.0 / (double)N;
    const double pmass = om / (double)np;
    const double H = or_hconf(a, fourpiG, om, ol, 0.0);
    double *ft = (double *)malloc(sizeof(double) * 2 * nk * 6);
    double *ker = (double *)malloc(sizeof(double) * 2 * nk);
    double *kL = (double *)malloc(sizeof(double) * npk), *amp = (double *)malloc(sizeof(double) * npk);
    or_pk_table(npk, k_hmpc, P_mpch3, boxsize, kL, amp);

    /* generateCICKernel standard branch: 6 renorm at the origin, -renorm at the 6 neighbours, renorm = N^2 */
    memset(source, 0, sizeof(double) * M);
    double renorm = (double)N * (double)N;
    source[0] = 6.0 * renorm;
    source[1] = -renorm; source[N - 1] = -renorm;
    source[N] = -renorm; source[(long)N * (N - 1)] = -renorm;
    source[(long)N * N] = -renorm; source[(long)N * N * (N - 1)] = -renorm;
    or_fft_r2c(N, source, ker);
    or_ic_displacement_ft(N, seed, ksphere, gr ? H * H : 0.0, npk, kL, amp, ker, ft);
    or_fft_c2r(N, ft, chi); /* chi = displacement potential */

    for (long z = 0; z < N; z++)
        for (long y = 0; y < N; y++)
            for (long x = 0; x < N; x++) {
                long i = x + N * (y + N * z);
                pos[i] = ((double)x + 0.5) / (double)N;
                pos[np + i] = ((double)y + 0.5) / (double)N;
                pos[2 * np + i] = ((double)z + 0.5) / (double)N;
                q[i] = q[np + i] = q[2 * np + i] = 0.0;
                id[i] = i;
            }
    double md = or_displace_particles(N, np, pos, chi, 1.0);
    if (max_disp) *max_disp = md;

    /* phi from the displaced CIC density (ic_basic.hpp:2406-2428) */
    memset(source, 0, sizeof(double) * M);
    or_project_T00(N, np, pos, q, pmass, source, a, NULL, 1.0);
    or_fft_r2c(N, source, ft);
    ft[0] = 0.0; ft[1] = 0.0;
    or_poisson(N, ft, ft, fourpiG / a, 3.0 * (double)gr * (H * H + fourpiG * om / a));
    or_fft_c2r(N, ft, phi);

    /* velocities from the linear approximation (ic_basic.hpp:2431-2437) */
    double om_a = om / (om + ol * a * a * a);
    double rescale = a / H / (1.5 * om_a);
    double vmax2 = or_initialize_q(N, np, pos, q, phi, rescale);
    double maxvel = sqrt(vmax2) / a;

    /* B (ic_basic.hpp:2469-2479) */
    memset(Bi, 0, sizeof(double) * 3 * M);
    or_project_T0i(N, np, pos, q, pmass, Bi, phi, 1.0);
    for (int c = 0; c < 3; c++) or_fft_r2c(N, Bi + c * M, ft + 2 * c * nk);
    or_project_ft_vector(N, ft, ft, fourpiG / (double)N / (double)N, 0.0);
    for (int c = 0; c < 3; c++) or_fft_c2r(N, ft + 2 * c * nk, Bi + c * M);

    /* chi (ic_basic.hpp:2481-2493) */
    memset(Sij, 0, sizeof(double) * 6 * M);
    or_project_Tij(N, np, pos, q, pmass, Sij, a, phi, 1.0);
    or_prepare_source_Tij(N, phi, Sij, Sij, 2.0 * fourpiG / a / (double)N / (double)N);
    for (int c = 0; c < 6; c++) or_fft_r2c(N, Sij + c * M, ft + 2 * c * nk);
    or_project_ft_scalar(N, ft, ker, 0);
    or_fft_c2r(N, ker, chi);

    if (gr) maxvel /= sqrt(maxvel * maxvel + 1.0);
    (void)dx;
    free(ft); free(ker); free(kL); free(amp);
    return maxvel;
}

/* ------------------------------------------------------------------------
 * One main-loop cycle in main.cu order (main.cu:608-1317), vector method
 * elliptic, one CDM species, no baryons/ncdm/radiation, no outputs.
 * sc: in/out {a, tau, dtau, dtau_old, maxvel, T00hom, phi_hom}.  dtau for
 * the cycle is sc[2]; on return sc[3] = dtau (the caller sets the next dtau).
 * Pinned by: linear growth of a single small mode against D(a), momentum
 * conservation (tests); every part is pinned on its own.
 * ------------------------------------------------------------------------ */
OR_API void or_step(int N, long np, int gr, double pmass, double fourpiG, double om, double ol, double z_relax,
                    double *pos, double *q, double *phi, double *chi, double *Bi, double *source, double *Sij,
                    double *sc)
{
    const long M = (long)N * N * N;
    const long P = N / 2 + 1, nk = P * N * N;
    const double dx = 1.0 / (double)N;
    double a = sc[0], tau = sc[1], dtau = sc[2], dtau_old = sc[3];
    double *ft = (double *)malloc(sizeof(double) * 2 * nk * 6);
    double *sft = (double *)malloc(sizeof(double) * 2 * nk);
    double maxv2;

    memset(source, 0, sizeof(double) * M);
    if (gr) {
        or_project_T00(N, np, pos, q, pmass, source, a, phi, 1.0);
        memset(Sij, 0, sizeof(double) * 6 * M);
        double H = or_hconf(a, fourpiG, om, ol, 0.0);
        if (dtau_old > 0.) {
            double s = or_prepare_source_T00(N, phi, chi, source, om, source, 3. * H * dx * dx / dtau_old,
                                             fourpiG * dx * dx / a, 3. * H * H * dx * dx);
            sc[5] = s / (double)M;
            or_fft_r2c(N, source, sft);
        }
        or_project_Tij(N, np, pos, q, pmass, Sij, a, phi, 1.0);
        if (dtau_old > 0.) {
            or_poisson(N, sft, sft, 1. / (dx * dx), 3. * H / dtau_old);
            or_fft_c2r(N, sft, phi);
            sc[6] = sft[0];
        }
        or_prepare_source_Tij(N, phi, Sij, Sij, 2. * fourpiG * dx * dx / a);
        for (int c = 0; c < 6; c++) or_fft_r2c(N, Sij + c * M, ft + 2 * c * nk);
        memset(Bi, 0, sizeof(double) * 3 * M);
        or_project_T0i(N, np, pos, q, pmass, Bi, phi, 1.0);
        or_project_ft_scalar(N, ft, sft, 0);
        or_fft_c2r(N, sft, chi);
        for (int c = 0; c < 3; c++) or_fft_r2c(N, Bi + c * M, ft + 2 * c * nk);
        or_project_ft_vector(N, ft, ft, fourpiG * dx * dx, 0.0);
        for (int c = 0; c < 3; c++) or_fft_c2r(N, ft + 2 * c * nk, Bi + c * M);

        int nf = (1. / a < z_relax + 1.) ? 3 : 2;
        maxv2 = or_update_q(N, np, pos, q, (dtau + dtau_old) / 2., phi, chi, Bi, nf, a, a * a * (double)N);
        a = or_rk4bg(a, fourpiG, om, ol, 0.0, 0.5 * dtau);
        nf = (1. / a < z_relax + 1.) ? 3 : 0;
        or_update_pos(N, np, pos, q, dtau, phi, chi, Bi, nf, a, a * a * (double)N);
        a = or_rk4bg(a, fourpiG, om, ol, 0.0, 0.5 * dtau);
        double mv = sqrt(maxv2);
        sc[4] = mv / sqrt(mv * mv + 1.0);
    } else {
        or_project_T00(N, np, pos, q, pmass, source, a, NULL, 1.0);
        or_fft_r2c(N, source, sft);
        or_poisson(N, sft, sft, fourpiG / a, 0.0);
        or_fft_c2r(N, sft, phi);
        sc[6] = sft[0];
        maxv2 = or_update_q_newton(N, np, pos, q, (dtau + dtau_old) / 2., phi, NULL, 1, a);
        a = or_rk4bg(a, fourpiG, om, ol, 0.0, 0.5 * dtau);
        or_update_pos_newton(N, np, pos, q, dtau, a);
        a = or_rk4bg(a, fourpiG, om, ol, 0.0, 0.5 * dtau);
        sc[4] = sqrt(maxv2);
    }
    sc[0] = a;
    sc[1] = tau + dtau;
    sc[3] = dtau;
    free(ft); free(sft);
}

/* ------------------------------------------------------------------------
 * extractCrossSpectrum (tools.hpp:65-330): f1FT, f2FT hold ncomp components of
 * N*N*(N/2+1) modes each (component c at c*N*N*(N/2+1)).  ktype 0 KTYPE_GRID
 * (|k|^2 = sum (2N sin(pi n/N))^2), 1 KTYPE_LINEAR (sum (2 pi |n|)^2).  The
 * mode's power p is Re(f1 conj f2) of components comp1, comp2 when both are
 * valid, else the symmetric-tensor contraction (6 components xx xy xz yy yz
 * zz: the off-diagonals twice) when symmetric, else the sum over components.
 * Pinned by: cross(F, F) equals or_power_spectrum; the polarization identity
 * cross(F, G) = (P(F + G) - P(F) - P(G)) / 2 on power and occupation; a
 * single GRID mode lands in the closed-form bin; the symmetric contraction
 * equals the 9-component sum over a full symmetric tensor (tests).
 * ------------------------------------------------------------------------ */
OR_API void or_cross_spectrum(int N, int ncomp, const double *f1FT, const double *f2FT, int symmetric, int comp1,
                              int comp2, int numbins, int deconvolve, int ktype, double *kbin, double *power,
                              double *kscatter, double *pscatter, int *occupation)
{
    const long P = N / 2 + 1, NK = (long)N * N * P;
    double *typek2 = (double *)malloc(sizeof(double) * N), *sinc = (double *)malloc(sizeof(double) * N);
    if (ktype == 0) {
        for (int i = 0; i < N; i++) { typek2[i] = 2. * (double)N * sin(M_PI * (double)i / (double)N); typek2[i] *= typek2[i]; }
    } else {
        for (int i = 0; i <= N / 2; i++) { typek2[i] = 2. * M_PI * i; typek2[i] *= typek2[i]; }
        for (int i = N / 2 + 1; i < N; i++) { typek2[i] = 2. * M_PI * (N - i); typek2[i] *= typek2[i]; }
    }
    sinc[0] = 1.;
    for (int i = 1; i <= N / 2; i++)
        sinc[i] = deconvolve ? sin(M_PI * (double)i / (double)N) * (double)N / (M_PI * (double)i) : 1.0;
    for (int i = N / 2 + 1; i < N; i++) sinc[i] = sinc[N - i];
    const double kmax = sqrt(3. * typek2[N / 2]);
    double *acc = (double *)calloc(4 * numbins, sizeof(double));
    long *occ = (long *)calloc(numbins, sizeof(long));
    for (long kz = 0; kz < N; kz++)
        for (long ky = 0; ky < N; ky++)
            for (long kx = 0; kx < P; kx++) {
                if (kx == 0 && ky == 0 && kz == 0) continue;
                const int weight = (kx == 0 || kx == N / 2) ? 1 : 2;
                const double k2 = typek2[kx] + typek2[ky] + typek2[kz];
                double s = sinc[kx] * sinc[ky] * sinc[kz];
                s *= s;
                const long i = kx + P * (ky + N * kz);
#define OR_RCP(a, b) (f1FT[2 * ((a) * NK + i)] * f2FT[2 * ((b) * NK + i)] + f1FT[2 * ((a) * NK + i) + 1] * f2FT[2 * ((b) * NK + i) + 1])
                double p;
                if (comp1 >= 0 && comp2 >= 0 && comp1 < ncomp && comp2 < ncomp) {
                    p = OR_RCP(comp1, comp2);
                } else if (symmetric) {
                    p = 2. * (OR_RCP(1, 1) + OR_RCP(2, 2) + OR_RCP(4, 4));
                    p += OR_RCP(0, 0) + OR_RCP(3, 3) + OR_RCP(5, 5);
                } else {
                    p = 0.;
                    for (int c = 0; c < ncomp; c++) p += OR_RCP(c, c);
                }
#undef OR_RCP
                const double sk = sqrt(k2);
                const int bin = (int)floor((double)numbins * sk / kmax);
                if (bin < numbins) {
                    acc[bin] += weight * sk;
                    acc[numbins + bin] += weight * k2;
                    acc[2 * numbins + bin] += weight * p * k2 * sk / s;
                    acc[3 * numbins + bin] += weight * p * p * k2 * k2 * k2 / s / s;
                    occ[bin] += weight;
                }
            }
    for (int b = 0; b < numbins; b++) {
        occupation[b] = (int)occ[b];
        if (occ[b] > 0) {
            kbin[b] = acc[b];
            kscatter[b] = sqrt(acc[numbins + b] * occ[b] - kbin[b] * kbin[b]) / occ[b];
            if (!isfinite(kscatter[b])) kscatter[b] = 0.;
            kbin[b] /= occ[b];
            power[b] = acc[2 * numbins + b] / occ[b];
            pscatter[b] = sqrt(acc[3 * numbins + b] / occ[b] - power[b] * power[b]);
            if (!isfinite(pscatter[b])) pscatter[b] = 0.;
        } else {
            kbin[b] = power[b] = kscatter[b] = pscatter[b] = 0.;
        }
    }
    free(typek2); free(sinc); free(acc); free(occ);
}

/* ------------------------------------------------------------------------
 * extractPowerSpectrum (tools.hpp:65-342), KTYPE_LINEAR, one component.
 * Pinned by: single-mode field lands in the expected bin with the expected
 * weight (tests).
 * ------------------------------------------------------------------------ */
OR_API void or_power_spectrum(int N, const double *fFT, int numbins, int deconvolve, double *kbin, double *power,
                              double *kscatter, double *pscatter, int *occupation)
{
    const long P = N / 2 + 1;
    double *typek2 = (double *)malloc(sizeof(double) * N), *sinc = (double *)malloc(sizeof(double) * N);
    for (int i = 0; i <= N / 2; i++) { typek2[i] = 2. * M_PI * i; typek2[i] *= typek2[i]; }
    for (int i = N / 2 + 1; i < N; i++) { typek2[i] = 2. * M_PI * (N - i); typek2[i] *= typek2[i]; }
    sinc[0] = 1.;
    for (int i = 1; i <= N / 2; i++)
        sinc[i] = deconvolve ? sin(M_PI * (double)i / (double)N) * (double)N / (M_PI * (double)i) : 1.0;
    for (int i = N / 2 + 1; i < N; i++) sinc[i] = sinc[N - i];
    double kmax = sqrt(3. * typek2[N / 2]);
    double *acc = (double *)calloc(4 * numbins, sizeof(double));
    long *occ = (long *)calloc(numbins, sizeof(long));
    for (long kz = 0; kz < N; kz++)
        for (long ky = 0; ky < N; ky++)
            for (long kx = 0; kx < P; kx++) {
                if (kx == 0 && ky == 0 && kz == 0) continue;
                int weight = (kx == 0 || kx == N / 2) ? 1 : 2;
                double k2 = typek2[kx] + typek2[ky] + typek2[kz];
                double s = sinc[kx] * sinc[ky] * sinc[kz];
                s *= s;
                long i = kx + P * (ky + N * kz);
                double re = fFT[2 * i], im = fFT[2 * i + 1];
                double p = re * re + im * im;
                double sk = sqrt(k2);
                int bin = (int)floor((double)numbins * sk / kmax);
                if (bin < numbins) {
                    acc[bin] += weight * sk;
                    acc[numbins + bin] += weight * k2;
                    acc[2 * numbins + bin] += weight * p * k2 * sk / s;
                    acc[3 * numbins + bin] += weight * p * p * k2 * k2 * k2 / s / s;
                    occ[bin] += weight;
                }
            }
    for (int b = 0; b < numbins; b++) {
        occupation[b] = (int)occ[b];
        if (occ[b] > 0) {
            kbin[b] = acc[b];
            kscatter[b] = sqrt(acc[numbins + b] * occ[b] - kbin[b] * kbin[b]) / occ[b];
            if (!isfinite(kscatter[b])) kscatter[b] = 0.;
            kbin[b] /= occ[b];
            power[b] = acc[2 * numbins + b] / occ[b];
            pscatter[b] = sqrt(acc[3 * numbins + b] / occ[b] - power[b] * power[b]);
            if (!isfinite(pscatter[b])) pscatter[b] = 0.;
        } else {
            kbin[b] = power[b] = kscatter[b] = pscatter[b] = 0.;
        }
    }
    free(typek2); free(sinc); free(acc); free(occ);
}

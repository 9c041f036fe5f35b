// gev_kspace.hip -- fused k-space kernels: one thread per Fourier mode, each
// reading its inputs once and writing its outputs once (HBM-bound).  The
// per-mode factors are evaluated in FP64 for both precision forms; the mixed
// form stores FP32 but builds the chi constraint combination in FP64.
#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_kops.hpp"

namespace gev {

constexpr int KS_NT = 256;

struct kidx {
    int kx, ky, kz;
    bool valid;
};

__device__ __forceinline__ kidx mode_of(int64_t e, int P, int N)
{
    kidx k;
    k.kx = (int)(e % P);
    int64_t r = e / P;
    k.ky = (int)(r % N);
    k.kz = (int)(r / N);
    k.valid = k.kx <= N / 2;
    return k;
}

template <typename T> __device__ __forceinline__ cx<double> ld(const cx<T> *p, int64_t i)
{
    cx<T> v = p[i];
    return mk((double)v.x, (double)v.y);
}
template <typename T> __device__ __forceinline__ void st(cx<T> *p, int64_t i, cx<double> v)
{
    p[i] = mk((T)v.x, (T)v.y);
}

// solveModifiedPoissonFT, gevolution.hpp:806-852
template <typename T>
__global__ __launch_bounds__(KS_NT) void k_poisson(const cx<T> *__restrict__ src, cx<T> *__restrict__ pot, int N,
                                                   int P, int64_t NK, const double *__restrict__ gk2, double coeff,
                                                   double modif)
{
    int64_t e = (int64_t)blockIdx.x * KS_NT + threadIdx.x;
    if (e >= NK) return;
    kidx k = mode_of(e, P, N);
    if (!k.valid) return;
    if (modif == 0.0 && k.kx == 0 && k.ky == 0 && k.kz == 0) {
        pot[e] = mk((T)0, (T)0);
        return;
    }
    st(pot, e, kop_poisson(ld(src, e), gk2[k.kx] + gk2[k.ky] + gk2[k.kz], coeff, modif));
}

// projectFTscalar, gevolution.hpp:251-340
template <typename T>
__global__ __launch_bounds__(KS_NT) void k_project_ft_scalar(const cx<T> *__restrict__ S, cx<T> *__restrict__ chi,
                                                             int N, int P, int64_t NK, const double *__restrict__ gk2,
                                                             const cx<double> *__restrict__ ks, int add)
{
    int64_t e = (int64_t)blockIdx.x * KS_NT + threadIdx.x;
    if (e >= NK) return;
    kidx k = mode_of(e, P, N);
    if (!k.valid) return;
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
        chi[e] = mk((T)0, (T)0);
        return;
    }
    double g0 = gk2[k.kx], g1 = gk2[k.ky], g2 = gk2[k.kz];
    cx<double> k0 = ks[k.kx], k1 = ks[k.ky], k2 = ks[k.kz];
    cx<double> acc = mk(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < 6; c++) acc = acc + kop_scalar_weight(c, g0, g1, g2, k0, k1, k2) * ld(S, e + c * NK);
    cx<double> r = acc * kop_scalar_norm(g0 + g1 + g2, N);
    if (add) r = r + ld(chi, e);
    st(chi, e, r);
}

// projectFTvector, gevolution.hpp:449-515 (in place allowed)
template <typename T>
__global__ __launch_bounds__(KS_NT) void k_project_ft_vector(const cx<T> *__restrict__ S, cx<T> *__restrict__ B,
                                                             int N, int P, int64_t NK, const double *__restrict__ gk2,
                                                             const cx<double> *__restrict__ ks, double coeff,
                                                             double modif)
{
    int64_t e = (int64_t)blockIdx.x * KS_NT + threadIdx.x;
    if (e >= NK) return;
    kidx k = mode_of(e, P, N);
    if (!k.valid) return;
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
        B[e] = B[e + NK] = B[e + 2 * NK] = mk((T)0, (T)0);
        return;
    }
    double kk = gk2[k.kx] + gk2[k.ky] + gk2[k.kz];
    cx<double> k0 = ks[k.kx], k1 = ks[k.ky], k2 = ks[k.kz];
    const cx<double> sv[3] = {ld(S, e), ld(S, e + NK), ld(S, e + 2 * NK)};
    cx<double> b[3];
    kop_vector(sv, kk, k0, k1, k2, coeff, modif, b);
    st(B, e, b[0]);
    st(B, e + NK, b[1]);
    st(B, e + 2 * NK, b[2]);
}

// evolveFTvector, gevolution.hpp:359-429
template <typename T>
__global__ __launch_bounds__(KS_NT) void k_evolve_ft_vector(const cx<T> *__restrict__ S, cx<T> *__restrict__ B,
                                                            int N, int P, int64_t NK, const double *__restrict__ gk2,
                                                            const cx<double> *__restrict__ ks, double a2dtau)
{
    int64_t e = (int64_t)blockIdx.x * KS_NT + threadIdx.x;
    if (e >= NK) return;
    kidx k = mode_of(e, P, N);
    if (!k.valid) return;
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
        B[e] = B[e + NK] = B[e + 2 * NK] = mk((T)0, (T)0);
        return;
    }
    double g0 = gk2[k.kx], g1 = gk2[k.ky], g2 = gk2[k.kz];
    cx<double> k0 = ks[k.kx], k1 = ks[k.ky], k2 = ks[k.kz];
    double k4 = g0 + g1 + g2;
    k4 *= k4;
    cx<double> pre = mk(0.0, -2.0 * a2dtau / k4);
    cx<double> S00 = ld(S, e), S01 = ld(S, e + NK), S02 = ld(S, e + 2 * NK), S11 = ld(S, e + 3 * NK),
               S12 = ld(S, e + 4 * NK), S22 = ld(S, e + 5 * NK);
    cx<double> t0 = conj(k0) * (S00 * (g1 + g2) - S11 * g1 - S22 * g2 - (k1 * k2 * S12) * 2.0) +
                    (k1 * S01 + k2 * S02) * (g1 + g2 - g0);
    cx<double> t1 = conj(k1) * (S11 * (g0 + g2) - S00 * g0 - S22 * g2 - (k0 * k2 * S02) * 2.0) +
                    (k0 * S01 + k2 * S12) * (g0 + g2 - g1);
    cx<double> t2 = conj(k2) * (S22 * (g0 + g1) - S00 * g0 - S11 * g1 - (k0 * k1 * S01) * 2.0) +
                    (k0 * S02 + k1 * S12) * (g0 + g1 - g2);
    st(B, e, ld(B, e) + pre * t0);
    st(B, e + NK, ld(B, e + NK) + pre * t1);
    st(B, e + 2 * NK, ld(B, e + 2 * NK) + pre * t2);
}

// ------------------------------------------------------------------ IC ----
// Philox-4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    }
}

__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo)
{
    uint64_t v = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}

// natural cubic spline, knots x, values y, second derivatives m (host-solved)
__device__ double spline_eval(int n, const double *x, const double *y, const double *m, double xv)
{
    int lo = 0, hi = n - 1;
    if (xv <= x[0]) hi = 1;
    else if (xv >= x[n - 1]) lo = n - 2;
    else {
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (x[mid] > xv) hi = mid; else lo = mid;
        }
    }
    lo = (hi - 1 < lo) ? lo : hi - 1;
    double h = x[lo + 1] - x[lo], t = xv - x[lo];
    double m0 = m[lo], m1 = m[lo + 1];
    double slope = (y[lo + 1] - y[lo]) / h - h * (2.0 * m0 + m1) / 6.0;
    return y[lo] + t * (slope + t * (0.5 * m0 + t * (m1 - m0) / (6.0 * h)));
}

__device__ __forceinline__ double sinc_n(int N, int n)
{
    if (n == 0) return 1.0;
    double x = M_PI * (double)n / (double)N;
    return sin(x) / x;
}

// generateDisplacementField, ic_basic.hpp:1164-1270 (Philox realisation;
// kx = 0 Hermitian pairs drawn once from the canonical member).
template <typename T>
__global__ __launch_bounds__(KS_NT) void k_ic_displacement(int N, int P, int64_t NK, uint32_t seed, int ksphere,
                                                           double coeff, int npk, const double *__restrict__ sx,
                                                           const double *__restrict__ sy,
                                                           const double *__restrict__ sm, const cx<T> *__restrict__ ker,
                                                           cx<T> *__restrict__ pot)
{
    int64_t e = (int64_t)blockIdx.x * KS_NT + threadIdx.x;
    if (e >= NK) return;
    kidx k = mode_of(e, P, N);
    if (!k.valid) return;
    const int kmax = N / 2 - 1;
    const int64_t kmax2 = (int64_t)kmax * kmax;
    int ny = (k.ky <= N / 2) ? k.ky : k.ky - N;
    int nz = (k.kz <= N / 2) ? k.kz : k.kz - N;
    int64_t k2i = (int64_t)k.kx * k.kx + (int64_t)ny * ny + (int64_t)nz * nz;
    if (k2i == 0 || k.kx >= kmax || abs(ny) >= kmax || abs(nz) >= kmax || (ksphere && k2i >= kmax2)) {
        pot[e] = mk((T)0, (T)0);
        return;
    }
    double sc = sinc_n(N, k.kx) * sinc_n(N, ny) * sinc_n(N, nz);
    double k2 = 4.0 * M_PI * M_PI * (double)k2i;
    int cy = ny, cz = nz, cj = 0;
    if (k.kx == 0 && (ny < 0 || (ny == 0 && nz < 0))) { cy = -ny; cz = -nz; cj = 1; }
    uint32_t c[4] = {(uint32_t)k.kx, (uint32_t)cy, (uint32_t)cz, 0u};
    philox4x32_10(c, seed, 0x67657631u);
    double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]);
    double r = sqrt(-2.0 * log(u1));
    double re = r * cos(2.0 * M_PI * u2);
    double im = (cj ? -1.0 : 1.0) * r * sin(2.0 * M_PI * u2);
    double mag = spline_eval(npk, sx, sy, sm, sqrt(k2)) * sc * (1.0 + 7.5 * coeff / k2);
    re *= mag;
    im *= mag;
    double kr = (double)ker[e].x, ki = (double)ker[e].y;
    double den = kr * kr + ki * ki;
    pot[e] = mk((T)((re * kr + im * ki) / den), (T)((im * kr - re * ki) / den));
}

template <typename T> __global__ void k_set_k0(cx<T> *ft, double re, double im) { ft[0] = mk((T)re, (T)im); }

// ------------------------------------------------------------ launchers ---
static inline unsigned nblk(int64_t n) { return (unsigned)((n + KS_NT - 1) / KS_NT); }

template <typename T> int poisson_impl(gev_ctx *ctx, const cx<T> *src, cx<T> *pot, double coeff, double modif)
{
    coeff /= -((double)ctx->N * (double)ctx->N * (double)ctx->N);
    hipLaunchKernelGGL((k_poisson<T>), dim3(nblk(ctx->NK)), dim3(KS_NT), 0, ctx->stream, src, pot, ctx->N, ctx->P,
                       ctx->NK, ctx->gk2, coeff, modif);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int project_ft_scalar_impl(gev_ctx *ctx, const cx<T> *S, cx<T> *chi, int add)
{
    hipLaunchKernelGGL((k_project_ft_scalar<T>), dim3(nblk(ctx->NK)), dim3(KS_NT), 0, ctx->stream, S, chi, ctx->N,
                       ctx->P, ctx->NK, ctx->gk2, ctx->kshift, add);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int project_ft_vector_impl(gev_ctx *ctx, const cx<T> *S, cx<T> *B, double coeff, double modif)
{
    hipLaunchKernelGGL((k_project_ft_vector<T>), dim3(nblk(ctx->NK)), dim3(KS_NT), 0, ctx->stream, S, B, ctx->N,
                       ctx->P, ctx->NK, ctx->gk2, ctx->kshift, coeff, modif);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int evolve_ft_vector_impl(gev_ctx *ctx, const cx<T> *S, cx<T> *B, double a2dtau)
{
    hipLaunchKernelGGL((k_evolve_ft_vector<T>), dim3(nblk(ctx->NK)), dim3(KS_NT), 0, ctx->stream, S, B, ctx->N,
                       ctx->P, ctx->NK, ctx->gk2, ctx->kshift, a2dtau);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T>
int ic_displacement_ft_impl(gev_ctx *ctx, uint32_t seed, int ksphere, double coeff, int npk, const double *dx,
                            const double *dy, const double *dm, const cx<T> *ker, cx<T> *pot)
{
    hipLaunchKernelGGL((k_ic_displacement<T>), dim3(nblk(ctx->NK)), dim3(KS_NT), 0, ctx->stream, ctx->N, ctx->P,
                       ctx->NK, seed, ksphere, coeff, npk, dx, dy, dm, ker, pot);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int set_k0_impl(gev_ctx *ctx, cx<T> *ft, double re, double im)
{
    hipLaunchKernelGGL((k_set_k0<T>), dim3(1), dim3(1), 0, ctx->stream, ft, re, im);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

#define INST(T)                                                                                              \
    template int poisson_impl<T>(gev_ctx *, const cx<T> *, cx<T> *, double, double);                        \
    template int project_ft_scalar_impl<T>(gev_ctx *, const cx<T> *, cx<T> *, int);                          \
    template int project_ft_vector_impl<T>(gev_ctx *, const cx<T> *, cx<T> *, double, double);               \
    template int evolve_ft_vector_impl<T>(gev_ctx *, const cx<T> *, cx<T> *, double);                        \
    template int ic_displacement_ft_impl<T>(gev_ctx *, uint32_t, int, double, int, const double *,           \
                                            const double *, const double *, const cx<T> *, cx<T> *);         \
    template int set_k0_impl<T>(gev_ctx *, cx<T> *, double, double);
INST(double)
INST(float)
#undef INST

}  // namespace gev

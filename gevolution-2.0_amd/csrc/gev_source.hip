// gev_source.hip -- site-parallel real-space source construction
// (prepareFTsource T00 and Tij forms with the PHINONLINEAR terms) and
// deterministic FP64 reductions.
#include "gev_common.hpp"
#include "gev_kernels.hpp"

namespace gev {

constexpr int SRC_NT = 256;
constexpr int RED_MAX_BLOCKS = 2048;

// block-sum NV doubles into partials[blockIdx.x*NV + v] (fixed tree: deterministic)
template <int NV> __device__ __forceinline__ void block_sum_store(double (&v)[NV], double *partials)
{
    __shared__ double sh[NV][SRC_NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double s = wave_sum(v[i]);
        if (lane == 0) sh[i][w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) {
            double s = 0.0;
            for (int j = 0; j < SRC_NT / 64; j++) s += sh[i][j];
            partials[(int64_t)blockIdx.x * NV + i] = s;
        }
    }
}

// prepareFTsource T00 form, gevolution.hpp:160-171
template <typename T>
__global__ __launch_bounds__(SRC_NT) void k_prepare_T00(const T *__restrict__ phi, const T *__restrict__ chi,
                                                        const T *src, T *res, int N, int64_t M, double bg,
                                                        double c1, double c2, double c3, double *partials)
{
    double acc[2] = {0.0, 0.0};
    const int lg = __ffs(N) - 1;
    for (int64_t i = (int64_t)blockIdx.x * SRC_NT + threadIdx.x; i < M; i += (int64_t)gridDim.x * SRC_NT) {
        int x = (int)(i & (N - 1)), y = (int)((i >> lg) & (N - 1)), z = (int)(i >> (2 * lg));
        T s = src[i];
        T p = phi[i];
        acc[0] += (double)s;
        T r = (T)c2 * (s - (T)bg);
        r *= (T)1 - (T)2 * p * ((T)1 - p);
        T d0 = phi[site(N, x - 1, y, z)] - phi[site(N, x + 1, y, z)];
        T d1 = phi[site(N, x, y - 1, z)] - phi[site(N, x, y + 1, z)];
        T d2 = phi[site(N, x, y, z - 1)] - phi[site(N, x, y, z + 1)];
        r += (T)0.125 * d0 * d0;
        r += (T)0.125 * d1 * d1;
        r += (T)0.125 * d2 * d2;
        r += ((T)c3 * ((T)1 - (T)3 * p) - (T)c1) * p - (T)c3 * chi[i];
        res[i] = r;
        acc[1] += (double)r;
    }
    block_sum_store<2>(acc, partials);
}

// prepareFTsource Tij form, gevolution.hpp:64-107 (Sij may alias Tij)
template <typename T>
__global__ __launch_bounds__(SRC_NT) void k_prepare_Tij(const T *__restrict__ phi, const T *Tij, T *Sij, int N,
                                                        int64_t M, double coeff)
{
    const int lg = __ffs(N) - 1;
    const T c = (T)coeff;
    for (int64_t i = (int64_t)blockIdx.x * SRC_NT + threadIdx.x; i < M; i += (int64_t)gridDim.x * SRC_NT) {
        int x = (int)(i & (N - 1)), y = (int)((i >> lg) & (N - 1)), z = (int)(i >> (2 * lg));
        T p = phi[i];
        T px = phi[site(N, x + 1, y, z)], py = phi[site(N, x, y + 1, z)], pz = phi[site(N, x, y, z + 1)];
        T mx = phi[site(N, x - 1, y, z)], my = phi[site(N, x, y - 1, z)], mz = phi[site(N, x, y, z - 1)];
        T pxy = phi[site(N, x + 1, y + 1, z)], pxz = phi[site(N, x + 1, y, z + 1)], pyz = phi[site(N, x, y + 1, z + 1)];
        T t00 = Tij[i], t01 = Tij[i + M], t02 = Tij[i + 2 * M], t11 = Tij[i + 3 * M], t12 = Tij[i + 4 * M],
          t22 = Tij[i + 5 * M];
        Sij[i] = (T)0.5 * (px - mx) * (px - mx) + c * t00;
        Sij[i + 3 * M] = (T)0.5 * (py - my) * (py - my) + c * t11;
        Sij[i + 5 * M] = (T)0.5 * (pz - mz) * (pz - mz) + c * t22;
        Sij[i + M] = px * py - p * pxy + (T)0.5 * (p * p - px * px - py * py + pxy * pxy) + c * t01;
        Sij[i + 2 * M] = px * pz - p * pxz + (T)0.5 * (p * p - px * px - pz * pz + pxz * pxz) + c * t02;
        Sij[i + 4 * M] = py * pz - p * pyz + (T)0.5 * (p * p - py * py - pz * pz + pyz * pyz) + c * t12;
    }
}

template <typename T>
__global__ __launch_bounds__(SRC_NT) void k_field_sum(const T *__restrict__ f, int64_t n, double *partials)
{
    double acc[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * SRC_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * SRC_NT)
        acc[0] += (double)f[i];
    block_sum_store<1>(acc, partials);
}

template <typename T> __global__ void k_fill(T *f, int64_t n, T v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        f[i] = v;
}

// sums nparts rows of nvals partials in a fixed order
__global__ void k_reduce_partials(const double *partials, int nparts, int nvals, double *out)
{
    __shared__ double sh[256];
    for (int v = 0; v < nvals; v++) {
        double s = 0.0;
        for (int j = threadIdx.x; j < nparts; j += 256) s += partials[(int64_t)j * nvals + v];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[v] = sh[0];
        __syncthreads();
    }
}

static unsigned grid_for(int64_t n)
{
    int64_t b = (n + SRC_NT - 1) / SRC_NT;
    if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
    if (b < 1) b = 1;
    return (unsigned)b;
}

int reduce_partials(gev_ctx *ctx, const double *partials, int nparts, int nvals, double *out_dev)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, ctx->stream, partials, nparts, nvals, out_dev);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

// ctx->red layout: [0..63] results, [64..) partials
template <typename T>
int prepare_T00_impl(gev_ctx *ctx, const T *phi, const T *chi, const T *src, double bg, T *res, double c1, double c2,
                     double c3, double *sums)
{
    unsigned g = grid_for(ctx->M);
    double *parts = ctx->red + 64;
    hipLaunchKernelGGL((k_prepare_T00<T>), dim3(g), dim3(SRC_NT), 0, ctx->stream, phi, chi, src, res, ctx->N, ctx->M,
                       bg, c1, c2, c3, parts);
    GEV_HIP(hipGetLastError());
    int rc = reduce_partials(ctx, parts, (int)g, 2, ctx->red);
    if (rc) return rc;
    GEV_HIP(hipMemcpyAsync(ctx->red_host, ctx->red, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GEV_HIP(hipStreamSynchronize(ctx->stream));
    sums[0] = ctx->red_host[0];
    sums[1] = ctx->red_host[1];
    return GEV_OK;
}

template <typename T> int prepare_Tij_impl(gev_ctx *ctx, const T *phi, const T *Tij, T *Sij, double coeff)
{
    unsigned g = grid_for(ctx->M);
    hipLaunchKernelGGL((k_prepare_Tij<T>), dim3(g), dim3(SRC_NT), 0, ctx->stream, phi, Tij, Sij, ctx->N, ctx->M,
                       coeff);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int field_sum_impl(gev_ctx *ctx, const T *f, int64_t count, double *sum)
{
    unsigned g = grid_for(count);
    double *parts = ctx->red + 64;
    hipLaunchKernelGGL((k_field_sum<T>), dim3(g), dim3(SRC_NT), 0, ctx->stream, f, count, parts);
    GEV_HIP(hipGetLastError());
    int rc = reduce_partials(ctx, parts, (int)g, 1, ctx->red);
    if (rc) return rc;
    GEV_HIP(hipMemcpyAsync(ctx->red_host, ctx->red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GEV_HIP(hipStreamSynchronize(ctx->stream));
    *sum = ctx->red_host[0];
    return GEV_OK;
}

template <typename T> int field_fill_impl(gev_ctx *ctx, T *f, int64_t count, T v)
{
    int64_t b = (count + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    hipLaunchKernelGGL((k_fill<T>), dim3((unsigned)b), dim3(256), 0, ctx->stream, f, count, v);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

#define INST(T)                                                                                                 \
    template int prepare_T00_impl<T>(gev_ctx *, const T *, const T *, const T *, double, T *, double, double,   \
                                     double, double *);                                                         \
    template int prepare_Tij_impl<T>(gev_ctx *, const T *, const T *, T *, double);                             \
    template int field_sum_impl<T>(gev_ctx *, const T *, int64_t, double *);                                    \
    template int field_fill_impl<T>(gev_ctx *, T *, int64_t, T);
INST(double)
INST(float)
#undef INST
template int field_fill_impl<int32_t>(gev_ctx *, int32_t *, int64_t, int32_t);

}  // namespace gev

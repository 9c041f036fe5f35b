// gev_particles.hip -- particle container kernels for CDNA4:
//   * re-binning: counting sort by cell (rank via atomic count, 3-phase
//     exclusive scan, scatter into ping-pong buffers);
//   * CIC projections of T00 / T0i / Tij as a two-phase gather: per tile of
//     sites, every cell of the tile plus its lower halo walks its (cell-sorted)
//     particles and accumulates the per-corner contributions in registers,
//     parks them in LDS, then each site sums the 8 (or 4, 2) cells that touch
//     it and issues one plain coalesced store.  No global atomics, no zeroing
//     pass, and the result does not depend on thread scheduling;
//   * kick / drift / IC gathers: one thread per particle in cell order, so a
//     wavefront's stencil reads hit the same cache lines.
#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_pcommon.hpp"

namespace gev {


// ------------------------------------------------------------- lattice ----
template <typename T> __global__ void k_lattice(PV<T> v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n) return;
    int x, y, z;
    cell_coords((int32_t)i, v.N, v.lg, x, y, z);
    if constexpr (sizeof(T) == 8) {
        v.p0[i] = ((double)x + 0.5) / (double)v.N;
        v.p1[i] = ((double)y + 0.5) / (double)v.N;
        v.p2[i] = ((double)z + 0.5) / (double)v.N;
    } else {
        v.p0[i] = v.p1[i] = v.p2[i] = 0.5f;
    }
    v.q0[i] = v.q1[i] = v.q2[i] = (T)0;
    v.id[i] = i;
    v.cell[i] = (int32_t)i;
}

// ---------------------------------------------------------- load/store ----
template <typename T>
__global__ void k_load(PV<T> v, const double *__restrict__ pos, const double *__restrict__ q, const int64_t *__restrict__ id)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n) return;
    const int64_t n = v.n;
    const int N = v.N;
    double p[3] = {wrap_pos(pos[i]), wrap_pos(pos[n + i]), wrap_pos(pos[2 * n + i])};
    int c[3] = {cell1(p[0], N), cell1(p[1], N), cell1(p[2], N)};
    if constexpr (sizeof(T) == 8) {
        v.p0[i] = p[0]; v.p1[i] = p[1]; v.p2[i] = p[2];
    } else {
        float r[3];
        for (int d = 0; d < 3; d++) {
            double o = p[d] * (double)N - (double)c[d];
            float of = (float)o;
            if (of >= 1.0f) { of = 0.0f; c[d] = (c[d] + 1) & (N - 1); }
            if (of < 0.0f) of = 0.0f;
            r[d] = of;
        }
        v.p0[i] = r[0]; v.p1[i] = r[1]; v.p2[i] = r[2];
    }
    v.q0[i] = (T)q[i]; v.q1[i] = (T)q[n + i]; v.q2[i] = (T)q[2 * n + i];
    v.id[i] = id[i];
    v.cell[i] = (int32_t)((int64_t)c[0] + (int64_t)N * ((int64_t)c[1] + (int64_t)N * c[2]));
}

template <typename T> __global__ void k_store(PV<T> v, double *__restrict__ pos, double *__restrict__ q, int64_t *__restrict__ id)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n) return;
    const int64_t n = v.n;
    if constexpr (sizeof(T) == 8) {
        pos[i] = v.p0[i]; pos[n + i] = v.p1[i]; pos[2 * n + i] = v.p2[i];
    } else {
        int x, y, z;
        cell_coords(v.cell[i], v.N, v.lg, x, y, z);
        const double dx = 1.0 / (double)v.N;
        pos[i] = ((double)x + (double)v.p0[i]) * dx;
        pos[n + i] = ((double)y + (double)v.p1[i]) * dx;
        pos[2 * n + i] = ((double)z + (double)v.p2[i]) * dx;
    }
    q[i] = (double)v.q0[i]; q[n + i] = (double)v.q1[i]; q[2 * n + i] = (double)v.q2[i];
    id[i] = v.id[i];
}

// ------------------------------------------------------------- sorting ----
__global__ void k_count(const int32_t *__restrict__ cell, int32_t *__restrict__ rank, int32_t *count, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rank[i] = atomicAdd(&count[cell[i]], 1);
}

constexpr int SCAN_NT = 256, SCAN_PER = 8, SCAN_CHUNK = SCAN_NT * SCAN_PER;

__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t &total)
{
    __shared__ int32_t sh[SCAN_NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    int32_t off = 0, tot = 0;
    for (int j = 0; j < SCAN_NT / 64; j++) {
        if (j < w) off += sh[j];
        tot += sh[j];
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_reduce(const int32_t *__restrict__ cnt, int64_t M, int32_t *bsum)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK;
    int32_t s = 0;
    for (int j = 0; j < SCAN_PER; j++) {
        int64_t i = base + threadIdx.x + (int64_t)j * SCAN_NT;
        if (i < M) s += cnt[i];
    }
    int32_t tot;
    block_excl_scan(s, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_mid(int32_t *bsum, int64_t nb)
{
    // one block of 1024 threads, each owns a contiguous run
    __shared__ int32_t sh[1024];
    int64_t per = (nb + 1023) / 1024;
    int64_t b0 = threadIdx.x * per;
    int32_t s = 0;
    for (int64_t j = 0; j < per; j++)
        if (b0 + j < nb) s += bsum[b0 + j];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int32_t y = (threadIdx.x >= (unsigned)o) ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += y;
        __syncthreads();
    }
    int32_t run = sh[threadIdx.x] - s;
    for (int64_t j = 0; j < per; j++)
        if (b0 + j < nb) {
            int32_t t = bsum[b0 + j];
            bsum[b0 + j] = run;
            run += t;
        }
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(const int32_t *__restrict__ cnt, int64_t M,
                                                        const int32_t *__restrict__ bsum, int32_t *__restrict__ start,
                                                        int32_t total)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_PER;
    int32_t v[SCAN_PER];
    int32_t s = 0;
    for (int j = 0; j < SCAN_PER; j++) {
        int64_t i = base + j;
        v[j] = (i < M) ? cnt[i] : 0;
        s += v[j];
    }
    int32_t tot;
    int32_t off = block_excl_scan(s, tot) + bsum[blockIdx.x];
    for (int j = 0; j < SCAN_PER; j++) {
        int64_t i = base + j;
        if (i < M) start[i] = off;
        off += v[j];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[M] = total;
}

template <typename T> __global__ void k_scatter(PV<T> s, PV<T> d, const int32_t *__restrict__ rank)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int32_t c = s.cell[i];
    int64_t j = (int64_t)s.start[c] + rank[i];
    d.p0[j] = s.p0[i]; d.p1[j] = s.p1[i]; d.p2[j] = s.p2[i];
    d.q0[j] = s.q0[i]; d.q1[j] = s.q1[i]; d.q2[j] = s.q2[i];
    d.id[j] = s.id[i];
    d.cell[j] = c;
}

// ------------------------------------------------------------ deposits ----

// projection_T00 (functor gevolution.hpp:1729-1854): 8 corner values per cell
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(PT_NT) void k_dep_T00(PV<T> v, T *out, const T *__restrict__ phi, double mass_d,
                                                   double a_d, int accumulate)
{
    constexpr int CX = TX + 1, CY = TY + 1, CZ = TZ + 1, NC = CX * CY * CZ;
    __shared__ T lds[8][NC];
    const int N = v.N;
    int x0, y0, z0;
    tile_origin<TX, TY, TZ>(N, x0, y0, z0);
    const T a = (T)a_d;
    for (int c = threadIdx.x; c < NC; c += PT_NT) {
        int lx = c % CX, ly = (c / CX) % CY, lz = c / (CX * CY);
        int gx = (x0 - 1 + lx) & (N - 1), gy = (y0 - 1 + ly) & (N - 1), gz = (z0 - 1 + lz) & (N - 1);
        int64_t g = site(N, gx, gy, gz);
        T acc[8];
#pragma unroll
        for (int b = 0; b < 8; b++) acc[b] = (T)0;
        int32_t p0 = v.start[g], p1 = v.start[g + 1];
        if (phi) {
            T ph[8];
#pragma unroll
            for (int b = 0; b < 8; b++) ph[b] = phi[site(N, gx + ((b >> 2) & 1), gy + ((b >> 1) & 1), gz + (b & 1))];
            for (int32_t p = p0; p < p1; p++) {
                T r[3];
                ref_dist(v, p, gx, gy, gz, r);
                T q0 = v.q0[p], q1 = v.q1[p], q2 = v.q2[p];
                T f2 = q0 * q0 + q1 * q1 + q2 * q2;
                T e = sqrt(f2 + a * a);
                T f = (T)3 * e + f2 / e;
                f2 = ((T)0.5 * f * f + f2 / ((T)1 + f2 / a / a)) / e;
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    T w = (((b >> 2) & 1) ? r[0] : (T)1 - r[0]) * (((b >> 1) & 1) ? r[1] : (T)1 - r[1]) *
                          ((b & 1) ? r[2] : (T)1 - r[2]);
                    acc[b] += w * (e + f * ph[b] + f2 * ph[b] * ph[b]);
                }
            }
        } else {
            for (int32_t p = p0; p < p1; p++) {
                T r[3];
                ref_dist(v, p, gx, gy, gz, r);
#pragma unroll
                for (int b = 0; b < 8; b++)
                    acc[b] += (((b >> 2) & 1) ? r[0] : (T)1 - r[0]) * (((b >> 1) & 1) ? r[1] : (T)1 - r[1]) *
                              ((b & 1) ? r[2] : (T)1 - r[2]);
            }
        }
#pragma unroll
        for (int b = 0; b < 8; b++) lds[b][c] = acc[b];
    }
    __syncthreads();
    const T mass = (T)mass_d;
    for (int s = threadIdx.x; s < TX * TY * TZ; s += PT_NT) {
        int sx = s % TX, sy = (s / TX) % TY, sz = s / (TX * TY);
        T val = (T)0;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
            int lc = (sx + 1 - bx) + CX * ((sy + 1 - by) + CY * (sz + 1 - bz));
            val += lds[b][lc];
        }
        val *= mass;
        int64_t gs = site(N, x0 + sx, y0 + sy, z0 + sz);
        out[gs] = accumulate ? out[gs] + val : val;
    }
}

// projection_T0i (functor gevolution.hpp:2031-2201): 12 link values per cell.
// g[k] lands on component comp[k] of the site cell + off[k].
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(PT_NT) void k_dep_T0i(PV<T> v, T *out, const T *__restrict__ phi, double mass_d,
                                                   int64_t M, int accumulate)
{
    constexpr int CX = TX + 1, CY = TY + 1, CZ = TZ + 1, NC = CX * CY * CZ;
    __shared__ T lds[12][NC];
    const int N = v.N;
    int x0, y0, z0;
    tile_origin<TX, TY, TZ>(N, x0, y0, z0);
    for (int c = threadIdx.x; c < NC; c += PT_NT) {
        int lx = c % CX, ly = (c / CX) % CY, lz = c / (CX * CY);
        int gx = (x0 - 1 + lx) & (N - 1), gy = (y0 - 1 + ly) & (N - 1), gz = (z0 - 1 + lz) & (N - 1);
        int64_t g = site(N, gx, gy, gz);
        T acc[12];
#pragma unroll
        for (int k = 0; k < 12; k++) acc[k] = (T)0;
        T fac[12];
        if (phi) {
            T ph[8];
#pragma unroll
            for (int b = 0; b < 8; b++) ph[b] = phi[site(N, gx + ((b >> 2) & 1), gy + ((b >> 1) & 1), gz + (b & 1))];
            fac[0] = (T)1 + ph[0] + ph[4]; fac[1] = (T)1 + ph[0] + ph[2]; fac[2] = (T)1 + ph[0] + ph[1];
            fac[3] = (T)1 + ph[4] + ph[6]; fac[4] = (T)1 + ph[4] + ph[5]; fac[5] = (T)1 + ph[2] + ph[6];
            fac[6] = (T)1 + ph[2] + ph[3]; fac[7] = (T)1 + ph[1] + ph[5]; fac[8] = (T)1 + ph[1] + ph[3];
            fac[9] = (T)1 + ph[3] + ph[7]; fac[10] = (T)1 + ph[5] + ph[7]; fac[11] = (T)1 + ph[6] + ph[7];
        }
        int32_t p0 = v.start[g], p1 = v.start[g + 1];
        for (int32_t p = p0; p < p1; p++) {
            T r[3];
            ref_dist(v, p, gx, gy, gz, r);
            T v0 = v.q0[p], v1 = v.q1[p], v2 = v.q2[p];
            T s[12];
            s[0] = ((T)1 - r[1]) * ((T)1 - r[2]) * v0;
            s[1] = ((T)1 - r[0]) * ((T)1 - r[2]) * v1;
            s[2] = ((T)1 - r[0]) * ((T)1 - r[1]) * v2;
            s[3] = r[0] * ((T)1 - r[2]) * v1;
            s[4] = r[0] * ((T)1 - r[1]) * v2;
            s[5] = r[1] * ((T)1 - r[2]) * v0;
            s[6] = r[1] * ((T)1 - r[0]) * v2;
            s[7] = r[2] * ((T)1 - r[1]) * v0;
            s[8] = r[2] * ((T)1 - r[0]) * v1;
            s[9] = r[1] * r[2] * v0;
            s[10] = r[0] * r[2] * v1;
            s[11] = r[0] * r[1] * v2;
#pragma unroll
            for (int k = 0; k < 12; k++) acc[k] += s[k];
        }
        if (phi) {
#pragma unroll
            for (int k = 0; k < 12; k++) acc[k] *= fac[k];
        }
#pragma unroll
        for (int k = 0; k < 12; k++) lds[k][c] = acc[k];
    }
    __syncthreads();
    const T mass = (T)mass_d;
    for (int s = threadIdx.x; s < TX * TY * TZ; s += PT_NT) {
        int sx = s % TX, sy = (s / TX) % TY, sz = s / (TX * TY);
        auto L = [&](int k, int ox, int oy, int oz) {
            return lds[k][(sx + 1 - ox) + CX * ((sy + 1 - oy) + CY * (sz + 1 - oz))];
        };
        // component 0: g0 (0,0,0), g5 (0,1,0), g7 (0,0,1), g9 (0,1,1)
        T c0 = L(0, 0, 0, 0) + L(5, 0, 1, 0) + L(7, 0, 0, 1) + L(9, 0, 1, 1);
        // component 1: g1 (0,0,0), g3 (1,0,0), g8 (0,0,1), g10 (1,0,1)
        T c1 = L(1, 0, 0, 0) + L(3, 1, 0, 0) + L(8, 0, 0, 1) + L(10, 1, 0, 1);
        // component 2: g2 (0,0,0), g4 (1,0,0), g6 (0,1,0), g11 (1,1,0)
        T c2 = L(2, 0, 0, 0) + L(4, 1, 0, 0) + L(6, 0, 1, 0) + L(11, 1, 1, 0);
        int64_t gs = site(N, x0 + sx, y0 + sy, z0 + sz);
        c0 *= mass; c1 *= mass; c2 *= mass;
        if (accumulate) { c0 += out[gs]; c1 += out[gs + M]; c2 += out[gs + 2 * M]; }
        out[gs] = c0; out[gs + M] = c1; out[gs + 2 * M] = c2;
    }
}

// projection_Tij (functor gevolution.hpp:2582-2965, no CIC_PROJECT_TIJ):
// per cell 24 diagonal corner values + 6 off-diagonal edge values.
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(PT_NT) void k_dep_Tij(PV<T> v, T *out, const T *__restrict__ phi, double mass_d,
                                                   double a_d, int64_t M, int accumulate)
{
    constexpr int CX = TX + 1, CY = TY + 1, CZ = TZ + 1, NC = CX * CY * CZ;
    __shared__ T lds[30][NC];
    const int N = v.N;
    int x0, y0, z0;
    tile_origin<TX, TY, TZ>(N, x0, y0, z0);
    const T a = (T)a_d;
    for (int c = threadIdx.x; c < NC; c += PT_NT) {
        int lx = c % CX, ly = (c / CX) % CY, lz = c / (CX * CY);
        int gx = (x0 - 1 + lx) & (N - 1), gy = (y0 - 1 + ly) & (N - 1), gz = (z0 - 1 + lz) & (N - 1);
        int64_t g = site(N, gx, gy, gz);
        T acc[30];
#pragma unroll
        for (int k = 0; k < 30; k++) acc[k] = (T)0;
        T ph[8];
        if (phi) {
#pragma unroll
            for (int b = 0; b < 8; b++) ph[b] = phi[site(N, gx + ((b >> 2) & 1), gy + ((b >> 1) & 1), gz + (b & 1))];
        }
        int32_t p0 = v.start[g], p1 = v.start[g + 1];
        for (int32_t p = p0; p < p1; p++) {
            T r[3];
            ref_dist(v, p, gx, gy, gz, r);
            T vv[3] = {v.q0[p], v.q1[p], v.q2[p]};
            T f = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
            T e = sqrt(f + a * a);
            if (phi) {
                f = (T)4 + a * a / (f + a * a);
                f *= ((T)1 - r[0]) * (((T)1 - r[1]) * (((T)1 - r[2]) * ph[0] + r[2] * ph[1]) +
                                      r[1] * (((T)1 - r[2]) * ph[2] + r[2] * ph[3])) +
                     r[0] * (((T)1 - r[1]) * (((T)1 - r[2]) * ph[4] + r[2] * ph[5]) +
                             r[1] * (((T)1 - r[2]) * ph[6] + r[2] * ph[7]));
                f += (T)1;
            } else {
                f = (T)1;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                T w = f * vv[d] * vv[d] / e;
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    T wt = w * (((b >> 2) & 1) ? r[0] : (T)1 - r[0]) * (((b >> 1) & 1) ? r[1] : (T)1 - r[1]) *
                           ((b & 1) ? r[2] : (T)1 - r[2]);
                    acc[d * 8 + b] += wt;
                }
            }
            T w = f * vv[0] * vv[1] / e;
            acc[24] += w * ((T)1 - r[2]);
            acc[25] += w * r[2];
            w = f * vv[0] * vv[2] / e;
            acc[26] += w * ((T)1 - r[1]);
            acc[27] += w * r[1];
            w = f * vv[1] * vv[2] / e;
            acc[28] += w * ((T)1 - r[0]);
            acc[29] += w * r[0];
        }
#pragma unroll
        for (int k = 0; k < 30; k++) lds[k][c] = acc[k];
    }
    __syncthreads();
    const T mass = (T)mass_d;
    for (int s = threadIdx.x; s < TX * TY * TZ; s += PT_NT) {
        int sx = s % TX, sy = (s / TX) % TY, sz = s / (TX * TY);
        auto L = [&](int k, int ox, int oy, int oz) {
            return lds[k][(sx + 1 - ox) + CX * ((sy + 1 - oy) + CY * (sz + 1 - oz))];
        };
        T d[3];
#pragma unroll
        for (int dd = 0; dd < 3; dd++) {
            T val = (T)0;
#pragma unroll
            for (int b = 0; b < 8; b++) val += L(dd * 8 + b, (b >> 2) & 1, (b >> 1) & 1, b & 1);
            d[dd] = val * mass;
        }
        T o01 = (L(24, 0, 0, 0) + L(25, 0, 0, 1)) * mass;
        T o02 = (L(26, 0, 0, 0) + L(27, 0, 1, 0)) * mass;
        T o12 = (L(28, 0, 0, 0) + L(29, 1, 0, 0)) * mass;
        int64_t gs = site(N, x0 + sx, y0 + sy, z0 + sz);
        T vals[6] = {d[0], o01, o02, d[1], o12, d[2]};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            T o = vals[k];
            if (accumulate) o += out[gs + k * M];
            out[gs + k * M] = o;
        }
    }
}


// displace_pcls_ic_basic, ic_basic.hpp:65-96 (one field)
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_displace(PV<T> v, const T *__restrict__ xi, double coeff,
                                                    unsigned long long *maxslot)
{
    int64_t i = (int64_t)blockIdx.x * PT_NT + threadIdx.x;
    double ret = 0.0;
    if (i < v.n) {
        const int N = v.N;
        int x, y, z;
        cell_coords(v.cell[i], N, v.lg, x, y, z);
        T r[3];
        ref_dist(v, i, x, y, z, r);
        T g[3];
        grad1(FieldAt<T>{xi, N, x, y, z}, r, g);
        const T Nn = (T)N;
        g[0] *= Nn; g[1] *= Nn; g[2] *= Nn;
        ret = coeff * sqrt((double)(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]));
        double d[3] = {coeff * (double)g[0], coeff * (double)g[1], coeff * (double)g[2]};
        move_by(v, i, x, y, z, d);
    }
    block_max_to(ret, maxslot);
}

// initialize_q_ic_basic, ic_basic.hpp:133-168 (one field)
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_initialize_q(PV<T> v, const T *__restrict__ pot, double coeff,
                                                        unsigned long long *maxslot)
{
    int64_t i = (int64_t)blockIdx.x * PT_NT + threadIdx.x;
    double ret = 0.0;
    if (i < v.n) {
        const int N = v.N;
        int x, y, z;
        cell_coords(v.cell[i], N, v.lg, x, y, z);
        T r[3];
        ref_dist(v, i, x, y, z, r);
        T g[3];
        grad1(FieldAt<T>{pot, N, x, y, z}, r, g);
        const T Nn = (T)N, c = (T)coeff;
        T q[3];
        T v2 = (T)0;
#pragma unroll
        for (int l = 0; l < 3; l++) {
            q[l] = -(g[l] * Nn) * c;
            v2 += q[l] * q[l];
        }
        v.q0[i] = q[0]; v.q1[i] = q[1]; v.q2[i] = q[2];
        ret = (double)v2;
    }
    block_max_to(ret, maxslot);
}

// ============================================================ host side ===
static inline unsigned pblocks(int64_t n) { return (unsigned)((n + PT_NT - 1) / PT_NT); }

// scan + scatter, given per-cell counts and per-particle ranks
template <typename T> int pcls_sort_counted_impl(gev_pcls *p)
{
    gev_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const int64_t M = ctx->M, n = p->n;
    int64_t nb = (M + SCAN_CHUNK - 1) / SCAN_CHUNK;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_NT), 0, s, p->cell_count, M, ctx->scan_blocks);
    hipLaunchKernelGGL(k_scan_mid, dim3(1), dim3(1024), 0, s, ctx->scan_blocks, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_NT), 0, s, p->cell_count, M, ctx->scan_blocks,
                       p->cell_start, (int32_t)n);
    if (n > 0) {
        PV<T> src = make_pv<T>(p);
        PV<T> dst = src;
        dst.p0 = (T *)p->pos2[0]; dst.p1 = (T *)p->pos2[1]; dst.p2 = (T *)p->pos2[2];
        dst.q0 = (T *)p->q2[0]; dst.q1 = (T *)p->q2[1]; dst.q2 = (T *)p->q2[2];
        dst.id = p->id2;
        dst.cell = p->cell2;
        hipLaunchKernelGGL((k_scatter<T>), dim3(pblocks(n)), dim3(PT_NT), 0, s, src, dst, p->rank);
        for (int d = 0; d < 3; d++) {
            std::swap(p->pos[d], p->pos2[d]);
            std::swap(p->q[d], p->q2[d]);
        }
        std::swap(p->id, p->id2);
        std::swap(p->cell, p->cell2);
    }
    GEV_HIP(hipGetLastError());
    p->sorted = true;
    return GEV_OK;
}

template <typename T> int pcls_sort_impl(gev_pcls *p)
{
    gev_ctx *ctx = p->ctx;
    const int64_t M = ctx->M, n = p->n;
    int rc = field_fill_impl<int32_t>(ctx, p->cell_count, M, 0);
    if (rc) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_count, dim3(pblocks(n)), dim3(PT_NT), 0, ctx->stream, p->cell, p->rank, p->cell_count, n);
    GEV_HIP(hipGetLastError());
    return pcls_sort_counted_impl<T>(p);
}

template <typename T> int pcls_lattice_impl(gev_pcls *p)
{
    hipLaunchKernelGGL((k_lattice<T>), dim3(pblocks(p->n)), dim3(PT_NT), 0, p->ctx->stream, make_pv<T>(p));
    GEV_HIP(hipGetLastError());
    return pcls_sort_impl<T>(p);
}

template <typename T> int pcls_load_impl(gev_pcls *p, const double *pos, const double *q, const int64_t *id)
{
    hipLaunchKernelGGL((k_load<T>), dim3(pblocks(p->n)), dim3(PT_NT), 0, p->ctx->stream, make_pv<T>(p), pos, q, id);
    GEV_HIP(hipGetLastError());
    return pcls_sort_impl<T>(p);
}

template <typename T> int pcls_store_impl(gev_pcls *p, double *pos, double *q, int64_t *id)
{
    hipLaunchKernelGGL((k_store<T>), dim3(pblocks(p->n)), dim3(PT_NT), 0, p->ctx->stream, make_pv<T>(p), pos, q, id);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T, int TX, int TY, int TZ> static unsigned tiles(int N)
{
    return (unsigned)((int64_t)(N / TX) * (N / TY) * (N / TZ));
}

// tile shapes per deposit and precision (LDS = values * cells * sizeof(T))
template <typename T> int dep_T00_impl(gev_pcls *p, T *out, const T *phi, double mass, double a, int acc)
{
    gev_ctx *c = p->ctx;
    const int N = c->N;
    if (N >= 8)
        hipLaunchKernelGGL((k_dep_T00<T, 8, 8, 8>), dim3(tiles<T, 8, 8, 8>(N)), dim3(PT_NT), 0, c->stream,
                           make_pv<T>(p), out, phi, mass, a, acc);
    else
        hipLaunchKernelGGL((k_dep_T00<T, 4, 4, 4>), dim3(tiles<T, 4, 4, 4>(N)), dim3(PT_NT), 0, c->stream,
                           make_pv<T>(p), out, phi, mass, a, acc);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int dep_T0i_impl(gev_pcls *p, T *out, const T *phi, double mass, int acc)
{
    gev_ctx *c = p->ctx;
    const int N = c->N;
    if (N >= 8)
        hipLaunchKernelGGL((k_dep_T0i<T, 8, 8, 4>), dim3(tiles<T, 8, 8, 4>(N)), dim3(PT_NT), 0, c->stream,
                           make_pv<T>(p), out, phi, mass, c->M, acc);
    else
        hipLaunchKernelGGL((k_dep_T0i<T, 4, 4, 4>), dim3(tiles<T, 4, 4, 4>(N)), dim3(PT_NT), 0, c->stream,
                           make_pv<T>(p), out, phi, mass, c->M, acc);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int dep_Tij_impl(gev_pcls *p, T *out, const T *phi, double mass, double a, int acc)
{
    gev_ctx *c = p->ctx;
    const int N = c->N;
    if (N >= 8)
        hipLaunchKernelGGL((k_dep_Tij<T, 8, 4, 4>), dim3(tiles<T, 8, 4, 4>(N)), dim3(PT_NT), 0, c->stream,
                           make_pv<T>(p), out, phi, mass, a, c->M, acc);
    else
        hipLaunchKernelGGL((k_dep_Tij<T, 4, 4, 4>), dim3(tiles<T, 4, 4, 4>(N)), dim3(PT_NT), 0, c->stream,
                           make_pv<T>(p), out, phi, mass, a, c->M, acc);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

static int read_max(gev_ctx *c, double *out)
{
    GEV_HIP(hipMemcpyAsync(c->red_host, c->red + 8, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GEV_HIP(hipStreamSynchronize(c->stream));
    *out = c->red_host[0];
    return GEV_OK;
}

static int clear_max(gev_ctx *c)
{
    return field_fill_impl<double>(c, c->red + 8, 1, 0.0);
}

template <typename T> int displace_impl(gev_pcls *p, const T *xi, double coeff, double *max_disp)
{
    gev_ctx *c = p->ctx;
    int rc = clear_max(c);
    if (rc) return rc;
    if (p->n > 0)
        hipLaunchKernelGGL((k_displace<T>), dim3(pblocks(p->n)), dim3(PT_NT), 0, c->stream, make_pv<T>(p), xi, coeff,
                           (unsigned long long *)(c->red + 8));
    GEV_HIP(hipGetLastError());
    if (max_disp) {
        rc = read_max(c, max_disp);
        if (rc) return rc;
    }
    return pcls_sort_impl<T>(p);
}

template <typename T> int initialize_q_impl(gev_pcls *p, const T *pot, double coeff, double *max_ret)
{
    gev_ctx *c = p->ctx;
    int rc = clear_max(c);
    if (rc) return rc;
    if (p->n > 0)
        hipLaunchKernelGGL((k_initialize_q<T>), dim3(pblocks(p->n)), dim3(PT_NT), 0, c->stream, make_pv<T>(p), pot,
                           coeff, (unsigned long long *)(c->red + 8));
    GEV_HIP(hipGetLastError());
    return max_ret ? read_max(c, max_ret) : GEV_OK;
}

#define INST(T)                                                                                                   \
    template int pcls_sort_impl<T>(gev_pcls *);                                                                   \
    template int pcls_sort_counted_impl<T>(gev_pcls *);                                                           \
    template int pcls_lattice_impl<T>(gev_pcls *);                                                                \
    template int pcls_load_impl<T>(gev_pcls *, const double *, const double *, const int64_t *);                  \
    template int pcls_store_impl<T>(gev_pcls *, double *, double *, int64_t *);                                   \
    template int dep_T00_impl<T>(gev_pcls *, T *, const T *, double, double, int);                                \
    template int dep_T0i_impl<T>(gev_pcls *, T *, const T *, double, int);                                        \
    template int dep_Tij_impl<T>(gev_pcls *, T *, const T *, double, double, int);                                \
    template int displace_impl<T>(gev_pcls *, const T *, double, double *);                                       \
    template int initialize_q_impl<T>(gev_pcls *, const T *, double, double *);
INST(double)
INST(float)
#undef INST

}  // namespace gev

// gev_adjoint.hip -- reverse-mode kernels of the Newtonian particle-mesh cycle
// (update_q_Newton kick, update_pos_Newton drift, CIC mass deposit, Poisson
// solve; gevolution.hpp:1051-1118, 1270-1273, 1902-1905, 806-852) and of the
// lattice Zel'dovich initial map.
//
// One forward cycle, with K = dtau_k a_k N and C = dtau_d / a_d:
//   rho   = m' sum_p W(x_p)                      (CIC, W the 8 corner weights)
//   phi   = G rho                                (G: r2c, -coeff/(N^3 k^2), c2r)
//   q*    = q - K g(x; phi),  g_i = 1/2 sum_c w_c (phi(c+e_i) - phi(c-e_i))
//   x'    = x + C q*,  q' = q*
// Its adjoint, given (lx', lq'):
//   lq*   = lq' + C lx'
//   u_i   = sum_p w_c(x_p) lq*_{p,i}                                  (k_adj_kick_deposit)
//   lphi  = -K/2 sum_i (u_i(s - e_i) - u_i(s + e_i))                  (k_adj_divergence)
//   lrho  = G lphi                        (G is symmetric: the same Poisson solve)
//   lx    = lx' - K N/2 sum_i lq*_i grad1(D_i phi)(x) + m' N grad1(lrho)(x)   (k_adj_gather)
//   lq    = lq*
// with grad1 the CIC edge gradient (the exact derivative of the CIC
// interpolant inside a cell) and D_i phi(c) = phi(c+e_i) - phi(c-e_i).
// Adjoint per-particle arrays are indexed by particle ID, so they survive the
// re-binning of the container between cycles.
#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_pcommon.hpp"

namespace gev {

// lq* = lq + C lx (stored back), u_i += w_c lq*_i at the 8 corners
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_adj_kick_deposit(PV<T> v, const T *__restrict__ lx, T *lq, double C, T *u,
                                                            int64_t M)
{
    const int64_t i = (int64_t)blockIdx.x * PT_NT + threadIdx.x;
    if (i >= v.n) return;
    const int N = v.N;
    const int64_t n = v.n;
    const int64_t id = v.id[i];
    int x, y, z;
    cell_coords(v.cell[i], N, v.lg, x, y, z);
    T r[3];
    ref_dist(v, i, x, y, z, r);
    T l[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        l[d] = lq[d * n + id] + (T)C * lx[d * n + id];
        lq[d * n + id] = l[d];
    }
    const T o = (T)1;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
        const T w = (bx ? r[0] : o - r[0]) * (by ? r[1] : o - r[1]) * (bz ? r[2] : o - r[2]);
        const int64_t s = site(N, x + bx, y + by, z + bz);
#pragma unroll
        for (int d = 0; d < 3; d++) atomicAdd(&u[d * M + s], w * l[d]);
    }
}

// lphi(s) = -K/2 sum_i (u_i(s - e_i) - u_i(s + e_i))
template <typename T>
__global__ __launch_bounds__(256) void k_adj_divergence(const T *__restrict__ u, T *__restrict__ lphi, int N, int64_t M,
                                                        double K)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const int x = (int)(s % N), y = (int)((s / N) % N), z = (int)(s / ((int64_t)N * N));
    const T *u0 = u, *u1 = u + M, *u2 = u + 2 * M;
    const T d = (u0[site(N, x - 1, y, z)] - u0[site(N, x + 1, y, z)]) + (u1[site(N, x, y - 1, z)] - u1[site(N, x, y + 1, z)]) +
                (u2[site(N, x, y, z - 1)] - u2[site(N, x, y, z + 1)]);
    lphi[s] = (T)(-0.5 * K) * d;
}

// D_i phi at offset (dx,dy,dz) of the cell: phi(c+e_i) - phi(c-e_i)
template <typename T> struct CentralDiff {
    const T *f;
    int N, x, y, z, i;
    __device__ __forceinline__ T operator()(int dx, int dy, int dz) const
    {
        const int ex = i == 0, ey = i == 1, ez = i == 2;
        return f[site(N, x + dx + ex, y + dy + ey, z + dz + ez)] - f[site(N, x + dx - ex, y + dy - ey, z + dz - ez)];
    }
};

// lx += -K N/2 sum_i lq*_i grad1(D_i phi)(x) + mN grad1(lrho)(x)
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_adj_gather(PV<T> v, const T *__restrict__ phi, const T *__restrict__ lrho,
                                                      const T *__restrict__ lq, T *lx, double K, double mN)
{
    const int64_t i = (int64_t)blockIdx.x * PT_NT + threadIdx.x;
    if (i >= v.n) return;
    const int N = v.N;
    const int64_t n = v.n;
    const int64_t id = v.id[i];
    int x, y, z;
    cell_coords(v.cell[i], N, v.lg, x, y, z);
    T r[3];
    ref_dist(v, i, x, y, z, r);
    T acc[3] = {(T)0, (T)0, (T)0};
    if (lrho) {
        T g[3];
        grad1(FieldAt<T>{lrho, N, x, y, z}, r, g);
#pragma unroll
        for (int j = 0; j < 3; j++) acc[j] = (T)mN * g[j];
    }
    if (phi) {
        const T l[3] = {lq[id], lq[n + id], lq[2 * n + id]};
        const T kf = (T)(-0.5 * K * (double)N);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            T h[3];
            grad1(CentralDiff<T>{phi, N, x, y, z, k}, r, h);
#pragma unroll
            for (int j = 0; j < 3; j++) acc[j] += kf * l[k] * h[j];
        }
    }
#pragma unroll
    for (int j = 0; j < 3; j++) lx[j * n + id] += acc[j];
}

// adjoint of the lattice Zel'dovich map (x = x_lat + cx N grad1(psi), q = cq N
// grad1(psi), both at the cell centre r = 1/2 of particle ID = cell):
// lpsi(s) = N/4 sum_j sum_(a,b) [v_j(s - e_j - a e_k - b e_l) - v_j(s - a e_k - b e_l)],
// v = cx lx + cq lq
template <typename T>
__global__ __launch_bounds__(256) void k_adj_ic_lattice(const T *__restrict__ lx, const T *__restrict__ lq, int N,
                                                        int64_t M, double cx, double cq, T *__restrict__ lpsi)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const int x = (int)(s % N), y = (int)((s / N) % N), z = (int)(s / ((int64_t)N * N));
    auto vv = [&](int j, int cx_, int cy_, int cz_) {
        const int64_t c = site(N, cx_, cy_, cz_);
        return (T)cx * lx[j * M + c] + (T)cq * lq[j * M + c];
    };
    T acc = (T)0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            acc += vv(0, x - 1, y - a, z - b) - vv(0, x, y - a, z - b);
            acc += vv(1, x - a, y - 1, z - b) - vv(1, x - a, y, z - b);
            acc += vv(2, x - a, y - b, z - 1) - vv(2, x - a, y - b, z);
        }
    lpsi[s] = (T)(0.25 * (double)N) * acc;
}

static inline unsigned nblocks(int64_t n, int nt) { return (unsigned)((n + nt - 1) / nt); }

template <typename T> int adj_kick_deposit_impl(gev_pcls *p, const T *lx, T *lq, double C, T *u)
{
    gev_ctx *c = p->ctx;
    int rc = field_fill_impl<T>(c, u, 3 * c->M, (T)0);
    if (rc) return rc;
    if (p->n > 0)
        hipLaunchKernelGGL((k_adj_kick_deposit<T>), dim3(nblocks(p->n, PT_NT)), dim3(PT_NT), 0, c->stream, make_pv<T>(p),
                           lx, lq, C, u, c->M);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int adj_divergence_impl(gev_ctx *c, const T *u, T *lphi, double K)
{
    hipLaunchKernelGGL((k_adj_divergence<T>), dim3(nblocks(c->M, 256)), dim3(256), 0, c->stream, u, lphi, c->N, c->M, K);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int adj_gather_impl(gev_pcls *p, const T *phi, const T *lrho, const T *lq, T *lx, double K, double mN)
{
    gev_ctx *c = p->ctx;
    if (p->n > 0)
        hipLaunchKernelGGL((k_adj_gather<T>), dim3(nblocks(p->n, PT_NT)), dim3(PT_NT), 0, c->stream, make_pv<T>(p), phi,
                           lrho, lq, lx, K, mN);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int adj_ic_lattice_impl(gev_ctx *c, const T *lx, const T *lq, double cx, double cq, T *lpsi)
{
    hipLaunchKernelGGL((k_adj_ic_lattice<T>), dim3(nblocks(c->M, 256)), dim3(256), 0, c->stream, lx, lq, c->N, c->M, cx,
                       cq, lpsi);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

#define INST(T)                                                                                   \
    template int adj_kick_deposit_impl<T>(gev_pcls *, const T *, T *, double, T *);               \
    template int adj_divergence_impl<T>(gev_ctx *, const T *, T *, double);                       \
    template int adj_gather_impl<T>(gev_pcls *, const T *, const T *, const T *, T *, double, double); \
    template int adj_ic_lattice_impl<T>(gev_ctx *, const T *, const T *, double, double, T *);
INST(double)
INST(float)
#undef INST

}  // namespace gev

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
//   lq*   = lq' + C lx'          (done by the gather that finalised lx', see below)
//   u_i   = sum_p w_c(x_p) lq*_{p,i}          (the tiled CIC deposit, per-ID weights)
//   lphi  = -K/2 sum_i (u_i(s - e_i) - u_i(s + e_i))                  (k_adj_divergence)
//   lrho  = G lphi                        (G is symmetric: the same Poisson solve)
//   lx    = lx' - K N/2 sum_i lq*_i grad1(D_i phi)(x) + m' N grad1(lrho)(x)   (k_adj_gather)
//   lq    = lq*
// The gather also applies the next (earlier) cycle's lq <- lq + C_prev lx once
// lx is final, so every per-particle update happens exactly once.
// with grad1 the CIC edge gradient (the exact derivative of the CIC
// interpolant inside a cell) and D_i phi(c) = phi(c+e_i) - phi(c-e_i).
// Adjoint per-particle arrays are indexed by particle ID, so they survive the
// re-binning of the container between cycles.
#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_pcommon.hpp"

namespace gev {

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

// lx += -K N/2 sum_i lq_i grad1(D_i phi)(x) + mN grad1(lrho)(x); then
// lq += C_prev lx (the drift adjoint of the cycle processed next).
// sum_i lq_i grad1(D_i phi) = grad1(F) with F(c) = sum_i lq_i (phi(c+e_i) -
// phi(c-e_i)) at the 8 corners, since lq is constant over the particle.
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_adj_gather(PV<T> v, const T *__restrict__ phi, const T *__restrict__ lrho,
                                                      T *lq, T *lx, double K, double mN, double C_prev)
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
    T l[3] = {lq[id], lq[n + id], lq[2 * n + id]};
    if (phi) {
        T Fc[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int cx = x + ((b >> 2) & 1), cy = y + ((b >> 1) & 1), cz = z + (b & 1);
            Fc[b] = l[0] * (phi[site(N, cx + 1, cy, cz)] - phi[site(N, cx - 1, cy, cz)]) +
                    l[1] * (phi[site(N, cx, cy + 1, cz)] - phi[site(N, cx, cy - 1, cz)]) +
                    l[2] * (phi[site(N, cx, cy, cz + 1)] - phi[site(N, cx, cy, cz - 1)]);
        }
        auto F = [&](int dx, int dy, int dz) { return Fc[(dx << 2) | (dy << 1) | dz]; };
        T h[3];
        grad1(F, r, h);
        const T kf = (T)(-0.5 * K * (double)N);
#pragma unroll
        for (int j = 0; j < 3; j++) acc[j] += kf * h[j];
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const T nx = lx[j * n + id] + acc[j];
        lx[j * n + id] = nx;
        if (C_prev != 0.0) lq[j * n + id] = l[j] + (T)C_prev * nx;
    }
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

template <typename T> int adj_divergence_impl(gev_ctx *c, const T *u, T *lphi, double K)
{
    hipLaunchKernelGGL((k_adj_divergence<T>), dim3(nblocks(c->M, 256)), dim3(256), 0, c->stream, u, lphi, c->N, c->M, K);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T>
int adj_gather_impl(gev_pcls *p, const T *phi, const T *lrho, T *lq, T *lx, double K, double mN, double C_prev)
{
    gev_ctx *c = p->ctx;
    if (p->n > 0)
        hipLaunchKernelGGL((k_adj_gather<T>), dim3(nblocks(p->n, PT_NT)), dim3(PT_NT), 0, c->stream, make_pv<T>(p), phi,
                           lrho, lq, lx, K, mN, C_prev);
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
    template int adj_divergence_impl<T>(gev_ctx *, const T *, T *, double);                       \
    template int adj_gather_impl<T>(gev_pcls *, const T *, const T *, T *, T *, double, double, double); \
    template int adj_ic_lattice_impl<T>(gev_ctx *, const T *, const T *, double, double, T *);
INST(double)
INST(float)
#undef INST

}  // namespace gev

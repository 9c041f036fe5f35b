// gev_deposit.hip -- CIC projections of T00, T0i and Tij onto the sites of a
// tile, particle-parallel.
//
// One workgroup owns a tile of TX x TY x TZ sites.  Its sites receive
// contributions from the cells [x0-1, x0+TX-1] x ... (the tile plus one lower
// halo layer), whose particles lie in 2 (TY+1)(TZ+1) contiguous runs of the
// cell-sorted arrays (per halo row: the x0-1 cell, then the TX cells of the
// tile; the split keeps a run contiguous when x0-1 wraps).  The runs are
// flattened and walked one particle per thread, so the loads are coalesced and
// every lane does useful work whatever the occupation of the cells.  Each
// particle evaluates the reference functor and adds the corner / link values
// that land inside the tile to per-site LDS accumulators with LDS atomics; the
// tile then issues one plain store per site and component.  The particles of
// the halo layer are evaluated by both neighbouring tiles, which trades
// (TX+1)(TY+1)(TZ+1)/(TX TY TZ) particle reads for having no global atomics
// and no zeroing pass.  The summation order inside a site follows LDS atomic
// arbitration, as the reference's atomicAdd projections do.
#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_pcommon.hpp"

namespace gev {

// DEP_T00TIJ: T00 and Tij of the same particles and phi in one pass (7 outputs)
// DEP_VEC3: CIC deposit of a per-particle 3-vector indexed by particle ID
// (the adjoint's u_i = sum_p w_c lq_{p,i}), weights read from `aux`
enum { DEP_T00 = 0, DEP_T0I = 1, DEP_TIJ = 2, DEP_T00TIJ = 3, DEP_VEC3 = 4 };
template <int KIND> struct DepC {
    static constexpr int C =
        KIND == DEP_T00 ? 1 : (KIND == DEP_T0I || KIND == DEP_VEC3 ? 3 : (KIND == DEP_TIJ ? 6 : 7));
};

// FP64 LDS accumulators in both precisions: ds_add_f64 is the fast LDS atomic
// on this part, and the per-site sums carry the full FP64 mantissa.
__device__ __forceinline__ void lds_add(double *a, double v)
{
    __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The values one particle adds, as NSLOT (component, corner) slots: slot k
// goes to accumulator comp(k) of the site cell + (bx, by, bz)(k).
template <int KIND> struct DepSlot {
    static constexpr int N = KIND == DEP_T00 ? 8 : KIND == DEP_T0I ? 12 : KIND == DEP_TIJ ? 30 : KIND == DEP_T00TIJ ? 38 : 24;
    // Tij slot j (0..29) -> component and corner code (bx<<2 | by<<1 | bz)
    static constexpr int tij_comp(int j) { return j < 24 ? (j < 8 ? 0 : (j < 16 ? 3 : 5)) : (j < 26 ? 1 : (j < 28 ? 2 : 4)); }
    static constexpr int tij_corner(int j)
    {
        return j < 24 ? (j & 7) : (j == 25 ? 1 : (j == 27 ? 2 : (j == 29 ? 4 : 0)));
    }
    // T0i link k: component and corner (gevolution.hpp:2060-2071 staggering)
    static constexpr int t0i_comp(int k) { return (k == 0 || k == 5 || k == 7 || k == 9) ? 0 : ((k == 1 || k == 3 || k == 8 || k == 10) ? 1 : 2); }
    static constexpr int t0i_corner(int k)
    {
        return k <= 2 ? 0 : (k <= 4 ? 4 : (k <= 6 ? 2 : (k <= 8 ? 1 : (k == 9 ? 3 : (k == 10 ? 5 : 6)))));
    }
    static constexpr int comp(int k)
    {
        return KIND == DEP_T00 ? 0
             : KIND == DEP_T0I ? t0i_comp(k)
             : KIND == DEP_TIJ ? tij_comp(k)
             : KIND == DEP_T00TIJ ? (k < 8 ? 0 : 1 + tij_comp(k - 8))
                                  : k / 8;
    }
    static constexpr int corner(int k)
    {
        return KIND == DEP_T00 ? k
             : KIND == DEP_T0I ? t0i_corner(k)
             : KIND == DEP_TIJ ? tij_corner(k)
             : KIND == DEP_T00TIJ ? (k < 8 ? k : tij_corner(k - 8))
                                  : k % 8;
    }
    static constexpr int bx(int k) { return (corner(k) >> 2) & 1; }
    static constexpr int by(int k) { return (corner(k) >> 1) & 1; }
    static constexpr int bz(int k) { return corner(k) & 1; }
};

// per-particle values of the reference functors; rd: position in the cell,
// q: momentum (DEP_VEC3: the deposited vector), P: phi box at the cell with
// strides (1, SY, SZ)
template <typename T, int KIND>
__device__ __forceinline__ void dep_values(const T rd[3], const T q[3], const T *P, int SY, int SZ, bool use_phi, T a,
                                           T ia2, T *val)
{
    const T o = (T)1;
    auto w = [&](int b) {
        const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
        return (bx ? rd[0] : o - rd[0]) * (by ? rd[1] : o - rd[1]) * (bz ? rd[2] : o - rd[2]);
    };
    auto ph = [&](int b) { return P[((b >> 2) & 1) + SY * ((b >> 1) & 1) + SZ * (b & 1)]; };
    if constexpr (KIND == DEP_VEC3) {
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const T wb = w(b);
            val[b] = wb * q[0]; val[8 + b] = wb * q[1]; val[16 + b] = wb * q[2];
        }
    }
    if constexpr (KIND == DEP_T00 || KIND == DEP_T00TIJ) {
        // projection_T00 functor, gevolution.hpp:1729-1854
        if (use_phi) {
            T f2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
            const T e = fsqrt(f2 + a * a), ie = frcp(e);
            const T f = (T)3 * e + f2 * ie;
            f2 = ((T)0.5 * f * f + f2 * frcp(o + f2 * ia2)) * ie;
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const T p = ph(b);
                val[b] = w(b) * (e + f * p + f2 * p * p);
            }
        } else {
#pragma unroll
            for (int b = 0; b < 8; b++) val[b] = w(b);
        }
    }
    if constexpr (KIND == DEP_T0I) {
        // projection_T0i functor, gevolution.hpp:2031-2201: 12 links per cell,
        // each weighted by (1 + phi_a + phi_b) of its two end points
        const T v0 = q[0], v1 = q[1], v2 = q[2];
        val[0] = (o - rd[1]) * (o - rd[2]) * v0; val[1] = (o - rd[0]) * (o - rd[2]) * v1;
        val[2] = (o - rd[0]) * (o - rd[1]) * v2; val[3] = rd[0] * (o - rd[2]) * v1;
        val[4] = rd[0] * (o - rd[1]) * v2;       val[5] = rd[1] * (o - rd[2]) * v0;
        val[6] = rd[1] * (o - rd[0]) * v2;       val[7] = rd[2] * (o - rd[1]) * v0;
        val[8] = rd[2] * (o - rd[0]) * v1;       val[9] = rd[1] * rd[2] * v0;
        val[10] = rd[0] * rd[2] * v1;            val[11] = rd[0] * rd[1] * v2;
        if (use_phi) {
            T p[8];
#pragma unroll
            for (int b = 0; b < 8; b++) p[b] = ph(b);
            val[0] *= o + p[0] + p[4]; val[1] *= o + p[0] + p[2]; val[2] *= o + p[0] + p[1];
            val[3] *= o + p[4] + p[6]; val[4] *= o + p[4] + p[5]; val[5] *= o + p[2] + p[6];
            val[6] *= o + p[2] + p[3]; val[7] *= o + p[1] + p[5]; val[8] *= o + p[1] + p[3];
            val[9] *= o + p[3] + p[7]; val[10] *= o + p[5] + p[7]; val[11] *= o + p[6] + p[7];
        }
    }
    if constexpr (KIND == DEP_TIJ || KIND == DEP_T00TIJ) {
        // projection_Tij functor, gevolution.hpp:2582-2965 (no CIC_PROJECT_TIJ):
        // diagonal parts CIC onto the 8 corners, off-diagonal parts onto 2 sites
        constexpr int S0 = KIND == DEP_T00TIJ ? 8 : 0;
        T f = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
        const T e2 = f + a * a, ie = frsqrt(e2);
        if (use_phi) {
            f = (T)4 + a * a * frcp(e2);
            f *= (o - rd[0]) * ((o - rd[1]) * ((o - rd[2]) * P[0] + rd[2] * P[SZ]) +
                                rd[1] * ((o - rd[2]) * P[SY] + rd[2] * P[SY + SZ])) +
                 rd[0] * ((o - rd[1]) * ((o - rd[2]) * P[1] + rd[2] * P[1 + SZ]) +
                          rd[1] * ((o - rd[2]) * P[1 + SY] + rd[2] * P[1 + SY + SZ]));
            f += o;
        } else {
            f = o;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const T wd = f * q[d] * q[d] * ie;
#pragma unroll
            for (int b = 0; b < 8; b++) val[S0 + 8 * d + b] = wd * w(b);
        }
        T wo = f * q[0] * q[1] * ie;
        val[S0 + 24] = wo * (o - rd[2]); val[S0 + 25] = wo * rd[2];
        wo = f * q[0] * q[2] * ie;
        val[S0 + 26] = wo * (o - rd[1]); val[S0 + 27] = wo * rd[1];
        wo = f * q[1] * q[2] * ie;
        val[S0 + 28] = wo * (o - rd[0]); val[S0 + 29] = wo * rd[0];
    }
}

template <typename T, int KIND, int TX, int TY, int TZ>
__global__ __launch_bounds__(PT_NT) void k_deposit(PV<T> v, T *out, T *out2, int64_t M, const T *__restrict__ phi,
                                                   double mass_d, double mass2_d, double a_d, int accumulate,
                                                   const T *__restrict__ aux)
{
    constexpr int C = DepC<KIND>::C;
    constexpr int NSLOT = DepSlot<KIND>::N;
    constexpr int NS = TX * TY * TZ;                    // owned sites
    constexpr int HY = TY + 1, NR = HY * (TZ + 1);      // halo rows of cells
    constexpr int NSEG = 2 * NR;                        // runs: halo cell, tile cells
    constexpr int PX = TX + 2, PY = TY + 2, PZ = TZ + 2;  // phi box x0-1 .. x0+TX
    constexpr int NPH = PX * PY * PZ;
    static_assert(NSEG <= PT_NT, "one run per thread in the scan");
    __shared__ double acc[C][NS];
    __shared__ T phs[NPH];
    __shared__ int32_t rbeg[NSEG + 1];
    __shared__ int32_t rst[NSEG];
    __shared__ int32_t wsum[PT_NT / 64];
    const int N = v.N;
    const uint32_t msk = (uint32_t)N - 1;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int x0, y0, z0;
    tile_origin<TX, TY, TZ>(N, x0, y0, z0);
    const bool use_phi = phi != nullptr;

    for (int e = t; e < NS; e += PT_NT) {
#pragma unroll
        for (int c = 0; c < C; c++) acc[c][e] = 0.0;
    }
    if (use_phi) {
        for (int e = t; e < NPH; e += PT_NT) {
            const int px = e % PX, py = (e / PX) % PY, pz = e / (PX * PY);
            const uint32_t g = ((uint32_t)(x0 - 1 + px) & msk) +
                               (uint32_t)N * (((uint32_t)(y0 - 1 + py) & msk) + (uint32_t)N * ((uint32_t)(z0 - 1 + pz) & msk));
            phs[e] = phi[g];
        }
    }
    // runs and their lengths; exclusive scan over NSEG <= PT_NT entries
    int32_t len = 0;
    if (t < NSEG) {
        const int r = t >> 1, ly = r % HY, lz = r / HY;
        const uint32_t row = (uint32_t)N * (((uint32_t)(y0 - 1 + ly) & msk) + (uint32_t)N * ((uint32_t)(z0 - 1 + lz) & msk));
        const uint32_t c0 = (t & 1) ? row + (uint32_t)x0 : row + ((uint32_t)(x0 - 1) & msk);
        const uint32_t c1 = (t & 1) ? c0 + TX : c0 + 1;
        const int32_t s0 = v.start[c0];
        rst[t] = s0;
        len = v.start[c1] - s0;
    }
    int32_t incl = len;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int32_t off = 0;
    for (int j = 0; j < wv; j++) off += wsum[j];
    if (t < NSEG) rbeg[t] = off + incl - len;
    if (t == NSEG - 1) rbeg[NSEG] = off + incl;
    __syncthreads();
    const int32_t total = rbeg[NSEG];
    const T a = (T)a_d, ia2 = (T)(1.0 / (a_d * a_d));

    // U particles per thread and pass: their loads are issued together
    constexpr int U = (KIND == DEP_TIJ || KIND == DEP_T00TIJ) ? 1 : 2;
    const bool need_q = KIND != DEP_T00 || use_phi;
    for (int32_t i0 = t; i0 < total; i0 += U * PT_NT) {
        int run[U];
        int64_t pu[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int32_t i = i0 + u * PT_NT;
            run[u] = -1;
            pu[u] = 0;
            if (i < total) {
                // run holding flattened index i
                int lo = 0, hi = NSEG - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (rbeg[mid] <= i) lo = mid;
                    else hi = mid - 1;
                }
                run[u] = lo;
                pu[u] = (int64_t)rst[lo] + (i - rbeg[lo]);
            }
        }
        T xu[U][3], qu[U][3];
        uint32_t cu[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (run[u] < 0) continue;
            const int64_t p = pu[u];
            cu[u] = (run[u] & 1) ? (uint32_t)v.cell[p] : 0u;
            xu[u][0] = v.p0[p]; xu[u][1] = v.p1[p]; xu[u][2] = v.p2[p];
            if constexpr (KIND == DEP_VEC3) {
                const int64_t id = v.id[p];
                qu[u][0] = aux[id]; qu[u][1] = aux[v.n + id]; qu[u][2] = aux[2 * v.n + id];
            } else {
                if (need_q) { qu[u][0] = v.q0[p]; qu[u][1] = v.q1[p]; qu[u][2] = v.q2[p]; }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool active = run[u] >= 0;
            const int rr = active ? run[u] : 0;
            const int r = rr >> 1, ly = r % HY, lz = r / HY;
            const int lx = (rr & 1) ? 1 + (int)((cu[u] & msk) - (uint32_t)x0) : 0;
            const int gx = (x0 - 1 + lx) & (int)msk, gy = (y0 - 1 + ly) & (int)msk, gz = (z0 - 1 + lz) & (int)msk;
            // site of corner (bx,by,bz) in tile coordinates and whether the tile owns it
            const int sx = lx - 1, sy = ly - 1, sz = lz - 1;
            const bool inx0 = sx >= 0, inx1 = sx + 1 < TX;
            const bool iny0 = sy >= 0, iny1 = sy + 1 < TY;
            const bool inz0 = sz >= 0, inz1 = sz + 1 < TZ;
            const int sb = sx + TX * (sy + TY * sz);
            auto own = [&](int bx, int by, int bz) {
                return (bx ? inx1 : inx0) && (by ? iny1 : iny0) && (bz ? inz1 : inz0);
            };
            auto sidx = [&](int bx, int by, int bz) { return sb + bx + TX * (by + TY * bz); };
            T val[NSLOT];
            if (active) {
                T rd[3];
                cell_rel(xu[u], gx, gy, gz, N, rd);
                dep_values<T, KIND>(rd, qu[u], phs + lx + PX * (ly + PY * lz), PX, PX * PY, use_phi, a, ia2, val);
            } else {
#pragma unroll
                for (int k = 0; k < NSLOT; k++) val[k] = (T)0;
            }
            // lanes of one cell (consecutive in the cell-sorted runs) first sum
            // their values, so a dense cell issues one LDS atomic per slot
            const int key = active ? rr * (TX + 2) + lx : -1;
            const int prevkey = __shfl_up(key, 1, 64);
            const unsigned long long act = __ballot(active);
            const bool same = active && lane > 0 && ((act >> (lane - 1)) & 1ull) && prevkey == key;
            const unsigned long long dup = __ballot(same);
            bool emit = active;
            if (dup) {
                // a lane starts a run unless it repeats its left neighbour; h: its run's first lane
                const unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
                const int h = 63 - __clzll(~dup & le);
                // segmented inclusive scan, as many steps as the longest run needs
                for (int o = 1; o < 64 && __ballot(lane - h >= o); o <<= 1) {
#pragma unroll
                    for (int k = 0; k < NSLOT; k++) {
                        const T y = __shfl_up(val[k], o, 64);
                        if (lane - o >= h) val[k] += y;
                    }
                }
                // the last lane of each run holds the run's sums
                const bool next_same = lane < 63 && ((dup >> (lane + 1)) & 1ull);
                emit = active && !next_same;
            }
            if (emit) {
#pragma unroll
                for (int k = 0; k < NSLOT; k++) {
                    const int c = DepSlot<KIND>::comp(k), bx = DepSlot<KIND>::bx(k), by = DepSlot<KIND>::by(k),
                              bz = DepSlot<KIND>::bz(k);
                    if (own(bx, by, bz)) lds_add(&acc[c][sidx(bx, by, bz)], (double)val[k]);
                }
            }
        }
    }
    __syncthreads();

    for (int s = t; s < NS; s += PT_NT) {
        const int sx = s % TX, sy = (s / TX) % TY, sz = s / (TX * TY);
        const uint32_t gs = (uint32_t)(x0 + sx) + (uint32_t)N * ((uint32_t)(y0 + sy) + (uint32_t)N * (uint32_t)(z0 + sz));
#pragma unroll
        for (int c = 0; c < C; c++) {
            T *oc = (KIND == DEP_T00TIJ && c > 0) ? out2 + (c - 1) * M : out + c * M;
            T val = (T)(acc[c][s] * ((KIND == DEP_T00TIJ && c > 0) ? mass2_d : mass_d));
            if (accumulate) val += oc[gs];
            oc[gs] = val;
        }
    }
}

template <typename T, int KIND, int TX, int TY, int TZ>
static int launch_dep(gev_pcls *p, T *out, T *out2, const T *phi, double mass, double mass2, double a, int acc,
                      const T *aux)
{
    gev_ctx *c = p->ctx;
    const int N = c->N;
    const unsigned nb = (unsigned)((int64_t)(N / TX) * (N / TY) * (N / TZ));
    hipLaunchKernelGGL((k_deposit<T, KIND, TX, TY, TZ>), dim3(nb), dim3(PT_NT), 0, c->stream, make_pv<T>(p), out, out2,
                       c->M, phi, mass, mass2, a, acc, aux);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T, int KIND>
static int dispatch_dep(gev_pcls *p, T *out, const T *phi, double mass, double a, int acc, T *out2 = nullptr,
                        double mass2 = 0.0, const T *aux = nullptr)
{
    const int N = p->ctx->N;
    // Tij keeps 6 (T00+Tij 7) FP64 accumulators per site: the smaller tile holds 4 blocks per CU
    constexpr bool big = KIND != DEP_TIJ && KIND != DEP_T00TIJ;
    if (KIND == DEP_T0I && N >= 32) return launch_dep<T, KIND, 32, 4, 4>(p, out, out2, phi, mass, mass2, a, acc, aux);
    if (big && N >= 16) return launch_dep<T, KIND, 16, 8, 8>(p, out, out2, phi, mass, mass2, a, acc, aux);
    if (KIND == DEP_T00TIJ && N >= 16) return launch_dep<T, KIND, 16, 4, 4>(p, out, out2, phi, mass, mass2, a, acc, aux);
    if (N >= 8) return launch_dep<T, KIND, 8, 8, 8>(p, out, out2, phi, mass, mass2, a, acc, aux);
    return launch_dep<T, KIND, 4, 4, 4>(p, out, out2, phi, mass, mass2, a, acc, aux);
}

// u = CIC deposit of the per-ID 3-vector w ([3][n]), unit mass
template <typename T> int dep_vec3_impl(gev_pcls *p, const T *w, T *u)
{
    return dispatch_dep<T, DEP_VEC3>(p, u, nullptr, 1.0, 1.0, 0, nullptr, 0.0, w);
}

template <typename T> int dep_T00_impl(gev_pcls *p, T *out, const T *phi, double mass, double a, int acc)
{
    return dispatch_dep<T, DEP_T00>(p, out, phi, mass, a, acc);
}

template <typename T> int dep_T0i_impl(gev_pcls *p, T *out, const T *phi, double mass, int acc)
{
    return dispatch_dep<T, DEP_T0I>(p, out, phi, mass, 1.0, acc);
}

template <typename T> int dep_Tij_impl(gev_pcls *p, T *out, const T *phi, double mass, double a, int acc)
{
    return dispatch_dep<T, DEP_TIJ>(p, out, phi, mass, a, acc);
}

template <typename T>
int dep_T00_Tij_impl(gev_pcls *p, T *T00, T *Tij, const T *phi, double mass00, double massij, double a, int acc)
{
    return dispatch_dep<T, DEP_T00TIJ>(p, T00, phi, mass00, a, acc, Tij, massij);
}

#define INST(T)                                                                   \
    template int dep_vec3_impl<T>(gev_pcls *, const T *, T *);                       \
    template int dep_T00_Tij_impl<T>(gev_pcls *, T *, T *, const T *, double, double, double, int); \
    template int dep_T00_impl<T>(gev_pcls *, T *, const T *, double, double, int); \
    template int dep_T0i_impl<T>(gev_pcls *, T *, const T *, double, int);         \
    template int dep_Tij_impl<T>(gev_pcls *, T *, const T *, double, double, int);
INST(double)
INST(float)
#undef INST

}  // namespace gev

// gev_pupdate.hip -- tiled particle updates (kick, drift, and both fused).
//
// One workgroup owns a tile of TX x TY x TZ cells.  Every field value the
// update_q / update_pos stencils can touch (offsets -1..+2 around the tile)
// is staged in LDS once; the tile's particles -- TY*TZ contiguous row
// segments of the cell-sorted container -- are then spread over the threads
// and read their ~60 stencil values from LDS instead of global memory.  The
// fused form also counts the re-binning ranks of the drifted particles, so a
// full cycle touches each particle's data once before the scatter.
#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_pcommon.hpp"

namespace gev {

enum { UPD_KICK_GR = 1, UPD_KICK_NEWTON = 2, UPD_DRIFT_GR = 4, UPD_DRIFT_NEWTON = 8, UPD_COUNT = 16 };

struct UpdParams {
    double dtau_k, a_k, bc_k;  // kick: dtau, a, a^2 N
    double dtau_d, a_d, bc_d;  // drift
    int nf_k, nf_d, flags;
};

// threads per workgroup
#ifndef UPD_NT_F32
#define UPD_NT_F32 512
#endif
#ifndef UPD_NT_F64
#define UPD_NT_F64 256
#endif
template <typename T> struct UpdShape {
    static constexpr int NT = sizeof(T) == 4 ? UPD_NT_F32 : UPD_NT_F64;
};


// position of an idle lane's stand-in particle: the centre of cell (x0, y0, z0)
// (FP64 stores box coordinates, FP32 offsets in the cell)
template <typename T> __device__ __forceinline__ void idle_pos(int x0, int y0, int z0, int N, T x[3])
{
    if constexpr (sizeof(T) == 8) {
        x[0] = (x0 + 0.5) / N; x[1] = (y0 + 0.5) / N; x[2] = (z0 + 0.5) / N;
    } else {
        x[0] = x[1] = x[2] = (T)0.5;
    }
}

// One particle of a kick / drift pass: the stencil values come from the
// staged box fields[f * FS + ...] of a tile with origin (x0, y0, z0) and box
// pitch SX, SY.  Writes q, position and cell; returns the new cell.
template <typename T, int F, int SX, int SY>
__device__ __forceinline__ int32_t upd_particle(const PV<T> &v, const UpdParams &u, int64_t p, int32_t cc, T xk[3],
                                                T q[3], const T *fields, int x0, int y0, int z0, int FS, double &ret)
{
    constexpr bool kick = F & (UPD_KICK_GR | UPD_KICK_NEWTON);
    constexpr bool drift = F & (UPD_DRIFT_GR | UPD_DRIFT_NEWTON);
    const int N = v.N;
    const int nf_k = u.nf_k;
    const int nf_d = (F & UPD_DRIFT_GR) ? u.nf_d : 0;
    const T Nn = (T)N;
    int32_t newc = 0;
    const int x = cc & (N - 1), y = (cc >> v.lg) & (N - 1), z = cc >> (2 * v.lg);
    const int o = (x - x0 + 1) + SX * ((y - y0 + 1) + SY * (z - z0 + 1));
    T r[3];
    cell_rel(xk, x, y, z, N, r);
    FieldLds<T> Fp{fields, o, SX, SX * SY}, Fc{fields + FS, o, SX, SX * SY};
    FieldLds<T> B0{fields + 2 * FS, o, SX, SX * SY}, B1{fields + 3 * FS, o, SX, SX * SY},
        B2{fields + 4 * FS, o, SX, SX * SY};
    // FP32: kick and drift share weights and stencil reads; FP64 keeps
    // them apart (the shared form needs more registers than it saves)
    if constexpr ((F & (UPD_KICK_GR | UPD_DRIFT_GR)) == (UPD_KICK_GR | UPD_DRIFT_GR) && sizeof(T) == 4) {
        double d[3];
        T rv = phys_kick_drift<T>(r, q, Fp, Fc, B0, B1, B2, nf_k, (T)u.a_k, (T)u.dtau_k, (T)u.bc_k, Nn, nf_d, (T)u.a_d,
                                  (T)u.bc_d, u.dtau_d, d);
        ret = fmax(ret, (double)rv);
        v.q0[p] = q[0]; v.q1[p] = q[1]; v.q2[p] = q[2];
        newc = move_by(v, p, xk, x, y, z, d);
    } else {
        if constexpr (F & UPD_KICK_GR) {
            T rv = phys_update_q<T>(r, q, Fp, Fc, B0, B1, B2, nf_k, (T)u.a_k, (T)u.dtau_k, (T)u.bc_k, Nn);
            ret = fmax(ret, (double)rv);
        } else if constexpr (F & UPD_KICK_NEWTON) {
            T rv = phys_update_q_newton<T>(r, q, Fp, Fc, nf_k, (T)u.a_k, (T)u.dtau_k, Nn);
            ret = fmax(ret, (double)rv);
        }
        if (kick) { v.q0[p] = q[0]; v.q1[p] = q[1]; v.q2[p] = q[2]; }
        if (drift) {
            double d[3];
            if constexpr (F & UPD_DRIFT_GR) {
                phys_update_pos<T>(r, q, Fp, Fc, B0, B1, B2, nf_d, (T)u.a_d, (T)u.bc_d, u.dtau_d, d);
            } else {
                d[0] = u.dtau_d * (double)q[0] / u.a_d;
                d[1] = u.dtau_d * (double)q[1] / u.a_d;
                d[2] = u.dtau_d * (double)q[2] / u.a_d;
            }
            newc = move_by(v, p, xk, x, y, z, d);
        }
    }
    return newc;
}

template <typename T, int TX, int TY, int TZ, int F, int NT = UpdShape<T>::NT>
__global__ __launch_bounds__(NT) void k_update_tiled(PV<T> v, UpdParams u, const T *__restrict__ phi,
                                                        const T *__restrict__ chi, const T *__restrict__ B0g,
                                                        const T *__restrict__ B1g, const T *__restrict__ B2g,
                                                        unsigned long long *maxslot, int32_t *count, int32_t *rank,
                                                        uint32_t *__restrict__ qb)
{
    constexpr int SX = TX + 3, SY = TY + 3, SZ = TZ + 3, NS = SX * SY * SZ, NROW = TY * TZ;
    static_assert(NROW <= 64 && SX <= 64 && (NROW & (NROW - 1)) == 0, "tile shape unsupported by the wave-level staging");
    __shared__ T fld[5][NS];
    const int N = v.N;
    const uint32_t msk = (uint32_t)N - 1;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int x0, y0, z0;
    tile_origin<TX, TY, TZ>(N, x0, y0, z0);
    const int32_t cell0 = (int32_t)((uint32_t)x0 + (uint32_t)N * ((uint32_t)y0 + (uint32_t)N * (uint32_t)z0));
    T idle_x[3];
    idle_pos<T>(x0, y0, z0, N, idle_x);
    // F: the kick / drift forms (compile time); UPD_COUNT in u.flags (run time)
    constexpr bool kick = F & (UPD_KICK_GR | UPD_KICK_NEWTON);
    constexpr bool drift = F & (UPD_DRIFT_GR | UPD_DRIFT_NEWTON);
    const int nf_k = u.nf_k;
    const int nf_d = (F & UPD_DRIFT_GR) ? u.nf_d : 0;
    const int nfmax = kick ? (nf_k > nf_d ? nf_k : nf_d) : nf_d;
    const bool use_phi = kick || nf_d >= 1, use_chi = nfmax >= 2, use_B = nfmax >= 3;
    // stage [x0-1, x0+TX+1] x ... : each wave covers RPW box rows per pass, lane -> (row, x)
    {
        constexpr int RPW = 64 / SX, NR = SY * SZ, NW = NT / 64;
        const int sub = lane / SX, lx = lane - sub * SX;
        const uint32_t gx = (uint32_t)(x0 - 1 + lx) & msk;
        if (sub < RPW) {
            for (int row = wv * RPW + sub; row < NR; row += NW * RPW) {
                const int ly = row % SY, lz = row / SY;
                const uint32_t g = gx + (uint32_t)N * (((uint32_t)(y0 - 1 + ly) & msk) +
                                                       (uint32_t)N * ((uint32_t)(z0 - 1 + lz) & msk));
                const int e = lx + SX * row;
                if (use_phi) fld[0][e] = phi[g];
                if (use_chi) fld[1][e] = chi[g];
                if (use_B) {
                    fld[2][e] = B0g[g];
                    fld[3][e] = B1g[g];
                    fld[4][e] = B2g[g];
                }
            }
        }
    }
    // qb (FP32 form, N >= 16): the largest |q|^2 of the updated particles per
    // block of 8^3 cells (gev_pcls::qbound), for the fixed-point deposits.
    // The blocks of cells [x0-1, x0+TX] x ... (where every particle lands
    // that moved at most one cell) are gathered in LDS with integer max on
    // the float bits (monotone for values >= 0) and merged into qb at the
    // end; a particle that landed elsewhere goes to qb directly.
    constexpr int QWX = TX / 8 + 2, QWY = TY / 8 + 2, QWZ = TZ / 8 + 2, QWN = QWX * QWY * QWZ;
    __shared__ uint32_t qwin[QWN];
    const bool qbon = sizeof(T) == 4 && qb != nullptr;
    const int qx0 = (x0 - 1 + N) >> QB_LG, qy0 = (y0 - 1 + N) >> QB_LG, qz0 = (z0 - 1 + N) >> QB_LG;
    const int nbq = N >> QB_LG;
    if (qbon)
        for (int e = tid; e < QWN; e += NT) qwin[e] = 0u;
    // particle ranges of the tile's TY*TZ rows, computed by every wave for
    // itself (no barrier before the first particle loads)
    int32_t rb0, incl;
    tile_rows_load<TX, TY, NROW>(v.start, N, x0, y0, z0, lane, rb0, incl);
    incl = wave_row_scan(incl, lane);
    const int total = __shfl(incl, NROW - 1, 64);
    const int niter = (total + NT - 1) / NT;  // the same for every lane of the workgroup
    auto locate = [&](int f) { return tile_locate<NROW>(rb0, incl, f); };
    // the next particle of this lane, loaded one pass ahead of its use
    int64_t pn = locate(tid);
    int32_t cn = 0;
    T xn[3] = {(T)0, (T)0, (T)0}, qn[3] = {(T)0, (T)0, (T)0};
    // idle lanes load the sink slot: the loads, like the stores below, run
    // without a branch, which keeps the compiler's memory-counter waits exact
    auto fetch = [&]() {
        const int64_t lp = pn >= 0 ? pn : v.sink;
        cn = v.cell[lp];
        xn[0] = v.p0[lp]; xn[1] = v.p1[lp]; xn[2] = v.p2[lp];
        qn[0] = v.q0[lp]; qn[1] = v.q1[lp]; qn[2] = v.q2[lp];
    };
    fetch();
    __syncthreads();  // staged fields
    double ret = 0.0;
    RankPending rpend{0, 0};
    int64_t ppend = -1;
    for (int it = 0; it < niter; it++) {
        const int64_t p = pn;
        const bool active = p >= 0;
        // idle lanes update a particle at rest in the tile's first cell and
        // store it to the sink slot
        const int32_t cc = active ? cn : cell0;
        T xk[3], q[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            xk[d] = active ? xn[d] : idle_x[d];
            q[d] = active ? qn[d] : (T)0;
        }
        pn = locate(tid + (it + 1) * NT);
        fetch();
        // the stand-in particle's return value is dropped, so maxvel is the
        // maximum over real particles only (a select, no branch)
        double rit = 0.0;
        const int32_t newc =
            upd_particle<T, F, SX, SY>(v, u, active ? p : v.sink, cc, xk, q, fld[0], x0, y0, z0, NS, rit);
        ret = active ? fmax(ret, rit) : ret;
        if (qbon && active) {
            const int32_t cq = drift ? newc : cc;
            const uint32_t qq = __float_as_uint((float)(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]));
            int cx, cy, cz;
            cell_coords(cq, N, v.lg, cx, cy, cz);
            const uint32_t rx = (uint32_t)(cx - x0 + 1) & msk, ry = (uint32_t)(cy - y0 + 1) & msk,
                           rz = (uint32_t)(cz - z0 + 1) & msk;
            if (rx <= (uint32_t)TX + 1 && ry <= (uint32_t)TY + 1 && rz <= (uint32_t)TZ + 1) {
                const int wx = ((x0 - 1 + (int)rx + N) >> QB_LG) - qx0, wy = ((y0 - 1 + (int)ry + N) >> QB_LG) - qy0,
                          wz = ((z0 - 1 + (int)rz + N) >> QB_LG) - qz0;
                // values only grow: the atomic is needed only above what the
                // window holds (most lanes of a wave share one block)
                uint32_t *w = &qwin[wx + QWX * (wy + QWY * wz)];
                if (qq > __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) atomicMax(w, qq);
            } else {
                atomicMax(&qb[(cx >> QB_LG) + nbq * ((cy >> QB_LG) + nbq * (cz >> QB_LG))], qq);
            }
        }
        // re-binning rank in the new cell, one counter atomic per run of equal
        // cells; each pass stores the rank of the previous one, so the
        // atomic's round trip overlaps a pass of work
        if (drift && (u.flags & UPD_COUNT)) {
            if (it > 0) {
                const int32_t rk = wave_rank_finish(rpend);
                rank[ppend >= 0 ? ppend : v.sink] = rk;
            }
            rpend = wave_rank_issue(count, newc, active);
            ppend = p;
        }
    }
    if (drift && (u.flags & UPD_COUNT) && niter > 0) {
        const int32_t rk = wave_rank_finish(rpend);
        rank[ppend >= 0 ? ppend : v.sink] = rk;
    }
    if (kick) block_max_to<NT>(ret, maxslot);
    if (qbon) {
        __syncthreads();  // every lane's window updates
        for (int e = tid; e < QWN; e += NT) {
            const uint32_t val = qwin[e];
            const int wx = e % QWX, wy = (e / QWX) % QWY, wz = e / (QWX * QWY);
            if (val)
                atomicMax(&qb[((qx0 + wx) & (nbq - 1)) + nbq * (((qy0 + wy) & (nbq - 1)) + nbq * ((qz0 + wz) & (nbq - 1)))],
                          val);
        }
    }
}


template <typename T> int pcls_sort_impl(gev_pcls *p);
template <typename T> int pcls_sort_counted_impl(gev_pcls *p);

template <typename T, int TX, int TY, int TZ, int F>
static void launch_tile(gev_pcls *p, const UpdParams &u, const T *phi, const T *chi, const T *B0, const T *B1,
                        const T *B2, unsigned long long *slot)
{
    gev_ctx *c = p->ctx;
    const int N = c->N;
    const unsigned nb = (unsigned)((int64_t)(N / TX) * (N / TY) * (N / TZ));
    if constexpr (!(sizeof(T) == 8 && TX == 16))  // the large tile exists for FP32 only
        hipLaunchKernelGGL((k_update_tiled<T, TX, TY, TZ, F>), dim3(nb), dim3(UpdShape<T>::NT), 0, c->stream,
                           make_pv<T>(p), u, phi, chi, B0, B1, B2, slot, p->cell_count, p->rank,
                           sizeof(T) == 4 ? p->qbound : nullptr);
}

template <typename T, int TX, int TY, int TZ>
static void launch_mode(gev_pcls *p, const UpdParams &u, const T *phi, const T *chi, const T *B0, const T *B1,
                        const T *B2, unsigned long long *slot)
{
    switch (u.flags & ~UPD_COUNT) {
    case UPD_KICK_GR: launch_tile<T, TX, TY, TZ, UPD_KICK_GR>(p, u, phi, chi, B0, B1, B2, slot); break;
    case UPD_KICK_NEWTON: launch_tile<T, TX, TY, TZ, UPD_KICK_NEWTON>(p, u, phi, chi, B0, B1, B2, slot); break;
    case UPD_DRIFT_GR: launch_tile<T, TX, TY, TZ, UPD_DRIFT_GR>(p, u, phi, chi, B0, B1, B2, slot); break;
    case UPD_DRIFT_NEWTON: launch_tile<T, TX, TY, TZ, UPD_DRIFT_NEWTON>(p, u, phi, chi, B0, B1, B2, slot); break;
    case UPD_KICK_GR | UPD_DRIFT_GR:
        launch_tile<T, TX, TY, TZ, UPD_KICK_GR | UPD_DRIFT_GR>(p, u, phi, chi, B0, B1, B2, slot);
        break;
    case UPD_KICK_NEWTON | UPD_DRIFT_NEWTON:
        launch_tile<T, TX, TY, TZ, UPD_KICK_NEWTON | UPD_DRIFT_NEWTON>(p, u, phi, chi, B0, B1, B2, slot);
        break;
    }
}

template <typename T> static int launch_update(gev_pcls *p, const UpdParams &u, const T *phi, const T *chi, const T *Bi)
{
    gev_ctx *c = p->ctx;
    const int N = c->N;
    if (u.flags & (UPD_KICK_GR | UPD_KICK_NEWTON)) {
        int rc = field_fill_impl<double>(c, c->red + 8, 1, 0.0);
        if (rc) return rc;
    }
    if (u.flags & UPD_COUNT) {
        int rc = pcls_clear_counts(p);
        if (rc) return rc;
    }
    // the |q|^2 block bounds are rebuilt by every FP32 update (N >= 16)
    p->qbound_valid = false;
    if (sizeof(T) == 4 && p->qbound) {
        const int64_t nb = (int64_t)(N >> QB_LG);
        int rc = field_fill_impl<int32_t>(c, (int32_t *)p->qbound, nb * nb * nb, 0);
        if (rc) return rc;
    }
    unsigned long long *slot = (unsigned long long *)(c->red + 8);
    const T *B0 = Bi, *B1 = Bi ? Bi + c->M : nullptr, *B2 = Bi ? Bi + 2 * c->M : nullptr;
    // FP32 fits the 16x8x8 tile (5 staged fields, 46 KB of LDS); FP64 keeps 8x8x4
    bool big = false;
    if constexpr (sizeof(T) == 4) big = N >= 16;
    if (big) {
        launch_mode<T, 16, 8, 8>(p, u, phi, chi, B0, B1, B2, slot);
    } else if (N >= 8) {
        launch_mode<T, 8, 8, 4>(p, u, phi, chi, B0, B1, B2, slot);
    } else {
        launch_mode<T, 4, 4, 4>(p, u, phi, chi, B0, B1, B2, slot);
    }
    GEV_HIP(hipGetLastError());
    p->qbound_valid = sizeof(T) == 4 && p->qbound != nullptr;
    return GEV_OK;
}

int read_max_slot(gev_ctx *c, double *out)
{
    GEV_HIP(hipMemcpyAsync(c->red_host, c->red + 8, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GEV_HIP(hipStreamSynchronize(c->stream));
    *out = c->red_host[0];
    return GEV_OK;
}

template <typename T>
int update_q_impl(gev_pcls *p, double dtau, const T *phi, const T *chi, const T *Bi, int nfield, double a,
                  double bcoeff, double *max_ret)
{
    UpdParams u = {dtau, a, bcoeff, 0, 1, 1, nfield, 0, UPD_KICK_GR};
    int rc = launch_update<T>(p, u, phi, chi, Bi);
    if (rc) return rc;
    return max_ret ? read_max_slot(p->ctx, max_ret) : GEV_OK;
}

template <typename T>
int update_q_newton_impl(gev_pcls *p, double dtau, const T *psi, const T *chi, int nfield, double a, double *max_ret)
{
    UpdParams u = {dtau, a, 1, 0, 1, 1, nfield, 0, UPD_KICK_NEWTON};
    int rc = launch_update<T>(p, u, psi, chi, nullptr);
    if (rc) return rc;
    return max_ret ? read_max_slot(p->ctx, max_ret) : GEV_OK;
}

template <typename T>
int update_pos_impl(gev_pcls *p, double dtau, const T *phi, const T *chi, const T *Bi, int nfield, double a,
                    double bcoeff)
{
    UpdParams u = {0, 1, 1, dtau, a, bcoeff, 0, nfield, UPD_DRIFT_GR | UPD_COUNT};
    int rc = launch_update<T>(p, u, phi, chi, Bi);
    if (rc) return rc;
    return pcls_sort_counted_impl<T>(p);
}

template <typename T> int update_pos_newton_impl(gev_pcls *p, double dtau, double a)
{
    UpdParams u = {0, 1, 1, dtau, a, 1, 0, 0, UPD_DRIFT_NEWTON | UPD_COUNT};
    int rc = launch_update<T>(p, u, nullptr, nullptr, nullptr);
    if (rc) return rc;
    return pcls_sort_counted_impl<T>(p);
}

template <typename T>
int kick_drift_impl(gev_pcls *p, int gr, double dtau_k, double dtau_d, const T *phi, const T *chi, const T *Bi,
                    int nf_k, int nf_d, double a_k, double a_d, double bc_k, double bc_d, double *max_ret, int rebin)
{
    UpdParams u = {dtau_k, a_k, bc_k, dtau_d, a_d, bc_d, nf_k, nf_d,
                   (gr ? (UPD_KICK_GR | UPD_DRIFT_GR) : (UPD_KICK_NEWTON | UPD_DRIFT_NEWTON)) | UPD_COUNT};
    int rc = launch_update<T>(p, u, phi, chi, Bi);
    if (rc) return rc;
    if (max_ret) {
        rc = read_max_slot(p->ctx, max_ret);
        if (rc) return rc;
    }
    if (!rebin) {
        p->sorted = false;  // counts and ranks are ready; the scatter is pending
        return GEV_OK;
    }
    return pcls_sort_counted_impl<T>(p);
}

#define INST(T)                                                                                                   \
    template int update_q_impl<T>(gev_pcls *, double, const T *, const T *, const T *, int, double, double,       \
                                  double *);                                                                      \
    template int update_q_newton_impl<T>(gev_pcls *, double, const T *, const T *, int, double, double *);       \
    template int update_pos_impl<T>(gev_pcls *, double, const T *, const T *, const T *, int, double, double);    \
    template int update_pos_newton_impl<T>(gev_pcls *, double, double);                                          \
    template int kick_drift_impl<T>(gev_pcls *, int, double, double, const T *, const T *, const T *, int, int,   \
                                    double, double, double, double, double *, int);
INST(double)
INST(float)
#undef INST

}  // namespace gev

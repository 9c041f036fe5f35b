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
// and no zeroing pass.  FP64 tiles sum in FP64 with LDS atomics, whose order
// inside a site follows atomic arbitration, as the reference's atomicAdd
// projections do.  FP32 tiles sum in 32-bit fixed point (integer LDS
// atomics, which this part executes several times faster than FP32 or FP64
// ones, and whose sums do not depend on the order): each tile bounds its
// per-site sums by (largest particle total of the 4 halo rows around one of
// its sites) x (largest per-particle value, from |q| and |phi|, separately for
// the T00, Tij and T0i components, whose sizes differ by the ratio of
// kinetic to rest energy) and scales so that bound is 2^30.  |q| is bounded
// by the largest |q|^2 per 8^3-cell block that the last particle update
// recorded (gev_pcls::qbound), or else by a pass over the tile's particles.
#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_pcommon.hpp"

namespace gev {

// DEP_T00TIJ: T00 and Tij of the same particles and phi in one pass (7 outputs)
// DEP_T00TIJL: DEP_T00TIJ plus the particle part of T0i on each link (10 outputs):
// projection_T0i weights the 12 link values of a particle by 1 + phi_a + phi_b
// of the link's end points, a factor common to every particle on the link, so
// the links can be summed here (with the old phi's T00 and Tij pass) and
// multiplied by the new phi's factor later, inside the B solve
// DEP_VEC3: CIC deposit of a per-particle 3-vector indexed by particle ID
// (the adjoint's u_i = sum_p w_c lq_{p,i}), weights read from `aux`
enum { DEP_T00 = 0, DEP_T0I = 1, DEP_TIJ = 2, DEP_T00TIJ = 3, DEP_VEC3 = 4, DEP_T00TIJL = 5 };
template <int KIND> struct DepC {
    static constexpr int C = KIND == DEP_T00 ? 1
                             : (KIND == DEP_T0I || KIND == DEP_VEC3 ? 3
                                                                    : (KIND == DEP_TIJ ? 6 : (KIND == DEP_T00TIJ ? 7 : 10)));
};

// LDS accumulation: FP64 sums (ds_add_f64) in the FP64 form; in the FP32 form
// 32-bit fixed point (ds_add_u32) of values already multiplied by the tile's
// scale, rounded to nearest.
__device__ __forceinline__ void lds_add(double *a, double v)
{
    __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(int32_t *a, float v)
{
    // floor(v + 1/2) in one instruction (v_cvt_rpi_i32_f32; the compiler emits
    // two for round-to-nearest): the deposit is bound by its VALU issue rate
    int32_t r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    __hip_atomic_fetch_add(a, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// accumulator of a kind and precision: fixed point for the FP32 form of the
// physical projections (their values are bounded by the particles' energies);
// the adjoint's vector deposit keeps FP64
template <typename T, int KIND> struct DepAcc { using type = double; };
template <int KIND> struct DepAcc<float, KIND> { using type = int32_t; };
template <> struct DepAcc<float, 4> { using type = double; };  // DEP_VEC3

#ifndef DEP_CHUNK_X4
#define DEP_CHUNK_X4 5  // FP32 single-particle kinds: as DEP_SPREAD_X4, per-lane chunks
#endif
#ifndef DEP_SPREAD_X4
#define DEP_SPREAD_X4 5  // four times the mean particles per halo-box cell above which a tile spreads
#endif

template <typename T, int KIND, int TX, int TY, int TZ>
__global__ __launch_bounds__(PT_NT) void k_deposit(PV<T> v, T *out, T *out2, T *out3, int64_t M,
                                                   const T *__restrict__ phi, double mass_d, double mass2_d,
                                                   double mass3_d, double a_d, double ia2_d, int accumulate,
                                                   const T *__restrict__ aux, const uint32_t *__restrict__ qb, int vec4)
{
    constexpr int C = DepC<KIND>::C;
    constexpr int NS = TX * TY * TZ;                    // owned sites
    constexpr int HY = TY + 1, NR = HY * (TZ + 1);      // halo rows of cells
    constexpr int NSEG = 2 * NR;                        // runs: halo cell, tile cells
    constexpr int PX = TX + 2, PY = TY + 2, PZ = TZ + 2;  // phi box x0-1 .. x0+TX
    constexpr int NPH = PX * PY * PZ;
    static_assert(NSEG <= PT_NT, "one run per thread in the scan");
    using AccT = typename DepAcc<T, KIND>::type;
    constexpr bool FIX = sizeof(AccT) == 4;
    constexpr int NW = PT_NT / 64;
    static_assert(!FIX || (TY * TZ <= 64 && NS % 4 == 0 && TX % 4 == 0), "fixed-point tile shape");
    __shared__ __align__(16) AccT acc[C][NS];
    __shared__ uint32_t wpm[NW];  // per wave: the largest |phi| of its share of the phi box (float bits)
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
    const bool need_q_ = KIND != DEP_T00 || use_phi;
    // the particle updates' largest |q|^2 over the blocks of 8^3 cells that
    // the halo box touches (per dimension the block of the halo layer and
    // those of the tile), loaded by every wave for itself first, so the loads
    // overlap the staging below (float bits: non-negative values order as
    // their bit patterns)
    uint32_t qbits = 0;
    if (FIX && qb && need_q_) {
        const int nbq = N >> QB_LG;
        const int bx0 = x0 >> QB_LG, by0 = y0 >> QB_LG, bz0 = z0 >> QB_LG;
        const int nx = ((x0 + TX - 1) >> QB_LG) - bx0 + 2, ny = ((y0 + TY - 1) >> QB_LG) - by0 + 2,
                  nz = ((z0 + TZ - 1) >> QB_LG) - bz0 + 2;
        for (int e = lane; e < nx * ny * nz; e += 64) {
            const int ix = e % nx, iy = (e / nx) % ny, iz = e / (nx * ny);
            const int bx = ix == 0 ? (((x0 - 1) & (int)msk) >> QB_LG) : bx0 + ix - 1;
            const int by = iy == 0 ? (((y0 - 1) & (int)msk) >> QB_LG) : by0 + iy - 1;
            const int bz = iz == 0 ? (((z0 - 1) & (int)msk) >> QB_LG) : bz0 + iz - 1;
            qbits = max(qbits, qb[bx + nbq * (by + nbq * bz)]);
        }
    }

    if constexpr (FIX) {
        int4 *a4 = reinterpret_cast<int4 *>(&acc[0][0]);
        for (int e = t; e < C * NS / 4; e += PT_NT) a4[e] = make_int4(0, 0, 0, 0);
    } else {
        for (int e = t; e < NS; e += PT_NT) {
#pragma unroll
            for (int c = 0; c < C; c++) acc[c][e] = (AccT)0;
        }
    }
    uint32_t pbits = 0;
    if (use_phi) {
        // all of a thread's loads issued before the first LDS write
        constexpr int NPI = (NPH + PT_NT - 1) / PT_NT;
        T ph[NPI];
#pragma unroll
        for (int k = 0; k < NPI; k++) {
            const int e = t + k * PT_NT;
            const int px = e % PX, py = (e / PX) % PY, pz = e / (PX * PY);
            const uint32_t g = ((uint32_t)(x0 - 1 + px) & msk) +
                               (uint32_t)N * (((uint32_t)(y0 - 1 + py) & msk) + (uint32_t)N * ((uint32_t)(z0 - 1 + pz) & msk));
            ph[k] = e < NPH ? phi[g] : (T)0;
        }
#pragma unroll
        for (int k = 0; k < NPI; k++) {
            const int e = t + k * PT_NT;
            if (e < NPH) phs[e] = ph[k];
            if constexpr (FIX) pbits = max(pbits, __float_as_uint(fabsf((float)ph[k])));
        }
    }
    if constexpr (FIX) {
        // the wave's maxima (all lanes are active here); the scan's barriers
        // below publish wpm
        pbits = wave_max_u32(pbits);
        if (FIX && qb && need_q_) qbits = wave_max_u32(qbits);
        if (lane == 0) wpm[wv] = pbits;
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
    const T a = (T)a_d, o = (T)1, ia2 = (T)ia2_d;  // ia2_d = 1 / a^2
    // flattened particle index -> container slot
    auto locate = [&](int32_t i, int &run) -> int64_t {
        int lo = 0, hi = NSEG - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (rbeg[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        run = lo;
        return (int64_t)rst[lo] + (i - rbeg[lo]);
    };
    // fixed-point scale of the FP32 form (see the top of the file): S maps the
    // bound of every per-site sum of the tile to 2^30
    // S0, S1, S2: scales of the T00, Tij and T0i accumulators (1 in FP64)
    float S0 = 1.0f, S1 = 1.0f, S2 = 1.0f;
    if constexpr (FIX) {
        // a site takes the particles of two cells in each of four halo rows,
        // so the four rows' totals bound its particle count (run lengths of the
        // scan; looser than the count of its eight cells by at most the row
        // length, which costs a few of the 30 bits)
        uint32_t cb = 0;
        if (lane < TY * TZ) {
            const int sy = lane % TY, sz = lane / TY;
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int r = (sy + (d & 1)) + HY * (sz + (d >> 1));
                cb += (uint32_t)(rbeg[2 * r + 2] - rbeg[2 * r]);
            }
        }
        const float cmax = (float)wave_max_u32(cb);
        float pmax = 0.0f;
#pragma unroll
        for (int j = 0; j < NW; j++) pmax = fmaxf(pmax, __uint_as_float(wpm[j]));
        float emax = 0.0f, tmax = 0.0f, vmax = 0.0f;
        const float a2 = (float)(a_d * a_d);
        if (need_q_ && qb) {
            // e, q^2/e and |q_i| all increase with |q|^2
            const float qq = __uint_as_float(qbits);
            emax = sqrtf(qq + a2);
            tmax = qq / emax;
            vmax = sqrtf(qq);
        } else if (need_q_) {
            // no valid block bounds: a pass over the tile's particles
            __shared__ uint32_t wq[3][NW];
            uint32_t eb = 0, tb = 0, vb = 0;
            for (int32_t i = t; i < total; i += PT_NT) {
                int run;
                const int64_t p = locate(i, run);
                const float q0 = (float)v.q0[p], q1 = (float)v.q1[p], q2 = (float)v.q2[p];
                const float qq = q0 * q0 + q1 * q1 + q2 * q2, ee = sqrtf(qq + a2);
                eb = max(eb, __float_as_uint(ee));
                tb = max(tb, __float_as_uint(qq / ee));
                vb = max(vb, __float_as_uint(fmaxf(fabsf(q0), fmaxf(fabsf(q1), fabsf(q2)))));
            }
            eb = wave_max_u32(eb);
            tb = wave_max_u32(tb);
            vb = wave_max_u32(vb);
            if (lane == 0) { wq[0][wv] = eb; wq[1][wv] = tb; wq[2][wv] = vb; }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < NW; j++) {
                emax = fmaxf(emax, __uint_as_float(wq[0][j]));
                tmax = fmaxf(tmax, __uint_as_float(wq[1][j]));
                vmax = fmaxf(vmax, __uint_as_float(wq[2][j]));
            }
        }
        // largest |value| one particle adds to one site, per accumulator group:
        // T00's e + f phi + f2 phi^2 <= e (1 + 4|phi| + 9 phi^2) (f <= 4e, f2 <= 9e;
        // 1 for the mass deposit), Tij's f q_i q_j / e <= (1 + 5|phi|) q^2 / e,
        // T0i's (1 + phi_a + phi_b) q_i <= (1 + 2|phi|) max|q_i|; each rounded up
        // by 1 % against the FP32 rounding of the values.  The tile's per-site
        // sums stay below cmax times these.
        // (in FP32 with the hardware reciprocal: any scale at or below the
        // bound is exact to use, since the stores divide by the same value)
        if (cmax > 0.0f) {
            const float pm = pmax, top = 1073741824.0f;
            const float B0 = need_q_ ? emax * (1.0f + 4.0f * pm + 9.0f * pm * pm) * 1.01f : 1.0f;
            const float B1 = tmax * (1.0f + 5.0f * pm) * 1.01f, B2 = vmax * (1.0f + 2.0f * pm) * 1.01f;
            S0 = top * __builtin_amdgcn_rcpf(cmax * B0);
            if (B1 > 0.0f) S1 = top * __builtin_amdgcn_rcpf(cmax * B1);
            if (B2 > 0.0f) S2 = top * __builtin_amdgcn_rcpf(cmax * B2);
        }
    }

    // U particles per thread and pass: their loads are issued together
    constexpr bool T00TIJ = KIND == DEP_T00TIJ || KIND == DEP_T00TIJL;
    constexpr int U = (KIND == DEP_TIJ || T00TIJ) ? 1 : 2;
    constexpr int TB = T00TIJ ? 1 : 0;  // first Tij accumulator
    const bool need_q = KIND != DEP_T00 || use_phi;
    // A dense tile (on average DEP_SPREAD_X4 / 4 or more particles per cell of
    // its halo box) hands consecutive particles to different waves, so the members
    // of one crowded cell contend in PT_NT/64 LDS atomic instructions at a
    // time instead of all in one; sparse tiles keep the lane-contiguous order,
    // whose loads coalesce fully.
    constexpr int NCELL = (TX + 1) * (TY + 1) * (TZ + 1);
    const bool spread = DEP_SPREAD_X4 > 0 && 4 * total >= DEP_SPREAD_X4 * NCELL;
    const int slot = spread ? lane * NW + wv : t;
    // In the FP32 form a crowded tile of a one-particle-per-pass kind instead
    // gives every lane its own contiguous chunk of Q particles, so the lanes of
    // one instruction sit Q particles apart, in different cells (late-time
    // T00+Tij+T0i links 10.7 -> 9.7 ms at 512^3; FP64 and two-particle kinds
    // measured slower with it)
    const bool chunk = sizeof(T) == 4 && U == 1 && DEP_CHUNK_X4 > 0 && 4 * total >= DEP_CHUNK_X4 * NCELL;
    const int32_t Q = (total + 63) / 64;
    for (int32_t i0 = 0; i0 < total; i0 += U * PT_NT) {
        int run[U];
        int64_t pu[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int32_t sq = (i0 / PT_NT + u) * NW + wv;  // position inside the lane's chunk
            const int32_t i = chunk ? (sq < Q ? lane * Q + sq : total) : i0 + u * PT_NT + slot;
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
        if (run[u] < 0) continue;
        const int r = run[u] >> 1, ly = r % HY, lz = r / HY;
        const int lx = (run[u] & 1) ? 1 + (int)((cu[u] & msk) - (uint32_t)x0) : 0;
        const int gx = (x0 - 1 + lx) & (int)msk, gy = (y0 - 1 + ly) & (int)msk, gz = (z0 - 1 + lz) & (int)msk;
        T rd[3];
        cell_rel(xu[u], gx, gy, gz, N, rd);
        // CIC weight of corner b = 4 bx + 2 by + bz.  FP32 forms the eight
        // once from the (y,z) pair products and shares them between the
        // components (its kernel is bound by the VALU issue rate); FP64 forms
        // each product where it is used (eight more live FP64 values would
        // cost it a wave per SIMD)
        constexpr bool SHARE = sizeof(T) == 4;
        const T ox = o - rd[0], oy = o - rd[1], oz = o - rd[2];
        T pyz[4], W8[8];
        if constexpr (SHARE) {
            pyz[0] = oy * oz; pyz[1] = oy * rd[2]; pyz[2] = rd[1] * oz; pyz[3] = rd[1] * rd[2];
#pragma unroll
            for (int b = 0; b < 8; b++) W8[b] = ((b >> 2) ? rd[0] : ox) * pyz[b & 3];
        }
        // v times the weight of corner b
        auto Wv = [&](int b, T v) -> T {
            if constexpr (SHARE) return v * W8[b];
            else return ((b >> 2) ? rd[0] : ox) * (((b >> 1) & 1) ? rd[1] : oy) * ((b & 1) ? rd[2] : oz) * v;
        };
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
        const int pb = lx + PX * (ly + PY * lz);
        if constexpr (KIND == DEP_VEC3) {
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
                if (own(bx, by, bz)) {
                    const int si = sidx(bx, by, bz);
                    lds_add(&acc[0][si], Wv(b, qu[u][0]));
                    lds_add(&acc[1][si], Wv(b, qu[u][1]));
                    lds_add(&acc[2][si], Wv(b, qu[u][2]));
                }
            }
        }
        if constexpr (KIND == DEP_T00 || T00TIJ) {
            // projection_T00 functor, gevolution.hpp:1729-1854
            if (use_phi) {
                const T q0 = qu[u][0], q1 = qu[u][1], q2 = qu[u][2];
                T f2 = q0 * q0 + q1 * q1 + q2 * q2;
                const T e = fsqrt(f2 + a * a), ie = frcp(e);
                const T f = (T)3 * e + f2 * ie;
                f2 = ((T)0.5 * f * f + f2 * frcp(o + f2 * ia2)) * ie;
                const T es = FIX ? e * (T)S0 : e, fs = FIX ? f * (T)S0 : f, f2s = FIX ? f2 * (T)S0 : f2;
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
                    if (own(bx, by, bz)) {
                        const T ph = phs[pb + bx + PX * (by + PY * bz)];
                        lds_add(&acc[0][sidx(bx, by, bz)], Wv(b, es + fs * ph + f2s * ph * ph));
                    }
                }
            } else {
                const T sc = FIX ? (T)S0 : o;
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
                    if (own(bx, by, bz)) lds_add(&acc[0][sidx(bx, by, bz)], Wv(b, sc));
                }
            }
        }
        if constexpr (KIND == DEP_T0I || KIND == DEP_T00TIJL) {
            // projection_T0i functor, gevolution.hpp:2031-2201: 12 links per cell,
            // each weighted by (1 + phi_a + phi_b) of its two end points (DEP_T0I;
            // DEP_T00TIJL leaves the factor to the B solve)
            constexpr int L0 = KIND == DEP_T0I ? 0 : 7;  // first T0i accumulator
            const T sq = FIX ? (T)S2 : o;
            const T v0 = qu[u][0] * sq, v1 = qu[u][1] * sq, v2 = qu[u][2] * sq;
            // link k: the pair product of the two axes across it times v
            T g[12];
            if constexpr (SHARE) {
                const T pxz[4] = {ox * oz, ox * rd[2], rd[0] * oz, rd[0] * rd[2]};
                const T pxy[4] = {ox * oy, ox * rd[1], rd[0] * oy, rd[0] * rd[1]};
                const T gs[12] = {pyz[0] * v0, pxz[0] * v1, pxy[0] * v2, pxz[2] * v1, pxy[2] * v2, pyz[2] * v0,
                                  pxy[1] * v2, pyz[1] * v0, pxz[1] * v1, pyz[3] * v0, pxz[3] * v1, pxy[3] * v2};
#pragma unroll
                for (int k = 0; k < 12; k++) g[k] = gs[k];
            } else {
                const T gs[12] = {oy * oz * v0,       ox * oz * v1,       ox * oy * v2,       rd[0] * oz * v1,
                                  rd[0] * oy * v2,    rd[1] * oz * v0,    rd[1] * ox * v2,    rd[2] * oy * v0,
                                  rd[2] * ox * v1,    rd[1] * rd[2] * v0, rd[0] * rd[2] * v1, rd[0] * rd[1] * v2};
#pragma unroll
                for (int k = 0; k < 12; k++) g[k] = gs[k];
            }
            if (KIND == DEP_T0I && use_phi) {
                T ph[8];
#pragma unroll
                for (int b = 0; b < 8; b++) ph[b] = phs[pb + ((b >> 2) & 1) + PX * (((b >> 1) & 1) + PY * (b & 1))];
                g[0] *= o + ph[0] + ph[4]; g[1] *= o + ph[0] + ph[2]; g[2] *= o + ph[0] + ph[1];
                g[3] *= o + ph[4] + ph[6]; g[4] *= o + ph[4] + ph[5]; g[5] *= o + ph[2] + ph[6];
                g[6] *= o + ph[2] + ph[3]; g[7] *= o + ph[1] + ph[5]; g[8] *= o + ph[1] + ph[3];
                g[9] *= o + ph[3] + ph[7]; g[10] *= o + ph[5] + ph[7]; g[11] *= o + ph[6] + ph[7];
            }
            // link k lands on component comp[k] of the site cell + offset[k]
            if (own(0, 0, 0)) {
                lds_add(&acc[L0 + 0][sidx(0, 0, 0)], g[0]);
                lds_add(&acc[L0 + 1][sidx(0, 0, 0)], g[1]);
                lds_add(&acc[L0 + 2][sidx(0, 0, 0)], g[2]);
            }
            if (own(1, 0, 0)) { lds_add(&acc[L0 + 1][sidx(1, 0, 0)], g[3]); lds_add(&acc[L0 + 2][sidx(1, 0, 0)], g[4]); }
            if (own(0, 1, 0)) { lds_add(&acc[L0 + 0][sidx(0, 1, 0)], g[5]); lds_add(&acc[L0 + 2][sidx(0, 1, 0)], g[6]); }
            if (own(0, 0, 1)) { lds_add(&acc[L0 + 0][sidx(0, 0, 1)], g[7]); lds_add(&acc[L0 + 1][sidx(0, 0, 1)], g[8]); }
            if (own(0, 1, 1)) lds_add(&acc[L0 + 0][sidx(0, 1, 1)], g[9]);
            if (own(1, 0, 1)) lds_add(&acc[L0 + 1][sidx(1, 0, 1)], g[10]);
            if (own(1, 1, 0)) lds_add(&acc[L0 + 2][sidx(1, 1, 0)], g[11]);
        }
        if constexpr (KIND == DEP_TIJ || T00TIJ) {
            // projection_Tij functor, gevolution.hpp:2582-2965 (no CIC_PROJECT_TIJ):
            // diagonal parts CIC onto the 8 corners, off-diagonal parts onto 2 sites
            const T vv[3] = {qu[u][0], qu[u][1], qu[u][2]};
            T f = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
            const T e2 = f + a * a, ie = frsqrt(e2);
            if (use_phi) {
                const T *P = phs + pb;
                f = (T)4 + a * a * frcp(e2);
                T pint = (T)0;  // phi at the particle, CIC-interpolated
                if constexpr (SHARE) {
#pragma unroll
                    for (int b = 0; b < 8; b++) pint += Wv(b, P[((b >> 2) & 1) + PX * (((b >> 1) & 1) + PY * (b & 1))]);
                } else {
                    pint = ox * (oy * (oz * P[0] + rd[2] * P[PX * PY]) + rd[1] * (oz * P[PX] + rd[2] * P[PX + PX * PY])) +
                           rd[0] * (oy * (oz * P[1] + rd[2] * P[1 + PX * PY]) +
                                    rd[1] * (oz * P[1 + PX] + rd[2] * P[1 + PX + PX * PY]));
                }
                f *= pint;
                f += o;
            } else {
                f = o;
            }
            if constexpr (FIX) f *= (T)S1;
            constexpr int dcomp[3] = {TB + 0, TB + 3, TB + 5};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const T w = f * vv[d] * vv[d] * ie;
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const int bx = (b >> 2) & 1, by = (b >> 1) & 1, bz = b & 1;
                    if (own(bx, by, bz)) lds_add(&acc[dcomp[d]][sidx(bx, by, bz)], Wv(b, w));
                }
            }
            T w = f * vv[0] * vv[1] * ie;
            if (own(0, 0, 0)) lds_add(&acc[TB + 1][sidx(0, 0, 0)], w * (o - rd[2]));
            if (own(0, 0, 1)) lds_add(&acc[TB + 1][sidx(0, 0, 1)], w * rd[2]);
            w = f * vv[0] * vv[2] * ie;
            if (own(0, 0, 0)) lds_add(&acc[TB + 2][sidx(0, 0, 0)], w * (o - rd[1]));
            if (own(0, 1, 0)) lds_add(&acc[TB + 2][sidx(0, 1, 0)], w * rd[1]);
            w = f * vv[1] * vv[2] * ie;
            if (own(0, 0, 0)) lds_add(&acc[TB + 4][sidx(0, 0, 0)], w * (o - rd[0]));
            if (own(1, 0, 0)) lds_add(&acc[TB + 4][sidx(1, 0, 0)], w * rd[0]);
        }
        }
    }
    __syncthreads();

    // per component group (T00 and the mass deposit, Tij, T0i): the mass over
    // the group's fixed-point scale, divided once
    double fac[3] = {mass_d, mass_d, mass_d};
    if constexpr (FIX) {
        const double m1 = T00TIJ ? mass2_d : mass_d, m2 = KIND == DEP_T00TIJL ? mass3_d : mass_d;
        fac[0] = mass_d / (double)S0;
        fac[1] = m1 / (double)S1;
        fac[2] = m2 / (double)S2;
    }
    // component c's output array and the factor of its group
    auto out_of = [&](int c) -> T * {
        if (T00TIJ && c > 0) return (KIND == DEP_T00TIJL && c >= 7) ? out3 + (c - 7) * M : out2 + (c - 1) * M;
        return out + c * M;
    };
    auto grp_of = [&](int c) {
        return KIND == DEP_T00 ? 0 : KIND == DEP_TIJ ? 1 : KIND == DEP_T0I ? 2 : (c == 0 ? 0 : (c < 7 ? 1 : 2));
    };
    if (FIX && vec4) {
        // four consecutive sites of one component per thread and pass: one
        // 16-byte LDS read and one 16-byte store (outputs 16-byte aligned)
        const float ff[3] = {(float)fac[0], (float)fac[1], (float)fac[2]};
        constexpr int NV = NS / 4;
        for (int w = t; w < C * NV; w += PT_NT) {
            // NV a multiple of the wave size: the component is the same across the wave
            const int c = NV % 64 == 0 ? __builtin_amdgcn_readfirstlane(w / NV) : w / NV;
            const int s = (w - c * NV) * 4;
            const int sx = s % TX, sy = (s / TX) % TY, sz = s / (TX * TY);
            const uint32_t gs = (uint32_t)(x0 + sx) + (uint32_t)N * ((uint32_t)(y0 + sy) + (uint32_t)N * (uint32_t)(z0 + sz));
            const int g = grp_of(c);
            const float f = g == 0 ? ff[0] : (g == 1 ? ff[1] : ff[2]);
            const int4 r = *reinterpret_cast<const int4 *>(&acc[c][s]);
            float4 val = make_float4((float)r.x * f, (float)r.y * f, (float)r.z * f, (float)r.w * f);
            float4 *dst = reinterpret_cast<float4 *>(out_of(c) + gs);
            if (accumulate) {
                const float4 o4 = *dst;
                val.x += o4.x; val.y += o4.y; val.z += o4.z; val.w += o4.w;
            }
            *dst = val;
        }
        return;
    }
    for (int s = t; s < NS; s += PT_NT) {
        const int sx = s % TX, sy = (s / TX) % TY, sz = s / (TX * TY);
        const uint32_t gs = (uint32_t)(x0 + sx) + (uint32_t)N * ((uint32_t)(y0 + sy) + (uint32_t)N * (uint32_t)(z0 + sz));
#pragma unroll
        for (int c = 0; c < C; c++) {
            T *oc = out_of(c);
            const double ms = (T00TIJ && c > 0) ? ((KIND == DEP_T00TIJL && c >= 7) ? mass3_d : mass2_d) : mass_d;
            T val = FIX ? (T)((double)acc[c][s] * fac[grp_of(c)]) : (T)(acc[c][s] * ms);
            if (accumulate) val += oc[gs];
            oc[gs] = val;
        }
    }
}

template <typename T, int KIND, int TX, int TY, int TZ>
static int launch_dep(gev_pcls *p, T *out, T *out2, const T *phi, double mass, double mass2, double a, int acc,
                      const T *aux, T *out3, double mass3)
{
    gev_ctx *c = p->ctx;
    const int N = c->N;
    const unsigned nb = (unsigned)((int64_t)(N / TX) * (N / TY) * (N / TZ));
    // the updates' |q|^2 block bounds replace the per-particle bound pass
    // of the fixed-point tiles while they are valid (not for the adjoint's
    // vector deposit, whose values are not momenta)
    const uint32_t *qb = (sizeof(T) == 4 && KIND != DEP_VEC3 && p->qbound_valid) ? p->qbound : nullptr;
    // 16-byte stores of four sites when every output is 16-byte aligned
    const int vec4 = (((uintptr_t)out | (uintptr_t)out2 | (uintptr_t)out3) & 15) == 0;
    hipLaunchKernelGGL((k_deposit<T, KIND, TX, TY, TZ>), dim3(nb), dim3(PT_NT), 0, c->stream, make_pv<T>(p), out, out2,
                       out3, c->M, phi, mass, mass2, mass3, a, 1.0 / (a * a), acc, aux, qb, vec4);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T, int KIND>
static int dispatch_dep(gev_pcls *p, T *out, const T *phi, double mass, double a, int acc, T *out2 = nullptr,
                        double mass2 = 0.0, const T *aux = nullptr, T *out3 = nullptr, double mass3 = 0.0)
{
    const int N = p->ctx->N;
    // Tij keeps 6 (T00+Tij 7) FP64 accumulators per site: the smaller tile holds 4 blocks per CU
    constexpr bool big = KIND != DEP_TIJ && KIND != DEP_T00TIJ && KIND != DEP_T00TIJL;
    if (KIND == DEP_T0I && N >= 32) return launch_dep<T, KIND, 32, 4, 4>(p, out, out2, phi, mass, mass2, a, acc, aux, out3, mass3);
    if (big && N >= 16) return launch_dep<T, KIND, 16, 8, 8>(p, out, out2, phi, mass, mass2, a, acc, aux, out3, mass3);
#ifndef DEP_F32_TY
#define DEP_F32_TY 4
#endif
    // the FP32 form's fixed-point accumulators take half the LDS of FP64 ones
    if ((KIND == DEP_T00TIJ || KIND == DEP_T00TIJL) && N >= 16 && sizeof(T) == 4 && DEP_F32_TY != 4)
        return launch_dep<T, KIND, 16, DEP_F32_TY, 4>(p, out, out2, phi, mass, mass2, a, acc, aux, out3, mass3);
    if ((KIND == DEP_T00TIJ || KIND == DEP_T00TIJL) && N >= 16)
        return launch_dep<T, KIND, 16, 4, 4>(p, out, out2, phi, mass, mass2, a, acc, aux, out3, mass3);
    if (N >= 8) return launch_dep<T, KIND, 8, 8, 8>(p, out, out2, phi, mass, mass2, a, acc, aux, out3, mass3);
    return launch_dep<T, KIND, 4, 4, 4>(p, out, out2, phi, mass, mass2, a, acc, aux, out3, mass3);
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

// T00 and Tij as dep_T00_Tij_impl, and in links[3][M] the T0i link sums
// without their 1 + phi_a + phi_b factor (mass0i per particle)
template <typename T>
int dep_T00_Tij_links_impl(gev_pcls *p, T *T00, T *Tij, T *links, const T *phi, double mass00, double massij,
                           double mass0i, double a)
{
    return dispatch_dep<T, DEP_T00TIJL>(p, T00, phi, mass00, a, 0, Tij, massij, nullptr, links, mass0i);
}

#define INST(T)                                                                   \
    template int dep_vec3_impl<T>(gev_pcls *, const T *, T *);                       \
    template int dep_T00_Tij_links_impl<T>(gev_pcls *, T *, T *, T *, const T *, double, double, double, double); \
    template int dep_T00_Tij_impl<T>(gev_pcls *, T *, T *, const T *, double, double, double, int); \
    template int dep_T00_impl<T>(gev_pcls *, T *, const T *, double, double, int); \
    template int dep_T0i_impl<T>(gev_pcls *, T *, const T *, double, int);         \
    template int dep_Tij_impl<T>(gev_pcls *, T *, const T *, double, double, int);
INST(double)
INST(float)
#undef INST

}  // namespace gev

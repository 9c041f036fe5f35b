// gev_particles.hip -- particle container kernels for CDNA4:
//   * load / store / cell-centred lattice of the container (FP64 absolute
//     positions in, FP64 or FP32 cell offsets plus an int32 cell key inside);
//   * re-binning by cell: k_count ranks each particle in its cell with one
//     counter atomic (the fused kick/drift in gev_pupdate.hip takes the same
//     ranks itself), a three-kernel exclusive scan turns the counts into
//     cell_start and clears the counters, k_perm writes each particle's old
//     slot and cell key at its new slot, and k_gather copies the rest of the
//     particle into the ping-pong arrays with coalesced stores;
//   * the IC displacement and velocity gathers (one thread per particle in
//     cell order).
// The order of particles inside a cell follows the counter atomics, so it can
// change from run to run; IDs, cell keys and cell_start do not.  The CIC
// deposits live in gev_deposit.hip (LDS atomics: the FP64 sums are not
// bit-reproducible from run to run, the FP32 form's 32-bit fixed-point sums
// are) and the kick/drift in gev_pupdate.hip.
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
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n;
    const int32_t r = wave_rank_add(count, active ? cell[i] : 0, active);
    if (active) rank[i] = r;
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
    // one block of 1024 threads, each owns a contiguous run (read 16 bytes at
    // a time when the runs are whole multiples of four)
    __shared__ int32_t sh[1024];
    int64_t per = (nb + 1023) / 1024;
    int64_t b0 = threadIdx.x * per;
    int32_t s = 0;
    if (per % 4 == 0 && b0 + per <= nb) {
        const int4 *v = reinterpret_cast<const int4 *>(bsum + b0);
        for (int64_t j = 0; j < per / 4; j++) {
            const int4 q = v[j];
            s += q.x + q.y + q.z + q.w;
        }
    } else {
        for (int64_t j = 0; j < per; j++)
            if (b0 + j < nb) s += bsum[b0 + j];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int32_t y = (threadIdx.x >= (unsigned)o) ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += y;
        __syncthreads();
    }
    int32_t run = sh[threadIdx.x] - s;
    if (per % 4 == 0 && b0 + per <= nb) {
        int4 *v = reinterpret_cast<int4 *>(bsum + b0);
        for (int64_t j = 0; j < per / 4; j++) {
            const int4 q = v[j];
            int4 o;
            o.x = run; run += q.x;
            o.y = run; run += q.y;
            o.z = run; run += q.z;
            o.w = run; run += q.w;
            v[j] = o;
        }
    } else {
        for (int64_t j = 0; j < per; j++)
            if (b0 + j < nb) {
                int32_t t = bsum[b0 + j];
                bsum[b0 + j] = run;
                run += t;
            }
    }
}

// writes the exclusive prefix of cnt to start and clears cnt for the next count;
// start[M] = total, or the sum of all counts when total < 0
__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(int32_t *__restrict__ cnt, int64_t M,
                                                        const int32_t *__restrict__ bsum, int32_t *__restrict__ start,
                                                        int32_t total)
{
    static_assert(SCAN_PER == 8, "two 16-byte vectors per thread");
    // each thread owns SCAN_PER consecutive cells: 16-byte loads and stores
    // where the whole run is inside [0, M) (always when M is a multiple of 8)
    const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_PER;
    const bool full = base + SCAN_PER <= M;
    int32_t v[SCAN_PER];
    if (full) {
        int4 *c4 = (int4 *)(cnt + base);
        const int4 a = c4[0], b = c4[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        c4[0] = make_int4(0, 0, 0, 0);
        c4[1] = make_int4(0, 0, 0, 0);
    } else {
        for (int j = 0; j < SCAN_PER; j++) {
            const int64_t i = base + j;
            v[j] = (i < M) ? cnt[i] : 0;
            if (i < M) cnt[i] = 0;
        }
    }
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_PER; j++) s += v[j];
    int32_t tot;
    int32_t off = block_excl_scan(s, tot) + bsum[blockIdx.x];
    int32_t o[SCAN_PER];
#pragma unroll
    for (int j = 0; j < SCAN_PER; j++) {
        o[j] = off;
        off += v[j];
    }
    if (full) {
        int4 *s4 = (int4 *)(start + base);
        s4[0] = make_int4(o[0], o[1], o[2], o[3]);
        s4[1] = make_int4(o[4], o[5], o[6], o[7]);
    } else {
        for (int j = 0; j < SCAN_PER; j++)
            if (base + j < M) start[base + j] = o[j];
    }
    if (total >= 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) start[M] = total;
    } else {
#pragma unroll
        for (int j = 0; j < SCAN_PER; j++)
            if (base + j == M - 1) start[M] = o[j] + v[j];
    }
}

// Re-binning in two passes: k_perm scatters only the old slot and the cell
// key of each particle to its new slot (start of its cell + its rank), then
// k_gather copies the rest of every particle to its new slot with coalesced
// stores.  Particles mostly keep
// their place, so the gathered loads stay close to sequential.  Scattering the
// 36 or 60 bytes of each particle directly breaks the stores of a wave into
// partial lines and measured about 1.5x slower.
__global__ __launch_bounds__(PT_NT) void k_perm(const int32_t *__restrict__ cell, const int32_t *__restrict__ start,
                                                const int32_t *__restrict__ rank, int32_t *__restrict__ perm,
                                                int32_t *__restrict__ cell_out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * PT_NT + threadIdx.x;
    if (i >= n) return;
    const int32_t c = cell[i], j = start[c] + rank[i];
    perm[j] = (int32_t)i;
    cell_out[j] = c;
}

template <typename T>
__global__ __launch_bounds__(PT_NT) void k_gather(PV<T> s, PV<T> d, const int32_t *__restrict__ perm)
{
    const int64_t j = (int64_t)blockIdx.x * PT_NT + threadIdx.x;
    if (j >= s.n) return;
    const int64_t i = perm[j];
    d.p0[j] = s.p0[i]; d.p1[j] = s.p1[i]; d.p2[j] = s.p2[i];
    d.q0[j] = s.q0[i]; d.q1[j] = s.q1[i]; d.q2[j] = s.q2[i];
    d.id[j] = s.id[i];
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
        const T xs[3] = {v.p0[i], v.p1[i], v.p2[i]};
        T r[3];
        cell_rel(xs, x, y, z, N, r);
        T g[3];
        grad1(FieldAt<T>{xi, N, x, y, z}, r, g);
        const T Nn = (T)N;
        g[0] *= Nn; g[1] *= Nn; g[2] *= Nn;
        ret = coeff * sqrt((double)(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]));
        double d[3] = {coeff * (double)g[0], coeff * (double)g[1], coeff * (double)g[2]};
        move_by(v, i, xs, x, y, z, d);
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

// counters must start at zero: normally the last scan left them so
int pcls_clear_counts(gev_pcls *p)
{
    if (!p->counts_zero) {
        int rc = field_fill_impl<int32_t>(p->ctx, p->cell_count, p->ctx->M, 0);
        if (rc) return rc;
    }
    p->counts_zero = false;  // the caller is about to count
    return GEV_OK;
}

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
    p->counts_zero = true;
    if (n > 0) {
        PV<T> src = make_pv<T>(p);
        PV<T> dst = src;
        dst.p0 = (T *)p->pos2[0]; dst.p1 = (T *)p->pos2[1]; dst.p2 = (T *)p->pos2[2];
        dst.q0 = (T *)p->q2[0]; dst.q1 = (T *)p->q2[1]; dst.q2 = (T *)p->q2[2];
        dst.id = p->id2;
        dst.cell = p->cell2;
        hipLaunchKernelGGL(k_perm, dim3(pblocks(n)), dim3(PT_NT), 0, s, p->cell, p->cell_start, p->rank, p->perm,
                           p->cell2, n);
        hipLaunchKernelGGL((k_gather<T>), dim3(pblocks(n)), dim3(PT_NT), 0, s, src, dst, p->perm);
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

int scan_exclusive_counts(gev_ctx *ctx, int32_t *cnt, int64_t M, int32_t *start)
{
    const int64_t nb = (M + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (M <= 0 || nb > ctx->scan_blocks_n) {
        gev_set_error("scan_exclusive_counts: length outside the context's scan space");
        return GEV_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_NT), 0, s, cnt, M, ctx->scan_blocks);
    hipLaunchKernelGGL(k_scan_mid, dim3(1), dim3(1024), 0, s, ctx->scan_blocks, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_NT), 0, s, cnt, M, ctx->scan_blocks, start,
                       (int32_t)-1);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T> int pcls_sort_impl(gev_pcls *p)
{
    gev_ctx *ctx = p->ctx;
    const int64_t n = p->n;
    int rc = pcls_clear_counts(p);
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
    template int displace_impl<T>(gev_pcls *, const T *, double, double *);                                       \
    template int initialize_q_impl<T>(gev_pcls *, const T *, double, double *);
INST(double)
INST(float)
#undef INST

}  // namespace gev

// gev_lightcone.hip -- lightcone particle output for one GPU: the lightcone
// form of perfParticles_gevolution::saveGadget2 (device path,
// Particles_gevolution.hpp:1270-1560 kernels, :1566-2317 driver), the vertex
// search findIntersectingLightcones (tools.hpp:619-862) and the ID backlog of
// particles written near the inner edge of the previous cycle's shell
// (lightcone_id_backlog.hpp, the IDprelog / IDbacklog hand-over of
// writeLightcones, output.hpp:1017-1185, 2166-2880, for one rank).
//
// Selection, per particle (FP64, from the stored state):
//   ID % tracer_factor == 0; the first vertex v (periodic copy of the
//   observer) with inner <= d < outer, d = |x - v|, and, unless
//   opening == -1, (x - v).direction / d > opening; dropped when
//   outer - d <= 2 ZONE dtau_old (the check zone) and the ID is in the
//   backlog.  inner = dist - dtau/2, outer = dist + (1/2 + ZONE) dtau_old.
// Output: the particle moved to the crossing, dtau_pos = (dist - d)/(1 + vlos),
// kicked by dtau_pos + dtau_old/2 at a' = a + dtau_pos dadtau with the CIC
// value and edge gradient of phi (bufferTracerParticle with
// EXACT_OUTPUT_REDSHIFTS, :500-592), shifted by (lightcone vertex - v) and
// scaled by BoxSize.  IDs with d - inner < 2 ZONE dtau go to the ID log's
// prelog, the next cycle's backlog.
//
// On the device: a counting pass writes one count per 1024 particles, an
// exclusive scan turns the counts into output offsets, and a second pass
// writes the selected particles in container order (deterministic, unlike the
// reference's atomic slots) into the container's ping-pong arrays.  The
// selection arithmetic is compiled without contraction so that it is the
// same IEEE arithmetic as a plain C evaluation of the same expressions.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_pcommon.hpp"

struct gev_idlog {
    gev_ctx *ctx = nullptr;
    std::vector<int64_t> backlog;  // sorted, unique
    std::vector<int64_t> prelog;   // IDs of this cycle's outputs near the inner edge
    int64_t *dev = nullptr;        // device copy of backlog
    int64_t dev_cap = 0;
};

namespace gev {

constexpr double LC_ZONE = GEV_LIGHTCONE_IDCHECK_ZONE;
constexpr double LC_GVC = 3.335640952e-6;  // GADGET_VELOCITY_CONVERSION, metadata.hpp:117
constexpr int LC_NT = 256, LC_PER = 4, LC_CHUNK = LC_NT * LC_PER;

struct LcArgs {
    double vert[GEV_LIGHTCONE_MAX_VERTICES][3];
    double lcv[3], dir[3];
    double opening, dist, inner, outer, zc, zp, dtau_old, dadtau, a, boxsize;
    int64_t tf;
    int nv;
};

struct LcOut {
    float *px, *py, *pz, *vx, *vy, *vz;
    int64_t *id;
    int32_t *prelog;
};

template <typename T> struct FieldAtD {
    const T *f;
    int N, x, y, z;
    __device__ __forceinline__ double operator()(int dx, int dy, int dz) const
    {
        return (double)f[site(N, x + dx, y + dy, z + dz)];
    }
};

template <typename T>
__device__ __forceinline__ void lc_position(const PV<T> &v, int64_t i, double P[3], int &x, int &y, int &z)
{
    cell_coords(v.cell[i], v.N, v.lg, x, y, z);
    if constexpr (sizeof(T) == 8) {
        P[0] = v.p0[i]; P[1] = v.p1[i]; P[2] = v.p2[i];
    } else {
        const double dx = 1.0 / (double)v.N;
        P[0] = ((double)x + (double)v.p0[i]) * dx;
        P[1] = ((double)y + (double)v.p1[i]) * dx;
        P[2] = ((double)z + (double)v.p2[i]) * dx;
    }
}

__device__ __forceinline__ bool lc_backlog_has(const int64_t *__restrict__ b, int64_t nb, int64_t id)
{
    int64_t lo = 0, hi = nb;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (b[mid] < id) lo = mid + 1;
        else hi = mid;
    }
    return lo < nb && b[lo] == id;
}

// the selection of one particle; hit = vertex index, dh = its distance
template <typename T>
__device__ __forceinline__ bool lc_select(const LcArgs &g, const PV<T> &v, int64_t i, const int64_t *__restrict__ bl,
                                          int64_t nbl, int &hit, double &dh)
{
#pragma clang fp contract(off)
    hit = -1;
    dh = 0.0;
    if (i >= v.n) return false;
    const int64_t id = v.id[i];
    if (id % g.tf != 0) return false;
    double P[3];
    int x, y, z;
    lc_position(v, i, P, x, y, z);
    for (int k = 0; k < g.nv; k++) {
        const double r0 = P[0] - g.vert[k][0], r1 = P[1] - g.vert[k][1], r2 = P[2] - g.vert[k][2];
        const double d = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
        if (!(d >= g.inner && d < g.outer)) continue;
        if (g.opening == -1.0 || (r0 * g.dir[0] + r1 * g.dir[1] + r2 * g.dir[2]) / d > g.opening) {
            hit = k;
            dh = d;
            break;
        }
    }
    if (hit < 0) return false;
    if (g.outer - dh <= g.zc && nbl > 0 && lc_backlog_has(bl, nbl, id)) return false;
    return true;
}

template <typename T>
__global__ __launch_bounds__(LC_NT) void k_lc_count(PV<T> v, LcArgs g, const int64_t *__restrict__ bl, int64_t nbl,
                                                    int32_t *__restrict__ bcount)
{
    __shared__ int32_t wsum[LC_NT / 64];
    const int64_t base = (int64_t)blockIdx.x * LC_CHUNK;
    int c = 0;
    for (int j = 0; j < LC_PER; j++) {
        int h;
        double dh;
        c += lc_select(g, v, base + j * LC_NT + threadIdx.x, bl, nbl, h, dh) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t s = 0;
        for (int w = 0; w < LC_NT / 64; w++) s += wsum[w];
        bcount[blockIdx.x] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(LC_NT) void k_lc_write(PV<T> v, LcArgs g, const int64_t *__restrict__ bl, int64_t nbl,
                                                    const int32_t *__restrict__ bstart, const T *__restrict__ phi,
                                                    LcOut o)
{
    __shared__ int32_t wc[LC_PER][LC_NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * LC_CHUNK;
    bool sel[LC_PER];
    int hit[LC_PER], pre[LC_PER];
    double dh[LC_PER];
#pragma unroll
    for (int j = 0; j < LC_PER; j++) {
        sel[j] = lc_select(g, v, base + j * LC_NT + threadIdx.x, bl, nbl, hit[j], dh[j]);
        const unsigned long long m = __ballot(sel[j]);
        pre[j] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wc[j][w] = __popcll(m);
    }
    __syncthreads();
    int32_t run = bstart[blockIdx.x];
    int32_t slot[LC_PER];
#pragma unroll
    for (int j = 0; j < LC_PER; j++) {
        int32_t s = run;
        for (int w2 = 0; w2 < LC_NT / 64; w2++) {
            if (w2 < w) s += wc[j][w2];
            run += wc[j][w2];
        }
        slot[j] = s + pre[j];
    }
    const int N = v.N;
#pragma unroll
    for (int j = 0; j < LC_PER; j++) {
        if (!sel[j]) continue;
        const int64_t i = base + j * LC_NT + threadIdx.x;
        double P[3];
        int x, y, z;
        lc_position(v, i, P, x, y, z);
        const double *vx = g.vert[hit[j]];
        const double d = dh[j];
        const double Q[3] = {(double)v.q0[i], (double)v.q1[i], (double)v.q2[i]};
        const double q2 = Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2];
        const double vlos = (Q[0] * (P[0] - vx[0]) + Q[1] * (P[1] - vx[1]) + Q[2] * (P[2] - vx[2])) / d /
                            sqrt(q2 + g.a * (g.a + g.dtau_old * g.dadtau));
        const double dtp = (g.dist - d) / (1.0 + vlos), dtv = dtp + 0.5 * g.dtau_old, ap = g.a + dtp * g.dadtau;
        double phip = 0.0, gr[3] = {0.0, 0.0, 0.0};
        if (phi) {
            const FieldAtD<T> F{phi, N, x, y, z};
            double fr[3];
            if constexpr (sizeof(T) == 8) {
                cell_rel(P, x, y, z, N, fr);
            } else {
                fr[0] = (double)v.p0[i]; fr[1] = (double)v.p1[i]; fr[2] = (double)v.p2[i];
            }
            phip = cic<double>(F, fr);
            grad1<double>(F, fr, gr);
        }
        const double e2 = q2 + ap * ap, e = sqrt(e2), wq = q2 + e2;
        const double vfac = 1.0 + (4.0 - wq / e2) * phip;
        const double rescale = 1.0 / sqrt(ap) / LC_GVC;
        float *po[3] = {o.px, o.py, o.pz}, *vo[3] = {o.vx, o.vy, o.vz};
        const int32_t s = slot[j];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            po[c][s] = (float)((P[c] + dtp * Q[c] * vfac / e + (g.lcv[c] - vx[c])) * g.boxsize);
            vo[c][s] = (float)((Q[c] - dtv * wq * gr[c] * (double)N / e) * rescale / ap);
        }
        o.id[s] = v.id[i];
        o.prelog[s] = (d - g.inner < g.zp) ? 1 : 0;
    }
}

template <typename T>
static int lc_run(gev_pcls *p, const LcArgs &g, const gev_idlog *log, const T *phi, std::vector<float> &pos,
                  std::vector<float> &vel, std::vector<int64_t> &id, std::vector<int32_t> &pre)
{
    gev_ctx *c = p->ctx;
    hipStream_t s = c->stream;
    const int64_t n = p->n;
    pos.clear(); vel.clear(); id.clear(); pre.clear();
    if (n == 0) return GEV_OK;
    const int64_t nb = (n + LC_CHUNK - 1) / LC_CHUNK;
    if (nb > c->M) {
        gev_set_error("gev_pcls_save_gadget2_lightcone: more than 1024 particles per lattice site");
        return GEV_ERR_ARG;
    }
    // block counts in the cell counters (left zero again by the scan), offsets in rank
    if (!p->counts_zero) {
        int rc = pcls_clear_counts(p);
        if (rc) return rc;
    }
    p->counts_zero = false;
    int32_t *bcount = p->cell_count, *bstart = p->rank;
    const int64_t *bl = log ? log->dev : nullptr;
    const int64_t nbl = log ? (int64_t)log->backlog.size() : 0;
    PV<T> v = make_pv<T>(p);
    hipLaunchKernelGGL((k_lc_count<T>), dim3((unsigned)nb), dim3(LC_NT), 0, s, v, g, bl, nbl, bcount);
    GEV_HIP(hipGetLastError());
    int rc = scan_exclusive_counts(c, bcount, nb, bstart);
    if (rc) return rc;
    p->counts_zero = true;
    int32_t total = 0;
    GEV_HIP(hipMemcpyAsync(&total, bstart + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GEV_HIP(hipStreamSynchronize(s));
    if (total == 0) return GEV_OK;
    LcOut o;
    o.px = (float *)p->pos2[0]; o.py = (float *)p->pos2[1]; o.pz = (float *)p->pos2[2];
    o.vx = (float *)p->q2[0]; o.vy = (float *)p->q2[1]; o.vz = (float *)p->q2[2];
    o.id = p->id2;
    o.prelog = p->cell2;
    hipLaunchKernelGGL((k_lc_write<T>), dim3((unsigned)nb), dim3(LC_NT), 0, s, v, g, bl, nbl, bstart, phi, o);
    GEV_HIP(hipGetLastError());
    const size_t m = (size_t)total;
    pos.resize(3 * m);
    vel.resize(3 * m);
    id.resize(m);
    pre.resize(m);
    std::vector<float> tmp(m);
    const float *src[6] = {o.px, o.py, o.pz, o.vx, o.vy, o.vz};
    for (int k = 0; k < 6; k++) {
        GEV_HIP(hipMemcpyAsync(tmp.data(), src[k], sizeof(float) * m, hipMemcpyDeviceToHost, s));
        GEV_HIP(hipStreamSynchronize(s));
        std::vector<float> &dst = k < 3 ? pos : vel;
        for (size_t j = 0; j < m; j++) dst[3 * j + (k % 3)] = tmp[j];
    }
    GEV_HIP(hipMemcpyAsync(id.data(), o.id, sizeof(int64_t) * m, hipMemcpyDeviceToHost, s));
    GEV_HIP(hipMemcpyAsync(pre.data(), o.prelog, sizeof(int32_t) * m, hipMemcpyDeviceToHost, s));
    GEV_HIP(hipStreamSynchronize(s));
    return GEV_OK;
}

static double dist3(const double a[3], const double b[3])
{
    return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

static int idlog_upload(gev_idlog *log)
{
    const int64_t n = (int64_t)log->backlog.size();
    if (n > log->dev_cap) {
        if (log->dev) GEV_HIP(hipFree(log->dev));
        log->dev = nullptr;
        log->dev_cap = 0;
        const int64_t cap = std::max<int64_t>(n, 1024);
        GEV_HIP(hipMalloc(&log->dev, sizeof(int64_t) * cap));
        log->dev_cap = cap;
    }
    if (n > 0)
        GEV_HIP(hipMemcpyAsync(log->dev, log->backlog.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                               log->ctx->stream));
    GEV_HIP(hipStreamSynchronize(log->ctx->stream));
    return GEV_OK;
}

}  // namespace gev

using namespace gev;

extern "C" {

int gev_lightcone_vertices(const gev_lightcone_geometry *lc, double outer, double inner, const double domain[6],
                           double *vertex, int cap, int *count)
{
    if (!lc || !domain || !count || (cap > 0 && !vertex)) {
        gev_set_error("gev_lightcone_vertices: NULL argument");
        return GEV_ERR_ARG;
    }
    if (!(outer >= 0.0) || !std::isfinite(outer) || outer > 64.0) {
        gev_set_error("gev_lightcone_vertices: outer must be in [0, 64] box lengths");
        return GEV_ERR_ARG;
    }
    const double *lo = domain, *hi = domain + 3;
    double corner[8][3];
    for (int k = 0; k < 8; k++) {
        corner[k][0] = (k & 1) ? hi[0] : lo[0];
        corner[k][1] = (k & 2) ? hi[1] : lo[1];
        corner[k][2] = (k & 4) ? hi[2] : lo[2];
    }
    const double centre[3] = {0.5 * lo[0] + 0.5 * hi[0], 0.5 * lo[1] + 0.5 * hi[1], 0.5 * lo[2] + 0.5 * hi[2]};
    const double rdom = 0.5 * std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                        (hi[2] - lo[2]) * (hi[2] - lo[2]));
    const int range = (int)std::ceil(outer) + 1;
    int n = 0;
    for (int u = -range; u <= range; u++)
        for (int v = -range; v <= range; v++)
            for (int w = -range; w <= range; w++) {
                const double vx[3] = {lc->vertex[0] + u, lc->vertex[1] + v, lc->vertex[2] + w};
                // outside the outer sphere: distance to the nearest point of the domain
                double s2 = 0.0;
                for (int d = 0; d < 3; d++) {
                    const double e = vx[d] < lo[d] ? lo[d] - vx[d] : (vx[d] > hi[d] ? vx[d] - hi[d] : 0.0);
                    s2 += e * e;
                }
                if (std::sqrt(s2) > outer) continue;
                // inside the inner sphere: every corner nearer than inner
                bool inside = true;
                for (int k = 0; k < 8 && inside; k++) inside = dist3(corner[k], vx) < inner;
                if (inside) continue;
                const double dist = dist3(centre, vx);
                bool take = dist <= rdom;
                if (!take) {
                    const double cosang = ((centre[0] - vx[0]) * lc->direction[0] +
                                           (centre[1] - vx[1]) * lc->direction[1] +
                                           (centre[2] - vx[2]) * lc->direction[2]) / dist;
                    if (cosang >= lc->opening)
                        take = true;
                    else if (dist > outer && std::acos(cosang) - std::acos(lc->opening) <=
                                                 std::acos((outer * outer + dist * dist - rdom * rdom) /
                                                           (2. * outer * dist)))
                        take = true;
                    else if (dist <= outer && std::acos(cosang) - std::acos(lc->opening) <= std::asin(rdom / dist))
                        take = true;
                }
                if (take) {
                    if (n < cap)
                        for (int d = 0; d < 3; d++) vertex[3 * n + d] = vx[d];
                    n++;
                }
            }
    *count = n;
    return GEV_OK;
}

int gev_idlog_create(gev_ctx *ctx, gev_idlog **out)
{
    if (!ctx || !out) { gev_set_error("gev_idlog_create: NULL argument"); return GEV_ERR_ARG; }
    gev_idlog *log = new gev_idlog;
    log->ctx = ctx;
    *out = log;
    return GEV_OK;
}

int gev_idlog_destroy(gev_idlog *log)
{
    if (!log) return GEV_OK;
    if (log->dev) (void)hipFree(log->dev);
    delete log;
    return GEV_OK;
}

int gev_idlog_sizes(const gev_idlog *log, int64_t *nbacklog, int64_t *nprelog)
{
    if (!log) { gev_set_error("gev_idlog_sizes: NULL log"); return GEV_ERR_ARG; }
    if (nbacklog) *nbacklog = (int64_t)log->backlog.size();
    if (nprelog) *nprelog = (int64_t)log->prelog.size();
    return GEV_OK;
}

int gev_idlog_get(const gev_idlog *log, int which, int64_t *ids)
{
    if (!log || (which != 0 && which != 1)) { gev_set_error("gev_idlog_get: bad argument"); return GEV_ERR_ARG; }
    const std::vector<int64_t> &src = which == 0 ? log->backlog : log->prelog;
    if (!src.empty()) {
        if (!ids) { gev_set_error("gev_idlog_get: NULL output"); return GEV_ERR_ARG; }
        memcpy(ids, src.data(), sizeof(int64_t) * src.size());
    }
    return GEV_OK;
}

int gev_idlog_set_backlog(gev_idlog *log, const int64_t *ids, int64_t n)
{
    if (!log || n < 0 || (n > 0 && !ids)) { gev_set_error("gev_idlog_set_backlog: bad argument"); return GEV_ERR_ARG; }
    log->backlog.assign(ids, ids + n);
    std::sort(log->backlog.begin(), log->backlog.end());
    log->backlog.erase(std::unique(log->backlog.begin(), log->backlog.end()), log->backlog.end());
    return idlog_upload(log);
}

int gev_idlog_commit(gev_idlog *log)
{
    if (!log) { gev_set_error("gev_idlog_commit: NULL log"); return GEV_ERR_ARG; }
    std::vector<int64_t> ids;
    ids.swap(log->prelog);
    return gev_idlog_set_backlog(log, ids.data(), (int64_t)ids.size());
}

int gev_pcls_save_gadget2_lightcone(gev_pcls *p, const char *path, gev_gadget2_header *hdr,
                                    const gev_lightcone_geometry *lc, double dist, double dtau, double dtau_old,
                                    double dadtau, const double *vertex, int vertexcount, gev_idlog *log,
                                    const void *phi, int tracer_factor, int64_t *written)
{
    const char *fn = "gev_pcls_save_gadget2_lightcone";
    if (!p || !path || !hdr || !lc || (vertexcount > 0 && !vertex)) {
        gev_set_error(std::string(fn) + ": NULL argument");
        return GEV_ERR_ARG;
    }
    if (tracer_factor < 1) { gev_set_error(std::string(fn) + ": tracer_factor must be >= 1"); return GEV_ERR_ARG; }
    if (hdr->num_files != 1) { gev_set_error(std::string(fn) + ": only single-file output (num_files = 1)"); return GEV_ERR_ARG; }
    if (vertexcount < 0 || vertexcount > GEV_LIGHTCONE_MAX_VERTICES) {
        gev_set_error(std::string(fn) + ": vertexcount must be in [0, GEV_LIGHTCONE_MAX_VERTICES]");
        return GEV_ERR_ARG;
    }
    if (log && log->ctx != p->ctx) { gev_set_error(std::string(fn) + ": the ID log belongs to another context"); return GEV_ERR_ARG; }
    if (!p->sorted) { gev_set_error(std::string(fn) + ": particles are not binned"); return GEV_ERR_STATE; }
    LcArgs g;
    memset(&g, 0, sizeof(g));
    for (int k = 0; k < vertexcount; k++)
        for (int d = 0; d < 3; d++) g.vert[k][d] = vertex[3 * k + d];
    for (int d = 0; d < 3; d++) {
        g.lcv[d] = lc->vertex[d];
        g.dir[d] = lc->direction[d];
    }
    g.nv = vertexcount;
    g.opening = lc->opening;
    g.dist = dist;
    g.inner = dist - 0.5 * dtau;
    g.outer = dist + (0.5 + LC_ZONE) * dtau_old;
    g.zc = 2. * LC_ZONE * dtau_old;
    g.zp = 2. * LC_ZONE * dtau;
    g.dtau_old = dtau_old;
    g.dadtau = dadtau;
    g.a = hdr->time;
    g.boxsize = hdr->BoxSize;
    g.tf = tracer_factor;
    std::vector<float> pos, vel;
    std::vector<int64_t> id;
    std::vector<int32_t> pre;
    const int rc = p->ctx->prec == GEV_FP64
                       ? lc_run<double>(p, g, log, (const double *)phi, pos, vel, id, pre)
                       : lc_run<float>(p, g, log, (const float *)phi, pos, vel, id, pre);
    if (rc) return rc;
    const uint64_t m = id.size();
    if (written) *written = (int64_t)m;
    hdr->npart[1] = (uint32_t)(m % (1ull << 32));
    hdr->npartTotal[1] = (uint32_t)(m % (1ull << 32));
    hdr->npartTotalHW[1] = (uint32_t)(m >> 32);
    if (log)
        for (uint64_t j = 0; j < m; j++)
            if (pre[j]) log->prelog.push_back(id[j]);
    if (m == 0) return GEV_OK;  // no file, as the reference writes none for an empty shell
    if (m * 12 > 0xffffffffull) { gev_set_error(std::string(fn) + ": record exceeds 4 GiB"); return GEV_ERR_ARG; }
    FILE *f = fopen(path, "wb");
    if (!f) { gev_set_error(std::string(fn) + ": cannot open " + path); return GEV_ERR_ARG; }
    const uint32_t b3 = (uint32_t)(12 * m), b1 = (uint32_t)(8 * m);
    bool ok = gadget_put_record(f, hdr, sizeof(gev_gadget2_header)) && gadget_put_record(f, pos.data(), b3) &&
              gadget_put_record(f, vel.data(), b3) && gadget_put_record(f, id.data(), b1);
    ok = (fclose(f) == 0) && ok;
    if (!ok) { gev_set_error(std::string(fn) + ": write failed on " + path); return GEV_ERR_ARG; }
    return GEV_OK;
}

}  // extern "C"

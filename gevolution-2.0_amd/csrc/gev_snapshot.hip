// gev_snapshot.hip -- Gadget-2 particle snapshots (single file, type-1 particles):
// perfParticles_gevolution::saveGadget2 / loadGadget2
// (Particles_gevolution.hpp:597-800, 2455-2640) with the header of
// metadata.hpp:159-178 and the unit conventions of metadata.hpp:110-120.
//
// File layout (Fortran-style records, each framed by its byte count):
//   [256] header [256]  [12 n] positions float3 [12 n]  [12 n] velocities float3 [12 n]
//   [8 n] IDs int64 [8 n]
// Positions are in the header's length unit (BoxSize per box), velocities
// are q / (a^1.5 GADGET_VELOCITY_CONVERSION).  The writer selects the
// particles whose ID is a multiple of the tracer factor and, as the reference
// built with EXACT_OUTPUT_REDSHIFTS (config/defaults.mk:21) does, moves them to
// the snapshot time: positions drift by dtau_pos with the GR velocity
// q (1 + (4 - (2q^2 + a^2)/e^2) phi) / e, velocities kick by dtau_vel with
// (2q^2 + a^2)/e times the CIC edge gradient of phi (phi optional).
//
// The selection and conversion run on the device: a count per workgroup, the
// exclusive scan of the counts, then each selected particle written at its
// rank into the container's ping-pong arrays (free between cycles).  Only the
// selected particles are copied to the host, which writes the records in
// container (cell) order.
//
// The express format (saveExpress / loadExpress, Particles_gevolution.hpp:
// 2642-2900) is the raw restart state of one rank: a 120-byte express header,
// positions (box units) and momenta as Real triplets, IDs as int64, then the
// Gadget-2 header.  Real is the container's precision (4 bytes in the mixed
// form, as in the reference's single-precision build); this build is one rank
// of one, so the file is <name>.0.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "gev_common.hpp"
#include "gev_kernels.hpp"
#include "gev_pcommon.hpp"

namespace gev {

constexpr double GADGET_VELOCITY_CONVERSION = 3.335640952e-6;  // Gadget velocity unit / c (metadata.hpp:117)
static_assert(sizeof(gev_gadget2_header) == 256, "the Gadget-2 header is 256 bytes");

constexpr int GS_NT = 256, GS_PER = 4, GS_CHUNK = GS_NT * GS_PER;

// tracers (ID a multiple of the factor) among the GS_CHUNK particles of each workgroup
__global__ __launch_bounds__(GS_NT) void k_gadget_count(const int64_t *__restrict__ id, int64_t n, int tracer,
                                                        int32_t *__restrict__ bcount)
{
    __shared__ int32_t wsum[GS_NT / 64];
    const int64_t base = (int64_t)blockIdx.x * GS_CHUNK;
    int c = 0;
#pragma unroll
    for (int j = 0; j < GS_PER; j++) {
        const int64_t i = base + j * GS_NT + threadIdx.x;
        c += (i < n && id[i] % tracer == 0) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t t = 0;
        for (int w = 0; w < GS_NT / 64; w++) t += wsum[w];
        bcount[blockIdx.x] = t;
    }
}

// Each selected particle's snapshot position (length units) and velocity,
// written at its rank among the selected ones in container order: the
// workgroup's offset from the scan of k_gadget_count, then wave ballots.
template <typename T>
__global__ __launch_bounds__(GS_NT) void k_gadget_convert(PV<T> v, int tracer, const int32_t *__restrict__ bstart,
                                                          double dtau_pos, double dtau_vel, double a, double boxsize,
                                                          const T *__restrict__ phi, float *px, float *py, float *pz,
                                                          float *vx, float *vy, float *vz, int64_t *oid)
{
    __shared__ int32_t wc[GS_PER][GS_NT / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * GS_CHUNK;
    bool sel[GS_PER];
    int pre[GS_PER];
#pragma unroll
    for (int j = 0; j < GS_PER; j++) {
        const int64_t i = base + j * GS_NT + threadIdx.x;
        sel[j] = i < v.n && v.id[i] % tracer == 0;
        const unsigned long long m = __ballot(sel[j]);
        pre[j] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wc[j][wv] = __popcll(m);
    }
    __syncthreads();
    int32_t run = bstart[blockIdx.x];
    const int N = v.N;
#pragma unroll
    for (int j = 0; j < GS_PER; j++) {
        int32_t slot = run + pre[j];
        for (int w2 = 0; w2 < GS_NT / 64; w2++) {
            if (w2 < wv) slot += wc[j][w2];
            run += wc[j][w2];
        }
        if (!sel[j]) continue;
        const int64_t i = base + j * GS_NT + threadIdx.x;
        int x, y, z;
        cell_coords(v.cell[i], N, v.lg, x, y, z);
        const T xs[3] = {v.p0[i], v.p1[i], v.p2[i]};
        T fr[3];
        cell_rel(xs, x, y, z, N, fr);
        const double q[3] = {(double)v.q0[i], (double)v.q1[i], (double)v.q2[i]};
        double phip = 0.0, g[3] = {0.0, 0.0, 0.0};
        if (phi) {
            // CIC value and edge gradient (differences, not yet divided by dx)
            auto P = [&](int dx_, int dy_, int dz_) { return (double)phi[site(N, x + dx_, y + dy_, z + dz_)]; };
            const double f0 = fr[0], f1 = fr[1], f2 = fr[2], m0 = 1.0 - f0, m1 = 1.0 - f1, m2 = 1.0 - f2;
            phip = P(0, 0, 0) * m0 * m1 * m2 + P(1, 0, 0) * f0 * m1 * m2 + P(0, 1, 0) * m0 * f1 * m2 +
                   P(1, 1, 0) * f0 * f1 * m2 + P(0, 0, 1) * m0 * m1 * f2 + P(1, 0, 1) * f0 * m1 * f2 +
                   P(0, 1, 1) * m0 * f1 * f2 + P(1, 1, 1) * f0 * f1 * f2;
            g[0] = m1 * m2 * (P(1, 0, 0) - P(0, 0, 0)) + f1 * m2 * (P(1, 1, 0) - P(0, 1, 0)) +
                   m1 * f2 * (P(1, 0, 1) - P(0, 0, 1)) + f1 * f2 * (P(1, 1, 1) - P(0, 1, 1));
            g[1] = m0 * m2 * (P(0, 1, 0) - P(0, 0, 0)) + f0 * m2 * (P(1, 1, 0) - P(1, 0, 0)) +
                   m0 * f2 * (P(0, 1, 1) - P(0, 0, 1)) + f0 * f2 * (P(1, 1, 1) - P(1, 0, 1));
            g[2] = m0 * m1 * (P(0, 0, 1) - P(0, 0, 0)) + f0 * m1 * (P(1, 0, 1) - P(1, 0, 0)) +
                   m0 * f1 * (P(0, 1, 1) - P(0, 1, 0)) + f0 * f1 * (P(1, 1, 1) - P(1, 1, 0));
        }
        const double q2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
        const double e2 = q2 + a * a, e = sqrt(e2);
        const double w = 2.0 * q2 + a * a;  // q^2 + e^2
        const double vfac = 1.0 + (4.0 - w / e2) * phip;
        const double rescale = 1.0 / sqrt(a) / GADGET_VELOCITY_CONVERSION;
        const double cell[3] = {(double)x, (double)y, (double)z};
        float *po[3] = {px, py, pz}, *vo[3] = {vx, vy, vz};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            // absolute position in box units, drifted, wrapped to [0, 1)
            const double p = (cell[d] + (double)fr[d]) / N + dtau_pos * q[d] * vfac / e;
            double wrapped = p - floor(p);
            if (wrapped >= 1.0) wrapped = 0.0;
            po[d][slot] = (float)(wrapped * boxsize);
            vo[d][slot] = (float)((q[d] - dtau_vel * w * g[d] * N / e) * rescale / a);
        }
        oid[slot] = v.id[i];
    }
}

// the selected particles' snapshot records, interleaved on the host
template <typename T>
static int gadget_convert(gev_pcls *p, int tracer, double dtau_pos, double dtau_vel, double a, double boxsize,
                          const T *phi, std::vector<float> &pos, std::vector<float> &vel, std::vector<int64_t> &id)
{
    gev_ctx *c = p->ctx;
    hipStream_t s = c->stream;
    const int64_t n = p->n;
    pos.clear(); vel.clear(); id.clear();
    if (n == 0) return GEV_OK;
    const int64_t nb = (n + GS_CHUNK - 1) / GS_CHUNK;
    if (nb > c->M) {
        gev_set_error("gev_pcls_save_gadget2: more than 1024 particles per lattice site");
        return GEV_ERR_ARG;
    }
    // workgroup counts in the cell counters (left zero again by the scan), offsets in rank
    int rc = pcls_clear_counts(p);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gadget_count, dim3((unsigned)nb), dim3(GS_NT), 0, s, p->id, n, tracer, p->cell_count);
    GEV_HIP(hipGetLastError());
    rc = scan_exclusive_counts(c, p->cell_count, nb, p->rank);
    if (rc) return rc;
    p->counts_zero = true;
    int32_t total = 0;
    GEV_HIP(hipMemcpyAsync(&total, p->rank + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GEV_HIP(hipStreamSynchronize(s));
    if (total == 0) return GEV_OK;
    // staging: the ping-pong arrays hold n values each
    float *st[6] = {(float *)p->pos2[0], (float *)p->pos2[1], (float *)p->pos2[2],
                    (float *)p->q2[0],   (float *)p->q2[1],   (float *)p->q2[2]};
    hipLaunchKernelGGL((k_gadget_convert<T>), dim3((unsigned)nb), dim3(GS_NT), 0, s, make_pv<T>(p), tracer, p->rank,
                       dtau_pos, dtau_vel, a, boxsize, phi, st[0], st[1], st[2], st[3], st[4], st[5], p->id2);
    GEV_HIP(hipGetLastError());
    const size_t m = (size_t)total;
    pos.resize(3 * m);
    vel.resize(3 * m);
    id.resize(m);
    std::vector<float> tmp(m);
    for (int k = 0; k < 6; k++) {
        GEV_HIP(hipMemcpyAsync(tmp.data(), st[k], sizeof(float) * m, hipMemcpyDeviceToHost, s));
        GEV_HIP(hipStreamSynchronize(s));
        std::vector<float> &dst = k < 3 ? pos : vel;
        for (size_t j = 0; j < m; j++) dst[3 * j + (k % 3)] = tmp[j];
    }
    GEV_HIP(hipMemcpyAsync(id.data(), p->id2, sizeof(int64_t) * m, hipMemcpyDeviceToHost, s));
    GEV_HIP(hipStreamSynchronize(s));
    return GEV_OK;
}

bool gadget_put_record(FILE *f, const void *data, uint32_t bytes)
{
    return fwrite(&bytes, 4, 1, f) == 1 && (bytes == 0 || fwrite(data, 1, bytes, f) == bytes) &&
           fwrite(&bytes, 4, 1, f) == 1;
}

static bool get_record(FILE *f, void *data, uint32_t bytes)
{
    uint32_t b0 = 0, b1 = 0;
    if (fread(&b0, 4, 1, f) != 1 || b0 != bytes) return false;
    if (bytes > 0 && fread(data, 1, bytes, f) != bytes) return false;
    return fread(&b1, 4, 1, f) == 1 && b1 == bytes;
}


// express_header (Particles_gevolution.hpp:98-120), 120 bytes
struct ExpressHeader {
    char magic[8];
    uint32_t version, header_size, endian, sizeof_real, sizeof_long, sizeof_gadget2_header;
    int32_t rank, size, grid_size[2], lat_size[3], lat_size_local[3], coord_skip[2];
    uint64_t local_npart, global_npart, position_bytes, momentum_bytes, id_bytes;
};
static_assert(sizeof(ExpressHeader) == 120, "express header layout");
static const char EXPRESS_MAGIC[8] = {'G', 'E', 'X', 'P', 'C', 'D', 'M', '\0'};
constexpr uint32_t EXPRESS_VERSION = 1, EXPRESS_ENDIAN = 0x01020304u;

// express_rank_filename (Particles_gevolution.hpp:125-144) for rank 0
static std::string express_name(const char *path)
{
    std::string f(path);
    const size_t dot = f.find_last_of('.');
    if (dot != std::string::npos && dot + 1 < f.size() &&
        f.find_first_not_of("0123456789", dot + 1) == std::string::npos)
        f.resize(dot);
    return f + ".0";
}

template <typename R> static void put_triplets(std::vector<char> &buf, const std::vector<double> &a, int64_t n)
{
    // a: [3][n] -> R[n][3]
    const size_t o = buf.size();
    buf.resize(o + sizeof(R) * 3 * n);
    R *d = (R *)(buf.data() + o);
    for (int64_t j = 0; j < n; j++)
        for (int k = 0; k < 3; k++) d[3 * j + k] = (R)a[k * n + j];
}
template <typename R> static void get_triplets(const char *src, std::vector<double> &a, int64_t n)
{
    const R *s = (const R *)src;
    for (int64_t j = 0; j < n; j++)
        for (int k = 0; k < 3; k++) a[k * n + j] = (double)s[3 * j + k];
}

}  // namespace gev

using namespace gev;

extern "C" {

int gev_pcls_save_gadget2(gev_pcls *p, const char *path, const gev_gadget2_header *hdr, int tracer_factor,
                          double dtau_pos, double dtau_vel, const void *phi)
{
    if (!p || !path || !hdr) { gev_set_error("gev_pcls_save_gadget2: NULL argument"); return GEV_ERR_ARG; }
    if (tracer_factor < 1) { gev_set_error("gev_pcls_save_gadget2: tracer_factor must be >= 1"); return GEV_ERR_ARG; }
    if (hdr->num_files != 1) { gev_set_error("gev_pcls_save_gadget2: only single-file snapshots (num_files = 1)"); return GEV_ERR_ARG; }
    if (!p->sorted) { gev_set_error("gev_pcls_save_gadget2: particles are not binned"); return GEV_ERR_STATE; }
    std::vector<float> pos, vel;
    std::vector<int64_t> id;
    // tracer selection (ID multiple of the factor, Particles_gevolution.hpp:444,494) in container order
    const int rc = p->ctx->prec == GEV_FP64
                       ? gadget_convert<double>(p, tracer_factor, dtau_pos, dtau_vel, hdr->time, hdr->BoxSize,
                                                (const double *)phi, pos, vel, id)
                       : gadget_convert<float>(p, tracer_factor, dtau_pos, dtau_vel, hdr->time, hdr->BoxSize,
                                               (const float *)phi, pos, vel, id);
    if (rc) return rc;
    const int64_t m = (int64_t)id.size();
    const uint64_t declared = (uint64_t)hdr->npart[1] + ((uint64_t)hdr->npartTotalHW[1] << 32);
    if (declared != (uint64_t)m || hdr->npartTotalHW[1] != 0) {
        gev_set_error("gev_pcls_save_gadget2: header npart[1] does not match the selected particles (or exceeds 2^32-1)");
        return GEV_ERR_ARG;
    }
    if ((uint64_t)m * 12 > 0xffffffffull) { gev_set_error("gev_pcls_save_gadget2: record exceeds 4 GiB"); return GEV_ERR_ARG; }
    FILE *f = fopen(path, "wb");
    if (!f) { gev_set_error(std::string("gev_pcls_save_gadget2: cannot open ") + path); return GEV_ERR_ARG; }
    const uint32_t b3 = (uint32_t)(12 * m), b1 = (uint32_t)(8 * m);
    bool ok = gadget_put_record(f, hdr, sizeof(gev_gadget2_header)) && gadget_put_record(f, pos.data(), b3) &&
              gadget_put_record(f, vel.data(), b3) && gadget_put_record(f, id.data(), b1);
    ok = (fclose(f) == 0) && ok;
    if (!ok) { gev_set_error(std::string("gev_pcls_save_gadget2: write failed on ") + path); return GEV_ERR_ARG; }
    return GEV_OK;
}

int gev_pcls_load_gadget2(gev_pcls *p, const char *path, gev_gadget2_header *hdr_out)
{
    if (!p || !path) { gev_set_error("gev_pcls_load_gadget2: NULL argument"); return GEV_ERR_ARG; }
    FILE *f = fopen(path, "rb");
    if (!f) { gev_set_error(std::string("gev_pcls_load_gadget2: cannot open ") + path); return GEV_ERR_ARG; }
    gev_gadget2_header hdr;
    bool ok = get_record(f, &hdr, sizeof(hdr));
    const int64_t n = ok ? (int64_t)hdr.npart[1] : 0;
    if (ok && n > p->cap) {
        fclose(f);
        gev_set_error("gev_pcls_load_gadget2: the snapshot holds more particles than the container's capacity");
        return GEV_ERR_ARG;
    }
    std::vector<float> pos(3 * (size_t)n), vel(3 * (size_t)n);
    std::vector<int64_t> id((size_t)n);
    ok = ok && get_record(f, pos.data(), (uint32_t)(12 * n)) && get_record(f, vel.data(), (uint32_t)(12 * n)) &&
         get_record(f, id.data(), (uint32_t)(8 * n));
    fclose(f);
    if (!ok) { gev_set_error(std::string("gev_pcls_load_gadget2: not a single-file Gadget-2 snapshot: ") + path); return GEV_ERR_ARG; }
    // units back to the box (loadGadget2, Particles_gevolution.hpp:2563-2569)
    const double rescale = 1.0 / GADGET_VELOCITY_CONVERSION / sqrt(hdr.time);
    std::vector<double> dp(3 * (size_t)n), dq(3 * (size_t)n);
    for (size_t j = 0; j < 3 * (size_t)n; j++) {
        float x = (float)((double)pos[j] / hdr.BoxSize);  // the reference divides a float in place
        while (x < 0.0f) x += 1.0f;
        while (x >= 1.0f) x -= 1.0f;
        dp[j] = (double)x;
        dq[j] = (double)(float)((double)vel[j] * (hdr.time / rescale));
    }
    // [3][n] layout for gev_pcls_load
    std::vector<double> P(3 * (size_t)n), Q(3 * (size_t)n);
    for (int64_t j = 0; j < n; j++)
        for (int d = 0; d < 3; d++) { P[d * n + j] = dp[3 * j + d]; Q[d * n + j] = dq[3 * j + d]; }
    const int rc = gev_pcls_load(p, n, P.data(), Q.data(), id.data(), 0);
    if (rc) return rc;
    if (hdr_out) *hdr_out = hdr;
    return GEV_OK;
}

int gev_pcls_save_express(gev_pcls *p, const char *path, gev_gadget2_header *hdr)
{
    if (!p || !path || !hdr) { gev_set_error("gev_pcls_save_express: NULL argument"); return GEV_ERR_ARG; }
    if (hdr->num_files != 1) { gev_set_error("gev_pcls_save_express: one file per task, and this build runs one task"); return GEV_ERR_ARG; }
    const int64_t n = p->n;
    std::vector<double> pos(3 * (size_t)n), q(3 * (size_t)n);
    std::vector<int64_t> id((size_t)n);
    if (n > 0) {
        const int rc = gev_pcls_store(p, pos.data(), q.data(), id.data(), 0);
        if (rc) return rc;
    }
    hdr->npart[1] = (uint32_t)((uint64_t)n & 0xffffffffu);
    hdr->npartTotal[1] = hdr->npart[1];
    hdr->npartTotalHW[1] = (uint32_t)((uint64_t)n >> 32);
    const int N = p->ctx->N;
    const uint32_t rs = p->ctx->prec == GEV_FP64 ? 8 : 4;
    ExpressHeader e;
    memset(&e, 0, sizeof(e));
    memcpy(e.magic, EXPRESS_MAGIC, sizeof(e.magic));
    e.version = EXPRESS_VERSION;
    e.header_size = sizeof(e);
    e.endian = EXPRESS_ENDIAN;
    e.sizeof_real = rs;
    e.sizeof_long = sizeof(int64_t);
    e.sizeof_gadget2_header = sizeof(gev_gadget2_header);
    e.rank = 0;
    e.size = 1;
    e.grid_size[0] = e.grid_size[1] = 1;
    for (int i = 0; i < 3; i++) e.lat_size[i] = e.lat_size_local[i] = N;
    e.local_npart = e.global_npart = (uint64_t)n;
    e.position_bytes = e.momentum_bytes = 3 * (uint64_t)n * rs;
    e.id_bytes = (uint64_t)n * sizeof(int64_t);
    std::vector<char> buf((const char *)&e, (const char *)&e + sizeof(e));
    if (rs == 8) { put_triplets<double>(buf, pos, n); put_triplets<double>(buf, q, n); }
    else { put_triplets<float>(buf, pos, n); put_triplets<float>(buf, q, n); }
    buf.insert(buf.end(), (const char *)id.data(), (const char *)(id.data() + n));
    buf.insert(buf.end(), (const char *)hdr, (const char *)hdr + sizeof(gev_gadget2_header));
    const std::string name = express_name(path);
    FILE *f = fopen(name.c_str(), "wb");
    if (!f) { gev_set_error("gev_pcls_save_express: cannot open " + name); return GEV_ERR_ARG; }
    const bool ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if ((fclose(f) != 0) || !ok) { gev_set_error("gev_pcls_save_express: write failed on " + name); return GEV_ERR_ARG; }
    return GEV_OK;
}

int gev_pcls_load_express(gev_pcls *p, const char *path, gev_gadget2_header *hdr_out)
{
    if (!p || !path) { gev_set_error("gev_pcls_load_express: NULL argument"); return GEV_ERR_ARG; }
    const std::string name = express_name(path);
    FILE *f = fopen(name.c_str(), "rb");
    if (!f) { gev_set_error("gev_pcls_load_express: cannot open " + name); return GEV_ERR_ARG; }
    ExpressHeader e;
    bool ok = fread(&e, sizeof(e), 1, f) == 1;
    const int N = p->ctx->N;
    const uint32_t rs = p->ctx->prec == GEV_FP64 ? 8 : 4;
    // the layout checks of loadExpress (Particles_gevolution.hpp:2786-2808)
    ok = ok && memcmp(e.magic, EXPRESS_MAGIC, sizeof(e.magic)) == 0 && e.version == EXPRESS_VERSION &&
         e.header_size == sizeof(e) && e.endian == EXPRESS_ENDIAN && e.sizeof_real == rs &&
         e.sizeof_long == sizeof(int64_t) && e.sizeof_gadget2_header == sizeof(gev_gadget2_header) && e.rank == 0 &&
         e.size == 1 && e.grid_size[0] == 1 && e.grid_size[1] == 1 && e.coord_skip[0] == 0 && e.coord_skip[1] == 0;
    for (int i = 0; ok && i < 3; i++) ok = e.lat_size[i] == N && e.lat_size_local[i] == N;
    ok = ok && e.position_bytes == 3 * e.local_npart * rs && e.momentum_bytes == 3 * e.local_npart * rs &&
         e.id_bytes == e.local_npart * sizeof(int64_t);
    if (!ok) {
        fclose(f);
        gev_set_error("gev_pcls_load_express: " + name + " is incompatible with this lattice, precision or ABI");
        return GEV_ERR_ARG;
    }
    const int64_t n = (int64_t)e.local_npart;
    if (n > p->cap) {
        fclose(f);
        gev_set_error("gev_pcls_load_express: the file holds more particles than the container's capacity");
        return GEV_ERR_ARG;
    }
    std::vector<char> raw(2 * e.position_bytes + e.id_bytes);
    gev_gadget2_header hdr;
    ok = fread(raw.data(), 1, raw.size(), f) == raw.size() && fread(&hdr, sizeof(hdr), 1, f) == 1;
    fclose(f);
    const uint64_t hdr_global = (uint64_t)hdr.npartTotal[1] + ((uint64_t)hdr.npartTotalHW[1] << 32);
    if (!ok || hdr_global != e.global_npart || hdr.num_files != e.size) {
        gev_set_error("gev_pcls_load_express: " + name + " is truncated or its Gadget-2 metadata does not match");
        return GEV_ERR_ARG;
    }
    std::vector<double> pos(3 * (size_t)n), q(3 * (size_t)n);
    std::vector<int64_t> id((size_t)n);
    if (rs == 8) { get_triplets<double>(raw.data(), pos, n); get_triplets<double>(raw.data() + e.position_bytes, q, n); }
    else { get_triplets<float>(raw.data(), pos, n); get_triplets<float>(raw.data() + e.position_bytes, q, n); }
    memcpy(id.data(), raw.data() + 2 * e.position_bytes, e.id_bytes);
    const int rc = gev_pcls_load(p, n, pos.data(), q.data(), id.data(), 0);
    if (rc) return rc;
    if (hdr_out) *hdr_out = hdr;
    return GEV_OK;
}

}  // extern "C"

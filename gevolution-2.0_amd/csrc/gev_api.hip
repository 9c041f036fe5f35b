// gev_api.hip -- the C ABI (include/gev_pm.h): argument checking, context and
// particle-container ownership, dispatch on the precision form.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "gev_common.hpp"
#include "gev_kernels.hpp"

namespace gev {
template <typename T> int pcls_sort_impl(gev_pcls *p);
template <typename T> int pcls_lattice_impl(gev_pcls *p);
template <typename T> int pcls_load_impl(gev_pcls *p, const double *pos, const double *q, const int64_t *id);
template <typename T> int pcls_store_impl(gev_pcls *p, double *pos, double *q, int64_t *id);
template <typename T> int dep_T00_impl(gev_pcls *p, T *out, const T *phi, double mass, double a, int acc);
template <typename T>
int dep_T00_Tij_impl(gev_pcls *p, T *T00, T *Tij, const T *phi, double mass00, double massij, double a, int acc);
template <typename T> int dep_T0i_impl(gev_pcls *p, T *out, const T *phi, double mass, int acc);
template <typename T> int dep_Tij_impl(gev_pcls *p, T *out, const T *phi, double mass, double a, int acc);
template <typename T>
int dep_T00_Tij_links_impl(gev_pcls *p, T *T00, T *Tij, T *links, const T *phi, double mass00, double massij,
                           double mass0i, double a);
template <typename T>
int update_q_impl(gev_pcls *p, double dtau, const T *phi, const T *chi, const T *Bi, int nfield, double a,
                  double bcoeff, double *max_ret);
template <typename T>
int update_q_newton_impl(gev_pcls *p, double dtau, const T *psi, const T *chi, int nfield, double a, double *max_ret);
template <typename T>
int update_pos_impl(gev_pcls *p, double dtau, const T *phi, const T *chi, const T *Bi, int nfield, double a,
                    double bcoeff);
template <typename T> int update_pos_newton_impl(gev_pcls *p, double dtau, double a);
template <typename T>
int kick_drift_impl(gev_pcls *p, int gr, double dtau_k, double dtau_d, const T *phi, const T *chi, const T *Bi,
                    int nf_k, int nf_d, double a_k, double a_d, double bc_k, double bc_d, double *max_ret, int rebin);
template <typename T> int pcls_sort_counted_impl(gev_pcls *p);
template <typename T> int displace_impl(gev_pcls *p, const T *xi, double coeff, double *max_disp);
template <typename T> int initialize_q_impl(gev_pcls *p, const T *pot, double coeff, double *max_ret);
template <typename T> int dep_vec3_impl(gev_pcls *p, const T *w, T *u);
template <typename T> int adj_divergence_impl(gev_ctx *c, const T *u, T *lphi, double K);
template <typename T>
int adj_gather_impl(gev_pcls *p, const T *phi, const T *lrho, T *lq, T *lx, double K, double mN, double C_prev);
template <typename T> int adj_ic_lattice_impl(gev_ctx *c, const T *lx, const T *lq, double cx, double cq, T *lpsi);
}  // namespace gev

using namespace gev;

static thread_local std::string g_err;

void gev_set_error(const std::string &msg) { g_err = msg; }

#define CHECK(cond, msg)                      \
    do {                                      \
        if (!(cond)) {                        \
            gev_set_error(msg);               \
            return GEV_ERR_ARG;               \
        }                                     \
    } while (0)

#define PREC_DISPATCH(ctx, CALL_D, CALL_F) ((ctx)->prec == GEV_FP64 ? (CALL_D) : (CALL_F))

extern "C" {

const char *gev_last_error(void) { return g_err.c_str(); }

int gev_version(void) { return 1; }

int gev_ctx_create(int N, int precision, void *hip_stream, gev_ctx **out)
{
    CHECK(out != nullptr, "gev_ctx_create: out is NULL");
    *out = nullptr;
    CHECK(N >= 4 && N <= 1024 && (N & (N - 1)) == 0, "gev_ctx_create: N must be a power of two in [4,1024]");
    CHECK(precision == GEV_FP64 || precision == GEV_MIXED, "gev_ctx_create: unknown precision");
    gev_ctx *c = new gev_ctx();
    c->N = N;
    c->prec = precision;
    const int TX = (precision == GEV_FP64) ? 8 : 16;
    c->P = ((N / 2 + 1 + TX - 1) / TX) * TX;
    c->M = (int64_t)N * N * N;
    c->NK = (int64_t)c->P * N * N;
    GEV_HIP(hipGetDevice(&c->device));
    GEV_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    if (hip_stream) {
        c->stream = (hipStream_t)hip_stream;
    } else {
        GEV_HIP(hipStreamCreate(&c->stream));
        c->own_stream = true;
    }
    GEV_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));  // no implicit sync with the null stream
    GEV_HIP(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
    GEV_HIP(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
    // twiddles exp(-2 pi i m/N) and k-space tables (gevolution.hpp:269-274)
    std::vector<cx<double>> tw(N), ks(N);
    std::vector<cx<float>> twf(N);
    std::vector<double> gk2(N);
    for (int m = 0; m < N; m++) {
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)N;
        tw[m] = mk((double)cosl(ang), (double)-sinl(ang));
        twf[m] = mk((float)cosl(ang), (float)-sinl(ang));
        double s = 2.0 * (double)N * sin(M_PI * (double)m / (double)N);
        ks[m] = mk(s * cos(M_PI * (double)m / (double)N), -s * sin(M_PI * (double)m / (double)N));
        gk2[m] = s * s;
    }
    GEV_HIP(hipMalloc(&c->tw64, sizeof(cx<double>) * N));
    GEV_HIP(hipMalloc(&c->tw32, sizeof(cx<float>) * N));
    GEV_HIP(hipMalloc(&c->gk2, sizeof(double) * N));
    GEV_HIP(hipMalloc(&c->kshift, sizeof(cx<double>) * N));
    GEV_HIP(hipMemcpy(c->tw64, tw.data(), sizeof(cx<double>) * N, hipMemcpyHostToDevice));
    GEV_HIP(hipMemcpy(c->tw32, twf.data(), sizeof(cx<float>) * N, hipMemcpyHostToDevice));
    GEV_HIP(hipMemcpy(c->gk2, gk2.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    GEV_HIP(hipMemcpy(c->kshift, ks.data(), sizeof(cx<double>) * N, hipMemcpyHostToDevice));
    // results, then workgroup partials: the field sums' 2048 x 8 and the
    // summing x pass's two per workgroup (N^2 / NL workgroups, NL >= 4096 / N)
    const int64_t nparts = std::max<int64_t>(2048 * 8, 2 * ((int64_t)N * N * N / 4096 + 1));
    GEV_HIP(hipMalloc(&c->red, sizeof(double) * (64 + nparts)));
    GEV_HIP(hipHostMalloc(&c->red_host, sizeof(double) * 64, hipHostMallocDefault));
    c->scan_blocks_n = (c->M + 2047) / 2048 + 1;
    GEV_HIP(hipMalloc(&c->scan_blocks, sizeof(int32_t) * c->scan_blocks_n));
    *out = c;
    return GEV_OK;
}

int gev_ctx_destroy(gev_ctx *c)
{
    if (!c) return GEV_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->tw64);
    (void)hipFree(c->tw32);
    (void)hipFree(c->gk2);
    (void)hipFree(c->kshift);
    (void)hipFree(c->red);
    (void)hipHostFree(c->red_host);
    (void)hipFree(c->scan_blocks);
    for (int i = 0; i < gev_ctx::MAX_MARKS; i++)
        if (c->marks[i]) (void)hipEventDestroy(c->marks[i]);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GEV_OK;
}

int gev_ctx_sync(gev_ctx *c)
{
    CHECK(c, "gev_ctx_sync: NULL context");
    GEV_HIP(hipStreamSynchronize(c->stream));
    return GEV_OK;
}

int gev_ctx_kpitch(const gev_ctx *c) { return c ? c->P : 0; }

// ------------------------------------------------------------ particles ---
int gev_pcls_create(gev_ctx *c, int64_t capacity, double mass, gev_pcls **out)
{
    CHECK(c && out, "gev_pcls_create: NULL argument");
    CHECK(capacity >= 0 && capacity < (int64_t)INT32_MAX, "gev_pcls_create: capacity out of range");
    *out = nullptr;
    gev_pcls *p = new gev_pcls();
    p->ctx = c;
    p->cap = capacity;
    p->mass = mass;
    const size_t es = (c->prec == GEV_FP64) ? 8 : 4;
    // one slot past the capacity: the sink that the particle kernels read
    // and write for idle lanes, so their memory instructions need no branch
    const size_t cap = (size_t)std::max<int64_t>(capacity, 1) + 1;
    for (int d = 0; d < 3; d++) {
        GEV_HIP(hipMalloc(&p->pos[d], es * cap));
        GEV_HIP(hipMalloc(&p->q[d], es * cap));
        GEV_HIP(hipMalloc(&p->pos2[d], es * cap));
        GEV_HIP(hipMalloc(&p->q2[d], es * cap));
    }
    GEV_HIP(hipMalloc(&p->id, 8 * cap));
    GEV_HIP(hipMalloc(&p->id2, 8 * cap));
    GEV_HIP(hipMalloc(&p->cell, 4 * cap));
    GEV_HIP(hipMalloc(&p->cell2, 4 * cap));
    GEV_HIP(hipMalloc(&p->rank, 4 * cap));
    GEV_HIP(hipMalloc(&p->perm, 4 * cap));
    GEV_HIP(hipMalloc(&p->cell_start, 4 * (size_t)(c->M + 1)));
    GEV_HIP(hipMalloc(&p->cell_count, 4 * (size_t)c->M));
    if (c->prec == GEV_MIXED && c->N >= 16) {
        const size_t nb = (size_t)(c->N >> QB_LG);
        GEV_HIP(hipMalloc(&p->qbound, 4 * nb * nb * nb));
    }
    *out = p;
    return GEV_OK;
}

int gev_pcls_destroy(gev_pcls *p)
{
    if (!p) return GEV_OK;
    (void)hipStreamSynchronize(p->ctx->stream);
    for (int d = 0; d < 3; d++) {
        (void)hipFree(p->pos[d]); (void)hipFree(p->q[d]);
        (void)hipFree(p->pos2[d]); (void)hipFree(p->q2[d]);
    }
    (void)hipFree(p->id); (void)hipFree(p->id2);
    (void)hipFree(p->cell); (void)hipFree(p->cell2);
    (void)hipFree(p->rank);
    (void)hipFree(p->perm);
    (void)hipFree(p->cell_start);
    (void)hipFree(p->cell_count);
    (void)hipFree(p->qbound);
    delete p;
    return GEV_OK;
}

int gev_pcls_load(gev_pcls *p, int64_t n, const double *pos, const double *q, const int64_t *id, int on_device)
{
    CHECK(p && pos && q && id, "gev_pcls_load: NULL argument");
    CHECK(n >= 0 && n <= p->cap, "gev_pcls_load: n exceeds the container capacity");
    gev_ctx *c = p->ctx;
    p->n = n;
    p->sorted = false;
    p->qbound_valid = false;
    if (n == 0) return PREC_DISPATCH(c, pcls_sort_impl<double>(p), pcls_sort_impl<float>(p));
    const double *dpos = pos, *dq = q;
    const int64_t *did = id;
    // device staging for the host path, freed on every return (hipFree waits
    // for the work queued on it)
    struct Staging {
        double *p = nullptr;
        ~Staging() { if (p) (void)hipFree(p); }
    } tmp;
    if (!on_device) {
        GEV_HIP(hipMalloc(&tmp.p, sizeof(double) * 7 * n));
        GEV_HIP(hipMemcpyAsync(tmp.p, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
        GEV_HIP(hipMemcpyAsync(tmp.p + 3 * n, q, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
        GEV_HIP(hipMemcpyAsync(tmp.p + 6 * n, id, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->stream));
        dpos = tmp.p; dq = tmp.p + 3 * n; did = (const int64_t *)(tmp.p + 6 * n);
    }
    int rc = PREC_DISPATCH(c, pcls_load_impl<double>(p, dpos, dq, did), pcls_load_impl<float>(p, dpos, dq, did));
    if (tmp.p && rc == GEV_OK) GEV_HIP(hipStreamSynchronize(c->stream));
    return rc;
}

int gev_pcls_store(gev_pcls *p, double *pos, double *q, int64_t *id, int on_device)
{
    CHECK(p && pos && q && id, "gev_pcls_store: NULL argument");
    gev_ctx *c = p->ctx;
    const int64_t n = p->n;
    if (n == 0) return GEV_OK;
    double *dpos = pos, *dq = q;
    int64_t *did = id;
    // host staging for the host path, freed on every return (hipFree waits for
    // the work queued on it)
    struct Staging {
        double *p = nullptr;
        ~Staging() { if (p) (void)hipFree(p); }
    } tmp;
    if (!on_device) {
        GEV_HIP(hipMalloc(&tmp.p, sizeof(double) * 7 * n));
        dpos = tmp.p; dq = tmp.p + 3 * n; did = (int64_t *)(tmp.p + 6 * n);
    }
    int rc = PREC_DISPATCH(c, pcls_store_impl<double>(p, dpos, dq, did), pcls_store_impl<float>(p, dpos, dq, did));
    if (rc) return rc;
    if (tmp.p) {
        GEV_HIP(hipMemcpyAsync(pos, dpos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        GEV_HIP(hipMemcpyAsync(q, dq, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        GEV_HIP(hipMemcpyAsync(id, did, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
        GEV_HIP(hipStreamSynchronize(c->stream));
    }
    return GEV_OK;
}

int gev_pcls_view_get(gev_pcls *p, gev_pcls_view *v)
{
    CHECK(p && v, "gev_pcls_view_get: NULL argument");
    v->n = p->n;
    v->precision = p->ctx->prec;
    for (int d = 0; d < 3; d++) { v->pos[d] = p->pos[d]; v->q[d] = p->q[d]; }
    v->id = p->id;
    v->cell = p->cell;
    v->cell_start = p->cell_start;
    return GEV_OK;
}

int gev_pcls_copy_index(gev_pcls *p, int32_t *cell, int32_t *cell_start)
{
    CHECK(p, "gev_pcls_copy_index: NULL argument");
    if (cell && p->n > 0)
        GEV_HIP(hipMemcpyAsync(cell, p->cell, sizeof(int32_t) * p->n, hipMemcpyDeviceToDevice, p->ctx->stream));
    if (cell_start)
        GEV_HIP(hipMemcpyAsync(cell_start, p->cell_start, sizeof(int32_t) * (p->ctx->M + 1), hipMemcpyDeviceToDevice,
                               p->ctx->stream));
    return GEV_OK;
}

int gev_pcls_sort(gev_pcls *p)
{
    CHECK(p, "gev_pcls_sort: NULL argument");
    p->qbound_valid = false;  // cells are recomputed from the positions, which a caller may have written
    return PREC_DISPATCH(p->ctx, pcls_sort_impl<double>(p), pcls_sort_impl<float>(p));
}

int gev_pcls_init_lattice(gev_pcls *p)
{
    CHECK(p, "gev_pcls_init_lattice: NULL argument");
    CHECK(p->cap >= p->ctx->M, "gev_pcls_init_lattice: capacity below N^3");
    p->n = p->ctx->M;
    p->qbound_valid = false;
    return PREC_DISPATCH(p->ctx, pcls_lattice_impl<double>(p), pcls_lattice_impl<float>(p));
}

#define NEED_SORTED(p) CHECK((p)->sorted, "particles are not binned (call gev_pcls_sort)")

// ------------------------------------------------------------ deposits ----
int gev_project_T00(gev_pcls *p, void *T00, double a, const void *phi, double coeff, int accumulate)
{
    CHECK(p && T00, "gev_project_T00: NULL argument");
    NEED_SORTED(p);
    const double dx = 1.0 / (double)p->ctx->N;
    double mass = p->mass * coeff / (dx * dx * dx);
    if (phi) mass /= a;
    return PREC_DISPATCH(p->ctx, dep_T00_impl<double>(p, (double *)T00, (const double *)phi, mass, a, accumulate),
                         dep_T00_impl<float>(p, (float *)T00, (const float *)phi, mass, a, accumulate));
}

int gev_project_mass(gev_pcls *p, void *rho, double coeff, int accumulate)
{
    return gev_project_T00(p, rho, 1.0, nullptr, coeff, accumulate);
}

int gev_project_T0i(gev_pcls *p, void *T0i, const void *phi, double coeff, int accumulate)
{
    CHECK(p && T0i, "gev_project_T0i: NULL argument");
    NEED_SORTED(p);
    const double dx = 1.0 / (double)p->ctx->N;
    double mass = p->mass * coeff / (dx * dx * dx);
    return PREC_DISPATCH(p->ctx, dep_T0i_impl<double>(p, (double *)T0i, (const double *)phi, mass, accumulate),
                         dep_T0i_impl<float>(p, (float *)T0i, (const float *)phi, mass, accumulate));
}

int gev_project_Tij(gev_pcls *p, void *Tij, double a, const void *phi, double coeff, int accumulate)
{
    CHECK(p && Tij, "gev_project_Tij: NULL argument");
    NEED_SORTED(p);
    const double dx = 1.0 / (double)p->ctx->N;
    double mass = p->mass * coeff / (dx * dx * dx) / a;
    return PREC_DISPATCH(p->ctx, dep_Tij_impl<double>(p, (double *)Tij, (const double *)phi, mass, a, accumulate),
                         dep_Tij_impl<float>(p, (float *)Tij, (const float *)phi, mass, a, accumulate));
}

int gev_project_T00_Tij(gev_pcls *p, void *T00, void *Tij, double a, const void *phi, double coeff, int accumulate)
{
    CHECK(p && T00 && Tij, "gev_project_T00_Tij: NULL argument");
    NEED_SORTED(p);
    const double dx = 1.0 / (double)p->ctx->N;
    const double m = p->mass * coeff / (dx * dx * dx);
    const double m00 = phi ? m / a : m, mij = m / a;  // as gev_project_T00 / gev_project_Tij
    return PREC_DISPATCH(p->ctx,
                         dep_T00_Tij_impl<double>(p, (double *)T00, (double *)Tij, (const double *)phi, m00, mij, a,
                                                  accumulate),
                         dep_T00_Tij_impl<float>(p, (float *)T00, (float *)Tij, (const float *)phi, m00, mij, a,
                                                 accumulate));
}

int gev_project_T00_Tij_T0i_links(gev_pcls *p, void *T00, void *Tij, void *links, double a, const void *phi,
                                  double coeff)
{
    CHECK(p && T00 && Tij && links, "gev_project_T00_Tij_T0i_links: NULL argument");
    NEED_SORTED(p);
    const double dx = 1.0 / (double)p->ctx->N;
    const double m = p->mass * coeff / (dx * dx * dx);
    const double m00 = phi ? m / a : m, mij = m / a;  // as gev_project_T00 / gev_project_Tij
    return PREC_DISPATCH(p->ctx,
                         dep_T00_Tij_links_impl<double>(p, (double *)T00, (double *)Tij, (double *)links,
                                                        (const double *)phi, m00, mij, m, a),
                         dep_T00_Tij_links_impl<float>(p, (float *)T00, (float *)Tij, (float *)links,
                                                       (const float *)phi, m00, mij, m, a));
}

// -------------------------------------------------------------- sources ---
int gev_prepare_source_T00(gev_ctx *c, const void *phi, const void *chi, const void *source, double bg, void *result,
                           double c1, double c2, double c3, double *sum_source, double *sum_result)
{
    CHECK(c && phi && chi && source && result, "gev_prepare_source_T00: NULL argument");
    double sums[2] = {0, 0};
    int rc = PREC_DISPATCH(c,
                           prepare_T00_impl<double>(c, (const double *)phi, (const double *)chi,
                                                    (const double *)source, bg, (double *)result, c1, c2, c3, sums),
                           prepare_T00_impl<float>(c, (const float *)phi, (const float *)chi, (const float *)source,
                                                   bg, (float *)result, c1, c2, c3, sums));
    if (sum_source) *sum_source = sums[0];
    if (sum_result) *sum_result = sums[1];
    return rc;
}

int gev_prepare_source_Tij(gev_ctx *c, const void *phi, const void *Tij, void *Sij, double coeff)
{
    CHECK(c && phi && Tij && Sij, "gev_prepare_source_Tij: NULL argument");
    return PREC_DISPATCH(c, prepare_Tij_impl<double>(c, (const double *)phi, (const double *)Tij, (double *)Sij, coeff),
                         prepare_Tij_impl<float>(c, (const float *)phi, (const float *)Tij, (float *)Sij, coeff));
}

// ------------------------------------------------------------------ FFT ---
int gev_fft_r2c(gev_ctx *c, const void *in, void *out, int ncomp)
{
    CHECK(c && in && out && ncomp >= 1, "gev_fft_r2c: bad argument");
    return PREC_DISPATCH(c, fft_r2c_impl<double>(c, (const double *)in, (cx<double> *)out, ncomp, c->tw64),
                         fft_r2c_impl<float>(c, (const float *)in, (cx<float> *)out, ncomp, c->tw32));
}

int gev_fft_c2r(gev_ctx *c, void *in, void *out, int ncomp)
{
    CHECK(c && in && out && ncomp >= 1, "gev_fft_c2r: bad argument");
    return PREC_DISPATCH(c, fft_c2r_impl<double>(c, (cx<double> *)in, (double *)out, ncomp, c->tw64),
                         fft_c2r_impl<float>(c, (cx<float> *)in, (float *)out, ncomp, c->tw32));
}

// -------------------------------------------------------------- k-space ---
int gev_poisson(gev_ctx *c, const void *src, void *pot, double coeff, double modif)
{
    CHECK(c && src && pot, "gev_poisson: NULL argument");
    return PREC_DISPATCH(c, poisson_impl<double>(c, (const cx<double> *)src, (cx<double> *)pot, coeff, modif),
                         poisson_impl<float>(c, (const cx<float> *)src, (cx<float> *)pot, coeff, modif));
}

int gev_project_ft_scalar(gev_ctx *c, const void *S, void *chi, int add)
{
    CHECK(c && S && chi, "gev_project_ft_scalar: NULL argument");
    return PREC_DISPATCH(c, project_ft_scalar_impl<double>(c, (const cx<double> *)S, (cx<double> *)chi, add),
                         project_ft_scalar_impl<float>(c, (const cx<float> *)S, (cx<float> *)chi, add));
}

int gev_project_ft_vector(gev_ctx *c, const void *S, void *B, double coeff, double modif)
{
    CHECK(c && S && B, "gev_project_ft_vector: NULL argument");
    return PREC_DISPATCH(c, project_ft_vector_impl<double>(c, (const cx<double> *)S, (cx<double> *)B, coeff, modif),
                         project_ft_vector_impl<float>(c, (const cx<float> *)S, (cx<float> *)B, coeff, modif));
}

int gev_evolve_ft_vector(gev_ctx *c, const void *S, void *B, double a2dtau)
{
    CHECK(c && S && B, "gev_evolve_ft_vector: NULL argument");
    return PREC_DISPATCH(c, evolve_ft_vector_impl<double>(c, (const cx<double> *)S, (cx<double> *)B, a2dtau),
                         evolve_ft_vector_impl<float>(c, (const cx<float> *)S, (cx<float> *)B, a2dtau));
}

// ---- container checkpoints (device memory of gev_pcls_state_bytes bytes):
// a 256-byte header holding the count, then positions, momenta, IDs, cells
// and cell starts, each at a 256-byte aligned offset, copied by one kernel
// with 16-byte accesses (a save is one launch and no host synchronisation)
namespace gev {
struct CopySegs {
    char *dst[9];
    const char *src[9];
    int64_t bytes[9];  // multiples of 4
    int nseg;
    int64_t count;     // written to *count_dst when that is set (a save)
    int64_t *count_dst;
};

__global__ __launch_bounds__(256) void k_copy_segs(CopySegs L)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (int64_t)gridDim.x * blockDim.x;
    if (L.count_dst && t == 0) *L.count_dst = L.count;
    for (int g = 0; g < L.nseg; g++) {
        const int64_t nb = L.bytes[g], nv = nb >> 4;
        const uint4 *src = (const uint4 *)L.src[g];
        uint4 *dst = (uint4 *)L.dst[g];
        int64_t i = t;
        for (; i + 3 * st < nv; i += 4 * st) {
            const uint4 a = src[i], b = src[i + st], c = src[i + 2 * st], d = src[i + 3 * st];
            dst[i] = a;
            dst[i + st] = b;
            dst[i + 2 * st] = c;
            dst[i + 3 * st] = d;
        }
        for (; i < nv; i += st) dst[i] = src[i];
        const int64_t tail = (nb - (nv << 4)) >> 2;  // up to three 4-byte words
        if (t < tail) ((uint32_t *)(L.dst[g] + (nv << 4)))[t] = ((const uint32_t *)(L.src[g] + (nv << 4)))[t];
    }
}
}  // namespace gev

static int64_t ckpt_round(int64_t b) { return (b + 255) & ~(int64_t)255; }

// the container arrays and their offsets in a checkpoint of n particles
static int pcls_state_segs(gev_pcls *p, int64_t n, char *blob, bool save, gev::CopySegs &L)
{
    const int64_t es = (p->ctx->prec == GEV_FP64) ? 8 : 4;
    char *arr[9] = {(char *)p->pos[0], (char *)p->pos[1], (char *)p->pos[2], (char *)p->q[0], (char *)p->q[1],
                    (char *)p->q[2], (char *)p->id, (char *)p->cell, (char *)p->cell_start};
    const int64_t nb[9] = {n * es, n * es, n * es, n * es, n * es, n * es, 8 * n, 4 * n, 4 * (p->ctx->M + 1)};
    int64_t off = 256;
    L.nseg = 9;
    for (int g = 0; g < 9; g++) {
        L.dst[g] = save ? blob + off : arr[g];
        L.src[g] = save ? arr[g] : blob + off;
        L.bytes[g] = nb[g];
        if (((uintptr_t)L.dst[g] | (uintptr_t)L.src[g]) & 15) return GEV_ERR_ARG;
        off += ckpt_round(nb[g]);
    }
    return GEV_OK;
}

static int64_t pcls_state_bytes(const gev_pcls *p, int64_t n)
{
    const int64_t es = (p->ctx->prec == GEV_FP64) ? 8 : 4;
    return 256 + 6 * ckpt_round(n * es) + ckpt_round(8 * n) + ckpt_round(4 * n) + ckpt_round(4 * (p->ctx->M + 1));
}

static int launch_copy_segs(gev_pcls *p, const gev::CopySegs &L)
{
    int64_t most = 0;
    for (int g = 0; g < L.nseg; g++) most = L.bytes[g] > most ? L.bytes[g] : most;
    // four 16-byte vectors per thread and pass, at most eight workgroups per CU
    int64_t nbk = (most / 16 + 4 * 256 - 1) / (4 * 256);
    nbk = nbk < 1 ? 1 : (nbk > 2048 ? 2048 : nbk);
    hipLaunchKernelGGL(gev::k_copy_segs, dim3((unsigned)nbk), dim3(256), 0, p->ctx->stream, L);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

int gev_pcls_state_bytes(gev_pcls *p, int64_t *bytes)
{
    CHECK(p && bytes, "gev_pcls_state_bytes: NULL argument");
    *bytes = pcls_state_bytes(p, p->n);
    return GEV_OK;
}

int gev_pcls_save(gev_pcls *p, void *dst, int64_t bytes)
{
    CHECK(p && dst, "gev_pcls_save: NULL argument");
    NEED_SORTED(p);
    CHECK(bytes >= pcls_state_bytes(p, p->n), "gev_pcls_save: buffer smaller than gev_pcls_state_bytes");
    gev::CopySegs L{};
    CHECK(pcls_state_segs(p, p->n, (char *)dst, true, L) == GEV_OK, "gev_pcls_save: buffer not 16-byte aligned");
    L.count = p->n;
    L.count_dst = (int64_t *)dst;
    return launch_copy_segs(p, L);
}

int gev_pcls_restore(gev_pcls *p, const void *src, int64_t bytes)
{
    CHECK(p && src, "gev_pcls_restore: NULL argument");
    CHECK(bytes >= 256, "gev_pcls_restore: buffer smaller than a checkpoint header");
    p->qbound_valid = false;
    hipStream_t s = p->ctx->stream;
    int64_t n = 0;
    GEV_HIP(hipMemcpyAsync(&n, src, 8, hipMemcpyDeviceToHost, s));
    GEV_HIP(hipStreamSynchronize(s));
    CHECK(n >= 0 && n <= p->cap, "gev_pcls_restore: saved state does not fit this container");
    CHECK(bytes >= pcls_state_bytes(p, n), "gev_pcls_restore: buffer smaller than the saved state");
    gev::CopySegs L{};
    CHECK(pcls_state_segs(p, n, (char *)src, false, L) == GEV_OK, "gev_pcls_restore: buffer not 16-byte aligned");
    const int rc = launch_copy_segs(p, L);
    if (rc) return rc;
    p->n = n;
    p->sorted = true;
    return GEV_OK;
}

// ---- adjoint of the Newtonian cycle (gev_adjoint.hip)
int gev_adj_deposit(gev_pcls *p, const void *lq, void *u)
{
    CHECK(p && lq && u, "gev_adj_deposit: NULL argument");
    NEED_SORTED(p);
    return PREC_DISPATCH(p->ctx, dep_vec3_impl<double>(p, (const double *)lq, (double *)u),
                         dep_vec3_impl<float>(p, (const float *)lq, (float *)u));
}

int gev_adj_divergence(gev_ctx *c, const void *u, void *lphi, double K)
{
    CHECK(c && u && lphi, "gev_adj_divergence: NULL argument");
    return PREC_DISPATCH(c, adj_divergence_impl<double>(c, (const double *)u, (double *)lphi, K),
                         adj_divergence_impl<float>(c, (const float *)u, (float *)lphi, K));
}

int gev_adj_gather(gev_pcls *p, const void *phi, const void *lrho, void *lq, void *lx, double K, double mN,
                   double C_prev)
{
    CHECK(p && lx && lq && (lrho || phi), "gev_adj_gather: NULL argument");
    NEED_SORTED(p);
    return PREC_DISPATCH(p->ctx,
                         adj_gather_impl<double>(p, (const double *)phi, (const double *)lrho, (double *)lq,
                                                 (double *)lx, K, mN, C_prev),
                         adj_gather_impl<float>(p, (const float *)phi, (const float *)lrho, (float *)lq, (float *)lx,
                                                K, mN, C_prev));
}

int gev_adj_ic_lattice(gev_ctx *c, const void *lx, const void *lq, double cx, double cq, void *lpsi)
{
    CHECK(c && lx && lq && lpsi, "gev_adj_ic_lattice: NULL argument");
    return PREC_DISPATCH(c, adj_ic_lattice_impl<double>(c, (const double *)lx, (const double *)lq, cx, cq, (double *)lpsi),
                         adj_ic_lattice_impl<float>(c, (const float *)lx, (const float *)lq, cx, cq, (float *)lpsi));
}

// ---- adjoint of the weak-field GR cycle (gev_adjoint_gr.hip)
int gev_adj_gr_kick_drift(gev_pcls *p, const void *phi, const void *chi, const void *Bi, void *lx, void *lq,
                          void *lphi, void *lchi, void *lB, int nfield_kick, int nfield_drift, double a_kick,
                          double a_drift, double dtau_kick, double dtau_drift)
{
    CHECK(p && phi && chi && Bi && lx && lq && lphi && lchi && lB, "gev_adj_gr_kick_drift: NULL argument");
    CHECK(nfield_kick >= 1 && nfield_kick <= 3 && nfield_drift >= 0 && nfield_drift <= 3,
          "gev_adj_gr_kick_drift: nfield out of range");
    NEED_SORTED(p);
    return PREC_DISPATCH(p->ctx,
                         adj_gr_kick_drift_impl<double>(p, (const double *)phi, (const double *)chi,
                                                        (const double *)Bi, (double *)lx, (double *)lq,
                                                        (double *)lphi, (double *)lchi, (double *)lB, nfield_kick,
                                                        nfield_drift, a_kick, a_drift, dtau_kick, dtau_drift),
                         adj_gr_kick_drift_impl<float>(p, (const float *)phi, (const float *)chi, (const float *)Bi,
                                                       (float *)lx, (float *)lq, (float *)lphi, (float *)lchi,
                                                       (float *)lB, nfield_kick, nfield_drift, a_kick, a_drift,
                                                       dtau_kick, dtau_drift));
}

int gev_adj_gr_deposit(gev_pcls *p, const void *phi, const void *lT00, const void *lTij, const void *lL, void *lx,
                       void *lq, void *lphi, double a, double tij_scale)
{
    CHECK(p && lT00 && lTij && lx && lq && (!phi || lphi), "gev_adj_gr_deposit: NULL argument");
    NEED_SORTED(p);
    return PREC_DISPATCH(p->ctx,
                         adj_gr_deposit_impl<double>(p, (const double *)phi, (const double *)lT00,
                                                     (const double *)lTij, (const double *)lL, (double *)lx,
                                                     (double *)lq, (double *)lphi, a, tij_scale),
                         adj_gr_deposit_impl<float>(p, (const float *)phi, (const float *)lT00, (const float *)lTij,
                                                    (const float *)lL, (float *)lx, (float *)lq, (float *)lphi, a,
                                                    tij_scale));
}

int gev_adj_source_T00(gev_ctx *c, const void *phi, const void *T00, const void *lsrc, void *lT00, void *lphi,
                       void *lchi, double bg, double c1, double c2, double c3)
{
    CHECK(c && phi && T00 && lsrc && lT00 && lphi && lchi, "gev_adj_source_T00: NULL argument");
    CHECK(lT00 != lsrc, "gev_adj_source_T00: lT00 must not alias lsrc");
    return PREC_DISPATCH(c,
                         adj_source_T00_impl<double>(c, (const double *)phi, (const double *)T00, (const double *)lsrc,
                                                     (double *)lT00, (double *)lphi, (double *)lchi, bg, c1, c2, c3),
                         adj_source_T00_impl<float>(c, (const float *)phi, (const float *)T00, (const float *)lsrc,
                                                    (float *)lT00, (float *)lphi, (float *)lchi, bg, c1, c2, c3));
}

int gev_adj_source_Tij(gev_ctx *c, const void *phi, const void *lS, void *lphi)
{
    CHECK(c && phi && lS && lphi, "gev_adj_source_Tij: NULL argument");
    return PREC_DISPATCH(c, adj_source_Tij_phi_impl<double>(c, (const double *)phi, (const double *)lS, (double *)lphi),
                         adj_source_Tij_phi_impl<float>(c, (const float *)phi, (const float *)lS, (float *)lphi));
}

int gev_adj_link_factor(gev_ctx *c, const void *phi, const void *L, void *u, void *lphi)
{
    CHECK(c && phi && L && u && lphi, "gev_adj_link_factor: NULL argument");
    return PREC_DISPATCH(c,
                         adj_link_factor_impl<double>(c, (const double *)phi, (const double *)L, (double *)u,
                                                      (double *)lphi),
                         adj_link_factor_impl<float>(c, (const float *)phi, (const float *)L, (float *)u,
                                                     (float *)lphi));
}

int gev_adj_project_ft_scalar(gev_ctx *c, const void *lchiFT, void *lSijFT)
{
    CHECK(c && lchiFT && lSijFT, "gev_adj_project_ft_scalar: NULL argument");
    return PREC_DISPATCH(c, adj_project_ft_scalar_impl<double>(c, (const cx<double> *)lchiFT, (cx<double> *)lSijFT),
                         adj_project_ft_scalar_impl<float>(c, (const cx<float> *)lchiFT, (cx<float> *)lSijFT));
}

// ---- fused real -> k-space -> real solves (the operator runs inside the z pass)
int gev_solve_poisson(gev_ctx *c, const void *src, void *pot, void *workFT, double coeff, double modif, int set_k0,
                      double k0_re)
{
    CHECK(c && src && pot && workFT, "gev_solve_poisson: NULL argument");
    if (zsolve_supported(c))
        return PREC_DISPATCH(c,
                             solve_poisson_fused_impl<double>(c, (const double *)src, (double *)pot,
                                                              (cx<double> *)workFT, coeff, modif, set_k0, k0_re),
                             solve_poisson_fused_impl<float>(c, (const float *)src, (float *)pot, (cx<float> *)workFT,
                                                             coeff, modif, set_k0, k0_re));
    int rc = gev_fft_r2c(c, src, workFT, 1);
    if (!rc && set_k0) rc = gev_set_k0(c, workFT, k0_re, 0.0);
    if (!rc) rc = gev_poisson(c, workFT, workFT, coeff, modif);
    if (!rc) rc = gev_fft_c2r(c, workFT, pot, 1);
    return rc;
}

int gev_solve_chi(gev_ctx *c, const void *Sij, void *chi, void *SijFT, void *chiFT)
{
    CHECK(c && Sij && chi && SijFT && chiFT, "gev_solve_chi: NULL argument");
    if (zsolve_supported(c))
        return PREC_DISPATCH(c,
                             solve_chi_fused_impl<double>(c, (const double *)Sij, (double *)chi, (cx<double> *)SijFT,
                                                          (cx<double> *)chiFT),
                             solve_chi_fused_impl<float>(c, (const float *)Sij, (float *)chi, (cx<float> *)SijFT,
                                                         (cx<float> *)chiFT));
    int rc = gev_fft_r2c(c, Sij, SijFT, 6);
    if (!rc) rc = gev_project_ft_scalar(c, SijFT, chiFT, 0);
    if (!rc) rc = gev_fft_c2r(c, chiFT, chi, 1);
    return rc;
}

int gev_solve_chi_Tij(gev_ctx *c, const void *phi, void *Tij, double coeff, void *chi, void *SijFT, void *chiFT)
{
    CHECK(c && phi && Tij && chi && SijFT && chiFT, "gev_solve_chi_Tij: NULL argument");
    if (chi_from_Tij_supported(c))
        return PREC_DISPATCH(c,
                             solve_chi_Tij_fused_impl<double>(c, (const double *)phi, (const double *)Tij, coeff,
                                                              (double *)chi, (cx<double> *)SijFT, (cx<double> *)chiFT),
                             solve_chi_Tij_fused_impl<float>(c, (const float *)phi, (const float *)Tij, coeff,
                                                             (float *)chi, (cx<float> *)SijFT, (cx<float> *)chiFT));
    // small meshes: Sij in place over Tij, then the chi solve
    int rc = gev_prepare_source_Tij(c, phi, Tij, Tij, coeff);
    if (!rc) rc = gev_solve_chi(c, Tij, chi, SijFT, chiFT);
    return rc;
}

int gev_solve_vector(gev_ctx *c, void *Bi, void *BiFT, double coeff, double modif)
{
    CHECK(c && Bi && BiFT, "gev_solve_vector: NULL argument");
    if (zsolve_supported(c))
        return PREC_DISPATCH(c, solve_vector_fused_impl<double>(c, (double *)Bi, (cx<double> *)BiFT, coeff, modif),
                             solve_vector_fused_impl<float>(c, (float *)Bi, (cx<float> *)BiFT, coeff, modif));
    int rc = gev_fft_r2c(c, Bi, BiFT, 3);
    if (!rc) rc = gev_project_ft_vector(c, BiFT, BiFT, coeff, modif);
    if (!rc) rc = gev_fft_c2r(c, BiFT, Bi, 3);
    return rc;
}

int gev_solve_poisson_T00(gev_ctx *c, const void *phi, const void *chi, void *source, double bg, double c1,
                          double c2, double c3, void *pot, void *work, double coeff, double modif, int k0set,
                          double *sums)
{
    CHECK(c && phi && chi && source && pot && work, "gev_solve_poisson_T00: NULL argument");
    if (chi_from_Tij_supported(c)) {
        int rc = PREC_DISPATCH(c,
                               solve_poisson_T00_fused_impl<double>(c, (const double *)phi, (const double *)chi,
                                                                    (const double *)source, bg, c1, c2, c3,
                                                                    (double *)pot, (cx<double> *)work, coeff, modif,
                                                                    k0set),
                               solve_poisson_T00_fused_impl<float>(c, (const float *)phi, (const float *)chi,
                                                                   (const float *)source, bg, c1, c2, c3,
                                                                   (float *)pot, (cx<float> *)work, coeff, modif,
                                                                   k0set));
        if (rc || !sums) return rc;
        return gev_ctx_sums(c, sums);
    }
    // small meshes: the source pass in place, then the solve
    double s2[2] = {0.0, 0.0};
    int rc = gev_prepare_source_T00(c, phi, chi, source, bg, source, c1, c2, c3, &s2[0], &s2[1]);
    if (!rc) rc = gev_solve_poisson(c, source, pot, work, coeff, modif, k0set, s2[1]);
    if (!rc && sums) { sums[0] = s2[0]; sums[1] = s2[1]; }
    return rc;
}

int gev_ctx_sums(gev_ctx *c, double sums[2])
{
    CHECK(c && sums, "gev_ctx_sums: NULL argument");
    GEV_HIP(hipMemcpyAsync(c->red_host, c->red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GEV_HIP(hipStreamSynchronize(c->stream));
    sums[0] = c->red_host[0];
    sums[1] = c->red_host[1];
    return GEV_OK;
}

int gev_solve_vector_T0i_links(gev_ctx *c, void *Bi, const void *phi, void *BiFT, double coeff, double modif)
{
    CHECK(c && Bi && phi && BiFT, "gev_solve_vector_T0i_links: NULL argument");
    if (chi_from_Tij_supported(c))
        return PREC_DISPATCH(c,
                             solve_vector_links_fused_impl<double>(c, (double *)Bi, (const double *)phi,
                                                                   (cx<double> *)BiFT, coeff, modif),
                             solve_vector_links_fused_impl<float>(c, (float *)Bi, (const float *)phi,
                                                                  (cx<float> *)BiFT, coeff, modif));
    int rc = PREC_DISPATCH(c, t0i_link_factor_impl<double>(c, (double *)Bi, (const double *)phi),
                           t0i_link_factor_impl<float>(c, (float *)Bi, (const float *)phi));
    if (!rc) rc = gev_solve_vector(c, Bi, BiFT, coeff, modif);
    return rc;
}

int gev_set_k0(gev_ctx *c, void *ft, double re, double im)
{
    CHECK(c && ft, "gev_set_k0: NULL argument");
    return PREC_DISPATCH(c, set_k0_impl<double>(c, (cx<double> *)ft, re, im),
                         set_k0_impl<float>(c, (cx<float> *)ft, re, im));
}

// ---------------------------------------------------------- kick / drift --
int gev_update_q(gev_pcls *p, double dtau, const void *phi, const void *chi, const void *Bi, int nfield, double a,
                 double bcoeff, double *max_ret)
{
    CHECK(p && phi, "gev_update_q: NULL argument");
    CHECK(nfield < 2 || chi, "gev_update_q: nfield >= 2 needs chi");
    CHECK(nfield < 3 || Bi, "gev_update_q: nfield >= 3 needs Bi");
    NEED_SORTED(p);
    return PREC_DISPATCH(p->ctx,
                         update_q_impl<double>(p, dtau, (const double *)phi, (const double *)chi, (const double *)Bi,
                                               nfield, a, bcoeff, max_ret),
                         update_q_impl<float>(p, dtau, (const float *)phi, (const float *)chi, (const float *)Bi,
                                              nfield, a, bcoeff, max_ret));
}

int gev_update_q_newton(gev_pcls *p, double dtau, const void *psi, const void *chi, int nfield, double a,
                        double *max_ret)
{
    CHECK(p && psi, "gev_update_q_newton: NULL argument");
    CHECK(nfield < 2 || chi, "gev_update_q_newton: nfield >= 2 needs chi");
    NEED_SORTED(p);
    return PREC_DISPATCH(
        p->ctx, update_q_newton_impl<double>(p, dtau, (const double *)psi, (const double *)chi, nfield, a, max_ret),
        update_q_newton_impl<float>(p, dtau, (const float *)psi, (const float *)chi, nfield, a, max_ret));
}

int gev_update_pos(gev_pcls *p, double dtau, const void *phi, const void *chi, const void *Bi, int nfield, double a,
                   double bcoeff)
{
    CHECK(p, "gev_update_pos: NULL argument");
    CHECK(nfield < 1 || phi, "gev_update_pos: nfield >= 1 needs phi");
    CHECK(nfield < 2 || chi, "gev_update_pos: nfield >= 2 needs chi");
    CHECK(nfield < 3 || Bi, "gev_update_pos: nfield >= 3 needs Bi");
    NEED_SORTED(p);
    return PREC_DISPATCH(p->ctx,
                         update_pos_impl<double>(p, dtau, (const double *)phi, (const double *)chi,
                                                 (const double *)Bi, nfield, a, bcoeff),
                         update_pos_impl<float>(p, dtau, (const float *)phi, (const float *)chi, (const float *)Bi,
                                                nfield, a, bcoeff));
}

int gev_update_pos_newton(gev_pcls *p, double dtau, double a)
{
    CHECK(p, "gev_update_pos_newton: NULL argument");
    NEED_SORTED(p);
    return PREC_DISPATCH(p->ctx, update_pos_newton_impl<double>(p, dtau, a), update_pos_newton_impl<float>(p, dtau, a));
}

int gev_kick_drift(gev_pcls *p, int gr, double dtau_kick, double dtau_drift, const void *phi, const void *chi,
                   const void *Bi, int nfield_kick, int nfield_drift, double a_kick, double a_drift,
                   double bcoeff_kick, double bcoeff_drift, double *max_ret, int rebin)
{
    CHECK(p && phi, "gev_kick_drift: NULL argument");
    const int nmax = nfield_kick > nfield_drift ? nfield_kick : nfield_drift;
    CHECK(nmax < 2 || chi, "gev_kick_drift: nfield >= 2 needs chi");
    CHECK(nmax < 3 || Bi, "gev_kick_drift: nfield >= 3 needs Bi");
    NEED_SORTED(p);
    return PREC_DISPATCH(p->ctx,
                         kick_drift_impl<double>(p, gr, dtau_kick, dtau_drift, (const double *)phi, (const double *)chi,
                                                 (const double *)Bi, nfield_kick, nfield_drift, a_kick, a_drift,
                                                 bcoeff_kick, bcoeff_drift, max_ret, rebin),
                         kick_drift_impl<float>(p, gr, dtau_kick, dtau_drift, (const float *)phi, (const float *)chi,
                                                (const float *)Bi, nfield_kick, nfield_drift, a_kick, a_drift,
                                                bcoeff_kick, bcoeff_drift, max_ret, rebin));
}

int gev_pcls_rebin_counted(gev_pcls *p)
{
    CHECK(p, "gev_pcls_rebin_counted: NULL argument");
    return PREC_DISPATCH(p->ctx, pcls_sort_counted_impl<double>(p), pcls_sort_counted_impl<float>(p));
}

// ------------------------------------------------------------------- IC ---
int gev_ic_displacement_ft(gev_ctx *c, uint32_t seed, int ksphere, double coeff, int npk, const double *kL,
                           const double *amp, const void *kerFT, void *potFT)
{
    CHECK(c && kL && amp && kerFT && potFT, "gev_ic_displacement_ft: NULL argument");
    CHECK(npk >= 2, "gev_ic_displacement_ft: need at least two table points");
    for (int i = 1; i < npk; i++) CHECK(kL[i] > kL[i - 1], "gev_ic_displacement_ft: k not strictly increasing");
    // natural cubic spline second derivatives (gsl cspline semantics)
    std::vector<double> m(npk, 0.0);
    if (npk >= 3) {
        int k = npk - 2;
        std::vector<double> A(k), B(k), Cc(k), D(k);
        for (int i = 1; i <= k; i++) {
            double h0 = kL[i] - kL[i - 1], h1 = kL[i + 1] - kL[i];
            A[i - 1] = h0; B[i - 1] = 2.0 * (h0 + h1); Cc[i - 1] = h1;
            D[i - 1] = 6.0 * ((amp[i + 1] - amp[i]) / h1 - (amp[i] - amp[i - 1]) / h0);
        }
        for (int i = 1; i < k; i++) {
            double w = A[i] / B[i - 1];
            B[i] -= w * Cc[i - 1];
            D[i] -= w * D[i - 1];
        }
        m[k] = D[k - 1] / B[k - 1];
        for (int i = k - 2; i >= 0; i--) m[i + 1] = (D[i] - Cc[i] * m[i + 2]) / B[i];
    }
    double *tab = nullptr;
    GEV_HIP(hipMalloc(&tab, sizeof(double) * 3 * npk));
    GEV_HIP(hipMemcpyAsync(tab, kL, sizeof(double) * npk, hipMemcpyHostToDevice, c->stream));
    GEV_HIP(hipMemcpyAsync(tab + npk, amp, sizeof(double) * npk, hipMemcpyHostToDevice, c->stream));
    GEV_HIP(hipMemcpyAsync(tab + 2 * npk, m.data(), sizeof(double) * npk, hipMemcpyHostToDevice, c->stream));
    int rc = PREC_DISPATCH(c,
                           ic_displacement_ft_impl<double>(c, seed, ksphere, coeff, npk, tab, tab + npk, tab + 2 * npk,
                                                           (const cx<double> *)kerFT, (cx<double> *)potFT),
                           ic_displacement_ft_impl<float>(c, seed, ksphere, coeff, npk, tab, tab + npk, tab + 2 * npk,
                                                          (const cx<float> *)kerFT, (cx<float> *)potFT));
    GEV_HIP(hipStreamSynchronize(c->stream));
    GEV_HIP(hipFree(tab));
    return rc;
}

int gev_displace_particles(gev_pcls *p, const void *xi, double coeff, double *max_disp)
{
    CHECK(p && xi, "gev_displace_particles: NULL argument");
    NEED_SORTED(p);
    p->qbound_valid = false;
    return PREC_DISPATCH(p->ctx, displace_impl<double>(p, (const double *)xi, coeff, max_disp),
                         displace_impl<float>(p, (const float *)xi, coeff, max_disp));
}

int gev_initialize_q(gev_pcls *p, const void *pot, double coeff, double *max_ret)
{
    CHECK(p && pot, "gev_initialize_q: NULL argument");
    NEED_SORTED(p);
    p->qbound_valid = false;
    return PREC_DISPATCH(p->ctx, initialize_q_impl<double>(p, (const double *)pot, coeff, max_ret),
                         initialize_q_impl<float>(p, (const float *)pot, coeff, max_ret));
}

// ------------------------------------------------------------ utilities ---
int gev_field_sum(gev_ctx *c, const void *f, int ncomp, double *sum)
{
    CHECK(c && f && sum && ncomp >= 1, "gev_field_sum: bad argument");
    return PREC_DISPATCH(c, field_sum_impl<double>(c, (const double *)f, c->M * ncomp, sum),
                         field_sum_impl<float>(c, (const float *)f, c->M * ncomp, sum));
}

int gev_field_fill(gev_ctx *c, void *f, int64_t count, double value)
{
    CHECK(c && f && count >= 0, "gev_field_fill: bad argument");
    return PREC_DISPATCH(c, field_fill_impl<double>(c, (double *)f, count, value),
                         field_fill_impl<float>(c, (float *)f, count, (float)value));
}

}  // extern "C"

// gev_spectra.hip -- binned power and cross spectra of Fourier fields on the
// device: extractCrossSpectrum / extractPowerSpectrum (tools.hpp:65-342) with
// KTYPE_GRID or KTYPE_LINEAR bins, scalar, vector, symmetric-tensor or
// single-component products, optionally with the CIC window divided out (the
// deconvolved density spectrum of output.hpp:3006-3033), sums in FP64.
//
// One thread owns one segment of a (kx, ky) column of the half spectrum at a
// time and walks it along the slowest index kz, so the lanes of a wavefront read
// consecutive kx of the same row (coalesced).  Along the column |k| rises for
// kz = 0..N/2 and falls after, so the bin changes only a few times per
// segment: the thread keeps the five sums of its current bin in registers and
// adds them to its workgroup's LDS bins when the bin changes.  A second
// kernel adds the workgroups' bins in a fixed order.  The LDS adds inside one
// workgroup come in no fixed order, so repeated calls may differ in the last
// bits of the sums.
#include <math.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "gev_common.hpp"
#include "gev_kernels.hpp"

namespace gev {

constexpr int PK_NT = 256;
constexpr int PK_MAX_BINS = 1024;

// Per bin: occupation, sum k, sum k^2, sum P k^3 / W^2, sum (P k^3 / W^2)^2,
// with the weight 2 of a mode whose conjugate is not stored (tools.hpp:180-290).
// The product of a mode (tools.hpp:200-225): Re(F1 conj F2) of components c1,
// c2 when both are valid, else the symmetric contraction of 6 components
// (xx xy xz yy yz zz, off-diagonals twice) when sym, else the sum over the
// ncomp components; components lie NK modes apart.  ktype 0 KTYPE_GRID,
// 1 KTYPE_LINEAR (tools.hpp:96-116).
struct PkArgs {
    const void *F1, *F2;
    int64_t NK;
    int ncomp, sym, c1, c2, ktype;
};
// grid: x over columns (PK_NT per workgroup; a grid narrower than the columns
// strides over them), y over kz segments of zlen;
// partials[j * groups + g], g = blockIdx.y * gridDim.x + blockIdx.x
template <typename T>
__global__ __launch_bounds__(PK_NT) void k_pk_columns(PkArgs pa, int N, int P, int numbins, int deconvolve, int zlen,
                                                      double *__restrict__ partials)
{
    const cx<T> *__restrict__ F1 = (const cx<T> *)pa.F1;
    const cx<T> *__restrict__ F2 = (const cx<T> *)pa.F2;
    const int64_t NK = pa.NK;
    const bool pick = pa.c1 >= 0 && pa.c2 >= 0 && pa.c1 < pa.ncomp && pa.c2 < pa.ncomp;
    extern __shared__ double sh[];
    double *bins = sh;               // [5][numbins]
    double *tk2 = sh + 5 * numbins;  // |k|^2 per index: (2N sin(pi n / N))^2 or (2 pi |n|)^2
    double *isn = tk2 + N;           // pi |n| / (N sin(pi |n| / N)), or 1
    const int H = N / 2;
    for (int j = threadIdx.x; j < 5 * numbins; j += PK_NT) bins[j] = 0.0;
    for (int i = threadIdx.x; i < N; i += PK_NT) {
        const int lin = i <= H ? i : N - i;
        const double k = pa.ktype == 0 ? 2.0 * (double)N * sin(M_PI * (double)i / (double)N) : 2.0 * M_PI * lin;
        tk2[i] = k * k;
        isn[i] = (deconvolve && lin > 0) ? M_PI * (double)lin / (sin(M_PI * (double)lin / (double)N) * (double)N) : 1.0;
    }
    __syncthreads();
    const double kmax = sqrt(3.0 * tk2[H]), inv_kmax = 1.0 / kmax;
    const int64_t ncols = (int64_t)N * (H + 1);
    for (int64_t col = (int64_t)blockIdx.x * PK_NT + threadIdx.x; col < ncols; col += (int64_t)gridDim.x * PK_NT) {
        const int kx = (int)(col % (H + 1)), ky = (int)(col / (H + 1));
        const double kxy2 = tk2[kx] + tk2[ky], ixy = isn[kx] * isn[ky];
        const double w = (kx == 0 || kx == H) ? 1.0 : 2.0;
        const int64_t fo = (int64_t)ky * P + kx;
        const int64_t zstride = (int64_t)N * P;
        int cur = -1;
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
        const int z0 = blockIdx.y * zlen, z1 = min(N, z0 + zlen);
        for (int kz = z0; kz < z1; kz++) {
            if (kz == 0 && kx == 0 && ky == 0) continue;
            const int64_t m = fo + kz * zstride;
            auto rc = [&](int a, int b) {
                const cx<T> u = F1[a * NK + m], w = F2[b * NK + m];
                return (double)u.x * (double)w.x + (double)u.y * (double)w.y;
            };
            double p;
            if (pick) {
                p = rc(pa.c1, pa.c2);
            } else if (pa.sym) {
                p = 2.0 * (rc(1, 1) + rc(2, 2) + rc(4, 4));
                p += rc(0, 0) + rc(3, 3) + rc(5, 5);
            } else {
                p = 0.0;
                for (int c = 0; c < pa.ncomp; c++) p += rc(c, c);
            }
            const double k2 = kxy2 + tk2[kz];
            const double sk = sqrt(k2);
            // the bin is floor(numbins sk / kmax) with a correctly rounded
            // division; the product with 1 / kmax agrees except within a few
            // ulp of a bin edge, where the division is done
            const double q = (double)numbins * sk;
            double r = q * inv_kmax, fl = floor(r);
            if (r - fl < 1e-9 || fl + 1.0 - r < 1e-9) fl = floor(q / kmax);
            const int b = (int)fl;
            if (b >= numbins) continue;
            if (b != cur) {
                if (cur >= 0) {
                    atomicAdd(&bins[cur], a0);
                    atomicAdd(&bins[numbins + cur], a1);
                    atomicAdd(&bins[2 * numbins + cur], a2);
                    atomicAdd(&bins[3 * numbins + cur], a3);
                    atomicAdd(&bins[4 * numbins + cur], a4);
                }
                a0 = a1 = a2 = a3 = a4 = 0.0;
                cur = b;
            }
            double is = ixy * isn[kz];
            is *= is;
            const double pk = p * k2 * sk * is;
            a0 += w;
            a1 += w * sk;
            a2 += w * k2;
            a3 += w * pk;
            a4 += w * (pk * pk);
        }
        if (cur >= 0) {
            atomicAdd(&bins[cur], a0);
            atomicAdd(&bins[numbins + cur], a1);
            atomicAdd(&bins[2 * numbins + cur], a2);
            atomicAdd(&bins[3 * numbins + cur], a3);
            atomicAdd(&bins[4 * numbins + cur], a4);
        }
    }
    __syncthreads();
    const int64_t groups = (int64_t)gridDim.x * gridDim.y, g = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    for (int j = threadIdx.x; j < 5 * numbins; j += PK_NT) partials[j * groups + g] = bins[j];
}

// out[j] = sum over g of partials[j * groups + g]: one workgroup per j, each
// thread a strided run of g, then the runs added in thread order
__global__ __launch_bounds__(PK_NT) void k_pk_reduce(const double *__restrict__ partials, int groups,
                                                     double *__restrict__ out)
{
    __shared__ double part[PK_NT / 64];
    const double *row = partials + (int64_t)blockIdx.x * groups;
    double s = 0.0;
    for (int g = threadIdx.x; g < groups; g += PK_NT) s += row[g];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < PK_NT / 64; w++) t += part[w];
        out[blockIdx.x] = t;
    }
}

// Launch shape of one call: nbx workgroups across the columns, nseg kz
// segments; every workgroup writes all 5 numbins of its partials.
struct PkShape {
    int nbx, nseg, zlen;
    int64_t groups() const { return (int64_t)nbx * nseg; }
};

static PkShape pk_shape(const gev_ctx *c)
{
    const int N = c->N;
    const int64_t cols = (int64_t)N * (N / 2 + 1);
    PkShape s;
    s.nbx = (int)((cols + PK_NT - 1) / PK_NT);
    // split the columns into kz segments until the launch holds about 2048
    // threads per compute unit, keeping segments of at least 16 modes
    int nseg = (int)std::min<int64_t>(std::max<int64_t>(1, (int64_t)c->ncu * 2048 / cols), std::max(1, N / 16));
    s.zlen = (N + nseg - 1) / nseg;
    s.nseg = (N + s.zlen - 1) / s.zlen;
    return s;
}

// buf: groups * 5 numbins partials followed by the 5 numbins sums
template <typename T>
static int power_spectrum_launch(gev_ctx *c, const PkArgs &pa, int numbins, int deconvolve, const PkShape &sh,
                                 double *buf, double *sums_host)
{
    const int nvals = 5 * numbins;
    const int64_t nb = sh.groups();
    double *out = buf + (size_t)nb * nvals;
    const size_t shm = sizeof(double) * (5 * (size_t)numbins + 2 * (size_t)c->N);
    hipLaunchKernelGGL((k_pk_columns<T>), dim3(sh.nbx, sh.nseg), dim3(PK_NT), shm, c->stream, pa, c->N, c->P, numbins,
                       deconvolve, sh.zlen, buf);
    hipLaunchKernelGGL(k_pk_reduce, dim3(nvals), dim3(PK_NT), 0, c->stream, buf, (int)nb, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(sums_host, out, sizeof(double) * nvals, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) {
        gev_set_error(std::string("gev_power_spectrum: ") + hipGetErrorString(e));
        return GEV_ERR_HIP;
    }
    return GEV_OK;
}

static int pk_lds_check(const gev_ctx *c, int numbins)
{
    const size_t shm = sizeof(double) * (5 * (size_t)numbins + 2 * (size_t)c->N);
    if (shm > 65536) {
        gev_set_error("gev_power_spectrum: numbins too large for this mesh (LDS)");
        return GEV_ERR_ARG;
    }
    return GEV_OK;
}

template <typename T>
static int power_spectrum_impl(gev_ctx *c, const PkArgs &pa, int numbins, int deconvolve, double *sums_host)
{
    if (int rc = pk_lds_check(c, numbins)) return rc;
    const PkShape sh = pk_shape(c);
    const int nvals = 5 * numbins;
    double *buf = nullptr;
    hipError_t e = hipMallocAsync((void **)&buf, sizeof(double) * ((size_t)sh.groups() + 1) * nvals, c->stream);
    if (e != hipSuccess) {
        gev_set_error(std::string("gev_power_spectrum: ") + hipGetErrorString(e));
        return GEV_ERR_NOMEM;
    }
    int rc = power_spectrum_launch<T>(c, pa, numbins, deconvolve, sh, buf, sums_host);
    e = hipFreeAsync(buf, c->stream);
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = e2;
    if (rc == GEV_OK && e != hipSuccess) {
        gev_set_error(std::string("gev_power_spectrum: ") + hipGetErrorString(e));
        rc = GEV_ERR_HIP;
    }
    return rc;
}

// ------------------------------------------------ allocation-free variant --
// Workspace of gev_power_spectrum_noalloc: the per-workgroup bin partials and
// the reduced sums, PK_SCRATCH_DOUBLES doubles (4 MiB) in every device's
// memory from the library's load until the process ends.  Host code reaches
// it only through the address that hipGetSymbolAddress returns for the
// current device (scratch_on below); the symbol itself is a host-side stand-in.
constexpr size_t PK_SCRATCH_DOUBLES = (size_t)1 << 19;
static __device__ __align__(256) double scratch[PK_SCRATCH_DOUBLES];

constexpr int PK_MAX_DEVICES = 64;
static double *g_scratch_ptr[PK_MAX_DEVICES];
static std::mutex g_scratch_mu;

// the device address of `scratch` on `device`, looked up once per device with
// that device current (the caller's current device is restored)
static int scratch_on(int device, double **out)
{
    if (device < 0 || device >= PK_MAX_DEVICES) {
        gev_set_error("gev_power_spectrum_noalloc: device ordinal out of range");
        return GEV_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int prev = -1;
        GEV_HIP(hipGetDevice(&prev));
        if (prev != device) GEV_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) GEV_HIP(hipSetDevice(prev));
        if (e != hipSuccess || !p) {
            gev_set_error(std::string("gev_power_spectrum_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return GEV_ERR_HIP;
        }
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return GEV_OK;
}

// Same sums as power_spectrum_impl, with the partials in the device's static
// workspace: the column grid is narrowed (the kernel strides over the
// columns) until the partials of all workgroups and the sums fit.  Every
// partial the reduction reads is written by the same call, so nothing in the
// workspace needs clearing; two calls on one device that overlap in time
// share the workspace and race.
template <typename T>
static int power_spectrum_noalloc_impl(gev_ctx *c, const PkArgs &pa, int numbins, int deconvolve, double *sums_host)
{
    if (int rc = pk_lds_check(c, numbins)) return rc;
    const int64_t nvals = 5 * (int64_t)numbins;
    const int64_t max_groups = (int64_t)(PK_SCRATCH_DOUBLES / (size_t)nvals) - 1;
    if (max_groups < 1) {
        gev_set_error("gev_power_spectrum_noalloc: 5 * numbins * 2 doubles exceed the 4 MiB workspace");
        return GEV_ERR_ARG;
    }
    PkShape sh = pk_shape(c);
    if (sh.groups() > max_groups) {
        sh.nseg = 1;
        sh.zlen = c->N;
        sh.nbx = (int)std::min<int64_t>(sh.nbx, max_groups);
    }
    double *buf = nullptr;
    if (int rc = scratch_on(c->device, &buf)) return rc;
    int rc = power_spectrum_launch<T>(c, pa, numbins, deconvolve, sh, buf, sums_host);
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GEV_OK && e != hipSuccess) {
        gev_set_error(std::string("gev_power_spectrum_noalloc: ") + hipGetErrorString(e));
        rc = GEV_ERR_HIP;
    }
    return rc;
}

}  // namespace gev

using namespace gev;

// per-bin normalisation of extractPowerSpectrum (tools.hpp:300-342)
static void pk_normalise(const double *s, int numbins, double *kbin, double *power, double *kscatter,
                         double *pscatter, int64_t *occupation)
{
    for (int b = 0; b < numbins; b++) {
        const double occ = s[b], ks = s[numbins + b], k2s = s[2 * numbins + b], ps = s[3 * numbins + b],
                     p2s = s[4 * numbins + b];
        occupation[b] = (int64_t)llround(occ);
        if (occ > 0) {
            kbin[b] = ks / occ;
            kscatter[b] = sqrt(fmax(k2s * occ - ks * ks, 0.0)) / occ;
            power[b] = ps / occ;
            pscatter[b] = sqrt(fmax(p2s / occ - power[b] * power[b], 0.0));
        } else {
            kbin[b] = power[b] = kscatter[b] = pscatter[b] = 0.0;
        }
    }
}

static int pk_args(const char *fn, gev_ctx *c, const void *F, int numbins, const double *kbin, const double *power,
                   const double *kscatter, const double *pscatter, const int64_t *occupation)
{
    if (!c || !F || !kbin || !power || !kscatter || !pscatter || !occupation) {
        gev_set_error(std::string(fn) + ": NULL argument");
        return GEV_ERR_ARG;
    }
    if (numbins < 1 || numbins > PK_MAX_BINS) {
        gev_set_error(std::string(fn) + ": numbins must be in [1, 1024]");
        return GEV_ERR_ARG;
    }
    return GEV_OK;
}

// one scalar field with itself, KTYPE_LINEAR (extractPowerSpectrum's default)
static PkArgs pk_scalar(const gev_ctx *c, const void *F) { return PkArgs{F, F, c->NK, 1, 0, -1, -1, 1}; }

extern "C" int gev_power_spectrum(gev_ctx *c, const void *F, int numbins, int deconvolve, double *kbin,
                                  double *power, double *kscatter, double *pscatter, int64_t *occupation)
{
    if (int rc = pk_args("gev_power_spectrum", c, F, numbins, kbin, power, kscatter, pscatter, occupation)) return rc;
    std::vector<double> s(5 * (size_t)numbins);
    const PkArgs pa = pk_scalar(c, F);
    const int rc = c->prec == GEV_FP64 ? power_spectrum_impl<double>(c, pa, numbins, deconvolve, s.data())
                                       : power_spectrum_impl<float>(c, pa, numbins, deconvolve, s.data());
    if (rc) return rc;
    pk_normalise(s.data(), numbins, kbin, power, kscatter, pscatter, occupation);
    return GEV_OK;
}

extern "C" int gev_cross_spectrum(gev_ctx *c, const void *F1, const void *F2, int ncomp, int symmetric, int comp1,
                                  int comp2, int numbins, int deconvolve, int ktype, double *kbin, double *power,
                                  double *kscatter, double *pscatter, int64_t *occupation)
{
    const char *fn = "gev_cross_spectrum";
    if (int rc = pk_args(fn, c, F1, numbins, kbin, power, kscatter, pscatter, occupation)) return rc;
    if (!F2) { gev_set_error(std::string(fn) + ": NULL argument"); return GEV_ERR_ARG; }
    if (ncomp < 1) { gev_set_error(std::string(fn) + ": ncomp must be >= 1"); return GEV_ERR_ARG; }
    if (symmetric && ncomp != 6) {
        gev_set_error(std::string(fn) + ": a symmetric tensor field has 6 components");
        return GEV_ERR_ARG;
    }
    if (ktype != 0 && ktype != 1) {
        gev_set_error(std::string(fn) + ": ktype must be 0 (KTYPE_GRID) or 1 (KTYPE_LINEAR)");
        return GEV_ERR_ARG;
    }
    std::vector<double> s(5 * (size_t)numbins);
    const PkArgs pa{F1, F2, c->NK, ncomp, symmetric ? 1 : 0, comp1, comp2, ktype};
    const int rc = c->prec == GEV_FP64 ? power_spectrum_impl<double>(c, pa, numbins, deconvolve, s.data())
                                       : power_spectrum_impl<float>(c, pa, numbins, deconvolve, s.data());
    if (rc) return rc;
    pk_normalise(s.data(), numbins, kbin, power, kscatter, pscatter, occupation);
    return GEV_OK;
}

extern "C" int gev_power_spectrum_noalloc(gev_ctx *c, const void *F, int numbins, int deconvolve, double *kbin,
                                          double *power, double *kscatter, double *pscatter, int64_t *occupation)
{
    if (int rc = pk_args("gev_power_spectrum_noalloc", c, F, numbins, kbin, power, kscatter, pscatter, occupation))
        return rc;
    double s[5 * PK_MAX_BINS];  // the sums come back into the host stack, not a heap buffer
    const PkArgs pa = pk_scalar(c, F);
    const int rc = c->prec == GEV_FP64 ? power_spectrum_noalloc_impl<double>(c, pa, numbins, deconvolve, s)
                                       : power_spectrum_noalloc_impl<float>(c, pa, numbins, deconvolve, s);
    if (rc) return rc;
    pk_normalise(s, numbins, kbin, power, kscatter, pscatter, occupation);
    return GEV_OK;
}

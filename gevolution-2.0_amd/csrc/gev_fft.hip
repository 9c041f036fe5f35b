// gev_fft.hip -- 3D real<->complex FFT for CDNA4 as three LDS-staged pencil
// passes: x (real<->half-complex, contiguous lines), then y and z (complex,
// strided columns read as TX-wide coalesced row segments).  Radix-8/4/2
// Stockham stages in LDS, twiddles from a per-size table in global memory.
#include "gev_common.hpp"
#include "gev_fft_engine.hpp"
#include "gev_kernels.hpp"

namespace gev {

// Stockham autosort stages over NLINES lines of length L stored in LDS with
// line pitch LDL.  tw = exp(-2 pi i m / (L*TWM)), m in [0, L*TWM).
template <typename T, int L, int NS, int NLINES, int NT, int LDL, int TWM>
__device__ __forceinline__ void fft_stages(cx<T> *buf, int dir, const cx<T> *__restrict__ tw)
{
    if constexpr (NS < L) {
        constexpr int REM = L / NS;
        constexpr int R = (REM >= 8) ? 8 : REM;
        constexpr int NB = L / R;
        constexpr int TOTAL = NB * NLINES;
        constexpr int PER = (TOTAL + NT - 1) / NT;
        cx<T> v[PER][R];
#pragma unroll
        for (int p = 0; p < PER; p++) {
            int b = threadIdx.x + p * NT;
            if (TOTAL % NT == 0 || b < TOTAL) {
                int line = b / NB, j = b % NB;
                int k = j % NS;
                const cx<T> *ln = buf + line * LDL;
#pragma unroll
                for (int r = 0; r < R; r++) v[p][r] = ln[j + r * NB];
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        cx<T> w = tw[(r * k * (L / (NS * R))) * TWM];
                        if (dir > 0) w.y = -w.y;
                        v[p][r] = v[p][r] * w;
                    }
                }
                dftR<T, R>(v[p], dir);
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PER; p++) {
            int b = threadIdx.x + p * NT;
            if (TOTAL % NT == 0 || b < TOTAL) {
                int line = b / NB, j = b % NB;
                int k = j % NS;
                cx<T> *ln = buf + line * LDL;
                int base = (j / NS) * NS * R + k;
#pragma unroll
                for (int r = 0; r < R; r++) ln[base + r * NS] = v[p][r];
            }
        }
        __syncthreads();
        fft_stages<T, L, NS * R, NLINES, NT, LDL, TWM>(buf, dir, tw);
    }
}

constexpr int FFT_NT = 256;

// pad a line so consecutive lines start in different LDS banks
template <typename T, int L> constexpr int ldl_of() { return L + (L >= 16 ? 1 : 0); }

// ------------------------------------------------------------ x pass r2c --
// NLINES consecutive (y,z) lines of N reals -> N/2+1 complex (pitch P).
template <typename T, int N, int NLINES>
__global__ __launch_bounds__(FFT_NT) void k_r2c_x(const T *__restrict__ in, cx<T> *__restrict__ out, int P,
                                                  int64_t nlines, const cx<T> *__restrict__ tw)
{
    constexpr int H = N / 2;
    constexpr int LDL = ldl_of<T, H>();
    __shared__ cx<T> buf[NLINES * LDL];
    const int64_t line0 = (int64_t)blockIdx.x * NLINES;
    const cx<T> *src = reinterpret_cast<const cx<T> *>(in) + line0 * H;
    for (int e = threadIdx.x; e < NLINES * H; e += FFT_NT) {
        int l = e / H, n = e % H;
        buf[l * LDL + n] = (line0 + l < nlines) ? src[e] : mk((T)0, (T)0);
    }
    __syncthreads();
    fft_stages<T, H, 1, NLINES, FFT_NT, LDL, 2>(buf, -1, tw);
    cx<T> *dst = out + line0 * P;
    for (int e = threadIdx.x; e < NLINES * P; e += FFT_NT) {
        int l = e / P, k = e % P;
        if (line0 + l >= nlines) continue;
        cx<T> X = mk((T)0, (T)0);
        if (k <= H) {
            cx<T> Zk = buf[l * LDL + (k % H)];
            cx<T> Zm = conj(buf[l * LDL + ((H - k) % H)]);
            cx<T> Fe = (Zk + Zm) * (T)0.5;
            cx<T> Fo = mul_i(Zk - Zm, -1) * (T)0.5;
            X = Fe + tw[k] * Fo;
        }
        dst[e] = X;
    }
}

// ------------------------------------------------------------ x pass c2r --
template <typename T, int N, int NLINES>
__global__ __launch_bounds__(FFT_NT) void k_c2r_x(const cx<T> *__restrict__ in, T *__restrict__ out, int P,
                                                  int64_t nlines, const cx<T> *__restrict__ tw)
{
    constexpr int H = N / 2;
    constexpr int LDL = ldl_of<T, H>();
    constexpr int LDX = H + 1;
    __shared__ cx<T> xin[NLINES * LDX];
    __shared__ cx<T> buf[NLINES * LDL];
    const int64_t line0 = (int64_t)blockIdx.x * NLINES;
    const cx<T> *src = in + line0 * P;
    for (int e = threadIdx.x; e < NLINES * P; e += FFT_NT) {
        int l = e / P, k = e % P;
        if (k <= H) xin[l * LDX + k] = (line0 + l < nlines) ? src[e] : mk((T)0, (T)0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NLINES * H; e += FFT_NT) {
        int l = e / H, k = e % H;
        cx<T> Xk = xin[l * LDX + k];
        cx<T> Xm = xin[l * LDX + (H - k)];
        if (k == 0) { Xk.y = 0; Xm.y = 0; } /* DC and Nyquist are real on the x axis */
        cx<T> cm = conj(Xm);
        cx<T> E = Xk + cm;
        cx<T> O = (Xk - cm) * conj(tw[k]);
        buf[l * LDL + k] = E + mul_i(O, +1);
    }
    __syncthreads();
    fft_stages<T, H, 1, NLINES, FFT_NT, LDL, 2>(buf, +1, tw);
    cx<T> *dst = reinterpret_cast<cx<T> *>(out) + line0 * H;
    for (int e = threadIdx.x; e < NLINES * H; e += FFT_NT) {
        int l = e / H, n = e % H;
        if (line0 + l < nlines) dst[e] = buf[l * LDL + n];
    }
}

// ------------------------------------------------ y / z pass (complex) ----
// One workgroup = TX consecutive kx columns x the full line along y (axis 1)
// or z (axis 2) at a fixed other index.  In place.
template <typename T, int L, int TX>
__global__ __launch_bounds__(FFT_NT) void k_c2c_col(cx<T> *__restrict__ data, int P, int axis, int dir,
                                                    const cx<T> *__restrict__ tw)
{
    constexpr int LDL = L + 1;
    __shared__ cx<T> buf[TX * LDL];
    const int ntiles = P / TX;
    const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int ct = (int)(b % ntiles);
    const int64_t o = b / ntiles;
    const int64_t stride = (axis == 1) ? (int64_t)P : (int64_t)P * L;
    const int64_t base = (int64_t)ct * TX + ((axis == 1) ? (int64_t)P * L * o : (int64_t)P * o);
    for (int e = threadIdx.x; e < TX * L; e += FFT_NT) {
        int j = e / TX, c = e % TX;
        buf[c * LDL + j] = data[base + j * stride + c];
    }
    __syncthreads();
    fft_stages<T, L, 1, TX, FFT_NT, LDL, 1>(buf, dir, tw);
    for (int e = threadIdx.x; e < TX * L; e += FFT_NT) {
        int j = e / TX, c = e % TX;
        data[base + j * stride + c] = buf[c * LDL + j];
    }
}

// ============================================ register-staged passes ====
template <typename T> constexpr int tx_of() { return sizeof(T) == 8 ? 8 : 16; }
template <typename T, int L> constexpr int col_tx() { return (tx_of<T>() * L / 8 > 1024) ? 8192 / L : tx_of<T>(); }
template <typename T, int N> constexpr int xlines() { return (sizeof(T) == 8 ? 4096 : 8192) / N; }

// y / z column pass, in place: TX adjacent kx columns x one full line.
template <typename T, int L, int TX>
__global__ __launch_bounds__(TX * L / 8) void k_col_reg(cx<T> *__restrict__ data, int P, int axis, int dir,
                                                        const cx<T> *__restrict__ tw)
{
    constexpr int NT = TX * L / 8;
    __shared__ cx<T> lds[TX * (L + 1)];
    __shared__ cx<T> twl[L];
    for (int m = threadIdx.x; m < L; m += NT) twl[m] = tw[m];
    const int ntiles = P / TX;
    const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int ct = (int)(b % ntiles);
    const int64_t o = b / ntiles;
    const int64_t stride = (axis == 1) ? (int64_t)P : (int64_t)P * L;
    cx<T> *base = data + (int64_t)ct * TX + ((axis == 1) ? (int64_t)P * L * o : (int64_t)P * o);
    auto ld = [&](int line, int row) { return base[(int64_t)row * stride + line]; };
    auto st = [&](int line, int row, cx<T> v) { base[(int64_t)row * stride + line] = v; };
    __syncthreads();
    reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, dir, ld, st);
}

// x pass r2c: NL consecutive lines of N reals -> N/2+1 complex (pitch P)
template <typename T, int N, int NL>
__global__ __launch_bounds__(NL * N / 16) void k_r2c_x_reg(const T *__restrict__ in, cx<T> *__restrict__ out, int P,
                                                           int64_t nlines, const cx<T> *__restrict__ tw)
{
    constexpr int H = N / 2, NT = NL * H / 8, LDL = H + 1;
    __shared__ cx<T> lds[NL * LDL];
    __shared__ cx<T> twl[N];
    for (int m = threadIdx.x; m < N; m += NT) twl[m] = tw[m];
    const int64_t line0 = (int64_t)blockIdx.x * NL;
    const cx<T> *src = reinterpret_cast<const cx<T> *>(in) + line0 * H;
    auto ld = [&](int line, int n) { return (line0 + line < nlines) ? src[line * H + n] : mk((T)0, (T)0); };
    auto st = [&](int line, int k, cx<T> v) { lds[line * LDL + k] = v; };
    __syncthreads();
    reg_fft<T, H, NL, false, 2, true, 1>(lds, twl, -1, ld, st);
    __syncthreads();
    cx<T> *dst = out + line0 * P;
    for (int e = threadIdx.x; e < NL * P; e += NT) {
        int l = e / P, k = e % P;
        if (line0 + l >= nlines) continue;
        cx<T> X = mk((T)0, (T)0);
        if (k <= H) {
            cx<T> Zk = lds[l * LDL + (k % H)];
            cx<T> Zm = conj(lds[l * LDL + ((H - k) % H)]);
            cx<T> Fe = (Zk + Zm) * (T)0.5;
            cx<T> Fo = mul_i(Zk - Zm, -1) * (T)0.5;
            X = Fe + twl[k] * Fo;
        }
        dst[e] = X;
    }
}

// x pass c2r: N/2+1 complex (pitch P) -> N reals; DC and Nyquist taken as real
template <typename T, int N, int NL>
__global__ __launch_bounds__(NL * N / 16) void k_c2r_x_reg(const cx<T> *__restrict__ in, T *__restrict__ out, int P,
                                                           int64_t nlines, const cx<T> *__restrict__ tw)
{
    constexpr int H = N / 2, NT = NL * H / 8, LDL = H + 1;
    __shared__ cx<T> xin[NL * LDL];
    __shared__ cx<T> lds[NL * LDL];
    __shared__ cx<T> twl[N];
    for (int m = threadIdx.x; m < N; m += NT) twl[m] = tw[m];
    const int64_t line0 = (int64_t)blockIdx.x * NL;
    const cx<T> *src = in + line0 * P;
    for (int e = threadIdx.x; e < NL * P; e += NT) {
        int l = e / P, k = e % P;
        if (k <= H) xin[l * LDL + k] = (line0 + l < nlines) ? src[e] : mk((T)0, (T)0);
    }
    __syncthreads();
    auto ld = [&](int line, int k) {
        cx<T> Xk = xin[line * LDL + k];
        cx<T> Xm = xin[line * LDL + (H - k)];
        if (k == 0) { Xk.y = 0; Xm.y = 0; }
        cx<T> cm = conj(Xm);
        cx<T> E = Xk + cm;
        cx<T> O = (Xk - cm) * conj(twl[k]);
        return E + mul_i(O, +1);
    };
    cx<T> *dst = reinterpret_cast<cx<T> *>(out) + line0 * H;
    auto st = [&](int line, int n, cx<T> v) {
        if (line0 + line < nlines) dst[line * H + n] = v;
    };
    reg_fft<T, H, NL, false, 2, false, 1>(lds, twl, +1, ld, st);
}

// ---------------------------------------------------------- dispatchers ---

template <typename T, int N>
static void launch_r2c(const T *in, cx<T> *out, int P, int64_t nlines, const cx<T> *tw, hipStream_t s)
{
    if constexpr (N >= 128) {
        constexpr int NL = xlines<T, N>();
        int64_t nb = (nlines + NL - 1) / NL;
        hipLaunchKernelGGL((k_r2c_x_reg<T, N, NL>), dim3((unsigned)nb), dim3(NL * N / 16), 0, s, in, out, P, nlines,
                           tw);
    } else {
        constexpr int NLINES = (4096 / N > 1024) ? 1024 : 4096 / N;
        int64_t nb = (nlines + NLINES - 1) / NLINES;
        hipLaunchKernelGGL((k_r2c_x<T, N, NLINES>), dim3((unsigned)nb), dim3(FFT_NT), 0, s, in, out, P, nlines, tw);
    }
}

template <typename T, int N>
static void launch_c2r(const cx<T> *in, T *out, int P, int64_t nlines, const cx<T> *tw, hipStream_t s)
{
    if constexpr (N >= 128) {
        constexpr int NL = xlines<T, N>();
        int64_t nb = (nlines + NL - 1) / NL;
        hipLaunchKernelGGL((k_c2r_x_reg<T, N, NL>), dim3((unsigned)nb), dim3(NL * N / 16), 0, s, in, out, P, nlines,
                           tw);
    } else {
        constexpr int NLINES = (4096 / N > 1024) ? 1024 : 4096 / N;
        int64_t nb = (nlines + NLINES - 1) / NLINES;
        hipLaunchKernelGGL((k_c2r_x<T, N, NLINES>), dim3((unsigned)nb), dim3(FFT_NT), 0, s, in, out, P, nlines, tw);
    }
}

template <typename T, int N>
static void launch_col(cx<T> *data, int P, int axis, int dir, const cx<T> *tw, hipStream_t s)
{
    if constexpr (N >= 64) {
        constexpr int TX = col_tx<T, N>();
        int64_t nb = (int64_t)(P / TX) * N;
        hipLaunchKernelGGL((k_col_reg<T, N, TX>), dim3((unsigned)nb), dim3(TX * N / 8), 0, s, data, P, axis, dir, tw);
    } else {
        constexpr int TX = tx_of<T>();
        int64_t nb = (int64_t)(P / TX) * N;
        hipLaunchKernelGGL((k_c2c_col<T, N, TX>), dim3((unsigned)nb), dim3(FFT_NT), 0, s, data, P, axis, dir, tw);
    }
}

#define GEV_FOR_EACH_N(X) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)

template <typename T>
int fft_r2c_impl(gev_ctx *ctx, const T *in, cx<T> *out, int ncomp, const cx<T> *tw)
{
    const int N = ctx->N, P = ctx->P;
    const int64_t M = ctx->M, NK = ctx->NK, nlines = (int64_t)N * N;
    for (int c = 0; c < ncomp; c++) {
        const T *ic = in + c * M;
        cx<T> *oc = out + c * NK;
        switch (N) {
#define X(n)                                                          \
    case n:                                                           \
        launch_r2c<T, n>(ic, oc, P, nlines, tw, ctx->stream);         \
        launch_col<T, n>(oc, P, 1, -1, tw, ctx->stream);              \
        launch_col<T, n>(oc, P, 2, -1, tw, ctx->stream);              \
        break;
            GEV_FOR_EACH_N(X)
#undef X
        default: gev_set_error("unsupported N"); return GEV_ERR_ARG;
        }
    }
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T>
int fft_c2r_impl(gev_ctx *ctx, cx<T> *in, T *out, int ncomp, const cx<T> *tw)
{
    const int N = ctx->N, P = ctx->P;
    const int64_t M = ctx->M, NK = ctx->NK, nlines = (int64_t)N * N;
    for (int c = 0; c < ncomp; c++) {
        cx<T> *ic = in + c * NK;
        T *oc = out + c * M;
        switch (N) {
#define X(n)                                                          \
    case n:                                                           \
        launch_col<T, n>(ic, P, 2, +1, tw, ctx->stream);              \
        launch_col<T, n>(ic, P, 1, +1, tw, ctx->stream);              \
        launch_c2r<T, n>(ic, oc, P, nlines, tw, ctx->stream);         \
        break;
            GEV_FOR_EACH_N(X)
#undef X
        default: gev_set_error("unsupported N"); return GEV_ERR_ARG;
        }
    }
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template int fft_r2c_impl<double>(gev_ctx *, const double *, cx<double> *, int, const cx<double> *);
template int fft_r2c_impl<float>(gev_ctx *, const float *, cx<float> *, int, const cx<float> *);
template int fft_c2r_impl<double>(gev_ctx *, cx<double> *, double *, int, const cx<double> *);
template int fft_c2r_impl<float>(gev_ctx *, cx<float> *, float *, int, const cx<float> *);

}  // namespace gev

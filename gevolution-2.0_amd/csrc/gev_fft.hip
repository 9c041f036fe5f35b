// gev_fft.hip -- 3D real<->complex FFT for CDNA4 as three LDS-staged pencil
// passes: x (real<->half-complex, contiguous lines), then y and z (complex,
// strided columns read as TX-wide coalesced row segments).  Radix-8/4/2
// Stockham stages in LDS, twiddles from a per-size table in global memory.
#include <type_traits>

#include "gev_common.hpp"
#include "gev_fft_engine.hpp"
#include "gev_kernels.hpp"
#include "gev_kops.hpp"

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
    auto st = [&](int line, int row, cx<T> v, int) { base[(int64_t)row * stride + line] = v; };
    __syncthreads();
    reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, dir, ld, st);
}

// Sources of the r2c x pass: the pair of reals (2n, 2n+1) of line L (= y + N z)
template <typename T, int N> struct SrcPlain {
    static constexpr bool kSums = false;
    const T *in;
    __device__ __forceinline__ cx<T> operator()(int64_t L, int n) const
    {
        return reinterpret_cast<const cx<T> *>(in)[L * (N / 2) + n];
    }
};
// component C of the Sij source of prepareFTsource_Tij (gevolution.hpp:64-107),
// computed from Tij and phi while the pass reads it (the same expressions as
// k_prepare_Tij); C: 0 xx, 1 xy, 2 xz, 3 yy, 4 yz, 5 zz
template <typename T, int N, int C> struct SrcSij {
    static constexpr bool kSums = false;
    const T *phi, *Tc;  // Tc: component C of Tij
    T coeff;
    __device__ __forceinline__ cx<T> operator()(int64_t L, int n) const
    {
        constexpr int lg = __builtin_ctz(N), M1 = N - 1;
        const int y = (int)(L & M1), z = (int)(L >> lg);
        const int x = 2 * n, xm = (x - 1) & M1, xp2 = (x + 2) & M1;
        auto row = [&](int dy, int dz) {
            return phi + (int64_t)((y + dy) & M1) * N + (int64_t)((z + dz) & M1) * N * N;
        };
        const cx<T> t = reinterpret_cast<const cx<T> *>(Tc + L * N)[n];
        const T *r0 = phi + L * N;
        const cx<T> pp = reinterpret_cast<const cx<T> *>(r0)[n];  // phi at x, x+1
        const T h = (T)0.5;
        cx<T> out;
        if constexpr (C == 0) {
            const T pm = r0[xm], p2 = r0[xp2];
            out.x = h * (pp.y - pm) * (pp.y - pm) + coeff * t.x;
            out.y = h * (p2 - pp.x) * (p2 - pp.x) + coeff * t.y;
        } else if constexpr (C == 3 || C == 5) {
            const T *ra = C == 3 ? row(1, 0) : row(0, 1), *rb = C == 3 ? row(-1, 0) : row(0, -1);
            const cx<T> a = reinterpret_cast<const cx<T> *>(ra)[n], b = reinterpret_cast<const cx<T> *>(rb)[n];
            out.x = h * (a.x - b.x) * (a.x - b.x) + coeff * t.x;
            out.y = h * (a.y - b.y) * (a.y - b.y) + coeff * t.y;
        } else if constexpr (C == 1 || C == 2) {
            // px py - p pxy + (p^2 - px^2 - py^2 + pxy^2)/2 with y -> z for C = 2
            const T *r1 = C == 1 ? row(1, 0) : row(0, 1);
            const cx<T> q = reinterpret_cast<const cx<T> *>(r1)[n];
            const T p2 = r0[xp2], q2 = r1[xp2];
            auto f = [&](T p, T px, T py, T pxy) { return px * py - p * pxy + h * (p * p - px * px - py * py + pxy * pxy); };
            out.x = f(pp.x, pp.y, q.x, q.y) + coeff * t.x;
            out.y = f(pp.y, p2, q.y, q2) + coeff * t.y;
        } else {
            // py pz - p pyz + (p^2 - py^2 - pz^2 + pyz^2)/2
            const cx<T> a = reinterpret_cast<const cx<T> *>(row(1, 0))[n], b = reinterpret_cast<const cx<T> *>(row(0, 1))[n],
                        d = reinterpret_cast<const cx<T> *>(row(1, 1))[n];
            auto f = [&](T p, T py, T pz, T pyz) { return py * pz - p * pyz + h * (p * p - py * py - pz * pz + pyz * pyz); };
            out.x = f(pp.x, a.x, b.x, d.x) + coeff * t.x;
            out.y = f(pp.y, a.y, b.y, d.y) + coeff * t.y;
        }
        return out;
    }
};

// component C of T0i from its link sums s (the particle parts of
// projection_T0i, gevolution.hpp:2031-2201) and the link factor
// 1 + phi(x) + phi(x + e_C) of the link from site x along axis C
template <typename T, int N, int C> struct SrcT0iLink {
    static constexpr bool kSums = false;
    const T *phi, *s;  // s: component C of the link sums
    __device__ __forceinline__ cx<T> operator()(int64_t L, int n) const
    {
        constexpr int lg = __builtin_ctz(N), M1 = N - 1;
        const int y = (int)(L & M1), z = (int)(L >> lg);
        const cx<T> sv = reinterpret_cast<const cx<T> *>(s + L * N)[n];
        const T *r0 = phi + L * N;
        const cx<T> p = reinterpret_cast<const cx<T> *>(r0)[n];  // phi at x, x+1
        cx<T> e;  // phi at the far end of the links from x and x+1
        if constexpr (C == 0) {
            e = mk(p.y, r0[(2 * n + 2) & M1]);
        } else {
            const int64_t row = C == 1 ? (int64_t)((y + 1) & M1) + (int64_t)z * N : (int64_t)y + (int64_t)((z + 1) & M1) * N;
            e = reinterpret_cast<const cx<T> *>(phi + row * N)[n];
        }
        return mk(sv.x * ((T)1 + p.x + e.x), sv.y * ((T)1 + p.y + e.y));
    }
};

// prepareFTsource's T00 form (gevolution.hpp:160-171, the same expressions as
// k_prepare_T00) evaluated while the Poisson solve's x pass reads it; every
// copy sums the source and the result it produced (FP64) for the k=0 mode
template <typename T, int N> struct SrcT00 {
    static constexpr bool kSums = true;
    const T *phi, *chi, *src;
    T bg, c1, c2, c3;
    double acc[2];
    __device__ __forceinline__ cx<T> operator()(int64_t L, int n)
    {
        constexpr int lg = __builtin_ctz(N), M1 = N - 1;
        const int y = (int)(L & M1), z = (int)(L >> lg);
        auto pair = [&](const T *f, int64_t line) { return reinterpret_cast<const cx<T> *>(f + line * N)[n]; };
        const int64_t ym = (int64_t)((y - 1) & M1) + (int64_t)z * N, yp = (int64_t)((y + 1) & M1) + (int64_t)z * N;
        const int64_t zm = (int64_t)y + (int64_t)((z - 1) & M1) * N, zp = (int64_t)y + (int64_t)((z + 1) & M1) * N;
        const T *r0 = phi + L * N;
        const cx<T> p = pair(phi, L), s = pair(src, L), ch = pair(chi, L);
        const cx<T> a = pair(phi, ym), b = pair(phi, yp), d = pair(phi, zm), e = pair(phi, zp);
        const T pm = r0[(2 * n - 1) & M1], p2 = r0[(2 * n + 2) & M1];
        auto res = [&](T sv, T pv, T xm, T xp, T ym_, T yp_, T zm_, T zp_, T cv) {
            T r = c2 * (sv - bg);
            r *= (T)1 - (T)2 * pv * ((T)1 - pv);
            const T d0 = xm - xp, d1 = ym_ - yp_, d2 = zm_ - zp_;
            r += (T)0.125 * d0 * d0;
            r += (T)0.125 * d1 * d1;
            r += (T)0.125 * d2 * d2;
            r += (c3 * ((T)1 - (T)3 * pv) - c1) * pv - c3 * cv;
            return r;
        };
        const T r0v = res(s.x, p.x, pm, p.y, a.x, b.x, d.x, e.x, ch.x);
        const T r1v = res(s.y, p.y, p.x, p2, a.y, b.y, d.y, e.y, ch.y);
        acc[0] += (double)s.x + (double)s.y;
        acc[1] += (double)r0v + (double)r1v;
        return mk(r0v, r1v);
    }
};

// x pass r2c: NL consecutive lines of N reals -> N/2+1 complex (pitch P)
template <typename T, int N, int NL, typename Src>
__global__ __launch_bounds__(NL * N / 16) void k_r2c_x_reg(Src src, cx<T> *__restrict__ out, int P, int64_t nlines,
                                                           const cx<T> *__restrict__ tw, double *partials)
{
    constexpr int H = N / 2, NT = NL * H / 8, LDL = H + 1;
    __shared__ cx<T> lds[NL * LDL];
    __shared__ cx<T> twl[N];
    for (int m = threadIdx.x; m < N; m += NT) twl[m] = tw[m];
    const int64_t line0 = (int64_t)blockIdx.x * NL;
    Src sr = src;  // a source that sums keeps its sums per thread
    if constexpr (Src::kSums) sr.acc[0] = sr.acc[1] = 0.0;
    auto ld = [&](int line, int n) { return (line0 + line < nlines) ? sr(line0 + line, n) : mk((T)0, (T)0); };
    auto st = [&](int line, int k, cx<T> v, int) { lds[line * LDL + k] = v; };
    __syncthreads();
    reg_fft<T, H, NL, false, 2, true, 1>(lds, twl, -1, ld, st);
    if constexpr (Src::kSums) {
        // workgroup partials of the two sums, one slot pair per workgroup
        __shared__ double ws[NT / 64][2];
        double v0 = sr.acc[0], v1 = sr.acc[1];
        for (int o = 32; o > 0; o >>= 1) {
            v0 += __shfl_xor(v0, o, 64);
            v1 += __shfl_xor(v1, o, 64);
        }
        if ((threadIdx.x & 63) == 0) { ws[threadIdx.x >> 6][0] = v0; ws[threadIdx.x >> 6][1] = v1; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double t0 = 0.0, t1 = 0.0;
            for (int w = 0; w < NT / 64; w++) { t0 += ws[w][0]; t1 += ws[w][1]; }
            partials[2 * (int64_t)blockIdx.x] = t0;
            partials[2 * (int64_t)blockIdx.x + 1] = t1;
        }
    }
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

// x pass r2c of NC components of the same lines in one workgroup (out + c NK
// for component c): the sources of all components read the same phi rows,
// which the later components then find in the caches.  Src::get<C>(L, n)
// gives component C's pair n of line L.
template <typename T, int N, int NL, int NC, typename Src>
__global__ __launch_bounds__(NL * N / 16) void k_r2c_x_multi(Src src, cx<T> *__restrict__ out, int P, int64_t NK,
                                                             int64_t nlines, const cx<T> *__restrict__ tw)
{
    constexpr int H = N / 2, NT = NL * H / 8, LDL = H + 1;
    __shared__ cx<T> lds[NL * LDL];
    __shared__ cx<T> twl[N];
    for (int m = threadIdx.x; m < N; m += NT) twl[m] = tw[m];
    const int64_t line0 = (int64_t)blockIdx.x * NL;
    auto comp = [&](auto cc) {
        constexpr int C = decltype(cc)::value;
        auto ld = [&](int line, int n) {
            return (line0 + line < nlines) ? src.template get<C>(line0 + line, n) : mk((T)0, (T)0);
        };
        auto st = [&](int line, int k, cx<T> v, int) { lds[line * LDL + k] = v; };
        __syncthreads();  // twiddles loaded; the previous component's stores have read lds
        reg_fft<T, H, NL, false, 2, true, 1>(lds, twl, -1, ld, st);
        __syncthreads();
        cx<T> *dst = out + C * NK + line0 * P;
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
    };
    comp(std::integral_constant<int, 0>{});
    if constexpr (NC > 1) comp(std::integral_constant<int, 1>{});
    if constexpr (NC > 2) comp(std::integral_constant<int, 2>{});
    if constexpr (NC > 3) comp(std::integral_constant<int, 3>{});
    if constexpr (NC > 4) comp(std::integral_constant<int, 4>{});
    if constexpr (NC > 5) comp(std::integral_constant<int, 5>{});
}

// the six Sij components (Tij at Tij + C M)
template <typename T, int N> struct SrcSij6 {
    const T *phi, *Tij;
    int64_t M;
    T coeff;
    template <int C> __device__ __forceinline__ cx<T> get(int64_t L, int n) const
    {
        return SrcSij<T, N, C>{phi, Tij + C * M, coeff}(L, n);
    }
};

template <typename T, int N, int NC, typename Src>
static void launch_r2c_multi(const Src &src, cx<T> *out, int P, int64_t NK, int64_t nlines, const cx<T> *tw,
                             hipStream_t s)
{
    constexpr int NL = xlines<T, N>();
    int64_t nb = (nlines + NL - 1) / NL;
    hipLaunchKernelGGL((k_r2c_x_multi<T, N, NL, NC, Src>), dim3((unsigned)nb), dim3(NL * N / 16), 0, s, src, out, P, NK,
                       nlines, tw);
}

// x pass c2r: N/2+1 complex (pitch P) -> N reals; DC and Nyquist taken as real.
// The lines are staged in LDS with one coalesced read per element, the first
// stage pairs X[k] with X[H-k] from there, and the last stage's outputs go
// back through LDS so that the global stores are consecutive per lane group.
template <typename T, int N, int NL>
__global__ __launch_bounds__(NL * N / 16) void k_c2r_x_reg(const cx<T> *__restrict__ in, T *__restrict__ out, int P,
                                                           int64_t nlines, const cx<T> *__restrict__ tw)
{
    constexpr int H = N / 2, NT = NL * H / 8, LDL = H + 1;
    __shared__ cx<T> lds[NL * LDL];
    __shared__ cx<T> twl[N];
    for (int m = threadIdx.x; m < N; m += NT) twl[m] = tw[m];
    const int64_t line0 = (int64_t)blockIdx.x * NL;
    const cx<T> *src = in + line0 * P;
    static_assert(LDL == H + 1, "a staged line holds X[0..H]");
    for (int e = threadIdx.x; e < NL * LDL; e += NT) {
        const int l = e / LDL;
        lds[e] = (line0 + l < nlines) ? src[(int64_t)l * P + (e - l * LDL)] : mk((T)0, (T)0);
    }
    __syncthreads();
    auto ld = [&](int line, int k) {
        const cx<T> *row = lds + line * LDL;
        cx<T> Xk = row[k];
        cx<T> Xm = row[H - k];
        if (k == 0) { Xk.y = 0; Xm.y = 0; }
        cx<T> cm = conj(Xm);
        cx<T> E = Xk + cm;
        cx<T> O = (Xk - cm) * conj(twl[k]);
        return E + mul_i(O, +1);
    };
    cx<T> *dst = reinterpret_cast<cx<T> *>(out) + line0 * H;
    auto st = [&](int line, int n, cx<T> v, int) { lds[line * LDL + n] = v; };
    reg_fft<T, H, NL, false, 2, true, 1>(lds, twl, +1, ld, st);
    __syncthreads();
    for (int e = threadIdx.x; e < NL * H; e += NT) {
        const int l = e / H;
        if (line0 + l < nlines) dst[e] = lds[l * LDL + (e - l * H)];
    }
}

// ---------------------------------------------------------- dispatchers ---

template <typename T, int N, typename Src>
static void launch_r2c_src(const Src &src, cx<T> *out, int P, int64_t nlines, const cx<T> *tw, hipStream_t s,
                           double *partials = nullptr)
{
    constexpr int NL = xlines<T, N>();
    int64_t nb = (nlines + NL - 1) / NL;
    hipLaunchKernelGGL((k_r2c_x_reg<T, N, NL, Src>), dim3((unsigned)nb), dim3(NL * N / 16), 0, s, src, out, P, nlines,
                       tw, partials);
}

template <typename T, int N>
static void launch_r2c(const T *in, cx<T> *out, int P, int64_t nlines, const cx<T> *tw, hipStream_t s)
{
    if constexpr (N >= 128) {
        launch_r2c_src<T, N>(SrcPlain<T, N>{in}, out, P, nlines, tw, s);
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

// ===================================== z pass fused with a k-space solve ====
// Forward z transforms of the NIN input components of TX adjacent kx columns
// (one ky), the per-mode operator applied in registers, inverse z transforms
// of the NOUT outputs.  Replaces [z pass, k-space kernel, inverse z pass]:
// every mode is read once and written once instead of three times each.
// In place (in == out) is allowed: a workgroup reads all its columns before
// writing them and no two workgroups share a column.
// ZOP_SCALAR3: projectFTscalar from the three groups A, B, C that k_col_comb
// forms in the y pass (below) instead of the six Sij components
enum { ZOP_POISSON = 0, ZOP_SCALAR = 1, ZOP_VECTOR = 2, ZOP_SCALAR3 = 3 };

struct ZParams {
    const double *gk2;
    const cx<double> *ks;
    double coeff, modif;  // Poisson: coeff pre-scaled by -1/N^3; vector: as projectFTvector
    double k0re, k0im;    // Poisson: value taken for the k=0 mode when k0set
    int k0set;
    const double *k0dev;  // when set, the real part of that mode is read here instead
};

// ============================ y pass of Sij forming the groups of the z solve ==
// The y transforms of the six Sij components (x-transformed, in place at
// data + c NK) of TX adjacent kx columns and one z, each output mode added
// with its (kx, ky) weight of projectFTscalar into the three groups that the
// z solve transforms (gevolution.hpp:317-323 with g = gridk2, k = kshift):
//   A = (g1 - 2 g0) S_xx - 6 k0 k1 S_xy + (g0 - 2 g1) S_yy + (g0 + g1) S_zz
//   B = S_xx + S_yy - 2 S_zz,   C = -6 (k0 S_xz + k1 S_yz)
// written over components 0..2 (every column of the workgroup is read before
// any is written; no other workgroup touches them).  Padding columns
// (kx > N/2) get 0.  Replaces six y passes and the z solve's reads of six
// components by reads of three.
template <typename T, int L, int TX>
__global__ __launch_bounds__(TX * L / 8) void k_col_comb(cx<T> *data, int P, int64_t NK, const cx<T> *__restrict__ tw,
                                                         const double *__restrict__ gk2, const cx<double> *__restrict__ ks)
{
    constexpr int NT = TX * L / 8;
    __shared__ cx<T> lds[TX * (L + 1)];
    __shared__ cx<T> twl[L];
    // the k tables in the transform's precision (the weights are applied in it)
    __shared__ T gkl[L];
    __shared__ cx<T> ksl[L];
    for (int m = threadIdx.x; m < L; m += NT) {
        twl[m] = tw[m];
        gkl[m] = (T)gk2[m];
        ksl[m] = mk((T)ks[m].x, (T)ks[m].y);
    }
    const int ntiles = P / TX;
    const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int ct = (int)(b % ntiles);
    const int64_t z = b / ntiles;
    cx<T> *base = data + (int64_t)ct * TX + (int64_t)P * L * z;
    // byte offset of (row, column) inside a component's (kx, y) plane of this
    // z: 32 bits, so every access takes the workgroup's base from scalar
    // registers instead of a 64-bit address per element
    auto boff = [&](int row, int ln) -> uint32_t { return ((uint32_t)row * (uint32_t)P + (uint32_t)ln) * (uint32_t)sizeof(cx<T>); };
    const int line = threadIdx.x % TX;  // the column every slot of this thread lands in
    const int kx = ct * TX + line;
    const bool valid = kx <= L / 2;
    const T g0 = valid ? (T)gk2[kx] : (T)0;
    const cx<T> k0 = valid ? mk((T)ks[kx].x, (T)ks[kx].y) : mk((T)0, (T)0);
    const cx<T> k0m = k0 * (T)-6;
    cx<T> A[8], B[8], C[8], v[8];
    int rows[8];
    auto comp = [&](auto cc) {
        constexpr int c = decltype(cc)::value;
        const cx<T> *src = base + c * NK;
        auto ld = [&](int ln, int row) { return *reinterpret_cast<const cx<T> *>(reinterpret_cast<const char *>(src) + boff(row, ln)); };
        auto st = [&](int, int row, cx<T> val, int slot) {
            if constexpr (c == 0) rows[slot] = row;
            v[slot] = val;
        };
        __syncthreads();  // tables loaded; the previous transform is done with lds
        reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, -1, ld, st);
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int row = rows[s];
            if constexpr (c == 0) {
                A[s] = v[s] * (gkl[row] - (T)2 * g0);
                B[s] = v[s];
            } else if constexpr (c == 1) {
                A[s] = A[s] + (k0m * ksl[row]) * v[s];
            } else if constexpr (c == 2) {
                C[s] = k0m * v[s];
            } else if constexpr (c == 3) {
                A[s] = A[s] + v[s] * (g0 - (T)2 * gkl[row]);
                B[s] = B[s] + v[s];
            } else if constexpr (c == 4) {
                C[s] = C[s] + (ksl[row] * (T)-6) * v[s];
            } else {
                A[s] = A[s] + v[s] * (g0 + gkl[row]);
                B[s] = B[s] - v[s] * (T)2;
            }
        }
    };
    comp(std::integral_constant<int, 0>{});
    comp(std::integral_constant<int, 1>{});
    comp(std::integral_constant<int, 2>{});
    comp(std::integral_constant<int, 3>{});
    comp(std::integral_constant<int, 4>{});
    comp(std::integral_constant<int, 5>{});
    // (selects on the scalar parts: a select of whole structs becomes a select
    // of their addresses, which keeps the arrays out of registers)
    auto put = [&](cx<T> *dst, int s, cx<T> val) {
        *reinterpret_cast<cx<T> *>(reinterpret_cast<char *>(dst) + boff(rows[s], line)) =
            mk(valid ? val.x : (T)0, valid ? val.y : (T)0);
    };
#pragma unroll
    for (int s = 0; s < 8; s++) {
        put(base, s, A[s]);
        put(base + NK, s, B[s]);
        put(base + 2 * NK, s, C[s]);
    }
}

template <typename T, int L> constexpr int comb_tx() { return 4096 / L < 8 ? 4096 / L : 8; }

template <typename T, int N>
static void launch_col_comb(cx<T> *data, int P, int64_t NK, const cx<T> *tw, const ZParams &zp, hipStream_t s)
{
    constexpr int TX = comb_tx<T, N>();
    const int64_t nb = (int64_t)(P / TX) * N;
    hipLaunchKernelGGL((k_col_comb<T, N, TX>), dim3((unsigned)nb), dim3(TX * N / 8), 0, s, data, P, NK, tw, zp.gk2,
                       zp.ks);
}

// FP32 vector projection from L = 128 up: 4 waves per SIMD (two workgroups
// per CU) at the price of a spill of up to 80 bytes, 4.5 -> 4.1 ms for the
// 512^3 B solve; the scalar projection and FP64 measured no better that way
// ZS2_DIRECT: the vector solve reads each component in its transform's first
// stage and applies the projection per mode with the last one (1.31 against
// 1.34 ms at 512^3, spill 56 against 80 bytes); 0 keeps the all-operands-first form
#ifndef ZS2_DIRECT
#define ZS2_DIRECT 1
#endif
#ifndef ZS2_WAVES
#define ZS2_WAVES 4
#endif
template <typename T, int L, int OP> constexpr int zs_waves() { return (OP == 2 && sizeof(T) == 4 && L >= 128) ? ZS2_WAVES : 1; }

template <typename T, int L, int TX, int OP>
__global__ __launch_bounds__(TX * L / 8, (zs_waves<T, L, OP>())) void k_zsolve(const cx<T> *in, cx<T> *out, int P, int64_t NK,
                                                       const cx<T> *__restrict__ tw, ZParams zp)
{
    constexpr int NT = TX * L / 8, LDL = L + 1;
    constexpr int NOUT = OP == ZOP_VECTOR ? 3 : 1;
    constexpr int NV = OP == ZOP_VECTOR ? 3 : 1;  // inputs held at once
    __shared__ cx<T> lds[TX * LDL];
    __shared__ cx<T> twl[L];
    for (int m = threadIdx.x; m < L; m += NT) twl[m] = tw[m];
    const int ntiles = P / TX;
    const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int ct = (int)(b % ntiles);
    const int ky = (int)(b / ntiles);
    const int64_t stride = (int64_t)P * L;
    const int64_t off = (int64_t)ct * TX + (int64_t)P * ky;
    const int H = L / 2;
    int lines[8], poss[8];
    cx<T> vin[NV][8];
    cx<double> acc[8];
#pragma unroll
    for (int s = 0; s < 8; s++) acc[s] = mk(0.0, 0.0);
    const double gy = zp.gk2[ky];
    const cx<double> kyv = zp.ks[ky];
    // forward z transform of input component c (a compile-time index when the
    // operator holds all inputs at once, so they stay in registers)
    auto forward = [&](int c, auto held) {
        constexpr int CH = decltype(held)::value;
        const cx<T> *src = in + c * NK + off;
        auto ld = [&](int line, int row) { return src[(int64_t)row * stride + line]; };
        auto st = [&](int line, int row, cx<T> v, int slot) {
            lines[slot] = line;
            poss[slot] = row;
            if constexpr (OP == ZOP_SCALAR) {
                const int kx = ct * TX + line;
                if (kx <= H) {
                    const cx<double> w =
                        kop_scalar_weight(c, zp.gk2[kx], gy, zp.gk2[row], zp.ks[kx], kyv, zp.ks[row]);
                    acc[slot] = acc[slot] + w * mk((double)v.x, (double)v.y);
                }
            } else {
                vin[CH][slot] = v;
            }
        };
        __syncthreads();  // twiddles loaded; previous transform done with lds
        reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, -1, ld, st);
    };
    // ZOP_SCALAR3: the k tables in LDS (FP64), read per mode by the
    // combination and the operator
    constexpr bool TAB = OP == ZOP_SCALAR3;
    __shared__ double gkd[TAB ? L : 1];
    __shared__ cx<double> ksd[TAB ? L : 1];
    if constexpr (TAB)
        for (int m = threadIdx.x; m < L; m += NT) {
            gkd[m] = zp.gk2[m];
            ksd[m] = zp.ks[m];
        }
    if constexpr (OP == ZOP_SCALAR3) {
        // chi from the three kz groups (formed by the y pass):
        //   sum_c w_c S_c = FFT_z(A) + g2 FFT_z(B) + k2 FFT_z(C)
        // each group read by its transform's first stage straight from
        // memory, so only the FP64 sums stay live between the transforms
        auto fft = [&](auto kind) {
            constexpr int K = decltype(kind)::value;
            const cx<T> *src = in + K * NK + off;
            auto ld = [&](int line, int row) { return src[(int64_t)row * stride + line]; };
            auto st = [&](int ln, int row, cx<T> v, int slot) {
                lines[slot] = ln;
                poss[slot] = row;
                const cx<double> d = mk((double)v.x, (double)v.y);
                if constexpr (K == 0) {
                    acc[slot] = d;
                } else if constexpr (K == 1) {
                    acc[slot] = acc[slot] + d * gkd[row];
                } else {
                    // the last group completes the mode: the operator here,
                    // so each slot's FP64 sum dies at once (eight of them
                    // held through an operator loop doubled the registers)
                    const cx<double> sum = acc[slot] + ksd[row] * d;
                    const int kx = ct * TX + ln;
                    cx<double> r = mk(0.0, 0.0);
                    if (kx <= H && !(kx == 0 && ky == 0 && row == 0))
                        r = sum * kop_scalar_norm(gkd[kx] + gy + gkd[row], L);
                    vin[0][slot] = mk((T)r.x, (T)r.y);
                }
            };
            __syncthreads();  // tables loaded; previous transform done with lds
            reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, -1, ld, st);
        };
        fft(std::integral_constant<int, 0>{});
        fft(std::integral_constant<int, 1>{});
        fft(std::integral_constant<int, 2>{});
    } else if constexpr (OP == ZOP_VECTOR && ZS2_DIRECT) {
        // each component read by its transform's first stage straight from
        // memory; the projection applied per mode when the last component
        // arrives, so no operator loop holds eight modes' FP64 values
        auto fwd = [&](auto held) {
            constexpr int CH = decltype(held)::value;
            const cx<T> *src = in + CH * NK + off;
            auto ld = [&](int line, int row) { return src[(int64_t)row * stride + line]; };
            auto st = [&](int ln, int row, cx<T> v, int slot) {
                lines[slot] = ln;
                poss[slot] = row;
                if constexpr (CH < 2) {
                    vin[CH][slot] = v;
                } else {
                    const int kx = ct * TX + ln;
                    cx<double> bo[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
                    if (kx <= H && !(kx == 0 && ky == 0 && row == 0)) {
                        const cx<double> sv[3] = {mk((double)vin[0][slot].x, (double)vin[0][slot].y),
                                                  mk((double)vin[1][slot].x, (double)vin[1][slot].y),
                                                  mk((double)v.x, (double)v.y)};
                        kop_vector(sv, zp.gk2[kx] + gy + zp.gk2[row], zp.ks[kx], kyv, zp.ks[row], zp.coeff, zp.modif,
                                   bo);
                    }
#pragma unroll
                    for (int j = 0; j < 3; j++) vin[j][slot] = mk((T)bo[j].x, (T)bo[j].y);
                }
            };
            __syncthreads();  // twiddles loaded; previous transform done with lds
            reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, -1, ld, st);
        };
        fwd(std::integral_constant<int, 0>{});
        fwd(std::integral_constant<int, 1>{});
        fwd(std::integral_constant<int, 2>{});
    } else if constexpr (OP == ZOP_VECTOR) {
        // all 24 first-stage operands (3 components x the 8 elements reg_fft's
        // first stage reads: line t % TX, rows t / TX + r L/8) are loaded
        // before the first transform
        constexpr int NB = L / 8;
        const int line = threadIdx.x % TX, jt = threadIdx.x / TX;
        cx<T> pre[3][8];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 8; r++) pre[c][r] = in[c * NK + off + (int64_t)(jt + r * NB) * stride + line];
        auto fwd = [&](auto held) {
            constexpr int CH = decltype(held)::value;
            auto ld = [&](int, int row) { return pre[CH][(row - jt) / NB]; };
            auto st = [&](int ln, int row, cx<T> v, int slot) {
                lines[slot] = ln;
                poss[slot] = row;
                vin[CH][slot] = v;
            };
            __syncthreads();  // twiddles loaded; previous transform done with lds
            reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, -1, ld, st);
        };
        fwd(std::integral_constant<int, 0>{});
        fwd(std::integral_constant<int, 1>{});
        fwd(std::integral_constant<int, 2>{});
    } else if constexpr (OP == ZOP_POISSON) {
        forward(0, std::integral_constant<int, 0>{});
    } else {
        // projectFTscalar: chi = sum_c w_c S_c, and every w_c is linear in the
        // kz factors g2 = gk2[kz] and k2 = ks[kz]; three combinations formed
        // before the z transform carry the whole projection:
        //   A = (g1 - 2 g0) S_xx - 6 k0 k1 S_xy + (g0 - 2 g1) S_yy + (g0 + g1) S_zz
        //   B = S_xx + S_yy - 2 S_zz,   C = -6 (k0 S_xz + k1 S_yz)
        //   sum_c w_c S_c = FFT_z(A) + g2 FFT_z(B) + k2 FFT_z(C)
        // Each thread combines the 8 elements reg_fft's first stage reads
        // (line t % TX, rows t / TX + r L/8).
        constexpr int NB = L / 8;
        const int line = threadIdx.x % TX, jt = threadIdx.x / TX;
        const int kx = ct * TX + line;
        const bool valid = kx <= H;  // padding columns give 0 below
        const double g0 = valid ? zp.gk2[kx] : 0.0;
        const cx<double> k0 = valid ? zp.ks[kx] : mk(0.0, 0.0);
        const cx<double> k01 = (k0 * kyv) * -6.0, k0m = k0 * -6.0, k1m = kyv * -6.0;
        const T wa0 = (T)(gy - 2.0 * g0), wa3 = (T)(g0 - 2.0 * gy), wa5 = (T)(g0 + gy);
        const cx<T> wa1 = mk((T)k01.x, (T)k01.y), wc2 = mk((T)k0m.x, (T)k0m.y), wc4 = mk((T)k1m.x, (T)k1m.y);
        cx<T> A[8], B[8], C[8];
        const cx<T> *src = in + off + line;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int64_t e = (int64_t)(jt + r * NB) * stride;
            const cx<T> s0 = src[e], s1 = src[NK + e], s2 = src[2 * NK + e], s3 = src[3 * NK + e],
                        s4 = src[4 * NK + e], s5 = src[5 * NK + e];
            A[r] = s0 * wa0 + wa1 * s1 + s3 * wa3 + s5 * wa5;
            B[r] = s0 + s3 - s5 * (T)2;
            C[r] = wc2 * s2 + wc4 * s4;
        }
        auto fft = [&](const cx<T> *pre, auto kind) {
            constexpr int K = decltype(kind)::value;
            auto ld = [&](int, int row) { return pre[(row - jt) / NB]; };
            auto st = [&](int ln, int row, cx<T> v, int slot) {
                lines[slot] = ln;
                poss[slot] = row;
                const cx<double> d = mk((double)v.x, (double)v.y);
                if constexpr (K == 0) acc[slot] = d;
                else if constexpr (K == 1) acc[slot] = acc[slot] + d * zp.gk2[row];
                else acc[slot] = acc[slot] + zp.ks[row] * d;
            };
            __syncthreads();  // twiddles loaded; previous transform done with lds
            reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, -1, ld, st);
        };
        fft(A, std::integral_constant<int, 0>{});
        fft(B, std::integral_constant<int, 1>{});
        fft(C, std::integral_constant<int, 2>{});
    }
    // the operator, per owned mode; the outputs replace the inputs in vin
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int kx = ct * TX + lines[s], kz = poss[s];
        const bool pad = kx > H;
        const bool zero_mode = (kx == 0 && ky == 0 && kz == 0);
        const double kk = (pad || TAB) ? 1.0 : zp.gk2[kx] + gy + zp.gk2[kz];
        if constexpr (OP == ZOP_POISSON) {
            cx<double> r = mk(0.0, 0.0);
            if (!pad) {
                cx<double> sv = mk((double)vin[0][s].x, (double)vin[0][s].y);
                if (zero_mode && zp.k0set) sv = mk(zp.k0dev ? *zp.k0dev : zp.k0re, zp.k0im);
                if (!(zero_mode && zp.modif == 0.0)) r = kop_poisson(sv, kk, zp.coeff, zp.modif);
            }
            vin[0][s] = mk((T)r.x, (T)r.y);
        } else if constexpr (OP == ZOP_SCALAR3) {
            // applied with the last forward transform
        } else if constexpr (OP == ZOP_SCALAR) {
            cx<double> r = mk(0.0, 0.0);
            if (!pad && !zero_mode) r = acc[s] * kop_scalar_norm(kk, L);
            vin[0][s] = mk((T)r.x, (T)r.y);
        } else if constexpr (!ZS2_DIRECT) {
            cx<double> bo[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
            if (!pad && !zero_mode) {
                const cx<double> sv[3] = {mk((double)vin[0][s].x, (double)vin[0][s].y),
                                          mk((double)vin[1][s].x, (double)vin[1][s].y),
                                          mk((double)vin[2][s].x, (double)vin[2][s].y)};
                kop_vector(sv, kk, zp.ks[kx], kyv, zp.ks[kz], zp.coeff, zp.modif, bo);
            }
#pragma unroll
            for (int j = 0; j < 3; j++) vin[j][s] = mk((T)bo[j].x, (T)bo[j].y);
        }
    }
    // inverse z transform of output component j
    auto inverse = [&](int j, auto held) {
        constexpr int CH = decltype(held)::value;
        __syncthreads();  // previous transform done with lds
#pragma unroll
        for (int s = 0; s < 8; s++) lds[lines[s] * LDL + poss[s]] = vin[CH][s];
        __syncthreads();
        cx<T> *dst = out + j * NK + off;
        auto ld2 = [&](int line, int row) { return lds[line * LDL + row]; };
        auto st2 = [&](int line, int row, cx<T> v, int) { dst[(int64_t)row * stride + line] = v; };
        reg_fft<T, L, TX, true, 1, false, 1>(lds, twl, +1, ld2, st2);
    };
    inverse(0, std::integral_constant<int, 0>{});
    if constexpr (NOUT == 3) {
        inverse(1, std::integral_constant<int, 1>{});
        inverse(2, std::integral_constant<int, 2>{});
    }
}

// the vector and scalar operators keep 3 x 8 values per thread in registers:
// at most 512 threads per workgroup there; Poisson takes 512 too up to L = 512
// (three workgroups per CU instead of one, 4 % faster at L = 512)
// threads per workgroup of the scalar-projection and vector z solves: the
// scalar one (150+ registers, one 512-thread workgroup per CU) runs three
// 256-thread workgroups per CU (0.90 against 1.02 ms at 512^3, FP32); the
// vector one measured slower that way (1.75 against 1.34 ms)
#ifndef ZS3_THREADS
#define ZS3_THREADS 256
#endif
#ifndef ZS2_THREADS
#define ZS2_THREADS 512
#endif
template <typename T, int L, int OP> constexpr int zs_tx()
{
    constexpr bool narrow = OP != ZOP_POISSON || L <= 512;
    constexpr int tx = (narrow && col_tx<T, L>() * L / 8 > 512) ? 4096 / L : col_tx<T, L>();
    constexpr int cap = OP == ZOP_SCALAR3 ? ZS3_THREADS : (OP == ZOP_VECTOR ? ZS2_THREADS : 1 << 30);
    return (tx * L / 8 > cap && cap * 8 / L >= 1) ? cap * 8 / L : tx;
}

template <typename T, int N, int OP>
static void launch_zsolve(const cx<T> *in, cx<T> *out, int P, int64_t NK, const cx<T> *tw, const ZParams &zp,
                          hipStream_t s)
{
    constexpr int TX = zs_tx<T, N, OP>();
    const int64_t nb = (int64_t)(P / TX) * N;
    hipLaunchKernelGGL((k_zsolve<T, N, TX, OP>), dim3((unsigned)nb), dim3(TX * N / 8), 0, s, in, out, P, NK, tw, zp);
}

template <typename T> static const cx<T> *twiddles_of(gev_ctx *c)
{
    if constexpr (sizeof(T) == 8) return c->tw64;
    else return c->tw32;
}

// forward x and y passes (r2c without the z pass) / inverse y and x passes
template <typename T, int N> static void fwd_xy(gev_ctx *c, const T *in, cx<T> *out)
{
    launch_r2c<T, N>(in, out, c->P, (int64_t)N * N, twiddles_of<T>(c), c->stream);
    launch_col<T, N>(out, c->P, 1, -1, twiddles_of<T>(c), c->stream);
}
template <typename T, int N> static void inv_yx(gev_ctx *c, cx<T> *in, T *out)
{
    launch_col<T, N>(in, c->P, 1, +1, twiddles_of<T>(c), c->stream);
    launch_c2r<T, N>(in, out, c->P, (int64_t)N * N, twiddles_of<T>(c), c->stream);
}

template <typename T, int N>
static void zsolve_chain(gev_ctx *c, int op, const T *rin, int nin, cx<T> *ftin, cx<T> *ftout, T *rout, int nout,
                         const ZParams &zp)
{
    if (op == ZOP_SCALAR) {
        // x passes, then the y pass that forms the three groups of projectFTscalar
        for (int i = 0; i < nin; i++)
            launch_r2c<T, N>(rin + i * c->M, ftin + i * c->NK, c->P, (int64_t)N * N, twiddles_of<T>(c), c->stream);
        launch_col_comb<T, N>(ftin, c->P, c->NK, twiddles_of<T>(c), zp, c->stream);
        launch_zsolve<T, N, ZOP_SCALAR3>(ftin, ftout, c->P, c->NK, twiddles_of<T>(c), zp, c->stream);
    } else {
        for (int i = 0; i < nin; i++) fwd_xy<T, N>(c, rin + i * c->M, ftin + i * c->NK);
        if (op == ZOP_POISSON)
            launch_zsolve<T, N, ZOP_POISSON>(ftin, ftout, c->P, c->NK, twiddles_of<T>(c), zp, c->stream);
        else
            launch_zsolve<T, N, ZOP_VECTOR>(ftin, ftout, c->P, c->NK, twiddles_of<T>(c), zp, c->stream);
    }
    for (int j = 0; j < nout; j++) inv_yx<T, N>(c, ftout + j * c->NK, rout + j * c->M);
}

template <typename T>
static int zsolve_dispatch(gev_ctx *c, int op, const T *rin, int nin, cx<T> *ftin, cx<T> *ftout, T *rout, int nout,
                           const ZParams &zp)
{
    switch (c->N) {
    case 64: zsolve_chain<T, 64>(c, op, rin, nin, ftin, ftout, rout, nout, zp); break;
    case 128: zsolve_chain<T, 128>(c, op, rin, nin, ftin, ftout, rout, nout, zp); break;
    case 256: zsolve_chain<T, 256>(c, op, rin, nin, ftin, ftout, rout, nout, zp); break;
    case 512: zsolve_chain<T, 512>(c, op, rin, nin, ftin, ftout, rout, nout, zp); break;
    case 1024: zsolve_chain<T, 1024>(c, op, rin, nin, ftin, ftout, rout, nout, zp); break;
    default: gev_set_error("fused z-pass solve needs 64 <= N <= 1024"); return GEV_ERR_ARG;
    }
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

bool zsolve_supported(const gev_ctx *c) { return c->N >= 64 && c->N <= 1024; }

// pot = c2r(poisson(r2c(src))) with the k=0 mode optionally replaced first
template <typename T>
int solve_poisson_fused_impl(gev_ctx *c, const T *src, T *pot, cx<T> *work, double coeff, double modif, int k0set,
                             double k0re)
{
    ZParams zp{c->gk2, c->kshift, coeff / -((double)c->N * (double)c->N * (double)c->N), modif, k0re, 0.0, k0set,
               nullptr};
    return zsolve_dispatch<T>(c, ZOP_POISSON, src, 1, work, work, pot, 1, zp);
}

// pot = c2r(poisson(r2c(prepareFTsource_T00(phi, chi, src)))) with the source
// formed inside the x pass; its workgroup sums go to ctx->red + 64 and are
// reduced to ctx->red[0] (source) and ctx->red[1] (result) before the z pass,
// which takes the k=0 mode from ctx->red[1] when zp.k0set
template <typename T, int N>
static void poisson_T00_chain(gev_ctx *c, const SrcT00<T, N> &src, T *pot, cx<T> *work, const ZParams &zp)
{
    constexpr int NL = xlines<T, N>();
    const int64_t nl = (int64_t)N * N, nb = (nl + NL - 1) / NL;
    const cx<T> *tw = twiddles_of<T>(c);
    launch_r2c_src<T, N>(src, work, c->P, nl, tw, c->stream, c->red + 64);
    launch_col<T, N>(work, c->P, 1, -1, tw, c->stream);
    reduce_partials(c, c->red + 64, (int)nb, 2, c->red);
    launch_zsolve<T, N, ZOP_POISSON>(work, work, c->P, c->NK, tw, zp, c->stream);
    inv_yx<T, N>(c, work, pot);
}

template <typename T>
int solve_poisson_T00_fused_impl(gev_ctx *c, const T *phi, const T *chi, const T *src, double bg, double c1,
                                 double c2, double c3, T *pot, cx<T> *work, double coeff, double modif, int k0set)
{
    ZParams zp{c->gk2, c->kshift, coeff / -((double)c->N * (double)c->N * (double)c->N), modif, 0.0, 0.0, k0set,
               k0set ? c->red + 1 : nullptr};
    switch (c->N) {
#define X(n)                                                                                                     \
    case n:                                                                                                      \
        poisson_T00_chain<T, n>(c, SrcT00<T, n>{phi, chi, src, (T)bg, (T)c1, (T)c2, (T)c3, {0.0, 0.0}}, pot, work, \
                                zp);                                                                             \
        break;
        X(128) X(256) X(512) X(1024)
#undef X
    default: gev_set_error("prepareFTsource T00 inside the Poisson solve needs 128 <= N <= 1024"); return GEV_ERR_ARG;
    }
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

// chi = c2r(projectFTscalar(r2c(Sij)))
template <typename T> int solve_chi_fused_impl(gev_ctx *c, const T *Sij, T *chi, cx<T> *SijFT, cx<T> *chiFT)
{
    ZParams zp{c->gk2, c->kshift, 0.0, 0.0, 0.0, 0.0, 0, nullptr};
    return zsolve_dispatch<T>(c, ZOP_SCALAR, Sij, 6, SijFT, chiFT, chi, 1, zp);
}

// chi from Tij: the x passes compute Sij = prepareFTsource(phi, Tij, coeff)
// on the fly (N >= 128; the caller runs the unfused sequence below that)
template <typename T, int N>
static void chi_from_Tij_chain(gev_ctx *c, const T *phi, const T *Tij, T coeff, cx<T> *SijFT, cx<T> *chiFT, T *chi)
{
    const int64_t nl = (int64_t)N * N;
    const cx<T> *tw = twiddles_of<T>(c);
    launch_r2c_multi<T, N, 6>(SrcSij6<T, N>{phi, Tij, c->M, coeff}, SijFT, c->P, c->NK, nl, tw, c->stream);
    ZParams zp{c->gk2, c->kshift, 0.0, 0.0, 0.0, 0.0, 0, nullptr};
    launch_col_comb<T, N>(SijFT, c->P, c->NK, tw, zp, c->stream);
    launch_zsolve<T, N, ZOP_SCALAR3>(SijFT, chiFT, c->P, c->NK, tw, zp, c->stream);
    inv_yx<T, N>(c, chiFT, chi);
}

bool chi_from_Tij_supported(const gev_ctx *c) { return c->N >= 128 && c->N <= 1024; }

template <typename T>
int solve_chi_Tij_fused_impl(gev_ctx *c, const T *phi, const T *Tij, double coeff, T *chi, cx<T> *SijFT, cx<T> *chiFT)
{
    switch (c->N) {
    case 128: chi_from_Tij_chain<T, 128>(c, phi, Tij, (T)coeff, SijFT, chiFT, chi); break;
    case 256: chi_from_Tij_chain<T, 256>(c, phi, Tij, (T)coeff, SijFT, chiFT, chi); break;
    case 512: chi_from_Tij_chain<T, 512>(c, phi, Tij, (T)coeff, SijFT, chiFT, chi); break;
    case 1024: chi_from_Tij_chain<T, 1024>(c, phi, Tij, (T)coeff, SijFT, chiFT, chi); break;
    default: gev_set_error("chi from Tij needs 128 <= N <= 1024"); return GEV_ERR_ARG;
    }
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

// Bi = c2r(projectFTvector(r2c(Bi))), in place on Bi
template <typename T> int solve_vector_fused_impl(gev_ctx *c, T *Bi, cx<T> *BiFT, double coeff, double modif)
{
    ZParams zp{c->gk2, c->kshift, coeff, modif, 0.0, 0.0, 0, nullptr};
    return zsolve_dispatch<T>(c, ZOP_VECTOR, Bi, 3, BiFT, BiFT, Bi, 3, zp);
}

// the same from T0i link sums: the x passes apply the link factor with phi
template <typename T, int N>
static void vector_from_links_chain(gev_ctx *c, T *Bi, const T *phi, cx<T> *BiFT, const ZParams &zp)
{
    const int64_t nl = (int64_t)N * N, M = c->M;
    const cx<T> *tw = twiddles_of<T>(c);
    auto xy = [&](const auto &src, int i) {
        launch_r2c_src<T, N>(src, BiFT + i * c->NK, c->P, nl, tw, c->stream);
        launch_col<T, N>(BiFT + i * c->NK, c->P, 1, -1, tw, c->stream);
    };
    xy(SrcT0iLink<T, N, 0>{phi, Bi}, 0);
    xy(SrcT0iLink<T, N, 1>{phi, Bi + M}, 1);
    xy(SrcT0iLink<T, N, 2>{phi, Bi + 2 * M}, 2);
    launch_zsolve<T, N, ZOP_VECTOR>(BiFT, BiFT, c->P, c->NK, tw, zp, c->stream);
    for (int j = 0; j < 3; j++) inv_yx<T, N>(c, BiFT + j * c->NK, Bi + j * M);
}

// the link factor alone, in place (meshes below the fused chain's range)
template <typename T> __global__ void k_t0i_link_factor(T *Bi, const T *__restrict__ phi, int N, int64_t M)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int lg = __builtin_ctz(N), M1 = N - 1;
    const int x = (int)(i & M1), y = (int)((i >> lg) & M1), z = (int)(i >> (2 * lg));
    const T p = phi[i];
    Bi[i] *= (T)1 + p + phi[site(N, x + 1, y, z)];
    Bi[M + i] *= (T)1 + p + phi[site(N, x, y + 1, z)];
    Bi[2 * M + i] *= (T)1 + p + phi[site(N, x, y, z + 1)];
}

template <typename T> int t0i_link_factor_impl(gev_ctx *c, T *Bi, const T *phi)
{
    hipLaunchKernelGGL((k_t0i_link_factor<T>), dim3((unsigned)((c->M + 255) / 256)), dim3(256), 0, c->stream, Bi, phi,
                       c->N, c->M);
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

template <typename T>
int solve_vector_links_fused_impl(gev_ctx *c, T *Bi, const T *phi, cx<T> *BiFT, double coeff, double modif)
{
    ZParams zp{c->gk2, c->kshift, coeff, modif, 0.0, 0.0, 0, nullptr};
    switch (c->N) {
    case 128: vector_from_links_chain<T, 128>(c, Bi, phi, BiFT, zp); break;
    case 256: vector_from_links_chain<T, 256>(c, Bi, phi, BiFT, zp); break;
    case 512: vector_from_links_chain<T, 512>(c, Bi, phi, BiFT, zp); break;
    case 1024: vector_from_links_chain<T, 1024>(c, Bi, phi, BiFT, zp); break;
    default: gev_set_error("B from T0i links needs 128 <= N <= 1024"); return GEV_ERR_ARG;
    }
    GEV_HIP(hipGetLastError());
    return GEV_OK;
}

#define INSTZ(T)                                                                                       \
    template int solve_chi_Tij_fused_impl<T>(gev_ctx *, const T *, const T *, double, T *, cx<T> *, cx<T> *); \
    template int solve_poisson_fused_impl<T>(gev_ctx *, const T *, T *, cx<T> *, double, double, int, double); \
    template int solve_chi_fused_impl<T>(gev_ctx *, const T *, T *, cx<T> *, cx<T> *);                  \
    template int solve_vector_fused_impl<T>(gev_ctx *, T *, cx<T> *, double, double);                    \
    template int solve_vector_links_fused_impl<T>(gev_ctx *, T *, const T *, cx<T> *, double, double);    \
    template int t0i_link_factor_impl<T>(gev_ctx *, T *, const T *);                                      \
    template int solve_poisson_T00_fused_impl<T>(gev_ctx *, const T *, const T *, const T *, double, double, double, \
                                                 double, T *, cx<T> *, double, double, int);
INSTZ(double)
INSTZ(float)
#undef INSTZ

template int fft_r2c_impl<double>(gev_ctx *, const double *, cx<double> *, int, const cx<double> *);
template int fft_r2c_impl<float>(gev_ctx *, const float *, cx<float> *, int, const cx<float> *);
template int fft_c2r_impl<double>(gev_ctx *, cx<double> *, double *, int, const cx<double> *);
template int fft_c2r_impl<float>(gev_ctx *, cx<float> *, float *, int, const cx<float> *);

}  // namespace gev

// gev_fft_engine.hpp -- register-staged Stockham FFT for one workgroup.
//
// A workgroup transforms NL lines of length L (power of two, L >= 64) with
// NL*L/8 threads; every thread owns one radix-8 butterfly slot per stage
// (two radix-4 or four radix-2 butterflies in a trailing stage).  The first
// stage reads its operands straight from a caller functor (global memory or
// LDS) into registers, intermediate stages exchange through one LDS buffer,
// and the last stage hands its results to a store functor.  For L = 512 that
// is two LDS round trips per transform instead of four.
#pragma once
#include "gev_common.hpp"

namespace gev {

template <typename T> __device__ __forceinline__ void dft2(cx<T> *v)
{
    cx<T> t = v[0] - v[1];
    v[0] = v[0] + v[1];
    v[1] = t;
}

template <typename T> __device__ __forceinline__ void dft4(cx<T> *v, int dir)
{
    cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2], a2 = v[1] + v[3], a3 = mul_i(v[1] - v[3], dir);
    v[0] = a0 + a2;
    v[2] = a0 - a2;
    v[1] = a1 + a3;
    v[3] = a1 - a3;
}

template <typename T> __device__ __forceinline__ void dft8(cx<T> *v, int dir)
{
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e, dir);
    dft4(o, dir);
    const T r = (T)0.707106781186547524400844362104849039;
    const T d = (T)dir;
    cx<T> t0 = o[0];
    cx<T> t1 = mk((o[1].x - d * o[1].y) * r, (o[1].y + d * o[1].x) * r);
    cx<T> t2 = mul_i(o[2], dir);
    cx<T> t3 = mk((-o[3].x - d * o[3].y) * r, (-o[3].y + d * o[3].x) * r);
    v[0] = e[0] + t0; v[4] = e[0] - t0;
    v[1] = e[1] + t1; v[5] = e[1] - t1;
    v[2] = e[2] + t2; v[6] = e[2] - t2;
    v[3] = e[3] + t3; v[7] = e[3] - t3;
}

template <typename T, int R> __device__ __forceinline__ void dftR(cx<T> *v, int dir)
{
    if constexpr (R == 8) dft8(v, dir);
    else if constexpr (R == 4) dft4(v, dir);
    else dft2(v);
}

// LF (line-fast): consecutive threads walk lines first (coalesced when lines
// are adjacent columns in memory); otherwise they walk along a line.
// The store functor is called as st(line, pos, value, slot) with slot in
// [0, 8) a compile-time constant per call site: the thread's slot-th output,
// always at the same (line, pos) for a given thread and L, NL, LF.
// twl[m * TWS] = exp(-2 pi i m / L).  ST_LDS: the store functor writes the
// same LDS buffer, so the last stage waits for all reads first.
template <typename T, int L, int NL, bool LF, int TWS, bool ST_LDS, int NS, class LD, class ST>
__device__ __forceinline__ void reg_fft(cx<T> *lds, const cx<T> *twl, int dir, LD &ld, ST &st)
{
    constexpr int REM = L / NS;
    constexpr int R = REM >= 8 ? 8 : REM;
    constexpr int BPT = 8 / R;
    constexpr int NB = L / R;
    constexpr bool FIRST = (NS == 1), LAST = (NS * R == L);
    constexpr int LDL = L + 1;
    static_assert(L >= 64, "reg_fft needs L >= 64");
    const int t = threadIdx.x;
    const int line = LF ? t % NL : t / (L / 8);
    const int jt = LF ? t / NL : t % (L / 8);
    cx<T> v[BPT][R];
#pragma unroll
    for (int u = 0; u < BPT; u++) {
        const int j = jt * BPT + u;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if constexpr (FIRST) v[u][r] = ld(line, j + r * NB);
            else v[u][r] = lds[line * LDL + j + r * NB];
        }
        if constexpr (NS > 1) {
            const int k = j % NS;
#pragma unroll
            for (int r = 1; r < R; r++) {
                cx<T> w = twl[(r * k * (L / (NS * R))) * TWS];
                if (dir > 0) w.y = -w.y;
                v[u][r] = v[u][r] * w;
            }
        }
        dftR<T, R>(v[u], dir);
    }
    if constexpr (LAST) {
        if constexpr (ST_LDS) __syncthreads();
#pragma unroll
        for (int u = 0; u < BPT; u++) {
            const int j = jt * BPT + u;
#pragma unroll
            for (int r = 0; r < R; r++) st(line, (j / NS) * NS * R + (j % NS) + r * NS, v[u][r], u * R + r);
        }
    } else {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < BPT; u++) {
            const int j = jt * BPT + u;
#pragma unroll
            for (int r = 0; r < R; r++) lds[line * LDL + (j / NS) * NS * R + (j % NS) + r * NS] = v[u][r];
        }
        __syncthreads();
        reg_fft<T, L, NL, LF, TWS, ST_LDS, NS * R, LD, ST>(lds, twl, dir, ld, st);
    }
}

}  // namespace gev

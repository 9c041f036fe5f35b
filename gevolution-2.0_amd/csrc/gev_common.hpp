// gev_common.hpp -- shared device helpers and the library context for the
// MI355X (gfx950) particle-mesh step.  HIP only, written for CDNA4 wave64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/gev_pm.h"

namespace gev {

// ---------------------------------------------------------------- complex --
template <typename T> struct cx { T x, y; };

template <typename T> __host__ __device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b)
{
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T> __host__ __device__ __forceinline__ cx<T> conj(cx<T> a) { return mk(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse): dir = -1 / +1
template <typename T> __host__ __device__ __forceinline__ cx<T> mul_i(cx<T> a, int dir)
{
    return dir < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// ----------------------------------------------------------- lattice index --
__host__ __device__ __forceinline__ int64_t site(int N, int x, int y, int z)
{
    const int m = N - 1;
    return (int64_t)(x & m) + (int64_t)N * ((int64_t)(y & m) + (int64_t)N * (int64_t)(z & m));
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Remap so that each XCD walks a contiguous range of logical tiles,
// keeping neighbouring tiles (which share halo data) inside one XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nblocks)
{
    const int64_t nx = 8;
    if (nblocks < nx * 2) return b;
    int64_t per = nblocks / nx, rem = nblocks % nx;
    int64_t xcd = b % nx, loc = b / nx;
    // blocks beyond per*nx (the remainder) keep their index
    if (b >= per * nx) return b;
    (void)rem;
    return xcd * per + loc;
}

// -------------------------------------------------------------- atomics ----
// max over non-negative doubles via integer max on the IEEE bit pattern
// (monotone for x >= 0); order independent, hence deterministic.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *addr, double v)
{
    atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

// wave64 reductions
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
// max over the wave of 32-bit unsigned values (also the max of non-negative
// floats through their bit patterns), through DPP row shifts and broadcasts;
// the result is uniform
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
    // row_shr:1,2,4,8 leave each row's maximum in its lane 15 (lanes shifted
    // in from outside the row read 0); row_bcast:15 and row_bcast:31 carry
    // rows 0-2 into lane 63
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xf, 0xf, true));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace gev

// ------------------------------------------------------------ context ------
struct gev_ctx {
    int N = 0;
    int prec = 0;          // GEV_FP64 or GEV_MIXED
    int P = 0;             // kx pitch of Fourier arrays (>= N/2+1)
    int64_t M = 0;         // N^3
    int64_t NK = 0;        // P*N*N
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // second stream of gev_step (default priority, like the first): the T0i
    // deposit and the B solve run on it beside the chi solve
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    int device = 0;
    int ncu = 0;           // compute units of the device
    // twiddles exp(-2 pi i m / N), m in [0,N): double2 and float2 copies
    gev::cx<double> *tw64 = nullptr;
    gev::cx<float> *tw32 = nullptr;
    // k-space tables: gridk2[i] (double) and kshift[i] (complex double), i in [0,N)
    double *gk2 = nullptr;
    gev::cx<double> *kshift = nullptr;
    // reductions: [0..7] double accumulators, [8..15] uint64 max slots
    double *red = nullptr;
    double *red_host = nullptr;  // pinned
    // scan scratch
    int32_t *scan_blocks = nullptr;
    int64_t scan_blocks_n = 0;
    // optional per-phase timing of gev_step (events recorded on the stream)
    static constexpr int MAX_MARKS = 24;
    bool timing = false;
    int nmarks = 0;
    hipEvent_t marks[MAX_MARKS] = {};
    int mark_id[MAX_MARKS] = {};
    std::string err;
};

struct gev_pcls {
    gev_ctx *ctx = nullptr;
    int64_t n = 0, cap = 0;
    double mass = 0.0;
    // current arrays (SoA).  FP64 form: pos = absolute box coordinates
    // (double); mixed form: pos = cell-relative offsets in [0,1) (float).
    void *pos[3] = {nullptr, nullptr, nullptr};
    void *q[3] = {nullptr, nullptr, nullptr};
    int64_t *id = nullptr;
    int32_t *cell = nullptr;
    // ping-pong partners used by the re-binning scatter
    void *pos2[3] = {nullptr, nullptr, nullptr};
    void *q2[3] = {nullptr, nullptr, nullptr};
    int64_t *id2 = nullptr;
    int32_t *cell2 = nullptr;
    int32_t *rank = nullptr;        // slot of each particle inside its cell
    int32_t *perm = nullptr;        // re-binning: perm[new slot] = old slot
    int32_t *cell_start = nullptr;  // M+1 offsets, valid when sorted
    int32_t *cell_count = nullptr;  // M counters
    bool sorted = false;
    bool counts_zero = false;       // cell_count is all zero (the scan clears what it consumes)
    // mixed form, N >= 16: the largest |q|^2 (float bits) of the particles in
    // each block of 8^3 cells, written by the particle updates and read by the
    // fixed-point deposits as their per-tile bound; valid from an update until
    // anything else writes positions or momenta
    uint32_t *qbound = nullptr;
    bool qbound_valid = false;
};

namespace gev {
constexpr int QB_LG = 3;  // log2 of the block edge of gev_pcls::qbound
}

#define GEV_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            gev_set_error(std::string(#call) + ": " + hipGetErrorString(e_));                \
            return GEV_ERR_HIP;                                                              \
        }                                                                                    \
    } while (0)

void gev_set_error(const std::string &msg);

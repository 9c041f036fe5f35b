// gev_pcommon.hpp -- device helpers shared by the particle kernels: the
// container view, cell/offset arithmetic, periodic moves, field accessors
// (global memory or an LDS-staged box), CIC interpolation / gradients and the
// per-particle functor physics of gevolution.hpp.
#pragma once
#include "gev_common.hpp"

namespace gev {

constexpr int PT_NT = 256;

// ---------------------------------------------------------------- views ---
template <typename T> struct PV {
    T *p0, *p1, *p2;  // FP64: absolute box coords; FP32: cell offsets
    T *q0, *q1, *q2;
    int64_t *id;
    int32_t *cell;
    const int32_t *start;
    int64_t n;
    int N, lg;
};

template <typename T> static PV<T> make_pv(gev_pcls *p)
{
    PV<T> v;
    v.p0 = (T *)p->pos[0]; v.p1 = (T *)p->pos[1]; v.p2 = (T *)p->pos[2];
    v.q0 = (T *)p->q[0]; v.q1 = (T *)p->q[1]; v.q2 = (T *)p->q[2];
    v.id = p->id;
    v.cell = p->cell;
    v.start = p->cell_start;
    v.n = p->n;
    v.N = p->ctx->N;
    v.lg = __builtin_ctz(p->ctx->N);
    return v;
}

__device__ __forceinline__ void cell_coords(int32_t c, int N, int lg, int &x, int &y, int &z)
{
    x = c & (N - 1);
    y = (c >> lg) & (N - 1);
    z = c >> (2 * lg);
}

// ref_dist of particle i in cell (x,y,z): FP64 form (p - c dx)/dx, exact for
// N a power of two; FP32 form stores it directly.
template <typename T> __device__ __forceinline__ void ref_dist(const PV<T> &v, int64_t i, int x, int y, int z, T r[3]);
template <> __device__ __forceinline__ void ref_dist<double>(const PV<double> &v, int64_t i, int x, int y, int z,
                                                             double r[3])
{
    // dividing by dx = 1/N is exactly multiplying by N (N a power of two)
    const double dx = 1.0 / (double)v.N, Nd = (double)v.N;
    r[0] = (v.p0[i] - (double)x * dx) * Nd;
    r[1] = (v.p1[i] - (double)y * dx) * Nd;
    r[2] = (v.p2[i] - (double)z * dx) * Nd;
}
template <> __device__ __forceinline__ void ref_dist<float>(const PV<float> &v, int64_t i, int x, int y, int z,
                                                            float r[3])
{
    r[0] = v.p0[i];
    r[1] = v.p1[i];
    r[2] = v.p2[i];
}

// the same from coordinates already loaded (x[] as stored in the container)
__device__ __forceinline__ void cell_rel(const double x[3], int cx, int cy, int cz, int N, double r[3])
{
    const double dx = 1.0 / (double)N, Nd = (double)N;
    r[0] = (x[0] - (double)cx * dx) * Nd;
    r[1] = (x[1] - (double)cy * dx) * Nd;
    r[2] = (x[2] - (double)cz * dx) * Nd;
}
__device__ __forceinline__ void cell_rel(const float x[3], int, int, int, int, float r[3])
{
    r[0] = x[0];
    r[1] = x[1];
    r[2] = x[2];
}

__device__ __forceinline__ double wrap_pos(double p)
{
    if (p >= 1.0) p -= 1.0;
    else if (p < 0.0) p += 1.0;
    if (p >= 1.0) p = 0.0;
    if (p < 0.0) p = 0.0;
    return p;
}

__device__ __forceinline__ int cell1(double p, int N)
{
    int c = (int)floor(p * (double)N);
    return c < 0 ? 0 : (c >= N ? N - 1 : c);
}

// move particle i by d (box units) and recompute its cell
template <typename T> __device__ __forceinline__ void move_by(const PV<T> &v, int64_t i, int x, int y, int z,
                                                              const double d[3]);
template <> __device__ __forceinline__ void move_by<double>(const PV<double> &v, int64_t i, int x, int y, int z,
                                                            const double d[3])
{
    double p0 = wrap_pos(v.p0[i] + d[0]), p1 = wrap_pos(v.p1[i] + d[1]), p2 = wrap_pos(v.p2[i] + d[2]);
    v.p0[i] = p0; v.p1[i] = p1; v.p2[i] = p2;
    const int N = v.N;
    v.cell[i] = (int32_t)((uint32_t)cell1(p0, N) + (uint32_t)N * ((uint32_t)cell1(p1, N) + (uint32_t)N * (uint32_t)cell1(p2, N)));
}
__device__ __forceinline__ void off_step(float &r, int &c, double dcells, int N)
{
    double rn = (double)r + dcells;
    double sh = floor(rn);
    rn -= sh;
    float rf = (float)rn;
    int s = (int)sh;
    if (rf >= 1.0f) { rf = 0.0f; s += 1; }
    r = rf;
    c = (c + s) & (N - 1);
}
template <> __device__ __forceinline__ void move_by<float>(const PV<float> &v, int64_t i, int x, int y, int z,
                                                           const double d[3])
{
    const int N = v.N;
    float r0 = v.p0[i], r1 = v.p1[i], r2 = v.p2[i];
    off_step(r0, x, d[0] * N, N);
    off_step(r1, y, d[1] * N, N);
    off_step(r2, z, d[2] * N, N);
    v.p0[i] = r0; v.p1[i] = r1; v.p2[i] = r2;
    v.cell[i] = (int32_t)((uint32_t)x + (uint32_t)N * ((uint32_t)y + (uint32_t)N * (uint32_t)z));
}

// tile -> (x0,y0,z0), XCD-aware
template <int TX, int TY, int TZ>
__device__ __forceinline__ void tile_origin(int N, int &x0, int &y0, int &z0)
{
    const int ntx = N / TX, nty = N / TY;
    int64_t b = xcd_remap(blockIdx.x, gridDim.x);
    x0 = (int)(b % ntx) * TX;
    y0 = (int)((b / ntx) % nty) * TY;
    z0 = (int)(b / ((int64_t)ntx * nty)) * TZ;
}

// ---------------------------------------------------------- kick / drift --
template <typename T> struct FieldAt {
    const T *f;
    int N, x, y, z;
    __device__ __forceinline__ T operator()(int dx, int dy, int dz) const { return f[site(N, x + dx, y + dy, z + dz)]; }
};

// field staged in LDS over a box starting at (bx,by,bz) of extent SX x SY (x fastest);
// o is the box-local index of the particle's cell
template <typename T> struct FieldLds {
    const T *f;
    int o, SX, SXY;
    __device__ __forceinline__ T operator()(int dx, int dy, int dz) const { return f[o + dx + SX * dy + SXY * dz]; }
};

template <typename T, class FA> __device__ __forceinline__ T cic(const FA &F_, const T r[3])
{
    const FA &Fv = F_;
    T s = Fv(0, 0, 0) * ((T)1 - r[0]) * ((T)1 - r[1]) * ((T)1 - r[2]);
    s += Fv(1, 0, 0) * r[0] * ((T)1 - r[1]) * ((T)1 - r[2]);
    s += Fv(0, 1, 0) * ((T)1 - r[0]) * r[1] * ((T)1 - r[2]);
    s += Fv(1, 1, 0) * r[0] * r[1] * ((T)1 - r[2]);
    s += Fv(0, 0, 1) * ((T)1 - r[0]) * ((T)1 - r[1]) * r[2];
    s += Fv(1, 0, 1) * r[0] * ((T)1 - r[1]) * r[2];
    s += Fv(0, 1, 1) * ((T)1 - r[0]) * r[1] * r[2];
    s += Fv(1, 1, 1) * r[0] * r[1] * r[2];
    return s;
}

// GRADIENT_ORDER = 2 (gevolution.hpp:930-941)
template <typename T, class FA> __device__ __forceinline__ void grad2(const FA &F, const T r[3], T g[3])
{
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int ex = (i == 0), ey = (i == 1), ez = (i == 2);
        T s;
        s = (T)0.5 * ((T)1 - r[0]) * ((T)1 - r[1]) * ((T)1 - r[2]) * (F(ex, ey, ez) - F(-ex, -ey, -ez));
        s += (T)0.5 * r[0] * ((T)1 - r[1]) * ((T)1 - r[2]) * (F(ex + 1, ey, ez) - F(-ex + 1, -ey, -ez));
        s += (T)0.5 * ((T)1 - r[0]) * r[1] * ((T)1 - r[2]) * (F(ex, ey + 1, ez) - F(-ex, -ey + 1, -ez));
        s += (T)0.5 * r[0] * r[1] * ((T)1 - r[2]) * (F(ex + 1, ey + 1, ez) - F(-ex + 1, -ey + 1, -ez));
        s += (T)0.5 * ((T)1 - r[0]) * ((T)1 - r[1]) * r[2] * (F(ex, ey, ez + 1) - F(-ex, -ey, -ez + 1));
        s += (T)0.5 * r[0] * ((T)1 - r[1]) * r[2] * (F(ex + 1, ey, ez + 1) - F(-ex + 1, -ey, -ez + 1));
        s += (T)0.5 * ((T)1 - r[0]) * r[1] * r[2] * (F(ex, ey + 1, ez + 1) - F(-ex, -ey + 1, -ez + 1));
        s += (T)0.5 * r[0] * r[1] * r[2] * (F(ex + 1, ey + 1, ez + 1) - F(-ex + 1, -ey + 1, -ez + 1));
        g[i] = s;
    }
}

// CIC edge gradient (gevolution.hpp:950-964; ic_basic.hpp:73-84)
template <typename T, class FA> __device__ __forceinline__ void grad1(const FA &F, const T r[3], T g[3])
{
    g[0] = ((T)1 - r[1]) * ((T)1 - r[2]) * (F(1, 0, 0) - F(0, 0, 0));
    g[1] = ((T)1 - r[0]) * ((T)1 - r[2]) * (F(0, 1, 0) - F(0, 0, 0));
    g[2] = ((T)1 - r[0]) * ((T)1 - r[1]) * (F(0, 0, 1) - F(0, 0, 0));
    g[0] += r[1] * ((T)1 - r[2]) * (F(1, 1, 0) - F(0, 1, 0));
    g[1] += r[0] * ((T)1 - r[2]) * (F(1, 1, 0) - F(1, 0, 0));
    g[2] += r[0] * ((T)1 - r[1]) * (F(1, 0, 1) - F(1, 0, 0));
    g[0] += ((T)1 - r[1]) * r[2] * (F(1, 0, 1) - F(0, 0, 1));
    g[1] += ((T)1 - r[0]) * r[2] * (F(0, 1, 1) - F(0, 0, 1));
    g[2] += ((T)1 - r[0]) * r[1] * (F(0, 1, 1) - F(0, 1, 0));
    g[0] += r[1] * r[2] * (F(1, 1, 1) - F(0, 1, 1));
    g[1] += r[0] * r[2] * (F(1, 1, 1) - F(1, 0, 1));
    g[2] += r[0] * r[1] * (F(1, 1, 1) - F(1, 1, 0));
}


// ------------------------------------------------------ functor physics --
// update_q, gevolution.hpp:888-1011 (GRADIENT_ORDER 2, no anisotropic
// expansion).  Updates vv in place; returns the functor value v2 / a^2.
template <typename T, class FA>
__device__ __forceinline__ T phys_update_q(const T r[3], T vv[3], const FA &Fp, const FA &Fc, const FA &B0,
                                           const FA &B1, const FA &B2, int nfield, T a, T dtau, T bcoeff, T Nn)
{
    T v2 = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
    T e2 = v2 + a * a;
    T phiint = cic<T>(Fp, r);
    phiint /= e2;
    phiint = (T)4 * phiint * v2 + ((T)1 + a * a * phiint) * (v2 + e2) / e2;
    T g[3];
    grad2<T>(Fp, r, g);
    g[0] *= phiint; g[1] *= phiint; g[2] *= phiint;
    if (nfield >= 2) {
        T gc[3];
        grad1<T>(Fc, r, gc);
        g[0] -= gc[0]; g[1] -= gc[1]; g[2] -= gc[2];
    }
    e2 = sqrt(e2);
    if (nfield >= 3) {
        const T o = (T)1;
        T pg[3];
        pg[0] = ((o - r[2]) * (B1(1, 0, 0) - B1(0, 0, 0)) + r[2] * (B1(1, 0, 1) - B1(0, 0, 1))) * vv[1];
        pg[0] += ((o - r[1]) * (B2(1, 0, 0) - B2(0, 0, 0)) + r[1] * (B2(1, 1, 0) - B2(0, 1, 0))) * vv[2];
        pg[0] += (o - r[1]) * (o - r[2]) * ((r[0] - o) * B0(-1, 0, 0) + (o - (T)2 * r[0]) * B0(0, 0, 0) + r[0] * B0(1, 0, 0)) * vv[0];
        pg[0] += r[1] * (o - r[2]) * ((r[0] - o) * B0(-1, 1, 0) + (o - (T)2 * r[0]) * B0(0, 1, 0) + r[0] * B0(1, 1, 0)) * vv[0];
        pg[0] += (o - r[1]) * r[2] * ((r[0] - o) * B0(-1, 0, 1) + (o - (T)2 * r[0]) * B0(0, 0, 1) + r[0] * B0(1, 0, 1)) * vv[0];
        pg[0] += r[1] * r[2] * ((r[0] - o) * B0(-1, 1, 1) + (o - (T)2 * r[0]) * B0(0, 1, 1) + r[0] * B0(1, 1, 1)) * vv[0];

        pg[1] = ((o - r[0]) * (B2(0, 1, 0) - B2(0, 0, 0)) + r[0] * (B2(1, 1, 0) - B2(1, 0, 0))) * vv[2];
        pg[1] += ((o - r[2]) * (B0(0, 1, 0) - B0(0, 0, 0)) + r[2] * (B0(0, 1, 1) - B0(0, 0, 1))) * vv[0];
        pg[1] += (o - r[0]) * (o - r[2]) * ((r[1] - o) * B1(0, -1, 0) + (o - (T)2 * r[1]) * B1(0, 0, 0) + r[1] * B1(0, 1, 0)) * vv[1];
        pg[1] += r[0] * (o - r[2]) * ((r[1] - o) * B1(1, -1, 0) + (o - (T)2 * r[1]) * B1(1, 0, 0) + r[1] * B1(1, 1, 0)) * vv[1];
        pg[1] += (o - r[0]) * r[2] * ((r[1] - o) * B1(0, -1, 1) + (o - (T)2 * r[1]) * B1(0, 0, 1) + r[1] * B1(0, 1, 1)) * vv[1];
        pg[1] += r[0] * r[2] * ((r[1] - o) * B1(1, -1, 1) + (o - (T)2 * r[1]) * B1(1, 0, 1) + r[1] * B1(1, 1, 1)) * vv[1];

        pg[2] = ((o - r[1]) * (B0(0, 0, 1) - B0(0, 0, 0)) + r[1] * (B0(0, 1, 1) - B0(0, 1, 0))) * vv[0];
        pg[2] += ((o - r[0]) * (B1(0, 0, 1) - B1(0, 0, 0)) + r[0] * (B1(1, 0, 1) - B1(1, 0, 0))) * vv[1];
        pg[2] += (o - r[0]) * (o - r[1]) * ((r[2] - o) * B2(0, 0, -1) + (o - (T)2 * r[2]) * B2(0, 0, 0) + r[2] * B2(0, 0, 1)) * vv[2];
        pg[2] += r[0] * (o - r[1]) * ((r[2] - o) * B2(1, 0, -1) + (o - (T)2 * r[2]) * B2(1, 0, 0) + r[2] * B2(1, 0, 1)) * vv[2];
        pg[2] += (o - r[0]) * r[1] * ((r[2] - o) * B2(0, 1, -1) + (o - (T)2 * r[2]) * B2(0, 1, 0) + r[2] * B2(0, 1, 1)) * vv[2];
        pg[2] += r[0] * r[1] * ((r[2] - o) * B2(1, 1, -1) + (o - (T)2 * r[2]) * B2(1, 1, 0) + r[2] * B2(1, 1, 1)) * vv[2];

        g[0] += pg[0] / bcoeff / e2;
        g[1] += pg[1] / bcoeff / e2;
        g[2] += pg[2] / bcoeff / e2;
    }
    v2 = (T)0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        vv[d] -= dtau * e2 * g[d] * Nn;
        v2 += vv[d] * vv[d];
    }
    return v2 / a / a;
}

// update_q_Newton, gevolution.hpp:1051-1118
template <typename T, class FA>
__device__ __forceinline__ T phys_update_q_newton(const T r[3], T q[3], const FA &Fp, const FA &Fc, int nfield, T a,
                                                  T dtau, T Nn)
{
    T g[3];
    grad2<T>(Fp, r, g);
    if (nfield >= 2) {
        T gc[3];
        grad1<T>(Fc, r, gc);
        g[0] -= gc[0]; g[1] -= gc[1]; g[2] -= gc[2];
    }
    T v2 = (T)0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        q[d] = q[d] - dtau * a * g[d] * Nn;
        v2 += q[d] * q[d];
    }
    return v2 / a / a;
}

// update_pos, gevolution.hpp:1162-1232: displacement d (box units) of one drift
template <typename T, class FA>
__device__ __forceinline__ void phys_update_pos(const T r[3], const T q[3], const FA &Fp, const FA &Fc, const FA &B0,
                                                const FA &B1, const FA &B2, int nfield, T a, T bcoeff, double dtau,
                                                double d[3])
{
    T v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    T e2 = v2 + a * a;
    T ph = (T)0, ch = (T)0;
    if (nfield >= 1) ph = cic<T>(Fp, r);
    if (nfield >= 2) ch = cic<T>(Fc, r);
    v2 = ((T)1 + ((T)3 - v2 / e2) * ph - ch) / sqrt(e2);
    T vel[3] = {q[0] * v2, q[1] * v2, q[2] * v2};
    if (nfield >= 3) {
        const T o = (T)1;
        T b[3];
        b[0] = B0(0, 0, 0) * (o - r[1]) * (o - r[2]);
        b[1] = B1(0, 0, 0) * (o - r[0]) * (o - r[2]);
        b[2] = B2(0, 0, 0) * (o - r[0]) * (o - r[1]);
        b[1] += B1(1, 0, 0) * r[0] * (o - r[2]);
        b[2] += B2(1, 0, 0) * r[0] * (o - r[1]);
        b[0] += B0(0, 1, 0) * r[1] * (o - r[2]);
        b[2] += B2(0, 1, 0) * (o - r[0]) * r[1];
        b[0] += B0(0, 0, 1) * (o - r[1]) * r[2];
        b[1] += B1(0, 0, 1) * (o - r[0]) * r[2];
        b[1] += B1(1, 0, 1) * r[0] * r[2];
        b[0] += B0(0, 1, 1) * r[1] * r[2];
        b[2] += B2(1, 1, 0) * r[0] * r[1];
#pragma unroll
        for (int l = 0; l < 3; l++) d[l] = dtau * (double)(vel[l] + b[l] / bcoeff);
    } else {
#pragma unroll
        for (int l = 0; l < 3; l++) d[l] = dtau * (double)vel[l];
    }
}

__device__ __forceinline__ void block_max_to(double v, unsigned long long *slot)
{
    __shared__ double sh[PT_NT / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        for (int j = 1; j < PT_NT / 64; j++) m = fmax(m, sh[j]);
        // all blocks hit one address: skip the atomic unless it can raise the maximum
        unsigned long long cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((unsigned long long)__double_as_longlong(m) > cur) atomic_max_nonneg(slot, m);
    }
}

}  // namespace gev

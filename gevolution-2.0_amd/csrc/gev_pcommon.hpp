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
    int64_t sink;  // the slot past the capacity, read and written by idle lanes
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
    v.sink = std::max<int64_t>(p->cap, 1);
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

// move particle i (stored coordinates x[]) by d (box units); writes the new
// coordinates and cell and returns the cell
template <typename T> __device__ __forceinline__ int32_t move_by(const PV<T> &v, int64_t i, const T x[3], int cx, int cy,
                                                                 int cz, const double d[3]);
template <> __device__ __forceinline__ int32_t move_by<double>(const PV<double> &v, int64_t i, const double x[3], int, int,
                                                               int, const double d[3])
{
    double p0 = wrap_pos(x[0] + d[0]), p1 = wrap_pos(x[1] + d[1]), p2 = wrap_pos(x[2] + d[2]);
    v.p0[i] = p0; v.p1[i] = p1; v.p2[i] = p2;
    const int N = v.N;
    const int32_t c = (int32_t)((uint32_t)cell1(p0, N) + (uint32_t)N * ((uint32_t)cell1(p1, N) + (uint32_t)N * (uint32_t)cell1(p2, N)));
    v.cell[i] = c;
    return c;
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
template <> __device__ __forceinline__ int32_t move_by<float>(const PV<float> &v, int64_t i, const float x[3], int cx, int cy,
                                                              int cz, const double d[3])
{
    const int N = v.N;
    float r0 = x[0], r1 = x[1], r2 = x[2];
    off_step(r0, cx, d[0] * N, N);
    off_step(r1, cy, d[1] * N, N);
    off_step(r2, cz, d[2] * N, N);
    v.p0[i] = r0; v.p1[i] = r1; v.p2[i] = r2;
    const int32_t c = (int32_t)((uint32_t)cx + (uint32_t)N * ((uint32_t)cy + (uint32_t)N * (uint32_t)cz));
    v.cell[i] = c;
    return c;
}

// tile -> (x0,y0,z0), XCD-aware
template <int TX, int TY, int TZ>
__device__ __forceinline__ void tile_origin(int N, int &x0, int &y0, int &z0)
{
    // N and the tile edges are powers of two (N <= 1024): shifts and masks in
    // 32 bits, the same tile order as xcd_remap
    const uint32_t nb = gridDim.x, b0 = blockIdx.x;
    const uint32_t per = nb >> 3;
    const uint32_t b = (nb < 16 || b0 >= (per << 3)) ? b0 : (b0 & 7u) * per + (b0 >> 3);
    const uint32_t lx = (uint32_t)__builtin_ctz((unsigned)(N / TX)), ly = (uint32_t)__builtin_ctz((unsigned)(N / TY));
    x0 = (int)(b & ((1u << lx) - 1u)) * TX;
    y0 = (int)((b >> lx) & ((1u << ly) - 1u)) * TY;
    z0 = (int)(b >> (lx + ly)) * TZ;
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

// reciprocal and square roots: FP32 uses the hardware approximations (1 ulp),
// FP64 the correctly rounded forms.  The functors' divisions by a per-particle
// quantity become one reciprocal and multiplications.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float frsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double frsqrt(double x) { return 1.0 / sqrt(x); }

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
    const T ie2 = frcp(e2);
    T phiint = cic<T>(Fp, r) * ie2;
    phiint = (T)4 * phiint * v2 + ((T)1 + a * a * phiint) * (v2 + e2) * ie2;
    T g[3];
    grad2<T>(Fp, r, g);
    g[0] *= phiint; g[1] *= phiint; g[2] *= phiint;
    if (nfield >= 2) {
        T gc[3];
        grad1<T>(Fc, r, gc);
        g[0] -= gc[0]; g[1] -= gc[1]; g[2] -= gc[2];
    }
    e2 = fsqrt(e2);
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

        const T s = frcp(bcoeff * e2);
        g[0] += pg[0] * s;
        g[1] += pg[1] * s;
        g[2] += pg[2] * s;
    }
    v2 = (T)0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        vv[d] -= dtau * e2 * g[d] * Nn;
        v2 += vv[d] * vv[d];
    }
    return v2 * frcp(a * a);
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
    return v2 * frcp(a * a);
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
    const T ie = frsqrt(e2);
    v2 = ((T)1 + ((T)3 - v2 * ie * ie) * ph - ch) * ie;
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
        const T ibc = frcp(bcoeff);
#pragma unroll
        for (int l = 0; l < 3; l++) d[l] = dtau * (double)(vel[l] + b[l] * ibc);
    } else {
#pragma unroll
        for (int l = 0; l < 3; l++) d[l] = dtau * (double)vel[l];
    }
}

// update_q (gevolution.hpp:888-1011) followed by update_pos (:1162-1232) of
// the same particle, as the fused GR cycle applies them: the field stencils
// are read once and the CIC corner and edge weights computed once for both.
// nf_d = 0 is the free-streaming drift of the unrelaxed phase.  Updates q,
// returns the kick's v^2/a^2 and the drift displacement d (box units).
template <typename T, class FA>
__device__ __forceinline__ T phys_kick_drift(const T r[3], T q[3], const FA &Fp, const FA &Fc, const FA &B0,
                                             const FA &B1, const FA &B2, int nf_k, T a_k, T dtau_k, T bc_k, T Nn,
                                             int nf_d, T a_d, T bc_d, double dtau_d, double d[3])
{
    const T o = (T)1, h = (T)0.5;
    const T r0 = r[0], r1 = r[1], r2 = r[2], s0 = o - r0, s1 = o - r1, s2 = o - r2;
    // edge weights: x-edges (y,z), y-edges (x,z), z-edges (x,y)
    const T ex00 = s1 * s2, ex10 = r1 * s2, ex01 = s1 * r2, ex11 = r1 * r2;
    const T ey00 = s0 * s2, ey10 = r0 * s2, ey01 = s0 * r2, ey11 = r0 * r2;
    const T ez00 = s0 * s1, ez10 = r0 * s1, ez01 = s0 * r1, ez11 = r0 * r1;
    // corner weights w[x][y][z]
    const T w000 = s0 * ex00, w100 = r0 * ex00, w010 = s0 * ex10, w110 = r0 * ex10;
    const T w001 = s0 * ex01, w101 = r0 * ex01, w011 = s0 * ex11, w111 = r0 * ex11;
    auto cic8 = [&](const FA &F) {
        return w000 * F(0, 0, 0) + w100 * F(1, 0, 0) + w010 * F(0, 1, 0) + w110 * F(1, 1, 0) + w001 * F(0, 0, 1) +
               w101 * F(1, 0, 1) + w011 * F(0, 1, 1) + w111 * F(1, 1, 1);
    };
    const T phic = cic8(Fp);

    // ---- kick
    T v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    T e2 = v2 + a_k * a_k;
    const T ie2 = frcp(e2);
    T phiint = phic * ie2;
    phiint = (T)4 * phiint * v2 + (o + a_k * a_k * phiint) * (v2 + e2) * ie2;
    T g[3];
    g[0] = w000 * (Fp(1, 0, 0) - Fp(-1, 0, 0)) + w100 * (Fp(2, 0, 0) - Fp(0, 0, 0)) +
           w010 * (Fp(1, 1, 0) - Fp(-1, 1, 0)) + w110 * (Fp(2, 1, 0) - Fp(0, 1, 0)) +
           w001 * (Fp(1, 0, 1) - Fp(-1, 0, 1)) + w101 * (Fp(2, 0, 1) - Fp(0, 0, 1)) +
           w011 * (Fp(1, 1, 1) - Fp(-1, 1, 1)) + w111 * (Fp(2, 1, 1) - Fp(0, 1, 1));
    g[1] = w000 * (Fp(0, 1, 0) - Fp(0, -1, 0)) + w100 * (Fp(1, 1, 0) - Fp(1, -1, 0)) +
           w010 * (Fp(0, 2, 0) - Fp(0, 0, 0)) + w110 * (Fp(1, 2, 0) - Fp(1, 0, 0)) +
           w001 * (Fp(0, 1, 1) - Fp(0, -1, 1)) + w101 * (Fp(1, 1, 1) - Fp(1, -1, 1)) +
           w011 * (Fp(0, 2, 1) - Fp(0, 0, 1)) + w111 * (Fp(1, 2, 1) - Fp(1, 0, 1));
    g[2] = w000 * (Fp(0, 0, 1) - Fp(0, 0, -1)) + w100 * (Fp(1, 0, 1) - Fp(1, 0, -1)) +
           w010 * (Fp(0, 1, 1) - Fp(0, 1, -1)) + w110 * (Fp(1, 1, 1) - Fp(1, 1, -1)) +
           w001 * (Fp(0, 0, 2) - Fp(0, 0, 0)) + w101 * (Fp(1, 0, 2) - Fp(1, 0, 0)) +
           w011 * (Fp(0, 1, 2) - Fp(0, 1, 0)) + w111 * (Fp(1, 1, 2) - Fp(1, 1, 0));
    const T hp = h * phiint;
    g[0] *= hp; g[1] *= hp; g[2] *= hp;
    if (nf_k >= 2) {
        g[0] -= ex00 * (Fc(1, 0, 0) - Fc(0, 0, 0)) + ex10 * (Fc(1, 1, 0) - Fc(0, 1, 0)) +
                ex01 * (Fc(1, 0, 1) - Fc(0, 0, 1)) + ex11 * (Fc(1, 1, 1) - Fc(0, 1, 1));
        g[1] -= ey00 * (Fc(0, 1, 0) - Fc(0, 0, 0)) + ey10 * (Fc(1, 1, 0) - Fc(1, 0, 0)) +
                ey01 * (Fc(0, 1, 1) - Fc(0, 0, 1)) + ey11 * (Fc(1, 1, 1) - Fc(1, 0, 1));
        g[2] -= ez00 * (Fc(0, 0, 1) - Fc(0, 0, 0)) + ez10 * (Fc(1, 0, 1) - Fc(1, 0, 0)) +
                ez01 * (Fc(0, 1, 1) - Fc(0, 1, 0)) + ez11 * (Fc(1, 1, 1) - Fc(1, 1, 0));
    }
    const T e = fsqrt(e2);
    if (nf_k >= 3) {
        const T t0 = (T)2 * r0, t1 = (T)2 * r1, t2 = (T)2 * r2;
        auto l0 = [&](int y, int z) { return (r0 - o) * B0(-1, y, z) + (o - t0) * B0(0, y, z) + r0 * B0(1, y, z); };
        auto l1 = [&](int x, int z) { return (r1 - o) * B1(x, -1, z) + (o - t1) * B1(x, 0, z) + r1 * B1(x, 1, z); };
        auto l2 = [&](int x, int y) { return (r2 - o) * B2(x, y, -1) + (o - t2) * B2(x, y, 0) + r2 * B2(x, y, 1); };
        T pg0 = (s2 * (B1(1, 0, 0) - B1(0, 0, 0)) + r2 * (B1(1, 0, 1) - B1(0, 0, 1))) * q[1];
        pg0 += (s1 * (B2(1, 0, 0) - B2(0, 0, 0)) + r1 * (B2(1, 1, 0) - B2(0, 1, 0))) * q[2];
        pg0 += (ex00 * l0(0, 0) + ex10 * l0(1, 0) + ex01 * l0(0, 1) + ex11 * l0(1, 1)) * q[0];
        T pg1 = (s0 * (B2(0, 1, 0) - B2(0, 0, 0)) + r0 * (B2(1, 1, 0) - B2(1, 0, 0))) * q[2];
        pg1 += (s2 * (B0(0, 1, 0) - B0(0, 0, 0)) + r2 * (B0(0, 1, 1) - B0(0, 0, 1))) * q[0];
        pg1 += (ey00 * l1(0, 0) + ey10 * l1(1, 0) + ey01 * l1(0, 1) + ey11 * l1(1, 1)) * q[1];
        T pg2 = (s1 * (B0(0, 0, 1) - B0(0, 0, 0)) + r1 * (B0(0, 1, 1) - B0(0, 1, 0))) * q[0];
        pg2 += (s0 * (B1(0, 0, 1) - B1(0, 0, 0)) + r0 * (B1(1, 0, 1) - B1(1, 0, 0))) * q[1];
        pg2 += (ez00 * l2(0, 0) + ez10 * l2(1, 0) + ez01 * l2(0, 1) + ez11 * l2(1, 1)) * q[2];
        const T sb = frcp(bc_k * e);
        g[0] += pg0 * sb;
        g[1] += pg1 * sb;
        g[2] += pg2 * sb;
    }
    const T kf = dtau_k * e * Nn;
    v2 = (T)0;
#pragma unroll
    for (int l = 0; l < 3; l++) {
        q[l] -= kf * g[l];
        v2 += q[l] * q[l];
    }
    const T ret = v2 * frcp(a_k * a_k);

    // ---- drift with the kicked momentum
    const T ed2 = v2 + a_d * a_d, ie = frsqrt(ed2);
    const T ph = nf_d >= 1 ? phic : (T)0;
    const T ch = nf_d >= 2 ? cic8(Fc) : (T)0;
    const T vf = (o + ((T)3 - v2 * ie * ie) * ph - ch) * ie;
    if (nf_d >= 3) {
        const T ibc = frcp(bc_d);
        const T b0 = B0(0, 0, 0) * ex00 + B0(0, 1, 0) * ex10 + B0(0, 0, 1) * ex01 + B0(0, 1, 1) * ex11;
        const T b1 = B1(0, 0, 0) * ey00 + B1(1, 0, 0) * ey10 + B1(0, 0, 1) * ey01 + B1(1, 0, 1) * ey11;
        const T b2 = B2(0, 0, 0) * ez00 + B2(1, 0, 0) * ez10 + B2(0, 1, 0) * ez01 + B2(1, 1, 0) * ez11;
        d[0] = dtau_d * (double)(q[0] * vf + b0 * ibc);
        d[1] = dtau_d * (double)(q[1] * vf + b1 * ibc);
        d[2] = dtau_d * (double)(q[2] * vf + b2 * ibc);
    } else {
#pragma unroll
        for (int l = 0; l < 3; l++) d[l] = dtau_d * (double)(q[l] * vf);
    }
    return ret;
}

// rank = atomicAdd(&count[key], 1) for every active lane, with one atomic per
// run of consecutive lanes that share the key (cell-sorted particles put the
// members of a dense cell in neighbouring lanes, which would otherwise queue
// on one counter).  Lanes outside the current execution mask count as
// inactive (their ballot bits are 0); inactive lanes get -1.
// The rank comes back with the counter atomic; wave_rank_issue starts it and
// wave_rank_finish (a wave-wide shuffle, so every lane calls it) reads it, so a
// caller can overlap the atomic's round trip with other work.
struct RankPending {
    int32_t old;  // the counter value before the run's add (valid in the run's first lane)
    int h;        // first lane of this lane's run
};
__device__ __forceinline__ RankPending wave_rank_issue(int32_t *count, int32_t key, bool active)
{
    const int lane = threadIdx.x & 63;
    const int32_t prev = __shfl_up(key, 1, 64);
    const unsigned long long act = __ballot(active);
    const bool prev_active = lane > 0 && ((act >> (lane - 1)) & 1ull);
    const bool head = !active || lane == 0 || !prev_active || prev != key;
    const unsigned long long heads = __ballot(head) | ~act;
    const unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const int h = 63 - __clzll(heads & le);
    const unsigned long long above = heads & ~le;
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    int32_t old = 0;
    if (active && lane == h) old = atomicAdd(&count[key], next - h);
    return RankPending{old, h};
}
__device__ __forceinline__ int32_t wave_rank_finish(const RankPending &r)
{
    const int lane = threadIdx.x & 63;
    return __shfl(r.old, r.h, 64) + (lane - r.h);
}
__device__ __forceinline__ int32_t wave_rank_add(int32_t *count, int32_t key, bool active)
{
    const int32_t r = wave_rank_finish(wave_rank_issue(count, key, active));
    return active ? r : -1;
}

template <int NT = PT_NT> __device__ __forceinline__ void block_max_to(double v, unsigned long long *slot)
{
    __shared__ double sh[NT / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        for (int j = 1; j < NT / 64; j++) m = fmax(m, sh[j]);
        // all blocks hit one address: skip the atomic unless it can raise the maximum
        unsigned long long cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((unsigned long long)__double_as_longlong(m) > cur) atomic_max_nonneg(slot, m);
    }
}

// ---- particles of a tile in cell order (the tiled particle kernels)
// row prefix of a tile: lane r < NROW gets the first container index of row r
// and the inclusive prefix of the row lengths (after wave_row_scan)
template <int TX, int TY, int NROW>
__device__ __forceinline__ void tile_rows_load(const int32_t *start, int N, int x0, int y0, int z0, int lane,
                                               int32_t &rb0, int32_t &len)
{
    rb0 = 0;
    len = 0;
    if (lane < NROW) {
        const int ry = lane % TY, rz = lane / TY;
        const uint32_t rb = (uint32_t)x0 + (uint32_t)N * ((uint32_t)(y0 + ry) + (uint32_t)N * (uint32_t)(z0 + rz));
        rb0 = start[rb];
        len = start[rb + TX] - rb0;
    }
}
__device__ __forceinline__ int32_t wave_row_scan(int32_t len, int lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(len, o, 64);
        if (lane >= o) len += y;
    }
    return len;
}
// container index of flattened particle f of a tile with row prefix (rb0,
// incl); every lane runs the same steps (the shuffles need all lanes), lanes
// with f beyond the tile's particles get -1
template <int NROW> __device__ __forceinline__ int64_t tile_locate(int32_t rb0, int32_t incl, int f)
{
    const int32_t total = __shfl(incl, NROW - 1, 64);
    int lo = 0;
#pragma unroll
    for (int step = NROW / 2; step >= 1; step >>= 1) {
        const int j = lo + step < NROW ? lo + step : NROW - 1;
        const int32_t rsj = __shfl(incl, j - 1, 64);  // first flattened index of row j
        if (lo + step < NROW && rsj <= f) lo = j;
    }
    const int32_t rsl = __shfl(incl, lo > 0 ? lo - 1 : 0, 64);
    const int32_t rbl = __shfl(rb0, lo, 64);
    return f < total ? (int64_t)rbl + (f - (lo > 0 ? rsl : 0)) : (int64_t)-1;
}

}  // namespace gev

// gev_kops.hpp -- per-mode arithmetic of the k-space solvers, shared by the
// elementwise k-space kernels and the FFT z-pass kernels that apply them in
// registers.  Everything is FP64 (the mixed form converts at load / store).
//
// Mode tables of the context: gk2[n] = (2 N sin(pi n / N))^2 (the discrete
// Laplacian eigenvalue along one axis, box units) and ks[n] = 2 N sin(pi n /
// N) exp(-i pi n / N) (the half-cell shifted derivative used by the staggered
// tensor components, gevolution.hpp:272-274).
#pragma once
#include "gev_common.hpp"

namespace gev {

// solveModifiedPoissonFT (gevolution.hpp:806-852) with coeff already scaled by
// -1/N^3: pot = src * coeff / (k^2 + modif); the k=0 mode is 0 unless modif != 0
__device__ __forceinline__ cx<double> kop_poisson(cx<double> s, double kk, double coeff, double modif)
{
    return s * (coeff / (kk + modif));
}

// projectFTscalar (gevolution.hpp:251-340): chi = sum_c w_c(k) S_c / (2 k^4 N)
// over the components (00,01,02,11,12,22); k=0 gives 0
__device__ __forceinline__ cx<double> kop_scalar_weight(int c, double g0, double g1, double g2, cx<double> k0,
                                                        cx<double> k1, cx<double> k2)
{
    switch (c) {
    case 0: return mk(g1 + g2 - 2.0 * g0, 0.0);
    case 1: return (k0 * k1) * -6.0;
    case 2: return (k0 * k2) * -6.0;
    case 3: return mk(g0 + g2 - 2.0 * g1, 0.0);
    case 4: return (k1 * k2) * -6.0;
    default: return mk(g0 + g1 - 2.0 * g2, 0.0);
    }
}
__device__ __forceinline__ double kop_scalar_norm(double kk, int N) { return 1.0 / (2.0 * kk * kk * (double)N); }

// projectFTvector (gevolution.hpp:449-515): B_i = 4 coeff (S_i - conj(ks_i)
// (ks.S) / k^2) / (k^2 + modif); k=0 gives 0
__device__ __forceinline__ void kop_vector(const cx<double> s[3], double kk, cx<double> k0, cx<double> k1,
                                           cx<double> k2, double coeff, double modif, cx<double> b[3])
{
    const cx<double> tmp = (k0 * s[0] + k1 * s[1] + k2 * s[2]) * (1.0 / kk);
    const double f = 4.0 * coeff / (kk + modif);
    b[0] = (s[0] - conj(k0) * tmp) * f;
    b[1] = (s[1] - conj(k1) * tmp) * f;
    b[2] = (s[2] - conj(k2) * tmp) * f;
}

}  // namespace gev

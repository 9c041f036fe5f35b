// gev_kernels.hpp -- internal launchers shared between the HIP translation
// units of the library (not part of the C ABI).
#pragma once
#include <stdio.h>

#include "gev_common.hpp"

namespace gev {

template <typename T> int fft_r2c_impl(gev_ctx *ctx, const T *in, cx<T> *out, int ncomp, const cx<T> *tw);
template <typename T> int fft_c2r_impl(gev_ctx *ctx, cx<T> *in, T *out, int ncomp, const cx<T> *tw);

template <typename T> int poisson_impl(gev_ctx *ctx, const cx<T> *src, cx<T> *pot, double coeff, double modif);
template <typename T> int project_ft_scalar_impl(gev_ctx *ctx, const cx<T> *S, cx<T> *chi, int add);
template <typename T> int project_ft_vector_impl(gev_ctx *ctx, const cx<T> *S, cx<T> *B, double coeff, double modif);
template <typename T> int evolve_ft_vector_impl(gev_ctx *ctx, const cx<T> *S, cx<T> *B, double a2dtau);
template <typename T>
int ic_displacement_ft_impl(gev_ctx *ctx, uint32_t seed, int ksphere, double coeff, int npk, const double *dx,
                            const double *dy, const double *dm, const cx<T> *ker, cx<T> *pot);

template <typename T> int set_k0_impl(gev_ctx *ctx, cx<T> *ft, double re, double im);

// r2c -> k-space operator -> c2r with the operator applied inside the z pass
// (64 <= N <= 1024); the FT arguments are work space of the usual layout
bool zsolve_supported(const gev_ctx *c);
bool chi_from_Tij_supported(const gev_ctx *c);
template <typename T>
int solve_poisson_fused_impl(gev_ctx *c, const T *src, T *pot, cx<T> *work, double coeff, double modif, int k0set,
                             double k0re);
template <typename T> int solve_chi_fused_impl(gev_ctx *c, const T *Sij, T *chi, cx<T> *SijFT, cx<T> *chiFT);
template <typename T>
int solve_chi_Tij_fused_impl(gev_ctx *c, const T *phi, const T *Tij, double coeff, T *chi, cx<T> *SijFT, cx<T> *chiFT);
template <typename T> int solve_vector_fused_impl(gev_ctx *c, T *Bi, cx<T> *BiFT, double coeff, double modif);
// Bi holds T0i link sums (no 1 + phi_a + phi_b factor) on entry, B on return
// (128 <= N <= 1024, as chi_from_Tij_supported)
template <typename T>
int solve_vector_links_fused_impl(gev_ctx *c, T *Bi, const T *phi, cx<T> *BiFT, double coeff, double modif);
// prepareFTsource T00 form formed inside the Poisson solve's x pass (128 <= N
// <= 1024); ctx->red[0] / ctx->red[1] receive the FP64 sums of the source and
// of the result, and the k=0 mode is taken from ctx->red[1] when k0set
template <typename T>
int solve_poisson_T00_fused_impl(gev_ctx *c, const T *phi, const T *chi, const T *src, double bg, double c1,
                                 double c2, double c3, T *pot, cx<T> *work, double coeff, double modif, int k0set);
// Bi[c] *= 1 + phi(x) + phi(x + e_c), in place
template <typename T> int t0i_link_factor_impl(gev_ctx *c, T *Bi, const T *phi);

template <typename T>
int prepare_T00_impl(gev_ctx *ctx, const T *phi, const T *chi, const T *src, double bg, T *res, double c1, double c2,
                     double c3, double *sums);
template <typename T> int prepare_Tij_impl(gev_ctx *ctx, const T *phi, const T *Tij, T *Sij, double coeff);
template <typename T> int field_sum_impl(gev_ctx *ctx, const T *f, int64_t count, double *sum);
template <typename T> int field_fill_impl(gev_ctx *ctx, T *f, int64_t count, T v);

// copy the last kick's maximum (FP64 slot) to the host; synchronises the stream
int read_max_slot(gev_ctx *c, double *out);
template <typename T>
int adj_gr_kick_drift_impl(gev_pcls *p, const T *phi, const T *chi, const T *Bi, T *lx, T *lq, T *lphi, T *lchi,
                           T *lB, int nf_k, int nf_d, double a_k, double a_d, double dtau_k, double dtau_d);
template <typename T>
int adj_gr_deposit_impl(gev_pcls *p, const T *phi, const T *lT00, const T *lTij, const T *lL, T *lx, T *lq, T *lphi,
                        double a, double tij_scale);
template <typename T>
int adj_source_T00_impl(gev_ctx *c, const T *phi, const T *T00, const T *lsrc, T *lT00, T *lphi, T *lchi, double bg,
                        double c1, double c2, double c3);
template <typename T> int adj_source_Tij_phi_impl(gev_ctx *c, const T *phi, const T *lS, T *lphi);
template <typename T> int adj_link_factor_impl(gev_ctx *c, const T *phi, const T *L, T *u, T *lphi);
template <typename T> int adj_project_ft_scalar_impl(gev_ctx *c, const cx<T> *lchi, cx<T> *lS);

// zero the cell counters unless the last scan already left them zero
int pcls_clear_counts(gev_pcls *p);

// start[0..M) = exclusive prefix of cnt[0..M), start[M] = the sum; clears cnt.
// M at most the context's lattice size.
int scan_exclusive_counts(gev_ctx *ctx, int32_t *cnt, int64_t M, int32_t *start);

// one Fortran-style record: byte count, data, byte count
bool gadget_put_record(FILE *f, const void *data, uint32_t bytes);

// deterministic two-level FP64 sum: per-block partials then one block
int reduce_partials(gev_ctx *ctx, const double *partials, int nparts, int nvals, double *out_dev);

}  // namespace gev

/*
 * gev_pm.h -- C ABI of the MI355X-native gevolution particle-mesh step.
 *
 * One context (gev_ctx) = one periodic lattice of N^3 sites (N a power of two,
 * 4 <= N <= 1024) in one precision form on one GPU, with one HIP stream.
 *
 * Memory layout of every buffer passed in (device pointers, caller-owned):
 *   real field     T[ncomp][N^3], site (x,y,z) at x + N*(y + N*z); no halo,
 *                  periodic indexing is done inside the kernels.  Symmetric
 *                  tensors (Tij, Sij) store components 00,01,02,11,12,22.
 *   Fourier field  complex T2[ncomp][N][N][P] (re,im interleaved); mode
 *                  (kx,ky,kz) at kx + P*(ky + N*kz), kx in [0,N/2], ky,kz in
 *                  [0,N); P = gev_ctx_kpitch() >= N/2+1 (padding columns are
 *                  zero after a forward transform and ignored otherwise).
 *   T = double for GEV_FP64 contexts, float for GEV_MIXED contexts.
 *   Forward transforms use exp(-2 pi i k.x/N); both directions are
 *   unnormalised, as LATfield2's PlanFFT (gevolution.hpp:817 divides by N^3).
 *
 * Particles live in a library-owned container (gev_pcls), the analogue of
 * LATfield2's perfParticles used by perfParticles_gevolution
 * (Particles_gevolution.hpp:176-211).  They are kept sorted by cell
 * (cell = x + N*(y + N*z) with x = floor(pos*N)).  GEV_FP64 stores absolute
 * positions in box units (double); GEV_MIXED stores FP32 cell-relative
 * offsets in [0,1) next to the int32 cell.  Load/store always exchange FP64
 * absolute positions.
 *
 * Every entry point returns GEV_OK (0) or a negative code; gev_last_error()
 * gives the message.  Calls are asynchronous on the context stream unless
 * they return a host value (reductions), which synchronises.  A context is
 * not re-entrant: one host thread at a time.
 */
#ifndef GEV_PM_H
#define GEV_PM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GEV_OK = 0, GEV_ERR_ARG = -1, GEV_ERR_HIP = -2, GEV_ERR_STATE = -3, GEV_ERR_NOMEM = -4 };
enum { GEV_FP64 = 0, GEV_MIXED = 1 };

typedef struct gev_ctx gev_ctx;
typedef struct gev_pcls gev_pcls;

/* raw device view of a particle container (valid until the next re-binning).
 * For reading: a caller that writes positions or momenta through it calls
 * gev_pcls_sort before the next deposit (the mixed form's deposits otherwise
 * bound their fixed-point sums by the momenta of the last particle update) */
typedef struct {
    int64_t n;
    int precision;
    void *pos[3];        /* double* (FP64, box units) or float* (MIXED, cell offsets) */
    void *q[3];          /* double* or float*: q/m ("v a", gevolution.hpp:860-863) */
    int64_t *id;
    int32_t *cell;
    int32_t *cell_start; /* [N^3 + 1] */
} gev_pcls_view;

const char *gev_last_error(void);
int gev_version(void);

/* ---- context (replaces LATfield2 Lattice + PlanFFT setup, main.cu:300-357) */
int gev_ctx_create(int N, int precision, void *hip_stream, gev_ctx **out);
int gev_ctx_destroy(gev_ctx *ctx);
int gev_ctx_sync(gev_ctx *ctx);
int gev_ctx_kpitch(const gev_ctx *ctx);
/* launch subsequent work on this stream (NULL = the legacy default stream) */
int gev_ctx_set_stream(gev_ctx *ctx, void *hip_stream);

/* ---- particles (perfParticles: initialize / addParticle_global /
 *      updateRowBuffers, ic_basic.hpp:1639,2206-2211; sampleParticles,
 *      unit_tests.cu:529,537) */
int gev_pcls_create(gev_ctx *ctx, int64_t capacity, double mass, gev_pcls **out);
int gev_pcls_destroy(gev_pcls *p);
int gev_pcls_load(gev_pcls *p, int64_t n, const double *pos, const double *q, const int64_t *id, int on_device);
int gev_pcls_store(gev_pcls *p, double *pos, double *q, int64_t *id, int on_device);
int gev_pcls_view_get(gev_pcls *p, gev_pcls_view *v);
int gev_pcls_sort(gev_pcls *p);
/* copy the cell index of every particle (n int32) and/or the cell offsets
 * (N^3+1 int32) into caller device buffers; either may be NULL */
int gev_pcls_copy_index(gev_pcls *p, int32_t *cell, int32_t *cell_start);
/* sc1-type lattice: one particle at every cell centre, ID = cell index
 * (initializeParticlePositions, ic_basic.hpp:1612-1644, tiling N/4 of the
 * 64-particle cell-centred template) */
int gev_pcls_init_lattice(gev_pcls *p);

/* ---- CIC projections (out = deposit, or out += deposit if accumulate) */
/* projection_T00_project, gevolution.hpp:1887-1899 (functor :1729-1854);
 * phi == NULL gives the plain mass density (nfield = 1 branch) */
int gev_project_T00(gev_pcls *p, void *T00, double a, const void *phi, double coeff, int accumulate);
/* scalarProjectionCIC_project / meshprojection_project, gevolution.hpp:1902-1905 */
int gev_project_mass(gev_pcls *p, void *rho, double coeff, int accumulate);
/* projection_T0i_project, gevolution.hpp:2232-2245 (functor :2031-2201) */
int gev_project_T0i(gev_pcls *p, void *T0i, const void *phi, double coeff, int accumulate);
/* projection_Tij_project, gevolution.hpp:2999-3033 (functor :2582-2965), no CIC_PROJECT_TIJ */
int gev_project_Tij(gev_pcls *p, void *Tij, double a, const void *phi, double coeff, int accumulate);
/* projection_T00_project and projection_Tij_project of the same particles
 * with the same phi and a (main.cu:623 and 756) in one pass over the
 * particles; no single reference function.  Results as the two calls. */
int gev_project_T00_Tij(gev_pcls *p, void *T00, void *Tij, double a, const void *phi, double coeff, int accumulate);
/* No single reference function: gev_project_T00_Tij (with phi) and, in
 * links[3][N^3], projection_T0i_project's link sums without their
 * 1 + phi_a + phi_b factor (gevolution.hpp:2031-2201 multiplies each of a
 * particle's 12 link values by the phi of the link's two end points, which is
 * the same for every particle on the link).  gev_solve_vector_T0i_links
 * applies the factor with the phi of its own call, so T0i can be deposited in
 * the same pass over the particles as T00 and Tij although main.cu projects
 * it after the phi solve (main.cu:858).  Overwrites its outputs. */
int gev_project_T00_Tij_T0i_links(gev_pcls *p, void *T00, void *Tij, void *links, double a, const void *phi,
                                  double coeff);

/* ---- real-space sources */
/* prepareFTsource (T00 form, PHINONLINEAR), gevolution.hpp:160-218.
 * *sum_source = sum of source before the update (the reference's return
 * value); *sum_result = FP64 sum of the result (used for the k=0 mode). */
int gev_prepare_source_T00(gev_ctx *ctx, const void *phi, const void *chi, const void *source, double bgmodel,
                           void *result, double c1, double c2, double c3, double *sum_source, double *sum_result);
/* prepareFTsource (Tij form, PHINONLINEAR), gevolution.hpp:64-138 */
int gev_prepare_source_Tij(gev_ctx *ctx, const void *phi, const void *Tij, void *Sij, double coeff);

/* ---- FFT (PlanFFT<Cplx>::execute FFT_FORWARD / FFT_BACKWARD, main.cu:742-948) */
int gev_fft_r2c(gev_ctx *ctx, const void *in, void *out, int ncomp);
int gev_fft_c2r(gev_ctx *ctx, void *in, void *out, int ncomp); /* clobbers in */

/* ---- k-space solvers */
int gev_poisson(gev_ctx *ctx, const void *src, void *pot, double coeff, double modif);  /* gevolution.hpp:806-852 */
int gev_project_ft_scalar(gev_ctx *ctx, const void *Sij, void *chi, int add);           /* gevolution.hpp:251-340 */
int gev_project_ft_vector(gev_ctx *ctx, const void *Si, void *Bi, double coeff, double modif); /* :449-515 */
int gev_evolve_ft_vector(gev_ctx *ctx, const void *Sij, void *Bi, double a2dtau);       /* gevolution.hpp:359-429 */
/* overwrite the k=0 mode (used to restore the FP64 mean in the mixed form) */
int gev_set_k0(gev_ctx *ctx, void *ft, double re, double im);

/* ---- fused real-space solves.  No single reference function: each is one
 *      main.cu call sequence with the k-space operator applied inside the z
 *      pass of the transforms (64 <= N <= 1024; smaller N runs the sequence
 *      unfused).  Work buffers are Fourier fields of the usual layout; their
 *      content is undefined afterwards. */
/* plan_source FFT_FORWARD, [k=0 := k0_re], solveModifiedPoissonFT, plan_phi
 * FFT_BACKWARD (main.cu:742-794) */
int gev_solve_poisson(gev_ctx *ctx, const void *src, void *pot, void *workFT, double coeff, double modif, int set_k0,
                      double k0_re);
/* plan_Sij FFT_FORWARD, projectFTscalar (add = 0), plan_chi FFT_BACKWARD (main.cu:829-895) */
int gev_solve_chi(gev_ctx *ctx, const void *Sij, void *chi, void *SijFT, void *chiFT);
/* prepareFTsource<Cplx>(phi, Sij, Sij, coeff) (gevolution.hpp:64-119,
 * main.cu:822-828) followed by gev_solve_chi: for 128 <= N <= 1024 the x
 * passes read Tij and phi and form Sij on the fly, so Sij is never stored;
 * Tij is left unchanged there and holds Sij on return for smaller N. */
int gev_solve_chi_Tij(gev_ctx *ctx, const void *phi, void *Tij, double coeff, void *chi, void *SijFT, void *chiFT);
/* plan_Bi FFT_FORWARD, projectFTvector, plan_Bi FFT_BACKWARD (main.cu:918-948), Bi in place */
int gev_solve_vector(gev_ctx *ctx, void *Bi, void *BiFT, double coeff, double modif);
/* prepareFTsource, T00 form (gevolution.hpp:160-183; main.cu:725), and
 * solveModifiedPoissonFT with its transforms (main.cu:742-794) in one chain,
 * pot = phi allowed: for 128 <= N <= 1024 the source is formed inside the
 * forward x pass and never stored, and source is only read; below that the
 * source pass runs in place over source first.  k0set: the k=0 mode of the
 * transformed source becomes the FP64 sum of the source terms (as
 * gev_solve_poisson with k0re = sums[1]).  sums (may be NULL): the FP64 sums
 * of source and of the source terms, as gev_prepare_source_T00 returns them;
 * with NULL the call does not wait for the GPU, and gev_ctx_sums reads them
 * later (until the context's next reduction). */
int gev_solve_poisson_T00(gev_ctx *ctx, const void *phi, const void *chi, void *source, double bg, double c1,
                          double c2, double c3, void *pot, void *work, double coeff, double modif, int k0set,
                          double *sums);
/* the two sums of the last gev_solve_poisson_T00 (waits for the context's stream) */
int gev_ctx_sums(gev_ctx *ctx, double sums[2]);
/* gev_solve_vector of T0i given as link sums (gev_project_T00_Tij_T0i_links):
 * Bi[c](x) *= 1 + phi(x) + phi(x + e_c), then the solve; for
 * 128 <= N <= 1024 the factor is applied inside the forward x passes. */
int gev_solve_vector_T0i_links(gev_ctx *ctx, void *Bi, const void *phi, void *BiFT, double coeff, double modif);

/* ---- particle updates (perfParticles::updateVel / moveParticles with the
 *      gevolution functors, main.cu:1227-1266).  *max_ret receives the
 *      maximum over particles of the functor's return value (v^2 / a^2). */
int gev_update_q(gev_pcls *p, double dtau, const void *phi, const void *chi, const void *Bi, int nfield, double a,
                 double bcoeff, double *max_ret);                            /* update_q, gevolution.hpp:888-1011 */
int gev_update_q_newton(gev_pcls *p, double dtau, const void *psi, const void *chi, int nfield, double a,
                        double *max_ret);                                    /* gevolution.hpp:1051-1118 */
int gev_update_pos(gev_pcls *p, double dtau, const void *phi, const void *chi, const void *Bi, int nfield, double a,
                   double bcoeff);                                           /* gevolution.hpp:1162-1232 + re-bin */
int gev_update_pos_newton(gev_pcls *p, double dtau, double a);               /* gevolution.hpp:1270-1273 + re-bin */

/* kick and drift fused (main.cu:1227-1266 with the background half-step in
 * between done by the caller): one pass reads each particle once, applies
 * update_q (gr=1) or update_q_Newton (gr=0) with (dtau_kick, a_kick,
 * bcoeff_kick, nfield_kick), then update_pos / update_pos_Newton with the
 * kicked momentum and (dtau_drift, a_drift, bcoeff_drift, nfield_drift), and
 * counts the re-binning ranks.  rebin=1 finishes the re-binning; rebin=0
 * leaves it to gev_pcls_rebin_counted (no other call may touch the container
 * in between).  *max_ret gets the kick's maximum functor value, which waits
 * for the stream; max_ret = NULL skips it (no wait). */
int gev_kick_drift(gev_pcls *p, int gr, double dtau_kick, double dtau_drift, const void *phi, const void *chi,
                   const void *Bi, int nfield_kick, int nfield_drift, double a_kick, double a_drift,
                   double bcoeff_kick, double bcoeff_drift, double *max_ret, int rebin);
int gev_pcls_rebin_counted(gev_pcls *p);

/* ---- initial conditions (generateIC_basic mPk path, ic_basic.hpp:1834-2505) */
/* displacement potential in k-space (generateDisplacementField, ic_basic.hpp:1164-1270);
 * kL/amp: power-spectrum table converted as loadPowerSpectrum (ic_basic.hpp:445-447) */
int gev_ic_displacement_ft(gev_ctx *ctx, uint32_t seed, int ksphere, double coeff, int npk, const double *kL,
                           const double *amp, const void *kerFT, void *potFT);
/* displace_pcls_ic_basic (ic_basic.hpp:65-96) + re-bin; *max_disp = max |displacement| */
int gev_displace_particles(gev_pcls *p, const void *xi, double coeff, double *max_disp);
/* initialize_q_ic_basic (ic_basic.hpp:133-168); *max_ret = max q^2 */
int gev_initialize_q(gev_pcls *p, const void *pot, double coeff, double *max_ret);

/* ---- container checkpoints (no reference counterpart: the state that a
 *      hibernation file holds for the particles, kept in device memory for
 *      the adjoint's recompute).  dst/src: 16-byte aligned device buffers of
 *      `bytes` bytes, at least gev_pcls_state_bytes of the container saved
 *      (a restore checks the saved count against the container's capacity
 *      and the buffer's size). */
int gev_pcls_state_bytes(gev_pcls *p, int64_t *bytes);
int gev_pcls_save(gev_pcls *p, void *dst, int64_t bytes);
int gev_pcls_restore(gev_pcls *p, const void *src, int64_t bytes);

/* ---- Gadget-2 particle snapshots, single file, particle type 1
 *      (perfParticles_gevolution::saveGadget2 / loadGadget2,
 *      Particles_gevolution.hpp:597-800 and 2455-2640).  The header is the
 *      reference's gadget2_header (metadata.hpp:159-178), 256 bytes. */
typedef struct gev_gadget2_header {
    uint32_t npart[6];
    double mass[6];
    double time;
    double redshift;
    int32_t flag_sfr;
    int32_t flag_feedback;
    uint32_t npartTotal[6];
    int32_t flag_cooling;
    int32_t num_files;
    double BoxSize;
    double Omega0;
    double OmegaLambda;
    double HubbleParam;
    int32_t flag_age;
    int32_t flag_metals;
    uint32_t npartTotalHW[6];
    char fill[64];
} gev_gadget2_header;
/* Writes the particles with ID % tracer_factor == 0 (hdr->npart[1] must equal
 * their number; hdr->num_files must be 1): positions p*BoxSize, velocities
 * q/(a^1.5 GADGET_VELOCITY_CONVERSION) with a = hdr->time, both moved to the
 * snapshot time as EXACT_OUTPUT_REDSHIFTS does (drift dtau_pos, kick dtau_vel
 * with the gradient of phi; phi may be NULL).  Records in container order. */
int gev_pcls_save_gadget2(gev_pcls *p, const char *path, const gev_gadget2_header *hdr, int tracer_factor,
                          double dtau_pos, double dtau_vel, const void *phi);
/* Replaces the container's particles with the snapshot's (positions wrapped
 * into the box, velocities converted back to q/m) and re-bins; *hdr (may be
 * NULL) receives the header. */
int gev_pcls_load_gadget2(gev_pcls *p, const char *path, gev_gadget2_header *hdr);
/* saveExpress / loadExpress (Particles_gevolution.hpp:2642-2900): the raw
 * particle state as one rank's file "<path>.0" (a numeric suffix of path is
 * replaced): express header, positions (box units) and momenta as Real
 * triplets (Real = 8 bytes in FP64, 4 in the mixed form), int64 IDs, then
 * the Gadget-2 header.  save fills hdr's particle counts, as the reference
 * does; load checks the file against this lattice and precision. */
int gev_pcls_save_express(gev_pcls *p, const char *path, gev_gadget2_header *hdr);
int gev_pcls_load_express(gev_pcls *p, const char *path, gev_gadget2_header *hdr);

/* ---- lightcone particle output (one rank) */
#define GEV_MAX_INTERSECTS 12            /* MAX_INTERSECTS, metadata.hpp:34-36 */
#define GEV_LIGHTCONE_MAX_VERTICES 32    /* vertices one output call accepts */
#define GEV_LIGHTCONE_IDCHECK_ZONE 0.05  /* LIGHTCONE_IDCHECK_ZONE, metadata.hpp:38-40 */
/* lightcone_geometry (metadata.hpp:205-212): vertex in box units, opening =
 * cos of the half-angle (-1: full sky), distance[2] the far and near edges. */
typedef struct gev_lightcone_geometry {
    double vertex[3];
    double z;
    double direction[3];
    double opening;
    double distance[2];
} gev_lightcone_geometry;
/* findIntersectingLightcones (tools.hpp:619-862): the periodic copies of the
 * observer whose look-back shell [inner, outer] and cone may meet the box
 * domain[0..2] (low corner) to domain[3..5] (high corner), box units, in the
 * reference's u-major order.  *count = the number found; the first cap are
 * stored in vertex[cap][3]. */
int gev_lightcone_vertices(const gev_lightcone_geometry *lc, double outer, double inner, const double domain[6],
                           double *vertex, int cap, int *count);
/* The ID log of one lightcone group: LightconeIDBacklog
 * (lightcone_id_backlog.hpp:16-125) plus the IDprelog vector writeLightcones
 * fills and hands over (output.hpp:1017-1185, 2166-2880).  The lightcone
 * output skips check-zone particles whose ID is in the backlog and appends
 * the IDs it writes near the inner edge to the prelog; commit makes the
 * prelog (sorted, unique) the next backlog.  which: 0 = backlog, 1 = prelog. */
typedef struct gev_idlog gev_idlog;
int gev_idlog_create(gev_ctx *ctx, gev_idlog **out);
int gev_idlog_destroy(gev_idlog *log);
int gev_idlog_sizes(const gev_idlog *log, int64_t *nbacklog, int64_t *nprelog);
int gev_idlog_get(const gev_idlog *log, int which, int64_t *ids);
int gev_idlog_set_backlog(gev_idlog *log, const int64_t *ids, int64_t n);
int gev_idlog_commit(gev_idlog *log);
/* saveGadget2, lightcone form (perfParticles_gevolution, Particles_gevolution.hpp:
 * 172,1566-2317, device path with EXACT_OUTPUT_REDSHIFTS): writes the
 * particles whose ID is a multiple of tracer_factor and that lie in the shell
 * dist - dtau/2 <= d < dist + (1/2 + ZONE) dtau_old around one of the
 * vertices (and inside the opening), moved to their light-cone crossing and
 * kicked with the gradient of phi (may be NULL), as a single-file Gadget-2
 * snapshot, in container order.  hdr->time is a, hdr->BoxSize the length
 * scale; hdr's type-1 counts are set to the number written (*written, may
 * be NULL).  No file is written when it is zero.  log may be NULL (no
 * backlog check, no prelog). */
int gev_pcls_save_gadget2_lightcone(gev_pcls *p, const char *path, gev_gadget2_header *hdr,
                                    const gev_lightcone_geometry *lc, double dist, double dtau, double dtau_old,
                                    double dadtau, const double *vertex, int vertexcount, gev_idlog *log,
                                    const void *phi, int tracer_factor, int64_t *written);

/* ---- power spectra (extractPowerSpectrum, tools.hpp:65-342, KTYPE_LINEAR;
 * the deconvolved density spectrum of writeSpectra, output.hpp:3006-3033).
 * F: one Fourier component in the context's layout (complex of the context's
 * precision, pitch gev_ctx_kpitch).  numbins in [1, 1024] linear bins up to
 * sqrt(3) * 2 pi N/2; deconvolve != 0 divides |F|^2 by the CIC window squared.
 * Per bin (host arrays of numbins): mean k, mean of |F|^2 k^3 / W^2, their
 * scatters, and the occupation (modes whose conjugate is not stored count
 * twice).  Sums in FP64; returns after the result is on the host. */
int gev_power_spectrum(gev_ctx *ctx, const void *F, int numbins, int deconvolve, double *kbin, double *power,
                       double *kscatter, double *pscatter, int64_t *occupation);
/* extractCrossSpectrum / extractPowerSpectrum (tools.hpp:65-342, the latter
 * with F1 = F2) of fields of ncomp components each (component c NK modes after
 * component 0, the library's Fourier layout): the product per mode is
 * Re(F1 conj F2) of components comp1, comp2 when both are in [0, ncomp), else
 * the symmetric-tensor contraction when symmetric (ncomp = 6, components
 * xx xy xz yy yz zz, the off-diagonals counted twice), else the sum over
 * components.  ktype 0 = KTYPE_GRID, 1 = KTYPE_LINEAR (tools.hpp:27-28).
 * Outputs as gev_power_spectrum. */
int gev_cross_spectrum(gev_ctx *ctx, const void *F1, const void *F2, int ncomp, int symmetric, int comp1, int comp2,
                       int numbins, int deconvolve, int ktype, double *kbin, double *power, double *kscatter,
                       double *pscatter, int64_t *occupation);

/* gev_power_spectrum without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results; the workgroup partials live in a static 4 MiB
 * device workspace that every device holds until the process ends (the
 * column grid narrows until 5 numbins (groups + 1) doubles fit; numbins as
 * above, larger requests are refused).  Calls that overlap in time on one
 * device share that workspace and race. */
int gev_power_spectrum_noalloc(gev_ctx *ctx, const void *F, int numbins, int deconvolve, double *kbin,
                               double *power, double *kscatter, double *pscatter, int64_t *occupation);

/* ---- adjoint of the Newtonian cycle (no reference counterpart; reverse mode
 *      of update_q_Newton, update_pos_Newton, the CIC mass deposit and the
 *      Poisson solve, gevolution.hpp:1051-1118, 1270-1273, 1902-1905, 806-852).
 *      Per-particle adjoints lx, lq are [3][n] arrays indexed by particle ID
 *      (IDs in [0, n)), in the context precision.  With K = dtau_kick a N and
 *      C = dtau_drift / a_drift of the forward cycle, one reverse cycle is
 *   (lq <- lq + C lx, applied by the gather that finished lx, see C_prev)
 *   gev_adj_deposit:    u (3 components, overwritten) = CIC deposit of lq
 *   gev_adj_divergence: lphi = -K/2 sum_i (u_i(s - e_i) - u_i(s + e_i))
 *   (lrho = the Poisson solve of lphi with the forward coefficients)
 *   gev_adj_gather:     lx += -K N/2 sum_i lq_i grad1(D_i phi) + mN grad1(lrho),
 *                       then lq += C_prev lx when C_prev != 0 (C of the cycle
 *                       reversed next); mN = particle mass / dx^3 * N; phi or
 *                       lrho may be NULL
 *   gev_adj_ic_lattice: lpsi from the adjoints (lx, lq) of the lattice map
 *                       x = x_lat + cx N grad1(psi), q = cq N grad1(psi), ID = cell */
int gev_adj_deposit(gev_pcls *p, const void *lq, void *u);
int gev_adj_divergence(gev_ctx *ctx, const void *u, void *lphi, double K);
int gev_adj_gather(gev_pcls *p, const void *phi, const void *lrho, void *lq, void *lx, double K, double mN,
                   double C_prev);
int gev_adj_ic_lattice(gev_ctx *ctx, const void *lx, const void *lq, double cx, double cq, void *lpsi);

/* ---- adjoint of the weak-field GR cycle (no reference counterpart; reverse
 *      mode of the T00 / Tij / T0i projections, prepareFTsource in both forms,
 *      solveModifiedPoissonFT, projectFTscalar, projectFTvector with the
 *      elliptic B, update_q and update_pos: gevolution.hpp:1729-2965, 64-218,
 *      806-852, 251-340, 449-515, 888-1232).  lx, lq are [3][n] per particle
 *      ID; field adjoints are in the context precision and accumulate (+=)
 *      unless stated.
 *   gev_adj_gr_kick_drift: the fused kick (nfield_kick, a_kick, dtau_kick,
 *      bcoeff a_kick^2 N) and drift (nfield_drift, a_drift, dtau_drift) of the
 *      particles in their pre-kick state: lx, lq (after the drift) are
 *      replaced by the adjoints before the kick; lphi, lchi, lB (3) += the
 *      fields' adjoints
 *   gev_adj_gr_deposit: the T00 + Tij + T0i-link deposits (gev_project_T00_Tij_
 *      T0i_links with coeff 1) given their adjoints lT00, tij_scale lTij (6),
 *      lL (3, may be NULL): lx, lq += ; lphi += phi's adjoint (phi may be
 *      NULL: no phi)
 *   gev_adj_source_T00: prepareFTsource T00 form (bg, c1, c2, c3 as in
 *      gev_prepare_source_T00) given lsrc: lT00 = (overwritten), lphi +=,
 *      lchi +=
 *   gev_adj_source_Tij: the phi stencil of prepareFTsource Tij form given lS
 *      (6): lphi += (the Tij part is lS times the coefficient)
 *   gev_adj_link_factor: given u = the B solve applied to lB (it is its own
 *      adjoint) and the link sums L (3): lphi += sum over links of L u at
 *      both ends; u is replaced by lL = (1 + phi_a + phi_b) u
 *   gev_adj_project_ft_scalar: k-space adjoint of projectFTscalar, lSijFT (6
 *      components) = conj(weights) lchiFT / (2 k^4 N) (overwritten) */
int gev_adj_gr_kick_drift(gev_pcls *p, const void *phi, const void *chi, const void *Bi, void *lx, void *lq,
                          void *lphi, void *lchi, void *lB, int nfield_kick, int nfield_drift, double a_kick,
                          double a_drift, double dtau_kick, double dtau_drift);
int gev_adj_gr_deposit(gev_pcls *p, const void *phi, const void *lT00, const void *lTij, const void *lL, void *lx,
                       void *lq, void *lphi, double a, double tij_scale);
int gev_adj_source_T00(gev_ctx *ctx, const void *phi, const void *T00, const void *lsrc, void *lT00, void *lphi,
                       void *lchi, double bg, double c1, double c2, double c3);
int gev_adj_source_Tij(gev_ctx *ctx, const void *phi, const void *lS, void *lphi);
int gev_adj_link_factor(gev_ctx *ctx, const void *phi, const void *L, void *u, void *lphi);
int gev_adj_project_ft_scalar(gev_ctx *ctx, const void *lchiFT, void *lSijFT);

/* ---- background (background.hpp:200-290), Omega_rad/ncdm/fld as given */
double gev_hconf(double a, double fourpiG, double Omega_m, double Omega_Lambda, double Omega_rad);
double gev_rk4bg(double a, double fourpiG, double Omega_m, double Omega_Lambda, double Omega_rad, double dtau);
double gev_particle_horizon(double a, double fourpiG, double Omega_m, double Omega_Lambda, double Omega_rad);

/* ---- the main-loop cycle (main.cu:608-1317) on one GPU.  Fields are
 *      caller-owned device buffers in the context precision. */
typedef struct {
    void *phi, *chi, *Bi, *source, *Sij; /* real: 1,1,3,1,6 components */
    void *scalarFT, *SijFT, *BiFT;       /* Fourier work: 1,6,3 components */
} gev_fields;

typedef struct {
    int gr;              /* gravity theory: 1 = GR (weak field), 0 = Newton */
    int vector_elliptic; /* vector method: 1 = elliptic, 0 = parabolic */
    double fourpiG, Omega_m, Omega_Lambda, Omega_rad, z_relax;
    double a, tau, dtau, dtau_old; /* in/out; dtau is the step to take */
    double maxvel, T00hom, phi_hom; /* out */
} gev_step_state;

int gev_step(gev_pcls *p, const gev_fields *f, gev_step_state *st);

/* per-phase timing of gev_step (measurement support, no reference
 * counterpart; the reference accumulates MPI_Wtime per phase under
 * -DBENCHMARK, main.cu:1340-1371).  When enabled, gev_step records HIP
 * events at its phase boundaries; gev_step_timings returns, for the last
 * step, ids[i] (a GEV_PHASE_* value) and the milliseconds ms[i] spent in
 * that phase.  Returns the number of phases written (<= max).  In the GR
 * elliptic cycle the T0i deposit and the B solve run on a second stream
 * beside the chi solve: there DEPOSIT_T0I marks the fork (~0 ms), SOLVE_CHI
 * is the chi solve running beside them and SOLVE_B the wait that follows. */
enum {
    GEV_PHASE_DEPOSIT_T00_TIJ = 0, /* GR: T00 and Tij in one pass; Newton: the mass deposit */
    GEV_PHASE_SOURCE_T00 = 1, GEV_PHASE_SOLVE_PHI = 2, GEV_PHASE_SOURCE_TIJ = 3, GEV_PHASE_DEPOSIT_T0I = 4,
    GEV_PHASE_SOLVE_CHI = 5, GEV_PHASE_SOLVE_B = 6, GEV_PHASE_KICK_DRIFT = 7, GEV_PHASE_REBIN = 8, GEV_PHASE_COUNT = 9
};
int gev_step_timing(gev_ctx *ctx, int enable);
int gev_step_timings(gev_ctx *ctx, int *ids, float *ms, int max);

/* basic IC generator, mPk path (generateIC_basic, ic_basic.hpp:1834-2505,
 * standard CIC kernel).  Requires a GEV_FP64 context; particles are placed
 * on the cell-centred lattice (N^3 of them).  k_hmpc/P_mpch3: the mPk table
 * (k in h/Mpc, P in (Mpc/h)^3).  On return st->a = 1/(1+z_in), st->tau =
 * particleHorizon(a), dtau_old = 0, maxvel set; *max_disp in box units. */
int gev_ic_generate(gev_pcls *p, const gev_fields *f, gev_step_state *st, uint32_t seed, int ksphere, double boxsize,
                    double z_in, int npk, const double *k_hmpc, const double *P_mpch3, double *max_disp);

/* ---- reductions / utilities */
int gev_field_sum(gev_ctx *ctx, const void *f, int ncomp, double *sum);
int gev_field_fill(gev_ctx *ctx, void *f, int64_t count, double value);

#ifdef __cplusplus
}
#endif
#endif

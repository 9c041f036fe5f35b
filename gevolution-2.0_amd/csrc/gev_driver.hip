// gev_driver.hip -- host-side sequencing on top of the kernels: background
// integration, one main-loop cycle in main.cu order, and the basic IC
// generator.  Everything stays on the context stream; host syncs only where
// a reduction result is needed on the host (T00hom, maxvel).
#include <math.h>

#include <vector>

#include "gev_common.hpp"
#include "gev_kernels.hpp"

#define RC(call)                 \
    do {                         \
        int rc_ = (call);        \
        if (rc_ != GEV_OK) return rc_; \
    } while (0)

static int mark(gev_ctx *c, int id)
{
    if (!c->timing || c->nmarks >= gev_ctx::MAX_MARKS) return GEV_OK;
    hipEvent_t &e = c->marks[c->nmarks];
    if (!e) GEV_HIP(hipEventCreate(&e));
    GEV_HIP(hipEventRecord(e, c->stream));
    c->mark_id[c->nmarks++] = id;
    return GEV_OK;
}

extern "C" {

int gev_step_timing(gev_ctx *c, int enable)
{
    if (!c) { gev_set_error("gev_step_timing: NULL context"); return GEV_ERR_ARG; }
    c->timing = enable != 0;
    c->nmarks = 0;
    return GEV_OK;
}

int gev_step_timings(gev_ctx *c, int *ids, float *ms, int max)
{
    if (!c || !ids || !ms) { gev_set_error("gev_step_timings: NULL argument"); return GEV_ERR_ARG; }
    if (c->nmarks < 2) return 0;
    GEV_HIP(hipEventSynchronize(c->marks[c->nmarks - 1]));
    int n = 0;
    for (int i = 1; i < c->nmarks && n < max; i++, n++) {
        GEV_HIP(hipEventElapsedTime(&ms[n], c->marks[i - 1], c->marks[i]));
        ids[n] = c->mark_id[i];
    }
    return n;
}

int gev_ctx_set_stream(gev_ctx *c, void *s)
{
    if (!c) { gev_set_error("gev_ctx_set_stream: NULL context"); return GEV_ERR_ARG; }
    if (c->own_stream && c->stream != (hipStream_t)s) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        c->own_stream = false;
    }
    c->stream = (hipStream_t)s;
    return GEV_OK;
}

// Hconf, background.hpp:200-208 (no ncdm, no CLASS spline, no fluid)
double gev_hconf(double a, double fourpiG, double om, double ol, double orad)
{
    return sqrt((2. * fourpiG / 3.) * ((om / a) + (ol * a * a) + (orad / a / a)));
}

// rungekutta4bg, background.hpp:235-245
double gev_rk4bg(double a, double fourpiG, double om, double ol, double orad, double dtau)
{
    double k1a = a * gev_hconf(a, fourpiG, om, ol, orad);
    double k2a = (a + k1a * dtau / 2.) * gev_hconf(a + k1a * dtau / 2., fourpiG, om, ol, orad);
    double k3a = (a + k2a * dtau / 2.) * gev_hconf(a + k2a * dtau / 2., fourpiG, om, ol, orad);
    double k4a = (a + k3a * dtau) * gev_hconf(a + k3a * dtau, fourpiG, om, ol, orad);
    return a + dtau * (k1a + 2. * k2a + 2. * k3a + k4a) / 6.;
}

// particleHorizon, background.hpp:268-290: tau = int da/(a Hconf) in s = sqrt(a),
// 16-point Gauss-Legendre on 512 panels (the reference uses GSL qng).
double gev_particle_horizon(double a, double fourpiG, double om, double ol, double orad)
{
    static const double x[8] = {0.0950125098376374, 0.2816035507792589, 0.4580167776572274, 0.6178762444026438,
                                0.7554044083550030, 0.8656312023878318, 0.9445750230732326, 0.9894009349916499};
    static const double w[8] = {0.1894506104550685, 0.1826034150449236, 0.1691565193950025, 0.1495959888165767,
                                0.1246289712555339, 0.0951585116824928, 0.0622535239386479, 0.0271524594117541};
    const double lo = sqrt(a) * 1.0e-7, hi = sqrt(a);
    const int panels = 512;
    const double h = (hi - lo) / panels;
    double sum = 0.0;
    for (int p = 0; p < panels; p++) {
        double c = lo + (p + 0.5) * h;
        for (int j = 0; j < 8; j++)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                double s = c + sgn * 0.5 * h * x[j];
                sum += w[j] * 0.5 * h * 2. / (s * gev_hconf(s * s, 1., om, ol, orad));
            }
    }
    return sum / sqrt(fourpiG);
}

// One main-loop cycle, main.cu:608-1317, one CDM species, no outputs.
int gev_step(gev_pcls *p, const gev_fields *f, gev_step_state *st)
{
    if (!p || !f || !st) { gev_set_error("gev_step: NULL argument"); return GEV_ERR_ARG; }
    gev_ctx *c = p->ctx;
    const int N = c->N;
    const int64_t M = c->M;
    const double dx = 1.0 / (double)N;
    const bool mixed = (c->prec == GEV_MIXED);
    const double om = st->Omega_m, ol = st->Omega_Lambda, orad = st->Omega_rad, fpg = st->fourpiG;
    double a = st->a, dtau = st->dtau, dtau_old = st->dtau_old;
    double maxv2 = 0.0;
    c->nmarks = 0;
    RC(mark(c, -1));

    if (st->gr) {
        // T00 (main.cu:622-651) and Tij (main.cu:750-758), both with the old phi,
        // in one pass over the particles (the outputs need no zeroing).  With
        // the elliptic B on 128 <= N <= 1024 the same pass sums T0i
        // (main.cu:858) on the links into Bi, whose old B no step before the B
        // solve reads; the new phi's link factor is applied in the B solve.
        const bool t0i_links = st->vector_elliptic && gev::chi_from_Tij_supported(c);
        if (t0i_links) RC(gev_project_T00_Tij_T0i_links(p, f->source, f->Sij, f->Bi, a, f->phi, 1.0));
        else RC(gev_project_T00_Tij(p, f->source, f->Sij, a, f->phi, 1.0, 0));
        RC(mark(c, GEV_PHASE_DEPOSIT_T00_TIJ));
        const double H = gev_hconf(a, fpg, om, ol, orad);
        // prepareFTsource (main.cu:725) and the phi solve (main.cu:742-794); the
        // mixed form takes the k=0 mode from the FP64 sum.  For 128 <= N <= 1024
        // the source is formed inside the solve's x pass and its sums stay on
        // the device until the end of the cycle (no wait here).
        bool t00_sums_pending = false;
        if (dtau_old > 0. && gev::chi_from_Tij_supported(c)) {
            RC(mark(c, GEV_PHASE_SOURCE_T00));
            RC(gev_solve_poisson_T00(c, f->phi, f->chi, f->source, om, 3. * H * dx * dx / dtau_old, fpg * dx * dx / a,
                                     3. * H * H * dx * dx, f->phi, f->scalarFT, 1. / (dx * dx), 3. * H / dtau_old,
                                     mixed ? 1 : 0, nullptr));
            t00_sums_pending = true;
            RC(mark(c, GEV_PHASE_SOLVE_PHI));
        } else if (dtau_old > 0.) {
            double s_src = 0, s_res = 0;
            RC(gev_prepare_source_T00(c, f->phi, f->chi, f->source, om, f->source, 3. * H * dx * dx / dtau_old,
                                      fpg * dx * dx / a, 3. * H * H * dx * dx, &s_src, &s_res));
            st->T00hom = s_src / (double)M;
            RC(mark(c, GEV_PHASE_SOURCE_T00));
            RC(gev_solve_poisson(c, f->source, f->phi, f->scalarFT, 1. / (dx * dx), 3. * H / dtau_old, mixed ? 1 : 0,
                                 s_res));
            RC(mark(c, GEV_PHASE_SOLVE_PHI));
        }
        // chi (main.cu:822-902) and B (main.cu:908-955)
        if (st->vector_elliptic) {
            // Sij is formed inside the chi solve's x passes.  The B solve (and
            // the T0i deposit where it is not summed on the links above) share
            // no buffer with the chi solve (Bi, BiFT against chi, SijFT,
            // scalarFT; both read phi): they run on the side stream while the
            // chi solve runs on the main one, which waits for them before the
            // kick.  Phases: deposit_T0i is the fork (~0), solve_chi the chi
            // solve beside them, solve_B the wait after it.
            RC(mark(c, GEV_PHASE_SOURCE_TIJ));
            GEV_HIP(hipEventRecord(c->fork, c->stream));
            GEV_HIP(hipStreamWaitEvent(c->side, c->fork, 0));
            hipStream_t main_stream = c->stream;
            c->stream = c->side;
            int rc_side;
            if (t0i_links) {
                rc_side = gev_solve_vector_T0i_links(c, f->Bi, f->phi, f->BiFT, fpg * dx * dx, 0.0);
            } else {
                rc_side = gev_project_T0i(p, f->Bi, f->phi, 1.0, 0);
                if (!rc_side) rc_side = gev_solve_vector(c, f->Bi, f->BiFT, fpg * dx * dx, 0.0);
            }
            c->stream = main_stream;
            GEV_HIP(hipEventRecord(c->join, c->side));
            RC(rc_side);
            RC(mark(c, GEV_PHASE_DEPOSIT_T0I));
            RC(gev_solve_chi_Tij(c, f->phi, f->Sij, 2. * fpg * dx * dx / a, f->chi, f->SijFT, f->scalarFT));
            RC(mark(c, GEV_PHASE_SOLVE_CHI));
            GEV_HIP(hipStreamWaitEvent(c->stream, c->join, 0));
        } else {
            RC(gev_prepare_source_Tij(c, f->phi, f->Sij, f->Sij, 2. * fpg * dx * dx / a));
            RC(mark(c, GEV_PHASE_SOURCE_TIJ));
            // parabolic: SijFT feeds evolveFTvector, B persists in k-space
            RC(gev_fft_r2c(c, f->Sij, f->SijFT, 6));
            RC(gev_project_ft_scalar(c, f->SijFT, f->scalarFT, 0));
            RC(gev_fft_c2r(c, f->scalarFT, f->chi, 1));
            RC(mark(c, GEV_PHASE_SOLVE_CHI));
            RC(gev_evolve_ft_vector(c, f->SijFT, f->BiFT, a * a * dtau_old));
            RC(gev_fft_c2r(c, f->BiFT, f->Bi, 3));
            RC(gev_fft_r2c(c, f->Bi, f->BiFT, 3));  // c2r clobbers its input
        }
        RC(mark(c, GEV_PHASE_SOLVE_B));
        // kick (main.cu:1227-1240), half background step, drift (:1253-1266):
        // one fused pass over the particles, then the re-binning scatter
        const int nf_k = (1. / a < st->z_relax + 1.) ? 3 : 2;
        const double a_d = gev_rk4bg(a, fpg, om, ol, orad, 0.5 * dtau);
        const int nf_d = (1. / a_d < st->z_relax + 1.) ? 3 : 0;
        RC(gev_kick_drift(p, 1, (dtau + dtau_old) / 2., dtau, f->phi, f->chi, f->Bi, nf_k, nf_d, a, a_d,
                          a * a * (double)N, a_d * a_d * (double)N, nullptr, 0));
        RC(mark(c, GEV_PHASE_KICK_DRIFT));
        RC(gev_pcls_rebin_counted(p));
        RC(mark(c, GEV_PHASE_REBIN));
        RC(gev::read_max_slot(c, &maxv2));  // the scatter leaves the kick's maximum alone
        if (t00_sums_pending) {
            double sums[2];
            RC(gev_ctx_sums(c, sums));  // nothing since the phi solve reduced into these slots
            st->T00hom = sums[0] / (double)M;
        }
        a = gev_rk4bg(a_d, fpg, om, ol, orad, 0.5 * dtau);
        double mv = sqrt(maxv2);
        st->maxvel = mv / sqrt(mv * mv + 1.0); // main.cu:1279-1283
    } else {
        RC(gev_project_mass(p, f->source, 1.0, 0));
        RC(mark(c, GEV_PHASE_DEPOSIT_T00_TIJ));
        RC(gev_fft_r2c(c, f->source, f->scalarFT, 1));
        RC(gev_poisson(c, f->scalarFT, f->scalarFT, fpg / a, 0.0)); // main.cu:765
        RC(gev_fft_c2r(c, f->scalarFT, f->phi, 1));
        RC(mark(c, GEV_PHASE_SOLVE_PHI));
        const double a_d = gev_rk4bg(a, fpg, om, ol, orad, 0.5 * dtau);
        RC(gev_kick_drift(p, 0, (dtau + dtau_old) / 2., dtau, f->phi, nullptr, nullptr, 1, 0, a, a_d, 1.0, 1.0,
                          nullptr, 0));
        RC(mark(c, GEV_PHASE_KICK_DRIFT));
        RC(gev_pcls_rebin_counted(p));
        RC(mark(c, GEV_PHASE_REBIN));
        RC(gev::read_max_slot(c, &maxv2));
        a = gev_rk4bg(a_d, fpg, om, ol, orad, 0.5 * dtau);
        st->maxvel = sqrt(maxv2);
    }
    st->a = a;
    st->tau += dtau;
    st->dtau_old = dtau;
    (void)M;
    return GEV_OK;
}

// generateIC_basic, mPk path with the standard CIC kernel (ic_basic.hpp:1834-2505)
int gev_ic_generate(gev_pcls *p, const gev_fields *f, gev_step_state *st, uint32_t seed, int ksphere, double boxsize,
                    double z_in, int npk, const double *k_hmpc, const double *P_mpch3, double *max_disp)
{
    if (!p || !f || !st || !k_hmpc || !P_mpch3) { gev_set_error("gev_ic_generate: NULL argument"); return GEV_ERR_ARG; }
    gev_ctx *c = p->ctx;
    if (c->prec != GEV_FP64) { gev_set_error("gev_ic_generate: needs a GEV_FP64 context"); return GEV_ERR_ARG; }
    const int N = c->N;
    const double a = 1. / (1. + z_in);
    const double om = st->Omega_m, ol = st->Omega_Lambda, orad = st->Omega_rad, fpg = st->fourpiG;
    const double H = gev_hconf(a, fpg, om, ol, orad);

    // loadPowerSpectrum conversion (ic_basic.hpp:445-447, 1910-1914)
    std::vector<double> kL(npk), amp(npk);
    for (int i = 0; i < npk; i++) {
        if (k_hmpc[i] < 0 || P_mpch3[i] < 0) { gev_set_error("gev_ic_generate: negative table entry"); return GEV_ERR_ARG; }
        kL[i] = k_hmpc[i] * boxsize;
        amp[i] = sqrt(0.5 * P_mpch3[i] * boxsize) / boxsize / boxsize;
    }
    // standard CIC kernel (ic_basic.hpp:810-838): renorm N^2 x 7-point stencil
    std::vector<double> ker(8, 0.0);
    const double renorm = (double)N * (double)N;
    double *src = (double *)f->source;
    RC(gev_field_fill(c, src, c->M, 0.0));
    struct { int64_t idx; double v; } pts[7] = {
        {0, 6.0 * renorm}, {1, -renorm}, {N - 1, -renorm}, {N, -renorm}, {(int64_t)N * (N - 1), -renorm},
        {(int64_t)N * N, -renorm}, {(int64_t)N * N * (N - 1), -renorm}};
    for (auto &pt : pts) GEV_HIP(hipMemcpyAsync(src + pt.idx, &pt.v, sizeof(double), hipMemcpyHostToDevice, c->stream));
    GEV_HIP(hipStreamSynchronize(c->stream));
    RC(gev_fft_r2c(c, f->source, f->SijFT, 1));   // kernel in k-space -> SijFT comp 0
    RC(gev_ic_displacement_ft(c, seed, ksphere, st->gr ? H * H : 0.0, npk, kL.data(), amp.data(), f->SijFT,
                              f->scalarFT));
    RC(gev_fft_c2r(c, f->scalarFT, f->chi, 1));   // chi = displacement potential

    RC(gev_pcls_init_lattice(p));
    double md = 0.0;
    RC(gev_displace_particles(p, f->chi, 1.0, &md));
    if (max_disp) *max_disp = md;

    // phi from the displaced CIC density (ic_basic.hpp:2406-2428)
    RC(gev_project_mass(p, f->source, 1.0, 0));
    RC(gev_fft_r2c(c, f->source, f->scalarFT, 1));
    RC(gev_set_k0(c, f->scalarFT, 0.0, 0.0));
    RC(gev_poisson(c, f->scalarFT, f->scalarFT, fpg / a, 3.0 * (double)st->gr * (H * H + fpg * om / a)));
    RC(gev_fft_c2r(c, f->scalarFT, f->phi, 1));

    // velocities (ic_basic.hpp:2431-2437), Omega_m(a) with Omega_rad = 0 here
    double om_a = om / (om + ol * a * a * a + orad / a);
    double rescale = a / H / (1.5 * om_a + 2.0 * (orad / (om * a + ol * a * a * a * a + orad)));
    double vmax2 = 0.0;
    RC(gev_initialize_q(p, f->phi, rescale, &vmax2));
    double maxvel = sqrt(vmax2) / a;

    // B (ic_basic.hpp:2469-2479)
    RC(gev_project_T0i(p, f->Bi, f->phi, 1.0, 0));
    RC(gev_fft_r2c(c, f->Bi, f->BiFT, 3));
    RC(gev_project_ft_vector(c, f->BiFT, f->BiFT, fpg / (double)N / (double)N, 0.0));
    RC(gev_fft_c2r(c, f->BiFT, f->Bi, 3));

    // chi (ic_basic.hpp:2481-2493)
    RC(gev_project_Tij(p, f->Sij, a, f->phi, 1.0, 0));
    RC(gev_prepare_source_Tij(c, f->phi, f->Sij, f->Sij, 2. * fpg / a / (double)N / (double)N));
    RC(gev_fft_r2c(c, f->Sij, f->SijFT, 6));
    RC(gev_project_ft_scalar(c, f->SijFT, f->scalarFT, 0));
    RC(gev_fft_c2r(c, f->scalarFT, f->chi, 1));

    if (st->gr) maxvel /= sqrt(maxvel * maxvel + 1.0); // main.cu:462-467
    st->a = a;
    st->tau = gev_particle_horizon(a, fpg, om, ol, orad);
    st->dtau_old = 0.0;
    st->maxvel = maxvel;
    return gev_ctx_sync(c);
}

}  // extern "C"

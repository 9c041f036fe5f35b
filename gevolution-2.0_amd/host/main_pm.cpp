// main_pm.cpp -- a main.cu-shaped driver over the C++ façade: basic mPk
// initial conditions, then the particle-mesh cycle of main.cu:608-1317
// (Newtonian or weak-field GR with the elliptic vector method), called
// function by function through the gevolution-named interfaces.
//
//   main_pm -n N -L boxsize -zin z -zend z -steps S [-newton] [-seed s]
//           [-zrelax z] -pk mpk.txt [-out final.bin] [-gadget snap] [-express restart]
//           [-lightcone prefix -lcvertex x,y,z -lcz z] [-spectra prefix [-numbins n] [-zpk z]]
//   main_pm -n N -template crystal.dat -out lattice.bin
//
// The mPk file has the loadPowerSpectrum format (ic_basic.hpp:370-449): two
// columns, k [h/Mpc] and P(k) [(Mpc/h)^3].  Steps are uniform in conformal
// time between z_in and z_end.  -out writes the final particles as
// int64 n, int64 id[n], double pos[3][n], double q[3][n].  -gadget writes them
// as a Gadget-2 snapshot with the header output.hpp:418-466 fills (Omega_cdm
// particles, no tracer thinning), -express as an express restart file.
// -lightcone writes, each cycle before the kick, the full-sky particle
// lightcone of an observer at lcvertex (box units) and redshift lcz as
// <prefix>_<cycle>_cdm (writeLightcones' Gadget branch, output.hpp:2094-2160,
// and the ID backlog hand-over, :2166-2880).  -spectra writes, at the end, the
// deconvolved CIC density spectrum as <prefix>_deltaN.dat in writeSpectra's
// delta_N form (output.hpp:3017-3021, writePowerSpectrum tools.hpp:370-452),
// and the spectra of phi and (GR) B of the last cycle as <prefix>_phi.dat and
// <prefix>_B.dat (output.hpp:3176-3184, 3456-3461).
// -zpk z adds one power-spectrum redshift (sim.z_pk): the same spectrum is
// written inside the cycle as <prefix>000_deltaN.dat as main.cu:1080-1131 does
// with EXACT_OUTPUT_REDSHIFTS (this driver is built with it): a plain spectrum
// at the last cycle before z, then at the first cycle past z the spectrum
// interpolated in redshift to z against that file.
// -template builds the particle lattice as generateIC_basic does before the
// displacement (loadHomogeneousTemplate, initialize with the reference's
// arguments, initializeParticlePositions, updateRowBuffers; ic_basic.hpp:
// 205-330, 1612-1644, 2197-2211), writes it with -out through
// sampleParticles(Real*, Real*, long*, n) and exits.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "gevolution_pm.hpp"

using namespace gevpm;

namespace {

constexpr double C_SPEED_OF_LIGHT = 2997.92458;  // metadata.hpp:97 (100 km/s units)
constexpr double C_RHO_CRIT = 2.77459457e11;     // metadata.hpp:98 [M_sun h^2 / Mpc^3]
constexpr double GADGET_LENGTH_CONVERSION = 0.001;  // metadata.hpp:111 (kpc/h)
constexpr double GADGET_MASS_CONVERSION = 1.0e10;   // metadata.hpp:114 (1e10 M_sun/h)

struct Options {
    int N = 32, steps = 5, newton = 0;
    uint32_t seed = 42;
    double L = 320.0, zin = 100.0, zend = 50.0, zrelax = -2.0;
    std::string pk, out, gadget, express, lightcone, spectra, tmpl;
    int numbins = 32;
    double lcvertex[3] = {0.5, 0.5, 0.5}, lcz = 0.0;
    double zpk = -1.0;  // -1: no power-spectrum redshift inside the cycle
};

Options parse(int argc, char **argv)
{
    Options o;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
            return argv[++i];
        };
        if (a == "-n") o.N = std::atoi(next());
        else if (a == "-L") o.L = std::atof(next());
        else if (a == "-zin") o.zin = std::atof(next());
        else if (a == "-zend") o.zend = std::atof(next());
        else if (a == "-steps") o.steps = std::atoi(next());
        else if (a == "-newton") o.newton = 1;
        else if (a == "-seed") o.seed = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "-zrelax") o.zrelax = std::atof(next());
        else if (a == "-pk") o.pk = next();
        else if (a == "-out") o.out = next();
        else if (a == "-gadget") o.gadget = next();
        else if (a == "-express") o.express = next();
        else if (a == "-lightcone") o.lightcone = next();
        else if (a == "-spectra") o.spectra = next();
        else if (a == "-numbins") o.numbins = std::atoi(next());
        else if (a == "-zpk") o.zpk = std::atof(next());
        else if (a == "-template") o.tmpl = next();
        else if (a == "-lcvertex") {
            if (std::sscanf(next(), "%lf,%lf,%lf", &o.lcvertex[0], &o.lcvertex[1], &o.lcvertex[2]) != 3)
                throw std::runtime_error("-lcvertex wants x,y,z");
        } else if (a == "-lcz") o.lcz = std::atof(next());
        else throw std::runtime_error("unknown option " + a);
    }
    if (!o.tmpl.empty()) {
        if (o.out.empty()) throw std::runtime_error("-template needs -out");
        return o;
    }
    if (o.pk.empty()) throw std::runtime_error("-pk <mPk file> is required");
    if (o.zpk >= 0.0 && o.spectra.empty()) throw std::runtime_error("-zpk needs -spectra");
    if (o.zrelax < -1.0) o.zrelax = o.zin;  // parser.hpp:1468 default
    return o;
}

void load_mpk(const std::string &path, std::vector<double> &k, std::vector<double> &P)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("unable to open mPk file " + path);
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        double a, b;
        if (ss >> a >> b) {
            k.push_back(a);
            P.push_back(b);
        }
    }
    if (k.size() < 4) throw std::runtime_error("mPk file " + path + " has fewer than 4 rows");
}

// FP64 field of the IC lattice -> working-precision field (host round trip, once)
template <typename S> void convert_field(Lattice &src_lat, void *src, Field<Real> &dst)
{
    std::vector<S> h((size_t)dst.count());
    src_lat.sync();
    hip_check(hipMemcpy(h.data(), src, sizeof(S) * h.size(), hipMemcpyDeviceToHost), "convert_field");
    std::vector<Real> r(h.begin(), h.end());
    dst.from_host(r);
}

void write_particles(const std::string &path, int64_t n, const std::vector<int64_t> &id, const std::vector<double> &pos,
                     const std::vector<double> &q)
{
    std::ofstream f(path, std::ios::binary);
    f.write((const char *)&n, sizeof(n));
    f.write((const char *)id.data(), sizeof(int64_t) * n);
    f.write((const char *)pos.data(), sizeof(double) * 3 * n);
    f.write((const char *)q.data(), sizeof(double) * 3 * n);
    if (!f) throw std::runtime_error("writing " + path + " failed");
}

// the particle lattice of generateIC_basic from a homogeneous template
int template_lattice(const Options &opt)
{
    const int N = opt.N;
    const int64_t M = (int64_t)N * N * N;
    long numpart = 0;
    float *pcldata = nullptr;
    loadHomogeneousTemplate(opt.tmpl.c_str(), numpart, pcldata);
    int numtile = (int)std::lround(std::cbrt((double)M / (double)numpart));
    if ((int64_t)numpart * numtile * numtile * numtile != M) {
        std::free(pcldata);
        throw std::runtime_error("template does not tile the N^3 lattice");
    }
    Lattice lat(N);
    Particles pcls;
    part_simple_info info;
    info.mass = 0.31 / (double)M;
    Real boxSize[3] = {1., 1., 1.};
    const uint64_t extra = 1024;  // PCL_EXTRA_CAPACITY-style slack
    pcls.initialize(info, &lat, boxSize, (uint64_t)M + extra, extra);
    initializeParticlePositions(numpart, pcldata, numtile, pcls);
    pcls.updateRowBuffers();
    std::free(pcldata);
    const int64_t n = pcls.num_particles();
    std::vector<Real> pos_s(3 * n), vel_s(3 * n);
    std::vector<long> id_s(n);
    if (pcls.sampleParticles(pos_s.data(), vel_s.data(), id_s.data(), (uint64_t)n) != n)
        throw std::runtime_error("sampleParticles returned too few particles");
    std::vector<int64_t> id(n);
    std::vector<double> pos(3 * n), q(3 * n);
    for (int64_t i = 0; i < n; i++) {
        id[i] = id_s[i];
        for (int d = 0; d < 3; d++) {
            pos[d * n + i] = pos_s[3 * i + d];
            q[d * n + i] = vel_s[3 * i + d];
        }
    }
    write_particles(opt.out, n, id, pos, q);
    std::printf("template: %ld particles tiled %d times per axis, %lld particles\n", numpart, numtile, (long long)n);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        const Options opt = parse(argc, argv);
        if (!opt.tmpl.empty()) return template_lattice(opt);
        const int N = opt.N;
        const int64_t M = (int64_t)N * N * N;
        const double dx = 1.0 / N;
        cosmology cosmo;
        const double fourpiG = 1.5 * opt.L * opt.L / C_SPEED_OF_LIGHT / C_SPEED_OF_LIGHT;  // main.cu:418
        std::vector<double> kt, Pt;
        load_mpk(opt.pk, kt, Pt);

        // ---- initial conditions (generateIC_basic, FP64) -> working precision
        Lattice lat(N);
        Particles pcls;
        part_simple_info info;
        info.mass = cosmo.Omega_m / (double)M;
        pcls.initialize(info, &lat, (uint64_t)M);
        Field<Real> phi(lat), chi(lat), Bi(lat, 3), source(lat), Sij(lat, 6, symmetric);
        Field<Cplx> scalarFT(lat), SijFT(lat, 6, symmetric), BiFT(lat, 3);
        double a, tau, dtau_old = 0.0, maxvel;
        {
            Lattice ic_lat(N, GEV_FP64);
            std::vector<void *> bufs;
            auto alloc = [&](int64_t bytes) {
                void *d = nullptr;
                hip_check(hipMalloc(&d, (size_t)bytes), "hipMalloc(IC)");
                hip_check(hipMemset(d, 0, (size_t)bytes), "hipMemset(IC)");
                bufs.push_back(d);
                return d;
            };
            const int64_t NK = ic_lat.ksites();
            gev_fields f{alloc(8 * M), alloc(8 * M), alloc(8 * 3 * M), alloc(8 * M), alloc(8 * 6 * M),
                         alloc(16 * NK), alloc(16 * 6 * NK), alloc(16 * 3 * NK)};
            gev_pcls *icp = nullptr;
            check(gev_pcls_create(ic_lat.ctx(), M, info.mass, &icp), "gev_pcls_create(IC)");
            gev_step_state st{};
            st.gr = opt.newton ? 0 : 1;
            st.vector_elliptic = 1;
            st.fourpiG = fourpiG;
            st.Omega_m = cosmo.Omega_m;
            st.Omega_Lambda = cosmo.Omega_Lambda;
            st.Omega_rad = cosmo.Omega_rad;
            st.z_relax = opt.zrelax;
            double max_disp = 0;
            check(gev_ic_generate(icp, &f, &st, opt.seed, 1, opt.L, opt.zin, (int)kt.size(), kt.data(), Pt.data(),
                                  &max_disp),
                  "generateIC_basic");
            std::vector<double> pos(3 * M), q(3 * M);
            std::vector<int64_t> id(M);
            check(gev_pcls_store(icp, pos.data(), q.data(), id.data(), 0), "IC store");
            check(gev_pcls_load(pcls.handle(), M, pos.data(), q.data(), id.data(), 0), "IC load");
            convert_field<double>(ic_lat, f.phi, phi);
            convert_field<double>(ic_lat, f.chi, chi);
            convert_field<double>(ic_lat, f.Bi, Bi);
            a = st.a;
            tau = st.tau;
            maxvel = st.maxvel;
            gev_pcls_destroy(icp);
            for (void *d : bufs) (void)hipFree(d);
            std::printf("IC: N=%d L=%g z_in=%g a=%.8f tau=%.8f max_disp=%.3e cells\n", N, opt.L, opt.zin, a, tau,
                        max_disp * N);
        }
        const double tau_end = particleHorizon(1.0 / (1.0 + opt.zend), fourpiG, cosmo);
        const double dtau = (tau_end - tau) / opt.steps;

        PlanFFT<Cplx> plan_source(&source, &scalarFT), plan_phi(&phi, &scalarFT), plan_chi(&chi, &scalarFT);
        PlanFFT<Cplx> plan_Sij(&Sij, &SijFT, 6), plan_Bi(&Bi, &BiFT, 3);
        // main.cu:312-403: the field lists and the parameter buffer of the functor calls
        Field<Real> *update_cdm_fields[3] = {&phi, &chi, &Bi};
        Field<Real> *project_Tij_fields[2] = {&Sij, &phi};
        Field<Real> *project_T0i_fields[2] = {&Bi, &phi};
        double f_params[7] = {0., 0., 0., 0., 0., 0., 0.};
        DeviceStagingBuffer<double> d_f_params(7);
        const double numpts3d = (double)M;
        const double Omega_bg = cosmo.Omega_cdm;  // cosmo.Omega_cdm + cosmo.Omega_b + bg_ncdm(a, cosmo), CDM only
        double T00hom = 0., phi_hom = 0.;
        (void)phi_hom;

        lightcone_geometry lc{};
        for (int d = 0; d < 3; d++) lc.vertex[d] = opt.lcvertex[d];
        lc.z = opt.lcz;
        lc.direction[2] = 1.0;
        lc.opening = -1.0;
        lc.distance[0] = 1e30;
        lc.distance[1] = 0.0;
        LightconeIDBacklog IDbacklog(&lat);
        auto write_lightcone = [&](int cycle) {
            if (opt.lightcone.empty()) return;
            const double d = particleHorizon(1. / (1. + lc.z), fourpiG, cosmo);  // output.hpp:1213
            const double shell = d - tau + 0.5 * dtau_old;
            if (lc.distance[0] > shell && lc.distance[1] <= shell && shell > 0.) {
                double domain[6], vertex[MAX_INTERSECTS][3];
                for (int k = 0; k < 3; k++) {
                    domain[k] = -0.5 / N;
                    domain[k + 3] = (N + 0.5) / N;
                }
                const int nv = findIntersectingLightcones(lc, d - tau + (0.5 + GEV_LIGHTCONE_IDCHECK_ZONE) * dtau_old,
                                                          d - tau - 0.5 * dtau, domain, vertex);
                if (nv > MAX_INTERSECTS) throw std::runtime_error("lightcone meets more than MAX_INTERSECTS copies");
                gadget2_header hdr{};
                hdr.num_files = 1;
                hdr.Omega0 = cosmo.Omega_m;
                hdr.OmegaLambda = cosmo.Omega_Lambda;
                hdr.HubbleParam = cosmo.h;
                hdr.BoxSize = opt.L / GADGET_LENGTH_CONVERSION;
                hdr.time = a;
                hdr.redshift = 1. / a - 1.;
                hdr.mass[1] = C_RHO_CRIT * cosmo.Omega_cdm * opt.L * opt.L * opt.L / (double)M / GADGET_MASS_CONVERSION;
                char name[32];
                std::snprintf(name, sizeof(name), "_%04d_cdm", cycle);
                pcls.saveGadget2(opt.lightcone + name, hdr, lc, d - tau, dtau, dtau_old, a * Hconf(a, fourpiG, cosmo),
                                 vertex, nv, IDbacklog, &phi, 1);
            }
            IDbacklog.finalize();
        };

        // writeSpectra's particle branch (output.hpp:2997-3022): the
        // deconvolved CIC density spectrum as delta_N, z_target as sim.z_pk
        auto write_deltaN = [&](const std::string &filename, double z_target) {
            projection_init(&source);
            scalarProjectionCIC_project(&pcls, &source);
            plan_source.execute(FFT_FORWARD);
            const int nb = opt.numbins;
            std::vector<Real> kbin(nb), power(nb), kscatter(nb), pscatter(nb);
            std::vector<int> occupation(nb);
            extractPowerSpectrum(scalarFT, kbin.data(), power.data(), kscatter.data(), pscatter.data(),
                                 occupation.data(), nb, true, KTYPE_LINEAR);
            const Real np3 = (Real)M;
            writePowerSpectrum(kbin.data(), power.data(), kscatter.data(), pscatter.data(), occupation.data(), nb,
                               (Real)opt.L, np3 * np3 * (Real)(2. * M_PI * M_PI * cosmo.Omega_m * cosmo.Omega_m),
                               filename.c_str(), "power spectrum of delta_N", a, z_target);
        };
        const int num_pk = opt.zpk >= 0.0 ? 1 : 0;
        int pkcount = 0;
        const std::string pkfile = opt.spectra + "000_deltaN.dat";  // "%s%s%03d_deltaN.dat" (output.hpp:3019)

        // The cycle of main.cu:608-1317 for one CDM species (no baryons, ncdm,
        // radiation, fluid or tensor evolution; vector method elliptic), call
        // for call in main.cu's order.  Ghost-layer folds and halo updates
        // are kept as calls: they have nothing to do on this lattice.
        const int gr_flag = opt.newton ? 0 : 1;
        Particles *pcls_cdm = &pcls;
        for (int cycle = 0; cycle < opt.steps; cycle++) {
            // construct stress-energy tensor (main.cu:614-665)
            projection_init(&source);  // thrust::fill_n(source.data(), ..., 0)
            if (gr_flag > 0)
                projection_T00_project(pcls_cdm, &source, a, &phi);
            else
                scalarProjectionCIC_project(pcls_cdm, &source);
            projection_T00_comm(&source);
            projection_init(&Sij);  // main.cu:678-681

            // solve phi (main.cu:719-806)
            double T00_terms = 0.;
            if (gr_flag > 0 && dtau_old > 0.) {
                T00hom = prepareFTsource(phi, chi, source, Omega_bg, source, 3. * Hconf(a, fourpiG, cosmo) * dx * dx / dtau_old,
                                         fourpiG * dx * dx / a, 3. * Hconf(a, fourpiG, cosmo) * Hconf(a, fourpiG, cosmo) * dx * dx,
                                         &T00_terms);  // main.cu:725
                T00hom /= numpts3d;
            }
            if (gr_flag == 0 || dtau_old > 0.) {
                plan_source.execute(FFT_FORWARD);  // main.cu:742
#ifdef SINGLE
                // the mixed form restores the FP64 mean of the source (no reference counterpart)
                if (gr_flag > 0) set_k0(scalarFT, T00_terms);
#endif
            }
            if (gr_flag > 0) {
                f_params[0] = a;
                f_params[1] = 1.;
                d_f_params.copy_from_host(f_params, 7);
                projection_Tij_project_Async(pcls_cdm, project_Tij_fields, 2, d_f_params.data());  // main.cu:756
            }
            if (gr_flag == 0)
                solveModifiedPoissonFT(scalarFT, scalarFT, fourpiG / a);  // main.cu:765
            else if (dtau_old > 0.)
                solveModifiedPoissonFT(scalarFT, scalarFT, 1. / (dx * dx), 3. * Hconf(a, fourpiG, cosmo) / dtau_old);  // :769
            if (gr_flag > 0) {
                hip_check(hipDeviceSynchronize(), "projection_Tij_project_Async");  // main.cu:776
                projection_Tij_comm(&Sij);                                         // main.cu:784
            }
            if (gr_flag == 0 || dtau_old > 0.) {
                plan_phi.execute(FFT_BACKWARD);  // main.cu:794
                phi.updateHalo();                // main.cu:802
            }
            // phi_hom = scalarFT(k = 0) (main.cu:808-809) is read only for the
            // background file, which this driver does not write

            if (gr_flag > 0) {
                // solve chi (main.cu:810-906)
                prepareFTsource(phi, Sij, Sij, 2. * fourpiG * dx * dx / a);  // main.cu:822
                plan_Sij.execute(FFT_FORWARD);                               // main.cu:829
                projection_init(&Bi);                                        // main.cu:841
                f_params[0] = 1.;
                d_f_params.copy_from_host(f_params, 1);
                projection_T0i_project_Async(pcls_cdm, project_T0i_fields, 2, d_f_params.data());  // main.cu:858
                projectFTscalar(SijFT, scalarFT);                                                   // main.cu:873
                hip_check(hipDeviceSynchronize(), "projection_T0i_project_Async");                 // main.cu:879
                projection_T0i_comm(&Bi);                                                           // main.cu:887
                plan_chi.execute(FFT_BACKWARD);                                                     // main.cu:895
                chi.updateHalo();                                                                   // main.cu:902

                // solve B (main.cu:908-957)
                plan_Bi.execute(FFT_FORWARD);                    // main.cu:918
                projectFTvector(BiFT, BiFT, fourpiG * dx * dx);  // main.cu:925
                plan_Bi.execute(FFT_BACKWARD);                   // main.cu:948
                Bi.updateHalo();                                 // main.cu:955
            }

            write_lightcone(cycle);  // main.cu:1064-1070, outputs before the particle update

            // power spectra (main.cu:1080-1103)
            if (pkcount < num_pk && 1. / a < opt.zpk + 1.) {
                std::printf(" writing power spectra at z = %f (cycle %d)\n", 1. / a - 1., cycle);
                write_deltaN(pkfile, opt.zpk);
                pkcount++;
            }
#ifdef EXACT_OUTPUT_REDSHIFTS
            {  // main.cu:1106-1131: the spectrum before the crossing, for the interpolation
                double tmp = a;
                rungekutta4bg(tmp, fourpiG, cosmo, 0.5 * dtau);
                rungekutta4bg(tmp, fourpiG, cosmo, 0.5 * dtau);
                if (pkcount < num_pk && 1. / tmp < opt.zpk + 1.) write_deltaN(pkfile, opt.zpk);
            }
#endif

            // cdm particle update (main.cu:1225-1283)
            f_params[0] = a;
            f_params[1] = a * a * N;  // a * a * sim.numpts
            d_f_params.copy_from_host(f_params, 7);
            if (gr_flag > 0)
                maxvel = pcls_cdm->updateVel(update_q_functor(), (dtau + dtau_old) / 2., update_cdm_fields,
                                             (1. / a < opt.zrelax + 1. ? 3 : 2), d_f_params.data());
            else
                maxvel = pcls_cdm->updateVel(update_q_Newton_functor(), (dtau + dtau_old) / 2., update_cdm_fields, 1,
                                             d_f_params.data());
            rungekutta4bg(a, fourpiG, cosmo, 0.5 * dtau);  // evolve background by half a time step
            f_params[0] = a;
            f_params[1] = a * a * N;
            d_f_params.copy_from_host(f_params, 7);
            if (gr_flag > 0)
                pcls_cdm->moveParticles(update_pos_functor(), dtau, update_cdm_fields, (1. / a < opt.zrelax + 1. ? 3 : 0),
                                        d_f_params.data());
            else
                pcls_cdm->moveParticles(update_pos_Newton_functor(), dtau, NULL, 0, d_f_params.data());
            rungekutta4bg(a, fourpiG, cosmo, 0.5 * dtau);  // evolve background by half a time step
            if (gr_flag > 0) maxvel /= std::sqrt(maxvel * maxvel + 1.0);  // main.cu:1279-1283
            tau += dtau;
            dtau_old = dtau;
            std::printf("cycle %d z=%.6f a=%.8f tau=%.8f maxvel=%.6e T00hom=%.10e\n", cycle, 1. / a - 1., a, tau, maxvel,
                        T00hom);
        }
        lat.sync();

        if (!opt.spectra.empty()) {  // at the final time
            write_deltaN(opt.spectra + "_deltaN.dat", -1.0);
            // the metric spectra of writeSpectra (output.hpp:3176-3184, 3456-3461)
            // for the fields of the last cycle
            const int nb = opt.numbins;
            std::vector<Real> kbin(nb), power(nb), kscatter(nb), pscatter(nb);
            std::vector<int> occupation(nb);
            const Real np3 = (Real)M;
            plan_phi.execute(FFT_FORWARD);
            extractPowerSpectrum(scalarFT, kbin.data(), power.data(), kscatter.data(), pscatter.data(),
                                 occupation.data(), nb, false, KTYPE_LINEAR);
            writePowerSpectrum(kbin.data(), power.data(), kscatter.data(), pscatter.data(), occupation.data(), nb,
                               (Real)opt.L, np3 * np3 * (Real)(2. * M_PI * M_PI), (opt.spectra + "_phi.dat").c_str(),
                               "power spectrum of phi", a);
            if (gr_flag > 0) {
                plan_Bi.execute(FFT_FORWARD);
                extractPowerSpectrum(BiFT, kbin.data(), power.data(), kscatter.data(), pscatter.data(),
                                     occupation.data(), nb, false, KTYPE_LINEAR);
                writePowerSpectrum(kbin.data(), power.data(), kscatter.data(), pscatter.data(), occupation.data(), nb,
                                   (Real)opt.L, (Real)(a * a * a * a * N * N * 2. * M_PI * M_PI),
                                   (opt.spectra + "_B.dat").c_str(), "power spectrum of B", a);
            }
        }

        if (!opt.out.empty()) {
            std::vector<part_simple> ps;
            pcls.sampleParticles(ps);
            const int64_t n = (int64_t)ps.size();
            std::vector<int64_t> id(n);
            std::vector<double> pos(3 * n), q(3 * n);
            for (int64_t i = 0; i < n; i++) {
                id[i] = ps[i].ID;
                for (int d = 0; d < 3; d++) {
                    pos[d * n + i] = ps[i].pos[d];
                    q[d * n + i] = ps[i].vel[d];
                }
            }
            write_particles(opt.out, n, id, pos, q);
        }
        if (!opt.gadget.empty() || !opt.express.empty()) {
            // output.hpp:418-466 at the current time (no exact-redshift step)
            gadget2_header hdr{};
            const int64_t n = pcls.num_particles();
            hdr.num_files = 1;
            hdr.Omega0 = cosmo.Omega_m;
            hdr.OmegaLambda = cosmo.Omega_Lambda;
            hdr.HubbleParam = cosmo.h;
            hdr.BoxSize = opt.L / GADGET_LENGTH_CONVERSION;
            hdr.time = a;
            hdr.redshift = 1. / a - 1.;
            hdr.npart[1] = hdr.npartTotal[1] = (uint32_t)(n % (1ll << 32));
            hdr.npartTotalHW[1] = (uint32_t)(n / (1ll << 32));
            hdr.mass[1] = C_RHO_CRIT * cosmo.Omega_cdm * opt.L * opt.L * opt.L / (double)n / GADGET_MASS_CONVERSION;
            if (!opt.gadget.empty()) pcls.saveGadget2(opt.gadget, hdr, 1, 0., 0., &phi);
            if (!opt.express.empty()) pcls.saveExpress(opt.express, hdr);
        }
        std::printf("done: %d cycles, final z=%.6f\n", opt.steps, 1. / a - 1.);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "main_pm: %s\n", e.what());
        return 1;
    }
}

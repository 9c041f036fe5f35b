// main_pm.cpp -- a main.cu-shaped driver over the C++ façade: basic mPk
// initial conditions, then the particle-mesh cycle of main.cu:608-1317
// (Newtonian or weak-field GR with the elliptic vector method), called
// function by function through the gevolution-named interfaces.
//
//   main_pm -n N -L boxsize -zin z -zend z -steps S [-newton] [-seed s]
//           [-zrelax z] -pk mpk.txt [-out final.bin] [-gadget snap] [-express restart]
//           [-lightcone prefix -lcvertex x,y,z -lcz z] [-spectra prefix [-numbins n]]
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
// delta_N form (output.hpp:3017-3021, writePowerSpectrum tools.hpp:370-452).
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
    std::string pk, out, gadget, express, lightcone, spectra;
    int numbins = 32;
    double lcvertex[3] = {0.5, 0.5, 0.5}, lcz = 0.0;
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
        else if (a == "-lcvertex") {
            if (std::sscanf(next(), "%lf,%lf,%lf", &o.lcvertex[0], &o.lcvertex[1], &o.lcvertex[2]) != 3)
                throw std::runtime_error("-lcvertex wants x,y,z");
        } else if (a == "-lcz") o.lcz = std::atof(next());
        else throw std::runtime_error("unknown option " + a);
    }
    if (o.pk.empty()) throw std::runtime_error("-pk <mPk file> is required");
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

}  // namespace

int main(int argc, char **argv)
{
    try {
        const Options opt = parse(argc, argv);
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
        Field<Real> phi(lat), chi(lat), Bi(lat, 3), source(lat), Sij(lat, 6);
        Field<Cplx> scalarFT(lat), SijFT(lat, 6), BiFT(lat, 3);
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
        Field<Real> *update_fields[3] = {&phi, &chi, &Bi};
        double f_params[2];

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

        for (int cycle = 0; cycle < opt.steps; cycle++) {
            if (!opt.newton) {
                // T00 (main.cu:615-651) with the old phi
                projection_init(&source);
                projection_T00_project(&pcls, &source, a, &phi);
                const double H = Hconf(a, fourpiG, cosmo);
                if (dtau_old > 0.) {
                    double rsum = 0;
                    prepareFTsource(phi, chi, source, cosmo.Omega_m, source, 3. * H * dx * dx / dtau_old,
                                    fourpiG * dx * dx / a, 3. * H * H * dx * dx, &rsum);  // main.cu:725
                    plan_source.execute(FFT_FORWARD);                                    // main.cu:742
                    if (kPrecision == GEV_MIXED) set_k0(scalarFT, rsum);
                }
                // Tij (main.cu:680-758) with the old phi
                projection_init(&Sij);
                projection_Tij_project(&pcls, &Sij, a, &phi);
                if (dtau_old > 0.) {
                    solveModifiedPoissonFT(scalarFT, scalarFT, 1. / (dx * dx), 3. * H / dtau_old);  // main.cu:769
                    plan_phi.execute(FFT_BACKWARD);                                               // main.cu:794
                }
                prepareFTsource(phi, Sij, Sij, 2. * fourpiG * dx * dx / a);  // main.cu:822
                plan_Sij.execute(FFT_FORWARD);                               // main.cu:829
                projection_init(&Bi);
                projection_T0i_project(&pcls, &Bi, &phi);  // main.cu:858
                projectFTscalar(SijFT, scalarFT);          // main.cu:873
                plan_chi.execute(FFT_BACKWARD);            // main.cu:895
                plan_Bi.execute(FFT_FORWARD);              // main.cu:918
                projectFTvector(BiFT, BiFT, fourpiG * dx * dx);  // main.cu:925
                plan_Bi.execute(FFT_BACKWARD);                   // main.cu:948
                write_lightcone(cycle);  // main.cu:1013-1131 outputs, before the particle update
                // kick (main.cu:1227-1240), background half step, drift (:1253-1266)
                f_params[0] = a;
                f_params[1] = a * a * N;
                maxvel = pcls.updateVel(update_q_functor(), (dtau + dtau_old) / 2., update_fields,
                                        (1. / a < opt.zrelax + 1. ? 3 : 2), f_params);
                rungekutta4bg(a, fourpiG, cosmo, 0.5 * dtau);
                f_params[0] = a;
                f_params[1] = a * a * N;
                pcls.moveParticles(update_pos_functor(), dtau, update_fields, (1. / a < opt.zrelax + 1. ? 3 : 0),
                                   f_params);
                rungekutta4bg(a, fourpiG, cosmo, 0.5 * dtau);
                maxvel /= std::sqrt(maxvel * maxvel + 1.0);  // main.cu:1279-1283
            } else {
                projection_init(&source);
                scalarProjectionCIC_project(&pcls, &source);           // main.cu:655
                plan_source.execute(FFT_FORWARD);                      // main.cu:742
                solveModifiedPoissonFT(scalarFT, scalarFT, fourpiG / a);  // main.cu:765
                plan_phi.execute(FFT_BACKWARD);                           // main.cu:794
                write_lightcone(cycle);
                f_params[0] = a;
                maxvel = pcls.updateVel(update_q_Newton_functor(), (dtau + dtau_old) / 2., update_fields, 1, f_params);
                rungekutta4bg(a, fourpiG, cosmo, 0.5 * dtau);
                f_params[0] = a;
                pcls.moveParticles(update_pos_Newton_functor(), dtau, nullptr, 0, f_params);
                rungekutta4bg(a, fourpiG, cosmo, 0.5 * dtau);
            }
            tau += dtau;
            dtau_old = dtau;
            std::printf("cycle %d z=%.6f a=%.8f tau=%.8f maxvel=%.6e\n", cycle, 1. / a - 1., a, tau, maxvel);
        }
        lat.sync();

        if (!opt.spectra.empty()) {
            // writeSpectra's particle branch (output.hpp:2997-3022) at the final time:
            // the deconvolved CIC density spectrum as delta_N
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
                               (opt.spectra + "_deltaN.dat").c_str(), "power spectrum of delta_N", a);
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
            std::ofstream f(opt.out, std::ios::binary);
            f.write((const char *)&n, sizeof(n));
            f.write((const char *)id.data(), sizeof(int64_t) * n);
            f.write((const char *)pos.data(), sizeof(double) * 3 * n);
            f.write((const char *)q.data(), sizeof(double) * 3 * n);
            if (!f) throw std::runtime_error("writing " + opt.out + " failed");
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

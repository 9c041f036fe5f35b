// gevolution_pm.hpp -- C++ host façade over the C ABI (include/gev_pm.h) with
// the names, argument meaning and error behaviour of gevolution-2.0's
// particle-mesh interfaces, so a main.cu-shaped driver compiles against it.
//
//   Lattice                       LATfield2 Lattice + the GPU context (main.cu:300-357)
//   Field<Real> / Field<Cplx>     LATfield2 Field in real / Fourier layout
//   PlanFFT<Cplx>                 LATfield2 PlanFFT (execute FFT_FORWARD / FFT_BACKWARD)
//   perfParticles_gevolution<>    Particles_gevolution.hpp:176-211 (+ perfParticles base:
//                                 initialize, addParticle_global, updateRowBuffers,
//                                 projectParticles[_Async], meshprojection_project,
//                                 updateVel, moveParticles, num_particles, sampleParticles,
//                                 saveGadget2, loadGadget2, saveExpress, loadExpress)
//   projection_* (+ _Async, _comm) / prepareFTsource / solveModifiedPoissonFT /
//   projectFTscalar / projectFTvector / evolveFTvector
//                                 gevolution.hpp free functions, same defaults
//   loadHomogeneousTemplate / initializeParticlePositions
//                                 ic_basic.hpp:205-330, 1612-1644
//   DeviceStagingBuffer<T>        cuda_staging.hpp:18-63 (parameter arrays)
//   Hconf / rungekutta4bg / particleHorizon
//                                 background.hpp:200-290
//   extractCrossSpectrum / extractPowerSpectrum / writePowerSpectrum
//                                 tools.hpp:65-342, 370-452
//
// Errors: any non-zero status of the library becomes std::runtime_error with
// the library's message (the reference throws std::runtime_error on GPU
// errors, cuda_staging.hpp:12-16).  Precision: compile with -DSINGLE for the
// mixed form (Real = float, as the reference's -DSINGLE build), otherwise FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "gev_pm.h"

namespace gevpm {

#ifdef SINGLE
using Real = float;
constexpr int kPrecision = GEV_MIXED;
#else
using Real = double;
constexpr int kPrecision = GEV_FP64;
#endif
using Cplx = std::complex<Real>;  // (re, im) interleaved, the library's Fourier layout

inline void check(int rc, const char *what)
{
    if (rc != GEV_OK) throw std::runtime_error(std::string(what) + ": " + gev_last_error());
}

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

enum { FFT_FORWARD = 0, FFT_BACKWARD = 1 };

// ------------------------------------------------------------------ lattice
class Lattice {
  public:
    explicit Lattice(int N, int precision = kPrecision, hipStream_t stream = nullptr) : N_(N), prec_(precision)
    {
        check(gev_ctx_create(N, precision, (void *)stream, &ctx_), "gev_ctx_create");
        P_ = gev_ctx_kpitch(ctx_);
    }
    ~Lattice() { gev_ctx_destroy(ctx_); }
    Lattice(const Lattice &) = delete;
    Lattice &operator=(const Lattice &) = delete;

    int size() const { return N_; }
    int precision() const { return prec_; }
    int kpitch() const { return P_; }
    int64_t sites() const { return (int64_t)N_ * N_ * N_; }
    int64_t ksites() const { return (int64_t)N_ * N_ * P_; }
    gev_ctx *ctx() const { return ctx_; }
    void sync() const { check(gev_ctx_sync(ctx_), "gev_ctx_sync"); }

  private:
    int N_, prec_, P_ = 0;
    gev_ctx *ctx_ = nullptr;
};

// -------------------------------------------------------------------- field
// Field<Real>: components x N^3 reals; Field<Cplx>: components x N x N x P complex.
// Device memory owned by the field (RAII).
// symmetric: the six-component layout (11, 12, 13, 22, 23, 33) of LATfield2's
// symmetric rank-2 fields (initialize(lat, 3, 3, symmetric)); read by the
// spectra, which contract such fields over all nine components
enum { unsymmetric = 0, symmetric = 1 };
template <typename T> class Field {
  public:
    Field(Lattice &lat, int components = 1, int symmetry = unsymmetric)
        : lat_(&lat), ncomp_(components), sym_(symmetry)
    {
        if (symmetry == symmetric && components != 6)
            throw std::runtime_error("Field: a symmetric field has 6 components");
        n_ = (int64_t)components * (is_complex() ? lat.ksites() : lat.sites());
        hip_check(hipMalloc(&d_, sizeof(T) * (size_t)n_), "hipMalloc(Field)");
        hip_check(hipMemset(d_, 0, sizeof(T) * (size_t)n_), "hipMemset(Field)");
    }
    ~Field() { (void)hipFree(d_); }
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    static constexpr bool is_complex() { return !std::is_same<T, Real>::value; }
    T *data() { return d_; }
    const T *data() const { return d_; }
    int components() const { return ncomp_; }
    int symmetry() const { return sym_; }
    int64_t count() const { return n_; }
    Lattice &lattice() const { return *lat_; }

    // Field::updateHalo (main.cu:802,902,955): the fields here have no ghost
    // layers (the kernels index the periodic lattice directly), so there is
    // nothing to exchange; kept so that main.cu's call sequence compiles as is
    void updateHalo() {}

    void to_host(std::vector<T> &h) const
    {
        lat_->sync();
        h.resize((size_t)n_);
        hip_check(hipMemcpy(h.data(), d_, sizeof(T) * (size_t)n_, hipMemcpyDeviceToHost), "Field::to_host");
    }
    void from_host(const std::vector<T> &h)
    {
        if ((int64_t)h.size() != n_) throw std::runtime_error("Field::from_host: size mismatch");
        lat_->sync();
        hip_check(hipMemcpy(d_, h.data(), sizeof(T) * (size_t)n_, hipMemcpyHostToDevice), "Field::from_host");
    }

  private:
    Lattice *lat_;
    int ncomp_, sym_;
    int64_t n_ = 0;
    T *d_ = nullptr;
};

// ------------------------------------------------------------------ PlanFFT
template <typename T> class PlanFFT {
  public:
    PlanFFT(Field<Real> *rfield, Field<T> *kfield, int components = 1) : r_(rfield), k_(kfield), ncomp_(components)
    {
        static_assert(!std::is_same<T, Real>::value, "PlanFFT<Cplx> pairs a real and a Fourier field");
    }
    // FFT_BACKWARD overwrites the Fourier field (the library's c2r works in place)
    void execute(int dir)
    {
        gev_ctx *c = r_->lattice().ctx();
        if (dir == FFT_FORWARD) check(gev_fft_r2c(c, r_->data(), k_->data(), ncomp_), "PlanFFT::execute(FFT_FORWARD)");
        else check(gev_fft_c2r(c, k_->data(), r_->data(), ncomp_), "PlanFFT::execute(FFT_BACKWARD)");
    }

  private:
    Field<Real> *r_;
    Field<T> *k_;
    int ncomp_;
};

// ---------------------------------------------------- DeviceStagingBuffer
// cuda_staging.hpp:18-63: the reference copies small parameter arrays to the
// device before each functor launch.  The C ABI takes FP64 scalars by value,
// so the array stays in host memory here and data() returns it; the
// interface (count, copy_from_host, data) is the reference's, so main.cu's
// d_f_params.copy_from_host(f_params, 7) / d_f_params.data() compile as they are.
template <typename T> class DeviceStagingBuffer {
  public:
    DeviceStagingBuffer() = default;
    explicit DeviceStagingBuffer(size_t count) : v_(count) {}
    DeviceStagingBuffer(const T *host_data, size_t count) : v_(count) { copy_from_host(host_data, count); }
    T *data() { return v_.data(); }
    const T *data() const { return v_.data(); }
    size_t size() const { return v_.size(); }
    void copy_from_host(const T *host_data, size_t count)
    {
        if (count > v_.size()) throw std::runtime_error("DeviceStagingBuffer::copy_from_host: count exceeds capacity");
        std::copy(host_data, host_data + count, v_.begin());
    }

  private:
    std::vector<T> v_;
};

// ------------------------------------------------------------- cosmology
struct cosmology {
    double Omega_m = 0.31, Omega_Lambda = 0.69, Omega_rad = 0.0;
    double Omega_cdm = 0.31, h = 0.67;  // snapshot metadata only (output.hpp:423-457)
};
inline double Hconf(double a, double fourpiG, const cosmology &c)
{
    return gev_hconf(a, fourpiG, c.Omega_m, c.Omega_Lambda, c.Omega_rad);
}
inline void rungekutta4bg(double &a, double fourpiG, const cosmology &c, double dtau)
{
    a = gev_rk4bg(a, fourpiG, c.Omega_m, c.Omega_Lambda, c.Omega_rad, dtau);
}
inline double particleHorizon(double a, double fourpiG, const cosmology &c)
{
    return gev_particle_horizon(a, fourpiG, c.Omega_m, c.Omega_Lambda, c.Omega_rad);
}

// ---------------------------------------------------------------- particles
struct part_simple {
    int64_t ID;
    double pos[3];
    double vel[3];  // q/m
};
struct part_simple_info {
    double mass = 0.0;
    const char *type_name = "part_simple";
};

// metadata.hpp:159-178, byte for byte
using gadget2_header = gev_gadget2_header;
// metadata.hpp:205-212
using lightcone_geometry = gev_lightcone_geometry;
constexpr int MAX_INTERSECTS = GEV_MAX_INTERSECTS;

// LightconeIDBacklog (lightcone_id_backlog.hpp:16-125) together with the
// IDprelog vector writeLightcones keeps beside it (output.hpp:1017-1185):
// the lightcone saveGadget2 reads the backlog and appends to the prelog;
// finalize() makes the prelog, sorted and unique, the next backlog (the
// hand-over at the end of writeLightcones, output.hpp:2166-2880, one rank).
class LightconeIDBacklog {
  public:
    explicit LightconeIDBacklog(Lattice *lat) { check(gev_idlog_create(lat->ctx(), &h_), "LightconeIDBacklog"); }
    ~LightconeIDBacklog() { gev_idlog_destroy(h_); }
    LightconeIDBacklog(const LightconeIDBacklog &) = delete;
    LightconeIDBacklog &operator=(const LightconeIDBacklog &) = delete;

    size_t size() const { return (size_t)sizes().first; }
    bool empty() const { return size() == 0; }
    size_t prelog_size() const { return (size_t)sizes().second; }
    std::vector<int64_t> ids() const { return get(0); }
    std::vector<int64_t> prelog() const { return get(1); }
    bool contains(int64_t id) const
    {
        const std::vector<int64_t> v = ids();
        return std::binary_search(v.begin(), v.end(), id);
    }
    // clear + append + finalize of the reference class
    void assign(const std::vector<int64_t> &v) { check(gev_idlog_set_backlog(h_, v.data(), (int64_t)v.size()), "LightconeIDBacklog::assign"); }
    void clear() { assign({}); }
    void finalize() { check(gev_idlog_commit(h_), "LightconeIDBacklog::finalize"); }
    gev_idlog *handle() const { return h_; }

  private:
    std::pair<int64_t, int64_t> sizes() const
    {
        int64_t a = 0, b = 0;
        check(gev_idlog_sizes(h_, &a, &b), "LightconeIDBacklog::size");
        return {a, b};
    }
    std::vector<int64_t> get(int which) const
    {
        const auto s = sizes();
        std::vector<int64_t> v((size_t)(which == 0 ? s.first : s.second));
        check(gev_idlog_get(h_, which, v.data()), "LightconeIDBacklog::ids");
        return v;
    }
    gev_idlog *h_ = nullptr;
};

// findIntersectingLightcones (tools.hpp:619-862); returns the number found,
// of which the first MAX_INTERSECTS are stored
inline int findIntersectingLightcones(lightcone_geometry &lightcone, double outer, double inner, double *domain,
                                      double vertex[MAX_INTERSECTS][3])
{
    int n = 0;
    check(gev_lightcone_vertices(&lightcone, outer, inner, domain, &vertex[0][0], MAX_INTERSECTS, &n),
          "findIntersectingLightcones");
    return n;
}

// functor tags: which gevolution functor updateVel / moveParticles /
// projectParticles apply
struct update_q_functor {};         // gevolution.hpp:888-1011
struct update_q_Newton_functor {};  // gevolution.hpp:1051-1118
struct update_pos_functor {};       // gevolution.hpp:1162-1232
struct update_pos_Newton_functor {};  // gevolution.hpp:1270-1273
struct particle_T00_project_functor {};  // gevolution.hpp:1729-1854, params {a, coeff}
struct particle_T0i_project_functor {};  // gevolution.hpp:2031-2201, params {coeff}
struct particle_Tij_project_functor {};  // gevolution.hpp:2582-2965, params {a, coeff, hij_hom[5]}
struct displace_pcls_ic_basic_functor {};  // ic_basic.hpp:65-109, dtau = coeff
struct initialize_q_ic_basic_functor {};   // ic_basic.hpp:133-178, dtau = coeff
enum { SUM = 1, MIN = 2, MAX = 4 };  // reduction ops of the outputs tail (LATfield2's)

template <class part = part_simple, class part_info = part_simple_info> class perfParticles_gevolution {
  public:
    perfParticles_gevolution() = default;
    ~perfParticles_gevolution()
    {
        if (p_) gev_pcls_destroy(p_);
    }
    perfParticles_gevolution(const perfParticles_gevolution &) = delete;
    perfParticles_gevolution &operator=(const perfParticles_gevolution &) = delete;

    // perfParticles::initialize (ic_basic.hpp:2206, unit_tests.cu:541).
    // gevolution's particle coordinates are in box units, so boxSize must be
    // the unit box; capacity already includes the `extra` slack the reference
    // keeps for its row buffers (the container here re-bins in place).
    void initialize(const part_info &info, Lattice *lat, Real boxSize[3], uint64_t capacity, uint64_t extra)
    {
        for (int d = 0; d < 3; d++)
            if (boxSize[d] != (Real)1)
                throw std::runtime_error("perfParticles::initialize: particle coordinates are in box units (boxSize = 1)");
        if (extra > capacity) throw std::runtime_error("perfParticles::initialize: extra exceeds capacity");
        initialize(info, lat, capacity);
    }
    void initialize(const part_info &info, Lattice *lat, uint64_t capacity)
    {
        if (p_) throw std::runtime_error("perfParticles::initialize: already initialised");
        lat_ = lat;
        info_ = info;
        check(gev_pcls_create(lat->ctx(), (int64_t)capacity, info.mass, &p_), "perfParticles::initialize");
    }
    // addParticle_global (ic_basic.hpp:1639): the particle is queued on the
    // host; updateRowBuffers() (ic_basic.hpp:2211) hands the queue to the
    // device container and re-bins it.  A particle outside [0,1) is wrapped
    // into the box.
    void addParticle_global(const part &p) { pending_.push_back(p); }
    void addParticles_global(const std::vector<part> &ps) { pending_.insert(pending_.end(), ps.begin(), ps.end()); }
    void updateRowBuffers()
    {
        if (!flush()) check(gev_pcls_sort(p_), "updateRowBuffers");
    }

    int64_t num_particles()
    {
        flush();
        gev_pcls_view v;
        check(gev_pcls_view_get(p_, &v), "num_particles");
        return v.n;
    }
    // sampleParticles (unit_tests.cu:537): the first n particles in container
    // order (at most num_particles()), three values per particle in the order
    // of part_simple's pos[3] / vel[3]; returns how many were written
    int64_t sampleParticles(Real *pos, Real *vel, long *id, uint64_t n)
    {
        std::vector<part> ps;
        sampleParticles(ps);
        const int64_t m = std::min<int64_t>((int64_t)n, (int64_t)ps.size());
        for (int64_t i = 0; i < m; i++) {
            for (int d = 0; d < 3; d++) {
                pos[3 * i + d] = (Real)ps[i].pos[d];
                vel[3 * i + d] = (Real)ps[i].vel[d];
            }
            id[i] = (long)ps[i].ID;
        }
        return m;
    }
    // every particle, FP64, host memory
    void sampleParticles(std::vector<part> &out)
    {
        const int64_t n = num_particles();
        std::vector<double> pos(3 * n), q(3 * n);
        std::vector<int64_t> id(n);
        check(gev_pcls_store(p_, pos.data(), q.data(), id.data(), 0), "sampleParticles");
        out.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) {
            for (int d = 0; d < 3; d++) {
                out[i].pos[d] = pos[d * n + i];
                out[i].vel[d] = q[d * n + i];
            }
            out[i].ID = id[i];
        }
    }

    // projectParticles / projectParticles_Async (gevolution.hpp:1898, 2244,
    // 3027): the functor's deposit is added to fields[0]; fields[1] is phi
    // when nfield == 2.  Both forms are ordered on the lattice's stream; the
    // _Async form returns without waiting, as the reference's does.
    void projectParticles(particle_T00_project_functor, Field<Real> **fields, int nfield, const double *params)
    {
        flush();
        check(gev_project_T00(p_, fields[0]->data(), params[0], nfield >= 2 ? fields[1]->data() : nullptr, params[1], 1),
              "projectParticles(T00)");
    }
    void projectParticles(particle_T0i_project_functor, Field<Real> **fields, int nfield, const double *params)
    {
        flush();
        check(gev_project_T0i(p_, fields[0]->data(), nfield >= 2 ? fields[1]->data() : nullptr, params[0], 1),
              "projectParticles(T0i)");
    }
    void projectParticles(particle_Tij_project_functor, Field<Real> **fields, int nfield, const double *params)
    {
        flush();
        check(gev_project_Tij(p_, fields[0]->data(), params[0], nfield >= 2 ? fields[1]->data() : nullptr, params[1], 1),
              "projectParticles(Tij)");
    }
    template <class F> void projectParticles_Async(F f, Field<Real> **fields, int nfield, const double *params)
    {
        projectParticles(f, fields, nfield, params);
    }
    // meshprojection_project (gevolution.hpp:1904): adds the CIC mass density
    void meshprojection_project(Field<Real> *field)
    {
        flush();
        check(gev_project_mass(p_, field->data(), 1.0, 1), "meshprojection_project");
    }

    // updateVel: applies the kick functor and returns sqrt of the maximum of
    // the functor values (v^2/a^2), as main.cu:1232 then converts it to
    // maxvel.  The kick and drift functors write no outputs, so the reduction
    // tail (outputs, ops, nout) leaves outputs as they are.
    Real updateVel(update_q_functor, double dtau, Field<Real> **fields, int nfield, const double *params,
                   double *outputs = nullptr, int *ops = nullptr, int nout = 0)
    {
        (void)outputs, (void)ops, (void)nout;
        flush();
        double m = 0;
        check(gev_update_q(p_, dtau, fields[0]->data(), nfield >= 2 ? fields[1]->data() : nullptr,
                           nfield >= 3 ? fields[2]->data() : nullptr, nfield, params[0], params[1], &m),
              "updateVel(update_q)");
        return (Real)std::sqrt(m);
    }
    Real updateVel(update_q_Newton_functor, double dtau, Field<Real> **fields, int nfield, const double *params,
                   double *outputs = nullptr, int *ops = nullptr, int nout = 0)
    {
        (void)outputs, (void)ops, (void)nout;
        flush();
        double m = 0;
        check(gev_update_q_newton(p_, dtau, fields[0]->data(), nfield >= 2 ? fields[1]->data() : nullptr, nfield,
                                  params[0], &m),
              "updateVel(update_q_Newton)");
        return (Real)std::sqrt(m);
    }
    // initialize_q_ic_basic (ic_basic.hpp:133-178): q = -coeff grad(fields[0]),
    // coeff = dtau; returns sqrt of the maximum q^2 (one potential only: the
    // two-potential hybrid baryon form is not built)
    Real updateVel(initialize_q_ic_basic_functor, double dtau, Field<Real> **fields, int nfield,
                   const double *params = nullptr, double *outputs = nullptr, int *ops = nullptr, int nout = 0)
    {
        (void)params, (void)outputs, (void)ops, (void)nout;
        if (nfield != 1) throw std::runtime_error("updateVel(initialize_q_ic_basic): one velocity potential only");
        flush();
        double m = 0;
        check(gev_initialize_q(p_, fields[0]->data(), dtau, &m), "updateVel(initialize_q_ic_basic)");
        return (Real)std::sqrt(m);
    }
    Real updateVel(initialize_q_ic_basic_functor f, double dtau, Field<Real> *field, int nfield)
    {
        Field<Real> *fl[1] = {field};
        return updateVel(f, dtau, fl, nfield);
    }
    void moveParticles(update_pos_functor, double dtau, Field<Real> **fields, int nfield, const double *params,
                       double *outputs = nullptr, int *ops = nullptr, int nout = 0)
    {
        (void)outputs, (void)ops, (void)nout;
        flush();
        check(gev_update_pos(p_, dtau, nfield >= 1 ? fields[0]->data() : nullptr,
                             nfield >= 2 ? fields[1]->data() : nullptr, nfield >= 3 ? fields[2]->data() : nullptr,
                             nfield, params[0], params[1]),
              "moveParticles(update_pos)");
    }
    void moveParticles(update_pos_Newton_functor, double dtau, Field<Real> **, int, const double *params,
                       double *outputs = nullptr, int *ops = nullptr, int nout = 0)
    {
        (void)outputs, (void)ops, (void)nout;
        flush();
        check(gev_update_pos_newton(p_, dtau, params[0]), "moveParticles(update_pos_Newton)");
    }
    // displace_pcls_ic_basic (ic_basic.hpp:65-109): pos += coeff grad(fields[0])
    // with coeff = dtau; outputs[0] receives the maximum displacement when
    // nout > 0 (the reference reduces it with op MAX, ic_basic.hpp:2214-2219)
    void moveParticles(displace_pcls_ic_basic_functor, double dtau, Field<Real> **fields, int nfield,
                       const double *params = nullptr, double *outputs = nullptr, int *ops = nullptr, int nout = 0)
    {
        (void)params;
        if (nfield != 1) throw std::runtime_error("moveParticles(displace_pcls_ic_basic): one displacement field only");
        if (nout > 0 && ops && ops[0] != MAX) throw std::runtime_error("moveParticles(displace_pcls_ic_basic): MAX only");
        flush();
        double m = 0;
        check(gev_displace_particles(p_, fields[0]->data(), dtau, &m), "moveParticles(displace_pcls_ic_basic)");
        if (nout > 0 && outputs) outputs[0] = m;
    }

    // Particle files (Particles_gevolution.hpp:170-186): Gadget-2 snapshot,
    // single file, particle type 1, optional EXACT_OUTPUT_REDSHIFTS drift and
    // kick; express restart file "<filename>.0" with this lattice's layout
    void saveGadget2(std::string filename, gadget2_header &hdr, const int tracer_factor = 1, double dtau_pos = 0.,
                     double dtau_vel = 0., Field<Real> *phi = nullptr)
    {
        check(gev_pcls_save_gadget2(p_, filename.c_str(), &hdr, tracer_factor, dtau_pos, dtau_vel,
                                    phi ? phi->data() : nullptr),
              "saveGadget2");
    }
    void loadGadget2(std::string filename, gadget2_header &hdr)
    {
        check(gev_pcls_load_gadget2(p_, filename.c_str(), &hdr), "loadGadget2");
    }
    // saveGadget2, lightcone form (Particles_gevolution.hpp:172,1566-2317); the
    // prelog lives in IDbacklog (its prelog()) instead of a separate vector
    void saveGadget2(std::string filename, gadget2_header &hdr, lightcone_geometry &lightcone, double dist,
                     double dtau, double dtau_old, double dadtau, double vertex[][3], const int vertexcount,
                     LightconeIDBacklog &IDbacklog, Field<Real> *phi, const int tracer_factor = 1)
    {
        check(gev_pcls_save_gadget2_lightcone(p_, filename.c_str(), &hdr, &lightcone, dist, dtau, dtau_old, dadtau,
                                              vertexcount > 0 ? &vertex[0][0] : nullptr, vertexcount,
                                              IDbacklog.handle(), phi ? phi->data() : nullptr, tracer_factor,
                                              nullptr),
              "saveGadget2(lightcone)");
    }
    void saveExpress(std::string filename, gadget2_header &hdr)
    {
        check(gev_pcls_save_express(p_, filename.c_str(), &hdr), "saveExpress");
    }
    void loadExpress(std::string filename, gadget2_header &hdr)
    {
        check(gev_pcls_load_express(p_, filename.c_str(), &hdr), "loadExpress");
    }

    gev_pcls *handle()
    {
        flush();
        return p_;
    }
    Lattice *lattice() const { return lat_; }
    const part_info &parts_info() const { return info_; }

  private:
    // hands the queued addParticle_global particles to the device container
    // (after the particles it already holds) and re-bins; true if it did
    bool flush()
    {
        if (!p_) throw std::runtime_error("perfParticles: not initialised");
        if (pending_.empty()) return false;
        gev_pcls_view v;
        check(gev_pcls_view_get(p_, &v), "addParticle_global");
        const int64_t n0 = v.n, n = n0 + (int64_t)pending_.size();
        std::vector<double> pos(3 * n), q(3 * n), op(3 * n0), oq(3 * n0);
        std::vector<int64_t> id(n), oid(n0);
        if (n0 > 0) check(gev_pcls_store(p_, op.data(), oq.data(), oid.data(), 0), "addParticle_global");
        for (int64_t i = 0; i < n; i++) {
            for (int d = 0; d < 3; d++) {
                double x, u;
                if (i < n0) {
                    x = op[d * n0 + i];
                    u = oq[d * n0 + i];
                } else {
                    x = pending_[i - n0].pos[d];
                    u = pending_[i - n0].vel[d];
                    x -= std::floor(x);
                    if (x >= 1.0) x = 0.0;
                }
                pos[d * n + i] = x;
                q[d * n + i] = u;
            }
            id[i] = i < n0 ? oid[i] : pending_[i - n0].ID;
        }
        pending_.clear();
        check(gev_pcls_load(p_, n, pos.data(), q.data(), id.data(), 0), "addParticle_global");
        return true;
    }

    gev_pcls *p_ = nullptr;
    Lattice *lat_ = nullptr;
    part_info info_{};
    std::vector<part> pending_;
};

using Particles = perfParticles_gevolution<part_simple, part_simple_info>;

// loadHomogeneousTemplate (ic_basic.hpp:205-330): the type-1 particle
// positions of a single-file Gadget-2 template (sc0/sc1/bcc/fcc_crystal.dat),
// divided by the header's BoxSize so they lie in [0,1].  partdata is
// allocated with malloc, as the reference does, and freed by the caller.
// Every failure the reference aborts on throws std::runtime_error here.
inline void loadHomogeneousTemplate(const char *filename, long &numpart, float *&partdata)
{
    struct File {
        FILE *f;
        ~File() { if (f) std::fclose(f); }
    } tf{std::fopen(filename, "rb")};
    auto fail = [&](const std::string &why) -> void {
        throw std::runtime_error(std::string("loadHomogeneousTemplate: ") + why + " (" + filename + ")");
    };
    if (!tf.f) fail("unable to open template file");
    int32_t b1 = 0, b2 = 0;
    gadget2_header hdr;
    if (std::fread(&b1, 4, 1, tf.f) != 1 || b1 != (int32_t)sizeof(hdr)) fail("header not recognized");
    if (std::fread(&hdr, sizeof(hdr), 1, tf.f) != 1) fail("unable to read header");
    if (std::fread(&b2, 4, 1, tf.f) != 1 || b1 != b2) fail("block size mismatch while reading header");
    if (hdr.num_files != 1) fail("multiple input files are not supported");
    if (!(hdr.BoxSize > 0.)) fail("BoxSize must be positive");
    if (hdr.npart[1] == 0) fail("no particles declared");
    if (std::fread(&b1, 4, 1, tf.f) != 1) fail("unable to read particle data");
    if (hdr.npart[0] > 0 && std::fseek(tf.f, 3L * 4 * (long)hdr.npart[0], SEEK_CUR)) fail("unable to skip gas particles");
    std::vector<float> buf(3 * (size_t)hdr.npart[1]);
    if (std::fread(buf.data(), 4, buf.size(), tf.f) != buf.size()) fail("unable to read particle data");
    for (int i = 2; i < 6; i++)
        if (hdr.npart[i] > 0 && std::fseek(tf.f, 3L * 4 * (long)hdr.npart[i], SEEK_CUR)) fail("unable to skip particles");
    if (std::fread(&b2, 4, 1, tf.f) != 1 || b1 != b2) fail("block size mismatch while reading particles");
    for (float &x : buf) {
        x = (float)(x / hdr.BoxSize);
        if (x < 0.f || x > 1.f) fail("particle data corrupted");
    }
    partdata = (float *)std::malloc(sizeof(float) * buf.size());
    if (!partdata) fail("memory error");
    std::copy(buf.begin(), buf.end(), partdata);
    numpart = (long)hdr.npart[1];
}

// initializeParticlePositions (ic_basic.hpp:1612-1644) on one rank: the
// template tiled numtile times along each axis, particles at rest, ID =
// i + numpart (xtile + numtile (ytile + numtile ztile)), added with
// addParticle_global (call updateRowBuffers() afterwards, as the reference).
template <typename part_handler>
void initializeParticlePositions(const long numpart, const float *partdata, const int numtile, part_handler &pcls)
{
    part_simple part;
    part.vel[0] = part.vel[1] = part.vel[2] = 0.0;
    for (int ztile = 0; ztile < numtile; ztile++)
        for (int ytile = 0; ytile < numtile; ytile++)
            for (int xtile = 0; xtile < numtile; xtile++)
                for (long i = 0; i < numpart; i++) {
                    part.pos[0] = ((Real)xtile + partdata[3 * i]) / (Real)numtile;
                    part.pos[1] = ((Real)ytile + partdata[3 * i + 1]) / (Real)numtile;
                    part.pos[2] = ((Real)ztile + partdata[3 * i + 2]) / (Real)numtile;
                    part.ID = i + numpart * ((long)xtile + (long)numtile * ((long)ytile + (long)numtile * (long)ztile));
                    pcls.addParticle_global(part);
                }
}

// ------------------------------------------------------------ projections
// projection_init: zero a field before additive projections
inline void projection_init(Field<Real> *f)
{
    check(gev_field_fill(f->lattice().ctx(), f->data(), f->count(), 0.0), "projection_init");
}
// gevolution.hpp:1887 (functor :1729-1854); phi == nullptr: plain mass density
inline void projection_T00_project(Particles *pcls, Field<Real> *T00, double a = 1., Field<Real> *phi = nullptr,
                                   double coeff = 1.)
{
    check(gev_project_T00(pcls->handle(), T00->data(), a, phi ? phi->data() : nullptr, coeff, 1),
          "projection_T00_project");
}
// gevolution.hpp:1902
inline void scalarProjectionCIC_project(Particles *pcls, Field<Real> *rho) { pcls->meshprojection_project(rho); }
// gevolution.hpp:2232 (functor :2031-2201)
inline void projection_T0i_project(Particles *pcls, Field<Real> *T0i, Field<Real> *phi = nullptr, double coeff = 1.)
{
    check(gev_project_T0i(pcls->handle(), T0i->data(), phi ? phi->data() : nullptr, coeff, 1),
          "projection_T0i_project");
}
// gevolution.hpp:2999 (functor :2582-2965, no CIC_PROJECT_TIJ)
inline void projection_Tij_project(Particles *pcls, Field<Real> *Tij, double a = 1., Field<Real> *phi = nullptr,
                                   double coeff = 1.)
{
    check(gev_project_Tij(pcls->handle(), Tij->data(), a, phi ? phi->data() : nullptr, coeff, 1),
          "projection_Tij_project");
}

// projection_T0i_project_Async / projection_Tij_project_Async
// (gevolution.hpp:2242, 3030; main.cu:756, 858): fields {T0i or Tij, phi},
// nfield 1 or 2, params {coeff} (T0i) or {a, coeff, ...} (Tij); the deposit is
// added to fields[0], ordered on the lattice's stream, and the call returns
// without waiting.  main.cu's cudaDeviceSynchronize() after them is a plain
// hipDeviceSynchronize() here.
inline void projection_T0i_project_Async(Particles *pcls, Field<Real> **fields, int nfield, double *params)
{
    pcls->projectParticles_Async(particle_T0i_project_functor(), fields, nfield, params);
}
inline void projection_Tij_project_Async(Particles *pcls, Field<Real> **fields, int nfield, double *params)
{
    pcls->projectParticles_Async(particle_Tij_project_functor(), fields, nfield, params);
}
// projection_comm1 and its aliases projection_T00_comm, projection_T0i_comm,
// projection_Tij_comm (gevolution.hpp:1543-1620, 1726, 2028, 2573-2579) fold
// the upper ghost layers of a deposit into their owners.  The deposits here
// are gather-form tiles over a periodic lattice without ghost layers, so every
// site already holds its whole sum and the folds have nothing to do.
inline void projection_comm1(Field<Real> *) {}
inline void projection_T00_comm(Field<Real> *f) { projection_comm1(f); }
inline void projection_T0i_comm(Field<Real> *f) { projection_comm1(f); }
inline void projection_Tij_comm(Field<Real> *f) { projection_comm1(f); }

// ----------------------------------------------------------------- sources
// gevolution.hpp:183 (T00 form); returns the sum of source, as the reference
inline double prepareFTsource(Field<Real> &phi, Field<Real> &chi, Field<Real> &source, const double bgmodel,
                              Field<Real> &result, const double coeff, const double coeff2, const double coeff3,
                              double *result_sum = nullptr)
{
    double s = 0, r = 0;
    check(gev_prepare_source_T00(phi.lattice().ctx(), phi.data(), chi.data(), source.data(), bgmodel, result.data(),
                                 coeff, coeff2, coeff3, &s, &r),
          "prepareFTsource(T00)");
    if (result_sum) *result_sum = r;
    return s;
}
// gevolution.hpp:119 (Tij form)
inline void prepareFTsource(Field<Real> &phi, Field<Real> &Tij, Field<Real> &Sij, const double coeff)
{
    check(gev_prepare_source_Tij(phi.lattice().ctx(), phi.data(), Tij.data(), Sij.data(), coeff),
          "prepareFTsource(Tij)");
}

// ----------------------------------------------------------------- k-space
inline void solveModifiedPoissonFT(Field<Cplx> &sourceFT, Field<Cplx> &potFT, Real coeff, const Real modif = 0)
{
    check(gev_poisson(sourceFT.lattice().ctx(), sourceFT.data(), potFT.data(), coeff, modif),
          "solveModifiedPoissonFT");
}
inline void projectFTscalar(Field<Cplx> &SijFT, Field<Cplx> &chiFT, const int add = 0)
{
    check(gev_project_ft_scalar(SijFT.lattice().ctx(), SijFT.data(), chiFT.data(), add), "projectFTscalar");
}
inline void projectFTvector(Field<Cplx> &SiFT, Field<Cplx> &BiFT, const Real coeff = 1, const Real modif = 0)
{
    check(gev_project_ft_vector(SiFT.lattice().ctx(), SiFT.data(), BiFT.data(), coeff, modif), "projectFTvector");
}
inline void evolveFTvector(Field<Cplx> &SijFT, Field<Cplx> &BiFT, const Real a2dtau)
{
    check(gev_evolve_ft_vector(SijFT.lattice().ctx(), SijFT.data(), BiFT.data(), a2dtau), "evolveFTvector");
}
// extractCrossSpectrum (tools.hpp:65-337): the binned cross spectrum
// Re(F1 conj F2) of two Fourier fields of the same lattice.  As in the
// reference, the layout is fld1FT's: comp1 < fld1FT.components() and
// comp2 < fld2FT.components() (both >= 0) pick one component of each;
// otherwise a symmetric fld1FT is contracted with fld2FT over all nine
// components (off-diagonal pairs twice), and any other fld1FT is summed over
// its components against the same components of fld2FT (e.g. a scalar field
// against the first component of a tensor field, output.hpp:3088).  ktype
// KTYPE_GRID bins by the lattice momentum 2 N sin(pi n / N), KTYPE_LINEAR by
// 2 pi n.
constexpr int KTYPE_GRID = 0, KTYPE_LINEAR = 1;  // tools.hpp:27-28
inline void extractCrossSpectrum(Field<Cplx> &fld1FT, Field<Cplx> &fld2FT, Real *kbin, Real *power, Real *kscatter,
                                 Real *pscatter, int *occupation, const int numbins, const bool deconvolve = true,
                                 const int ktype = KTYPE_LINEAR, const int comp1 = -1, const int comp2 = -1)
{
    Lattice &lat = fld1FT.lattice();
    if (&lat != &fld2FT.lattice()) throw std::runtime_error("extractCrossSpectrum: the fields live on different lattices");
    const Cplx *F1 = fld1FT.data(), *F2 = fld2FT.data();
    int ncomp = fld1FT.components(), sym = 0;
    if (comp1 >= 0 && comp2 >= 0 && comp1 < fld1FT.components() && comp2 < fld2FT.components()) {
        F1 += (size_t)comp1 * (size_t)lat.ksites();  // one component of each
        F2 += (size_t)comp2 * (size_t)lat.ksites();
        ncomp = 1;
    } else if (fld1FT.symmetry() == symmetric) {
        if (fld2FT.symmetry() != symmetric)
            throw std::runtime_error("extractCrossSpectrum: a symmetric field needs a symmetric partner");
        sym = 1;
    } else if (fld2FT.components() < ncomp) {
        throw std::runtime_error("extractCrossSpectrum: the second field has fewer components than the first");
    }
    std::vector<double> kb(numbins), pw(numbins), ks(numbins), ps(numbins);
    std::vector<int64_t> oc(numbins);
    check(gev_cross_spectrum(lat.ctx(), F1, F2, ncomp, sym, -1, -1, numbins, deconvolve ? 1 : 0, ktype, kb.data(),
                             pw.data(), ks.data(), ps.data(), oc.data()),
          "extractCrossSpectrum");
    for (int i = 0; i < numbins; i++) {
        kbin[i] = (Real)kb[i];
        power[i] = (Real)pw[i];
        kscatter[i] = (Real)ks[i];
        pscatter[i] = (Real)ps[i];
        occupation[i] = (int)oc[i];
    }
}
// extractPowerSpectrum (tools.hpp:339-342): the cross spectrum of a field with itself
inline void extractPowerSpectrum(Field<Cplx> &fldFT, Real *kbin, Real *power, Real *kscatter, Real *pscatter,
                                 int *occupation, const int numbins, const bool deconvolve = true,
                                 const int ktype = KTYPE_LINEAR)
{
    extractCrossSpectrum(fldFT, fldFT, kbin, power, kscatter, pscatter, occupation, numbins, deconvolve, ktype);
}

// writePowerSpectrum (tools.hpp:370-452): the text table, one line per
// occupied bin.  Built with EXACT_OUTPUT_REDSHIFTS, a spectrum taken after the
// target redshift z_target was crossed (1/a < z_target + 1) is interpolated
// linearly in redshift against the one the previous call wrote to the same
// file, and written as the spectrum at z_target:
//   w = (z_prev - z_target) / (z_prev - z),  P = (1 - w) P_prev + w P,
// P_prev read from the file's occupied bins in order, w = 1 when the header
// cannot be parsed, and nothing interpolated when there is no such file.
// The other columns are this call's.  Failures to open or parse a file are
// reported on stdout and the call carries on, as the reference does.
inline void writePowerSpectrum(const Real *kbin, const Real *power, const Real *kscatter, const Real *pscatter,
                               const int *occupation, const int numbins, const Real rescalek, const Real rescalep,
                               const char *filename, const char *description, double a, const double z_target = -1)
{
    std::vector<Real> power2((size_t)numbins);
    for (int i = 0; i < numbins; i++) power2[i] = power[i] / rescalep;
#ifdef EXACT_OUTPUT_REDSHIFTS
    if (1. / a < z_target + 1.) {
        if (FILE *in = std::fopen(filename, "r")) {
            double zprev = 0., weight = 1.;
            char line[512];
            if (!std::fgets(line, sizeof(line), in) || !std::fgets(line, sizeof(line), in) ||
                std::sscanf(line, "# redshift z=%lf", &zprev) != 1) {
                std::printf(" error parsing power spectrum file header for interpolation (EXACT_OUTPUT_REDSHIFTS)\n");
            } else {
                weight = (zprev - z_target) / (1. + zprev - 1. / a);
                if (!std::fgets(line, sizeof(line), in))
                    std::printf(" error parsing power spectrum file header for interpolation (EXACT_OUTPUT_REDSHIFTS)\n");
                for (int i = 0; i < numbins; i++) {
                    if (occupation[i] <= 0) continue;
                    double k, p;
                    if (std::fscanf(in, " %le %le %*e %*e %*d", &k, &p) != 2) {
                        std::printf(" error parsing power spectrum file data %d for interpolation "
                                    "(EXACT_OUTPUT_REDSHIFTS)\n", i);
                        break;
                    }
                    power2[i] = (Real)p;
                }
            }
            std::fclose(in);
            for (int i = 0; i < numbins; i++) power2[i] = (Real)((1. - weight) * power2[i] + weight * power[i] / rescalep);
            a = 1. / (z_target + 1.);
        }
    }
#else
    (void)z_target;
#endif
    FILE *f = std::fopen(filename, "w");
    if (!f) {
        std::printf(" error opening file for power spectrum output!\n");
        return;
    }
    std::fprintf(f, "# %s\n", description);
    std::fprintf(f, "# redshift z=%f\n", 1. / a - 1.);
    std::fprintf(f, "# k              Pk             sigma(k)       sigma(Pk)      count\n");
    for (int i = 0; i < numbins; i++)
        if (occupation[i] > 0)
            std::fprintf(f, "  %e   %e   %e   %e   %d\n", kbin[i] / rescalek, power2[i], kscatter[i] / rescalek,
                         pscatter[i] / rescalep / std::sqrt((double)occupation[i]), occupation[i]);
    std::fclose(f);
}

// the mixed form's FP64 k=0 mode (no reference counterpart: FP64 mean of an FP32 field)
inline void set_k0(Field<Cplx> &ft, double re, double im = 0.0)
{
    check(gev_set_k0(ft.lattice().ctx(), ft.data(), re, im), "set_k0");
}

}  // namespace gevpm

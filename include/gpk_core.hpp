#pragma once
/*
 * gpk_core.hpp — C++ drop-in facade of gpk::GaussKernel / gpk::ThermoModel /
 * gpk::GPKGrid (include/gpk_core.hpp:30-328 in the GPK source tree) over the
 * HIP library's C ABI (gpk_c.h, libgpk_hip.so).  Header-only; link with
 * -lgpk_hip.
 *
 * Same constructor, step/run, diagnostics, idx/decode and public fields as
 * the reference.  The state lives on the GPU; the public std::vector fields
 * are host mirrors:
 *   - before the first step every mirror is uploaded (initial conditions
 *     written straight into the vectors work unchanged);
 *   - after host writes between steps call mark_dirty() (uploads before the
 *     next device call);
 *   - before reading the mirrors after steps call sync_host() (run() does it
 *     before each callback).
 * Extra constructor arguments select the arithmetic mode (strict32 = the
 * reference's float arithmetic, fp32, fp64) and the device.
 * Errors from the library are rethrown as std::runtime_error.
 */
#include "gpk_c.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace gpk {

class GPKGrid;

// Gaussian kernel weights (include/gpk_core.hpp:37-55): host-side, for callers
// that inspect the smoothing; the device tables are built by the library.
struct GaussKernel {
    std::vector<float> w;
    int radius;
    float sigma_cells;

    explicit GaussKernel(float sigma = 2.0f) : sigma_cells(std::max(sigma, 0.5f)) {
        const float s = sigma_cells;
        radius = static_cast<int>(std::ceil(3.5f * s));
        const int n = 2 * radius + 1;
        w.resize(n);
        float sum = 0.f;
        for (int i = 0; i < n; ++i) {
            const float x = float(i - radius);
            w[i] = std::exp(-x * x / (2.f * s * s));
            sum += w[i];
        }
        for (auto& v : w) v /= sum;
    }
};

// Constitutive model (include/gpk_core.hpp:72-200).  The public parameters
// are the reference's; device_descriptor() states the model to the library
// as a POD descriptor.  The base class has no device form (its default
// mobility is the Read-Shockley law, which no in-scope model uses); it is
// accepted for topology queries and rejected by step().
struct ThermoModel {
    float sigma0 = 0.5f;
    float eps4 = 0.0f;
    float M0 = 1e-8f;
    float Q = 0.0f;
    float Rgas = 8.314f;
    bool phase_selection = false;
    bool grain_network = false;
    bool stochastic = false;
    float noise_scale = 1.0f;
    float kB = 1.380649e-23f;
    bool solute_rejection = false;
    float k_partition = 1.0f;
    bool elastic_enabled = false;
    float elastic_modulus = 0.0f;
    float eigenstrain_coeff = 0.0f;
    float c_ref = 0.0f;
    int elastic_smooth_iters = 8;
    bool dislocation_enabled = false;
    float disloc_coeff = 0.0f;
    float rho_rex = 0.0f;
    float recovery_rate = 0.0f;
    bool magnetic_enabled = false;
    float Ku = 0.f, Ms = 0.f, H_ext = 0.f, H_theta = 0.f, g_easy_ref = 0.f, lambda_s = 0.f;

    virtual ~ThermoModel() = default;

    // Fills *d and returns true when the model has a device form.
    virtual bool device_descriptor(gpk_model_desc* d) const {
        (void)d;
        return false;
    }

  protected:
    // common ThermoModel fields into a descriptor of the given kind
    void fill_common(gpk_model_desc* d, int32_t kind) const {
        gpk_model_defaults(kind, d);
        d->sigma0 = sigma0;
        d->eps4 = eps4;
        d->M0 = M0;
        d->k_partition = k_partition;
        d->noise_scale = noise_scale;
        d->kB = kB;
        d->flags = (grain_network ? GPK_FLAG_GRAIN_NETWORK : 0u) |
                   (solute_rejection ? GPK_FLAG_SOLUTE_REJECTION : 0u) |
                   (stochastic ? GPK_FLAG_STOCHASTIC : 0u) |
                   (dislocation_enabled ? GPK_FLAG_DISLOCATION : 0u) |
                   (elastic_enabled ? GPK_FLAG_ELASTIC : 0u);
        d->elastic_modulus = elastic_modulus;
        d->eigenstrain_coeff = eigenstrain_coeff;
        d->c_ref = c_ref;
        d->elastic_smooth_iters = elastic_smooth_iters;
        d->disloc_coeff = disloc_coeff;
        d->rho_rex = rho_rex;
        d->recovery_rate = recovery_rate;
    }
    // couplings that exist in the reference but not on the device path
    // (stored-energy and elastic couplings are on it; phase selection only
    // through PrecipitationThermo; magnetostrictive eigenstrain is not)
    bool couplings_off() const {
        return !phase_selection && !magnetic_enabled && lambda_s == 0.f && Q == 0.f;
    }
};

class GPKGrid {
  public:
    const int Nz, Ny, Nx;
    const float dx;
    double time = 0.0;

    std::vector<int32_t> phi;
    std::vector<float> alpha, c, T, g, rho;
    std::vector<float> nx_f, ny_f, nz_f, kappa_f, vn_f, E_el;
    std::vector<int32_t> phi_comp;
    std::vector<float> g_comp;
    std::vector<uint8_t> mask;

    uint64_t topology_accepted_last = 0;
    uint64_t topology_deferred_last = 0;
    uint64_t gaussian_remap_exchanges_last = 0;

    GPKGrid(int Nz_, int Ny_, int Nx_, float dx_, std::shared_ptr<ThermoModel> thermo, float D = 0.f,
            float gaussian_sigma = 2.0f, int precision = GPK_PREC_STRICT32, int device = 0)
        : Nz(Nz_), Ny(Ny_), Nx(Nx_), dx(dx_),
          phi(size_t(Nz_) * Ny_ * Nx_, 0), alpha(phi.size(), 1.f), c(phi.size(), 0.f),
          T(phi.size(), 800.f), g(phi.size(), 0.f), rho(phi.size(), 0.f), nx_f(phi.size(), 0.f),
          ny_f(phi.size(), 0.f), nz_f(phi.size(), 0.f), kappa_f(phi.size(), 0.f), vn_f(phi.size(), 0.f),
          E_el(phi.size(), 0.f), phi_comp(phi.size(), 0), g_comp(phi.size(), 0.f), mask(phi.size(), 0),
          thermo_(std::move(thermo)), precision_(precision) {
        gpk_model_desc md;
        supported_ = thermo_ && thermo_->device_descriptor(&md);
        if (!supported_) gpk_model_defaults(GPK_MODEL_ISOTROPIC, &md);
        gpk_grid_desc gd{Nz_, Ny_, Nx_, dx_, D, gaussian_sigma, precision, device};
        check(gpk_create(&gd, &md, &h_));
        handle_.reset(h_, [](gpk_handle* p) { gpk_destroy(p); });
    }

    GPKGrid(const GPKGrid&) = delete;
    GPKGrid& operator=(const GPKGrid&) = delete;

    // src/gpk_simulation.cpp:31-54
    float step(float dt = -1.f, float dt_cap = 1.0e30f) {
        if (!supported_)
            throw std::runtime_error("GPKGrid: this ThermoModel has no device form");
        refresh_model();
        push();
        double used = 0.0;
        gpk_step_stats st{};
        check(gpk_step(h_, dt, dt_cap, &used, &st));
        time = st.time;
        last_time_ = time;
        stale_ = true;   // the device moved on: host mirrors are out of date
        if (sync_each_step) sync_host();
        topology_accepted_last = uint64_t(st.accepted);
        topology_deferred_last = uint64_t(st.deferred);
        gaussian_remap_exchanges_last = uint64_t(st.exchanges);
        return static_cast<float>(used);
    }

    // src/gpk_simulation.cpp:60-72
    int run(double t_end, float dt = -1.f, std::function<void(const GPKGrid&, int)> cb = nullptr,
            int cb_interval = 100) {
        int n = 0;
        while (time < t_end) {
            const float remaining = static_cast<float>(t_end - time);
            step(dt, remaining);
            ++n;
            if (cb && n % cb_interval == 0) {
                sync_host();
                cb(*this, n);
            }
        }
        return n;
    }

    float max_dt() const {
        double v = 0.0;
        check(gpk_max_dt(h_, &v));
        return static_cast<float>(v);
    }
    double phase_volume(int32_t label) const {
        push();
        double v = 0.0;
        check(gpk_phase_volume(h_, label, &v));
        return v;
    }
    double mean_grain_radius() const {
        push();
        double v = 0.0;
        check(gpk_mean_grain_radius(h_, &v));
        return v;
    }
    int n_grains() const {
        push();
        int v = 0;
        check(gpk_n_grains(h_, &v));
        return v;
    }
    float solid_fraction() const {
        push();
        float v = 0.f;
        check(gpk_solid_fraction(h_, &v));
        return v;
    }
    float interface_fraction() const {
        float v = 0.f;
        check(gpk_interface_fraction(h_, &v));
        return v;
    }
    std::size_t allocated_array_bytes() const {
        std::size_t b = 0;
        check(gpk_device_bytes(h_, &b));
        return b;
    }
    bool topology_simple_move(int cell, int32_t new_label) const {
        push();
        int r = 0;
        check(gpk_topology_simple_move(h_, cell, new_label, &r));
        return r != 0;
    }

    int N() const { return Nz * Ny * Nx; }
    inline int idx(int z, int y, int x) const { return z * Ny * Nx + y * Nx + x; }
    inline void decode(int i, int& z, int& y, int& x) const {
        x = i % Nx;
        y = (i / Nx) % Ny;
        z = i / (Ny * Nx);
    }

    // ---- host <-> device ----
    // The device owns the state.  After step() the host mirrors are stale
    // until sync_host(); reading them before that sees the previous state.
    // Callers that edit mirrors call sync_host() first, edit, then
    // mark_dirty(): the next device call uploads them.  Marking stale mirrors
    // dirty would upload old values over newer device state, so the upload
    // refuses it (std::logic_error).  sync_each_step = true downloads every
    // field after each step (drivers written for host-resident fields).
    bool sync_each_step = false;
    void mark_dirty() { dirty_ = true; }
    bool host_stale() const { return stale_; }
    // Uploads the host mirrors now (phi, alpha, c, T, g, rho), whatever the
    // device holds.
    void upload() const {
        dirty_ = true;
        stale_ = false;
        push();
    }
    // Downloads every field into the host mirrors.
    void sync_host() const {
        push();
        stale_ = false;
        auto& self = const_cast<GPKGrid&>(*this);
        get(GPK_FIELD_PHI, self.phi);
        get_real(GPK_FIELD_ALPHA, self.alpha);
        get_real(GPK_FIELD_C, self.c);
        get_real(GPK_FIELD_T, self.T);
        get_real(GPK_FIELD_G, self.g);
        get_real(GPK_FIELD_RHO, self.rho);
        get_real(GPK_FIELD_NX, self.nx_f);
        get_real(GPK_FIELD_NY, self.ny_f);
        get_real(GPK_FIELD_NZ, self.nz_f);
        get_real(GPK_FIELD_KAPPA, self.kappa_f);
        get_real(GPK_FIELD_VN, self.vn_f);
        get(GPK_FIELD_PHI_COMP, self.phi_comp);
        get_real(GPK_FIELD_G_COMP, self.g_comp);
        get(GPK_FIELD_MASK, self.mask);
    }
    // Downloads the labels only (phi); the other mirrors keep their state
    // (and stay stale after a step).
    void sync_labels() const {
        push();
        get(GPK_FIELD_PHI, const_cast<GPKGrid&>(*this).phi);
    }
    // Labels and orientations from voronoi_init (demos/gpk.cpp:51-189),
    // generated on the device; mirrors are refreshed.
    void voronoi_init(int n_seeds, uint32_t seed) {
        push();
        check(gpk_voronoi_init(h_, n_seeds, seed));
        sync_host();
    }

    gpk_handle* handle() const { return h_; }

  private:
    std::shared_ptr<ThermoModel> thermo_;
    int precision_;
    bool supported_ = false;
    gpk_handle* h_ = nullptr;
    std::shared_ptr<gpk_handle> handle_;
    mutable bool dirty_ = true;
    mutable bool stale_ = false;
    mutable double last_time_ = 0.0;

    static void check(int rc) {
        if (rc != GPK_OK) throw std::runtime_error(std::string("gpk: ") + gpk_last_error());
    }
    template <class V>
    void get(int f, V& v) const {
        check(gpk_download(h_, f, v.data(), v.size() * sizeof(v[0])));
    }
    template <class V>
    void put(int f, const V& v) const {
        check(gpk_upload(h_, f, v.data(), v.size() * sizeof(v[0])));
    }
    // upload host mirrors when they may differ from the device
    void push() const {
        if (time != last_time_) {
            check(gpk_set_time(h_, time));
            last_time_ = time;
        }
        if (!dirty_) return;
        if (stale_)
            throw std::logic_error(
                "GPKGrid: host fields were marked dirty after a step without sync_host(); "
                "uploading them would overwrite newer device state (call sync_host() before editing)");
        if (precision_ == GPK_PREC_FP64) {
            // the FP64 engine holds doubles: widen the float mirrors
            put(GPK_FIELD_PHI, phi);
            const std::vector<float>* f[5] = {&alpha, &c, &T, &g, &rho};
            const int id[5] = {GPK_FIELD_ALPHA, GPK_FIELD_C, GPK_FIELD_T, GPK_FIELD_G, GPK_FIELD_RHO};
            for (int k = 0; k < 5; ++k) {
                std::vector<double> dv(f[k]->begin(), f[k]->end());
                put(id[k], dv);
            }
        } else {
            put(GPK_FIELD_PHI, phi);
            put(GPK_FIELD_ALPHA, alpha);
            put(GPK_FIELD_C, c);
            put(GPK_FIELD_T, T);
            put(GPK_FIELD_G, g);
            put(GPK_FIELD_RHO, rho);
        }
        dirty_ = false;
    }
    // parameters may change between steps (e.g. TensorAnisotropyThermo::dG_bias)
    void refresh_model() {
        gpk_model_desc md;
        if (thermo_ && thermo_->device_descriptor(&md)) check(gpk_set_model(h_, &md));
    }

    // real field into a float mirror (the FP64 engine narrows on the host)
    void get_real(int f, std::vector<float>& v) const {
        if (precision_ != GPK_PREC_FP64) {
            get(f, v);
            return;
        }
        std::vector<double> dv(v.size());
        get(f, dv);
        std::copy(dv.begin(), dv.end(), v.begin());
    }
};

} // namespace gpk

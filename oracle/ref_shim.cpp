/*
 * ref_shim.cpp — C entry points around the reference's own GPKGrid, compiled
 * together with the reference sources (src/*.cpp) into oracle/_ref/libgpkref.so
 * by oracle/Makefile.  TEST INFRASTRUCTURE ONLY: tests use it to pin the
 * oracle restatement, and bench.py may time it as the CPU baseline.
 *
 * Only the reference's public API is used (include/gpk_core.hpp:205-326,
 * include/gpk_thermo.hpp model constructors).
 */
#include "gpk_core.hpp"
#include "gpk_thermo.hpp"
#include "../include/gpk_c.h"

#include <cstring>
#include <memory>

#include <omp.h>

namespace {

struct RefHandle {
    std::shared_ptr<gpk::ThermoModel> thermo;
    std::unique_ptr<gpk::GPKGrid> grid;
};

std::shared_ptr<gpk::ThermoModel> make_model(const gpk_model_desc& m) {
    std::shared_ptr<gpk::ThermoModel> t;
    switch (m.kind) {
    case GPK_MODEL_ISOTROPIC:
        t = std::make_shared<gpk::IsotropicGrowthThermo>(m.M0, m.sigma0); break;
    case GPK_MODEL_ANISO_GROWTH:
        t = std::make_shared<gpk::AnisotropicGrowthThermo>(m.M0, m.sigma0, m.eps4); break;
    case GPK_MODEL_SOLIDIFICATION:
        t = std::make_shared<gpk::SolidificationThermo>(m.M0, m.sigma0, m.eps4, m.m_liq,
                                                         m.k_part, m.T_melt, m.L); break;
    case GPK_MODEL_TENSOR_ANISO: {
        auto p = std::make_shared<gpk::TensorAnisotropyThermo>(m.M0, m.sigma0, m.bxx, m.byy, m.bzz);
        p->dG_bias = m.dG_bias;
        t = p; break;
    }
    case GPK_MODEL_ANISO_OSTWALD:
        t = std::make_shared<gpk::AnisoOstwaldThermo>(m.M0, m.sigma0, m.bxx, m.byy, m.bzz,
                                                       m.c_sat, m.Omega); break;
    case GPK_MODEL_RECRYSTALLISATION:
        t = std::make_shared<gpk::RecrystallisationThermo>(m.M0, m.sigma0, m.E_stored); break;
    case GPK_MODEL_PRECIPITATION: {
        const int n = m.n_phases;
        std::vector<float> a(m.pa, m.pa + n), b(m.pb, m.pb + n), c2(m.pc2, m.pc2 + n), dT(m.pdT, m.pdT + n);
        t = std::make_shared<gpk::PrecipitationThermo>(m.M0, m.sigma0, n, a, b, c2, dT, m.T_ref);
        break;
    }
    case GPK_MODEL_DISLOCATION:
        // constructed with unit mu, b; the coefficient is taken from the descriptor
        t = std::make_shared<gpk::DislocationThermo>(m.M0, m.sigma0, 1.f, 1.f); break;
    default:
        return nullptr;
    }
    // flags are set explicitly so that a descriptor can switch them either way
    t->grain_network = (m.flags & GPK_FLAG_GRAIN_NETWORK) != 0;
    t->solute_rejection = (m.flags & GPK_FLAG_SOLUTE_REJECTION) != 0;
    t->stochastic = (m.flags & GPK_FLAG_STOCHASTIC) != 0;
    t->k_partition = m.k_partition;
    t->noise_scale = m.noise_scale;
    t->kB = m.kB;
    t->dislocation_enabled = (m.flags & GPK_FLAG_DISLOCATION) != 0;
    t->disloc_coeff = m.disloc_coeff;
    t->rho_rex = m.rho_rex;
    t->recovery_rate = m.recovery_rate;
    t->elastic_enabled = (m.flags & GPK_FLAG_ELASTIC) != 0;
    t->elastic_modulus = m.elastic_modulus;
    t->eigenstrain_coeff = m.eigenstrain_coeff;
    t->c_ref = m.c_ref;
    t->elastic_smooth_iters = m.elastic_smooth_iters;
    return t;
}

template <class V>
int copy_in(V& v, const void* p, size_t b) {
    if (b != v.size() * sizeof(v[0])) return GPK_E_INVALID;
    std::memcpy(v.data(), p, b);
    return GPK_OK;
}
template <class V>
int copy_out(const V& v, void* p, size_t b) {
    if (b != v.size() * sizeof(v[0])) return GPK_E_INVALID;
    std::memcpy(p, v.data(), b);
    return GPK_OK;
}

} // namespace

extern "C" {

void* rf_create(int Nz, int Ny, int Nx, float dx, float D, float sigma, const gpk_model_desc* m) {
    if (!m) return nullptr;
    auto h = new RefHandle;
    h->thermo = make_model(*m);
    if (!h->thermo) { delete h; return nullptr; }
    h->grid.reset(new gpk::GPKGrid(Nz, Ny, Nx, dx, h->thermo, D, sigma));
    return h;
}

void rf_destroy(void* p) { delete static_cast<RefHandle*>(p); }

int rf_set(void* p, int f, const void* src, size_t b) {
    auto& g = *static_cast<RefHandle*>(p)->grid;
    switch (f) {
    case GPK_FIELD_PHI: return copy_in(g.phi, src, b);
    case GPK_FIELD_ALPHA: return copy_in(g.alpha, src, b);
    case GPK_FIELD_C: return copy_in(g.c, src, b);
    case GPK_FIELD_T: return copy_in(g.T, src, b);
    case GPK_FIELD_G: return copy_in(g.g, src, b);
    case GPK_FIELD_RHO: return copy_in(g.rho, src, b);
    default: return GPK_E_INVALID;
    }
}

int rf_get(void* p, int f, void* dst, size_t b) {
    auto& g = *static_cast<RefHandle*>(p)->grid;
    switch (f) {
    case GPK_FIELD_PHI: return copy_out(g.phi, dst, b);
    case GPK_FIELD_ALPHA: return copy_out(g.alpha, dst, b);
    case GPK_FIELD_C: return copy_out(g.c, dst, b);
    case GPK_FIELD_T: return copy_out(g.T, dst, b);
    case GPK_FIELD_G: return copy_out(g.g, dst, b);
    case GPK_FIELD_RHO: return copy_out(g.rho, dst, b);
    case GPK_FIELD_E_EL: return copy_out(g.E_el, dst, b);
    case GPK_FIELD_NX: return copy_out(g.nx_f, dst, b);
    case GPK_FIELD_NY: return copy_out(g.ny_f, dst, b);
    case GPK_FIELD_NZ: return copy_out(g.nz_f, dst, b);
    case GPK_FIELD_KAPPA: return copy_out(g.kappa_f, dst, b);
    case GPK_FIELD_VN: return copy_out(g.vn_f, dst, b);
    case GPK_FIELD_PHI_COMP: return copy_out(g.phi_comp, dst, b);
    case GPK_FIELD_G_COMP: return copy_out(g.g_comp, dst, b);
    case GPK_FIELD_MASK: return copy_out(g.mask, dst, b);
    default: return GPK_E_INVALID;
    }
}

double rf_step(void* p, float dt, float cap) {
    return static_cast<double>(static_cast<RefHandle*>(p)->grid->step(dt, cap));
}

int rf_run(void* p, double t_end, float dt) {
    return static_cast<RefHandle*>(p)->grid->run(t_end, dt);
}

double rf_time(void* p) { return static_cast<RefHandle*>(p)->grid->time; }
void rf_set_time(void* p, double t) { static_cast<RefHandle*>(p)->grid->time = t; }

void rf_counters(void* p, int64_t* out3) {
    auto& g = *static_cast<RefHandle*>(p)->grid;
    out3[0] = static_cast<int64_t>(g.topology_accepted_last);
    out3[1] = static_cast<int64_t>(g.topology_deferred_last);
    out3[2] = static_cast<int64_t>(g.gaussian_remap_exchanges_last);
}

double rf_max_dt(void* p) { return static_cast<RefHandle*>(p)->grid->max_dt(); }
double rf_phase_volume(void* p, int32_t l) { return static_cast<RefHandle*>(p)->grid->phase_volume(l); }
int rf_n_grains(void* p) { return static_cast<RefHandle*>(p)->grid->n_grains(); }
int rf_simple_move(void* p, int cell, int32_t nl) {
    return static_cast<RefHandle*>(p)->grid->topology_simple_move(cell, nl) ? 1 : 0;
}
int rf_threads(void) {
    int n = 1;
    #pragma omp parallel
    {
        #pragma omp single
        n = omp_get_num_threads();
    }
    return n;
}

} // extern "C"

#pragma once
/*
 * gpk_thermo.hpp — the constitutive models of the GPU path with the
 * reference's class names, constructors and parameters
 * (include/gpk_thermo.hpp in the GPK source tree: 18-85, 229-325).  Each
 * states itself to the HIP library through device_descriptor(); couplings the
 * device path does not evaluate (elastic, dislocation, magnetic, phase
 * selection, Arrhenius mobility) make device_descriptor() return false.
 */
#include "gpk_core.hpp"

#include <stdexcept>
#include <vector>

namespace gpk {

// Isotropic grain growth, dG = 0, M = M0 (include/gpk_thermo.hpp:18-27)
struct IsotropicGrowthThermo : ThermoModel {
    explicit IsotropicGrowthThermo(float M0_val, float sigma0_val = 0.5f) {
        M0 = M0_val;
        sigma0 = sigma0_val;
        eps4 = 0.f;
        grain_network = true;
    }
    bool device_descriptor(gpk_model_desc* d) const override {
        if (!couplings_off()) return false;
        fill_common(d, GPK_MODEL_ISOTROPIC);
        return true;
    }
};

// Anisotropic grain growth, cubic / 4-fold sigma (include/gpk_thermo.hpp:33-45)
struct AnisotropicGrowthThermo : ThermoModel {
    float eps6 = 0.f;   // 2-D six-fold term (extension; 0 = the reference law)
    AnisotropicGrowthThermo(float M0_val, float sigma0_val, float eps4_val) {
        M0 = M0_val;
        sigma0 = sigma0_val;
        eps4 = eps4_val;
        grain_network = true;
    }
    bool device_descriptor(gpk_model_desc* d) const override {
        if (!couplings_off()) return false;
        fill_common(d, GPK_MODEL_ANISO_GROWTH);
        d->eps6 = eps6;
        return true;
    }
};

// Binary-alloy solidification, 0 = liquid, 1 = solid (include/gpk_thermo.hpp:52-85)
struct SolidificationThermo : ThermoModel {
    float m_liq, k_part, T_melt, L;
    SolidificationThermo(float M0_val = 5e-8f, float sigma0_v = 0.2f, float eps4_v = 0.04f,
                         float m_liq_v = -500.f, float k_part_v = 0.3f, float T_melt_v = 933.f,
                         float L_v = 1e8f)
        : m_liq(m_liq_v), k_part(k_part_v), T_melt(T_melt_v), L(L_v) {
        M0 = M0_val;
        sigma0 = sigma0_v;
        eps4 = eps4_v;
    }
    bool device_descriptor(gpk_model_desc* d) const override {
        if (!couplings_off()) return false;
        fill_common(d, GPK_MODEL_SOLIDIFICATION);
        d->m_liq = m_liq;
        d->k_part = k_part;
        d->T_melt = T_melt;
        d->L = L;
        return true;
    }
};

// Tensor interfacial energy with an optional volume bias (include/gpk_thermo.hpp:229-263)
struct TensorAnisotropyThermo : ThermoModel {
    float bxx, byy, bzz;
    mutable float dG_bias = 0.f;
    TensorAnisotropyThermo(float M0_val, float sigma0_val, float bxx_val = 1.f, float byy_val = 1.f,
                           float bzz_val = 1.f)
        : bxx(bxx_val), byy(byy_val), bzz(bzz_val) {
        M0 = M0_val;
        sigma0 = sigma0_val;
        eps4 = 0.f;
    }
    bool device_descriptor(gpk_model_desc* d) const override {
        if (!couplings_off()) return false;
        fill_common(d, GPK_MODEL_TENSOR_ANISO);
        d->bxx = bxx; d->byy = byy; d->bzz = bzz;
        d->dG_bias = dG_bias;
        return true;
    }
};

// Tensor sigma + volume-conserving composition coupling (include/gpk_thermo.hpp:284-325)
struct AnisoOstwaldThermo : ThermoModel {
    float bxx, byy, bzz;
    float c_sat, Omega;
    AnisoOstwaldThermo(float M0_val, float sigma0_val, float bxx_val, float byy_val, float bzz_val,
                       float c_sat_val, float Omega_val)
        : bxx(bxx_val), byy(byy_val), bzz(bzz_val), c_sat(c_sat_val), Omega(Omega_val) {
        M0 = M0_val;
        sigma0 = sigma0_val;
        eps4 = 0.f;
    }
    bool device_descriptor(gpk_model_desc* d) const override {
        if (!couplings_off()) return false;
        fill_common(d, GPK_MODEL_ANISO_OSTWALD);
        d->bxx = bxx; d->byy = byy; d->bzz = bzz;
        d->c_sat = c_sat; d->Omega = Omega;
        return true;
    }
};

// include/gpk_thermo.hpp:90-111 (GPK source tree): label 0 is the deformed
// matrix, labels >= 1 recrystallised grains; E_stored drives 0 -> rex.
struct RecrystallisationThermo : ThermoModel {
    float E_stored;
    RecrystallisationThermo(float M0_val, float sigma0_v, float E_stored_v) : E_stored(E_stored_v) {
        M0 = M0_val;
        sigma0 = sigma0_v;
        eps4 = 0.f;
    }
    bool device_descriptor(gpk_model_desc* d) const override {
        if (!couplings_off()) return false;
        fill_common(d, GPK_MODEL_RECRYSTALLISATION);
        d->E_stored = E_stored;
        return true;
    }
};

// include/gpk_thermo.hpp:135-170: stored energy disloc_coeff * rho with
// disloc_coeff = 0.5 alphaT mu b^2, sweeping to rho_rex, static recovery.
struct DislocationThermo : ThermoModel {
    float mu, b, alphaT;
    DislocationThermo(float M0_val, float sigma0_val, float mu_val, float b_val, float alphaT_val = 0.5f,
                      float recovery_rate_val = 0.f, float rho_rex_val = 0.f)
        : mu(mu_val), b(b_val), alphaT(alphaT_val) {
        M0 = M0_val;
        sigma0 = sigma0_val;
        eps4 = 0.f;
        dislocation_enabled = true;
        disloc_coeff = 0.5f * alphaT * mu * b * b;
        recovery_rate = recovery_rate_val;
        rho_rex = rho_rex_val;
    }
    float stored_energy(float rho_val) const { return disloc_coeff * rho_val; }
    float critical_radius(float rho_val) const {
        const float E = disloc_coeff * rho_val;
        return (E > 0.f) ? sigma0 / E : 1e30f;
    }
    bool device_descriptor(gpk_model_desc* d) const override {
        if (!couplings_off()) return false;
        fill_common(d, GPK_MODEL_DISLOCATION);
        return true;
    }
};

// include/gpk_thermo.hpp:333-387: phases 0..n-1 with polynomial Gibbs
// energies G_q(c, T) = a_q + b_q c + c2_q c^2 + dT_q (T - T_ref); the product
// phase of a cell is chosen by local Gibbs minimisation on the device.
struct PrecipitationThermo : ThermoModel {
    int n_phases = 0;
    float T_ref = 900.f;
    std::vector<float> a, b, c2, dT;
    PrecipitationThermo(float M0_val, float sigma0_val, int n_phases_val, const std::vector<float>& a_val,
                        const std::vector<float>& b_val, const std::vector<float>& c2_val,
                        const std::vector<float>& dT_val, float T_ref_val = 900.f)
        : n_phases(n_phases_val), T_ref(T_ref_val), a(a_val), b(b_val), c2(c2_val), dT(dT_val) {
        M0 = M0_val;
        sigma0 = sigma0_val;
        eps4 = 0.f;
        phase_selection = true;
        if (int(a.size()) != n_phases || int(b.size()) != n_phases || int(c2.size()) != n_phases ||
            int(dT.size()) != n_phases)
            throw std::runtime_error("PrecipitationThermo: coefficient size mismatch");
    }
    float G(int q, float c_val, float T_val) const {
        return a[q] + b[q] * c_val + c2[q] * c_val * c_val + dT[q] * (T_val - T_ref);
    }
    bool device_descriptor(gpk_model_desc* d) const override {
        if (magnetic_enabled || lambda_s != 0.f || Q != 0.f || n_phases < 1 || n_phases > GPK_MAX_PHASES)
            return false;
        fill_common(d, GPK_MODEL_PRECIPITATION);
        d->n_phases = n_phases;
        d->T_ref = T_ref;
        for (int q = 0; q < n_phases; ++q) {
            d->pa[q] = a[q]; d->pb[q] = b[q]; d->pc2[q] = c2[q]; d->pdT[q] = dT[q];
        }
        return true;
    }
};

} // namespace gpk

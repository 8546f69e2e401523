/*
 * gpk_c.h — C ABI of the MI355X-native GPK per-step kinetics library
 * (libgpk_hip.so).
 *
 * Plain pointers and sizes only: no C++ or torch types cross this boundary.
 * Every entry point returns an int status (GPK_OK == 0, negative GPK_E_* on
 * failure) and never lets an exception escape; gpk_last_error() returns a
 * thread-local description of the last failure on the calling thread.
 *
 * Each function names the reference interface it replaces
 * (paths relative to the GPK source tree).
 *
 * Threading: a handle owns one HIP stream and is not thread-safe; distinct
 * handles may be driven from distinct host threads.  Work is asynchronous on
 * the handle's stream until a download, a diagnostic or gpk_sync().
 */
#ifndef GPK_C_H
#define GPK_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum gpk_status {
    GPK_OK                   =  0,
    GPK_E_INVALID            = -1,  /* bad argument (null, size mismatch, range) */
    GPK_E_HIP                = -2,  /* HIP runtime failure (message has details) */
    GPK_E_UNSUPPORTED_MODEL  = -3,  /* model kind/flag not available on device */
    GPK_E_LABEL_RANGE        = -4,  /* labels outside [0, GPK_MAX_LABEL] */
    GPK_E_OOM                = -5,  /* device allocation failed */
    GPK_E_NO_DEVICE          = -6,  /* no HIP device visible */
    GPK_E_STALLED            = -7   /* Voronoi connectivity normalisation stalled */
};

/* Labels are dense non-negative integers; per-label tables are sized to the
 * largest label present (plus headroom), capped at this value. */
#define GPK_MAX_LABEL (1 << 26)

/* ---- arithmetic modes ------------------------------------------------- */
enum gpk_precision {
    /* FP32 fields, geometry as the reference's direct masked stencil with the
     * reference's operation order: bit-exact with the reference CPU step. */
    GPK_PREC_STRICT32 = 0,
    /* FP32 fields, geometry through separable row sums with FMA
     * (fewer operations; compared to FP64 by error statistics). */
    GPK_PREC_FP32     = 1,
    /* FP64 fields and weights, the same separable geometry order as FP32. */
    GPK_PREC_FP64     = 2
};

/* ---- state fields ----------------------------------------------------- */
/* dtype: I32 = int32, U8 = uint8, R = float (STRICT32/FP32) or double (FP64) */
enum gpk_field {
    GPK_FIELD_PHI      = 0,   /* I32  grain / phase label            */
    GPK_FIELD_ALPHA    = 1,   /* R    sub-cell coordinate            */
    GPK_FIELD_C        = 2,   /* R    composition                    */
    GPK_FIELD_T        = 3,   /* R    temperature                    */
    GPK_FIELD_G        = 4,   /* R    orientation                    */
    GPK_FIELD_RHO      = 5,   /* R    dislocation density (carried)  */
    GPK_FIELD_NX       = 6,   /* R    outward normal (band cells)    */
    GPK_FIELD_NY       = 7,
    GPK_FIELD_NZ       = 8,
    GPK_FIELD_KAPPA    = 9,   /* R    curvature (band cells)         */
    GPK_FIELD_VN       = 10,  /* R    projected normal velocity      */
    GPK_FIELD_PHI_COMP = 11,  /* I32  competitor label               */
    GPK_FIELD_G_COMP   = 12,  /* R    competitor orientation         */
    GPK_FIELD_MASK     = 13,  /* U8   interface mask                 */
    GPK_FIELD_E_EL     = 14,  /* R    elastic energy density (GPK_FLAG_ELASTIC) */
    GPK_FIELD_COUNT    = 15
};

/* ---- constitutive models (POD replacement of the ThermoModel plugin) --- */
enum gpk_model_kind {
    GPK_MODEL_ISOTROPIC      = 0,  /* IsotropicGrowthThermo   include/gpk_thermo.hpp:18-27  */
    GPK_MODEL_ANISO_GROWTH   = 1,  /* AnisotropicGrowthThermo include/gpk_thermo.hpp:33-45  */
    GPK_MODEL_SOLIDIFICATION = 2,  /* SolidificationThermo    include/gpk_thermo.hpp:52-85  */
    GPK_MODEL_TENSOR_ANISO   = 3,  /* TensorAnisotropyThermo  include/gpk_thermo.hpp:229-263 */
    GPK_MODEL_ANISO_OSTWALD  = 4,  /* AnisoOstwaldThermo      include/gpk_thermo.hpp:284-325 */
    GPK_MODEL_RECRYSTALLISATION = 5, /* RecrystallisationThermo include/gpk_thermo.hpp:90-111 */
    GPK_MODEL_DISLOCATION    = 6,  /* DislocationThermo       include/gpk_thermo.hpp:135-170 */
    GPK_MODEL_PRECIPITATION  = 7   /* PrecipitationThermo     include/gpk_thermo.hpp:333-387 */
};

#define GPK_MAX_PHASES 8           /* PrecipitationThermo phases held by the descriptor */

/* flag bits of gpk_model_desc.flags (ThermoModel fields, include/gpk_core.hpp:72-123) */
#define GPK_FLAG_GRAIN_NETWORK    0x1u
#define GPK_FLAG_SOLUTE_REJECTION 0x2u
#define GPK_FLAG_STOCHASTIC       0x4u
#define GPK_FLAG_DISLOCATION      0x8u   /* dislocation_enabled: rho coupling, sweep, recovery */
#define GPK_FLAG_ELASTIC          0x10u  /* elastic_enabled: smoothed-eigenstrain energy coupling */

typedef struct gpk_model_desc {
    int32_t  kind;          /* gpk_model_kind */
    uint32_t flags;         /* GPK_FLAG_* */
    float sigma0;           /* baseline interfacial energy */
    float eps4;             /* 4-fold / cubic anisotropy */
    float eps6;             /* 6-fold 2-D anisotropy (extension; 0 = reference law) */
    float M0;               /* mobility */
    float k_partition;      /* solute partition coefficient (solute_rejection) */
    float noise_scale;      /* stochastic kinetics amplitude */
    float kB;               /* Boltzmann constant used by the noise term */
    /* SolidificationThermo */
    float m_liq, k_part, T_melt, L;
    /* TensorAnisotropyThermo / AnisoOstwaldThermo */
    float bxx, byy, bzz;
    float c_sat, Omega;     /* AnisoOstwaldThermo */
    float dG_bias;          /* TensorAnisotropyThermo (mutable between steps) */
    float E_stored;         /* RecrystallisationThermo */
    /* stored-energy coupling (include/gpk_core.hpp:101-110), GPK_FLAG_DISLOCATION */
    float disloc_coeff;     /* E_rho = disloc_coeff * rho */
    float rho_rex;          /* density left behind a swept boundary */
    float recovery_rate;    /* d(rho)/dt = -recovery_rate * rho */
    /* elastic coupling (src/gpk_physics.cpp:10-63), GPK_FLAG_ELASTIC:
     * eps* = eigenstrain_coeff (c - c_ref), smoothed elastic_smooth_iters times */
    float elastic_modulus, eigenstrain_coeff, c_ref;
    int32_t elastic_smooth_iters;
    /* PrecipitationThermo: G_q(c, T) = a_q + b_q c + c2_q c^2 + dT_q (T - T_ref),
     * q < n_phases <= GPK_MAX_PHASES; the product phase is selected per cell */
    int32_t n_phases;
    float T_ref;
    float pa[GPK_MAX_PHASES], pb[GPK_MAX_PHASES], pc2[GPK_MAX_PHASES], pdT[GPK_MAX_PHASES];
} gpk_model_desc;

/* Fills *m with the defaults of the reference ThermoModel base and the given
 * kind (gpk_core.hpp:73-123 and each model's constructor defaults). */
void gpk_model_defaults(int32_t kind, gpk_model_desc* m);

typedef struct gpk_grid_desc {
    int32_t Nz, Ny, Nx;     /* Nz == 1 selects the 2-D path */
    float   dx;             /* cell size */
    float   D;              /* solute diffusivity */
    float   gaussian_sigma; /* geometry Gaussian width in cells */
    int32_t precision;      /* gpk_precision */
    int32_t device;         /* HIP device ordinal */
} gpk_grid_desc;

/* Per-step counters (include/gpk_core.hpp:233-235 plus the band size). */
typedef struct gpk_step_stats {
    int64_t band_cells;     /* |A| after the mask pass at the start of the step */
    int64_t candidates;     /* cells whose alpha crossed the flip threshold */
    int64_t accepted;       /* topology_accepted_last */
    int64_t deferred;       /* topology_deferred_last */
    int64_t exchanges;      /* gaussian_remap_exchanges_last */
    int64_t eroded;         /* cells relabelled by unresolved-grain extinction */
    double  dt;             /* timestep used */
    double  time;           /* simulation time after the step */
    int64_t step_id;        /* steps taken by this handle */
    int64_t guarded;        /* FP32 mode: flip decisions taken again in FP64 because the FP32
                               alpha_new lay within 2.5e-7 of the threshold (0 otherwise) */
} gpk_step_stats;

typedef struct gpk_handle gpk_handle;

/* Thread-local text of the last error on this thread ("" if none). */
const char* gpk_last_error(void);

/* Library version string. */
const char* gpk_version(void);

/* Replaces GPKGrid::GPKGrid (include/gpk_core.hpp:237-240, src/gpk_grid.cpp:9-49):
 * allocates the device state with the reference's initial values
 * (phi 0, alpha 1, c 0, T 800, g 0, rho 0). */
int gpk_create(const gpk_grid_desc* grid, const gpk_model_desc* model,
               gpk_handle** out);

/* Releases every device buffer of the handle. */
int gpk_destroy(gpk_handle* h);

/* Replaces direct writes to the public GPKGrid vectors
 * (include/gpk_core.hpp:212-228): copies a full field from host memory.
 * bytes must equal N * sizeof(dtype).  Uploading PHI rebuilds the label
 * tables (counts) used by the flip and extinction passes. */
int gpk_upload(gpk_handle* h, int field, const void* host, size_t bytes);

/* Replaces reads of the public GPKGrid vectors: copies a full field to host
 * memory after the handle's stream has drained.  Band-only fields are
 * returned at full size with zeros outside the band, as the reference
 * holds them. */
int gpk_download(gpk_handle* h, int field, void* host, size_t bytes);

/* Device-to-device variants (for callers that keep state in HBM, e.g. torch):
 * ptr is a device address on the handle's device; the copy is ordered on the
 * handle's stream. */
int gpk_upload_device(gpk_handle* h, int field, const void* dev_ptr, size_t bytes);
int gpk_download_device(gpk_handle* h, int field, void* dev_ptr, size_t bytes);

/* Replaces GPKGrid::step (include/gpk_core.hpp:243, src/gpk_simulation.cpp:31-54).
 * dt < 0 selects the stable timestep.  Writes the timestep used to *dt_used
 * and the step counters to *stats (either may be NULL). */
int gpk_step(gpk_handle* h, float dt, float dt_cap, double* dt_used,
             gpk_step_stats* stats);

/* Runs nsteps consecutive step(dt, dt_cap) calls without returning to the
 * host in between (no per-step synchronisation).  If stats is non-NULL it
 * receives the last step's counters with band_cells, candidates, accepted,
 * deferred, exchanges and eroded summed over the nsteps. */
int gpk_step_many(gpk_handle* h, int nsteps, float dt, float dt_cap,
                  gpk_step_stats* stats);

/* Replaces GPKGrid::run (include/gpk_core.hpp:246-248, src/gpk_simulation.cpp:60-72)
 * without the callback: steps until time >= t_end; *steps_taken receives the
 * count.  Callbacks are driven by the host facade between calls. */
int gpk_run(gpk_handle* h, double t_end, float dt, int* steps_taken);

/* Updates the model parameters between steps (e.g. TensorAnisotropyThermo::
 * dG_bias, include/gpk_thermo.hpp:238).  The kind must not change. */
int gpk_set_model(gpk_handle* h, const gpk_model_desc* model);

/* GPKGrid::time accessors (include/gpk_core.hpp:209). */
int gpk_get_time(gpk_handle* h, double* t);
int gpk_set_time(gpk_handle* h, double t);

/* Diagnostics (src/gpk_simulation.cpp:78-123, include/gpk_core.hpp:250-255). */
int gpk_max_dt(gpk_handle* h, double* dt);
int gpk_phase_volume(gpk_handle* h, int32_t label, double* volume);
int gpk_n_grains(gpk_handle* h, int* n);
int gpk_mean_grain_radius(gpk_handle* h, double* radius);
int gpk_solid_fraction(gpk_handle* h, float* fraction);
int gpk_interface_fraction(gpk_handle* h, float* fraction);
/* Sum of c over the grid in FP64 (fixed order, deterministic); solute mass. */
int gpk_solute_sum(gpk_handle* h, double* sum);
/* Bytes of device memory held by the handle (cf. allocated_array_bytes,
 * src/gpk_simulation.cpp:125-144). */
int gpk_device_bytes(gpk_handle* h, size_t* bytes);

/* Replaces GPKGrid::topology_simple_move (include/gpk_core.hpp:264,
 * src/gpk_connectivity.cpp:197-206): *result = 1 when relabelling cell to
 * new_label passes the face-connectivity certificate on the current labels. */
int gpk_topology_simple_move(gpk_handle* h, int cell, int32_t new_label, int* result);

/* Replaces voronoi_init (demos/gpk.cpp:51-189): periodic Voronoi labels from
 * n_seeds unique sites drawn with std::mt19937(seed), orientations from
 * uniform_real<float>[0, pi/2), alpha = 1, then the face-connectivity
 * normalisation.  Writes PHI, G, ALPHA on the device. */
int gpk_voronoi_init(gpk_handle* h, int n_seeds, uint32_t seed);

/* Host-only form of the same generator: fills phi[N] and g[N] (float). */
int gpk_voronoi_host(int Nz, int Ny, int Nx, int n_seeds, uint32_t seed,
                     int32_t* phi, float* g);

/* Blocks until the handle's stream has drained. */
int gpk_sync(gpk_handle* h);

/* Kernel timing (no reference counterpart; measurement support).
 * gpk_set_profiling(h, 1) resets the counters and starts recording a HIP
 * event pair around every geometry launch on the handle's stream (no host
 * synchronisation is added).  gpk_kernel_timing(h, 0, ...) drains the stream
 * and returns the average launch duration in ms, the number of launches and
 * the algorithmic flops per launch of the reference-formulation stencil
 * (2 * window cells * components * band cells). */
int gpk_set_profiling(gpk_handle* h, int enable);
int gpk_kernel_timing(gpk_handle* h, int which, double* avg_ms, int64_t* launches,
                      double* flops_per_launch);

/* Velocity chain and its adjoint (no reference counterpart: the reference
 * has no derivatives).  theta = {eps4, eps6, M0} replaces those model
 * parameters for the call only; models ISOTROPIC and ANISO_GROWTH.
 *
 * gpk_velocity_field recomputes the band, geometry and constitutive velocity
 * on the current labels (src/gpk_physics.cpp:67-148), applies the pair
 * projection (src/gpk_physics.cpp:152-193) and writes the projected velocity
 * of every cell into vout_dev (device pointer, N values of the handle's real
 * type: float for STRICT32/FP32, double for FP64; 0 off the band).
 *
 * gpk_velocity_adjoint takes gvel_dev = dL/dv (device, N values of the real
 * type, read on band cells only) and returns dL/dtheta in grad[3] (FP64,
 * fixed-order reduction): the transposed projection applied to gvel, times
 * dv_raw/dtheta.  Normals, curvature and projection weights are recomputed.
 * Both calls return after the handle's stream has drained; the caller makes
 * sure gvel_dev is complete (and vout_dev unused) before the call. */
int gpk_velocity_field(gpk_handle* h, const double theta[3], void* vout_dev);
int gpk_velocity_adjoint(gpk_handle* h, const double theta[3], const void* gvel_dev, double grad[3]);
/* gpk_velocity_adjoint without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and result as gpk_velocity_adjoint, bit for bit; the scratch is a
 * static 16 MiB device array per GPU (held until the process ends), enough
 * for about 2.09 million band cells: a larger band returns GPK_E_INVALID, a
 * 3-D grid with geometry radius above 18 GPK_E_UNSUPPORTED_MODEL.  Calls on
 * one GPU share that array: calls that overlap in time on one GPU (several
 * host threads, or handles on different streams) race. */
int gpk_velocity_adjoint_noalloc(gpk_handle* h, const double theta[3], const void* gvel_dev, double grad[3]);

/* ---- Batched 2-D instances ------------------------------------------------
 * No counterpart in the reference, which steps the grids of a sweep one
 * after another (demos/gpk.cpp:1121-1259).  A batch holds n_inst 2-D grids
 * (Ny[b] x Nx[b], shared dx, sigma, precision and model) packed back to back
 * in instance order; one persistent kernel advances every instance by one
 * GPKGrid::step per launch, each with its own dt and time.  Fields are
 * uploaded / downloaded packed (instance 0's cells first, each instance in
 * its own z-y-x order) for GPK_FIELD_PHI, ALPHA, G, C, T, MASK, PHI_COMP,
 * G_COMP.  Scope: D = 0; no solute rejection, noise, stored-energy
 * (dislocation), precipitation or elastic terms (e.g. ISOTROPIC,
 * ANISO_GROWTH, RECRYSTALLISATION, ANISO_OSTWALD with c = c_sat,
 * TENSOR_ANISO); grain-network models (GPK_FLAG_GRAIN_NETWORK) run the
 * topology test and the extinction of unresolved grains per instance;
 * labels in [0, 1024) per instance (else GPK_E_LABEL_RANGE).  Other models
 * return GPK_E_UNSUPPORTED_MODEL. */
typedef struct gpk_batch gpk_batch;

typedef struct gpk_batch_stats {
    int64_t band_cells;   /* sum over instances and steps of |A| */
    int64_t accepted;     /* accepted flips */
    int64_t exchanges;    /* remap exchanges */
    int64_t proposals;    /* remap proposals */
    int32_t instances;
    int32_t workgroup_size, workgroups;   /* launch shape in use */
    int64_t guarded;      /* FP32 mode: flip decisions taken again in FP64 (guard band) */
    int64_t deferred;     /* grain-network models: candidates the topology test refused */
    int64_t eroded;       /* grain-network models: cells of unresolved grains relabelled */
} gpk_batch_stats;

int gpk_batch_create(int n_inst, const int* Ny, const int* Nx, float dx, float D, float gaussian_sigma,
                     int precision, int device, const gpk_model_desc* model, gpk_batch** out);
int gpk_batch_destroy(gpk_batch* h);
int gpk_batch_upload(gpk_batch* h, int field, const void* host, size_t bytes);
int gpk_batch_download(gpk_batch* h, int field, void* host, size_t bytes);
/* nsteps steps of every instance (dt < 0: each instance's stable dt; dt_cap
 * caps it); a step is one replay of a captured hipGraph. */
int gpk_batch_step(gpk_batch* h, int nsteps, float dt, float dt_cap);
int gpk_batch_stats_get(gpk_batch* h, gpk_batch_stats* stats);
/* Per instance (arrays of n_inst, any may be NULL): simulation time, device
 * duration of its last step in ms, band size of its last step. */
int gpk_batch_instances(gpk_batch* h, double* time, double* last_step_ms, int* band_cells);
/* Autotuner: times two steps for each workgroup size (256, 512, 1024) and
 * workgroup count (1, 2, 4, 8 x 256-thread slots per CU) on a copy of the
 * state, restores the state, keeps the fastest shape and returns it. */
int gpk_batch_autotune(gpk_batch* h, int* workgroup_size, int* workgroups, double* step_ms);
int gpk_batch_set_launch(gpk_batch* h, int workgroup_size, int workgroups);

#ifdef __cplusplus
}
#endif
#endif /* GPK_C_H */

/*
 * dgswe — MI355X (gfx950) implementation of the 2D nodal-DG shallow-water
 * `WetDry` time step of EmilLindfors/dgcom, behind a C ABI.
 *
 * The entry points replace the reference's device boundary for this path:
 *   - `PhysicsModule<SWESolution2D>`            src/physics/traits.rs:61-163
 *   - `LocalTimeStepping<SWESolution2D>`        src/time/multirate.rs:296-391
 *   - `Integrable` (the device state)           src/time/integrator.rs:54-106
 *   - `CaseDevice::run_2d` and its physics-as-data input `SWEDeviceForm2D`
 *                                               src/case/run.rs:179-191, src/physics/device.rs:27-143
 * Conventions (the reference's, src/physics/device.rs:79-143, src/case/run.rs:528):
 * every call returns 0 on success or a nonzero status whose reason is
 * `dg_last_error()` (the reference's `Result<_, String>`); no exception
 * crosses the ABI; calls on one context are issued by one host thread in
 * order and run on one HIP stream. Numerical failure is not a call error:
 * `dg_is_finite` and `dg_clip_count` report it (src/simulation/runner.rs:441-449,
 * src/physics/builder.rs:199-206).
 *
 * Layout: SoA FP64, node i of element k at k*NN + i, NN = (N+1)^2, node
 * i = j*(N+1) + a with r fastest (src/solver/state/system_solution.rs:135-147).
 */
#ifndef DGSWE_H
#define DGSWE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dg_ctx dg_ctx;
typedef struct dg_state dg_state;

/* Boundary kinds of a boundary face (face_tags). */
enum { DG_BC_WALL = 0, DG_BC_TIDE = 1 };

/* Levels of local time stepping below the coarsest that the device takes
 * (gpu/src/swe2d.rs:264 MAX_DEVICE_LEVELS). */
#define DG_MAX_LEVELS 8

/* Flags of dg_form.flags */
enum { DG_FRICTION = 1, DG_WET_DRY = 2, DG_SUBCELLS = 4,
       /* elevation-only tide data as a wave entering from still water (ElevationOnly::IncomingWave,
        * src/boundary/characteristic.rs:49-53,424-426); without it, at rest (ElevationOnly::AtRest) */
       DG_INCOMING_WAVE = 8,
       /* SWEFormulation2D::EntropyStable faces: the EC flux with entropy-stable dissipation and
        * the bed's interface term (swe_2d_split_form.rs:611-651, swe_2d_entropy.rs:166-203) instead
        * of the hydrostatic HLL; no subcells (DG_SUBCELLS must be off); no VJP */
       DG_ENTROPY_STABLE = 16 };

/* Schemes (src/time/integrator.rs:226-263) */
enum { DG_SSP_RK3 = 0, DG_SSP_RK43 = 1 };

/*
 * The physics as data: replaces `SWEDeviceForm2D` (src/physics/device.rs:27-71).
 * face_links uses the reference's 4-byte `FaceLink` (src/mesh/core/mesh2d.rs:54-107):
 * element << 4 | shared << 3 | left << 2 | face, 0xFFFFFFFF at a boundary, whose
 * kind is face_tags (the boundary tag's condition, src/case/run.rs:653-737,892).
 */
typedef struct dg_form {
    int order;                    /* N in 1..3 */
    int64_t n_elements;           /* K */
    const double* corners;        /* [K][4][2], counter-clockwise from (r,s) = (-1,-1) */
    const uint32_t* face_links;   /* [K][4] */
    const uint8_t* face_tags;     /* [K][4] */
    const double* bed;            /* [K*NN] nodal bed elevation B */
    const double* manning;        /* [K] per-element Manning n, or NULL: manning_uniform */
    double manning_uniform;
    double coriolis_f0;           /* f-plane */
    double g, h_min, h_dry;       /* 9.81, 1e-6, 1e-3 */
    double relaxation_time;       /* thin-layer tau, 1 s */
    double max_velocity;          /* 20 m/s */
    int flags;                    /* DG_FRICTION | DG_WET_DRY | DG_SUBCELLS | DG_INCOMING_WAVE */
} dg_form;

/* Uniform harmonic tide of the open boundary (src/boundary/harmonic_tide.rs:24-136). */
typedef struct dg_tide {
    double mean, ramp;            /* ramp <= 0: none */
    int n;                        /* constituents, at most 8 */
    double amplitude[8], omega[8], phase[8];
} dg_tide;

/* One element stage: out = a*base + b*input + c*L(input, time), then damping over c
 * and the wet/dry post-process (src/time/multirate.rs:137-200 ElementStage). */
typedef struct dg_stage_desc {
    double a, b, c, time;
} dg_stage_desc;

const char* dg_last_error(void);

/* Context on HIP device `device` (replaces the reference's `Gpu` + `GpuSWE2D`). */
int dg_create(int device, dg_ctx** ctx);
int dg_destroy(dg_ctx* ctx);

/* Upload the statics (GpuSWE2D::from_form, gpu/src/swe2d.rs:3360). */
int dg_upload_form(dg_ctx* ctx, const dg_form* form);
int dg_set_tide(dg_ctx* ctx, const dg_tide* tide);
/* The tide's nodal corrections through a run (HarmonicTide::with_nodal_corrections,
 * src/boundary/harmonic_tide.rs:80-87, from ModelClock::nodal_drift, src/time/clock.rs:108-117):
 * per constituent of the last dg_set_tide, f = f0 + f_rate t and the phase offset
 * (v0 + u0 + u_rate t) degrees, t in seconds since the epoch. dg_set_tide resets them. */
int dg_set_tide_nodal(dg_ctx* ctx, const double* f0, const double* f_rate, const double* v0_deg,
                      const double* u0_deg, const double* u_rate_deg);

/* The open boundary's tides per boundary face node (BoundaryTides, src/boundary/tidal_atlas.rs:858-1140; the
 * reference device's moment-and-sums form, notes/gpu.md:191): n_c (<= 21) constituents with the angular speed
 * omega' (rad/s, the nodal phase drift included) and the nodal factor's drift f_rate (1/s); per boundary face
 * node [K][4][N+1] its slot (UINT32_MAX: the uniform tide of dg_set_tide); per slot coef[n_slots][n_c][6]
 * (C, S of eta, u, v), mean[n_slots] (added unramped; zeros for none) and source_depth[n_slots]; the ramp (s,
 * <= 0: none); velocity: the tides carry u, v through the characteristic boundary (else elevation only, as
 * the form's flags say); max_ratio > 0: the velocity scaled to carry the source's transport, limited to that
 * factor either way (with_transport_scaling). n_c = 0 clears them. The ensembles and the VJP refuse them; under local
 * time stepping each element takes them at its level's stage time. */
int dg_set_boundary_tides(dg_ctx* ctx, int n_c, const double* omega, const double* f_rate, double ramp, int velocity,
                          double max_ratio, int64_t n_slots, const uint32_t* slot_of_face_node, const double* coef,
                          const double* mean, const double* source_depth);

/* The nesting band (NestingRelaxation2D towards an OceanModelState parent, src/boundary/ocean_nesting.rs:
 * 851-1060; the reference device's moment-and-sums form, notes/gpu.md:199-201): every node's slot
 * slot_of_node[K*NN] (UINT32_MAX: not relaxed); the parent's snapshots at times[n_times] (Unix seconds,
 * increasing; model time t is epoch + t, ModelClock::unix, time/clock.rs:67), interpolated linearly or
 * (cubic) by Lagrange over the four nearest (TimeStencil::new, io/field_series.rs:41); per slot series[n_slots][n_times][3] = (zeta, u, v) (f32), the band weight (1 on the boundary,
 * 0 outside the band) and the parent's still-water depth (NaN: unknown); relaxation 1/timescale, nothing
 * where either depth is below wet_depth (0.05 m); momentum_only; velocity: the parent has a depth-mean
 * velocity (else the node's own); max_ratio > 0: that velocity scaled to the parent's transport; the datum
 * shift reference_level; the ramp (s, <= 0: none). A stage time outside the snapshots is an error (a
 * single snapshot covers every time). n_slots = 0 clears it. The ensembles and the VJP refuse it; under
 * local time stepping each element takes it at its level's stage time. */
int dg_set_nesting(dg_ctx* ctx, int64_t n_slots, const uint32_t* slot_of_node, int n_times, const double* times,
                   double epoch, int cubic, const float* series, const double* weight, const double* parent_depth, double timescale,
                   double wet_depth, int momentum_only, int velocity, double max_ratio, double reference_level,
                   double ramp);

/* on != 0: the open boundary's external state is the nesting parent's at each face node's slot (the
 * characteristic boundary over OceanModelState, DeviceExternal2D::Parent, src/boundary/ocean_nesting.rs:823-831,
 * characteristic.rs:135-151): its level and, if it has one, its transport-scaled velocity, raised by the inverse
 * barometer if that is set; face nodes without a slot keep the tides. Acts while dg_set_nesting holds a parent. */
int dg_set_nesting_boundary(dg_ctx* ctx, int on);

/* Gridded weather (GriddedAtmosphere2D's wind stress and pressure gradient, src/source/swe_2d/
 * gridded_atmosphere.rs:137-196, 564-566, 689-731; the reference device's regridded-snapshot form,
 * DeviceSource2D::Atmosphere): per snapshot at times[n_times] (Unix seconds, increasing; model time t is
 * epoch + t; cubic: Lagrange over the four nearest) the fields at every node, f32, field-major
 * series[n_times][5][K*NN] = u, v (m/s, mesh axes), dp/dx, dp/dy (Pa/m), p - p_ref (Pa, unused here);
 * the switches wind and pressure, ramped together (ramp, s; <= 0: none); the drag coefficient
 * (DragCoefficient, wind.rs:134-158: 0 constant cd, 1 Large & Pond, 2 Wu, 3 Smith, 4 Yelland & Taylor),
 * rho_air, rho_water and the depth h_min (1 cm) below which a node is not forced. The momentum source
 * (tau - h grad p) / rho_water, tau = rho_air C_d(|U|) |U| U, is added after Coriolis and the nesting
 * band. A stage time outside the snapshots is an error. n_times = 0 clears it. The ensembles and the VJP
 * refuse it; under local time stepping each element takes it at its level's stage time. */
int dg_set_weather(dg_ctx* ctx, int n_times, const double* times, double epoch, int cubic, const float* series,
                   int wind, int pressure, int drag, double cd, double rho_air, double rho_water, double h_min,
                   double ramp);

/* The open boundary's external level raised by factor times the weather's inverse-barometer level at each
 * face node, -ramp (p - p_ref) / (rho_water 9.81) (InverseBarometer::with_factor over GriddedAtmosphere2D's
 * BoundaryLevel, src/boundary/characteristic.rs:277-345, gridded_atmosphere.rs:198-200, 873-890; the device's
 * `raise`, gpu/src/swe2d.rs:900-909): on the uniform tide and on the tides per node alike. 0 (the default):
 * not raised. Acts while dg_set_weather holds weather with its pressure switched on. */
int dg_set_inverse_barometer(dg_ctx* ctx, double factor);

/* Device states (DeviceState::upload / download, gpu/src/swe2d.rs:2384-2396). */
int dg_state_alloc(dg_ctx* ctx, dg_state** state);
int dg_state_free(dg_ctx* ctx, dg_state* state);
int dg_state_upload(dg_ctx* ctx, dg_state* state, const double* h, const double* hu, const double* hv);
int dg_state_download(dg_ctx* ctx, const dg_state* state, double* h, double* hu, double* hv);
int dg_state_copy(dg_ctx* ctx, dg_state* dst, const dg_state* src);

/* PhysicsModule::compute_rhs_into (src/physics/traits.rs:61-163): out = L(in, t). */
int dg_rhs(dg_ctx* ctx, const dg_state* in, double t, dg_state* out);

/* F* at every face node [K][4][N+1][3] (outward, each element's own side, face
 * nodes in the reference's face order), as the stage kernel evaluates it: the
 * face fluxes that compute_rhs_swe_2d_mass_fluxes_into reports
 * (src/solver/rhs/swe_2d.rs:1005). */
int dg_face_fluxes(dg_ctx* ctx, const dg_state* in, double t, double* out);

/* LocalTimeStepping::stage_where over all elements (src/physics/builder.rs:652-737),
 * with the post-process of finish_where fused into every stage (one level). */
int dg_stage(dg_ctx* ctx, const dg_state* base, const dg_state* input, dg_state* out, const dg_stage_desc* stage);

/* ------------------------------------------------------------------------
 * Local time stepping: the element-subset operations of LocalTimeStepping
 * (src/time/multirate.rs:296-391) as SWEPhysics2D implements them
 * (src/physics/builder.rs:546-783) and the reference GPU takes them
 * (gpu/src/swe2d.rs:4503-4660). The stepper that drives them (levels, stage
 * rates, the passes of a coarse step: MultirateStepper, multirate.rs:713-1330)
 * is the caller's; at most DG_MAX_LEVELS levels below the coarsest.
 * ------------------------------------------------------------------------ */

/* One stage pass of a coarse step (StagePlan, multirate.rs:203-243): the coarse
 * step (t, dt), its finest level, the finest substep and its depth, the stage
 * (0-based) with its Shu-Osher weights (ShuOsherStage: start, input, beta, c)
 * and whether it is the scheme's last stage. Element k of the pass takes
 * element_stage (multirate.rs:169-200) of its level and, for the last stage,
 * its rate. */
typedef struct dg_stage_plan {
    double t, dt;
    int finest, substep, depth, stage;
    double start, input, beta, c;
    int last;
} dg_stage_plan;

/* LocalTimeStepping::begin_coarse_step (multirate.rs:356-362): every element's
 * level level[K] and stage rates rates[n_stages][K] (row-major, rates >= level)
 * of the coarse step about to be taken; the passes' element lists are made here
 * (and kept while the levels and rates are the last ones). */
int dg_begin_coarse_step(dg_ctx* ctx, const uint8_t* level, const uint8_t* rates, int n_stages);

/* LocalTimeStepping::stage_where (multirate.rs:364-382; builder.rs:652-737): for
 * every element with rates[plan->stage][k] >= plan->depth, its RHS from `input` at
 * its own stage time, out = start*base + input*input + c*F, the damping over c,
 * then either acc <- w*out + keep*acc (plan->last; `out` is then not written) or
 * the wet/dry correction into out. Asynchronous on the context's stream. With a
 * forcing per node (boundary tides, nesting, weather) its per-level moments go up
 * through a ring of pinned host slots, so such a stage refuses to be captured into
 * a graph (an error, by name, when the stream is capturing). */
int dg_stage_where(dg_ctx* ctx, const dg_state* base, const dg_state* input, const dg_stage_plan* plan,
                   dg_state* out, dg_state* acc);

/* LocalTimeStepping::finish_where (multirate.rs:384-387; builder.rs:739-782): for
 * every element with level >= depth, state <- acc and the wet/dry correction.
 * Asynchronous on the context's stream. */
int dg_finish_where(dg_ctx* ctx, dg_state* state, const dg_state* acc, int depth);

/* Bit j of out[k] (host, K bytes): every node of element k deeper than d_j, for
 * LocalTimeStepping::line_local and coarse_step_bound (builder.rs:569-637; the
 * reference GPU's depth_flags, gpu/src/swe2d.rs:2030). */
int dg_depth_flags(dg_ctx* ctx, const dg_state* state, double d0, double d1, double d2, uint8_t* out);

/* The multirate stepper in the library (MultirateStepper, multirate.rs:713-1330,
 * with the face stencil, node-by-node rates over SWEPhysics2D's line-local
 * elements and its coarse-step bound, builder.rs:563-637): the caller's loop is
 * assign_levels then step per coarse step, as Simulation::run_with_callback's
 * (runner.rs:298-455); dg_multirate_run is that loop without callbacks. */
typedef struct dg_multirate dg_multirate;
typedef struct dg_multirate_stats {
    int64_t steps, element_evaluations, global_evaluations;
    int finest_level;
    double model_seconds;
} dg_multirate_stats;   /* MultirateStats (multirate.rs:856-889) */

/* At most max_levels (<= DG_MAX_LEVELS) levels below the coarsest; line_rates:
 * spread the rates of line-local elements node by node (the reference's default). */
int dg_multirate_create(dg_ctx* ctx, int max_levels, int scheme, int line_rates, dg_multirate** out);
int dg_multirate_destroy(dg_multirate* mr);
/* MultirateStepper::assign_levels (multirate.rs:1014-1043): levels from the
 * state's element time steps at `cfl`; the coarse step at most dt_cap (<= 0: none)
 * and the module's bound, fitted to `landing` (<= 0: none) into *dt; *stable the
 * step the elements allow (stable_step). */
int dg_multirate_assign_levels(dg_multirate* mr, const dg_state* state, double cfl, double dt_cap, double landing,
                               double* dt, double* stable);
/* MultirateStepper::step (multirate.rs:1298-1330): one coarse step of dt from t. */
int dg_multirate_step(dg_multirate* mr, dg_state* state, double t, double dt);
/* The levels [K], the stage rates [n_stages][K] and the finest level of the last
 * assignment (any pointer may be NULL). */
int dg_multirate_levels(dg_multirate* mr, uint8_t* level, uint8_t* rates, int* finest);
int dg_multirate_stats_read(dg_multirate* mr, dg_multirate_stats* out);
/* Coarse steps from t to t_end, each landing on the next multiple of `interval`
 * after t (<= 0: t_end only), dt_max (<= 0: none); *t_out, *coarse_steps. */
int dg_multirate_run(dg_multirate* mr, dg_state* state, double t, double t_end, double cfl, double dt_max,
                     double interval, double* t_out, int64_t* coarse_steps);

/* FNV-1a (64 bit) of n bytes continuing from h (0xcbf29ce484222325 to start):
 * the checksum and domain fingerprint of the reference's restart files
 * (src/io/restart.rs:90-91,660-697), for the restart writer above the ABI. */
uint64_t dg_fnv1a(const uint8_t* data, size_t n, uint64_t h);

/* LocalTimeStepping::element_dt per element (src/solver/rhs/swe_2d.rs:2018-2115) into
 * the host array out[K], and PhysicsModule::compute_dt_ssp (builder.rs:464-485). */
int dg_element_dt(dg_ctx* ctx, const dg_state* state, double cfl, int scheme, double* out);
int dg_dt_ssp(dg_ctx* ctx, const dg_state* state, double cfl, int scheme, double* dt);

/* One SSP step of dt from t in place (MultirateStepper::step, one level,
 * src/time/multirate.rs:1285-1311). */
int dg_step(dg_ctx* ctx, dg_state* state, double t, double dt, int scheme);

/* n_steps adaptive SSP steps from t (the dt loop of Simulation::run_with_callback,
 * src/simulation/runner.rs:264-483, without output landing); t_end receives the
 * end time, dts (optional, n_steps) each step's dt. */
int dg_run(dg_ctx* ctx, dg_state* state, double t, int n_steps, double cfl, int scheme, double* t_end, double* dts);

/* The stage's launch shape (no counterpart in the reference, which launches one
 * thread per element): dg_tune times every shape (the wet kernel's block mapping,
 * XCD-aware or linear; the subcell kernel's grid, 2048, 16384 or 65536
 * workgroups) on the state given, three stages each, and keeps the fastest for
 * the context and for its (device, order, size class floor(log2 K)): a form
 * uploaded later in that class starts from it. The shapes change no result; the
 * clip count is kept. dg_launch_shape reports the context's shape (by default
 * XCD-aware, 16384 subcell workgroups). */
int dg_tune(dg_ctx* ctx, const dg_state* state);
int dg_launch_shape(dg_ctx* ctx, int* xcd_mapping, int* subcell_groups);

/* Integrable::is_finite (integrator.rs:54-106) and the negative-depth clips. */
int dg_is_finite(dg_ctx* ctx, const dg_state* state, int* finite);
int dg_clip_count(dg_ctx* ctx, uint64_t* clips);

/* Integer wet/dry masks: node_dry[K*NN] = h < h_dry, elem_subcell[K] = the element
 * takes the subcell volume term (swe_2d_split_form.rs:711-713). */
int dg_masks(dg_ctx* ctx, const dg_state* state, uint8_t* node_dry, uint8_t* elem_subcell);

/* dg_masks without allocating, for callers that cannot call hipMalloc (tight serving loops, harnesses
 * that forbid allocation): the same masks, staged through a fixed area of DG_NOALLOC_MAX_BYTES in the
 * library's static device memory, one per GPU, held until the process ends. No counterpart in the
 * reference. K*(NN+1) bytes at most (the call is refused beyond); calls that overlap on one GPU (from
 * several threads or contexts) share that area and race. */
#define DG_NOALLOC_MAX_BYTES (32u << 20)
int dg_masks_noalloc(dg_ctx* ctx, const dg_state* state, uint8_t* node_dry, uint8_t* elem_subcell);

/* Measurement: per-launch HIP-event timing of the stage kernel on the context's
 * stream. dg_timing_read returns the summed kernel milliseconds and launches since
 * the last reset. dg_copy_bench times a plain device copy of `bytes` (reps times). */
int dg_timing_enable(dg_ctx* ctx, int on);
int dg_timing_read(dg_ctx* ctx, double* stage_ms, int64_t* stage_launches, double* dt_ms, int64_t* dt_launches);
int dg_sync(dg_ctx* ctx);
int dg_copy_bench(dg_ctx* ctx, size_t bytes, int reps, double* ms_per_copy);

/* ------------------------------------------------------------------------
 * Discrete adjoint. The reference has no adjoint of this path: these calls are
 * the vector-Jacobian products of the calls above, for a torch
 * autograd.Function (forward = dg_stage / dg_step, backward = the VJP) and for
 * gradients of a station misfit with respect to the nodal bed and per-element
 * Manning n. They are exact for the arithmetic as written: at a kink (a
 * max/min tie, a branch of the HLL speeds, the limiter, the wet/dry switches)
 * the derivative is that of the branch the forward takes, and time steps are
 * held at the values given. Cotangents are device states; the parameter
 * cotangents accumulate in the context until dg_param_grads_zero.
 * ------------------------------------------------------------------------ */
typedef struct dg_tape dg_tape;

/* state = 0 */
int dg_state_zero(dg_ctx* ctx, dg_state* state);
/* A state over caller-owned device memory of 3*K*NN doubles (h, hu, hv), for
 * tensors of a framework; dg_state_free releases the handle, not the memory. */
int dg_state_wrap(dg_ctx* ctx, double* device_ptr, dg_state** state);
/* Run the context's work on the caller's HIP stream (NULL: its own stream). */
int dg_set_stream(dg_ctx* ctx, void* hip_stream);
/* Replace the bed [K*NN] or the per-element Manning n [K] (host or device memory). */
int dg_set_bed(dg_ctx* ctx, const double* bed, int on_device);
int dg_set_manning(dg_ctx* ctx, const double* manning, int on_device);
/* Manning n per node [K*NN] (SpatiallyVaryingManning2D::from_nodal, src/source/swe_2d/friction.rs:463-553: the
 * point-implicit damping with the node's g n^2) instead of the per-element n; NULL: back to per element. The
 * ensembles and the VJP refuse it. */
int dg_set_manning_nodal(dg_ctx* ctx, const double* manning, int on_device);
/* A beta-plane f = f0 + beta (y - y0), y of each node by the element's bilinear map (CoriolisSource2D::beta_plane
 * and add_element, src/source/swe_2d/coriolis.rs:101-121,157-169), f0 the form's coriolis_f0; beta = 0: the
 * f-plane again. The ensembles and the VJP refuse it. */
int dg_set_coriolis_beta(dg_ctx* ctx, double beta, double y0);
/* The accumulated cotangents of the bed [K*NN] and of Manning n [K]. */
int dg_param_grads_zero(dg_ctx* ctx);
int dg_param_grads_read(dg_ctx* ctx, double* bed_grad, double* manning_grad, int on_device);

/* VJP of dg_rhs: cot_in += (dL/din)^T cot_out; the bed's part accumulates. */
int dg_rhs_vjp(dg_ctx* ctx, const dg_state* in, double t, const dg_state* cot_out, dg_state* cot_in);
/* VJP of dg_stage: cot_base += (dout/dbase)^T cot_out (cot_base may be NULL),
 * cot_input += (dout/dinput)^T cot_out; cot_base may be cot_input (a stage
 * whose base is its input); neither may be cot_out. Bed and Manning parts
 * accumulate in the context. */
int dg_stage_vjp(dg_ctx* ctx, const dg_state* base, const dg_state* input, const dg_stage_desc* stage,
                 const dg_state* cot_out, dg_state* cot_base, dg_state* cot_input);
/* VJP of dg_step (the stage values are recomputed from state): cot_state =
 * (dX_{n+1}/dX_n)^T cot_next, overwritten; cot_next may not be cot_state. */
int dg_step_vjp(dg_ctx* ctx, const dg_state* state, double t, double dt, int scheme, const dg_state* cot_next,
                dg_state* cot_state);

/* Stations (Probe2D, src/solver/probe.rs:1-30): the water level
 * eta = sum_i w_i (h_i + B_i) over the nodes of a station's element, with w the
 * element polynomial's Lagrange weights at the station point. */
int dg_set_stations(dg_ctx* ctx, int n, const int64_t* element, const double* weights /* [n][NN] */);
int dg_sample_stations(dg_ctx* ctx, const dg_state* state, double* eta /* host [n] */);

/* Station misfit J = sum over samples m and stations s of (eta_ms - obs_ms)^2,
 * sampled after every `every`-th step of an n_steps run at the given dts, and its
 * gradient by reverse sweeps over binomial checkpoints: the tape holds
 * `snapshots` states (the initial one included). dg_misfit_forward runs the
 * forward, keeping the checkpoints of the schedule; dg_misfit_backward zeroes the
 * parameter cotangents and fills them with scale * dJ/d(bed, n) (read them with
 * dg_param_grads_read); cot_x0 (optional) receives scale * dJ/dX_0. */
int dg_tape_create(dg_ctx* ctx, int snapshots, dg_tape** tape);
int dg_tape_destroy(dg_ctx* ctx, dg_tape* tape);
int dg_misfit_forward(dg_ctx* ctx, dg_tape* tape, const dg_state* x0, double t0, int n_steps, const double* dts,
                      int scheme, int every, const double* obs /* [n_steps/every][stations] */, double* misfit);
int dg_misfit_backward(dg_ctx* ctx, dg_tape* tape, double scale, dg_state* cot_x0);
/* Step records (no counterpart in the reference, which has no adjoint):
 * with on != 0 the tape keeps step records — a step's stage values and stage
 * RHS, 5 states for SSP-RK3 and 7 for SSP-RK(4,3) — so that no reverse step
 * recomputes its stages: snapshots + 1 of them at most (one per step), made by
 * the forward sweep (the step before each checkpoint it keeps, and its last
 * step) and by the last step of each advance of the schedule, each taken by the
 * reverse of its step. The same gradient bit for bit; the records are allocated
 * by dg_misfit_forward, and on = 0 frees them. dg_tape_records_used: the reverse
 * steps of the last backward that took their step's record. */
int dg_tape_keep_stages(dg_ctx* ctx, dg_tape* tape, int on);
int dg_tape_records_used(const dg_tape* tape, int64_t* used);
/* Forward steps and step VJPs the last forward/backward pair took, and the
 * forward steps a schedule takes (host only). */
int dg_tape_stats(const dg_tape* tape, int64_t* forward_steps, int64_t* step_vjps);
int dg_revolve_count(int n_steps, int snapshots, int64_t* forward_steps);

/* ------------------------------------------------------------------------
 * Ensembles: independent runs of different meshes, orders and forcings,
 * advanced together (the reference runs one case per process; no counterpart).
 * Each member is a context with its form and tide uploaded; each keeps its own
 * adaptive time step and time. A step of the ensemble is one launch per order
 * per kernel over all members' element blocks and one host synchronisation for
 * all the members' time steps; DG_BATCH_GRAPH replays the stages from a HIP
 * graph captured on the first step of each parity, DG_BATCH_TUNE times the
 * launch shapes per order on the first step and keeps the fastest. Every
 * member's run equals, bit for bit, its dg_run.
 * ------------------------------------------------------------------------ */
typedef struct dg_batch dg_batch;
enum { DG_BATCH_GRAPH = 1, DG_BATCH_TUNE = 2 };

int dg_batch_create(int n_members, dg_ctx* const* members, int scheme, dg_batch** batch);
int dg_batch_destroy(dg_batch* batch);
int dg_batch_upload(dg_batch* batch, int member, const double* h, const double* hu, const double* hv);
int dg_batch_download(dg_batch* batch, int member, double* h, double* hu, double* hv);
/* n_steps adaptive SSP steps of every member; t[member] is each member's time (in/out). */
int dg_batch_run(dg_batch* batch, int n_steps, double cfl, int options, double* t);
/* The launch shape the tuner chose for an order: XCD-aware block mapping (0/1), subcell grid. */
int dg_batch_tuning(const dg_batch* batch, int order, int* xcd_mapping, int* subcell_groups);

#ifdef __cplusplus
}
#endif

#endif /* DGSWE_H */

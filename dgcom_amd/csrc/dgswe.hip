// HIP kernels for gfx950 and the C ABI of include/dgswe.h: the fused SSP stage
// of the 2D shallow-water `WetDry` split form (volume, faces, sources,
// Shu-Osher combination, point-implicit damping and the wet/dry correction in
// one pass), the per-element time-step bound with its global minimum, and
// the state utilities.
//
// Work mapping: one 64-lane wavefront is one workgroup and holds
// EPB = 64 / (N+1)^2 elements, one lane per GLL node (N = 1: 16 elements,
// N = 2: 7, N = 3: 4). Loads and stores of the SoA state are contiguous over
// the wavefront; an element's node states are exchanged through LDS for the
// line (sum-factored) volume terms, and its 4(N+1) face nodes are spread over
// its lanes for the Riemann fluxes, each evaluated from the side that is the
// left side of the edge so that both neighbours get identical F* bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/dgswe.h"
#include "swe_kernels.hpp"
#include "swe_batch.hpp"
#include "multirate.hpp"

using namespace dgswe;

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

thread_local std::string g_err;

struct DeviceArray {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

// Workgroups of the subcell kernel (a grid-stride loop over its list)
constexpr unsigned int SUB_GROUPS = 16384;

// The fixed staging area of dg_masks_noalloc, one copy per GPU for the life of
// the process. Host code never passes `scratch` itself to a kernel or a copy:
// it is a host-side stand-in; its device address on each GPU is looked up once
// (scratch_address) and kept in scratch_ptr.
constexpr size_t SCRATCH_BYTES = (size_t)DG_NOALLOC_MAX_BYTES;
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
void* scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on `device`, which must be current.
hipError_t scratch_address(int device, void** out) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[device] = ptr;
    }
    *out = scratch_ptr[device];
    return hipSuccess;
}
}  // namespace

struct dg_state {
    double* d = nullptr;  // 3*K*NN: h, hu, hv
    bool owned = true;    // false: caller memory (dg_state_wrap)
};

// The stage's launch shape, per (device, order, size class): the wet kernel's
// block mapping (XCD-aware or linear) and the subcell kernel's grid
struct LaunchShape {
    unsigned int xcd = 1, sub_groups = SUB_GROUPS;
};

struct dg_ctx {
    int device = 0;
    LaunchShape shape;
    hipStream_t stream = nullptr;      // where the work runs (the context's own, or the caller's)
    hipStream_t own_stream = nullptr;
    int order = 0, nn = 0;
    int64_t K = 0;
    int flags = 0;
    double g = 9.81, h_min = 1e-6, h_dry = 1e-3, tau = 1.0, u_max = 20.0, f0 = 0.0;
    dg_tide tide{};
    // nodal corrections of the tide's constituents, linear in the time since the
    // epoch: f0, f_rate, V0 (deg), u0 (deg), u_rate (deg/s); identity by default
    double tide_nodal[5][8] = {{1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0}, {}, {}, {}, {}};
    double *corners = nullptr, *bed = nullptr, *manning = nullptr, *element_dt = nullptr;
    // static geometry cache (geometry_kernel): geo per node; fgeo per face [3][K*4], then per face node
    // [K][4][N+1] the outer node's distance
    double *geo = nullptr, *fgeo = nullptr;
    double* emass = nullptr;  // [K]: the positivity bound's static mass ratio (geometry_kernel)
    uint32_t* links = nullptr;
    uint8_t *tags = nullptr, *affine = nullptr;
    unsigned long long* clips = nullptr;
    uint32_t* queue = nullptr;
    uint8_t* queue_count = nullptr;
    double* dt_host = nullptr;   // pinned, mapped: the time step, written by slots_min_kernel itself
    double* dt_value = nullptr;  // its device address
    unsigned long long* dt_slots = nullptr;  // [DT_SLOTS]: the time step's running minima (dt_kernel)
    uint32_t* sub_list = nullptr;
    unsigned int *chunk_total = nullptr, *sub_count = nullptr;
    int* flag = nullptr;
    dg_state* work[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    uint8_t* scratch = nullptr;    // this GPU's copy of the static staging area (scratch_address)
    double stage_ms = 0.0, dt_ms = 0.0;
    int64_t stage_n = 0, dt_n = 0;
    // discrete adjoint
    double *gbed = nullptr, *gman = nullptr;          // parameter cotangents [K*NN], [K]
    double *facebuf = nullptr, *innerbuf = nullptr;  // [K][4(N+1)][4] neighbour cotangents
    double* facecot = nullptr;                       // [K][4(N+1)][3] F* cotangents
    uint32_t *vjp_edges = nullptr, *vjp_split = nullptr;  // the edge kernels' faces (vjp_edge_lists)
    uint32_t n_vjp_edges = 0, n_vjp_split = 0;
    uint8_t* kind = nullptr;                         // [K] subcell elements of the last VJP
    dg_state* adj[2] = {nullptr, nullptr};           // stage cotangents of dg_step_vjp
    dg_state* rhsbuf[4] = {nullptr, nullptr, nullptr, nullptr};  // stage RHS saved by the forward for the VJP
    // local time stepping: the coarse step's levels and last-stage rates [K], the
    // passes' element orders [n_stages + 1][K] (stage rates, then the level;
    // largest key first, element order within a key) and, on the host, how
    // many have a key >= d
    uint8_t *lts_level = nullptr, *lts_lrate = nullptr;
    uint32_t* lts_order = nullptr;
    int lts_stages = 0, lts_finest = -1;
    uint64_t lts_generation = 0;    // bumped whenever the lists are rebuilt
    std::vector<uint8_t> lts_keys;  // the host copy of the last levels and rates
    std::vector<int64_t> lts_count;  // [n_stages + 1][DG_MAX_LEVELS + 1]
    uint8_t* depth_flags = nullptr;
    // boundary tides per face node (dg_set_boundary_tides): device tables and the host's constituents
    uint32_t* tide_slot = nullptr;
    double *tide_coef = nullptr, *tide_mean = nullptr, *tide_src = nullptr;
    int atlas_nc = 0, atlas_velocity = 0;
    double atlas_omega[ATLAS_MAX_CONSTITUENTS] = {}, atlas_f_rate[ATLAS_MAX_CONSTITUENTS] = {};
    double atlas_ramp = 0.0, atlas_max_ratio = 0.0;
    // the nesting band (dg_set_nesting): device tables and the snapshot times on the host
    uint32_t* nest_slot = nullptr;
    float* nest_series = nullptr;
    double *nest_weight = nullptr, *nest_depth = nullptr;
    std::vector<double> nest_times;
    int nest_velocity = 0, nest_momentum_only = 0, nest_cubic = 0, nest_boundary = 0;
    std::vector<uint32_t> h_nest_slot;  // the host copy of the slot of every node (the boundary's coverage check)
    double nest_epoch = 0.0;
    // the gridded weather (dg_set_weather): the snapshots on the device, their times on the host
    float* wx_series = nullptr;
    std::vector<double> wx_times;
    double wx_epoch = 0.0, wx_ramp = 0.0, wx_cd = 0.0, wx_rho_air = 1.225, wx_rho_water = 1025.0, wx_h_min = 0.01;
    int wx_cubic = 0, wx_wind = 0, wx_pressure = 0, wx_drag = 1;
    double wx_raise = 0.0;   // dg_set_inverse_barometer
    // local time stepping with forcing: a ring of per-level moment tables, pinned on the host and on the
    // device; a slot's host copy is written again only after its event (its upload) has completed
    static constexpr int MOMENT_RING = 64;
    double* lvl_moment_dev = nullptr;
    double* lvl_moment_host = nullptr;
    hipEvent_t lvl_moment_ev[MOMENT_RING] = {};
    bool lvl_moment_used[MOMENT_RING] = {};
    int lvl_moment_next = 0;
    double nest_rate = 0.0, nest_wet_depth = 0.05, nest_max_ratio = 0.0, nest_reference = 0.0, nest_ramp = 0.0;
    // Manning n per node and a beta-plane's f per node (dg_set_manning_nodal, dg_set_coriolis_beta), or NULL
    double *manning_node = nullptr, *coriolis_node = nullptr;
    double beta = 0.0, beta_y0 = 0.0;
    // every entry point on the context holds this (the reference guards its device module's scratch with a
    // Mutex, gpu/src/swe2d.rs:3689): calls from several host threads are serialised
    std::recursive_mutex mu;
    // stations
    int n_st = 0;
    int64_t* st_elem = nullptr;
    double *st_w = nullptr, *st_eta = nullptr;
};

#define DG_TRY(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            g_err = std::string(#expr) + ": " + hipGetErrorString(e_);                            \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

// Every entry point on a context holds its mutex for the call (recursive: an entry point may call another)
#define DG_LOCK(ctx)                                                   \
    DG_REQUIRE((ctx) != nullptr, "null context");                      \
    std::lock_guard<std::recursive_mutex> dg_lock_((ctx)->mu)

#define DG_REQUIRE(cond, msg)       \
    do {                            \
        if (!(cond)) {              \
            g_err = (msg);          \
            return 2;               \
        }                           \
    } while (0)

static bool parallelogram(const double* v) {
    // geometric.rs:464-475
    double bx = v[0] - v[2] + v[4] - v[6], by = v[1] - v[3] + v[5] - v[7];
    double scale = 0.0;
    for (int i = 0; i < 4; ++i) {
        int j = (i + 1) % 4;
        scale = std::fmax(scale, std::hypot(v[2 * j] - v[2 * i], v[2 * j + 1] - v[2 * i + 1]));
    }
    return std::hypot(bx, by) <= 1.0e-12 * scale;
}

static double host_bilinear_det(const double* c, double r, double s) {
    double x_r = 0.25 * ((1.0 - s) * (c[2] - c[0]) + (1.0 + s) * (c[4] - c[6]));
    double y_r = 0.25 * ((1.0 - s) * (c[3] - c[1]) + (1.0 + s) * (c[5] - c[7]));
    double x_s = 0.25 * ((1.0 - r) * (c[6] - c[0]) + (1.0 + r) * (c[4] - c[2]));
    double y_s = 0.25 * ((1.0 - r) * (c[7] - c[1]) + (1.0 + r) * (c[5] - c[3]));
    return x_r * y_s - x_s * y_r;
}

// HarmonicTide::elevation (harmonic_tide.rs:124-136) with NodalDrift::at
// (astronomy.rs:252-259) and NodalCorrection::phase_offset_rad (astronomy.rs:213-215):
// eta = mean + R(t) sum (f0 + f' t) A cos(w t + phase + (V0 + u0 + u' t) pi/180);
// with the identity correction this is sum A cos(w t + phase) exactly
static double tide_elevation(const dg_tide& t, const double (&nod)[5][8], double time) {
    constexpr double DEG_TO_RAD = M_PI / 180.0;
    double tide = 0.0;
    for (int i = 0; i < t.n; ++i) {
        const double f = nod[0][i] + nod[1][i] * time;
        const double u = nod[3][i] + nod[4][i] * time;
        tide += f * t.amplitude[i] * std::cos(t.omega[i] * time + t.phase[i] + (nod[2][i] + u) * DEG_TO_RAD);
    }
    double rf = 1.0;
    if (t.ramp > 0.0) {
        double tau = std::min(std::max(time / t.ramp, 0.0), 1.0);
        rf = tau * tau * (3.0 - 2.0 * tau);
    }
    return t.mean + rf * tide;
}

static int free_statics(dg_ctx* c) {
    void* ptrs[] = {c->corners, c->bed, c->manning, c->element_dt, c->links, c->tags, c->affine, c->queue,
                    c->queue_count, c->dt_slots, c->sub_list, c->chunk_total, c->gbed, c->gman, c->facebuf,
                    c->innerbuf, c->facecot, c->vjp_edges, c->vjp_split, c->kind, c->st_elem, c->st_w, c->st_eta, c->geo, c->fgeo,
                    c->lts_level, c->lts_lrate, c->lts_order, c->depth_flags, c->tide_slot, c->tide_coef,
                    c->tide_mean, c->tide_src, c->nest_slot, c->nest_series, c->nest_weight, c->nest_depth,
                    c->wx_series, c->manning_node, c->coriolis_node, c->emass};
    for (void* p : ptrs)
        if (p) DG_TRY(hipFree(p));
    c->gbed = c->gman = c->facebuf = c->innerbuf = c->facecot = nullptr;
    c->vjp_edges = c->vjp_split = nullptr;
    c->n_vjp_edges = c->n_vjp_split = 0;
    c->geo = c->fgeo = c->emass = nullptr;
    c->lts_level = c->lts_lrate = c->depth_flags = nullptr;
    c->tide_slot = nullptr;
    c->tide_coef = c->tide_mean = c->tide_src = nullptr;
    c->manning_node = c->coriolis_node = nullptr;
    c->beta = c->beta_y0 = 0.0;
    c->atlas_nc = 0;
    c->nest_slot = nullptr;
    c->nest_series = nullptr;
    c->nest_weight = c->nest_depth = nullptr;
    c->nest_times.clear();
    c->wx_series = nullptr;
    c->wx_times.clear();
    c->lts_order = nullptr;
    c->lts_stages = 0;
    c->lts_finest = -1;
    c->lts_keys.clear();
    ++c->lts_generation;
    c->lts_count.clear();
    c->kind = nullptr;
    c->st_elem = nullptr;
    c->st_w = c->st_eta = nullptr;
    c->n_st = 0;
    for (auto& w : c->adj) {
        if (w) {
            DG_TRY(hipFree(w->d));
            delete w;
            w = nullptr;
        }
    }
    for (auto& w : c->rhsbuf) {
        if (w) {
            DG_TRY(hipFree(w->d));
            delete w;
            w = nullptr;
        }
    }
    c->corners = c->bed = c->manning = c->element_dt = nullptr;
    c->links = nullptr;
    c->tags = c->affine = nullptr;
    c->queue = nullptr;
    c->queue_count = nullptr;
    c->dt_slots = nullptr;
    c->sub_list = nullptr;
    c->chunk_total = nullptr;
    for (auto& w : c->work) {
        if (w) {
            DG_TRY(hipFree(w->d));
            delete w;
            w = nullptr;
        }
    }
    return 0;
}

// The static geometry cache of the stage kernels (after the form's upload and every bed change)
static int launch_geometry(dg_ctx* c) {
    const int epb = 64 / c->nn;
    dim3 grid((unsigned)((c->K + epb - 1) / epb));
    switch (c->order) {
        case 1: hipLaunchKernelGGL(geometry_kernel<1>, grid, dim3(64), 0, c->stream, c->corners, c->affine, c->links, c->K, c->geo, c->fgeo, c->emass); break;
        case 2: hipLaunchKernelGGL(geometry_kernel<2>, grid, dim3(64), 0, c->stream, c->corners, c->affine, c->links, c->K, c->geo, c->fgeo, c->emass); break;
        default: hipLaunchKernelGGL(geometry_kernel<3>, grid, dim3(64), 0, c->stream, c->corners, c->affine, c->links, c->K, c->geo, c->fgeo, c->emass); break;
    }
    DG_TRY(hipGetLastError());
    return 0;
}

// The boundary tides' tables and their factors at the stage time (BoundaryTides::moment,
// tidal_atlas.rs:995-1008): the ramp, then per constituent f(t)/f0 = 1 + f_rate t and
// cos, sin of omega' t, evaluated here on the host (the device sums the nodes' terms)
static void set_atlas(const dg_ctx* c, StageParams& p, double time) {
    if (!c->tide_slot) return;
    p.tide_slot = c->tide_slot;
    p.tide_coef = c->tide_coef;
    p.tide_mean = c->tide_mean;
    p.tide_src = c->tide_src;
    p.atlas_nc = c->atlas_nc;
    p.atlas_velocity = c->atlas_velocity;
    p.atlas_max_ratio = c->atlas_max_ratio;
    double rf = 1.0;
    if (c->atlas_ramp > 0.0) {
        double tau = std::min(std::max(time / c->atlas_ramp, 0.0), 1.0);
        rf = tau * tau * (3.0 - 2.0 * tau);
    }
    p.atlas_moment[0] = rf;
    for (int j = 0; j < c->atlas_nc; ++j) {
        // the reference's `(omega t).sin_cos()` (tidal_atlas.rs:966-970) is one call of the C
        // library's sincos (LLVM merges the pair), which differs from its separate sin and cos
        // by an ulp for some arguments
        double sn, cs;
        sincos(c->atlas_omega[j] * time, &sn, &cs);
        p.atlas_moment[1 + 3 * j] = 1.0 + c->atlas_f_rate[j] * time;
        p.atlas_moment[2 + 3 * j] = cs;
        p.atlas_moment[3 + 3 * j] = sn;
    }
}

// TimeStencil::new (io/field_series.rs:41-81) at time t over increasing snapshot times: a single snapshot, or
// one hit exactly, alone; else the Lagrange weights of the two bracketing snapshots (cubic: up to one more
// either side), each product taken in the reference's order. False outside the snapshots.
static bool time_stencil(const std::vector<double>& ts, double t, bool cubic, int& len, int* idx, double* w) {
    const size_t n = ts.size();
    len = 1;
    idx[0] = 0;
    w[0] = 1.0;
    if (n == 1) return true;
    if (!(t >= ts[0] && t <= ts[n - 1])) return false;
    size_t i1 = (size_t)(std::lower_bound(ts.begin(), ts.end(), t) - ts.begin());
    i1 = std::min(std::max(i1, (size_t)1), n - 1);
    const size_t i0 = i1 - 1;
    if (t == ts[i0]) { idx[0] = (int)i0; return true; }
    if (t == ts[i1]) { idx[0] = (int)i1; return true; }
    const size_t lo = cubic ? (i0 > 0 ? i0 - 1 : 0) : i0;
    const size_t hi = cubic ? std::min(i1 + 1, n - 1) : i1;
    len = (int)(hi - lo + 1);
    for (size_t a = lo; a <= hi; ++a) {
        double wa = 1.0;
        for (size_t b = lo; b <= hi; ++b)
            if (b != a) wa *= (t - ts[b]) / (ts[a] - ts[b]);
        idx[a - lo] = (int)a;
        w[a - lo] = wa;
    }
    return true;
}

// tidal_ramp (harmonic_tide.rs:24-32): a smoothstep over `duration` seconds, none for duration <= 0
static double ramp_factor(double time, double duration) {
    if (!(duration > 0.0)) return 1.0;
    const double tau = std::min(std::max(time / duration, 0.0), 1.0);
    return tau * tau * (3.0 - 2.0 * tau);
}

// The nesting band's tables and the parent's moment at the stage time (OceanModelState::moment,
// ocean_nesting.rs:673-682): the snapshot stencil at the parent's clock (epoch + time) and the ramp
static int set_nesting(const dg_ctx* c, StageParams& p, double time) {
    if (!c->nest_slot) return 0;
    DG_REQUIRE(time_stencil(c->nest_times, c->nest_epoch + time, c->nest_cubic != 0, p.nest_len, p.nest_idx, p.nest_w),
               "the nesting parent's snapshots do not cover the stage time (its forcing would freeze)");
    const size_t n = c->nest_times.size();
    p.nest_slot = c->nest_slot;
    p.nest_series = c->nest_series;
    p.nest_weight = c->nest_weight;
    p.nest_depth = c->nest_depth;
    p.nest_times = (int)n;
    p.nest_velocity = c->nest_velocity;
    p.nest_momentum_only = c->nest_momentum_only;
    p.nest_rate = c->nest_rate;
    p.nest_wet_depth = c->nest_wet_depth;
    p.nest_max_ratio = c->nest_max_ratio;
    p.nest_reference = c->nest_reference;
    p.nest_ramp = ramp_factor(time, c->nest_ramp);
    p.nest_boundary = c->nest_boundary;
    return 0;
}

// The weather's snapshots and its moment at the stage time (GriddedAtmosphere2D::moment,
// gridded_atmosphere.rs:564-566): the stencil at the clock (epoch + time), the switches times the ramp
static int set_weather(const dg_ctx* c, StageParams& p, double time) {
    if (!c->wx_series) return 0;
    DG_REQUIRE(time_stencil(c->wx_times, c->wx_epoch + time, c->wx_cubic != 0, p.wx_len, p.wx_idx, p.wx_w),
               "the weather's snapshots do not cover the stage time (its forcing would freeze)");
    const double ramp = ramp_factor(time, c->wx_ramp);
    p.wx_series = c->wx_series;
    p.wx_nodes = (int64_t)c->K * c->nn;
    p.wx_drag = c->wx_drag;
    p.wx_wind = c->wx_wind ? ramp : 0.0;
    p.wx_pres = c->wx_pressure ? ramp : 0.0;
    p.wx_cd = c->wx_cd;
    p.wx_rho_air = c->wx_rho_air;
    p.wx_rho_water = c->wx_rho_water;
    p.wx_h_min = c->wx_h_min;
    // GriddedAtmosphere2D's level is 0 without its pressure (gridded_atmosphere.rs:879-881)
    p.wx_raise = c->wx_pressure ? c->wx_raise : 0.0;
    return 0;
}

static int launch_stage(dg_ctx* c, const double* base, const double* input, double* out, double a, double b, double cc,
                        double time, int rhs_only, double* fstar_out = nullptr, double* rhs_save = nullptr) {
    size_t nt = (size_t)c->K * c->nn;
    StageParams p{};
    p.bh = base; p.bhu = base + nt; p.bhv = base + 2 * nt;
    p.ih = input; p.ihu = input + nt; p.ihv = input + 2 * nt;
    p.oh = out; p.ohu = out + nt; p.ohv = out + 2 * nt;
    p.bed = c->bed; p.corners = c->corners; p.manning = c->manning;
    p.links = c->links; p.tags = c->tags; p.affine = c->affine;
    p.n_elements = c->K;
    p.a = a; p.b = b; p.c = cc;
    p.eta_e = tide_elevation(c->tide, c->tide_nodal, time);
    p.g = c->g; p.h_min = c->h_min; p.h_dry = c->h_dry; p.u_max = c->u_max; p.tau = c->tau; p.f0 = c->f0;
    p.sqrt_g = std::sqrt(c->g);
    p.rhs_only = rhs_only;
    p.flags = c->flags;
    p.clips = c->clips;
    p.fstar_out = fstar_out;
    p.geo = c->geo;
    p.fgeo = c->fgeo;
    p.rh = rhs_save;
    p.rhu = rhs_save ? rhs_save + nt : nullptr;
    p.rhv = rhs_save ? rhs_save + 2 * nt : nullptr;
    p.queue = c->queue;
    p.queue_count = c->queue_count;
    set_atlas(c, p, time);
    if (int r = set_nesting(c, p, time)) return r;
    if (int r = set_weather(c, p, time)) return r;
    const int epb = 64 / c->nn;
    const unsigned int nblk = (unsigned int)((c->K + epb - 1) / epb);
    const unsigned int nchunk = (nblk + 1023) / 1024;
    p.xcd_stride = c->shape.xcd ? (nblk + 7) / 8 : 0;
    dim3 grid(p.xcd_stride ? 8 * p.xcd_stride : nblk), sub_grid(std::min(nblk, c->shape.sub_groups));
    // the forcing given per node compiled in only where some is set
    const bool force = c->tide_slot || c->nest_slot || c->wx_series || c->manning_node || c->coriolis_node;
    p.manning_node = c->manning_node;
    p.coriolis_node = c->coriolis_node;
    if (c->timing) DG_TRY(hipEventRecord(c->ev[0], c->stream));
    switch (c->order) {
        case 1: hipLaunchKernelGGL((force ? wet_stage_kernel<1, true> : wet_stage_kernel<1>), grid, dim3(64), 0, c->stream, p); break;
        case 2: hipLaunchKernelGGL((force ? wet_stage_kernel<2, true> : wet_stage_kernel<2>), grid, dim3(64), 0, c->stream, p); break;
        default: hipLaunchKernelGGL((force ? wet_stage_kernel<3, true> : wet_stage_kernel<3>), grid, dim3(64), 0, c->stream, p); break;
    }
    // (one chunk: the compaction's own scan is the whole offset, no chunk totals)
    if (nchunk > 1)
        hipLaunchKernelGGL(chunk_count_kernel, dim3(nchunk), dim3(256), 0, c->stream, c->queue_count, nblk, c->chunk_total);
    hipLaunchKernelGGL(compact_kernel, dim3(nchunk), dim3(256), 0, c->stream, c->queue_count, c->queue, nblk,
                       c->chunk_total, epb, c->sub_list, c->sub_count);
    switch (c->order) {
        case 1: hipLaunchKernelGGL((force ? subcell_stage_kernel<1, false, true> : subcell_stage_kernel<1>), sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
        case 2: hipLaunchKernelGGL((force ? subcell_stage_kernel<2, false, true> : subcell_stage_kernel<2>), sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
        default: hipLaunchKernelGGL((force ? subcell_stage_kernel<3, false, true> : subcell_stage_kernel<3>), sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
    }
    DG_TRY(hipGetLastError());
    if (c->timing) {
        DG_TRY(hipEventRecord(c->ev[1], c->stream));
        DG_TRY(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        DG_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->stage_ms += ms;
        c->stage_n += 1;
    }
    return 0;
}

// The time step's kernels (element steps to c->element_dt with element_out,
// the minimum to c->dt_value), not waited for
static int launch_dt_kernels(dg_ctx* c, const double* state, double cfl, int scheme, bool element_out) {
    size_t nt = (size_t)c->K * c->nn;
    DtParams p;
    p.h = state; p.hu = state + nt; p.hv = state + 2 * nt;
    p.corners = c->corners; p.emass = c->emass; p.links = c->links; p.affine = c->affine;
    p.n_elements = c->K;
    p.cfl = cfl; p.g = c->g; p.h_min = c->h_min; p.h_dry = c->h_dry;
    // PositivityBound::new (swe_2d.rs:2188-2203)
    double n = c->order;
    double ssp = scheme == DG_SSP_RK3 ? 1.0 : 2.0;
    static const double lin3[4] = {1.6, 1.04, 0.73, 0.56}, lin43[4] = {2.2, 1.51, 1.15, 0.93};
    static const double sub3[4] = {0.94, 0.65, 0.47, 0.36}, sub43[4] = {1.5, 1.1, 0.97, 0.75};
    p.pcfl = ssp * ((2.0 * n + 1.0) / (2.0 * n * (n + 1.0)));
    p.max_cfl = std::fmax(p.pcfl, 0.9 * (scheme == DG_SSP_RK3 ? lin3 : lin43)[c->order - 1]);
    p.max_cfl_sub = std::fmax(p.pcfl, 0.9 * (scheme == DG_SSP_RK3 ? sub3 : sub43)[c->order - 1]);
    p.positivity = (c->flags & DG_WET_DRY) ? 1 : 0;
    p.element_dt = element_out ? c->element_dt : nullptr;
    p.slots = c->dt_slots;
    const int64_t epb = 64 / c->nn, nblk = (c->K + epb - 1) / epb;
    p.xcd_stride = (unsigned int)((nblk + 7) / 8);
    dim3 grid(8 * p.xcd_stride);
    switch (c->order) {
        case 1: hipLaunchKernelGGL(dt_kernel<1>, grid, dim3(64), 0, c->stream, p); break;
        case 2: hipLaunchKernelGGL(dt_kernel<2>, grid, dim3(64), 0, c->stream, p); break;
        default: hipLaunchKernelGGL(dt_kernel<3>, grid, dim3(64), 0, c->stream, p); break;
    }
    hipLaunchKernelGGL(slots_min_kernel, dim3(1), dim3(DT_SLOTS), 0, c->stream, c->dt_slots, c->dt_value);
    DG_TRY(hipGetLastError());
    return 0;
}

static int launch_dt(dg_ctx* c, const double* state, double cfl, int scheme, bool element_out, double* dt) {
    if (c->timing) DG_TRY(hipEventRecord(c->ev[0], c->stream));
    if (int r = launch_dt_kernels(c, state, cfl, scheme, element_out)) return r;
    if (c->timing) DG_TRY(hipEventRecord(c->ev[1], c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    const double value = *(volatile double*)c->dt_host;
    if (c->timing) {
        float ms = 0.f;
        DG_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->dt_ms += ms;
        c->dt_n += 1;
    }
    *dt = value;
    return 0;
}

static int ensure_work(dg_ctx* c) {
    size_t bytes = (size_t)3 * c->K * c->nn * sizeof(double);
    for (auto& w : c->work) {
        if (!w) {
            w = new dg_state;
            DG_TRY(hipMalloc(&w->d, bytes));
        }
    }
    return 0;
}

// Shu-Osher stages (a, b, c, time) per scheme (integrator.rs:239-263)
static const double RK3[3][4] = {{1.0, 0.0, 1.0, 0.0}, {0.75, 0.25, 0.25, 1.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 0.5}};
static const double RK43[4][4] = {{1.0, 0.0, 0.5, 0.0}, {0.0, 1.0, 0.5, 0.5}, {2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 1.0},
                                  {0.0, 1.0, 0.5, 0.5}};

// The launch shapes tuned so far, per (device, order, size class: floor(log2 K));
// a form uploaded in a class that is here starts from its shape
static std::mutex g_shape_mutex;
static std::map<std::tuple<int, int, int>, LaunchShape> g_shapes;

static int size_class(int64_t K) {
    int q = 0;
    while (q < 62 && (int64_t(1) << (q + 1)) <= K) ++q;
    return q;
}

static LaunchShape shape_for(int device, int order, int64_t K) {
    std::lock_guard<std::mutex> lock(g_shape_mutex);
    const auto it = g_shapes.find(std::make_tuple(device, order, size_class(K)));
    return it == g_shapes.end() ? LaunchShape{} : it->second;
}

// The autotuner of the single-run stage: every launch shape (block mapping,
// subcell grid) timed on the caller's state, three stages each after one
// warm-up, into work buffers; the fastest is kept for the context and for its
// (device, order, size class). The shapes change no arithmetic.
static int tune_impl(dg_ctx* c, const dg_state* s) {
    if (int r = ensure_work(c)) return r;
    const bool timing = c->timing;
    c->timing = false;
    unsigned long long clips = 0;  // the caller's count, kept across the timing launches
    DG_TRY(hipMemcpyAsync(&clips, c->clips, sizeof(clips), hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    static const unsigned int subs[3] = {2048, SUB_GROUPS, 65536};
    // two passes over the shapes, five stages each after a warm-up; a shape's
    // time is its better pass (the passes interleave the shapes against drift)
    float ms_of[2][3];
    for (int pass = 0; pass < 2; ++pass)
        for (unsigned int x = 0; x < 2; ++x)
            for (int q = 0; q < 3; ++q) {
                c->shape = LaunchShape{1u - x, subs[q]};
                // stage 1 of SSP-RK3 at a small step, into a work buffer
                if (int r = launch_stage(c, s->d, s->d, c->work[0]->d, 1.0, 0.0, 1e-3, 0.0, 0)) return r;
                DG_TRY(hipEventRecord(c->ev[0], c->stream));
                for (int rep = 0; rep < 5; ++rep)
                    if (int r = launch_stage(c, s->d, s->d, c->work[0]->d, 1.0, 0.0, 1e-3, 0.0, 0)) return r;
                DG_TRY(hipEventRecord(c->ev[1], c->stream));
                DG_TRY(hipEventSynchronize(c->ev[1]));
                float ms = 0.f;
                DG_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                ms_of[x][q] = pass == 0 ? ms : std::min(ms_of[x][q], ms);
            }
    LaunchShape best = c->shape;
    float best_ms = 1e30f;
    for (unsigned int x = 0; x < 2; ++x)
        for (int q = 0; q < 3; ++q)
            if (ms_of[x][q] < best_ms) { best_ms = ms_of[x][q]; best = LaunchShape{1u - x, subs[q]}; }
    c->shape = best;
    c->timing = timing;
    DG_TRY(hipMemcpyAsync(c->clips, &clips, sizeof(clips), hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    std::lock_guard<std::mutex> lock(g_shape_mutex);
    g_shapes[std::make_tuple(c->device, c->order, size_class(c->K))] = best;
    return 0;
}

static int step_impl(dg_ctx* c, dg_state* s, double t, double dt, int scheme) {
    // stage values in their own buffers, the last one the new state
    // (multirate.rs:1285-1311 with one level)
    if (int r = ensure_work(c)) return r;
    int ns = scheme == DG_SSP_RK3 ? 3 : 4;
    const double(*st)[4] = scheme == DG_SSP_RK3 ? RK3 : RK43;
    const double* input = s->d;
    for (int q = 0; q < ns; ++q) {
        double* out = c->work[q]->d;
        if (int r = launch_stage(c, s->d, input, out, st[q][0], st[q][1], st[q][2] * dt, t + st[q][3] * dt, 0)) return r;
        input = out;
    }
    if (s->owned) std::swap(s->d, c->work[ns - 1]->d);
    else DG_TRY(hipMemcpyAsync(s->d, c->work[ns - 1]->d, (size_t)3 * c->K * c->nn * 8, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

static int zero_async(dg_ctx* c, double* p, size_t n) {
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0, c->stream, p, n);
    DG_TRY(hipGetLastError());
    return 0;
}

static int ensure_adjoint(dg_ctx* c) {
    if (c->facebuf) return 0;
    const size_t nf = 4 * (size_t)(c->order + 1), nt = (size_t)c->K * c->nn;
    DG_TRY(hipMalloc(&c->facebuf, (size_t)c->K * nf * 4 * sizeof(double)));
    DG_TRY(hipMalloc(&c->innerbuf, (size_t)c->K * nf * 4 * sizeof(double)));
    DG_TRY(hipMalloc(&c->facecot, (size_t)c->K * nf * 3 * sizeof(double)));
    {
        std::vector<uint32_t> links((size_t)c->K * 4);
        std::vector<double> normals((size_t)c->K * 8);  // the faces' nx, then ny (the head of fgeo)
        DG_TRY(hipStreamSynchronize(c->stream));
        DG_TRY(hipMemcpy(links.data(), c->links, links.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        DG_TRY(hipMemcpy(normals.data(), c->fgeo, normals.size() * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<uint32_t> edges, split;
        vjp_edge_lists(links.data(), normals.data(), normals.data() + (size_t)c->K * 4, c->K, edges, split);
        c->n_vjp_edges = (uint32_t)edges.size();
        c->n_vjp_split = (uint32_t)split.size();
        DG_TRY(hipMalloc(&c->vjp_edges, edges.size() * sizeof(uint32_t)));
        DG_TRY(hipMemcpy(c->vjp_edges, edges.data(), edges.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        if (!split.empty()) {
            DG_TRY(hipMalloc(&c->vjp_split, split.size() * sizeof(uint32_t)));
            DG_TRY(hipMemcpy(c->vjp_split, split.data(), split.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        }
    }
    DG_TRY(hipMalloc(&c->kind, (size_t)c->K));
    DG_TRY(hipMalloc(&c->gbed, nt * sizeof(double)));
    DG_TRY(hipMalloc(&c->gman, (size_t)c->K * sizeof(double)));
    for (auto& w : c->adj) {
        w = new dg_state;
        DG_TRY(hipMalloc(&w->d, 3 * nt * sizeof(double)));
    }
    for (auto& w : c->rhsbuf) {
        w = new dg_state;
        DG_TRY(hipMalloc(&w->d, 3 * nt * sizeof(double)));
    }
    if (int r = zero_async(c, c->gbed, nt)) return r;
    return zero_async(c, c->gman, (size_t)c->K);
}

// The VJP of one stage: gbase (may be null) and ginput are added to, and the
// parameter cotangents; the neighbour parts of ginput and the bed's are
// gathered after the element kernels. rhs: the stage's RHS as its forward
// saved it (not read for rhs_only).
static int launch_stage_vjp(dg_ctx* c, const double* base, const double* input, const double* rhs, const double* w,
                            double* gbase, double* ginput, double a, double b, double cc, double time, int rhs_only,
                            int input_write = 0) {
    DG_REQUIRE(!(c->flags & DG_ENTROPY_STABLE), "the EntropyStable faces have no VJP here (the adjoint is of the WetDry faces)");
    DG_REQUIRE(!c->tide_slot, "the boundary tides per node have no VJP here (the adjoint is of the uniform tide)");
    DG_REQUIRE(!c->nest_slot, "the nesting band has no VJP here");
    DG_REQUIRE(!c->wx_series, "the gridded weather has no VJP here");
    DG_REQUIRE(!c->manning_node, "Manning n per node has no VJP here (the gradient is per element n)");
    DG_REQUIRE(!c->coriolis_node, "the beta-plane has no VJP here (the adjoint is of the f-plane)");
    if (int r = ensure_adjoint(c)) return r;
    size_t nt = (size_t)c->K * c->nn;
    VjpParams P{};
    StageParams& p = P.s;
    p.bh = base; p.bhu = base + nt; p.bhv = base + 2 * nt;
    p.ih = input; p.ihu = input + nt; p.ihv = input + 2 * nt;
    p.oh = p.ohu = p.ohv = nullptr;
    p.bed = c->bed; p.corners = c->corners; p.manning = c->manning;
    p.links = c->links; p.tags = c->tags; p.affine = c->affine;
    p.n_elements = c->K;
    p.a = a; p.b = b; p.c = cc;
    p.eta_e = tide_elevation(c->tide, c->tide_nodal, time);
    p.g = c->g; p.h_min = c->h_min; p.h_dry = c->h_dry; p.u_max = c->u_max; p.tau = c->tau; p.f0 = c->f0;
    p.sqrt_g = std::sqrt(c->g);
    p.rhs_only = rhs_only;
    p.flags = c->flags;
    p.clips = c->clips;
    p.fstar_out = nullptr;
    p.geo = c->geo;
    p.fgeo = c->fgeo;
    p.queue = c->queue;
    p.queue_count = c->queue_count;
    p.rh = const_cast<double*>(rhs);
    p.rhu = rhs ? p.rh + nt : nullptr;
    p.rhv = rhs ? p.rh + 2 * nt : nullptr;
    P.wh = w; P.whu = w + nt; P.whv = w + 2 * nt;
    P.gbh = gbase; P.gbhu = gbase ? gbase + nt : nullptr; P.gbhv = gbase ? gbase + 2 * nt : nullptr;
    P.gih = ginput; P.gihu = ginput + nt; P.gihv = ginput + 2 * nt;
    P.input_write = input_write;
    P.gbed = c->gbed;
    P.gman = c->gman;
    P.facebuf = c->facebuf;
    P.innerbuf = c->innerbuf;
    P.facecot = c->facecot;
    P.kind = c->kind;
    const int epb = 64 / c->nn;
    const unsigned int nblk = (unsigned int)((c->K + epb - 1) / epb);
    const unsigned int nchunk = (nblk + 1023) / 1024;
    p.xcd_stride = (nblk + 7) / 8;
    dim3 grid(8 * p.xcd_stride), sub_grid(std::min(nblk, SUB_GROUPS)), ggrid((unsigned)((nt + 255) / 256)),
        fgrid((unsigned)(((size_t)c->n_vjp_edges * (c->order + 1) + 63) / 64));
    switch (c->order) {
        case 1: hipLaunchKernelGGL(wet_vjp_kernel<1>, grid, dim3(64), 0, c->stream, P); break;
        case 2: hipLaunchKernelGGL(wet_vjp_kernel<2>, grid, dim3(64), 0, c->stream, P); break;
        default: hipLaunchKernelGGL(wet_vjp_kernel<3>, grid, dim3(64), 0, c->stream, P); break;
    }
    // (one chunk: the compaction's own scan is the whole offset, no chunk totals)
    if (nchunk > 1)
        hipLaunchKernelGGL(chunk_count_kernel, dim3(nchunk), dim3(256), 0, c->stream, c->queue_count, nblk, c->chunk_total);
    hipLaunchKernelGGL(compact_kernel, dim3(nchunk), dim3(256), 0, c->stream, c->queue_count, c->queue, nblk,
                       c->chunk_total, epb, c->sub_list, c->sub_count);
    switch (c->order) {
        case 1: hipLaunchKernelGGL(subcell_vjp_kernel<1>, sub_grid, dim3(64), 0, c->stream, P, c->sub_list, c->sub_count); break;
        case 2: hipLaunchKernelGGL(subcell_vjp_kernel<2>, sub_grid, dim3(64), 0, c->stream, P, c->sub_list, c->sub_count); break;
        default: hipLaunchKernelGGL(subcell_vjp_kernel<3>, sub_grid, dim3(64), 0, c->stream, P, c->sub_list, c->sub_count); break;
    }
    switch (c->order) {
        case 1: hipLaunchKernelGGL(edge_vjp_kernel<1>, fgrid, dim3(64), 0, c->stream, P, c->vjp_edges, c->n_vjp_edges); break;
        case 2: hipLaunchKernelGGL(edge_vjp_kernel<2>, fgrid, dim3(64), 0, c->stream, P, c->vjp_edges, c->n_vjp_edges); break;
        default: hipLaunchKernelGGL(edge_vjp_kernel<3>, fgrid, dim3(64), 0, c->stream, P, c->vjp_edges, c->n_vjp_edges); break;
    }
    if (c->n_vjp_split) {
        const dim3 sgrid((unsigned)(((size_t)c->n_vjp_split * (c->order + 1) + 63) / 64));
        switch (c->order) {
            case 1: hipLaunchKernelGGL(edge_vjp_split_kernel<1>, sgrid, dim3(64), 0, c->stream, P, c->vjp_split, c->n_vjp_split); break;
            case 2: hipLaunchKernelGGL(edge_vjp_split_kernel<2>, sgrid, dim3(64), 0, c->stream, P, c->vjp_split, c->n_vjp_split); break;
            default: hipLaunchKernelGGL(edge_vjp_split_kernel<3>, sgrid, dim3(64), 0, c->stream, P, c->vjp_split, c->n_vjp_split); break;
        }
    }
    double *gh = ginput, *ghu = ginput + nt, *ghv = ginput + 2 * nt;
    switch (c->order) {
        case 1: hipLaunchKernelGGL(vjp_gather_kernel<1>, ggrid, dim3(256), 0, c->stream, c->links, c->facebuf, c->innerbuf, c->kind, c->K, gh, ghu, ghv, c->gbed); break;
        case 2: hipLaunchKernelGGL(vjp_gather_kernel<2>, ggrid, dim3(256), 0, c->stream, c->links, c->facebuf, c->innerbuf, c->kind, c->K, gh, ghu, ghv, c->gbed); break;
        default: hipLaunchKernelGGL(vjp_gather_kernel<3>, ggrid, dim3(256), 0, c->stream, c->links, c->facebuf, c->innerbuf, c->kind, c->K, gh, ghu, ghv, c->gbed); break;
    }
    DG_TRY(hipGetLastError());
    return 0;
}

// The VJP of one SSP step from state X: the stages are recomputed, saving
// their RHS, then reversed; lam = (dX'/dX)^T lam_next (overwritten). With a
// step record (the step's stage values 0..ns-2, then its ns stage RHS, as the
// forward kept them) nothing is recomputed.
static int step_vjp_impl(dg_ctx* c, const double* X, double t, double dt, int scheme, const double* lam_next,
                         double* lam, const double* record = nullptr) {
    if (int r = ensure_work(c)) return r;
    if (int r = ensure_adjoint(c)) return r;
    const int ns = scheme == DG_SSP_RK3 ? 3 : 4;
    const double(*st)[4] = scheme == DG_SSP_RK3 ? RK3 : RK43;
    const size_t n3 = (size_t)3 * c->K * c->nn;
    const double* input = X;
    if (!record) {
        for (int q = 0; q < ns; ++q) {
            if (int r = launch_stage(c, X, input, c->work[q]->d, st[q][0], st[q][1], st[q][2] * dt, t + st[q][3] * dt,
                                     0, nullptr, c->rhsbuf[q]->d))
                return r;
            input = c->work[q]->d;
        }
    }
    auto stage_value = [&](int q) -> const double* { return record ? record + (size_t)q * n3 : c->work[q]->d; };
    auto stage_rhs = [&](int q) -> const double* { return record ? record + (size_t)(ns - 1 + q) * n3 : c->rhsbuf[q]->d; };
    if (int r = zero_async(c, lam, n3)) return r;
    const double* w = lam_next;
    int tb = 0;
    for (int q = ns - 1; q >= 0; --q) {
        // the stage values' cotangents are fresh buffers, written (not added to)
        const double* inq = q == 0 ? X : stage_value(q - 1);
        double* gin = q > 0 ? c->adj[tb]->d : lam;
        if (int r = launch_stage_vjp(c, X, inq, stage_rhs(q), w, lam, gin, st[q][0], st[q][1], st[q][2] * dt,
                                     t + st[q][3] * dt, 0, q > 0 ? 1 : 0))
            return r;
        w = gin;
        tb ^= 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Binomial checkpointing. reverse(a, b) holds the state of step a in a slot
// and reverses steps b-1 .. a with `free` more slots: it advances to a split
// a+m, stores that state, reverses [a+m, b) with one slot fewer, then [a, a+m).
// The splits minimise the forward steps: T(l, s) = min_m m + T(l-m, s-1) + T(m, s)
// with T(1, s) = 0 and T(l, 0) = l(l-1)/2 (restarting from the slot for every
// step); G(l, s) = min_m T(m, s) + G(l-m, s-1) is the same for the right-most
// descent, whose checkpoints the first forward sweep takes (it runs anyway for
// the misfit). T and G are convex in l, so the best split does not move left
// as l grows and one pointer per s finds it: O(s n) set-up and memory (the
// slots the schedule uses are capped so that the tables stay under 64 M
// entries). With 20 slots over 2000 steps it takes 5980 forward steps, the
// binomial bound of C(s+r, s) steps at r repetitions (r = 4 here).
// ---------------------------------------------------------------------------
struct Revolve {
    int n = 0, S = 0;             // steps, free slots beyond the initial state's
    std::vector<int64_t> T, G;    // costs [s][l]
    std::vector<int32_t> AT, AG;  // best splits [s][l]
    size_t idx(int s, int l) const { return (size_t)s * (n + 1) + l; }

    void init(int n_, int S_) {
        n = n_;
        S = std::min(S_, std::max(n_ - 1, 0));  // more slots than steps change nothing
        while (S > 1 && (size_t)(S + 1) * (size_t)(n + 1) > (size_t)64 << 20) --S;
        T.assign((size_t)(S + 1) * (n + 1), 0);
        G = T;
        AT.assign(T.size(), 1);
        AG = AT;
        for (int l = 1; l <= n; ++l) T[idx(0, l)] = G[idx(0, l)] = (int64_t)l * (l - 1) / 2;
        for (int s = 1; s <= S; ++s) {
            int mt = 1, mg = 1;
            for (int l = 2; l <= n; ++l) {
                auto ft = [&](int m) { return m + T[idx(s - 1, l - m)] + T[idx(s, m)]; };
                auto fg = [&](int m) { return T[idx(s, m)] + G[idx(s - 1, l - m)]; };
                mt = std::min(mt, l - 1);
                mg = std::min(mg, l - 1);
                while (mt + 1 < l && ft(mt + 1) <= ft(mt)) ++mt;
                while (mg + 1 < l && fg(mg + 1) <= fg(mg)) ++mg;
                T[idx(s, l)] = ft(mt); AT[idx(s, l)] = mt;
                G[idx(s, l)] = fg(mg); AG[idx(s, l)] = mg;
            }
        }
    }
    int slots(int free) const { return std::min(free, S); }
    int split(int l, int s, bool prestored) const {
        s = std::min(s, S);
        return prestored ? AG[idx(s, l)] : AT[idx(s, l)];
    }
};

struct dg_tape {
    int snapshots = 0;
    std::vector<dg_state*> slot;
    std::vector<char> used;
    dg_state* cur = nullptr;
    dg_state* lam[2] = {nullptr, nullptr};
    std::vector<double> dts, t_at;
    int n = 0, scheme = 0, every = 0, n_samples = 0, n_st = 0;
    double* eta = nullptr;  // device [n_samples][n_st]
    double* obs = nullptr;  // device [n_samples][n_st]
    int64_t fwd = 0, vjps = 0;
    bool ready = false;
    int64_t K = 0;
    int nn = 0;
    Revolve rv;
    std::vector<std::pair<int, int>> pre;  // (step, slot) kept by the forward sweep
    int lam_cur = 0;
    double scale = 1.0;
    // Step records (dg_tape_keep_stages): a step's stage values 0..ns-2 and its ns
    // stage RHS, so that its reverse recomputes nothing. A pool of snapshots + 1
    // records (at most one per step): the forward sweep records the step before
    // each checkpoint it keeps and its last step, every advance of the schedule its
    // last step, and a reverse step takes its step's record and frees it. The
    // schedule's reverse steps always find theirs: a record is made for the step
    // whose state the advance (or the sweep) leaves in a slot or in the tape's
    // state, and each level of the schedule holds at most one outstanding.
    bool keep_stages = false, records = false;
    double* rec = nullptr;
    size_t rec_bytes = 0;
    int rec_per_step = 0;
    std::vector<int> rec_step;  // the step each record holds (-1: free)
    int64_t rec_used = 0;       // reverse steps of the last backward taken from a record
};

// A free record for step i (nullptr when none is free or records are off)
static double* rec_take(dg_ctx* c, dg_tape* tp, int i) {
    if (!tp->records) return nullptr;
    for (size_t r = 0; r < tp->rec_step.size(); ++r)
        if (tp->rec_step[r] < 0) {
            tp->rec_step[r] = i;
            return tp->rec + r * tp->rec_per_step * (size_t)3 * c->K * c->nn;
        }
    return nullptr;
}

// One step i of the tape from state x: step_impl's stages, the last one written
// into dst (a checkpoint slot, so the state needs no copy there) or else swapped
// into the tape's own state buffer; with a record (rec) the stage values and RHS
// go into it
static int sweep_step(dg_ctx* c, dg_tape* tp, int i, const double* x, double* dst, double* rec) {
    if (int r = ensure_work(c)) return r;
    const int ns = tp->scheme == DG_SSP_RK3 ? 3 : 4;
    const double(*st)[4] = tp->scheme == DG_SSP_RK3 ? RK3 : RK43;
    const size_t n3 = (size_t)3 * c->K * c->nn;
    const double t = tp->t_at[i], dt = tp->dts[i];
    const double* input = x;
    for (int q = 0; q < ns; ++q) {
        double* out = q < ns - 1 ? (rec ? rec + (size_t)q * n3 : c->work[q]->d) : (dst ? dst : c->work[ns - 1]->d);
        if (int r = launch_stage(c, x, input, out, st[q][0], st[q][1], st[q][2] * dt, t + st[q][3] * dt, 0, nullptr,
                                 rec ? rec + (size_t)(ns - 1 + q) * n3 : nullptr))
            return r;
        input = out;
    }
    if (!dst) std::swap(tp->cur->d, c->work[ns - 1]->d);  // the tape's own state buffer
    return 0;
}

static int tape_free_slot(dg_tape* tp) {
    for (size_t i = 0; i < tp->slot.size(); ++i)
        if (!tp->used[i]) return (int)i;
    return -1;
}

static int tape_copy(dg_ctx* c, dg_state* dst, const dg_state* src) {
    DG_TRY(hipMemcpyAsync(dst->d, src->d, (size_t)3 * c->K * c->nn * 8, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

// Advance tp->cur from step a to step b (with step records, the last step
// recorded: it is the next one the schedule reverses from this state or its slot)
static int tape_advance(dg_ctx* c, dg_tape* tp, int a, int b) {
    for (int i = a; i < b; ++i) {
        double* rec = i == b - 1 ? rec_take(c, tp, i) : nullptr;
        if (int r = rec ? sweep_step(c, tp, i, tp->cur->d, nullptr, rec)
                        : step_impl(c, tp->cur, tp->t_at[i], tp->dts[i], tp->scheme))
            return r;
        tp->fwd += 1;
    }
    return 0;
}

// Reverse step i from the state X_i in `xs`: seed the cotangent of X_{i+1} with
// the misfit's sample there, then lam <- (dX_{i+1}/dX_i)^T lam
static int tape_step_back(dg_ctx* c, dg_tape* tp, const dg_state* xs, int i) {
    const int after = i + 1;
    double* lam = tp->lam[tp->lam_cur]->d;
    if (after % tp->every == 0) {
        const int m = after / tp->every - 1;
        if (m < tp->n_samples) {
            hipLaunchKernelGGL(station_seed_kernel, dim3(1), dim3(1), 0, c->stream, c->st_elem, c->st_w, tp->n_st,
                               c->nn, tp->eta + (size_t)m * tp->n_st, tp->obs + (size_t)m * tp->n_st, tp->scale, lam,
                               c->gbed);
            DG_TRY(hipGetLastError());
        }
    }
    double* out = tp->lam[tp->lam_cur ^ 1]->d;
    const double* record = nullptr;
    for (size_t r = 0; r < tp->rec_step.size(); ++r)
        if (tp->rec_step[r] == i) {
            record = tp->rec + r * tp->rec_per_step * (size_t)3 * c->K * c->nn;
            tp->rec_step[r] = -1;  // reusable by later records (stream order: after this VJP's reads)
            tp->rec_used += 1;
            break;
        }
    if (int r = step_vjp_impl(c, xs->d, tp->t_at[i], tp->dts[i], tp->scheme, lam, out, record)) return r;
    tp->lam_cur ^= 1;
    tp->vjps += 1;
    return 0;
}

static int tape_reverse(dg_ctx* c, dg_tape* tp, int a, int b, int slot_a, int free, bool prestored) {
    const int l = b - a;
    if (l <= 0) return 0;
    if (l == 1) return tape_step_back(c, tp, tp->slot[slot_a], a);
    if (free == 0) {
        for (int i = b - 1; i >= a; --i) {
            if (i == a) {
                if (int r = tape_step_back(c, tp, tp->slot[slot_a], a)) return r;
            } else {
                if (int r = tape_copy(c, tp->cur, tp->slot[slot_a])) return r;
                if (int r = tape_advance(c, tp, a, i)) return r;
                if (int r = tape_step_back(c, tp, tp->cur, i)) return r;
            }
        }
        return 0;
    }
    const int m = tp->rv.split(l, free, prestored);
    int sm = -1;
    if (prestored) {
        for (auto& ps : tp->pre)
            if (ps.first == a + m) sm = ps.second;
        DG_REQUIRE(sm >= 0, "dg_misfit_backward: checkpoint schedule mismatch");
    } else {
        sm = tape_free_slot(tp);
        DG_REQUIRE(sm >= 0, "dg_misfit_backward: no free checkpoint slot");
        tp->used[sm] = 1;
        if (int r = tape_copy(c, tp->cur, tp->slot[slot_a])) return r;
        if (int r = tape_advance(c, tp, a, a + m)) return r;
        if (int r = tape_copy(c, tp->slot[sm], tp->cur)) return r;
    }
    if (int r = tape_reverse(c, tp, a + m, b, sm, free - 1, prestored)) return r;
    tp->used[sm] = 0;
    return tape_reverse(c, tp, a, a + m, slot_a, free, false);
}

// Forward steps a schedule takes (first sweep included), without a device
static int64_t revolve_count_rec(const Revolve& rv, int l, int free, bool prestored) {
    if (l <= 1) return 0;
    if (free == 0) return (int64_t)l * (l - 1) / 2;
    const int m = rv.split(l, free, prestored);
    return (prestored ? 0 : m) + revolve_count_rec(rv, l - m, free - 1, prestored) + revolve_count_rec(rv, m, free, false);
}

// ---------------------------------------------------------------------------
// Ensembles (dg_batch): per order, the blocks of every member of that order in
// one work list; per step one time-step launch per order, one reduction to the
// members' steps, one copy of them to the host (the tide and the stage
// coefficients of every member are evaluated there, as for a single run),
// then the stages: per order, the wet kernel, the list compaction and the
// subcell kernel, optionally captured once in a HIP graph per step parity.
// ---------------------------------------------------------------------------
struct BatchOrder {
    unsigned int n_blocks = 0;
    uint32_t* block_member = nullptr;
    uint32_t* queue = nullptr;
    uint8_t* queue_count = nullptr;
    uint32_t *list = nullptr, *list_member = nullptr;
    unsigned int *chunk_total = nullptr, *list_count = nullptr;
    double* block_min = nullptr;
    unsigned int xcd = 1;            // XCD-aware block mapping (tuned)
    unsigned int sub_groups = SUB_GROUPS;  // subcell grid (tuned)
};

struct dg_batch {
    dg_ctx* owner = nullptr;
    std::vector<dg_ctx*> members;
    int scheme = DG_SSP_RK3, ns = 3;
    double cfl = -1.0;
    std::vector<BatchMember> h_mem;
    BatchMember* d_mem = nullptr;
    BatchOrder ord[4];
    BatchCoef* d_coef = nullptr;
    BatchCoef* h_coef = nullptr;  // pinned
    double* d_member_dt = nullptr;
    double* h_member_dt = nullptr;  // pinned
    int parity = 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    bool tuned = false;
};

static BatchArgs batch_args(dg_batch* b, int o, int q, int parity) {
    BatchOrder& B = b->ord[o];
    BatchArgs A;
    A.mem = b->d_mem;
    A.block_member = B.block_member;
    A.coef = b->d_coef;
    A.queue = B.queue;
    A.queue_count = B.queue_count;
    A.block_min = B.block_min;
    A.n_blocks = B.n_blocks;
    A.xcd_stride = B.xcd ? (B.n_blocks + 7) / 8 : 0;
    A.q = q;
    A.ns = b->ns;
    A.parity = parity;
    return A;
}

static int batch_stage_order(dg_batch* b, int o, int q, int parity) {
    BatchOrder& B = b->ord[o];
    if (B.n_blocks == 0) return 0;
    hipStream_t st = b->owner->stream;
    const BatchArgs A = batch_args(b, o, q, parity);
    const int epb = 64 / ((o + 1) * (o + 1));
    const unsigned int nchunk = (B.n_blocks + 1023) / 1024;
    dim3 grid(A.xcd_stride ? 8 * A.xcd_stride : B.n_blocks), sub_grid(std::min(B.n_blocks, B.sub_groups));
    switch (o) {
        case 1: hipLaunchKernelGGL(batch_wet_kernel<1>, grid, dim3(64), 0, st, A); break;
        case 2: hipLaunchKernelGGL(batch_wet_kernel<2>, grid, dim3(64), 0, st, A); break;
        default: hipLaunchKernelGGL(batch_wet_kernel<3>, grid, dim3(64), 0, st, A); break;
    }
    // (one chunk: the compaction's own scan is the whole offset, no chunk totals)
    if (nchunk > 1)
        hipLaunchKernelGGL(chunk_count_kernel, dim3(nchunk), dim3(256), 0, st, B.queue_count, B.n_blocks, B.chunk_total);
    hipLaunchKernelGGL(batch_compact_kernel, dim3(nchunk), dim3(256), 0, st, B.queue_count, B.queue, B.block_member,
                       B.n_blocks, B.chunk_total, epb, B.list, B.list_member, B.list_count);
    switch (o) {
        case 1: hipLaunchKernelGGL(batch_subcell_kernel<1>, sub_grid, dim3(64), 0, st, A, B.list, B.list_member, B.list_count); break;
        case 2: hipLaunchKernelGGL(batch_subcell_kernel<2>, sub_grid, dim3(64), 0, st, A, B.list, B.list_member, B.list_count); break;
        default: hipLaunchKernelGGL(batch_subcell_kernel<3>, sub_grid, dim3(64), 0, st, A, B.list, B.list_member, B.list_count); break;
    }
    DG_TRY(hipGetLastError());
    return 0;
}

static int batch_stages(dg_batch* b, int parity) {
    for (int q = 0; q < b->ns; ++q)
        for (int o = 1; o <= 3; ++o)
            if (int r = batch_stage_order(b, o, q, parity)) return r;
    return 0;
}

static int batch_dt(dg_batch* b) {
    hipStream_t st = b->owner->stream;
    for (int o = 1; o <= 3; ++o) {
        BatchOrder& B = b->ord[o];
        if (B.n_blocks == 0) continue;
        const BatchArgs A = batch_args(b, o, 0, b->parity);
        dim3 grid(A.xcd_stride ? 8 * A.xcd_stride : B.n_blocks);
        switch (o) {
            case 1: hipLaunchKernelGGL(batch_dt_kernel<1>, grid, dim3(64), 0, st, A); break;
            case 2: hipLaunchKernelGGL(batch_dt_kernel<2>, grid, dim3(64), 0, st, A); break;
            default: hipLaunchKernelGGL(batch_dt_kernel<3>, grid, dim3(64), 0, st, A); break;
        }
    }
    const int M = (int)b->members.size();
    hipLaunchKernelGGL(batch_member_min_kernel, dim3(M), dim3(256), 0, st, b->d_mem, M, b->d_member_dt);
    DG_TRY(hipGetLastError());
    DG_TRY(hipMemcpyAsync(b->h_member_dt, b->d_member_dt, M * sizeof(double), hipMemcpyDeviceToHost, st));
    DG_TRY(hipStreamSynchronize(st));
    return 0;
}

static int batch_drop_graphs(dg_batch* b) {
    for (int q = 0; q < 2; ++q) {
        if (b->exec[q]) DG_TRY(hipGraphExecDestroy(b->exec[q]));
        if (b->graph[q]) DG_TRY(hipGraphDestroy(b->graph[q]));
        b->exec[q] = nullptr;
        b->graph[q] = nullptr;
    }
    return 0;
}

// The autotuner: per order, the wet kernel's block mapping (XCD-aware or
// linear) and the subcell grid, timed on this ensemble's first stage
static int batch_tune(dg_batch* b) {
    hipStream_t st = b->owner->stream;
    hipEvent_t e0 = b->owner->ev[0], e1 = b->owner->ev[1];
    // the members' clip counts, kept across the timing launches
    const size_t M = b->members.size();
    std::vector<unsigned long long> clips(M, 0);
    for (size_t m = 0; m < M; ++m)
        DG_TRY(hipMemcpyAsync(&clips[m], b->members[m]->clips, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    DG_TRY(hipStreamSynchronize(st));
    static const unsigned int subs[3] = {2048, SUB_GROUPS, 65536};
    for (int o = 1; o <= 3; ++o) {
        BatchOrder& B = b->ord[o];
        if (B.n_blocks == 0) continue;
        float best = 1e30f;
        unsigned int bx = 1, bs = SUB_GROUPS;
        for (unsigned int x = 0; x < 2; ++x)
            for (unsigned int sg : subs) {
                B.xcd = x;
                B.sub_groups = sg;
                if (int r = batch_stage_order(b, o, 0, b->parity)) return r;  // warm
                DG_TRY(hipEventRecord(e0, st));
                for (int rep = 0; rep < 3; ++rep)
                    if (int r = batch_stage_order(b, o, 0, b->parity)) return r;
                DG_TRY(hipEventRecord(e1, st));
                DG_TRY(hipEventSynchronize(e1));
                float ms = 0.f;
                DG_TRY(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; bx = x; bs = sg; }
            }
        B.xcd = bx;
        B.sub_groups = bs;
    }
    for (size_t m = 0; m < M; ++m)
        DG_TRY(hipMemcpyAsync(b->members[m]->clips, &clips[m], sizeof(unsigned long long), hipMemcpyHostToDevice, st));
    DG_TRY(hipStreamSynchronize(st));
    b->tuned = true;
    return batch_drop_graphs(b);
}

static int batch_free(dg_batch* b) {
    if (int r = batch_drop_graphs(b)) return r;
    for (auto& M : b->h_mem) {
        for (double* x : M.x)
            if (x) DG_TRY(hipFree(x));
        for (double* y : M.y)
            if (y) DG_TRY(hipFree(y));
    }
    for (int o = 1; o <= 3; ++o) {
        BatchOrder& B = b->ord[o];
        void* ptrs[] = {B.block_member, B.queue, B.queue_count, B.list, B.list_member, B.chunk_total, B.list_count,
                        B.block_min};
        for (void* p : ptrs)
            if (p) DG_TRY(hipFree(p));
    }
    if (b->d_mem) DG_TRY(hipFree(b->d_mem));
    if (b->d_coef) DG_TRY(hipFree(b->d_coef));
    if (b->d_member_dt) DG_TRY(hipFree(b->d_member_dt));
    if (b->h_coef) DG_TRY(hipHostFree(b->h_coef));
    if (b->h_member_dt) DG_TRY(hipHostFree(b->h_member_dt));
    return 0;
}

extern "C" {

const char* dg_last_error(void) { return g_err.c_str(); }

int dg_create(int device, dg_ctx** out) {
    DG_REQUIRE(out != nullptr, "dg_create: null output");
    DG_TRY(hipSetDevice(device));
    dg_ctx* c = new dg_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; g_err = hipGetErrorString(e); return 1; }
    c->stream = c->own_stream;
    static bool tables = false;
    if (!tables) {
        GllTables t[4];
        std::memset(&t[0], 0, sizeof(GllTables));
        for (int n = 1; n <= 3; ++n) t[n] = make_tables(n);
        DG_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_gll), t, sizeof(t)));
        tables = true;
    }
    DG_TRY(hipMalloc(&c->clips, sizeof(unsigned long long)));
    // (the time step's one value goes to the host without a copy launch: the
    // reduction kernel stores it in pinned memory the host reads after the sync)
    DG_TRY(hipHostMalloc((void**)&c->dt_host, sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    DG_TRY(hipHostGetDevicePointer((void**)&c->dt_value, c->dt_host, 0));
    DG_TRY(hipMalloc(&c->sub_count, sizeof(unsigned int)));
    DG_TRY(hipMalloc(&c->flag, sizeof(int)));
    DG_TRY(hipMemset(c->clips, 0, sizeof(unsigned long long)));
    DG_TRY(hipEventCreate(&c->ev[0]));
    DG_TRY(hipEventCreate(&c->ev[1]));
    DG_TRY(scratch_address(device, (void**)&c->scratch));
    *out = c;
    return 0;
}

int dg_destroy(dg_ctx* c) {
    if (!c) return 0;
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    if (int r = free_statics(c)) return r;
    DG_TRY(hipFree(c->clips));
    DG_TRY(hipHostFree(c->dt_host));
    DG_TRY(hipFree(c->sub_count));
    DG_TRY(hipFree(c->flag));
    DG_TRY(hipEventDestroy(c->ev[0]));
    DG_TRY(hipEventDestroy(c->ev[1]));
    if (c->lvl_moment_dev) {
        DG_TRY(hipFree(c->lvl_moment_dev));
        DG_TRY(hipHostFree(c->lvl_moment_host));
        for (int i = 0; i < dg_ctx::MOMENT_RING; ++i) DG_TRY(hipEventDestroy(c->lvl_moment_ev[i]));
    }
    DG_TRY(hipStreamDestroy(c->own_stream));
    delete c;
    return 0;
}

int dg_upload_form(dg_ctx* c, const dg_form* f) {
    DG_LOCK(c);
    DG_REQUIRE(c && f, "dg_upload_form: null argument");
    DG_REQUIRE(f->order >= 1 && f->order <= 3, "dg_upload_form: order must be 1, 2 or 3");
    DG_REQUIRE(f->n_elements > 0 && f->n_elements < (int64_t(1) << 28) - 1, "dg_upload_form: element count out of range");
    DG_REQUIRE(f->corners && f->face_links && f->face_tags && f->bed, "dg_upload_form: missing mesh, tags or bed");
    // the kernels address per-node arrays (8 (N+1)^2 K bytes) and the corners (64 K bytes) by 32-bit byte offsets
    DG_REQUIRE((uint64_t)f->n_elements * 8u * (uint64_t)((f->order + 1) * (f->order + 1)) < (uint64_t(1) << 32) &&
                   (uint64_t)f->n_elements * 64u < (uint64_t(1) << 32),
               "dg_upload_form: a per-node array of this mesh would exceed 4 GiB (at most 536,870,911 nodes and "
               "67,108,863 elements per context)");
    DG_REQUIRE(!((f->flags & DG_ENTROPY_STABLE) && (f->flags & DG_SUBCELLS)),
               "dg_upload_form: the EntropyStable faces take no subcells (DG_SUBCELLS off)");
    DG_TRY(hipSetDevice(c->device));
    if (int r = free_statics(c)) return r;
    c->shape = shape_for(c->device, f->order, f->n_elements);
    c->order = f->order;
    c->nn = (f->order + 1) * (f->order + 1);
    c->K = f->n_elements;
    c->flags = f->flags;
    c->g = f->g; c->h_min = f->h_min; c->h_dry = f->h_dry; c->tau = f->relaxation_time; c->u_max = f->max_velocity;
    c->f0 = f->coriolis_f0;
    const int64_t K = c->K;
    const int nn = c->nn;
    // Host-side checks: positive Jacobian at every node (geometric.rs:163-170), links in range
    GllTables t = make_tables(f->order);
    std::vector<uint8_t> aff(K);
    for (int64_t k = 0; k < K; ++k) {
        const double* v = f->corners + 8 * k;
        aff[k] = parallelogram(v) ? 1 : 0;
        for (int i = 0; i < nn; ++i) {
            double d = host_bilinear_det(v, t.x1[i % (f->order + 1)], t.x1[i / (f->order + 1)]);
            if (!(std::isfinite(d) && d > 0.0)) {
                g_err = "dg_upload_form: element " + std::to_string(k) + " has a non-positive Jacobian at node " +
                        std::to_string(i) + " (degenerate, non-convex or clockwise)";
                return 2;
            }
        }
        for (int q = 0; q < 4; ++q) {
            uint32_t l = f->face_links[4 * k + q];
            if (l != LINK_BOUNDARY && (int64_t)(l >> 4) >= K) {
                g_err = "dg_upload_form: face link of element " + std::to_string(k) + " points past the mesh";
                return 2;
            }
        }
    }
    std::vector<double> man(K, f->manning_uniform);
    if (f->manning) std::copy(f->manning, f->manning + K, man.begin());
    DG_TRY(hipMalloc(&c->corners, 8 * K * sizeof(double)));
    DG_TRY(hipMalloc(&c->bed, (size_t)K * nn * sizeof(double)));
    DG_TRY(hipMalloc(&c->manning, K * sizeof(double)));
    DG_TRY(hipMalloc(&c->element_dt, K * sizeof(double)));
    DG_TRY(hipMalloc(&c->links, 4 * K * sizeof(uint32_t)));
    DG_TRY(hipMalloc(&c->tags, 4 * K));
    DG_TRY(hipMalloc(&c->affine, K));
    {
        const int64_t epb = 64 / nn, nblk = (K + epb - 1) / epb;
        DG_TRY(hipMalloc(&c->queue, nblk * epb * sizeof(uint32_t)));
        DG_TRY(hipMalloc(&c->queue_count, nblk));
        DG_TRY(hipMalloc(&c->dt_slots, DT_SLOTS * sizeof(unsigned long long)));
        const std::vector<unsigned long long> empty(DT_SLOTS, DT_SLOT_EMPTY);
        DG_TRY(hipMemcpy(c->dt_slots, empty.data(), DT_SLOTS * sizeof(unsigned long long), hipMemcpyHostToDevice));
        DG_TRY(hipMalloc(&c->sub_list, K * sizeof(uint32_t)));
        DG_TRY(hipMalloc(&c->chunk_total, ((nblk + 1023) / 1024) * sizeof(unsigned int)));
    }
    DG_TRY(hipMemcpy(c->corners, f->corners, 8 * K * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->bed, f->bed, (size_t)K * nn * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->manning, man.data(), K * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->links, f->face_links, 4 * K * sizeof(uint32_t), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->tags, f->face_tags, 4 * K, hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->affine, aff.data(), K, hipMemcpyHostToDevice));
    DG_TRY(hipMalloc(&c->geo, (size_t)GEO_COUNT * K * nn * sizeof(double)));
    DG_TRY(hipMalloc(&c->fgeo, (size_t)(3 * 4 + 4 * (f->order + 1)) * K * sizeof(double)));
    DG_TRY(hipMalloc(&c->emass, (size_t)K * sizeof(double)));
    if (int r = launch_geometry(c)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_set_tide(dg_ctx* c, const dg_tide* t) {
    DG_LOCK(c);
    DG_REQUIRE(c && t, "dg_set_tide: null argument");
    DG_REQUIRE(t->n >= 0 && t->n <= 8, "dg_set_tide: at most 8 constituents");
    c->tide = *t;
    for (int i = 0; i < 8; ++i) {
        c->tide_nodal[0][i] = 1.0;
        for (int q = 1; q < 5; ++q) c->tide_nodal[q][i] = 0.0;
    }
    return 0;
}

int dg_set_tide_nodal(dg_ctx* c, const double* f0, const double* f_rate, const double* v0_deg, const double* u0_deg,
                      const double* u_rate_deg) {
    DG_LOCK(c);
    DG_REQUIRE(c && f0 && f_rate && v0_deg && u0_deg && u_rate_deg, "dg_set_tide_nodal: null argument");
    for (int i = 0; i < c->tide.n; ++i) {
        c->tide_nodal[0][i] = f0[i];
        c->tide_nodal[1][i] = f_rate[i];
        c->tide_nodal[2][i] = v0_deg[i];
        c->tide_nodal[3][i] = u0_deg[i];
        c->tide_nodal[4][i] = u_rate_deg[i];
    }
    return 0;
}

int dg_state_alloc(dg_ctx* c, dg_state** s) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && c->K > 0, "dg_state_alloc: upload a form first");
    DG_TRY(hipSetDevice(c->device));
    dg_state* st = new dg_state;
    hipError_t e = hipMalloc(&st->d, (size_t)3 * c->K * c->nn * sizeof(double));
    if (e != hipSuccess) { delete st; g_err = hipGetErrorString(e); return 1; }
    *s = st;
    return 0;
}

int dg_state_free(dg_ctx* c, dg_state* s) {
    DG_LOCK(c);
    if (!s) return 0;
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    if (s->owned) DG_TRY(hipFree(s->d));
    delete s;
    return 0;
}

int dg_state_upload(dg_ctx* c, dg_state* s, const double* h, const double* hu, const double* hv) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && h && hu && hv, "dg_state_upload: null argument");
    size_t nt = (size_t)c->K * c->nn;
    DG_TRY(hipMemcpyAsync(s->d, h, nt * 8, hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipMemcpyAsync(s->d + nt, hu, nt * 8, hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipMemcpyAsync(s->d + 2 * nt, hv, nt * 8, hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_state_download(dg_ctx* c, const dg_state* s, double* h, double* hu, double* hv) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && h && hu && hv, "dg_state_download: null argument");
    size_t nt = (size_t)c->K * c->nn;
    DG_TRY(hipMemcpyAsync(h, s->d, nt * 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipMemcpyAsync(hu, s->d + nt, nt * 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipMemcpyAsync(hv, s->d + 2 * nt, nt * 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_state_copy(dg_ctx* c, dg_state* dst, const dg_state* src) {
    DG_LOCK(c);
    DG_REQUIRE(c && dst && src, "dg_state_copy: null argument");
    DG_TRY(hipMemcpyAsync(dst->d, src->d, (size_t)3 * c->K * c->nn * 8, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

int dg_rhs(dg_ctx* c, const dg_state* in, double t, dg_state* out) {
    DG_LOCK(c);
    DG_REQUIRE(c && in && out && c->K > 0, "dg_rhs: null argument or no form");
    DG_REQUIRE(in != out, "dg_rhs: in and out must differ");
    DG_TRY(hipSetDevice(c->device));
    if (int r = launch_stage(c, in->d, in->d, out->d, 0.0, 0.0, 0.0, t, 1)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_face_fluxes(dg_ctx* c, const dg_state* in, double t, double* out) {
    DG_LOCK(c);
    DG_REQUIRE(c && in && out && c->K > 0, "dg_face_fluxes: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    size_t n = (size_t)c->K * 4 * (c->order + 1) * 3;
    double* d = nullptr;
    DG_TRY(hipMalloc(&d, n * sizeof(double)));
    if (int r = ensure_work(c)) return r;
    if (int r = launch_stage(c, in->d, in->d, c->work[0]->d, 0.0, 0.0, 0.0, t, 1, d)) return r;
    DG_TRY(hipMemcpyAsync(out, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    DG_TRY(hipFree(d));
    return 0;
}

int dg_stage(dg_ctx* c, const dg_state* base, const dg_state* input, dg_state* out, const dg_stage_desc* st) {
    DG_LOCK(c);
    DG_REQUIRE(c && base && input && out && st && c->K > 0, "dg_stage: null argument or no form");
    DG_REQUIRE(out != input && out != base, "dg_stage: out must differ from base and input");
    DG_TRY(hipSetDevice(c->device));
    if (int r = launch_stage(c, base->d, input->d, out->d, st->a, st->b, st->c, st->time, 0)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Local time stepping (LocalTimeStepping, src/time/multirate.rs:296-391, as
// SWEPhysics2D implements it, src/physics/builder.rs:546-783)
// ---------------------------------------------------------------------------
int dg_begin_coarse_step(dg_ctx* c, const uint8_t* level, const uint8_t* rates, int n_stages) {
    DG_LOCK(c);
    DG_REQUIRE(c && level && rates && c->K > 0, "dg_begin_coarse_step: null argument or no form");
    DG_REQUIRE(n_stages == 3 || n_stages == 4, "dg_begin_coarse_step: 3 or 4 stages");
    const int64_t K = c->K;
    const int nkey = n_stages + 1;
    // unchanged levels and rates keep their lists (gpu/src/swe2d.rs:3285)
    std::vector<uint8_t> keys((size_t)nkey * K);
    std::memcpy(keys.data(), rates, (size_t)n_stages * K);
    std::memcpy(keys.data() + (size_t)n_stages * K, level, (size_t)K);
    if (c->lts_order && c->lts_stages == n_stages && keys == c->lts_keys) return 0;
    int finest = 0;
    for (uint8_t v : keys) finest = std::max(finest, (int)v);
    DG_REQUIRE(finest <= DG_MAX_LEVELS, "dg_begin_coarse_step: more than DG_MAX_LEVELS levels");
    {
        namespace mr = dgswe::multirate;
        const int parts = mr::host_parts((size_t)K);
        std::vector<uint8_t> bad((size_t)parts, 0);
        mr::host_parallel((size_t)K, parts, [&](size_t b, size_t e, int q) {
            for (size_t k = b; k < e; ++k)
                for (int s = 0; s < n_stages; ++s)
                    if (rates[(size_t)s * K + k] < level[k]) bad[(size_t)q] = 1;
        });
        for (uint8_t x : bad) DG_REQUIRE(!x, "dg_begin_coarse_step: a rate below its element's level");
    }
    DG_TRY(hipSetDevice(c->device));
    if (!c->lts_order || c->lts_stages != n_stages) {
        if (c->lts_order) DG_TRY(hipFree(c->lts_order));
        if (!c->lts_level) DG_TRY(hipMalloc(&c->lts_level, (size_t)K));
        if (!c->lts_lrate) DG_TRY(hipMalloc(&c->lts_lrate, (size_t)K));
        DG_TRY(hipMalloc(&c->lts_order, (size_t)nkey * K * sizeof(uint32_t)));
    }
    // per key: elements by key, largest first, element order within a key
    // (counting sort), and how many have a key of at least d
    std::vector<uint32_t> order((size_t)nkey * K);
    std::vector<int64_t> count((size_t)nkey * (DG_MAX_LEVELS + 1), 0);
    // one host thread per key
    dgswe::multirate::host_parallel((size_t)nkey, nkey, [&](size_t q0, size_t q1, int) {
        for (int q = (int)q0; q < (int)q1; ++q) {
            const uint8_t* key = keys.data() + (size_t)q * K;
            int64_t hist[DG_MAX_LEVELS + 2] = {0};
            for (int64_t k = 0; k < K; ++k) ++hist[key[k]];
            int64_t at_least[DG_MAX_LEVELS + 2] = {0};
            for (int d = DG_MAX_LEVELS; d >= 0; --d) at_least[d] = at_least[d + 1] + hist[d];
            int64_t next[DG_MAX_LEVELS + 1];
            for (int v = 0; v <= DG_MAX_LEVELS; ++v) next[v] = at_least[v + 1];
            uint32_t* o = order.data() + (size_t)q * K;
            for (int64_t k = 0; k < K; ++k) o[next[key[k]]++] = (uint32_t)k;
            for (int d = 0; d <= DG_MAX_LEVELS; ++d) count[(size_t)q * (DG_MAX_LEVELS + 1) + d] = at_least[d];
        }
    });
    DG_TRY(hipMemcpyAsync(c->lts_order, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipMemcpyAsync(c->lts_level, level, (size_t)K, hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipMemcpyAsync(c->lts_lrate, rates + (size_t)(n_stages - 1) * K, (size_t)K, hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    c->lts_keys.swap(keys);
    c->lts_count.swap(count);
    ++c->lts_generation;
    c->lts_stages = n_stages;
    c->lts_finest = finest;
    return 0;
}

int dg_stage_where(dg_ctx* c, const dg_state* base, const dg_state* input, const dg_stage_plan* plan, dg_state* out,
                   dg_state* acc) {
    DG_LOCK(c);
    DG_REQUIRE(c && base && input && plan && out && c->K > 0, "dg_stage_where: null argument or no form");
    DG_REQUIRE(c->lts_order, "dg_stage_where: no coarse step begun (dg_begin_coarse_step)");
    DG_REQUIRE(plan->stage >= 0 && plan->stage < c->lts_stages, "dg_stage_where: stage out of range");
    DG_REQUIRE(plan->finest >= c->lts_finest && plan->finest <= DG_MAX_LEVELS,
               "dg_stage_where: the plan's finest level is not the coarse step's");
    DG_REQUIRE(plan->depth >= 0 && plan->depth <= plan->finest, "dg_stage_where: depth out of range");
    DG_REQUIRE(plan->substep >= 0 && plan->substep < (1 << plan->finest), "dg_stage_where: substep out of range");
    DG_REQUIRE(!plan->last || acc, "dg_stage_where: an accumulating stage needs the sum buffer");
    DG_REQUIRE(out != input && out != base && acc != input && acc != base && (!acc || acc != out),
               "dg_stage_where: out and acc must differ from base and input");
    const int64_t n_list = c->lts_count[(size_t)plan->stage * (DG_MAX_LEVELS + 1) + plan->depth];
    if (n_list == 0) return 0;
    DG_TRY(hipSetDevice(c->device));
    size_t nt = (size_t)c->K * c->nn;
    StageParams p{};
    std::memset(&p, 0, sizeof(p));
    p.bh = base->d; p.bhu = base->d + nt; p.bhv = base->d + 2 * nt;
    p.ih = input->d; p.ihu = input->d + nt; p.ihv = input->d + 2 * nt;
    p.oh = out->d; p.ohu = out->d + nt; p.ohv = out->d + 2 * nt;
    p.bed = c->bed; p.corners = c->corners; p.manning = c->manning;
    p.links = c->links; p.tags = c->tags; p.affine = c->affine;
    p.n_elements = c->K;
    p.a = plan->start; p.b = plan->input; p.c = 0.0; p.eta_e = 0.0;
    p.g = c->g; p.h_min = c->h_min; p.h_dry = c->h_dry; p.u_max = c->u_max; p.tau = c->tau; p.f0 = c->f0;
    p.sqrt_g = std::sqrt(c->g);
    p.flags = c->flags;
    p.clips = c->clips;
    p.geo = c->geo;
    p.fgeo = c->fgeo;
    p.queue = c->queue;
    p.queue_count = c->queue_count;
    // a pass over every element goes in mesh order (no list: contiguous blocks,
    // each XCD's share of the mesh in its own L2)
    p.elist = n_list == c->K ? nullptr : c->lts_order + (size_t)plan->stage * c->K;
    p.n_list = n_list;
    p.level = c->lts_level;
    p.lrate = c->lts_lrate;
    p.depth = plan->depth;
    if (plan->last) { p.ah = acc->d; p.ahu = acc->d + nt; p.ahv = acc->d + 2 * nt; }
    // element_stage (multirate.rs:169-200) of every level: its substep's start,
    // its stage time and c = beta dt_l; the tide at that time
    for (int l = 0; l <= DG_MAX_LEVELS; ++l) {
        const int lc = std::min(l, plan->finest);
        const double dt_l = plan->dt / (double)(1ull << lc);
        const uint64_t q = (uint64_t)plan->substep >> (plan->finest - lc);
        const double t_k = plan->t + (double)q * dt_l;
        const double time = t_k + plan->c * dt_l;
        p.lvl_c[l] = plan->beta * dt_l;
        p.lvl_eta[l] = tide_elevation(c->tide, c->tide_nodal, time);
    }
    // the forcing given per node: its tables, and per level the moments at the level's stage time
    const bool force = c->tide_slot || c->nest_slot || c->wx_series || c->manning_node || c->coriolis_node;
    p.manning_node = c->manning_node;
    p.coriolis_node = c->coriolis_node;
    if (c->tide_slot || c->nest_slot || c->wx_series) {
        // The per-level moment rows go up through a ring of pinned host slots, reused after
        // their upload's event: neither the event wait nor a copy from a slot that is written
        // again before a replay belongs in a captured graph
        hipStreamCaptureStatus capturing = hipStreamCaptureStatusNone;
        DG_TRY(hipStreamIsCapturing(c->stream, &capturing));
        DG_REQUIRE(capturing == hipStreamCaptureStatusNone,
                   "dg_stage_where: a stage forced per node (boundary tides, nesting or weather) cannot be "
                   "captured into a graph");
        constexpr int ROW = LVL_MOMENT, ROWS = DG_MAX_LEVELS + 1;
        if (!c->lvl_moment_dev) {
            DG_TRY(hipMalloc(&c->lvl_moment_dev, sizeof(double) * dg_ctx::MOMENT_RING * ROWS * ROW));
            DG_TRY(hipHostMalloc(&c->lvl_moment_host, sizeof(double) * dg_ctx::MOMENT_RING * ROWS * ROW));
            for (int i = 0; i < dg_ctx::MOMENT_RING; ++i)
                DG_TRY(hipEventCreateWithFlags(&c->lvl_moment_ev[i], hipEventDisableTiming));
        }
        const int slot = c->lvl_moment_next;
        c->lvl_moment_next = (slot + 1) % dg_ctx::MOMENT_RING;
        if (c->lvl_moment_used[slot]) DG_TRY(hipEventSynchronize(c->lvl_moment_ev[slot]));
        double* rows = c->lvl_moment_host + (size_t)slot * ROWS * ROW;
        for (int l = 0; l <= DG_MAX_LEVELS; ++l) {
            const int lc = std::min(l, plan->finest);
            const double dt_l = plan->dt / (double)(1ull << lc);
            const uint64_t q = (uint64_t)plan->substep >> (plan->finest - lc);
            const double time = plan->t + (double)q * dt_l + plan->c * dt_l;
            StageParams m{};
            set_atlas(c, m, time);
            if (int r = set_nesting(c, m, time)) return r;
            if (int r = set_weather(c, m, time)) return r;
            double* row = rows + (size_t)l * ROW;
            std::fill(row, row + ROW, 0.0);
            for (int j = 0; j < 1 + 3 * ATLAS_MAX_CONSTITUENTS; ++j) row[j] = m.atlas_moment[j];
            row[LM_NEST] = m.nest_ramp;
            row[LM_NEST + 2] = m.nest_len;
            row[LM_WX] = m.wx_wind;
            row[LM_WX + 1] = m.wx_pres;
            row[LM_WX + 2] = m.wx_len;
            for (int j = 0; j < 4; ++j) {
                row[LM_NEST + 3 + j] = m.nest_idx[j];
                row[LM_NEST + 7 + j] = m.nest_w[j];
                row[LM_WX + 3 + j] = m.wx_idx[j];
                row[LM_WX + 7 + j] = m.wx_w[j];
            }
            if (l == 0) {
                // the tables and constants, the same at every level
                set_atlas(c, p, time);
                if (int r = set_nesting(c, p, time)) return r;
                if (int r = set_weather(c, p, time)) return r;
            }
        }
        double* dev = c->lvl_moment_dev + (size_t)slot * ROWS * ROW;
        DG_TRY(hipMemcpyAsync(dev, rows, sizeof(double) * ROWS * ROW, hipMemcpyHostToDevice, c->stream));
        DG_TRY(hipEventRecord(c->lvl_moment_ev[slot], c->stream));
        c->lvl_moment_used[slot] = true;
        p.lvl_moment = dev;
    }
    const int epb = 64 / c->nn;
    const unsigned int nblk = (unsigned int)((n_list + epb - 1) / epb);
    const unsigned int nchunk = (nblk + 1023) / 1024;
    p.xcd_stride = (nblk + 7) / 8;
    // (most passes list no subcell element: shallow elements sit on coarse levels; a
    // subcell grid of 2048 workgroups measured best of 2048 / 4096 / 16384 on the
    // 400,000-element channel, by the wall time of the run)
    dim3 grid(8 * p.xcd_stride), sub_grid(std::min(nblk, 2048u));
    switch (c->order) {
        case 1: hipLaunchKernelGGL((force ? wet_lts_kernel<1, true> : wet_lts_kernel<1>), grid, dim3(64), 0, c->stream, p); break;
        case 2: hipLaunchKernelGGL((force ? wet_lts_kernel<2, true> : wet_lts_kernel<2>), grid, dim3(64), 0, c->stream, p); break;
        default: hipLaunchKernelGGL((force ? wet_lts_kernel<3, true> : wet_lts_kernel<3>), grid, dim3(64), 0, c->stream, p); break;
    }
    // (one chunk: the compaction's own scan is the whole offset, no chunk totals)
    if (nchunk > 1)
        hipLaunchKernelGGL(chunk_count_kernel, dim3(nchunk), dim3(256), 0, c->stream, c->queue_count, nblk, c->chunk_total);
    hipLaunchKernelGGL(compact_kernel, dim3(nchunk), dim3(256), 0, c->stream, c->queue_count, c->queue, nblk,
                       c->chunk_total, epb, c->sub_list, c->sub_count);
    switch (c->order) {
        case 1: hipLaunchKernelGGL((force ? subcell_stage_kernel<1, true, true> : subcell_stage_kernel<1, true>), sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
        case 2: hipLaunchKernelGGL((force ? subcell_stage_kernel<2, true, true> : subcell_stage_kernel<2, true>), sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
        default: hipLaunchKernelGGL((force ? subcell_stage_kernel<3, true, true> : subcell_stage_kernel<3, true>), sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
    }
    DG_TRY(hipGetLastError());
    return 0;
}

int dg_finish_where(dg_ctx* c, dg_state* state, const dg_state* acc, int depth) {
    DG_LOCK(c);
    DG_REQUIRE(c && state && acc && c->K > 0, "dg_finish_where: null argument or no form");
    DG_REQUIRE(c->lts_order, "dg_finish_where: no coarse step begun (dg_begin_coarse_step)");
    DG_REQUIRE(depth >= 0 && depth <= DG_MAX_LEVELS, "dg_finish_where: depth out of range");
    DG_REQUIRE(state != acc, "dg_finish_where: state and acc must differ");
    const int64_t n_list = c->lts_count[(size_t)c->lts_stages * (DG_MAX_LEVELS + 1) + depth];
    if (n_list == 0) return 0;
    DG_TRY(hipSetDevice(c->device));
    size_t nt = (size_t)c->K * c->nn;
    FinishParams p;
    p.elist = n_list == c->K ? nullptr : c->lts_order + (size_t)c->lts_stages * c->K;
    p.n_list = n_list;
    p.ah = acc->d; p.ahu = acc->d + nt; p.ahv = acc->d + 2 * nt;
    p.h = state->d; p.hu = state->d + nt; p.hv = state->d + 2 * nt;
    p.geo = c->geo;
    p.n_elements = c->K;
    p.h_dry = c->h_dry; p.u_max = c->u_max;
    p.wet_dry = (c->flags & DG_WET_DRY) ? 1 : 0;
    p.clips = c->clips;
    const int epb = 64 / c->nn;
    dim3 grid((unsigned int)((n_list + epb - 1) / epb));
    switch (c->order) {
        case 1: hipLaunchKernelGGL(finish_kernel<1>, grid, dim3(64), 0, c->stream, p); break;
        case 2: hipLaunchKernelGGL(finish_kernel<2>, grid, dim3(64), 0, c->stream, p); break;
        default: hipLaunchKernelGGL(finish_kernel<3>, grid, dim3(64), 0, c->stream, p); break;
    }
    DG_TRY(hipGetLastError());
    return 0;
}

// The depth flags into c->depth_flags, not waited for
static int launch_depth_flags(dg_ctx* c, const double* h, double d0, double d1, double d2) {
    if (!c->depth_flags) DG_TRY(hipMalloc(&c->depth_flags, (size_t)c->K));
    dim3 grid((unsigned)((c->K + 255) / 256));
    switch (c->order) {
        case 1: hipLaunchKernelGGL(depth_flags_kernel<1>, grid, dim3(256), 0, c->stream, h, c->K, d0, d1, d2, c->depth_flags); break;
        case 2: hipLaunchKernelGGL(depth_flags_kernel<2>, grid, dim3(256), 0, c->stream, h, c->K, d0, d1, d2, c->depth_flags); break;
        default: hipLaunchKernelGGL(depth_flags_kernel<3>, grid, dim3(256), 0, c->stream, h, c->K, d0, d1, d2, c->depth_flags); break;
    }
    DG_TRY(hipGetLastError());
    return 0;
}

int dg_depth_flags(dg_ctx* c, const dg_state* s, double d0, double d1, double d2, uint8_t* out) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && out && c->K > 0, "dg_depth_flags: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    if (int r = launch_depth_flags(c, s->d, d0, d1, d2)) return r;
    DG_TRY(hipMemcpyAsync(out, c->depth_flags, (size_t)c->K, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// The multirate stepper (MultirateStepper, multirate.rs:713-1330) over the
// passes above, on the host side of the library
// ---------------------------------------------------------------------------
struct dg_multirate {
    dg_ctx* c = nullptr;
    int scheme = 0;
    dgswe::multirate::Stage st[4];
    dgswe::multirate::Assigner as;
    // the device half of an assignment (mr_*_kernel): the extremes' slots, the
    // plan, the levels and line-local flags, the kept assignment's, and the
    // differs / largest-level pair; pinned host copies of the plan and the pair
    unsigned long long* d_slots = nullptr;
    dgswe::multirate::Plan* d_plan = nullptr;
    uint8_t *d_level = nullptr, *d_ll = nullptr, *d_memo_given = nullptr, *d_memo_ll = nullptr;
    int* d_out = nullptr;
    dgswe::multirate::Plan* h_plan = nullptr;
    int* h_out = nullptr;
    dg_state* values[3] = {nullptr, nullptr, nullptr};
    dg_state *last = nullptr, *acc = nullptr;
    uint64_t generation = ~0ull;   // the context's list generation this stepper made
    dg_multirate_stats stats{};
};

static int mr_free_states(dg_multirate* m) {
    dg_state* all[5] = {m->values[0], m->values[1], m->values[2], m->last, m->acc};
    for (dg_state* x : all)
        if (x) { if (int r = dg_state_free(m->c, x)) return r; }
    m->values[0] = m->values[1] = m->values[2] = m->last = m->acc = nullptr;
    return 0;
}

static int mr_free_device(dg_multirate* m) {
    void* dev[] = {m->d_slots, m->d_plan, m->d_level, m->d_ll, m->d_memo_given, m->d_memo_ll, m->d_out};
    for (void* p : dev)
        if (p) DG_TRY(hipFree(p));
    if (m->h_plan) DG_TRY(hipHostFree(m->h_plan));
    if (m->h_out) DG_TRY(hipHostFree(m->h_out));
    m->d_slots = nullptr; m->d_plan = nullptr; m->d_out = nullptr;
    m->d_level = m->d_ll = m->d_memo_given = m->d_memo_ll = nullptr;
    m->h_plan = nullptr; m->h_out = nullptr;
    return 0;
}

static int mr_alloc(dg_multirate* m, size_t K) {
    DG_TRY(hipMalloc(&m->d_slots, 4 * sizeof(unsigned long long)));
    DG_TRY(hipMalloc(&m->d_plan, sizeof(dgswe::multirate::Plan)));
    DG_TRY(hipMalloc(&m->d_out, 2 * sizeof(int)));
    for (uint8_t** p : {&m->d_level, &m->d_ll, &m->d_memo_given, &m->d_memo_ll}) DG_TRY(hipMalloc(p, K));
    DG_TRY(hipHostMalloc((void**)&m->h_plan, sizeof(dgswe::multirate::Plan), hipHostMallocDefault));
    DG_TRY(hipHostMalloc((void**)&m->h_out, 2 * sizeof(int), hipHostMallocDefault));
    return 0;
}

int dg_multirate_create(dg_ctx* c, int max_levels, int scheme, int line_rates, dg_multirate** out) {
    DG_LOCK(c);
    DG_REQUIRE(c && out && c->K > 0, "dg_multirate_create: null argument or no form");
    DG_REQUIRE(max_levels >= 0 && max_levels <= DG_MAX_LEVELS, "dg_multirate_create: 0 to DG_MAX_LEVELS levels");
    DG_REQUIRE(scheme == DG_SSP_RK3 || scheme == DG_SSP_RK43, "dg_multirate_create: unknown scheme");
    DG_TRY(hipSetDevice(c->device));
    const size_t K = (size_t)c->K;
    std::vector<uint32_t> links(4 * K);
    DG_TRY(hipMemcpy(links.data(), c->links, 4 * K * sizeof(uint32_t), hipMemcpyDeviceToHost));
    dg_multirate* m = new dg_multirate;
    m->c = c;
    m->scheme = scheme;
    m->as.max_levels = max_levels;
    m->as.line_rates = line_rates != 0;
    m->as.n1 = c->order + 1;
    m->as.n_stages = dgswe::multirate::stages(scheme == DG_SSP_RK3 ? 0 : 1, m->st);
    m->as.init(links.data(), K);
    if (int r = mr_alloc(m, K)) {
        (void)mr_free_device(m);
        delete m;
        return r;
    }
    *out = m;
    return 0;
}

int dg_multirate_destroy(dg_multirate* m) {
    if (!m) return 0;
    DG_LOCK(m->c);
    DG_TRY(hipSetDevice(m->c->device));
    if (int r = mr_free_states(m)) return r;
    if (int r = mr_free_device(m)) return r;
    delete m;
    return 0;
}

int dg_multirate_assign_levels(dg_multirate* m, const dg_state* s, double cfl, double dt_cap, double landing,
                               double* dt_out, double* stable_out) {
    DG_REQUIRE(m != nullptr, "dg_multirate_assign_levels: null stepper");
    DG_LOCK(m->c);
    namespace mr = dgswe::multirate;
    DG_REQUIRE(m && s && dt_out, "dg_multirate_assign_levels: null argument");
    dg_ctx* c = m->c;
    const size_t K = (size_t)c->K;
    DG_TRY(hipSetDevice(c->device));
    // LocalTimeStepping::element_dt on the device, and the depth flags of the
    // coarse-step bound and the line-local elements; then the extremes, the
    // plan and the levels there too, with whether they differ from the kept
    // assignment's: only the plan and that pair come back
    if (int r = launch_dt_kernels(c, s->d, cfl, m->scheme, true)) return r;
    if (int r = launch_depth_flags(c, s->d, mr::SUBCELL_CAP_DEPTH_FACTOR * c->h_dry, c->h_dry,
                                   c->h_dry + mr::LINE_LOCAL_HEADROOM)) return r;
    // few workgroups for the extremes (each folds its minimum and maxima in with
    // atomics on four addresses), up to 1024 for the levels
    const unsigned egrid = (unsigned)std::min<size_t>((K + 4095) / 4096, 128);
    const unsigned rgrid = (unsigned)std::min<size_t>((K + 255) / 256, 1024);
    const bool lines = m->as.line_rates && m->as.n1 >= 2;
    const bool kept = m->as.memo;
    hipLaunchKernelGGL(mr::mr_clear_kernel, dim3(1), dim3(64), 0, c->stream, m->d_slots, m->d_out);
    hipLaunchKernelGGL(mr::mr_extremes_kernel, dim3(egrid), dim3(256), 0, c->stream, c->element_dt, c->depth_flags,
                       c->K, m->d_slots);
    hipLaunchKernelGGL(mr::mr_plan_kernel, dim3(1), dim3(64), 0, c->stream, m->d_slots, (c->flags & DG_WET_DRY) ? 1 : 0,
                       dt_cap, landing, m->as.max_levels, m->d_plan);
    hipLaunchKernelGGL(mr::mr_levels_kernel, dim3(rgrid), dim3(256), 0, c->stream, c->element_dt, c->depth_flags, c->K,
                       m->d_plan, lines ? 1 : 0, kept ? m->d_memo_given : nullptr, kept ? m->d_memo_ll : nullptr,
                       m->d_level, m->d_ll, m->d_out);
    DG_TRY(hipGetLastError());
    DG_TRY(hipMemcpyAsync(m->h_plan, m->d_plan, sizeof(mr::Plan), hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipMemcpyAsync(m->h_out, m->d_out, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    const mr::Plan p = *m->h_plan;
    if (!m->as.device_kept(p, m->h_out[0], m->h_out[1])) {
        m->as.level.resize(K);
        m->as.ll.resize(K);
        DG_TRY(hipMemcpyAsync(m->as.level.data(), m->d_level, K, hipMemcpyDeviceToHost, c->stream));
        DG_TRY(hipMemcpyAsync(m->as.ll.data(), m->d_ll, K, hipMemcpyDeviceToHost, c->stream));
        DG_TRY(hipStreamSynchronize(c->stream));
        m->as.apply(m->as.finest);
        // the kept assignment's inputs for the next comparison on the device
        DG_TRY(hipMemcpyAsync(m->d_memo_given, m->as.memo_given.data(), K, hipMemcpyHostToDevice, c->stream));
        DG_TRY(hipMemcpyAsync(m->d_memo_ll, m->as.memo_ll.data(), K, hipMemcpyHostToDevice, c->stream));
        DG_TRY(hipStreamSynchronize(c->stream));
    }
    *dt_out = p.dt;
    if (stable_out) *stable_out = p.stable;
    return 0;
}

int dg_multirate_step(dg_multirate* m, dg_state* s, double t, double dt) {
    DG_REQUIRE(m != nullptr, "dg_multirate_step: null stepper");
    DG_LOCK(m->c);
    DG_REQUIRE(m && s, "dg_multirate_step: null argument");
    DG_REQUIRE(!m->as.level.empty(), "dg_multirate_step: no levels assigned (dg_multirate_assign_levels)");
    dg_ctx* c = m->c;
    const int ns = m->as.n_stages;
    DG_TRY(hipSetDevice(c->device));
    if (!m->acc) {
        for (int q = 0; q < ns - 1; ++q)
            if (int r = dg_state_alloc(c, &m->values[q])) return r;
        if (int r = dg_state_alloc(c, &m->last)) return r;
        if (int r = dg_state_alloc(c, &m->acc)) return r;
    }
    // the passes' lists are the last ones while the levels and rates are (and
    // nothing else has begun a coarse step on the context since)
    if (m->as.changed || m->generation != c->lts_generation) {
        if (int r = dg_begin_coarse_step(c, m->as.level.data(), m->as.rates.data(), ns)) return r;
        m->as.changed = false;
        m->generation = c->lts_generation;
    }
    const int finest = m->as.finest;
    for (int p = 0; p < (1 << finest); ++p) {
        const int d = dgswe::multirate::depth(p, finest);
        for (int q = 0; q < ns; ++q) {
            const bool last = q + 1 == ns;
            dg_stage_plan plan;
            plan.t = t; plan.dt = dt; plan.finest = finest; plan.substep = p; plan.depth = d; plan.stage = q;
            plan.start = m->st[q].start; plan.input = m->st[q].input; plan.beta = m->st[q].beta; plan.c = m->st[q].c;
            plan.last = last ? 1 : 0;
            const dg_state* input = q == 0 ? s : m->values[q - 1];
            if (int r = dg_stage_where(c, s, input, &plan, last ? m->last : m->values[q], last ? m->acc : nullptr)) return r;
        }
        if (int r = dg_finish_where(c, s, m->acc, dgswe::multirate::depth(p + 1, finest))) return r;
    }
    m->stats.steps += 1;
    m->stats.element_evaluations += m->as.evals_per_step;
    m->stats.global_evaluations += (int64_t)ns * c->K * (int64_t(1) << finest);
    m->stats.finest_level = std::max(m->stats.finest_level, finest);
    m->stats.model_seconds += dt;
    return 0;
}

int dg_multirate_levels(dg_multirate* m, uint8_t* level, uint8_t* rates, int* finest) {
    DG_REQUIRE(m != nullptr, "dg_multirate_levels: null stepper");
    DG_LOCK(m->c);
    DG_REQUIRE(m, "dg_multirate_levels: null argument");
    if (level && !m->as.level.empty()) std::memcpy(level, m->as.level.data(), m->as.level.size());
    if (rates && !m->as.rates.empty()) std::memcpy(rates, m->as.rates.data(), m->as.rates.size());
    if (finest) *finest = m->as.finest;
    return 0;
}

int dg_multirate_stats_read(dg_multirate* m, dg_multirate_stats* out) {
    DG_REQUIRE(m != nullptr, "dg_multirate_stats_read: null stepper");
    DG_LOCK(m->c);
    DG_REQUIRE(m && out, "dg_multirate_stats_read: null argument");
    *out = m->stats;
    return 0;
}

int dg_multirate_run(dg_multirate* m, dg_state* s, double t, double t_end, double cfl, double dt_max, double interval,
                     double* t_out, int64_t* coarse_steps) {
    DG_REQUIRE(m != nullptr, "dg_multirate_run: null stepper");
    DG_LOCK(m->c);
    namespace mr = dgswe::multirate;
    DG_REQUIRE(m && s, "dg_multirate_run: null argument");
    int64_t n = 0, k = 0;
    const double t_start = t;
    while (t < t_end) {
        const double target = interval > 0.0 ? std::fmin(t_start + (double)(k + 1) * interval, t_end) : t_end;
        double fitted = 0.0, dt = 0.0;
        if (int r = dg_multirate_assign_levels(m, s, cfl, dt_max, target - t, &fitted, &dt)) return r;
        if (dt_max > 0.0) dt = std::fmin(dt, dt_max);
        dt = std::fmin(dt, fitted);
        const bool landed = t + dt * (1.0 + mr::LANDING_SLACK) >= target;
        if (landed) dt = target - t;
        if (int r = dg_multirate_step(m, s, t, dt)) return r;
        t = landed ? target : t + dt;
        ++n;
        if (interval > 0.0 && t >= t_start + (double)(k + 1) * interval) ++k;
    }
    DG_TRY(hipStreamSynchronize(m->c->stream));
    if (t_out) *t_out = t;
    if (coarse_steps) *coarse_steps = n;
    return 0;
}

int dg_set_boundary_tides(dg_ctx* c, int n_c, const double* omega, const double* f_rate, double ramp, int velocity,
                          double max_ratio, int64_t n_slots, const uint32_t* slot_of_face_node, const double* coef,
                          const double* mean, const double* source_depth) {
    DG_LOCK(c);
    DG_REQUIRE(c && c->K > 0, "dg_set_boundary_tides: null context or no form");
    DG_REQUIRE(n_c >= 0 && n_c <= ATLAS_MAX_CONSTITUENTS, "dg_set_boundary_tides: 0 to 21 constituents");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    for (void* ptr : {(void*)c->tide_slot, (void*)c->tide_coef, (void*)c->tide_mean, (void*)c->tide_src})
        if (ptr) DG_TRY(hipFree(ptr));
    c->tide_slot = nullptr;
    c->tide_coef = c->tide_mean = c->tide_src = nullptr;
    c->atlas_nc = 0;
    if (n_c == 0) return 0;
    DG_REQUIRE(omega && f_rate && slot_of_face_node && coef && mean && source_depth && n_slots > 0,
               "dg_set_boundary_tides: missing tables");
    const size_t nf = (size_t)c->K * 4 * (c->order + 1);
    std::vector<uint8_t> tags((size_t)c->K * 4);
    DG_TRY(hipMemcpy(tags.data(), c->tags, tags.size(), hipMemcpyDeviceToHost));
    std::vector<uint32_t> links((size_t)c->K * 4);
    DG_TRY(hipMemcpy(links.data(), c->links, links.size() * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nf; ++i) {
        const uint32_t sl = slot_of_face_node[i];
        if (sl == 0xFFFFFFFFu) continue;
        const size_t kf = i / (size_t)(c->order + 1);
        DG_REQUIRE((int64_t)sl < n_slots, "dg_set_boundary_tides: a slot out of range");
        DG_REQUIRE(links[kf] == LINK_BOUNDARY && tags[kf] == DG_BC_TIDE,
                   "dg_set_boundary_tides: a slot on a face that is not an open boundary");
    }
    DG_TRY(hipMalloc(&c->tide_slot, nf * sizeof(uint32_t)));
    DG_TRY(hipMalloc(&c->tide_coef, (size_t)n_slots * n_c * 6 * sizeof(double)));
    DG_TRY(hipMalloc(&c->tide_mean, (size_t)n_slots * sizeof(double)));
    DG_TRY(hipMalloc(&c->tide_src, (size_t)n_slots * sizeof(double)));
    DG_TRY(hipMemcpy(c->tide_slot, slot_of_face_node, nf * sizeof(uint32_t), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->tide_coef, coef, (size_t)n_slots * n_c * 6 * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->tide_mean, mean, (size_t)n_slots * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->tide_src, source_depth, (size_t)n_slots * sizeof(double), hipMemcpyHostToDevice));
    c->atlas_nc = n_c;
    for (int j = 0; j < n_c; ++j) { c->atlas_omega[j] = omega[j]; c->atlas_f_rate[j] = f_rate[j]; }
    c->atlas_ramp = ramp;
    c->atlas_velocity = velocity ? 1 : 0;
    c->atlas_max_ratio = max_ratio;
    return 0;
}

static int check_parent_covers_boundary(dg_ctx* c);

int dg_set_nesting(dg_ctx* c, int64_t n_slots, const uint32_t* slot_of_node, int n_times, const double* times,
                   double epoch, int cubic, const float* series, const double* weight, const double* parent_depth, double timescale,
                   double wet_depth, int momentum_only, int velocity, double max_ratio, double reference_level,
                   double ramp) {
    DG_LOCK(c);
    DG_REQUIRE(c && c->K > 0, "dg_set_nesting: null context or no form");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    for (void* ptr : {(void*)c->nest_slot, (void*)c->nest_series, (void*)c->nest_weight, (void*)c->nest_depth})
        if (ptr) DG_TRY(hipFree(ptr));
    c->nest_slot = nullptr;
    c->nest_series = nullptr;
    c->nest_weight = c->nest_depth = nullptr;
    c->nest_times.clear();
    c->h_nest_slot.clear();
    if (n_slots == 0) {
        c->nest_boundary = 0;
        return 0;
    }
    DG_REQUIRE(slot_of_node && times && series && weight && parent_depth && n_slots > 0 && n_times >= 1,
               "dg_set_nesting: missing tables");
    DG_REQUIRE(timescale > 0.0, "dg_set_nesting: the relaxation timescale must be positive");
    for (int i = 1; i < n_times; ++i) DG_REQUIRE(times[i] > times[i - 1], "dg_set_nesting: snapshot times must increase");
    const size_t nt = (size_t)c->K * c->nn;
    for (size_t i = 0; i < nt; ++i)
        DG_REQUIRE(slot_of_node[i] == 0xFFFFFFFFu || (int64_t)slot_of_node[i] < n_slots, "dg_set_nesting: a slot out of range");
    DG_TRY(hipMalloc(&c->nest_slot, nt * sizeof(uint32_t)));
    DG_TRY(hipMalloc(&c->nest_series, (size_t)n_slots * n_times * 3 * sizeof(float)));
    DG_TRY(hipMalloc(&c->nest_weight, (size_t)n_slots * sizeof(double)));
    DG_TRY(hipMalloc(&c->nest_depth, (size_t)n_slots * sizeof(double)));
    DG_TRY(hipMemcpy(c->nest_slot, slot_of_node, nt * sizeof(uint32_t), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->nest_series, series, (size_t)n_slots * n_times * 3 * sizeof(float), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->nest_weight, weight, (size_t)n_slots * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->nest_depth, parent_depth, (size_t)n_slots * sizeof(double), hipMemcpyHostToDevice));
    c->nest_times.assign(times, times + n_times);
    c->h_nest_slot.assign(slot_of_node, slot_of_node + nt);
    if (c->nest_boundary)
        if (int r = check_parent_covers_boundary(c)) {
            c->nest_boundary = 0;
            return r;
        }
    c->nest_epoch = epoch;
    c->nest_cubic = cubic ? 1 : 0;
    c->nest_rate = 1.0 / timescale;
    c->nest_wet_depth = wet_depth;
    c->nest_momentum_only = momentum_only ? 1 : 0;
    c->nest_velocity = velocity ? 1 : 0;
    c->nest_max_ratio = max_ratio;
    c->nest_reference = reference_level;
    c->nest_ramp = ramp;
    return 0;
}

int dg_set_weather(dg_ctx* c, int n_times, const double* times, double epoch, int cubic, const float* series,
                   int wind, int pressure, int drag, double cd, double rho_air, double rho_water, double h_min,
                   double ramp) {
    DG_LOCK(c);
    DG_REQUIRE(c && c->K > 0, "dg_set_weather: null context or no form");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    if (c->wx_series) DG_TRY(hipFree(c->wx_series));
    c->wx_series = nullptr;
    c->wx_times.clear();
    if (n_times == 0) return 0;
    DG_REQUIRE(times && series && n_times > 0, "dg_set_weather: missing snapshots");
    DG_REQUIRE(drag >= 0 && drag <= 4, "dg_set_weather: drag is 0 (constant) to 4 (Yelland & Taylor)");
    DG_REQUIRE(rho_water > 0.0, "dg_set_weather: the water density must be positive");
    for (int i = 1; i < n_times; ++i) DG_REQUIRE(times[i] > times[i - 1], "dg_set_weather: snapshot times must increase");
    const size_t bytes = (size_t)n_times * 5 * (size_t)c->K * c->nn * sizeof(float);
    DG_TRY(hipMalloc(&c->wx_series, bytes));
    DG_TRY(hipMemcpy(c->wx_series, series, bytes, hipMemcpyHostToDevice));
    c->wx_times.assign(times, times + n_times);
    c->wx_epoch = epoch;
    c->wx_cubic = cubic ? 1 : 0;
    c->wx_wind = wind ? 1 : 0;
    c->wx_pressure = pressure ? 1 : 0;
    c->wx_drag = drag;
    c->wx_cd = cd;
    c->wx_rho_air = rho_air;
    c->wx_rho_water = rho_water;
    c->wx_h_min = h_min;
    c->wx_ramp = ramp;
    if (!wind && !pressure) {
        // GriddedAtmosphere2D::add_element adds nothing with both switches off
        DG_TRY(hipFree(c->wx_series));
        c->wx_series = nullptr;
        c->wx_times.clear();
    }
    return 0;
}

int dg_set_inverse_barometer(dg_ctx* c, double factor) {
    DG_LOCK(c);
    DG_REQUIRE(c, "dg_set_inverse_barometer: null context");
    DG_REQUIRE(std::isfinite(factor), "dg_set_inverse_barometer: the factor must be finite");
    c->wx_raise = factor;
    return 0;
}

// Every open-boundary face node must have a parent slot when the parent is the
// open boundary's external state (DeviceExternal2D::Parent maps each open face node
// to one, ocean_nesting.rs:543-562): an uncovered node is refused by name here
// instead of falling back to the tide.
static int check_parent_covers_boundary(dg_ctx* c) {
    const int n1 = c->order + 1, nn = c->nn;
    std::vector<uint8_t> tags((size_t)c->K * 4);
    DG_TRY(hipMemcpy(tags.data(), c->tags, tags.size(), hipMemcpyDeviceToHost));
    std::vector<uint32_t> links((size_t)c->K * 4);
    DG_TRY(hipMemcpy(links.data(), c->links, links.size() * 4, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < c->K; ++k)
        for (int f = 0; f < 4; ++f) {
            if (links[4 * k + f] != LINK_BOUNDARY || tags[4 * k + f] != DG_BC_TIDE) continue;
            for (int fi = 0; fi < n1; ++fi) {
                const int step = (f & 1) ? fi * n1 : fi;
                const int start = f == 0 ? 0 : f == 1 ? n1 - 1 : f == 2 ? nn - 1 : (n1 - 1) * n1;
                const int node = (f & 2) ? start - step : start + step;
                if (c->h_nest_slot[(size_t)k * nn + node] == 0xFFFFFFFFu) {
                    char msg[200];
                    std::snprintf(msg, sizeof msg,
                                  "dg_set_nesting_boundary: open face node %d of face %d of element %lld has no parent "
                                  "slot (the slot map must cover the open boundary)", fi, f, (long long)k);
                    g_err = msg;
                    return 2;
                }
            }
        }
    return 0;
}

int dg_set_nesting_boundary(dg_ctx* c, int on) {
    DG_LOCK(c);
    DG_REQUIRE(c, "dg_set_nesting_boundary: null context");
    if (on) {
        DG_REQUIRE(c->nest_slot, "dg_set_nesting_boundary: no parent (dg_set_nesting first)");
        if (int r = check_parent_covers_boundary(c)) return r;
    }
    c->nest_boundary = on ? 1 : 0;
    return 0;
}

uint64_t dg_fnv1a(const uint8_t* data, size_t n, uint64_t h) {
    for (size_t i = 0; i < n; ++i) h = (h ^ (uint64_t)data[i]) * 0x100000001b3ull;
    return h;
}

int dg_element_dt(dg_ctx* c, const dg_state* s, double cfl, int scheme, double* out) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && out && c->K > 0, "dg_element_dt: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    double dt;
    if (int r = launch_dt(c, s->d, cfl, scheme, true, &dt)) return r;
    DG_TRY(hipMemcpy(out, c->element_dt, c->K * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int dg_dt_ssp(dg_ctx* c, const dg_state* s, double cfl, int scheme, double* dt) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && dt && c->K > 0, "dg_dt_ssp: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    return launch_dt(c, s->d, cfl, scheme, false, dt);
}

int dg_step(dg_ctx* c, dg_state* s, double t, double dt, int scheme) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && c->K > 0, "dg_step: null argument or no form");
    DG_REQUIRE(scheme == DG_SSP_RK3 || scheme == DG_SSP_RK43, "dg_step: unknown scheme");
    DG_TRY(hipSetDevice(c->device));
    if (int r = step_impl(c, s, t, dt, scheme)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_run(dg_ctx* c, dg_state* s, double t, int n_steps, double cfl, int scheme, double* t_end, double* dts) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && c->K > 0, "dg_run: null argument or no form");
    DG_REQUIRE(scheme == DG_SSP_RK3 || scheme == DG_SSP_RK43, "dg_run: unknown scheme");
    DG_TRY(hipSetDevice(c->device));
    for (int q = 0; q < n_steps; ++q) {
        double dt;
        if (int r = launch_dt(c, s->d, cfl, scheme, false, &dt)) return r;
        if (!(dt > 0.0) || !std::isfinite(dt)) {
            g_err = "dg_run: no finite positive time step at step " + std::to_string(q) + " (dry domain or blow-up)";
            return 3;
        }
        if (dts) dts[q] = dt;
        if (int r = step_impl(c, s, t, dt, scheme)) return r;
        t += dt;
    }
    DG_TRY(hipStreamSynchronize(c->stream));
    if (t_end) *t_end = t;
    return 0;
}

int dg_tune(dg_ctx* c, const dg_state* s) {
    DG_REQUIRE(c != nullptr, "dg_tune: null context");
    DG_LOCK(c);
    DG_REQUIRE(s && c->K > 0, "dg_tune: null state or no form");
    DG_TRY(hipSetDevice(c->device));
    return tune_impl(c, s);
}

int dg_launch_shape(dg_ctx* c, int* xcd_mapping, int* subcell_groups) {
    DG_REQUIRE(c != nullptr, "dg_launch_shape: null context");
    DG_LOCK(c);
    if (xcd_mapping) *xcd_mapping = (int)c->shape.xcd;
    if (subcell_groups) *subcell_groups = (int)c->shape.sub_groups;
    return 0;
}

int dg_is_finite(dg_ctx* c, const dg_state* s, int* finite) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && finite, "dg_is_finite: null argument");
    DG_TRY(hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
    size_t n = (size_t)3 * c->K * c->nn;
    hipLaunchKernelGGL(finite_kernel, dim3(1024), dim3(256), 0, c->stream, s->d, n, c->flag);
    DG_TRY(hipGetLastError());
    int bad = 0;
    DG_TRY(hipMemcpyAsync(&bad, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    *finite = bad ? 0 : 1;
    return 0;
}

int dg_clip_count(dg_ctx* c, uint64_t* clips) {
    DG_LOCK(c);
    DG_REQUIRE(c && clips, "dg_clip_count: null argument");
    unsigned long long v = 0;
    DG_TRY(hipMemcpyAsync(&v, c->clips, 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    *clips = v;
    return 0;
}

int dg_masks(dg_ctx* c, const dg_state* s, uint8_t* node_dry, uint8_t* elem_subcell) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && node_dry && elem_subcell, "dg_masks: null argument");
    uint8_t *dn = nullptr, *de = nullptr;
    DG_TRY(hipMalloc(&dn, (size_t)c->K * c->nn));
    DG_TRY(hipMalloc(&de, (size_t)c->K));
    dim3 grid((unsigned)((c->K + 255) / 256));
    int sub = (c->flags & DG_SUBCELLS) ? 1 : 0;
    switch (c->order) {
        case 1: hipLaunchKernelGGL(masks_kernel<1>, grid, dim3(256), 0, c->stream, s->d, c->K, c->h_dry, sub, dn, de); break;
        case 2: hipLaunchKernelGGL(masks_kernel<2>, grid, dim3(256), 0, c->stream, s->d, c->K, c->h_dry, sub, dn, de); break;
        default: hipLaunchKernelGGL(masks_kernel<3>, grid, dim3(256), 0, c->stream, s->d, c->K, c->h_dry, sub, dn, de); break;
    }
    DG_TRY(hipGetLastError());
    DG_TRY(hipMemcpyAsync(node_dry, dn, (size_t)c->K * c->nn, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipMemcpyAsync(elem_subcell, de, (size_t)c->K, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    DG_TRY(hipFree(dn));
    DG_TRY(hipFree(de));
    return 0;
}

int dg_masks_noalloc(dg_ctx* c, const dg_state* s, uint8_t* node_dry, uint8_t* elem_subcell) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && node_dry && elem_subcell, "dg_masks_noalloc: null argument");
    const size_t nodes = (size_t)c->K * c->nn, need = nodes + (size_t)c->K;
    if (need > SCRATCH_BYTES) {
        g_err = "dg_masks_noalloc: the masks of " + std::to_string(c->K) + " elements take " + std::to_string(need) +
                " bytes, more than the " + std::to_string(SCRATCH_BYTES) + "-byte static staging area";
        return 2;
    }
    DG_REQUIRE(c->scratch, "dg_masks_noalloc: the context has no staging area");
    // every byte of [0, need) is written by the kernel: nothing to clear
    uint8_t *dn = c->scratch, *de = c->scratch + nodes;
    dim3 grid((unsigned)((c->K + 255) / 256));
    int sub = (c->flags & DG_SUBCELLS) ? 1 : 0;
    switch (c->order) {
        case 1: hipLaunchKernelGGL(masks_kernel<1>, grid, dim3(256), 0, c->stream, s->d, c->K, c->h_dry, sub, dn, de); break;
        case 2: hipLaunchKernelGGL(masks_kernel<2>, grid, dim3(256), 0, c->stream, s->d, c->K, c->h_dry, sub, dn, de); break;
        default: hipLaunchKernelGGL(masks_kernel<3>, grid, dim3(256), 0, c->stream, s->d, c->K, c->h_dry, sub, dn, de); break;
    }
    DG_TRY(hipGetLastError());
    DG_TRY(hipMemcpyAsync(node_dry, dn, nodes, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipMemcpyAsync(elem_subcell, de, (size_t)c->K, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_timing_enable(dg_ctx* c, int on) {
    DG_LOCK(c);
    DG_REQUIRE(c, "dg_timing_enable: null context");
    c->timing = on != 0;
    c->stage_ms = c->dt_ms = 0.0;
    c->stage_n = c->dt_n = 0;
    return 0;
}

int dg_timing_read(dg_ctx* c, double* stage_ms, int64_t* stage_n, double* dt_ms, int64_t* dt_n) {
    DG_LOCK(c);
    DG_REQUIRE(c, "dg_timing_read: null context");
    if (stage_ms) *stage_ms = c->stage_ms;
    if (stage_n) *stage_n = c->stage_n;
    if (dt_ms) *dt_ms = c->dt_ms;
    if (dt_n) *dt_n = c->dt_n;
    return 0;
}

int dg_sync(dg_ctx* c) {
    DG_LOCK(c);
    DG_REQUIRE(c, "dg_sync: null context");
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_copy_bench(dg_ctx* c, size_t bytes, int reps, double* ms_per_copy) {
    DG_LOCK(c);
    DG_REQUIRE(c && ms_per_copy && reps > 0 && bytes >= 32, "dg_copy_bench: bad argument");
    DG_TRY(hipSetDevice(c->device));
    size_t n4 = bytes / 32;
    double4 *src = nullptr, *dst = nullptr;
    DG_TRY(hipMalloc(&src, n4 * 32));
    DG_TRY(hipMalloc(&dst, n4 * 32));
    DG_TRY(hipMemsetAsync(src, 0, n4 * 32, c->stream));
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, c->stream, src, dst, n4);
    DG_TRY(hipGetLastError());
    DG_TRY(hipEventRecord(c->ev[0], c->stream));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, c->stream, src, dst, n4);
    DG_TRY(hipGetLastError());
    DG_TRY(hipEventRecord(c->ev[1], c->stream));
    DG_TRY(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    DG_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    *ms_per_copy = ms / reps;
    DG_TRY(hipFree(src));
    DG_TRY(hipFree(dst));
    return 0;
}

int dg_state_zero(dg_ctx* c, dg_state* s) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && c->K > 0, "dg_state_zero: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    return zero_async(c, s->d, (size_t)3 * c->K * c->nn);
}

int dg_state_wrap(dg_ctx* c, double* device_ptr, dg_state** s) {
    DG_LOCK(c);
    DG_REQUIRE(c && device_ptr && s && c->K > 0, "dg_state_wrap: null argument or no form");
    dg_state* st = new dg_state;
    st->d = device_ptr;
    st->owned = false;
    *s = st;
    return 0;
}

int dg_set_stream(dg_ctx* c, void* stream) {
    DG_LOCK(c);
    DG_REQUIRE(c, "dg_set_stream: null context");
    DG_TRY(hipStreamSynchronize(c->stream));
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    return 0;
}

int dg_set_bed(dg_ctx* c, const double* bed, int on_device) {
    DG_LOCK(c);
    DG_REQUIRE(c && bed && c->K > 0, "dg_set_bed: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipMemcpyAsync(c->bed, bed, (size_t)c->K * c->nn * 8, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));  // (the kernels form the bed-slope factors from the bed itself)
    return 0;
}

int dg_set_manning(dg_ctx* c, const double* manning, int on_device) {
    DG_LOCK(c);
    DG_REQUIRE(c && manning && c->K > 0, "dg_set_manning: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipMemcpyAsync(c->manning, manning, (size_t)c->K * 8,
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_set_manning_nodal(dg_ctx* c, const double* manning, int on_device) {
    DG_LOCK(c);
    DG_REQUIRE(c && c->K > 0, "dg_set_manning_nodal: null context or no form");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    if (!manning) {
        if (c->manning_node) DG_TRY(hipFree(c->manning_node));
        c->manning_node = nullptr;
        return 0;
    }
    const size_t nt = (size_t)c->K * c->nn;
    if (!on_device)
        for (size_t i = 0; i < nt; ++i)
            DG_REQUIRE(std::isfinite(manning[i]) && manning[i] >= 0.0,
                       "dg_set_manning_nodal: a Manning coefficient is negative or not finite (friction.rs:520-526)");
    if (!c->manning_node) DG_TRY(hipMalloc(&c->manning_node, nt * sizeof(double)));
    DG_TRY(hipMemcpyAsync(c->manning_node, manning, nt * sizeof(double),
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_set_coriolis_beta(dg_ctx* c, double beta, double y0) {
    DG_LOCK(c);
    DG_REQUIRE(c && c->K > 0, "dg_set_coriolis_beta: null context or no form");
    DG_REQUIRE(std::isfinite(beta) && std::isfinite(y0), "dg_set_coriolis_beta: beta and y0 must be finite");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    c->beta = beta;
    c->beta_y0 = y0;
    if (beta == 0.0) {   // the f-plane (CoriolisSource2D::add_element takes f0 when beta is 0)
        if (c->coriolis_node) DG_TRY(hipFree(c->coriolis_node));
        c->coriolis_node = nullptr;
        return 0;
    }
    const int64_t nt = c->K * c->nn;
    if (!c->coriolis_node) DG_TRY(hipMalloc(&c->coriolis_node, (size_t)nt * sizeof(double)));
    dim3 grid((unsigned)((nt + 255) / 256));
    switch (c->order) {
        case 1: hipLaunchKernelGGL(coriolis_node_kernel<1>, grid, dim3(256), 0, c->stream, c->corners, c->K, c->f0, beta, y0, c->coriolis_node); break;
        case 2: hipLaunchKernelGGL(coriolis_node_kernel<2>, grid, dim3(256), 0, c->stream, c->corners, c->K, c->f0, beta, y0, c->coriolis_node); break;
        default: hipLaunchKernelGGL(coriolis_node_kernel<3>, grid, dim3(256), 0, c->stream, c->corners, c->K, c->f0, beta, y0, c->coriolis_node); break;
    }
    DG_TRY(hipGetLastError());
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_param_grads_zero(dg_ctx* c) {
    DG_LOCK(c);
    DG_REQUIRE(c && c->K > 0, "dg_param_grads_zero: no form");
    DG_TRY(hipSetDevice(c->device));
    if (int r = ensure_adjoint(c)) return r;
    if (int r = zero_async(c, c->gbed, (size_t)c->K * c->nn)) return r;
    return zero_async(c, c->gman, (size_t)c->K);
}

int dg_param_grads_read(dg_ctx* c, double* gbed, double* gman, int on_device) {
    DG_LOCK(c);
    DG_REQUIRE(c && c->K > 0, "dg_param_grads_read: no form");
    DG_TRY(hipSetDevice(c->device));
    if (int r = ensure_adjoint(c)) return r;
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (gbed) DG_TRY(hipMemcpyAsync(gbed, c->gbed, (size_t)c->K * c->nn * 8, kind, c->stream));
    if (gman) DG_TRY(hipMemcpyAsync(gman, c->gman, (size_t)c->K * 8, kind, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_rhs_vjp(dg_ctx* c, const dg_state* in, double t, const dg_state* cot_out, dg_state* cot_in) {
    DG_LOCK(c);
    DG_REQUIRE(c && in && cot_out && cot_in && c->K > 0, "dg_rhs_vjp: null argument or no form");
    DG_REQUIRE(cot_in != cot_out, "dg_rhs_vjp: cot_in must differ from cot_out");
    DG_TRY(hipSetDevice(c->device));
    if (int r = launch_stage_vjp(c, in->d, in->d, nullptr, cot_out->d, nullptr, cot_in->d, 0.0, 0.0, 0.0, t, 1)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_stage_vjp(dg_ctx* c, const dg_state* base, const dg_state* input, const dg_stage_desc* st,
                 const dg_state* cot_out, dg_state* cot_base, dg_state* cot_input) {
    DG_LOCK(c);
    DG_REQUIRE(c && base && input && st && cot_out && cot_input && c->K > 0, "dg_stage_vjp: null argument or no form");
    DG_REQUIRE(cot_input != cot_out && cot_base != cot_out, "dg_stage_vjp: the cotangent outputs must differ from cot_out");
    DG_TRY(hipSetDevice(c->device));
    if (int r = ensure_work(c)) return r;
    if (int r = ensure_adjoint(c)) return r;
    if (int r = launch_stage(c, base->d, input->d, c->work[0]->d, st->a, st->b, st->c, st->time, 0, nullptr,
                             c->rhsbuf[0]->d))
        return r;
    if (int r = launch_stage_vjp(c, base->d, input->d, c->rhsbuf[0]->d, cot_out->d, cot_base ? cot_base->d : nullptr,
                                 cot_input->d, st->a, st->b, st->c, st->time, 0))
        return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_step_vjp(dg_ctx* c, const dg_state* s, double t, double dt, int scheme, const dg_state* cot_next,
                dg_state* cot_state) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && cot_next && cot_state && c->K > 0, "dg_step_vjp: null argument or no form");
    DG_REQUIRE(scheme == DG_SSP_RK3 || scheme == DG_SSP_RK43, "dg_step_vjp: unknown scheme");
    DG_REQUIRE(cot_next != cot_state && cot_next->d != cot_state->d, "dg_step_vjp: cot_next must differ from cot_state");
    DG_TRY(hipSetDevice(c->device));
    if (int r = step_vjp_impl(c, s->d, t, dt, scheme, cot_next->d, cot_state->d)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_set_stations(dg_ctx* c, int n, const int64_t* element, const double* weights) {
    DG_LOCK(c);
    DG_REQUIRE(c && c->K > 0 && n >= 0 && (n == 0 || (element && weights)), "dg_set_stations: bad argument or no form");
    for (int s = 0; s < n; ++s)
        DG_REQUIRE(element[s] >= 0 && element[s] < c->K, "dg_set_stations: station element out of range");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    if (c->st_elem) DG_TRY(hipFree(c->st_elem));
    if (c->st_w) DG_TRY(hipFree(c->st_w));
    if (c->st_eta) DG_TRY(hipFree(c->st_eta));
    c->st_elem = nullptr;
    c->st_w = c->st_eta = nullptr;
    c->n_st = n;
    if (n == 0) return 0;
    DG_TRY(hipMalloc(&c->st_elem, (size_t)n * sizeof(int64_t)));
    DG_TRY(hipMalloc(&c->st_w, (size_t)n * c->nn * sizeof(double)));
    DG_TRY(hipMalloc(&c->st_eta, (size_t)n * sizeof(double)));
    DG_TRY(hipMemcpy(c->st_elem, element, (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->st_w, weights, (size_t)n * c->nn * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int dg_sample_stations(dg_ctx* c, const dg_state* s, double* eta) {
    DG_LOCK(c);
    DG_REQUIRE(c && s && eta && c->n_st > 0, "dg_sample_stations: null argument or no stations");
    DG_TRY(hipSetDevice(c->device));
    hipLaunchKernelGGL(station_sample_kernel, dim3((c->n_st + 63) / 64), dim3(64), 0, c->stream, s->d, c->bed,
                       c->st_elem, c->st_w, c->n_st, c->nn, c->st_eta);
    DG_TRY(hipGetLastError());
    DG_TRY(hipMemcpyAsync(eta, c->st_eta, (size_t)c->n_st * 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_tape_create(dg_ctx* c, int snapshots, dg_tape** out) {
    DG_LOCK(c);
    DG_REQUIRE(c && out && c->K > 0, "dg_tape_create: null argument or no form");
    DG_REQUIRE(snapshots >= 1, "dg_tape_create: at least one snapshot (the initial state)");
    DG_TRY(hipSetDevice(c->device));
    dg_tape* tp = new dg_tape;
    tp->snapshots = snapshots;
    tp->K = c->K;
    tp->nn = c->nn;
    const size_t bytes = (size_t)3 * c->K * c->nn * sizeof(double);
    auto alloc = [&](dg_state*& st) -> int {
        st = new dg_state;
        DG_TRY(hipMalloc(&st->d, bytes));
        return 0;
    };
    tp->slot.resize(snapshots, nullptr);
    tp->used.assign(snapshots, 0);
    for (auto& st : tp->slot)
        if (int r = alloc(st)) return r;
    if (int r = alloc(tp->cur)) return r;
    if (int r = alloc(tp->lam[0])) return r;
    if (int r = alloc(tp->lam[1])) return r;
    *out = tp;
    return 0;
}

int dg_tape_destroy(dg_ctx* c, dg_tape* tp) {
    if (!tp) return 0;
    DG_LOCK(c);
    DG_REQUIRE(c, "dg_tape_destroy: null context");
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    auto fr = [&](dg_state* st) -> int {
        if (st) {
            DG_TRY(hipFree(st->d));
            delete st;
        }
        return 0;
    };
    for (auto* st : tp->slot)
        if (int r = fr(st)) return r;
    if (int r = fr(tp->cur)) return r;
    if (int r = fr(tp->lam[0])) return r;
    if (int r = fr(tp->lam[1])) return r;
    if (tp->eta) DG_TRY(hipFree(tp->eta));
    if (tp->obs) DG_TRY(hipFree(tp->obs));
    if (tp->rec) DG_TRY(hipFree(tp->rec));
    delete tp;
    return 0;
}

int dg_tape_keep_stages(dg_ctx* c, dg_tape* tp, int on) {
    DG_LOCK(c);
    DG_REQUIRE(c && tp, "dg_tape_keep_stages: null argument");
    DG_TRY(hipSetDevice(c->device));
    tp->keep_stages = on != 0;
    tp->ready = false;
    if (!tp->keep_stages && tp->rec) {
        DG_TRY(hipStreamSynchronize(c->stream));
        DG_TRY(hipFree(tp->rec));
        tp->rec = nullptr;
        tp->rec_bytes = 0;
    }
    return 0;
}

int dg_misfit_forward(dg_ctx* c, dg_tape* tp, const dg_state* x0, double t0, int n_steps, const double* dts, int scheme,
                      int every, const double* obs, double* misfit) {
    DG_LOCK(c);
    DG_REQUIRE(c && tp && x0 && dts && obs && misfit && c->K > 0, "dg_misfit_forward: null argument or no form");
    DG_REQUIRE(tp->K == c->K && tp->nn == c->nn, "dg_misfit_forward: the tape belongs to another form");
    DG_REQUIRE(n_steps >= 1 && every >= 1 && every <= n_steps, "dg_misfit_forward: need 1 <= every <= n_steps");
    DG_REQUIRE(scheme == DG_SSP_RK3 || scheme == DG_SSP_RK43, "dg_misfit_forward: unknown scheme");
    DG_REQUIRE(c->n_st > 0, "dg_misfit_forward: no stations (dg_set_stations)");
    DG_TRY(hipSetDevice(c->device));
    tp->ready = false;
    tp->n = n_steps;
    tp->scheme = scheme;
    tp->every = every;
    tp->n_st = c->n_st;
    tp->n_samples = n_steps / every;
    tp->dts.assign(dts, dts + n_steps);
    tp->t_at.resize(n_steps);
    double t = t0;
    for (int i = 0; i < n_steps; ++i) {
        tp->t_at[i] = t;
        t += dts[i];
    }
    if (tp->eta) DG_TRY(hipFree(tp->eta));
    if (tp->obs) DG_TRY(hipFree(tp->obs));
    tp->eta = tp->obs = nullptr;
    tp->records = false;
    tp->rec_step.clear();
    if (tp->keep_stages) {
        const int per = scheme == DG_SSP_RK3 ? 5 : 7;
        const int n_rec = std::min(tp->snapshots + 1, n_steps);
        const size_t need = (size_t)n_rec * per * 3 * c->K * c->nn * sizeof(double);
        if (tp->rec_bytes != need) {
            if (tp->rec) {
                DG_TRY(hipStreamSynchronize(c->stream));
                DG_TRY(hipFree(tp->rec));
                tp->rec = nullptr;
                tp->rec_bytes = 0;
            }
            DG_TRY(hipMalloc(&tp->rec, need));
            tp->rec_bytes = need;
        }
        tp->rec_per_step = per;
        tp->rec_step.assign(n_rec, -1);
        tp->records = true;
    }
    const size_t ns = (size_t)tp->n_samples * tp->n_st;
    DG_TRY(hipMalloc(&tp->eta, ns * sizeof(double)));
    DG_TRY(hipMalloc(&tp->obs, ns * sizeof(double)));
    DG_TRY(hipMemcpy(tp->obs, obs, ns * sizeof(double), hipMemcpyHostToDevice));
    // the schedule: the initial state in slot 0, the right-most descent's checkpoints
    // taken by this sweep
    tp->rv.init(n_steps, tp->snapshots - 1);
    std::fill(tp->used.begin(), tp->used.end(), 0);
    tp->pre.clear();
    {
        int a = 0, free = tp->snapshots - 1, sl = 1;
        while (n_steps - a >= 2 && free > 0) {
            a += tp->rv.split(n_steps - a, free, true);
            tp->pre.push_back({a, sl});
            tp->used[sl] = 1;
            ++sl;
            --free;
        }
    }
    tp->used[0] = 1;
    if (int r = tape_copy(c, tp->slot[0], x0)) return r;
    if (int r = tape_copy(c, tp->cur, x0)) return r;
    tp->fwd = 0;
    tp->vjps = 0;
    size_t pi = 0;
    const double* x = tp->cur->d;  // the current state: the tape's own buffer or the last checkpoint's slot
    for (int i = 0; i < n_steps; ++i) {
        const int after = i + 1;
        double* dst = nullptr;
        if (pi < tp->pre.size() && tp->pre[pi].first == after) dst = tp->slot[tp->pre[pi++].second]->d;
        double* rec = dst || after == n_steps ? rec_take(c, tp, i) : nullptr;
        if (int r = sweep_step(c, tp, i, x, dst, rec)) return r;
        x = dst ? dst : tp->cur->d;
        tp->fwd += 1;
        if (after % every == 0 && after / every <= tp->n_samples) {
            hipLaunchKernelGGL(station_sample_kernel, dim3((tp->n_st + 63) / 64), dim3(64), 0, c->stream, x, c->bed,
                               c->st_elem, c->st_w, tp->n_st, c->nn, tp->eta + (size_t)(after / every - 1) * tp->n_st);
            DG_TRY(hipGetLastError());
        }
    }
    std::vector<double> eta(ns);
    DG_TRY(hipMemcpyAsync(eta.data(), tp->eta, ns * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    double J = 0.0;
    for (size_t q = 0; q < ns; ++q) {
        const double d = eta[q] - obs[q];
        J += d * d;
    }
    *misfit = J;
    tp->ready = true;
    return 0;
}

int dg_misfit_backward(dg_ctx* c, dg_tape* tp, double scale, dg_state* cot_x0) {
    DG_LOCK(c);
    DG_REQUIRE(c && tp && c->K > 0, "dg_misfit_backward: null argument or no form");
    DG_REQUIRE(tp->ready, "dg_misfit_backward: run dg_misfit_forward first (once per backward)");
    DG_REQUIRE(c->n_st == tp->n_st, "dg_misfit_backward: the stations changed since the forward");
    DG_TRY(hipSetDevice(c->device));
    tp->ready = false;
    if (int r = ensure_adjoint(c)) return r;
    if (int r = zero_async(c, c->gbed, (size_t)c->K * c->nn)) return r;
    if (int r = zero_async(c, c->gman, (size_t)c->K)) return r;
    tp->lam_cur = 0;
    tp->rec_used = 0;
    if (int r = zero_async(c, tp->lam[0]->d, (size_t)3 * c->K * c->nn)) return r;
    tp->scale = scale;
    if (int r = tape_reverse(c, tp, 0, tp->n, 0, tp->snapshots - 1, true)) return r;
    if (cot_x0)
        DG_TRY(hipMemcpyAsync(cot_x0->d, tp->lam[tp->lam_cur]->d, (size_t)3 * c->K * c->nn * 8,
                              hipMemcpyDeviceToDevice, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_tape_records_used(const dg_tape* tp, int64_t* used) {
    DG_REQUIRE(tp && used, "dg_tape_records_used: null argument");
    *used = tp->rec_used;
    return 0;
}

int dg_tape_stats(const dg_tape* tp, int64_t* forward_steps, int64_t* step_vjps) {
    DG_REQUIRE(tp, "dg_tape_stats: null tape");
    if (forward_steps) *forward_steps = tp->fwd;
    if (step_vjps) *step_vjps = tp->vjps;
    return 0;
}

int dg_revolve_count(int n_steps, int snapshots, int64_t* forward_steps) {
    DG_REQUIRE(n_steps >= 1 && snapshots >= 1 && forward_steps, "dg_revolve_count: bad argument");
    Revolve rv;
    rv.init(n_steps, snapshots - 1);
    *forward_steps = n_steps + revolve_count_rec(rv, n_steps, snapshots - 1, true);
    return 0;
}

int dg_batch_create(int n, dg_ctx* const* members, int scheme, dg_batch** out) {
    DG_REQUIRE(n >= 1 && members && out, "dg_batch_create: need at least one member");
    DG_REQUIRE(scheme == DG_SSP_RK3 || scheme == DG_SSP_RK43, "dg_batch_create: unknown scheme");
    for (int m = 0; m < n; ++m) {
        DG_REQUIRE(members[m] && members[m]->K > 0, "dg_batch_create: every member needs an uploaded form");
        DG_REQUIRE(members[m]->device == members[0]->device, "dg_batch_create: members on different devices");
        DG_REQUIRE(!members[m]->tide_slot, "dg_batch_create: a member with boundary tides per node (not batched)");
        DG_REQUIRE(!members[m]->nest_slot, "dg_batch_create: a member with a nesting band (not batched)");
        DG_REQUIRE(!members[m]->wx_series, "dg_batch_create: a member with gridded weather (not batched)");
        DG_REQUIRE(!members[m]->manning_node && !members[m]->coriolis_node,
                   "dg_batch_create: a member with Manning n per node or a beta-plane (not batched)");
    }
    DG_TRY(hipSetDevice(members[0]->device));
    dg_batch* b = new dg_batch;
    b->owner = members[0];
    b->members.assign(members, members + n);
    b->scheme = scheme;
    b->ns = scheme == DG_SSP_RK3 ? 3 : 4;
    b->h_mem.resize(n);
    std::vector<uint32_t> bm[4];
    for (int m = 0; m < n; ++m) {
        dg_ctx* c = members[m];
        BatchMember& M = b->h_mem[m];
        std::memset(&M, 0, sizeof(M));
        const size_t bytes = (size_t)3 * c->K * c->nn * sizeof(double);
        for (int q = 0; q < 2; ++q) DG_TRY(hipMalloc(&M.x[q], bytes));
        for (int q = 0; q < b->ns; ++q) DG_TRY(hipMalloc(&M.y[q], bytes));
        StageParams& p = M.sp;
        p.bed = c->bed; p.corners = c->corners; p.manning = c->manning;
        p.links = c->links; p.tags = c->tags; p.affine = c->affine;
        p.n_elements = c->K;
        p.g = c->g; p.h_min = c->h_min; p.h_dry = c->h_dry; p.u_max = c->u_max; p.tau = c->tau; p.f0 = c->f0;
        p.sqrt_g = std::sqrt(c->g);
        p.geo = c->geo;
        p.fgeo = c->fgeo;
        p.flags = c->flags;
        p.clips = c->clips;
        DtParams& d = M.dp;
        d.corners = c->corners; d.emass = c->emass; d.links = c->links; d.affine = c->affine; d.n_elements = c->K;
        d.g = c->g; d.h_min = c->h_min; d.h_dry = c->h_dry;
        d.positivity = (c->flags & DG_WET_DRY) ? 1 : 0;
        const int o = c->order;
        const int epb = 64 / c->nn;
        M.n_blocks = (uint32_t)((c->K + epb - 1) / epb);
        M.block0 = (uint32_t)bm[o].size();
        bm[o].insert(bm[o].end(), M.n_blocks, (uint32_t)m);
    }
    for (int o = 1; o <= 3; ++o) {
        BatchOrder& B = b->ord[o];
        B.n_blocks = (unsigned int)bm[o].size();
        if (B.n_blocks == 0) continue;
        const int epb = 64 / ((o + 1) * (o + 1));
        const size_t nchunk = (B.n_blocks + 1023) / 1024;
        DG_TRY(hipMalloc(&B.block_member, B.n_blocks * sizeof(uint32_t)));
        DG_TRY(hipMalloc(&B.queue, (size_t)B.n_blocks * epb * sizeof(uint32_t)));
        DG_TRY(hipMalloc(&B.queue_count, B.n_blocks));
        DG_TRY(hipMalloc(&B.list, (size_t)B.n_blocks * epb * sizeof(uint32_t)));
        DG_TRY(hipMalloc(&B.list_member, (size_t)B.n_blocks * epb * sizeof(uint32_t)));
        DG_TRY(hipMalloc(&B.chunk_total, nchunk * sizeof(unsigned int)));
        DG_TRY(hipMalloc(&B.list_count, sizeof(unsigned int)));
        DG_TRY(hipMalloc(&B.block_min, B.n_blocks * sizeof(double)));
        DG_TRY(hipMemcpy(B.block_member, bm[o].data(), B.n_blocks * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    for (int m = 0; m < n; ++m) {
        b->h_mem[m].block_min = b->ord[members[m]->order].block_min;
        b->h_mem[m].dp.slots = nullptr;
    }
    DG_TRY(hipMalloc(&b->d_mem, n * sizeof(BatchMember)));
    DG_TRY(hipMalloc(&b->d_coef, (size_t)n * 4 * sizeof(BatchCoef)));
    DG_TRY(hipMalloc(&b->d_member_dt, n * sizeof(double)));
    DG_TRY(hipHostMalloc((void**)&b->h_coef, (size_t)n * 4 * sizeof(BatchCoef)));
    DG_TRY(hipHostMalloc((void**)&b->h_member_dt, n * sizeof(double)));
    DG_TRY(hipMemcpy(b->d_mem, b->h_mem.data(), n * sizeof(BatchMember), hipMemcpyHostToDevice));
    *out = b;
    return 0;
}

int dg_batch_destroy(dg_batch* b) {
    if (!b) return 0;
    DG_LOCK(b->owner);
    DG_TRY(hipSetDevice(b->owner->device));
    DG_TRY(hipStreamSynchronize(b->owner->stream));
    if (int r = batch_free(b)) return r;
    delete b;
    return 0;
}

int dg_batch_upload(dg_batch* b, int m, const double* h, const double* hu, const double* hv) {
    DG_REQUIRE(b != nullptr, "dg_batch_upload: null ensemble");
    DG_LOCK(b->owner);
    DG_REQUIRE(b && m >= 0 && m < (int)b->members.size() && h && hu && hv, "dg_batch_upload: bad argument");
    dg_ctx* c = b->members[m];
    const size_t nt = (size_t)c->K * c->nn;
    double* x = b->h_mem[m].x[b->parity];
    hipStream_t st = b->owner->stream;
    DG_TRY(hipMemcpyAsync(x, h, nt * 8, hipMemcpyHostToDevice, st));
    DG_TRY(hipMemcpyAsync(x + nt, hu, nt * 8, hipMemcpyHostToDevice, st));
    DG_TRY(hipMemcpyAsync(x + 2 * nt, hv, nt * 8, hipMemcpyHostToDevice, st));
    DG_TRY(hipStreamSynchronize(st));
    return 0;
}

int dg_batch_download(dg_batch* b, int m, double* h, double* hu, double* hv) {
    DG_REQUIRE(b != nullptr, "dg_batch_download: null ensemble");
    DG_LOCK(b->owner);
    DG_REQUIRE(b && m >= 0 && m < (int)b->members.size() && h && hu && hv, "dg_batch_download: bad argument");
    dg_ctx* c = b->members[m];
    const size_t nt = (size_t)c->K * c->nn;
    const double* x = b->h_mem[m].x[b->parity];
    hipStream_t st = b->owner->stream;
    DG_TRY(hipMemcpyAsync(h, x, nt * 8, hipMemcpyDeviceToHost, st));
    DG_TRY(hipMemcpyAsync(hu, x + nt, nt * 8, hipMemcpyDeviceToHost, st));
    DG_TRY(hipMemcpyAsync(hv, x + 2 * nt, nt * 8, hipMemcpyDeviceToHost, st));
    DG_TRY(hipStreamSynchronize(st));
    return 0;
}

int dg_batch_run(dg_batch* b, int n_steps, double cfl, int options, double* t) {
    DG_REQUIRE(b != nullptr, "dg_batch_run: null ensemble");
    DG_LOCK(b->owner);
    DG_REQUIRE(b && t && n_steps >= 0, "dg_batch_run: bad argument");
    DG_TRY(hipSetDevice(b->owner->device));
    const int M = (int)b->members.size();
    hipStream_t st = b->owner->stream;
    if (cfl != b->cfl) {
        // PositivityBound::new per member (swe_2d.rs:2188-2203), as launch_dt
        static const double lin3[4] = {1.6, 1.04, 0.73, 0.56}, lin43[4] = {2.2, 1.51, 1.15, 0.93};
        static const double sub3[4] = {0.94, 0.65, 0.47, 0.36}, sub43[4] = {1.5, 1.1, 0.97, 0.75};
        for (int m = 0; m < M; ++m) {
            DtParams& d = b->h_mem[m].dp;
            const int order = b->members[m]->order;
            const double nn = order;
            const double ssp = b->scheme == DG_SSP_RK3 ? 1.0 : 2.0;
            d.cfl = cfl;
            d.pcfl = ssp * ((2.0 * nn + 1.0) / (2.0 * nn * (nn + 1.0)));
            d.max_cfl = std::fmax(d.pcfl, 0.9 * (b->scheme == DG_SSP_RK3 ? lin3 : lin43)[order - 1]);
            d.max_cfl_sub = std::fmax(d.pcfl, 0.9 * (b->scheme == DG_SSP_RK3 ? sub3 : sub43)[order - 1]);
            d.xcd_stride = 0;
        }
        DG_TRY(hipMemcpyAsync(b->d_mem, b->h_mem.data(), M * sizeof(BatchMember), hipMemcpyHostToDevice, st));
        b->cfl = cfl;
    }
    const double(*tab)[4] = b->scheme == DG_SSP_RK3 ? RK3 : RK43;
    const bool use_graph = (options & DG_BATCH_GRAPH) != 0;
    for (int s = 0; s < n_steps; ++s) {
        if (int r = batch_dt(b)) return r;
        for (int m = 0; m < M; ++m) {
            const double dt = b->h_member_dt[m];
            if (!(dt > 0.0) || !std::isfinite(dt)) {
                g_err = "dg_batch_run: member " + std::to_string(m) + " has no finite positive time step at step " +
                        std::to_string(s);
                return 3;
            }
            for (int q = 0; q < b->ns; ++q) {
                BatchCoef& cf = b->h_coef[(size_t)m * 4 + q];
                cf.a = tab[q][0];
                cf.b = tab[q][1];
                cf.c = tab[q][2] * dt;
                cf.eta = tide_elevation(b->members[m]->tide, b->members[m]->tide_nodal, t[m] + tab[q][3] * dt);
            }
        }
        DG_TRY(hipMemcpyAsync(b->d_coef, b->h_coef, (size_t)M * 4 * sizeof(BatchCoef), hipMemcpyHostToDevice, st));
        if ((options & DG_BATCH_TUNE) && !b->tuned)
            if (int r = batch_tune(b)) return r;
        const int par = b->parity;
        if (use_graph) {
            if (!b->exec[par]) {
                DG_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                int r = batch_stages(b, par);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(st, &g);
                if (r) return r;
                DG_TRY(e);
                b->graph[par] = g;
                DG_TRY(hipGraphInstantiate(&b->exec[par], g, nullptr, nullptr, 0));
            }
            DG_TRY(hipGraphLaunch(b->exec[par], st));
        } else if (int r = batch_stages(b, par)) return r;
        b->parity ^= 1;
        for (int m = 0; m < M; ++m) t[m] += b->h_member_dt[m];
    }
    DG_TRY(hipStreamSynchronize(st));
    return 0;
}

int dg_batch_tuning(const dg_batch* b, int order, int* xcd_mapping, int* subcell_groups) {
    DG_REQUIRE(b != nullptr, "dg_batch_tuning: null ensemble");
    DG_LOCK(b->owner);
    DG_REQUIRE(b && order >= 1 && order <= 3, "dg_batch_tuning: bad argument");
    if (xcd_mapping) *xcd_mapping = (int)b->ord[order].xcd;
    if (subcell_groups) *subcell_groups = (int)b->ord[order].sub_groups;
    return 0;
}

}  // extern "C"

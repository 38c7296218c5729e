// ORACLE — test infrastructure only (see swe2d_oracle.hpp): a C ABI over the
// CPU restatement for the Python tests (ctypes), the smoke check and the
// benchmark's CPU-baseline leg.
#include <chrono>
#include <cstdio>
#include <string>

#include "swe2d_solver.hpp"

using oracle::Dual;
using oracle::Solver;

namespace {
thread_local std::string g_err;
struct Handle {
    Solver<double>* s = nullptr;
    std::vector<double> work;
};
}  // namespace

extern "C" {

const char* oc_last_error(void) { return g_err.c_str(); }

void oc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// flags: bit0 subcells, bit1 friction, bit2 wet_dry, bit3 manufactured source, bit4 incoming-wave tide
// surface: 0 hydrostatic HLL (WetDry), 1 entropy stable, 2 entropy conservative
void* oc_create(int order, int64_t K, const double* corners, const uint32_t* links, const uint8_t* tags,
                const double* bed, const double* manning, double g, double h_min, double h_dry, double f0,
                int surface, int flags) {
    if (order < 1 || order > 3) { g_err = "order must be 1..3"; return nullptr; }
    Handle* h = new Handle;
    h->s = new Solver<double>(order, K, corners, links, tags);
    std::string e = h->s->init_geometry();
    if (!e.empty()) { g_err = e; delete h->s; delete h; return nullptr; }
    auto& c = h->s->cfg;
    c.g = g; c.h_min = h_min; c.h_dry = h_dry; c.subcell_depth = h_dry; c.f0 = f0; c.surface = surface;
    c.subcells = flags & 1; c.friction = flags & 2; c.wet_dry = flags & 4; c.manufactured = flags & 8;
    c.incoming_wave = flags & 16;
    size_t n = (size_t)K * h->s->ops.nn;
    std::copy(bed, bed + n, h->s->bed.begin());
    if (manning) std::copy(manning, manning + K, h->s->manning.begin());
    return h;
}

void oc_destroy(void* p) {
    Handle* h = (Handle*)p;
    if (!h) return;
    delete h->s;
    delete h;
}

void oc_set_tide(void* p, double mean, double ramp, int n, const double* amp, const double* omega, const double* phase) {
    auto& t = ((Handle*)p)->s->tide;
    t.mean = mean; t.ramp = ramp;
    t.amplitude.assign(amp, amp + n); t.omega.assign(omega, omega + n); t.phase.assign(phase, phase + n);
    t.f0.clear(); t.f_rate.clear(); t.v0.clear(); t.u0.clear(); t.u_rate.clear();
}

void oc_set_tide_nodal(void* p, int n, const double* f0, const double* f_rate, const double* v0, const double* u0,
                       const double* u_rate) {
    auto& t = ((Handle*)p)->s->tide;
    t.f0.assign(f0, f0 + n); t.f_rate.assign(f_rate, f_rate + n); t.v0.assign(v0, v0 + n);
    t.u0.assign(u0, u0 + n); t.u_rate.assign(u_rate, u_rate + n);
}

double oc_tide_elevation(void* p, double t) { return ((Handle*)p)->s->tide.elevation(t); }

void oc_rhs(void* p, const double* h, const double* hu, const double* hv, double t, double* oh, double* ohu, double* ohv) {
    ((Handle*)p)->s->rhs(h, hu, hv, t, oh, ohu, ohv);
}

// F* mass component at every face node from each element's side ([K][4][n1]); boundary faces included
void oc_face_mass(void* p, const double* h, const double* hu, const double* hv, double t, double* out) {
    Solver<double>& s = *((Handle*)p)->s;
    s.face_pass(h, hu, hv);
    int n1 = s.ops.n1;
    double eta = s.tide.elevation(t);
    for (int64_t k = 0; k < s.K; ++k)
        for (int f = 0; f < 4; ++f)
            for (int fi = 0; fi < n1; ++fi) {
                size_t idx = ((size_t)k * 4 + f) * n1 + fi;
                uint32_t l = s.links[4 * k + f];
                if (oracle::link_is_boundary(l)) {
                    double nx, ny, sj;
                    s.normal(k, f, fi, nx, ny, sj);
                    out[idx] = s.boundary_flux(k, f, s.node(h, hu, hv, k, s.ops.face_nodes[f][fi]), nx, ny, eta, fi, t).h;
                } else out[idx] = s.faces[idx].h;
            }
}

void oc_stage(void* p, const double* bh, const double* bhu, const double* bhv, const double* ih, const double* ihu,
              const double* ihv, double* oh, double* ohu, double* ohv, double a, double b, double c, double time, int last) {
    ((Handle*)p)->s->stage(bh, bhu, bhv, ih, ihu, ihv, oh, ohu, ohv, a, b, c, time, last != 0);
}

// LocalTimeStepping::stage_where / finish_where over listed elements (builder.rs:652-782)
void oc_stage_where(void* p, const double* bh, const double* bhu, const double* bhv, const double* ih,
                    const double* ihu, const double* ihv, const uint32_t* elements, int64_t n_list, double t,
                    double dt, int finest, int substep, int depth, double start, double input, double beta, double c,
                    int last, const uint8_t* level, const uint8_t* rate, double* oh, double* ohu, double* ohv,
                    double* ah, double* ahu, double* ahv) {
    oracle::ShuOsher st{start, input, beta, c};
    ((Handle*)p)->s->stage_where(bh, bhu, bhv, ih, ihu, ihv, elements, n_list, t, dt, finest, substep, depth, st,
                                 last != 0, level, rate, oh, ohu, ohv, ah, ahu, ahv);
}

void oc_finish_where(void* p, double* h, double* hu, double* hv, const double* ah, const double* ahu,
                     const double* ahv, const uint32_t* elements, int64_t n_list) {
    ((Handle*)p)->s->finish_where(h, hu, hv, ah, ahu, ahv, elements, n_list);
}

// BoundaryTides per boundary face node (tidal_atlas.rs:858-1140)
void oc_set_boundary_tides(void* p, int n_c, const double* omega, const double* f_rate, double ramp, int velocity,
                           double max_ratio, int64_t n_slots, const uint32_t* slot, const double* coef,
                           const double* mean, const double* source_depth) {
    Solver<double>& s = *((Handle*)p)->s;
    auto& a = s.atlas;
    a.n_c = n_c;
    a.omega.assign(omega, omega + n_c);
    a.f_rate.assign(f_rate, f_rate + n_c);
    a.ramp = ramp;
    a.velocity = velocity != 0;
    a.max_ratio = max_ratio;
    a.slot.assign(slot, slot + (size_t)s.K * 4 * s.ops.n1);
    a.coef.assign(coef, coef + (size_t)n_slots * n_c * 6);
    a.mean.assign(mean, mean + n_slots);
    a.source_depth.assign(source_depth, source_depth + n_slots);
}

void oc_set_nesting(void* p, int64_t n_slots, const uint32_t* slot, int n_times, const double* times, double epoch,
                    int cubic, const float* series, const double* weight, const double* depth, double timescale,
                    double wet_depth, int momentum_only, int velocity, double max_ratio, double reference_level,
                    double ramp) {
    Solver<double>& s = *((Handle*)p)->s;
    auto& ns = s.nesting;
    ns.on = n_slots > 0;
    if (!ns.on) return;
    ns.slot.assign(slot, slot + (size_t)s.K * s.ops.nn);
    ns.times.assign(times, times + n_times);
    ns.series.assign(series, series + (size_t)n_slots * n_times * 3);
    ns.weight.assign(weight, weight + n_slots);
    ns.depth.assign(depth, depth + n_slots);
    ns.epoch = epoch;
    ns.cubic = cubic != 0;
    ns.rate = 1.0 / timescale;
    ns.wet_depth = wet_depth;
    ns.momentum_only = momentum_only != 0;
    ns.velocity = velocity != 0;
    ns.max_ratio = max_ratio;
    ns.reference_level = reference_level;
    ns.ramp = ramp;
}

void oc_set_weather(void* p, int n_times, const double* times, double epoch, int cubic, const float* series, int wind,
                    int pressure, int drag, double cd, double rho_air, double rho_water, double h_min, double ramp) {
    Solver<double>& s = *((Handle*)p)->s;
    auto& w = s.weather;
    w.on = n_times > 0;
    if (!w.on) return;
    w.times.assign(times, times + n_times);
    w.series.assign(series, series + (size_t)n_times * 5 * (size_t)s.K * s.ops.nn);
    w.epoch = epoch;
    w.cubic = cubic != 0;
    w.wind = wind != 0;
    w.pressure = pressure != 0;
    w.drag = drag;
    w.cd = cd;
    w.rho_air = rho_air;
    w.rho_water = rho_water;
    w.h_min = h_min;
    w.ramp = ramp;
}

void oc_set_inverse_barometer(void* p, double factor) { ((Handle*)p)->s->weather.raise = factor; }

void oc_set_nesting_boundary(void* p, int on) { ((Handle*)p)->s->nesting.boundary = on != 0; }

double oc_hypot(double x, double y) { return Solver<double>::ref_hypot(x, y); }

// The thin-layer relaxation time tau (WetDryConfig, wetting_drying.rs:82-89; 1 s by default)
void oc_set_relaxation_time(void* p, double tau) { ((Handle*)p)->s->cfg.relaxation_time = tau; }
// A beta-plane f = f0 + beta (y - y0) (CoriolisSource2D::beta_plane, coriolis.rs:101-121)
void oc_set_beta(void* p, double beta, double y0) {
    auto& c = ((Handle*)p)->s->cfg;
    c.beta = beta; c.beta_y0 = y0;
}
// Manning n per node [K*nn] (SpatiallyVaryingManning2D::from_nodal, friction.rs:512-537); NULL: per element
void oc_set_manning_nodal(void* p, const double* n) {
    Solver<double>& s = *((Handle*)p)->s;
    if (!n) { s.manning_node.clear(); return; }
    s.manning_node.assign(n, n + (size_t)s.K * s.ops.nn);
}

void oc_step(void* p, double* h, double* hu, double* hv, double t, double dt, int scheme) {
    Handle* hd = (Handle*)p;
    hd->s->step(h, hu, hv, t, dt, scheme, hd->work);
}

double oc_dt(void* p, const double* h, const double* hu, const double* hv, double cfl, int scheme) {
    return ((Handle*)p)->s->dt_ssp(h, hu, hv, cfl, scheme);
}

void oc_element_dt(void* p, const double* h, const double* hu, const double* hv, double cfl, int scheme, double* out) {
    Solver<double>& s = *((Handle*)p)->s;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < s.K; ++k) out[k] = s.element_dt(h, hu, hv, k, cfl, scheme, s.cfg.wet_dry);
}

// n_steps SSP steps from t at the adaptive dt (no landing); returns the end time.
// dts (optional, n_steps) receives each step's dt.
double oc_run(void* p, double* h, double* hu, double* hv, double t, int n_steps, double cfl, int scheme, double* dts) {
    Handle* hd = (Handle*)p;
    for (int i = 0; i < n_steps; ++i) {
        double dt = hd->s->dt_ssp(h, hu, hv, cfl, scheme);
        if (dts) dts[i] = dt;
        hd->s->step(h, hu, hv, t, dt, scheme, hd->work);
        t += dt;
    }
    return t;
}

// The same, timed: wall seconds of the step loop alone
double oc_run_timed(void* p, double* h, double* hu, double* hv, double t, int n_steps, double cfl, int scheme, double* seconds) {
    auto t0 = std::chrono::steady_clock::now();
    double te = oc_run(p, h, hu, hv, t, n_steps, cfl, scheme, nullptr);
    auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    return te;
}

uint64_t oc_clips(void* p) { return ((Handle*)p)->s->clips; }

// ---- Tangent-linear model: the same restatement on dual numbers ----
struct TlmHandle {
    Solver<Dual>* s = nullptr;
    std::vector<Dual> work;
};

void* oc_tlm_create(int order, int64_t K, const double* corners, const uint32_t* links, const uint8_t* tags,
                    const double* bed, const double* manning, double g, double h_min, double h_dry, double f0,
                    int surface, int flags) {
    if (order < 1 || order > 3) { g_err = "order must be 1..3"; return nullptr; }
    TlmHandle* h = new TlmHandle;
    h->s = new Solver<Dual>(order, K, corners, links, tags);
    std::string e = h->s->init_geometry();
    if (!e.empty()) { g_err = e; delete h->s; delete h; return nullptr; }
    auto& c = h->s->cfg;
    c.g = g; c.h_min = h_min; c.h_dry = h_dry; c.subcell_depth = h_dry; c.f0 = f0; c.surface = surface;
    c.subcells = flags & 1; c.friction = flags & 2; c.wet_dry = flags & 4; c.manufactured = flags & 8;
    c.incoming_wave = flags & 16;
    size_t n = (size_t)K * h->s->ops.nn;
    for (size_t i = 0; i < n; ++i) h->s->bed[i] = Dual(bed[i]);
    if (manning) for (int64_t k = 0; k < K; ++k) h->s->manning[k] = Dual(manning[k]);
    return h;
}

void oc_tlm_destroy(void* p) {
    TlmHandle* h = (TlmHandle*)p;
    if (!h) return;
    delete h->s;
    delete h;
}

void oc_tlm_set_tide(void* p, double mean, double ramp, int n, const double* amp, const double* omega, const double* phase) {
    auto& t = ((TlmHandle*)p)->s->tide;
    t.mean = mean; t.ramp = ramp;
    t.amplitude.assign(amp, amp + n); t.omega.assign(omega, omega + n); t.phase.assign(phase, phase + n);
}

void oc_tlm_set_tide_nodal(void* p, int n, const double* f0, const double* f_rate, const double* v0, const double* u0,
                           const double* u_rate) {
    auto& t = ((TlmHandle*)p)->s->tide;
    t.f0.assign(f0, f0 + n); t.f_rate.assign(f_rate, f_rate + n); t.v0.assign(v0, v0 + n);
    t.u0.assign(u0, u0 + n); t.u_rate.assign(u_rate, u_rate + n);
}

// Tangents of the bed and of Manning n (directions of the parameters)
void oc_tlm_set_param_tangents(void* p, const double* dbed, const double* dmanning) {
    Solver<Dual>& s = *((TlmHandle*)p)->s;
    size_t n = (size_t)s.K * s.ops.nn;
    for (size_t i = 0; i < n; ++i) s.bed[i].d = dbed ? dbed[i] : 0.0;
    for (int64_t k = 0; k < s.K; ++k) s.manning[k].d = dmanning ? dmanning[k] : 0.0;
}

static void pack(std::vector<Dual>& v, const double* x, const double* dx, size_t n) {
    v.resize(n);
    for (size_t i = 0; i < n; ++i) v[i] = Dual(x[i], dx ? dx[i] : 0.0);
}
static void unpack(const std::vector<Dual>& v, double* x, double* dx) {
    for (size_t i = 0; i < v.size(); ++i) { x[i] = v[i].v; dx[i] = v[i].d; }
}

// One fused stage on dual numbers: states (value, tangent) in, (value, tangent) out
void oc_tlm_stage(void* p, const double* const* base, const double* const* dbase, const double* const* input,
                  const double* const* dinput, double a, double b, double c, double time, int last,
                  double* const* out, double* const* dout) {
    Solver<Dual>& s = *((TlmHandle*)p)->s;
    size_t n = (size_t)s.K * s.ops.nn;
    std::vector<Dual> B[3], I[3], O[3];
    for (int v = 0; v < 3; ++v) { pack(B[v], base[v], dbase[v], n); pack(I[v], input[v], dinput[v], n); O[v].resize(n); }
    s.stage(B[0].data(), B[1].data(), B[2].data(), I[0].data(), I[1].data(), I[2].data(), O[0].data(), O[1].data(),
            O[2].data(), a, b, c, time, last != 0);
    for (int v = 0; v < 3; ++v) unpack(O[v], out[v], dout[v]);
}

// The element RHS on dual numbers
void oc_tlm_rhs(void* p, const double* const* input, const double* const* dinput, double time, double* const* out,
                double* const* dout) {
    Solver<Dual>& s = *((TlmHandle*)p)->s;
    size_t n = (size_t)s.K * s.ops.nn;
    std::vector<Dual> I[3], O[3];
    for (int v = 0; v < 3; ++v) { pack(I[v], input[v], dinput[v], n); O[v].resize(n); }
    s.rhs(I[0].data(), I[1].data(), I[2].data(), time, O[0].data(), O[1].data(), O[2].data());
    for (int v = 0; v < 3; ++v) unpack(O[v], out[v], dout[v]);
}

// n_steps SSP steps at the prescribed dts, in place on (value, tangent) states
double oc_tlm_run(void* p, double* const* x, double* const* dx, double t, int n_steps, const double* dts, int scheme) {
    TlmHandle* hd = (TlmHandle*)p;
    Solver<Dual>& s = *hd->s;
    size_t n = (size_t)s.K * s.ops.nn;
    std::vector<Dual> S[3];
    for (int v = 0; v < 3; ++v) pack(S[v], x[v], dx[v], n);
    for (int i = 0; i < n_steps; ++i) {
        s.step(S[0].data(), S[1].data(), S[2].data(), t, dts[i], scheme, hd->work);
        t += dts[i];
    }
    for (int v = 0; v < 3; ++v) unpack(S[v], x[v], dx[v]);
    return t;
}

// n_steps SSP steps at prescribed dts on plain doubles (the forward of a frozen-dt adjoint)
double oc_run_dts(void* p, double* h, double* hu, double* hv, double t, int n_steps, const double* dts, int scheme) {
    Handle* hd = (Handle*)p;
    for (int i = 0; i < n_steps; ++i) {
        hd->s->step(h, hu, hv, t, dts[i], scheme, hd->work);
        t += dts[i];
    }
    return t;
}

// Integer wet/dry masks: node h < h_dry, and the element takes the subcells
void oc_masks(void* p, const double* h, uint8_t* node_dry, uint8_t* elem_subcell) {
    Solver<double>& s = *((Handle*)p)->s;
    int nn = s.ops.nn;
    for (int64_t k = 0; k < s.K; ++k) {
        for (int i = 0; i < nn; ++i) node_dry[(size_t)k * nn + i] = h[(size_t)k * nn + i] < s.cfg.h_dry ? 1 : 0;
        elem_subcell[k] = s.takes_subcells(h, k) ? 1 : 0;
    }
}

// Operators and geometry, for the closed-form and metric-identity checks
void oc_operators(int order, double* x1d, double* w1d, double* d1) {
    oracle::Operators ops(order);
    std::copy(ops.x1d.begin(), ops.x1d.end(), x1d);
    std::copy(ops.w1d.begin(), ops.w1d.end(), w1d);
    std::copy(ops.d1.begin(), ops.d1.end(), d1);
}

void oc_geometry(void* p, double* det_j, double* mass, double* jrx, double* jry, double* jsx, double* jsy) {
    Solver<double>& s = *((Handle*)p)->s;
    size_t n = (size_t)s.K * s.ops.nn;
    for (size_t i = 0; i < n; ++i) {
        det_j[i] = s.geom.det_j[i];
        mass[i] = s.geom.mass[i];
        jrx[i] = s.geom.det_j[i] * s.geom.rx[i];
        jry[i] = s.geom.det_j[i] * s.geom.ry[i];
        jsx[i] = s.geom.det_j[i] * s.geom.sx[i];
        jsy[i] = s.geom.det_j[i] * s.geom.sy[i];
    }
}

double oc_cbrt(double x) { return oracle::friction_cbrt(x); }

}  // extern "C"

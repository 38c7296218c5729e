// ORACLE — test infrastructure only (see swe2d_oracle.hpp). The RHS, the
// fused SSP stage, the wet/dry post-process, the time-step bound and the
// step loop of the reference's `WetDry` 2D shallow-water path, restated on
// the CPU with OpenMP over elements.
#pragma once

#include <cstdio>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <cstdlib>
#include "swe2d_oracle.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

struct Config {
    double g = 9.81, h_min = 1e-6;           // equations/shallow_water_2d.rs:67-78
    double h_dry = 1e-3, subcell_depth = 1e-3;  // solver/algorithms/wetting_drying.rs:74; rhs/swe_2d.rs:296
    double relaxation_time = 1.0, max_velocity = 20.0;  // wetting_drying.rs:82-89
    double friction_h_min = 1e-6;            // source/swe_2d/friction.rs:203-209
    double f0 = 0.0;                          // f-plane Coriolis (source/swe_2d/coriolis.rs:150-171)
    double beta = 0.0, beta_y0 = 0.0;         // beta-plane f = f0 + beta (y - y0) (coriolis.rs:118-121,157-169)
    int surface = 0;                          // 0 hydrostatic HLL (WetDry), 1 entropy stable, 2 entropy conservative
    bool subcells = true;                     // WetDry subcell finite volumes
    bool friction = true;                     // Manning, point-implicit
    bool wet_dry = true;                      // wet/dry correction, thin layer, desingularised friction velocity
    bool manufactured = false;                // the manufactured-solution source of tests/convergence_test.rs:961-1020
    bool incoming_wave = false;               // tide data as a wave entering from still water (characteristic.rs:424-426)
};

// SSP schemes (time/integrator.rs:226-263)
struct ShuOsher { double start, input, beta, c; };
inline int scheme_stages(int scheme, ShuOsher* out) {
    static const ShuOsher rk3[3] = {{1.0, 0.0, 1.0, 0.0}, {0.75, 0.25, 0.25, 1.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 0.5}};
    static const ShuOsher rk43[4] = {{1.0, 0.0, 0.5, 0.0}, {0.0, 1.0, 0.5, 0.5}, {2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 1.0}, {0.0, 1.0, 0.5, 0.5}};
    if (scheme == 0) { for (int i = 0; i < 3; ++i) out[i] = rk3[i]; return 3; }
    for (int i = 0; i < 4; ++i) out[i] = rk43[i];
    return 4;
}
inline double ssp_coefficient(int scheme) { return scheme == 0 ? 1.0 : 2.0; }

// rhs/swe_2d.rs:1238-1302
inline double positivity_cfl(int order) { double n = order; return (2.0 * n + 1.0) / (2.0 * n * (n + 1.0)); }
inline double linear_cfl(int order, int scheme) {
    static const double rk3[4] = {1.6, 1.04, 0.73, 0.56}, rk43[4] = {2.2, 1.51, 1.15, 0.93};
    return scheme == 0 ? rk3[order - 1] : rk43[order - 1];
}
inline double linear_cfl_subcells(int order, int scheme) {
    static const double rk3[4] = {0.94, 0.65, 0.47, 0.36}, rk43[4] = {1.5, 1.1, 0.97, 0.75};
    return scheme == 0 ? rk3[order - 1] : rk43[order - 1];
}

template <class T>
struct Solver {
    Operators ops;
    Geometry geom;
    Config cfg;
    Tide tide;
    // BoundaryTides (boundary/tidal_atlas.rs:858-1140): per face node [K][4][N+1]
    // its slot (UINT32_MAX: the uniform tide), per slot [n_c][6] C, S of eta, u, v,
    // the mean level and the source model's depth
    struct Atlas {
        int n_c = 0;
        std::vector<double> omega, f_rate, coef, mean, source_depth;
        std::vector<uint32_t> slot;
        double ramp = 0.0, max_ratio = 0.0;
        bool velocity = false;
    } atlas;
    // NestingRelaxation2D (boundary/ocean_nesting.rs:891-1032): per node [K][nn] its slot (UINT32_MAX:
    // outside the band), the parent's snapshots (Unix s) and per slot [n_times][3] (zeta, u, v) as f32,
    // the band weight and the parent's depth
    struct Nesting {
        bool on = false;
        std::vector<uint32_t> slot;
        std::vector<double> times, weight, depth;
        std::vector<float> series;
        double epoch = 0.0, rate = 0.0, wet_depth = 0.05, max_ratio = 0.0, reference_level = 0.0, ramp = 0.0;
        bool cubic = false, momentum_only = false, velocity = false;
        bool boundary = false;   // the open boundary's external state is the parent's (DeviceExternal2D::Parent)
    } nesting;
    // GriddedAtmosphere2D (source/swe_2d/gridded_atmosphere.rs:137-196, 689-731): per snapshot (Unix s) the
    // regridded fields, field-major [n_times][5][K*nn] f32 (u, v, dp/dx, dp/dy, p - p_ref)
    struct Weather {
        bool on = false, cubic = false, wind = false, pressure = false;
        std::vector<double> times;
        std::vector<float> series;
        double epoch = 0.0, ramp = 0.0, cd = 0.0, rho_air = 1.225, rho_water = 1025.0, h_min = 0.01;
        int drag = 1;
        double raise = 0.0;   // InverseBarometer's factor over the weather's level (0: none)
    } weather;

    // InverseBarometer::external_state's raise (characteristic.rs:324-328) with GriddedAtmosphere2D::level
    // (gridded_atmosphere.rs:873-890): factor * -ramp p / (rho_w G) at the face node's flat index, G = 9.81
    bool barometer_raise(int64_t k, int f, int fi, double time, double& raise) const {
        if (!(weather.on && weather.pressure && weather.raise != 0.0 && fi >= 0)) return false;
        size_t idx[4];
        double w[4];
        const int len = snapshot_stencil(weather.times, weather.cubic, weather.epoch + time, idx, w);
        const size_t total = (size_t)K * ops.nn, n = (size_t)k * ops.nn + ops.face_nodes[f][fi];
        double p = 0.0;
        for (int q = 0; q < len; ++q) p += w[q] * (double)weather.series[(idx[q] * 5 + 4) * total + n];
        double ramp = 1.0;
        if (weather.ramp > 0.0) {
            double tau = std::min(std::max(time / weather.ramp, 0.0), 1.0);
            ramp = tau * tau * (3.0 - 2.0 * tau);
        }
        raise = weather.raise * (-ramp * p / (weather.rho_water * 9.81));
        return true;
    }
    int64_t K = 0;
    std::vector<double> corners;
    std::vector<uint32_t> links;
    std::vector<uint8_t> tags;
    std::vector<T> bed;           // K*nn
    std::vector<T> manning;       // K (per element Manning n)
    std::vector<T> manning_node;  // K*nn per node Manning n (SpatiallyVaryingManning2D, friction.rs:463-553), or empty
    std::vector<State<T>> faces;  // K*4*n1: F* of interior faces, each element's own side
    uint64_t clips = 0;

    Solver(int order, int64_t K_, const double* corners_, const uint32_t* links_, const uint8_t* tags_)
        : ops(order), K(K_) {
        corners.assign(corners_, corners_ + 8 * K);
        links.assign(links_, links_ + 4 * K);
        tags.assign(tags_, tags_ + 4 * K);
        bed.assign((size_t)K * ops.nn, T(0.0));
        manning.assign((size_t)K, T(0.0));
        faces.assign((size_t)K * 4 * ops.n1, State<T>{T(0.0), T(0.0), T(0.0)});
    }
    std::string init_geometry() { return geom.compute(ops, K, corners.data()); }

    NodeState<T> node(const T* h, const T* hu, const T* hv, int64_t k, int i) const {
        size_t n = (size_t)k * ops.nn + i;
        return node_state(h[n], hu[n], hv[n], bed[n], cfg.h_min);
    }

    // TimeStencil::new (io/field_series.rs:41-81) at Unix time t: the snapshots and their weights
    int nesting_stencil(double t, size_t idx[4], double w[4]) const {
        return snapshot_stencil(nesting.times, nesting.cubic, t, idx, w);
    }
    static int snapshot_stencil(const std::vector<double>& ts, bool cubic, double t, size_t idx[4], double w[4]) {
        const size_t n = ts.size();
        idx[0] = 0;
        w[0] = 1.0;
        if (n == 1) return 1;
        if (!(t >= ts[0] && t <= ts[n - 1])) {
            // the reference panics here (OceanModelState::moment_at)
            std::fprintf(stderr, "time %.17g outside the snapshots\n", t);
            std::abort();
        }
        size_t i1 = 0;
        while (i1 < n && ts[i1] < t) ++i1;
        i1 = std::min(std::max(i1, (size_t)1), n - 1);
        size_t i0 = i1 - 1;
        if (t == ts[i0]) { idx[0] = i0; return 1; }
        if (t == ts[i1]) { idx[0] = i1; return 1; }
        size_t lo = cubic ? (i0 == 0 ? 0 : i0 - 1) : i0;
        size_t hi = cubic ? std::min(i1 + 1, n - 1) : i1;
        for (size_t a = lo; a <= hi; ++a) {
            double p = 1.0;
            for (size_t b = lo; b <= hi; ++b)
                if (b != a) p = p * ((t - ts[b]) / (ts[a] - ts[b]));
            idx[a - lo] = a;
            w[a - lo] = p;
        }
        return (int)(hi - lo + 1);
    }

    // relaxation_source with evaluate / parent_term / parent_state (ocean_nesting.rs:583-600, 837-880,
    // 962-988) at a node of slot `slot`, state (qh, qhu, qhv) over bed `b`
    State<double> nesting_source(uint32_t slot, int len, const size_t* idx, const double* w, double ramp,
                                 double qh, double qhu, double qhv, double b) const {
        const Nesting& ns = nesting;
        const double weight = ns.weight[slot];
        if (weight <= 0.0 || qh < ns.wet_depth) return {0.0, 0.0, 0.0};
        double eta, u, v;
        const bool has_velocity = parent_eval(slot, len, idx, w, ramp, b, eta, u, v);
        if (!has_velocity) { u = qhu / qh; v = qhv / qh; }
        const double h = eta - b;
        if (h < ns.wet_depth) return {0.0, 0.0, 0.0};
        const double gamma = weight * ns.rate;
        const double dh = ns.momentum_only ? 0.0 : h - qh;
        const double depth = ns.momentum_only ? qh : h;
        return {gamma * dh, gamma * (depth * u - qhu), gamma * (depth * v - qhv)};
    }

    // OceanModelState::evaluate with parent_term and parent_state (ocean_nesting.rs:583-600, 837-880): the
    // parent's level and, if it has one, its (transport-scaled) velocity at a slot over a child bed b
    bool parent_eval(uint32_t slot, int len, const size_t* idx, const double* w, double ramp, double b, double& eta,
                     double& u, double& v) const {
        const Nesting& ns = nesting;
        double sum[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < len; ++q) {
            const float* v = ns.series.data() + ((size_t)slot * ns.times.size() + idx[q]) * 3;
            for (int c = 0; c < 3; ++c) sum[c] += w[q] * (double)v[c];
        }
        const double zeta = ramp * sum[0];
        eta = zeta + ramp * ns.reference_level;
        u = v = 0.0;
        if (!ns.velocity) return false;
        u = ramp * sum[1];
        v = ramp * sum[2];
        const double pd = ns.depth[slot];
        if (ns.max_ratio > 0.0 && std::isfinite(pd)) {
            const double dp = zeta + pd, dc = eta - b;
            if (dp > 0.0 && dc > 0.0) {
                double r = dp / dc;
                if (r < 1.0 / ns.max_ratio) r = 1.0 / ns.max_ratio;
                if (r > ns.max_ratio) r = ns.max_ratio;
                u *= r;
                v *= r;
            }
        }
        return true;
    }

    // crate::math::hypot (math.rs:205-249): musl's hypot, the squares split exactly (Dekker)
    static double ref_hypot(double x, double y) {
        auto square = [](double v, double& hi, double& lo) {
            const double split = 134217728.0 + 1.0;
            double vc = v * split, vh = v - vc + vc, vl = v - vh;
            hi = v * v;
            lo = vh * vh - hi + 2.0 * vh * vl + vl * vl;
        };
        uint64_t ux, uy;
        std::memcpy(&ux, &x, 8);
        std::memcpy(&uy, &y, 8);
        ux &= ~(1ull << 63);
        uy &= ~(1ull << 63);
        if (ux < uy) std::swap(ux, uy);
        const int64_t ex = (int64_t)(ux >> 52), ey = (int64_t)(uy >> 52);
        std::memcpy(&x, &ux, 8);
        std::memcpy(&y, &uy, 8);
        if (ey == 0x7ff) return y;
        if (ex == 0x7ff || uy == 0) return x;
        if (ex - ey > 64) return x + y;
        double z = 1.0;
        if (ex > 0x3ff + 510) {
            z = std::ldexp(1.0, 700); x *= std::ldexp(1.0, -700); y *= std::ldexp(1.0, -700);
        } else if (ey < 0x3ff - 450) {
            z = std::ldexp(1.0, -700); x *= std::ldexp(1.0, 700); y *= std::ldexp(1.0, 700);
        }
        double hx, lx, hy, ly;
        square(x, hx, lx);
        square(y, hy, ly);
        return z * std::sqrt(ly + lx + hy + hx);
    }

    // DragCoefficient::compute (source/swe_2d/wind.rs:134-158)
    static double drag_cd(int kind, double cd, double s) {
        if (kind == 1) return s <= 11.0 ? 1.2e-3 : (0.49 + 0.065 * std::min(s, 25.0)) * 1e-3;
        if (kind == 2) return (0.8 + 0.065 * s) * 1e-3;
        if (kind == 3) return (0.61 + 0.063 * s) * 1e-3;
        if (kind == 4) return s <= 6.0 ? 1.1e-3 : (0.50 + 0.071 * s) * 1e-3;
        return cd;
    }

    // The face normal at face node fi (the per-node field; a parallelogram's is its node-0 value)
    void normal(int64_t k, int f, int fi, double& nx, double& ny, double& sj) const {
        const ElementGeometry& e = geom.elements[k];
        if (e.affine) { nx = e.fnx[f]; ny = e.fny[f]; sj = e.fsj[f]; return; }
        size_t fidx = ((size_t)k * 4 + f) * ops.n1 + fi;
        nx = geom.nx[fidx]; ny = geom.ny[fidx]; sj = geom.surface_j[fidx];
    }

    // swe_2d_split_form.rs:611-651 surface_flux: (F* from a, F* from b along its own normal)
    void surface_flux(const NodeState<T>& qa, const NodeState<T>& qb, double nx, double ny, State<T>& fa, State<T>& fb) const {
        double g = cfg.g;
        if (cfg.surface == 0) {
            State<T> a, b;
            hydrostatic_hll(side_of(qa), side_of(qb), T(nx), T(ny), g, cfg.h_min, a, b);
            fa = a;
            fb = -b;
            return;
        }
        State<T> shared = wintermeyer_flux(qa, qb, T(nx), T(ny), g);
        if (cfg.surface == 1) shared = shared + es_dissipation(qa, qb, T(nx), T(ny), g);
        fa = shared + wintermeyer_bed_interface(qa.h, qa.b, qb.b, T(nx), T(ny), g);
        fb = -shared + wintermeyer_bed_interface(qb.h, qb.b, qa.b, T(-nx), T(-ny), g);
    }

    // swe_2d_split_form.rs:580-609: F* of both sides of every interior face, once per face
    void face_pass(const T* h, const T* hu, const T* hv) {
        const int n1 = ops.n1;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < K; ++k) {
            for (int f = 0; f < 4; ++f) {
                uint32_t l = links[4 * k + f];
                if (link_is_boundary(l) || !link_is_left(l)) continue;
                int64_t rk = link_element(l);
                int rf = link_face(l);
                for (int fi = 0; fi < n1; ++fi) {
                    double nx, ny, sj;
                    normal(k, f, fi, nx, ny, sj);
                    int rfi = n1 - 1 - fi;
                    NodeState<T> ql = node(h, hu, hv, k, ops.face_nodes[f][fi]);
                    NodeState<T> qr = node(h, hu, hv, rk, ops.face_nodes[rf][rfi]);
                    State<T> fl, fr;
                    surface_flux(ql, qr, nx, ny, fl, fr);
                    faces[((size_t)k * 4 + f) * n1 + fi] = fl;
                    faces[((size_t)rk * 4 + rf) * n1 + rfi] = fr;
                }
            }
        }
    }

    // rhs/swe_2d.rs:304-327 boundary_state + swe_2d_split_form.rs:1090-1123 boundary_flux
    // BoundaryTides::external_state (tidal_atlas.rs:1045-1060 with tide_term, tide_level, tide_state,
    // :1078-1125): the node's sum over constituents at time t, ramped, its mean, the transport scaling
    void atlas_state(uint32_t slot, double time, double bed, double& eta, double& u, double& v) const {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int j = 0; j < atlas.n_c; ++j) {
            // (omega t).sin_cos() (tidal_atlas.rs:966-970): one call of the C library's sincos,
            // as the reference's compiler makes of the pair
            double f = 1.0 + atlas.f_rate[j] * time, cs, sn;
            sincos(atlas.omega[j] * time, &sn, &cs);
            const double* c = atlas.coef.data() + ((size_t)slot * atlas.n_c + j) * 6;
            s0 += f * (c[0] * cs + c[1] * sn);
            s1 += f * (c[2] * cs + c[3] * sn);
            s2 += f * (c[4] * cs + c[5] * sn);
        }
        double ramp = 1.0;
        if (atlas.ramp > 0.0) {
            double tau = std::min(std::max(time / atlas.ramp, 0.0), 1.0);
            ramp = tau * tau * (3.0 - 2.0 * tau);
        }
        eta = atlas.mean[slot] + ramp * s0;
        u = ramp * s1;
        v = ramp * s2;
        if (atlas.velocity && atlas.max_ratio > 0.0) {
            double hs = atlas.source_depth[slot], ds = hs + eta, dc = eta - bed;
            if (hs > 0.0 && ds > 0.0 && dc > 0.0) {
                double r = ds / dc;
                r = r < 1.0 / atlas.max_ratio ? 1.0 / atlas.max_ratio : (r > atlas.max_ratio ? atlas.max_ratio : r);
                u *= r;
                v *= r;
            }
        }
    }

    State<T> boundary_flux(int64_t k, int f, const NodeState<T>& q, double nx, double ny, double eta_e,
                           int fi = -1, double time = 0.0) const {
        double g = cfg.g, hm = cfg.h_min;
        T zero(0.0);
        const uint32_t pslot = (nesting.on && nesting.boundary && fi >= 0)
                                   ? nesting.slot[(size_t)k * ops.nn + ops.face_nodes[f][fi]] : 0xFFFFFFFFu;
        if (tags[4 * k + f] == BC_TIDE && pslot != 0xFFFFFFFFu) {
            // CharacteristicOBC<OceanModelState>: the parent's state at the face node's slot (external_state,
            // ocean_nesting.rs:823-826), raised by the inverse barometer; with a parent velocity, u_n,e = u.n,
            // u_t,e = -u n_y + v n_x (characteristic.rs:411-438)
            size_t idx[4];
            double w[4];
            const int len = nesting_stencil(nesting.epoch + time, idx, w);
            double ramp = 1.0;
            if (nesting.ramp > 0.0) {
                double tau = std::min(std::max(time / nesting.ramp, 0.0), 1.0);
                ramp = tau * tau * (3.0 - 2.0 * tau);
            }
            double eta, ue, ve, raise;
            const bool has_velocity = parent_eval(pslot, len, idx, w, ramp, val(q.b), eta, ue, ve);
            if (barometer_raise(k, f, fi, time, raise)) eta += raise;
            T h_e = vmax(T(eta) - q.b, zero);
            T un_e, ut_e;
            if (has_velocity) {
                un_e = T(ue * nx + ve * ny);
                ut_e = T(-ue * ny + ve * nx);
            } else {
                T u = zero, v = zero;
                if (q.h > T(hm)) { T inv = T(1.0) / q.h; u = q.hu * inv; v = q.hv * inv; }
                un_e = cfg.incoming_wave ? T(2.0) * (sq_sqrt(T(g) * vmax(-q.b, zero)) - sq_sqrt(T(g) * h_e)) : zero;
                ut_e = -u * T(ny) + v * T(nx);
            }
            State<T> qb = characteristic_state(State<T>{q.h, q.hu, q.hv}, h_e, un_e, ut_e, T(nx), T(ny), g, hm);
            NodeState<T> nb = node_state(qb.h, qb.hu, qb.hv, q.b, hm);
            return wintermeyer_flux(nb, nb, T(nx), T(ny), g);
        }
        const uint32_t slot = (atlas.n_c > 0 && fi >= 0) ? atlas.slot[((size_t)k * 4 + f) * ops.n1 + fi] : 0xFFFFFFFFu;
        if (tags[4 * k + f] == BC_TIDE && slot != 0xFFFFFFFFu) {
            // CharacteristicOBC<BoundaryTides>: the node's own tide; with velocity, u_n,e = u.n,
            // u_t,e = -u n_y + v n_x (characteristic.rs:411-438)
            double eta, ue, ve, raise;
            atlas_state(slot, time, val(q.b), eta, ue, ve);
            if (barometer_raise(k, f, fi, time, raise)) eta += raise;
            T h_e = vmax(T(eta) - q.b, zero);
            T un_e, ut_e;
            if (atlas.velocity) {
                un_e = T(ue * nx + ve * ny);
                ut_e = T(-ue * ny + ve * nx);
            } else {
                T u = zero, v = zero;
                if (q.h > T(hm)) { T inv = T(1.0) / q.h; u = q.hu * inv; v = q.hv * inv; }
                un_e = cfg.incoming_wave ? T(2.0) * (sq_sqrt(T(g) * vmax(-q.b, zero)) - sq_sqrt(T(g) * h_e)) : zero;
                ut_e = -u * T(ny) + v * T(nx);
            }
            State<T> qb = characteristic_state(State<T>{q.h, q.hu, q.hv}, h_e, un_e, ut_e, T(nx), T(ny), g, hm);
            NodeState<T> nb = node_state(qb.h, qb.hu, qb.hv, q.b, hm);
            return wintermeyer_flux(nb, nb, T(nx), T(ny), g);
        }
        if (tags[4 * k + f] == BC_TIDE) {
            // CharacteristicOBC<HarmonicTide>, elevation-only data at rest or as an incoming
            // wave from still water, u_n,e = 2 (sqrt(g H) - sqrt(g h_e)), H = -B (characteristic.rs:411-438)
            double raise;
            if (barometer_raise(k, f, fi, time, raise)) eta_e += raise;
            T h_e = vmax(T(eta_e) - q.b, zero);
            T u = zero, v = zero;
            if (q.h > T(hm)) { T inv = T(1.0) / q.h; u = q.hu * inv; v = q.hv * inv; }
            T un_e = cfg.incoming_wave ? T(2.0) * (sq_sqrt(T(g) * vmax(-q.b, zero)) - sq_sqrt(T(g) * h_e)) : zero;
            T ut_e = -u * T(ny) + v * T(nx);
            State<T> qb = characteristic_state(State<T>{q.h, q.hu, q.hv}, h_e, un_e, ut_e, T(nx), T(ny), g, hm);
            NodeState<T> nb = node_state(qb.h, qb.hu, qb.hv, q.b, hm);
            return wintermeyer_flux(nb, nb, T(nx), T(ny), g);
        }
        // Reflective2D ghost (boundary_2d.rs:271-285): free-slip wall
        State<T> ghost{q.h, zero, zero};
        if (q.h > T(hm)) {
            T inv = T(1.0) / q.h;
            T u = q.hu * inv, v = q.hv * inv;
            T un = u * T(nx) + v * T(ny);
            T ug = u - T(2.0) * un * T(nx), vg = v - T(2.0) * un * T(ny);
            ghost = {q.h, q.h * ug, q.h * vg};
        }
        NodeState<T> qe = node_state(ghost.h, ghost.hu, ghost.hv, q.b, hm);
        State<T> fa, fb;
        surface_flux(q, qe, nx, ny, fa, fb);
        return fa;
    }

    // swe_2d_split_form.rs:843-894
    void line_volume(bool affine, const std::vector<NodeState<T>>& nodes, std::vector<State<T>>& rhs,
                     const int* idx, const double* mx, const double* my, double amx, double amy) const {
        const int n1 = ops.n1;
        const double* d1 = ops.d1.data();
        double g = cfg.g;
        for (int a = 0; a < n1; ++a) {
            int ia = idx[a];
            const NodeState<T>& qa = nodes[ia];
            double ma0 = affine ? amx : mx[a], ma1 = affine ? amy : my[a];
            State<T> acc = scale(2.0 * d1[a * n1 + a], wintermeyer_flux(qa, qa, T(ma0), T(ma1), g));
            for (int c = a + 1; c < n1; ++c) {
                int ic = idx[c];
                double m0, m1;
                if (affine) { m0 = ma0; m1 = ma1; }
                else { m0 = 0.5 * (ma0 + mx[c]); m1 = 0.5 * (ma1 + my[c]); }
                State<T> fl = wintermeyer_flux(qa, nodes[ic], T(m0), T(m1), g);
                acc = acc + scale(2.0 * d1[a * n1 + c], fl);
#ifdef ORACLE_MUTANT_TRANSPOSE
                rhs[ic] = rhs[ic] - scale(2.0 * d1[a * n1 + c], fl);  // planted mistake: D transposed
#else
                rhs[ic] = rhs[ic] - scale(2.0 * d1[c * n1 + a], fl);
#endif
            }
            // bathymetry is always set on this path
            T db(0.0), db_x(0.0), db_y(0.0);
            for (int c = 0; c < n1; ++c) {
                T d_b = T(d1[a * n1 + c]) * nodes[idx[c]].b;
                db += d_b;
                if (!affine) { db_x += d_b * T(mx[c]); db_y += d_b * T(my[c]); }
            }
            T bx, by;
            if (affine) { bx = db * T(ma0); by = db * T(ma1); }
            else { bx = T(0.5) * (db * T(ma0) + db_x); by = T(0.5) * (db * T(ma1) + db_y); }
            T force = T(g) * qa.h;
#ifdef ORACLE_MUTANT_NO_BED
            force = T(0.0);  // planted mistake: the bed-slope term dropped
#endif
            acc = acc + State<T>{T(0.0), force * bx, force * by};
            rhs[ia] = rhs[ia] - acc;
        }
    }

    // subcells.rs:63-90
    void telescoped_interfaces(const double* mx, const double* my, bool affine, double* ix, double* iy) const {
        const int n1 = ops.n1;
        if (affine) { for (int a = 0; a <= n1; ++a) { ix[a] = mx[0]; iy[a] = my[0]; } return; }
        ix[0] = mx[0]; iy[0] = my[0];
        double m0 = 0.0, m1 = 0.0;
        for (int a = 0; a < n1 - 1; ++a) {
            for (int c = 0; c < n1; ++c) {
                if (c == a) continue;
                double q2 = 2.0 * ops.w1d[a] * ops.d1[a * n1 + c];
                m0 += q2 * 0.5 * (mx[a] + mx[c]);
                m1 += q2 * 0.5 * (my[a] + my[c]);
            }
            ix[a + 1] = m0; iy[a + 1] = m1;
        }
        ix[n1] = mx[n1 - 1]; iy[n1] = my[n1 - 1];
    }

    struct Outer { bool some; NodeState<T> q; double d; };

    // swe_2d_split_form.rs:1038-1085
    Outer outer_node(const T* h, const T* hu, const T* hv, int64_t k, int face, int fi) const {
        const int n1 = ops.n1;
        uint32_t l = links[4 * k + face];
        if (link_is_boundary(l)) return Outer{false, {}, 0.0};
        int64_t nbk = link_element(l);
        int nbf = link_face(l);
        auto height = [&](int64_t e, int f, int fi_, int node_) {
            const ElementGeometry& eg = geom.elements[e];
            if (eg.affine) return eg.det_j / eg.fsj[f];
            return geom.det_j[(size_t)e * ops.nn + node_] / geom.surface_j[((size_t)e * 4 + f) * n1 + fi_];
        };
        int nb_node = ops.face_nodes[nbf][n1 - 1 - fi];
        int inner = nbf == 0 ? nb_node + n1 : nbf == 1 ? nb_node - 1 : nbf == 2 ? nb_node - n1 : nb_node + 1;
        int own_node = ops.face_nodes[face][fi];
        double distance = (ops.x1d[1] - ops.x1d[0]) * height(nbk, nbf, n1 - 1 - fi, nb_node) / height(k, face, fi, own_node);
        return Outer{true, node(h, hu, hv, nbk, inner), distance};
    }

    // swe_2d_split_form.rs:938-1035
    void line_subcells(bool affine, const std::vector<NodeState<T>>& nodes, std::vector<State<T>>& rhs, const int* idx,
                       const double* mx, const double* my, const Outer ends[2]) const {
        const int n1 = ops.n1;
        const double* w = ops.w1d.data();
        const double* xi = ops.x1d.data();
        double g = cfg.g, h_dry = cfg.h_dry;
        double ix[8], iy[8];
        telescoped_interfaces(mx, my, affine, ix, iy);
        int first = idx[0], last = idx[n1 - 1];
        rhs[first] = rhs[first] + scale(1.0 / w[0], wintermeyer_flux(nodes[first], nodes[first], T(mx[0]), T(my[0]), g));
        rhs[last] = rhs[last] - scale(1.0 / w[n1 - 1], wintermeyer_flux(nodes[last], nodes[last], T(mx[n1 - 1]), T(my[n1 - 1]), g));
        Side<T> fl[8], fr[8];
        double x_left = -1.0;
        for (int a = 0; a < n1; ++a) {
            Side<T> q = side_of(nodes[idx[a]]);
            double x_right = x_left + w[a];
            bool has_l, has_r;
            Side<T> l, r;
            double x_l = 0.0, x_r = 0.0;
            if (a > 0) { has_l = true; l = side_of(nodes[idx[a - 1]]); x_l = xi[a - 1]; }
            else { has_l = ends[0].some; if (has_l) { l = side_of(ends[0].q); x_l = -1.0 - ends[0].d; } }
            if (a + 1 < n1) { has_r = true; r = side_of(nodes[idx[a + 1]]); x_r = xi[a + 1]; }
            else { has_r = ends[1].some; if (has_r) { r = side_of(ends[1].q); x_r = 1.0 + ends[1].d; } }
            if (has_l && has_r) {
                Side<T> st[3] = {l, q, r};
                T xs[3] = {T(x_l), T(xi[a]), T(x_r)};
                reconstruct_subcell(st, xs, T(x_left), T(x_right), h_dry, fl[a], fr[a]);
            } else {
                fl[a] = q; fr[a] = q;
            }
            x_left = x_right;
        }
        for (int a = 0; a < n1 - 1; ++a) {
            int ia = idx[a], ib = idx[a + 1];
            State<T> f_a, f_b;
            hydrostatic_hll(fr[a], fl[a + 1], T(ix[a + 1]), T(iy[a + 1]), g, cfg.h_min, f_a, f_b);
            rhs[ia] = rhs[ia] - scale(1.0 / w[a], f_a);
            rhs[ib] = rhs[ib] + scale(1.0 / w[a + 1], f_b);
        }
        for (int a = 0; a < n1; ++a) {
            int ia = idx[a];
            const Side<T>& left = fl[a];
            const Side<T>& right = fr[a];
            T bedt = T(-0.5 * g) * (left.h + right.h) * (right.b - left.b) / T(w[a]);
            T fx, fy;
            if (affine) { fx = bedt * T(mx[0]); fy = bedt * T(my[0]); }
            else {
                T h_a = nodes[ia].h, b_a = nodes[ia].b;
                T metric = T(0.25 * g) * ((right.h + h_a) * (b_a - right.b) + (left.h + h_a) * (b_a - left.b)) / T(w[a]);
                fx = bedt * T(0.5) * T(ix[a] + ix[a + 1]) + metric * T(ix[a + 1] - ix[a]);
                fy = bedt * T(0.5) * T(iy[a] + iy[a + 1]) + metric * T(iy[a + 1] - iy[a]);
            }
            rhs[ia] = rhs[ia] + State<T>{T(0.0), fx, fy};
        }
    }

    bool takes_subcells(const T* h, int64_t k) const {
        if (!(cfg.surface == 0 && cfg.subcells)) return false;
        for (int i = 0; i < ops.nn; ++i)
            if (h[(size_t)k * ops.nn + i] < T(cfg.subcell_depth)) return true;
        return false;
    }

    // swe_2d_split_form.rs:665-820 element_rhs, then rhs/swe_2d.rs:929-973 sources
    void element_rhs(const T* h, const T* hu, const T* hv, int64_t k, double time, T* oh, T* ohu, T* ohv,
                     std::vector<NodeState<T>>& nodes, std::vector<State<T>>& rhs) const {
        const int n1 = ops.n1, nn = ops.nn;
        const double g = cfg.g;
        for (int i = 0; i < nn; ++i) nodes[i] = node(h, hu, hv, k, i);
        for (int i = 0; i < nn; ++i) rhs[i] = State<T>{T(0.0), T(0.0), T(0.0)};
        bool sub = takes_subcells(h, k);
        Outer outer[4 * 4];
        if (sub)
            for (int f = 0; f < 4; ++f)
                for (int fi = 0; fi < n1; ++fi) outer[f * n1 + fi] = outer_node(h, hu, hv, k, f, fi);
        const ElementGeometry& eg = geom.elements[k];
        bool affine = eg.affine;
        double arx = eg.det_j * eg.rx, ary = eg.det_j * eg.ry, asx = eg.det_j * eg.sx, asy = eg.det_j * eg.sy;
        int idx[8];
        double mx[8], my[8];
        for (int line = 0; line < n1; ++line) {
            int rev = n1 - 1 - line, fwd = line;
            for (int a = 0; a < n1; ++a) {
                idx[a] = line * n1 + a;
                if (affine) { mx[a] = arx; my[a] = ary; }
                else { double c0, c1, c2, c3; geom.contravariant(k, idx[a], c0, c1, c2, c3); mx[a] = c0; my[a] = c1; }
            }
            if (sub) {
                Outer ends[2] = {outer[3 * n1 + rev], outer[n1 + fwd]};
                line_subcells(affine, nodes, rhs, idx, mx, my, ends);
            } else line_volume(affine, nodes, rhs, idx, mx, my, arx, ary);
            for (int j = 0; j < n1; ++j) {
                idx[j] = j * n1 + line;
                if (affine) { mx[j] = asx; my[j] = asy; }
                else { double c0, c1, c2, c3; geom.contravariant(k, idx[j], c0, c1, c2, c3); mx[j] = c2; my[j] = c3; }
            }
            if (sub) {
                Outer ends[2] = {outer[fwd], outer[2 * n1 + rev]};
                line_subcells(affine, nodes, rhs, idx, mx, my, ends);
            } else line_volume(affine, nodes, rhs, idx, mx, my, asx, asy);
        }
        double eta_e = tide.elevation(time);
        for (int f = 0; f < 4; ++f) {
            uint32_t l = links[4 * k + f];
            for (int fi = 0; fi < n1; ++fi) {
                int nd = ops.face_nodes[f][fi];
                double nx, ny, sj;
                normal(k, f, fi, nx, ny, sj);
                const NodeState<T>& q = nodes[nd];
                State<T> fstar = link_is_boundary(l) ? boundary_flux(k, f, q, nx, ny, eta_e, fi, time)
                                                     : faces[((size_t)k * 4 + f) * n1 + fi];
                State<T> diff = wintermeyer_flux(q, q, T(nx), T(ny), g) - fstar;
#ifdef ORACLE_MUTANT_LIFT
                if (f == 1) diff = -diff;  // planted mistake: one face lifted with the wrong sign
#endif
                rhs[nd] = rhs[nd] + scale(sj * ops.lift[f][nd * n1 + fi], diff);
            }
        }
        for (int i = 0; i < nn; ++i) {
            double j_inv = affine ? eg.det_j_inv : geom.det_j_inv[(size_t)k * nn + i];
            size_t n = (size_t)k * nn + i;
            oh[n] = T(j_inv) * rhs[i].h;
            ohu[n] = T(j_inv) * rhs[i].hu;
            ohv[n] = T(j_inv) * rhs[i].hv;
        }
        if (cfg.f0 != 0.0 || cfg.beta != 0.0) {
            // coriolis.rs:150-171: f0 on an f-plane; on a beta-plane f0 + beta (y - y0) at the node's
            // position by the bilinear map (mesh2d.rs:578-595)
            const double* v = corners.data() + 8 * k;
            for (int i = 0; i < nn; ++i) {
                size_t n = (size_t)k * nn + i;
                double f = cfg.f0;
                if (cfg.beta != 0.0) {
                    const double r = ops.r[i], s = ops.s[i];
                    const double n0 = (1.0 - r) * (1.0 - s) / 4.0, n1 = (1.0 + r) * (1.0 - s) / 4.0;
                    const double n2 = (1.0 + r) * (1.0 + s) / 4.0, n3 = (1.0 - r) * (1.0 + s) / 4.0;
                    const double y = n0 * v[1] + n1 * v[3] + n2 * v[5] + n3 * v[7];
                    f = cfg.f0 + cfg.beta * (y - cfg.beta_y0);
                }
#ifdef ORACLE_MUTANT_CORIOLIS_SIGN
                f = -f;  // planted mistake: the rotation turned the wrong way
#endif
                ohu[n] += T(f) * hv[n];
                ohv[n] += T(-f) * hu[n];
            }
        }
        if (nesting.on) {
            // NestingRelaxation2D::add_element (ocean_nesting.rs:1003-1029), at the parent's moment
            size_t idx[4];
            double w[4];
            const int len = nesting_stencil(nesting.epoch + time, idx, w);
            double ramp = 1.0;
            if (nesting.ramp > 0.0) {
                double tau = std::min(std::max(time / nesting.ramp, 0.0), 1.0);
                ramp = tau * tau * (3.0 - 2.0 * tau);
            }
            for (int i = 0; i < nn; ++i) {
                size_t n = (size_t)k * nn + i;
                const uint32_t slot = nesting.slot[n];
                if (slot == 0xFFFFFFFFu) continue;
                State<double> src = nesting_source(slot, len, idx, w, ramp, val(h[n]), val(hu[n]), val(hv[n]),
                                                   val(bed[n]));
                oh[n] += T(src.h); ohu[n] += T(src.hu); ohv[n] += T(src.hv);
            }
        }
        if (weather.on && (weather.wind || weather.pressure)) {
            // GriddedAtmosphere2D::add_element (gridded_atmosphere.rs:703-731) with atmosphere_source (:169-196)
            size_t idx[4];
            double w[4];
            const int len = snapshot_stencil(weather.times, weather.cubic, weather.epoch + time, idx, w);
            double ramp = 1.0;
            if (weather.ramp > 0.0) {
                double tau = std::min(std::max(time / weather.ramp, 0.0), 1.0);
                ramp = tau * tau * (3.0 - 2.0 * tau);
            }
            const double fw = weather.wind ? ramp : 0.0, fp = weather.pressure ? ramp : 0.0;
            const size_t total = (size_t)K * nn;
            for (int i = 0; i < nn; ++i) {
                size_t n = (size_t)k * nn + i;
                const double hh = val(h[n]);
                if (hh < weather.h_min) continue;
                double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                for (int q = 0; q < len; ++q)
                    for (int f = 0; f < 5; ++f)
                        sum[f] += w[q] * (double)weather.series[(idx[q] * 5 + f) * total + n];
                const double u = fw * sum[0], v = fw * sum[1], px = fp * sum[2], py = fp * sum[3];
                const double speed = ref_hypot(u, v);
                const double stress = weather.rho_air * drag_cd(weather.drag, weather.cd, speed) * speed;
                ohu[n] += T((stress * u - hh * px) / weather.rho_water);
                ohv[n] += T((stress * v - hh * py) / weather.rho_water);
            }
        }
        if (cfg.manufactured) {
            for (int i = 0; i < nn; ++i) {
                size_t n = (size_t)k * nn + i;
                double x, y;
                bilinear_point(corners.data() + 8 * k, ops.r[i], ops.s[i], x, y);
                double sh, shu, shv;
                manufactured_source(x, y, time, sh, shu, shv);
                oh[n] += T(sh); ohu[n] += T(shu); ohv[n] += T(shv);
            }
        }
    }

    // tests/convergence_test.rs:961-1005
    static void manufactured_source(double x, double y, double t, double& sh, double& shu, double& shv) {
        const double G = 9.81, U = 0.5, V = 1.5, OX = 2.0 * M_PI * std::sqrt(2.0), OT = 2.0 * M_PI;
        double cx = std::cos(OX * x), sx = std::sin(OX * x), cy = std::cos(OX * y), syy = std::sin(OX * y);
        double ct = std::cos(OT * t), st = std::sin(OT * t);
        double h_t = -OT * cx * syy * st;
        double surface_x = -OX * sx * syy * ct;
        double surface_y = OX * cx * cy * ct;
        double c = 0.5 * std::sqrt(2.0) * M_PI;
        double b_x = c * std::cos(std::sqrt(2.0) * M_PI * x), b_y = c * std::cos(std::sqrt(2.0) * M_PI * y);
        double bed = 2.0 + 0.5 * std::sin(std::sqrt(2.0) * M_PI * x) + 0.5 * std::sin(std::sqrt(2.0) * M_PI * y);
        double h = 7.0 + cx * syy * ct - bed;
        double s_h = h_t + U * (surface_x - b_x) + V * (surface_y - b_y);
        sh = s_h;
        shu = U * s_h + G * h * surface_x;
        shv = V * s_h + G * h * surface_y;
    }

    // The whole RHS at one time (rhs/swe_2d.rs:1478-1624 compute_rhs_swe_2d_parallel_into)
    void rhs(const T* h, const T* hu, const T* hv, double time, T* oh, T* ohu, T* ohv) {
        face_pass(h, hu, hv);
#pragma omp parallel
        {
            std::vector<NodeState<T>> nodes(ops.nn);
            std::vector<State<T>> acc(ops.nn);
#pragma omp for schedule(static)
            for (int64_t k = 0; k < K; ++k) element_rhs(h, hu, hv, k, time, oh, ohu, ohv, nodes, acc);
        }
    }

    // --- Wet/dry correction and damping (solver/algorithms/wetting_drying.rs) ---
    void desingularized_velocity(T h, T hu, T hv, T& u, T& v) const {
        // wetting_drying.rs:115-128
        if (h <= T(0.0)) { u = T(0.0); v = T(0.0); return; }
        double eps = cfg.h_dry;
        if (h >= T(eps)) { T inv = T(1.0) / h; u = hu * inv; v = hv * inv; return; }
        T r = h / T(eps);
        T factor = T(M_SQRT2) * r / (T(eps) * sq_sqrt(T(1.0) + r * r * r * r));
        u = factor * hu; v = factor * hv;
    }
    T thin_layer_rate(T h) const {
        // wetting_drying.rs:133-146
        if (h >= T(cfg.h_dry)) return T(0.0);
        if (h <= T(0.0)) return T(std::numeric_limits<double>::infinity());
        T excess = T(cfg.h_dry) / h - T(1.0);
#ifdef ORACLE_MUTANT_THIN_LAYER
        return excess / T(cfg.relaxation_time);  // planted mistake: the excess not squared
#endif
        return excess * excess / T(cfg.relaxation_time);
    }
    void correct_node(T h, T& hu, T& hv) const {
        // wetting_drying.rs:148-167
        if (h <= T(0.0)) { hu = T(0.0); hv = T(0.0); return; }
        if (h < T(cfg.h_dry)) { T u, v; desingularized_velocity(h, hu, hv, u, v); hu = h * u; hv = h * v; }
        T momentum = sq_sqrt(hu * hu + hv * hv);
        if (momentum > T(cfg.max_velocity) * h) {
            T sc = T(cfg.max_velocity) * h / momentum;
            hu *= sc; hv *= sc;
        }
    }
    // wetting_drying.rs:324-363 damp_node (Manning through friction.rs:71-99,318-324)
    void damp_node(int64_t k, int i, T h, T& hu, T& hv, T fh, T fhu, T fhv, double dt) const {
        if (h <= T(0.0)) { hu = T(0.0); hv = T(0.0); return; }
        T rate(0.0);
        if (cfg.friction) {
            T u, v;
            if (cfg.wet_dry) desingularized_velocity(fh, fhu, fhv, u, v);
            else { u = T(0.0); v = T(0.0); if (fh > T(cfg.h_min)) { T inv = T(1.0) / fh; u = fhu * inv; v = fhv * inv; } }
            T speed = sq_sqrt(u * u + v * v);
            if (speed > T(0.0)) {
                T n = manning_node.empty() ? manning[k] : manning_node[(size_t)k * ops.nn + i];
                T g_n2 = T(cfg.g) * n * n;
                T hh = vmax(h, T(cfg.friction_h_min));
#ifdef ORACLE_MUTANT_MANNING_EXPONENT
                rate += g_n2 * speed / hh;  // planted mistake: h^-1 instead of h^-4/3
#else
                rate += g_n2 * speed / (friction_cbrt(hh) * hh);
#endif
            }
        }
        if (cfg.wet_dry) rate += thin_layer_rate(h);
        T factor = T(1.0) / (T(1.0) + T(dt) * rate);
        hu *= factor; hv *= factor;
    }
    // wetting_drying.rs:194-213 with limiters/swe_2d.rs:104-197
    bool wet_dry_element(int64_t k, T* h, T* hu, T* hv) const {
        const int nn = ops.nn;
        const double* mass = geom.mass.data() + (size_t)k * nn;
        double inv_total = 1.0 / geom.area[k];
        double h_dry = cfg.h_dry;
        bool needs = false;
        for (int i = 0; i < nn; ++i) needs = needs || (h[i] < T(h_dry));
        bool clipped = false;
        if (needs) {
            T sh(0.0), shu(0.0), shv(0.0);
            for (int i = 0; i < nn; ++i) { sh += T(mass[i]) * h[i]; shu += T(mass[i]) * hu[i]; shv += T(mass[i]) * hv[i]; }
            T ah = sh * T(inv_total), ahu = shu * T(inv_total), ahv = shv * T(inv_total);
            if (ah <= T(0.0)) {
                for (int i = 0; i < nn; ++i) { h[i] = T(0.0); hu[i] = T(0.0); hv[i] = T(0.0); }
                clipped = ah < T(0.0);
            } else {
                T hmin(std::numeric_limits<double>::infinity());
                for (int i = 0; i < nn; ++i) hmin = vmin(hmin, h[i]);
                if (hmin < T(0.0)) {
                    T theta = (hmin >= T(0.0)) ? T(1.0) : vmax(T(0.0), vmin(T(1.0), ah / (ah - hmin)));
                    for (int i = 0; i < nn; ++i) {
                        h[i] = theta * (h[i] - ah) + ah;
                        hu[i] = theta * (hu[i] - ahu) + ahu;
                        hv[i] = theta * (hv[i] - ahv) + ahv;
                    }
                    for (int i = 0; i < nn; ++i) h[i] = vmax(h[i], T(0.0));
                }
                if (ah < T(h_dry)) for (int i = 0; i < nn; ++i) { hu[i] = T(0.0); hv[i] = T(0.0); }
            }
        }
        for (int i = 0; i < nn; ++i) correct_node(h[i], hu[i], hv[i]);
        return clipped;
    }

    // builder.rs:652-737 stage_where (one level, all elements): out = a*base + b*input + c*F,
    // damping over c with rates from input, then the wet/dry post-process unless `last`
    // (the last stage's value is the substep result: acc = 1*out, then finish_where)
    void stage(const T* bh, const T* bhu, const T* bhv, const T* ih, const T* ihu, const T* ihv,
               T* oh, T* ohu, T* ohv, double a, double b, double c, double time, bool last) {
        rhs(ih, ihu, ihv, time, oh, ohu, ohv);
        const int nn = ops.nn;
        uint64_t local_clips = 0;
#pragma omp parallel for schedule(static) reduction(+ : local_clips)
        for (int64_t k = 0; k < K; ++k) {
            size_t n0 = (size_t)k * nn;
            const T* base[3] = {bh + n0, bhu + n0, bhv + n0};
            const T* inp[3] = {ih + n0, ihu + n0, ihv + n0};
            T* out[3] = {oh + n0, ohu + n0, ohv + n0};
            for (int var = 0; var < 3; ++var)
                for (int i = 0; i < nn; ++i) {
                    T v = (a != 0.0) ? T(a) * base[var][i] : T(0.0);
                    if (b != 0.0) v += T(b) * inp[var][i];
                    if (c != 0.0) v += T(c) * out[var][i];
                    out[var][i] = v;
                }
            if (cfg.friction || cfg.wet_dry)
                for (int i = 0; i < nn; ++i) damp_node(k, i, out[0][i], out[1][i], out[2][i], inp[0][i], inp[1][i], inp[2][i], c);
            if (last) {
                for (int var = 0; var < 3; ++var)
                    for (int i = 0; i < nn; ++i) out[var][i] = T(1.0) * out[var][i];
            }
            if (cfg.wet_dry) local_clips += wet_dry_element(k, out[0], out[1], out[2]) ? 1 : 0;
        }
        clips += local_clips;
    }

    // multirate.rs:169-200 element_stage: the stage of an element on level `level` whose stage RHS
    // follows level `rate`, in finest substep `substep` (of 2^finest) of depth `depth` of the coarse
    // step of dt from t
    struct ElementStage { double time, a, b, c; bool accumulate; double w, keep; };
    static ElementStage element_stage(double t, double dt, int finest, int substep, int depth, const ShuOsher& st,
                                      bool last, int level, int rate) {
        double dt_l = dt / (double)(1ull << level);
        uint64_t q = (uint64_t)substep >> (finest - level);
        double t_k = t + (double)q * dt_l;
        ElementStage e;
        e.time = t_k + st.c * dt_l;
        e.a = st.start;
        e.b = st.input;
        e.c = st.beta * dt_l;
        e.accumulate = last;
        e.w = last ? 1.0 / (double)(1ull << (rate - level)) : 0.0;
        e.keep = last ? (level >= depth ? 0.0 : 1.0) : 0.0;
        return e;
    }

    // builder.rs:652-737 stage_where: the listed elements' RHS from `input` at their own stage times,
    // out = a*base + b*input + c*F, the damping over c, then acc <- w*out + keep*acc (the last stage)
    // or the wet/dry post-process
    void stage_where(const T* bh, const T* bhu, const T* bhv, const T* ih, const T* ihu, const T* ihv,
                     const uint32_t* elements, int64_t n_list, double t, double dt, int finest, int substep,
                     int depth, const ShuOsher& st, bool last, const uint8_t* level, const uint8_t* rate,
                     T* oh, T* ohu, T* ohv, T* ah, T* ahu, T* ahv) {
        face_pass(ih, ihu, ihv);
        const int nn = ops.nn;
        uint64_t local_clips = 0;
#pragma omp parallel reduction(+ : local_clips)
        {
            std::vector<NodeState<T>> nodes(nn);
            std::vector<State<T>> racc(nn);
#pragma omp for schedule(static)
            for (int64_t li = 0; li < n_list; ++li) {
                const int64_t k = elements[li];
                ElementStage es = element_stage(t, dt, finest, substep, depth, st, last, level[k], rate[k]);
                element_rhs(ih, ihu, ihv, k, es.time, oh, ohu, ohv, nodes, racc);
                size_t n0 = (size_t)k * nn;
                const T* base[3] = {bh + n0, bhu + n0, bhv + n0};
                const T* inp[3] = {ih + n0, ihu + n0, ihv + n0};
                T* out[3] = {oh + n0, ohu + n0, ohv + n0};
                for (int var = 0; var < 3; ++var)
                    for (int i = 0; i < nn; ++i) {
                        T v = (es.a != 0.0) ? T(es.a) * base[var][i] : T(0.0);
                        if (es.b != 0.0) v += T(es.b) * inp[var][i];
                        if (es.c != 0.0) v += T(es.c) * out[var][i];
                        out[var][i] = v;
                    }
                if (cfg.friction || cfg.wet_dry)
                    for (int i = 0; i < nn; ++i)
                        damp_node(k, i, out[0][i], out[1][i], out[2][i], inp[0][i], inp[1][i], inp[2][i], es.c);
                if (es.accumulate) {
                    T* sum[3] = {ah + n0, ahu + n0, ahv + n0};
                    for (int var = 0; var < 3; ++var)
                        for (int i = 0; i < nn; ++i)
                            sum[var][i] = es.keep != 0.0 ? T(es.w) * out[var][i] + T(es.keep) * sum[var][i]
                                                         : T(es.w) * out[var][i];
                } else if (cfg.wet_dry) {
                    local_clips += wet_dry_element(k, out[0], out[1], out[2]) ? 1 : 0;
                }
            }
        }
        clips += local_clips;
    }

    // builder.rs:739-782 finish_where: the listed elements' state from acc, then the post-process
    void finish_where(T* h, T* hu, T* hv, const T* ah, const T* ahu, const T* ahv, const uint32_t* elements,
                      int64_t n_list) {
        const int nn = ops.nn;
        uint64_t local_clips = 0;
#pragma omp parallel for schedule(static) reduction(+ : local_clips)
        for (int64_t li = 0; li < n_list; ++li) {
            const int64_t k = elements[li];
            size_t n0 = (size_t)k * nn;
            std::copy(ah + n0, ah + n0 + nn, h + n0);
            std::copy(ahu + n0, ahu + n0 + nn, hu + n0);
            std::copy(ahv + n0, ahv + n0 + nn, hv + n0);
            if (cfg.wet_dry) local_clips += wet_dry_element(k, h + n0, hu + n0, hv + n0) ? 1 : 0;
        }
        clips += local_clips;
    }

    // multirate.rs:1285-1311 with one level: stage values in their own buffers, the last into acc
    void step(T* h, T* hu, T* hv, double t, double dt, int scheme, std::vector<T>& work) {
        ShuOsher st[4];
        int ns = scheme_stages(scheme, st);
        size_t n = (size_t)K * ops.nn;
        work.resize(3 * n * (size_t)ns);
        T* prev[3] = {h, hu, hv};
        for (int s = 0; s < ns; ++s) {
            T* out[3] = {work.data() + (3 * s) * n, work.data() + (3 * s + 1) * n, work.data() + (3 * s + 2) * n};
            bool last = s + 1 == ns;
            stage(h, hu, hv, prev[0], prev[1], prev[2], out[0], out[1], out[2], st[s].start, st[s].input,
                  st[s].beta * dt, t + st[s].c * dt, last);
            prev[0] = out[0]; prev[1] = out[1]; prev[2] = out[2];
        }
        std::copy(prev[0], prev[0] + n, h);
        std::copy(prev[1], prev[1] + n, hu);
        std::copy(prev[2], prev[2] + n, hv);
    }

    // --- Time step (rhs/swe_2d.rs:1309-1367, 2072-2258; builder.rs:464-485) ---
    double node_rate(double h, double hu, double hv, double grx, double gry, double gsx, double gsy) const {
        double u = 0.0, v = 0.0;
        if (h > cfg.h_min) { double inv = 1.0 / h; u = hu * inv; v = hv * inv; }
        double nr = std::sqrt(grx * grx + gry * gry), ns = std::sqrt(gsx * gsx + gsy * gsy);
        double c = std::sqrt(cfg.g * std::fmax(h, 0.0));
        double lr = std::fabs(u * grx + v * gry) + c * nr;
        double ls = std::fabs(u * gsx + v * gsy) + c * ns;
        return 0.25 * (lr + ls);
    }
    double element_dt(const T* h, const T* hu, const T* hv, int64_t k, double cfl, int scheme, bool positivity) const {
        const int nn = ops.nn, n1 = ops.n1;
        const ElementGeometry& eg = geom.elements[k];
        double rate = 0.0;
        for (int i = 0; i < nn; ++i) {
            size_t n = (size_t)k * nn + i;
            double hh = val(h[n]);
            double r = 0.0;
            if (hh > cfg.h_min) {
                if (eg.affine) r = node_rate(hh, val(hu[n]), val(hv[n]), eg.rx, eg.ry, eg.sx, eg.sy);
                else r = node_rate(hh, val(hu[n]), val(hv[n]), geom.rx[n], geom.ry[n], geom.sx[n], geom.sy[n]);
            }
            rate = std::fmax(rate, r);
        }
        for (int f = 0; f < 4; ++f) {
            uint32_t l = links[4 * k + f];
            if (link_is_boundary(l)) continue;
            int64_t nbk = link_element(l);
            int nbf = link_face(l);
            for (int fi = 0; fi < n1; ++fi) {
                int i = ops.face_nodes[f][fi];
                int j = ops.face_nodes[nbf][n1 - 1 - fi];
                size_t m = (size_t)nbk * nn + j;
                double hh = val(h[m]);
                if (hh <= cfg.h_min) continue;
                size_t n = (size_t)k * nn + i;
                rate = std::fmax(rate, node_rate(hh, val(hu[m]), val(hv[m]), geom.rx[n], geom.ry[n], geom.sx[n], geom.sy[n]));
            }
        }
        if (positivity) {
            double pcfl = ssp_coefficient(scheme) * positivity_cfl(ops.order);
            double max_cfl = std::fmax(pcfl, 0.9 * linear_cfl(ops.order, scheme));
            double max_cfl_sub = std::fmax(pcfl, 0.9 * linear_cfl_subcells(ops.order, scheme));
            double shallow = 100.0 * cfg.h_dry;
            double water = 0.0, water_b = 0.0, weight = 0.0, weight_b = 0.0;
            bool is_shallow = false;
            for (int i = 0; i < nn; ++i) {
                size_t n = (size_t)k * nn + i;
                double hh = val(h[n]), m = geom.mass[n];
                is_shallow = is_shallow || (hh <= shallow);
                double wi = m * std::fmax(hh, 0.0);
                water += wi;
                weight += m;
                int a = i % n1, j = i / n1;
                if (a == 0 || a == n1 - 1 || j == 0 || j == n1 - 1) { water_b += wi; weight_b += m; }
            }
            double ratio = std::fmin(water / water_b, weight / weight_b);
            double rho = std::fmax(0.9 * ratio, 1.0);
            double cap = is_shallow ? max_cfl_sub : max_cfl;
            cfl = std::fmin(cfl, std::fmin(pcfl * rho, cap));
        }
        if (rate < 1e-14) return std::numeric_limits<double>::infinity();
        return cfl / ((2.0 * ops.order + 1.0) * rate);
    }
    double dt_ssp(const T* h, const T* hu, const T* hv, double cfl, int scheme) const {
        double dt = std::numeric_limits<double>::infinity();
#pragma omp parallel for schedule(static) reduction(min : dt)
        for (int64_t k = 0; k < K; ++k) dt = std::fmin(dt, element_dt(h, hu, hv, k, cfl, scheme, cfg.wet_dry));
        return dt;
    }
};

}  // namespace oracle

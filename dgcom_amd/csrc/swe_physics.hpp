// Pointwise physics of the 2D shallow-water `WetDry` split form, for the HIP
// kernels (host + device). Each function names the reference routine it
// computes; the operation order follows the reference so that the kernels
// can be compared with a CPU restatement to the last bit when both are built
// without floating-point contraction.
#pragma once

#ifndef DGSWE_EMULATE
#include <hip/hip_runtime.h>
#endif
#include <math.h>
#include <stdint.h>

#define DG_HD __host__ __device__ __forceinline__

namespace dgswe {

struct Cons { double h, hu, hv; };            // SWEState2D
struct Node { double h, hu, hv, u, v, b; };   // SWENodeState2D
struct Side { double h, u, v, b; };           // SWESide2D

// Per-lane Side and Cons values in LDS as one array per field: a lane stride of
// 8 bytes keeps a wavefront's 64-bit accesses off each other's banks (a struct
// per lane, 24 or 32 bytes apart, puts several lanes on one bank)
template <int M> struct SideRows {
    double h[M], u[M], v[M], b[M];
    DG_HD Side get(int i) const { return {h[i], u[i], v[i], b[i]}; }
    DG_HD void put(int i, const Side& s) { h[i] = s.h; u[i] = s.u; v[i] = s.v; b[i] = s.b; }
};
template <int M> struct ConsRows {
    double h[M], hu[M], hv[M];
    DG_HD Cons get(int i) const { return {h[i], hu[i], hv[i]}; }
    DG_HD void put(int i, const Cons& c) { h[i] = c.h; hu[i] = c.hu; hv[i] = c.hv; }
};

DG_HD Cons cadd(Cons a, Cons b) { return {a.h + b.h, a.hu + b.hu, a.hv + b.hv}; }
DG_HD Cons csub(Cons a, Cons b) { return {a.h - b.h, a.hu - b.hu, a.hv - b.hv}; }
DG_HD Cons cneg(Cons a) { return {-a.h, -a.hu, -a.hv}; }
DG_HD Cons cscale(double c, Cons a) { return {c * a.h, c * a.hu, c * a.hv}; }

// The correctly rounded square root. On the device: the Newton-Raphson refinement of
// the hardware's reciprocal square root, step for step the one the compiler's IEEE
// sqrt expansion runs, without the expansion's rescaling of inputs below 2^-767 and
// its zero / infinity / NaN fix-ups, which run (the full expansion) only where some
// lane needs them: the same bits, about half the instructions on normal inputs.
DG_HD double sqrt_rn(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    if (!(x >= 0x1p-767 && x < __builtin_inf())) {
        // (an empty volatile statement keeps this a branch skipped by the wavefront when no
        // lane takes it, rather than a second sqrt computed everywhere and selected)
        __asm__ volatile("");
        g = __builtin_sqrt(x);
    }
    return g;
#else
    return sqrt(x);
#endif
}

// Rust's f64::max / f64::min on non-NaN operands
DG_HD double fmax_(double a, double b) { return fmax(a, b); }
DG_HD double fmin_(double a, double b) { return fmin(a, b); }

// SWENodeState2D::new / node_velocity (flux/swe_2d_entropy.rs:97-124): one reciprocal,
// zero velocity at or below h_min
DG_HD Node make_node(double h, double hu, double hv, double b, double h_min) {
    double inv_h = 1.0 / h;
    bool wet = h > h_min;
    Node q;
    q.h = h; q.hu = hu; q.hv = hv; q.b = b;
    q.u = wet ? hu * inv_h : 0.0;
    q.v = wet ? hv * inv_h : 0.0;
    return q;
}
DG_HD Side side(const Node& q) { return {q.h, q.u, q.v, q.b}; }

// wintermeyer_flux_2d (flux/swe_2d_entropy.rs:140-158): F#(l, r)·m, bitwise symmetric
DG_HD Cons ec_flux(const Node& l, const Node& r, double mx, double my, double g) {
    double mass = 0.5 * ((l.hu + r.hu) * mx + (l.hv + r.hv) * my);
    double pressure = (0.5 * g) * (l.h * r.h);
    double ua = 0.5 * (l.u + r.u);
    double va = 0.5 * (l.v + r.v);
    return {mass, mass * ua + pressure * mx, mass * va + pressure * my};
}

// The EntropyStable faces' dissipation (swe_2d_entropy.rs:166-203): -1/2 the
// larger of the two sides' |u.n| + sqrt(g h) times the jump of the surface and
// velocity, weighted by the average state; odd in the pair, even in n
DG_HD Cons es_dissipation(const Node& i, const Node& e, double nx, double ny, double g) {
    const double si = fabs(i.u * nx + i.v * ny) + sqrt_rn(g * fmax_(i.h, 0.0));
    const double se = fabs(e.u * nx + e.v * ny) + sqrt_rn(g * fmax_(e.h, 0.0));
    const double sc = -0.5 * fmax_(si, se);
    const double jump = (e.h + e.b) - (i.h + i.b);
    const double ha = 0.5 * (i.h + e.h), ua = 0.5 * (i.u + e.u), va = 0.5 * (i.v + e.v);
    return {sc * jump, sc * (ua * jump + ha * (e.u - i.u)), sc * (va * jump + ha * (e.v - i.v))};
}

// The bed's interface term of one side of an EntropyStable face (swe_2d_split_form.rs:626-651):
// 1/2 g h (B_outside - B_inside) n
DG_HD Cons bed_interface(double h, double b, double b_out, double nx, double ny, double g) {
    const double force = (0.5 * g) * h * (b_out - b);
    return {0.0, force * nx, force * ny};
}

// The nesting band's parent at a node (parent_state, ocean_nesting.rs:851-880):
// the ramped snapshot sum (zeta, u, v), the datum shift ramped with it, and with
// a parent velocity, that velocity scaled by D_parent / D_child (limited either
// way by max_ratio > 0, where the parent depth is known)
DG_HD void parent_state(double s0, double s1, double s2, double ramp, double reference_level, bool velocity,
                        double max_ratio, double parent_depth, double bed, double& eta, bool& has_velocity,
                        double& um, double& vm) {
    const double zeta = ramp * s0, u = ramp * s1, v = ramp * s2;
    eta = zeta + ramp * reference_level;
    has_velocity = velocity;
    um = u;
    vm = v;
    if (velocity && max_ratio > 0.0 && isfinite(parent_depth)) {
        const double d_parent = zeta + parent_depth, d_child = eta - bed;
        if (d_parent > 0.0 && d_child > 0.0) {
            const double ratio = fmin_(fmax_(d_parent / d_child, 1.0 / max_ratio), max_ratio);
            um *= ratio;
            vm *= ratio;
        }
    }
}

// The flow-relaxation source towards the parent (relaxation_source,
// ocean_nesting.rs:962-988): gamma = weight / tau times the parent's depth and
// transport less the node's, nothing where either is shallower than wet_depth
DG_HD Cons relaxation_source(double weight, double rate, double wet_depth, bool momentum_only, double qh, double qhu,
                             double qhv, double bed, double eta, bool has_velocity, double pu, double pv) {
    if (weight <= 0.0 || qh < wet_depth) return {0.0, 0.0, 0.0};
    const double h = eta - bed;
    if (h < wet_depth) return {0.0, 0.0, 0.0};
    const double gamma = weight * rate;
    const double u = has_velocity ? pu : qhu / qh, v = has_velocity ? pv : qhv / qh;
    const double dh = momentum_only ? 0.0 : h - qh;
    const double depth = momentum_only ? qh : h;
    return {gamma * dh, gamma * (depth * u - qhu), gamma * (depth * v - qhv)};
}

// hypot as the reference computes it (math.rs:205-249, musl's hypot): each square
// split exactly into hi + lo (Dekker, 2^27 + 1), scaled by 2^+-700 away from
// overflow and underflow, the square root of the four parts' sum; the same bits on
// every machine with a correctly rounded square root
DG_HD void exact_square(double x, double& hi, double& lo) {
    const double xc = x * 134217729.0;
    const double xh = x - xc + xc;
    const double xl = x - xh;
    hi = x * x;
    lo = xh * xh - hi + 2.0 * xh * xl + xl * xl;
}

DG_HD double hypot_exact(double x, double y) {
    uint64_t ux = __builtin_bit_cast(uint64_t, x) & 0x7FFFFFFFFFFFFFFFull;
    uint64_t uy = __builtin_bit_cast(uint64_t, y) & 0x7FFFFFFFFFFFFFFFull;
    if (ux < uy) { const uint64_t t = ux; ux = uy; uy = t; }
    const int64_t ex = (int64_t)(ux >> 52), ey = (int64_t)(uy >> 52);
    double a = __builtin_bit_cast(double, ux), b = __builtin_bit_cast(double, uy);
    if (ey == 0x7ff) return b;
    if (ex == 0x7ff || uy == 0) return a;
    if (ex - ey > 64) return a + b;
    double z = 1.0;
    if (ex > 0x3ff + 510) {
        z = 0x1p700; a *= 0x1p-700; b *= 0x1p-700;
    } else if (ey < 0x3ff - 450) {
        z = 0x1p-700; a *= 0x1p700; b *= 0x1p700;
    }
    double hx, lx, hy, ly;
    exact_square(a, hx, lx);
    exact_square(b, hy, ly);
    return z * sqrt_rn(ly + lx + hy + hx);
}

// DragCoefficient::compute (source/swe_2d/wind.rs:134-158): 0 a constant cd,
// 1 Large & Pond (wind capped at 25 m/s), 2 Wu, 3 Smith, 4 Yelland & Taylor
DG_HD double drag_coefficient(int kind, double cd, double speed) {
    switch (kind) {
        case 1: return speed <= 11.0 ? 1.2e-3 : (0.49 + 0.065 * fmin_(speed, 25.0)) * 1e-3;
        case 2: return (0.8 + 0.065 * speed) * 1e-3;
        case 3: return (0.61 + 0.063 * speed) * 1e-3;
        case 4: return speed <= 6.0 ? 1.1e-3 : (0.50 + 0.071 * speed) * 1e-3;
        default: return cd;
    }
}

// atmosphere_source (gridded_atmosphere.rs:169-196): (tau - h grad p) / rho_water
// with tau = rho_air C_d(|U|) |U| U, the wind and the gradient already ramped;
// nothing below h_min
DG_HD void atmosphere_source(double h, double u, double v, double px, double py, int drag, double cd,
                             double rho_air, double rho_water, double h_min, double& su, double& sv) {
    su = 0.0;
    sv = 0.0;
    if (h < h_min) return;
    const double speed = hypot_exact(u, v);
    const double stress = rho_air * drag_coefficient(drag, cd, speed) * speed;
    su = (stress * u - h * px) / rho_water;
    sv = (stress * v - h * py) / rho_water;
}

// hll_flux_face_aligned + einfeldt_speeds_2d (flux/swe_2d.rs:277-374), with the
// sides in face-aligned components (depth, normal / tangential momentum and velocity)
struct Hll { double h, hun, hut, un, ut; };

// sqg = sqrt(g), evaluated once by the caller
DG_HD Cons hll_aligned(const Hll& l, const Hll& r, double g, double sqg, double h_min) {
    const double hl = l.h, hr = r.h;
    bool dry = (hl <= h_min) && (hr <= h_min);
    double sql = sqrt_rn(fmax_(hl, 0.0)), sqr = sqrt_rn(fmax_(hr, 0.0));
    double cl = sqg * sql, cr = sqg * sqr;
    bool roe = (sql + sqr) > 1e-10;
    double h_roe = 0.5 * (hl + hr);
    double un_roe = roe ? (sql * l.un + sqr * r.un) / (sql + sqr) : 0.0;
    double c_roe = roe ? sqrt_rn(g * h_roe) : 0.0;
    double s_l = (hl > h_min) ? fmin_(l.un - cl, un_roe - c_roe) : r.un - 2.0 * cr;
    double s_r = (hr > h_min) ? fmax_(r.un + cr, un_roe + c_roe) : l.un + 2.0 * cl;
    double hg = 0.5 * g;
    Cons fl = {l.h * l.un, l.h * l.un * l.un + hg * l.h * l.h, l.h * l.un * l.ut};
    Cons fr = {r.h * r.un, r.h * r.un * r.un + hg * r.h * r.h, r.h * r.un * r.ut};
    double inv_ds = 1.0 / (s_r - s_l);
    double slr = s_l * s_r;
    Cons st = {inv_ds * (s_r * fl.h - s_l * fr.h + slr * (hr - hl)),
               inv_ds * (s_r * fl.hu - s_l * fr.hu + slr * (r.hun - l.hun)),
               inv_ds * (s_r * fl.hv - s_l * fr.hv + slr * (r.hut - l.hut))};
    bool from_l = s_l >= 0.0, from_r = s_r <= 0.0;
    Cons o;
    o.h = dry ? 0.0 : (from_l ? fl.h : (from_r ? fr.h : st.h));
    o.hu = dry ? 0.0 : (from_l ? fl.hu : (from_r ? fr.hu : st.hu));
    o.hv = dry ? 0.0 : (from_l ? fl.hv : (from_r ? fr.hv : st.hv));
    return o;
}

// wall_ram_pressure (swe_2d_split_form.rs:147-156)
DG_HD double ram_pressure(double h, double un, double sill, double g) {
    un = fmax_(un, 0.0);
    double head = un * un * (0.5 / g);
    double ram = (1.0 - head / sill) * h * un * (2.0 * un + sqrt_rn(g * h));
    return (sill > head) ? ram : 0.0;
}

// One side of hydrostatic_hll in face-aligned components: the Audusse depth
// h* = max(0, eta - B_face), the velocity (zero at or below h_min), and the
// velocity HLL's speeds use (zero where h* is at or below h_min)
DG_HD Hll hll_side(const Side& q, double b_face, double nx, double ny, double h_min) {
    double h_star = fmax_(q.h + q.b - b_face, 0.0);
    bool wet = q.h > h_min;
    double u = wet ? q.u : 0.0, v = wet ? q.v : 0.0;
    double un = u * nx + v * ny, ut = -u * ny + v * nx;
    bool star_wet = h_star > h_min;
    return {h_star, h_star * un, h_star * ut, star_wet ? un : 0.0, star_wet ? ut : 0.0};
}

// hydrostatic_hll (swe_2d_split_form.rs:336-402): Audusse states, HLL along m,
// the hydrostatic correction per side; fa as seen from a, fb from b (both along +m)
DG_HD void hydro_hll(const Side& qa, const Side& qb, double mx, double my, double g, double sqg, double h_min, Cons& fa,
                     Cons& fb) {
    double norm = sqrt_rn(mx * mx + my * my);
    double inv_norm = 1.0 / norm;
    double nx = mx * inv_norm, ny = my * inv_norm;
    double b_face = fmax_(qa.b, qb.b);
    const Hll sa = hll_side(qa, b_face, nx, ny, h_min), sb = hll_side(qb, b_face, nx, ny, h_min);
    Cons rot = hll_aligned(sa, sb, g, sqg, h_min);
    double px = rot.hu * nx - rot.hv * ny, py = rot.hu * ny + rot.hv * nx;
    double fh = norm * rot.h, fhu = norm * px, fhv = norm * py;
    bool hll_dry = (sa.h <= h_min) && (sb.h <= h_min);
    double hsa = hll_dry ? 0.0 : sa.h, hsb = hll_dry ? 0.0 : sb.h;
    double wall_a = 0.0, wall_b = 0.0;
    if (hll_dry) {
        double una = 1.0 * (qa.u * nx + qa.v * ny);
        double ra = ram_pressure(qa.h, una, b_face - (qa.h + qa.b), g);
        wall_a = (qa.h > h_min) ? ra : 0.0;
        double unb = -1.0 * (qb.u * nx + qb.v * ny);
        double rb = ram_pressure(qb.h, unb, b_face - (qb.h + qb.b), g);
        wall_b = (qb.h > h_min) ? rb : 0.0;
    }
    double hg = 0.5 * g;
    double dpa = hg * (qa.h * qa.h - hsa * hsa) + wall_a;
    double dpb = hg * (qb.h * qb.h - hsb * hsb) + wall_b;
    fa = {fh + 0.0, fhu + dpa * mx, fhv + dpa * my};
    fb = {fh + 0.0, fhu + dpb * mx, fhv + dpb * my};
}

// limited_slope with SlopeDistances (swe_2d_split_form.rs:168-203)
struct Inv3 { double central, minus, plus; };
DG_HD Inv3 slope_inv(double xm, double xc, double xp, double x_l, double x_r) {
    double fm = xc - x_l, fp = x_r - xc;
    Inv3 r;
    r.central = 1.0 / (xp - xm);
    r.minus = 1.0 / ((fm > 0.0) ? fm : 0.5 * (xc - xm));
    r.plus = 1.0 / ((fp > 0.0) ? fp : 0.5 * (xp - xc));
    return r;
}
DG_HD double lim_slope(double q0, double q1, double q2, const Inv3& inv) {
    double dm = q1 - q0, dp = q2 - q1;
    bool extremum = (dm * dp) <= 0.0;
    double central = (q2 - q0) * inv.central;
    double bound = fmin_(fabs(dm * inv.minus), fabs(dp * inv.plus));
    double sign = (central < 0.0) ? -1.0 : 1.0;
    return extremum ? 0.0 : sign * fmin_(fabs(central), bound);
}

// reconstruct_subcell (swe_2d_split_form.rs:279-317): face states at x_l and x_r
DG_HD void reconstruct(const Side& m, const Side& c, const Side& p, double xm, double xc, double xp, double x_l,
                       double x_r, double h_dry, Side& left, Side& right) {
    double eta0 = m.h + m.b, eta1 = c.h + c.b, eta2 = p.h + p.b;
    double ws = -INFINITY;
    if ((m.h >= h_dry) && (eta0 > ws)) ws = eta0;
    if ((c.h >= h_dry) && (eta1 > ws)) ws = eta1;
    if ((p.h >= h_dry) && (eta2 > ws)) ws = eta2;
    bool flooding = ((m.h < h_dry) && (eta0 < ws)) || ((c.h < h_dry) && (eta1 < ws)) || ((p.h < h_dry) && (eta2 < ws));
    Inv3 inv = slope_inv(xm, xc, xp, x_l, x_r);
    double s_h = lim_slope(m.h, c.h, p.h, inv);
    double s_eta = flooding ? 0.0 : lim_slope(eta0, eta1, eta2, inv);
    double s_u = lim_slope(m.u, c.u, p.u, inv);
    double s_v = lim_slope(m.v, c.v, p.v, inv);
    double d = x_l - xc;
    double hh = fmax_(c.h + s_h * d, 0.0);
    left = {hh, c.u + s_u * d, c.v + s_v * d, eta1 + s_eta * d - hh};
    d = x_r - xc;
    hh = fmax_(c.h + s_h * d, 0.0);
    right = {hh, c.u + s_u * d, c.v + s_v * d, eta1 + s_eta * d - hh};
}

// characteristic_state (boundary/characteristic.rs:481-520) for elevation-only
// data at rest, whose tangential velocity is the interior's (characteristic.rs:411-438)
// characteristic_state from the interior trace and the external depth and
// normal / tangential velocity (characteristic.rs:481-520)
DG_HD Cons characteristic_state(double h, double hu, double hv, double h_e, double un_e, double ut_e, double nx,
                                double ny, double g, double h_min) {
    double h_i = fmax_(h, 0.0);
    double un_i = 0.0, ut_i = 0.0;
    if (h_i > h_min) {
        double u = hu / h_i, v = hv / h_i;
        un_i = u * nx + v * ny;
        ut_i = -u * ny + v * nx;
    }
    double c_i = sqrt_rn(g * h_i), c_e = sqrt_rn(g * h_e);
    double w_out = un_i + 2.0 * c_i, w_in = un_e - 2.0 * c_e;
    bool wi = h_i > h_min, we = h_e > h_min;
    double h_b, un_b;
    if (!wi && !we) { h_b = 0.0; un_b = 0.0; }
    else if (wi && un_i >= c_i) { h_b = h_i; un_b = un_i; }
    else if (we && un_e <= -c_e) { h_b = h_e; un_b = un_e; }
    else if (wi && !we) { double c = fmax_(w_out / 3.0, 0.0); h_b = c * c / g; un_b = c; }
    else if (!wi && we) { double c = fmax_(-w_in / 3.0, 0.0); h_b = c * c / g; un_b = -c; }
    else { double c = fmax_(0.25 * (w_out - w_in), 0.0); h_b = c * c / g; un_b = 0.5 * (w_out + w_in); }
    double ut_b = (un_b > 0.0) ? ut_i : ut_e;
    double ub = un_b * nx - ut_b * ny, vb = un_b * ny + ut_b * nx;
    return {h_b, h_b * ub, h_b * vb};
}

DG_HD Cons characteristic(double h, double hu, double hv, double bed, double eta_e, double nx, double ny, double g,
                          double h_min, bool incoming) {
    double h_e = fmax_(eta_e - bed, 0.0);
    double ui = 0.0, vi = 0.0;
    if (h > h_min) { double inv = 1.0 / h; ui = hu * inv; vi = hv * inv; }
    // at rest, or a simple wave entering from still water at mean sea level
    // (characteristic.rs:424-426): u_n,e = 2 (sqrt(g H) - sqrt(g h_e)), H = -B
    double un_e = incoming ? 2.0 * (sqrt_rn(g * fmax_(-bed, 0.0)) - sqrt_rn(g * h_e)) : 0.0;
    double ut_e = -ui * ny + vi * nx;
    return characteristic_state(h, hu, hv, h_e, un_e, ut_e, nx, ny, g, h_min);
}

// The external state with a velocity (characteristic.rs:419-420): u_n,e = u.n,
// u_t,e = -u n_y + v n_x
DG_HD Cons characteristic_velocity(double h, double hu, double hv, double bed, double eta_e, double ue, double ve,
                                   double nx, double ny, double g, double h_min) {
    double h_e = fmax_(eta_e - bed, 0.0);
    return characteristic_state(h, hu, hv, h_e, ue * nx + ve * ny, -ue * ny + ve * nx, nx, ny, g, h_min);
}

// A boundary node's atlas tide (BoundaryTides, tidal_atlas.rs:1002-1140): the
// sum over constituents of f (C cos + S sin) of eta, u, v with the stage time's
// factors `moment` = [ramp, (f, cos, sin) per constituent], then the ramp, the
// node's mean level and, with a transport ratio limit, the velocity scaled to
// carry the source model's transport over the local bed
constexpr int ATLAS_MAX_CONSTITUENTS = 21;   // the reference's whole constituent set (astronomy.rs:458)
// (the moment: an array in the kernel's arguments, or a row of a table in device memory)
template <class Moment>
DG_HD void atlas_state(const double* c, int n_c, const Moment& moment, double mean,
                       bool velocity, double max_ratio, double source_depth, double bed, double& eta, double& u,
                       double& v) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    // (a fixed trip count, so the moment is read at constant offsets)
#pragma unroll
    for (int j = 0; j < ATLAS_MAX_CONSTITUENTS; ++j) {
        if (j >= n_c) break;
        const double f = moment[1 + 3 * j], cs = moment[2 + 3 * j], sn = moment[3 + 3 * j];
        const double* cj = c + 6 * j;
        s0 += f * (cj[0] * cs + cj[1] * sn);
        s1 += f * (cj[2] * cs + cj[3] * sn);
        s2 += f * (cj[4] * cs + cj[5] * sn);
    }
    const double ramp = moment[0];
    eta = mean + ramp * s0;
    u = ramp * s1;
    v = ramp * s2;
    if (velocity && max_ratio > 0.0) {
        const double d_source = source_depth + eta, d_child = eta - bed;
        if (source_depth > 0.0 && d_source > 0.0 && d_child > 0.0) {
            const double ratio = fmin_(fmax_(d_source / d_child, 1.0 / max_ratio), max_ratio);
            u *= ratio;
            v *= ratio;
        }
    }
}

// Reflective2D ghost state (boundary/boundary_2d.rs:271-285)
DG_HD Cons wall_ghost(double h, double hu, double hv, double nx, double ny, double h_min) {
    if (h <= h_min) return {h, 0.0, 0.0};
    double inv = 1.0 / h;
    double u = hu * inv, v = hv * inv;
    double un = u * nx + v * ny;
    double ug = u - 2.0 * un * nx, vg = v - 2.0 * un * ny;
    return {h, h * ug, h * vg};
}

// Cube root of the Manning rate (source/swe_2d/friction.rs:39-64): Halley from a
// bit-pattern guess; only called on normal positive inputs (depths floored at h_min)
DG_HD double manning_cbrt(double x) {
    uint64_t bits = __double_as_longlong(x);
    double y = __longlong_as_double((long long)(bits / 3 + 0x2A9F7893782DA1CEull));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double y3 = y * y * y;
        y += y * ((x - y3) / (2.0 * y3 + x));
    }
    return y;
}

// Kurganov-Petrova desingularised velocity (wetting_drying.rs:115-128)
DG_HD void desing_velocity(double h, double hu, double hv, double eps, double& u, double& v) {
    if (h <= 0.0) { u = 0.0; v = 0.0; return; }
    if (h >= eps) { double inv = 1.0 / h; u = hu * inv; v = hv * inv; return; }
    double r = h / eps;
    double factor = 1.4142135623730951 * r / (eps * sqrt_rn(1.0 + r * r * r * r));
    u = factor * hu; v = factor * hv;
}

// correct_node (wetting_drying.rs:148-167): desingularise below h_dry, cap the speed
DG_HD void correct_node(double h, double& hu, double& hv, double h_dry, double u_max) {
    if (h <= 0.0) { hu = 0.0; hv = 0.0; return; }
    if (h < h_dry) {
        double u, v;
        desing_velocity(h, hu, hv, h_dry, u, v);
        hu = h * u;
        hv = h * v;
    }
    double momentum = sqrt_rn(hu * hu + hv * hv);
    if (momentum > u_max * h) {
        double sc = u_max * h / momentum;
        hu *= sc;
        hv *= sc;
    }
}

}  // namespace dgswe

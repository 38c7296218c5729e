// Reverse (vector-Jacobian) forms of the pointwise physics of swe_physics.hpp,
// for the discrete adjoint of the wet stage. Each *_adj function recomputes
// what it needs of its forward and adds the cotangents of its inputs, given
// the cotangent of its output. At the kinks (max/min ties, select branches)
// the derivative is that of the branch the forward takes; at a tie of max(a, b)
// or min(a, b) it is a's. The adjoint is of the arithmetic as written (the
// Halley cube root's iterations included), so it is the transpose of the
// tangent-linear model of the same code.
#pragma once

#include "swe_physics.hpp"

namespace dgswe {

// Cotangent of a node state (h, hu, hv, u, v, b)
struct NodeBar { double h, hu, hv, u, v, b; };
struct SideBar { double h, u, v, b; };
struct HllBar { double h, hun, hut, un, ut; };

DG_HD NodeBar nodebar_zero() { return {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; }

// max(a, b) as the derivative sees it: true when a is taken (a > b, or a tie)
DG_HD bool takes_first_max(double a, double b) { return !(b > a); }
DG_HD bool takes_first_min(double a, double b) { return !(b < a); }

// wintermeyer_flux_2d (ec_flux) reverse, the metric constant
DG_HD void ec_flux_adj(const Node& l, const Node& r, double mx, double my, double g, const Cons& fb, NodeBar& lb,
                       NodeBar& rb) {
    const double mass = 0.5 * ((l.hu + r.hu) * mx + (l.hv + r.hv) * my);
    const double ua = 0.5 * (l.u + r.u), va = 0.5 * (l.v + r.v);
    const double massb = fb.h + fb.hu * ua + fb.hv * va;
    const double uab = fb.hu * mass, vab = fb.hv * mass;
    const double pb = (fb.hu * mx + fb.hv * my) * (0.5 * g);
    lb.h += pb * r.h;
    rb.h += pb * l.h;
    lb.u += 0.5 * uab; rb.u += 0.5 * uab;
    lb.v += 0.5 * vab; rb.v += 0.5 * vab;
    const double sx = 0.5 * massb * mx, sy = 0.5 * massb * my;
    lb.hu += sx; rb.hu += sx;
    lb.hv += sy; rb.hv += sy;
}

// make_node reverse: (h, hu, hv, b) cotangents from a node cotangent
DG_HD void make_node_adj(double h, double hu, double hv, double h_min, const NodeBar& nb, double& hb, double& hub,
                         double& hvb, double& bb) {
    hb += nb.h; hub += nb.hu; hvb += nb.hv; bb += nb.b;
    if (h > h_min) {
        const double inv = 1.0 / h;
        hub += nb.u * inv;
        hvb += nb.v * inv;
        const double invb = nb.u * hu + nb.v * hv;
        hb += -invb * inv * inv;
    }
}

// sqrt(max(x, 0)) reverse
DG_HD double sqrt_pos_adj(double x, double yb) {
    const double y = sqrt_rn(fmax_(x, 0.0));
    return (x >= 0.0 && y > 0.0) ? yb / (2.0 * y) : 0.0;
}

// hll_aligned reverse
DG_HD void hll_aligned_adj(const Hll& l, const Hll& r, double g, double h_min, const Cons& ob, HllBar& lb, HllBar& rb) {
    const double hl = l.h, hr = r.h;
    const bool dry = (hl <= h_min) && (hr <= h_min);
    if (dry) return;
    const double sql = sqrt_rn(fmax_(hl, 0.0)), sqr = sqrt_rn(fmax_(hr, 0.0));
    const double sqg = sqrt_rn(g);
    const double cl = sqg * sql, cr = sqg * sqr;
    const bool roe = (sql + sqr) > 1e-10;
    const double h_roe = 0.5 * (hl + hr);
    const double den = sql + sqr;
    const double un_roe = roe ? (sql * l.un + sqr * r.un) / den : 0.0;
    const double c_roe = roe ? sqrt_rn(g * h_roe) : 0.0;
    const double a1 = l.un - cl, a2 = un_roe - c_roe;
    const double b1 = r.un + cr, b2 = un_roe + c_roe;
    const bool lwet = hl > h_min, rwet = hr > h_min;
    const double s_l = lwet ? fmin_(a1, a2) : r.un - 2.0 * cr;
    const double s_r = rwet ? fmax_(b1, b2) : l.un + 2.0 * cl;
    const double hg = 0.5 * g;
    const Cons fl = {l.h * l.un, l.h * l.un * l.un + hg * l.h * l.h, l.h * l.un * l.ut};
    const Cons fr = {r.h * r.un, r.h * r.un * r.un + hg * r.h * r.h, r.h * r.un * r.ut};
    const bool from_l = s_l >= 0.0, from_r = s_r <= 0.0;
    Cons flb = {0.0, 0.0, 0.0}, frb = {0.0, 0.0, 0.0};
    double s_lb = 0.0, s_rb = 0.0;
    if (from_l) flb = ob;
    else if (from_r) frb = ob;
    else {
        const double inv_ds = 1.0 / (s_r - s_l);
        const double slr = s_l * s_r;
        const Cons A = {s_r * fl.h - s_l * fr.h + slr * (hr - hl), s_r * fl.hu - s_l * fr.hu + slr * (r.hun - l.hun),
                        s_r * fl.hv - s_l * fr.hv + slr * (r.hut - l.hut)};
        const double inv_dsb = ob.h * A.h + ob.hu * A.hu + ob.hv * A.hv;
        const Cons Ab = {inv_ds * ob.h, inv_ds * ob.hu, inv_ds * ob.hv};
        s_rb += Ab.h * fl.h + Ab.hu * fl.hu + Ab.hv * fl.hv;
        flb = {Ab.h * s_r, Ab.hu * s_r, Ab.hv * s_r};
        s_lb -= Ab.h * fr.h + Ab.hu * fr.hu + Ab.hv * fr.hv;
        frb = {-Ab.h * s_l, -Ab.hu * s_l, -Ab.hv * s_l};
        const double slrb = Ab.h * (hr - hl) + Ab.hu * (r.hun - l.hun) + Ab.hv * (r.hut - l.hut);
        rb.h += Ab.h * slr; lb.h -= Ab.h * slr;
        rb.hun += Ab.hu * slr; lb.hun -= Ab.hu * slr;
        rb.hut += Ab.hv * slr; lb.hut -= Ab.hv * slr;
        s_lb += slrb * s_r;
        s_rb += slrb * s_l;
        const double dsb = -inv_dsb * inv_ds * inv_ds;
        s_rb += dsb;
        s_lb -= dsb;
    }
    // physical fluxes: f.h = h un, f.hu = (h un) un + hg h h, f.hv = (h un) ut
    {
        const double p = l.h * l.un;
        const double pb = flb.h + flb.hu * l.un + flb.hv * l.ut;
        lb.un += flb.hu * p + pb * l.h;
        lb.ut += flb.hv * p;
        lb.h += pb * l.un + flb.hu * hg * 2.0 * l.h;
    }
    {
        const double p = r.h * r.un;
        const double pb = frb.h + frb.hu * r.un + frb.hv * r.ut;
        rb.un += frb.hu * p + pb * r.h;
        rb.ut += frb.hv * p;
        rb.h += pb * r.un + frb.hu * hg * 2.0 * r.h;
    }
    // Einfeldt speeds
    double clb = 0.0, crb = 0.0, un_roeb = 0.0, c_roeb = 0.0;
    if (lwet) {
        if (takes_first_min(a1, a2)) { lb.un += s_lb; clb -= s_lb; }
        else { un_roeb += s_lb; c_roeb -= s_lb; }
    } else { rb.un += s_lb; crb -= 2.0 * s_lb; }
    if (rwet) {
        if (takes_first_max(b1, b2)) { rb.un += s_rb; crb += s_rb; }
        else { un_roeb += s_rb; c_roeb += s_rb; }
    } else { lb.un += s_rb; clb += 2.0 * s_rb; }
    double sqlb = clb * sqg, sqrb = crb * sqg;
    if (roe) {
        if (c_roe > 0.0) {
            const double h_roeb = c_roeb * g / (2.0 * c_roe);
            lb.h += 0.5 * h_roeb;
            rb.h += 0.5 * h_roeb;
        }
        const double iden = 1.0 / den;
        const double numb = un_roeb * iden;
        const double denb = -un_roeb * un_roe * iden;
        sqlb += numb * l.un + denb;
        sqrb += numb * r.un + denb;
        lb.un += numb * sql;
        rb.un += numb * sqr;
    }
    lb.h += sqrt_pos_adj(hl, sqlb);
    rb.h += sqrt_pos_adj(hr, sqrb);
}

// ram_pressure reverse (its forward is nonzero only when sill > head)
DG_HD void ram_pressure_adj(double h, double un, double sill, double g, double outb, double& hb, double& unb,
                            double& sillb) {
    const double up = fmax_(un, 0.0);
    const double head = up * up * (0.5 / g);
    if (!(sill > head)) return;
    const double A = 1.0 - head / sill;
    const double sq = sqrt_rn(g * h);
    const double C = 2.0 * up + sq;
    // ram = ((A h) up) C
    const double Ab = outb * C * up * h;
    hb += outb * C * up * A;
    double upb = outb * C * A * h;
    const double Cb = outb * A * h * up;
    upb += 2.0 * Cb;
    if (sq > 0.0) hb += Cb * g / (2.0 * sq);
    const double headb = -Ab / sill;
    sillb += Ab * head / (sill * sill);
    upb += headb * 2.0 * up * (0.5 / g);
    if (takes_first_max(un, 0.0)) unb += upb;
}

// One side of hydrostatic_hll reverse: Side cotangent from its Hll cotangent
DG_HD void hll_side_adj(const Side& q, double b_face, double nx, double ny, double h_min, const HllBar& sb, SideBar& qb,
                        double& b_faceb) {
    const double x = q.h + q.b - b_face;
    const double h_star = fmax_(x, 0.0);
    const bool wet = q.h > h_min;
    const double u = wet ? q.u : 0.0, v = wet ? q.v : 0.0;
    const double un = u * nx + v * ny, ut = -u * ny + v * nx;
    const bool star_wet = h_star > h_min;
    double unb = sb.hun * h_star + (star_wet ? sb.un : 0.0);
    double utb = sb.hut * h_star + (star_wet ? sb.ut : 0.0);
    const double h_starb = sb.h + sb.hun * un + sb.hut * ut;
    if (wet) {
        qb.u += unb * nx - utb * ny;
        qb.v += unb * ny + utb * nx;
    }
    if (takes_first_max(x, 0.0)) {
        qb.h += h_starb;
        qb.b += h_starb;
        b_faceb -= h_starb;
    }
}

// hydrostatic_hll reverse: the cotangents of the two sides, given those of (fa, fb)
DG_HD void hydro_hll_adj(const Side& qa, const Side& qb, double mx, double my, double g, double h_min, const Cons& fab,
                         const Cons& fbb, SideBar& qab, SideBar& qbb) {
    const double norm = sqrt_rn(mx * mx + my * my);
    const double inv_norm = 1.0 / norm;
    const double nx = mx * inv_norm, ny = my * inv_norm;
    const double b_face = fmax_(qa.b, qb.b);
    const Hll sa = hll_side(qa, b_face, nx, ny, h_min), sb = hll_side(qb, b_face, nx, ny, h_min);
    const bool hll_dry = (sa.h <= h_min) && (sb.h <= h_min);
    const double hsa = hll_dry ? 0.0 : sa.h, hsb = hll_dry ? 0.0 : sb.h;
    const double hg = 0.5 * g;
    const double fhb = fab.h + fbb.h, fhub = fab.hu + fbb.hu, fhvb = fab.hv + fbb.hv;
    const double dpab = fab.hu * mx + fab.hv * my, dpbb = fbb.hu * mx + fbb.hv * my;
    double b_faceb = 0.0;
    qab.h += dpab * hg * 2.0 * qa.h;
    qbb.h += dpbb * hg * 2.0 * qb.h;
    HllBar sab = {0.0, 0.0, 0.0, 0.0, 0.0}, sbb = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!hll_dry) {
        sab.h += -dpab * hg * 2.0 * hsa;
        sbb.h += -dpbb * hg * 2.0 * hsb;
    } else {
        // walls: wall_a = ram(qa.h, +un_a, b_face - eta_a) where qa.h > h_min, likewise b with -un_b
        if (qa.h > h_min) {
            double hb = 0.0, unb = 0.0, sillb = 0.0;
            const double una = 1.0 * (qa.u * nx + qa.v * ny);
            ram_pressure_adj(qa.h, una, b_face - (qa.h + qa.b), g, dpab, hb, unb, sillb);
            qab.h += hb - sillb;
            qab.b -= sillb;
            b_faceb += sillb;
            qab.u += unb * nx;
            qab.v += unb * ny;
        }
        if (qb.h > h_min) {
            double hb = 0.0, unb = 0.0, sillb = 0.0;
            const double unbv = -1.0 * (qb.u * nx + qb.v * ny);
            ram_pressure_adj(qb.h, unbv, b_face - (qb.h + qb.b), g, dpbb, hb, unb, sillb);
            qbb.h += hb - sillb;
            qbb.b -= sillb;
            b_faceb += sillb;
            qbb.u += -unb * nx;
            qbb.v += -unb * ny;
        }
    }
    // F = norm * rotate_from_normal(rot)
    const double pxb = fhub * norm, pyb = fhvb * norm;
    const Cons rotb = {fhb * norm, pxb * nx + pyb * ny, -pxb * ny + pyb * nx};
    hll_aligned_adj(sa, sb, g, h_min, rotb, sab, sbb);
    hll_side_adj(qa, b_face, nx, ny, h_min, sab, qab, b_faceb);
    hll_side_adj(qb, b_face, nx, ny, h_min, sbb, qbb, b_faceb);
    if (takes_first_max(qa.b, qb.b)) qab.b += b_faceb;
    else qbb.b += b_faceb;
}

// Reflective2D ghost reverse: (h, hu, hv) cotangents from the ghost's
DG_HD void wall_ghost_adj(double h, double hu, double hv, double nx, double ny, double h_min, const Cons& gb,
                          double& hb, double& hub, double& hvb) {
    hb += gb.h;
    if (h <= h_min) return;
    const double inv = 1.0 / h;
    const double u = hu * inv, v = hv * inv;
    const double un = u * nx + v * ny;
    const double ug = u - 2.0 * un * nx, vg = v - 2.0 * un * ny;
    hb += gb.hu * ug + gb.hv * vg;
    const double ugb = gb.hu * h, vgb = gb.hv * h;
    const double unb = -2.0 * (ugb * nx + vgb * ny);
    const double ub = ugb + unb * nx, vb = vgb + unb * ny;
    hub += ub * inv;
    hvb += vb * inv;
    hb += -(ub * hu + vb * hv) * inv * inv;
}

// characteristic (elevation-only data at rest) reverse: cotangents of (h, hu, hv, bed)
// given the boundary state's
DG_HD void characteristic_adj(double h, double hu, double hv, double bed, double eta_e, double nx, double ny, double g,
                              double h_min, bool incoming, const Cons& qbb, double& hb, double& hub, double& hvb,
                              double& bedb) {
    const double xe = eta_e - bed;
    const double h_e = fmax_(xe, 0.0);
    double ui = 0.0, vi = 0.0, inv1 = 0.0;
    if (h > h_min) { inv1 = 1.0 / h; ui = hu * inv1; vi = hv * inv1; }
    const double c_0 = sqrt_rn(g * fmax_(-bed, 0.0));
    const double un_e = incoming ? 2.0 * (c_0 - sqrt_rn(g * h_e)) : 0.0;
    const double ut_e = -ui * ny + vi * nx;
    const double h_i = fmax_(h, 0.0);
    double un_i = 0.0, ut_i = 0.0, u2 = 0.0, v2 = 0.0;
    if (h_i > h_min) {
        u2 = hu / h_i; v2 = hv / h_i;
        un_i = u2 * nx + v2 * ny;
        ut_i = -u2 * ny + v2 * nx;
    }
    const double c_i = sqrt_rn(g * h_i), c_e = sqrt_rn(g * h_e);
    const double w_out = un_i + 2.0 * c_i, w_in = un_e - 2.0 * c_e;
    const bool wi = h_i > h_min, we = h_e > h_min;
    int branch;
    double h_b, un_b, cc = 0.0;
    if (!wi && !we) { branch = 0; h_b = 0.0; un_b = 0.0; }
    else if (wi && un_i >= c_i) { branch = 1; h_b = h_i; un_b = un_i; }
    else if (we && un_e <= -c_e) { branch = 2; h_b = h_e; un_b = un_e; }
    else if (wi && !we) { branch = 3; cc = fmax_(w_out / 3.0, 0.0); h_b = cc * cc / g; un_b = cc; }
    else if (!wi && we) { branch = 4; cc = fmax_(-w_in / 3.0, 0.0); h_b = cc * cc / g; un_b = -cc; }
    else { branch = 5; cc = fmax_(0.25 * (w_out - w_in), 0.0); h_b = cc * cc / g; un_b = 0.5 * (w_out + w_in); }
    const bool out_flow = un_b > 0.0;
    const double ut_b = out_flow ? ut_i : ut_e;
    const double ub = un_b * nx - ut_b * ny, vb = un_b * ny + ut_b * nx;
    // q_b = (h_b, h_b ub, h_b vb)
    double h_bb = qbb.h + qbb.hu * ub + qbb.hv * vb;
    const double ubb = qbb.hu * h_b, vbb = qbb.hv * h_b;
    double un_bb = ubb * nx + vbb * ny;
    const double ut_bb = -ubb * ny + vbb * nx;
    double ut_ib = 0.0, ut_eb = 0.0;
    if (out_flow) ut_ib += ut_bb; else ut_eb += ut_bb;
    double h_ib = 0.0, h_eb = 0.0, un_ib = 0.0, w_outb = 0.0, w_inb = 0.0, un_eb = 0.0;
    double ccb = 0.0;
    switch (branch) {
        case 1: h_ib += h_bb; un_ib += un_bb; break;
        case 2: h_eb += h_bb; un_eb += un_bb; break;
        case 3: ccb = h_bb * 2.0 * cc / g + un_bb;
            if (takes_first_max(w_out / 3.0, 0.0)) w_outb += ccb / 3.0;
            break;
        case 4: ccb = h_bb * 2.0 * cc / g - un_bb;
            if (takes_first_max(-w_in / 3.0, 0.0)) w_inb -= ccb / 3.0;
            break;
        case 5: ccb = h_bb * 2.0 * cc / g;
            if (takes_first_max(0.25 * (w_out - w_in), 0.0)) { w_outb += 0.25 * ccb; w_inb -= 0.25 * ccb; }
            w_outb += 0.5 * un_bb; w_inb += 0.5 * un_bb;
            break;
        default: break;
    }
    un_ib += w_outb;
    un_eb += w_inb;
    // the incoming wave's u_n,e = 2 (c_0 - c_e): c_0 = sqrt(g max(-B, 0)) and c_e
    double c_eb = -2.0 * w_inb;
    if (incoming) {
        c_eb += -2.0 * un_eb;
        if (c_0 > 0.0) bedb += 2.0 * un_eb * (-g) / (2.0 * c_0);
    }
    const double c_ib = 2.0 * w_outb;
    if (c_i > 0.0) h_ib += c_ib * g / (2.0 * c_i);
    if (c_e > 0.0) h_eb += c_eb * g / (2.0 * c_e);
    if (h_i > h_min) {
        const double u2b = un_ib * nx - ut_ib * ny, v2b = un_ib * ny + ut_ib * nx;
        hub += u2b / h_i;
        hvb += v2b / h_i;
        h_ib += -(u2b * u2 + v2b * v2) / h_i;
    }
    if (takes_first_max(h, 0.0)) hb += h_ib;
    // ut_e = -ui ny + vi nx
    if (h > h_min) {
        const double uib = -ut_eb * ny, vib = ut_eb * nx;
        hub += uib * inv1;
        hvb += vib * inv1;
        hb += -(uib * hu + vib * hv) * inv1 * inv1;
    }
    if (takes_first_max(xe, 0.0)) bedb -= h_eb;
}

// Halley cube root reverse (its three iterations from the bit-pattern guess,
// whose derivative is zero): dcbrt/dx seen through the iterations
DG_HD double manning_cbrt_adj(double x, double yb) {
    uint64_t bits = __double_as_longlong(x);
    double y[4];
    y[0] = __longlong_as_double((long long)(bits / 3 + 0x2A9F7893782DA1CEull));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double y3 = y[i] * y[i] * y[i];
        y[i + 1] = y[i] + y[i] * ((x - y3) / (2.0 * y3 + x));
    }
    double xb = 0.0;
#pragma unroll
    for (int i = 2; i >= 0; --i) {
        const double yi = y[i];
        const double y3 = yi * yi * yi;
        const double num = x - y3, den = 2.0 * y3 + x;
        const double q = num / den;
        // y_{i+1} = y_i + y_i q
        double yib = yb * (1.0 + q);
        const double qb = yb * yi;
        const double numb = qb / den, denb = -qb * q / den;
        xb += numb + denb;
        const double y3b = -numb + 2.0 * denb;
        yib += y3b * 3.0 * yi * yi;
        yb = yib;
    }
    return xb;  // the guess carries no derivative
}

// desing_velocity reverse: (h, hu, hv) cotangents from those of (u, v)
DG_HD void desing_velocity_adj(double h, double hu, double hv, double eps, double ub, double vb, double& hb,
                               double& hub, double& hvb) {
    if (h <= 0.0) return;
    if (h >= eps) {
        const double inv = 1.0 / h;
        hub += ub * inv;
        hvb += vb * inv;
        hb += -(ub * hu + vb * hv) * inv * inv;
        return;
    }
    const double r = h / eps;
    const double r4 = r * r * r * r;
    const double s = sqrt_rn(1.0 + r4);
    const double num = 1.4142135623730951 * r, den = eps * s;
    const double factor = num / den;
    const double factorb = ub * hu + vb * hv;
    hub += ub * factor;
    hvb += vb * factor;
    const double numb = factorb / den, denb = -factorb * factor / den;
    double rb = numb * 1.4142135623730951;
    const double r4b = denb * eps / (2.0 * s);
    rb += r4b * 4.0 * r * r * r;
    hb += rb / eps;
}

// correct_node reverse: the cotangents of (h, hu, hv) before the correction, in
// place of those after it
DG_HD void correct_node_adj(double h, double hu, double hv, double h_dry, double u_max, double& hb, double& hub,
                            double& hvb) {
    if (h <= 0.0) { hub = 0.0; hvb = 0.0; return; }
    double u = 0.0, v = 0.0, hu1 = hu, hv1 = hv;
    const bool thin = h < h_dry;
    if (thin) {
        desing_velocity(h, hu, hv, h_dry, u, v);
        hu1 = h * u;
        hv1 = h * v;
    }
    const double m = sqrt_rn(hu1 * hu1 + hv1 * hv1);
    double hu1b = hub, hv1b = hvb;
    if (m > u_max * h) {
        const double sc = u_max * h / m;
        const double scb = hub * hu1 + hvb * hv1;
        hu1b = hub * sc;
        hv1b = hvb * sc;
        const double numb = scb / m, mb = -scb * sc / m;
        hb += numb * u_max;
        hu1b += mb * hu1 / m;
        hv1b += mb * hv1 / m;
    }
    if (thin) {
        hb += hu1b * u + hv1b * v;
        double hb2 = 0.0, hub2 = 0.0, hvb2 = 0.0;
        desing_velocity_adj(h, hu, hv, h_dry, hu1b * h, hv1b * h, hb2, hub2, hvb2);
        hb += hb2;
        hub = hub2;
        hvb = hvb2;
    } else {
        hub = hu1b;
        hvb = hv1b;
    }
}

// lim_slope reverse: cotangents of (q0, q1, q2) given that of the slope
DG_HD void lim_slope_adj(double q0, double q1, double q2, const Inv3& inv, double sb, double& b0, double& b1,
                         double& b2) {
    const double dm = q1 - q0, dp = q2 - q1;
    if ((dm * dp) <= 0.0) return;
    const double central = (q2 - q0) * inv.central;
    const double t1 = dm * inv.minus, t2 = dp * inv.plus;
    const double a1 = fabs(t1), a2 = fabs(t2);
    const bool first = takes_first_min(a1, a2);
    const double bound = first ? a1 : a2;
    const double mb = ((central < 0.0) ? -1.0 : 1.0) * sb;
    // the slope is c * (q_hi - q_lo) for the chosen candidate; values, not
    // references, are selected (selecting a reference would put it in memory)
    const bool use_c = takes_first_min(fabs(central), bound);
    const double cc = use_c ? ((central < 0.0) ? -mb : mb) * inv.central
                            : first ? ((t1 < 0.0) ? -mb : mb) * inv.minus : ((t2 < 0.0) ? -mb : mb) * inv.plus;
    const double w0 = (use_c || first) ? -cc : 0.0;
    const double w1 = use_c ? 0.0 : (first ? cc : -cc);
    const double w2 = (use_c || !first) ? cc : 0.0;
    b0 += w0;
    b1 += w1;
    b2 += w2;
}

// reconstruct reverse: cotangents of the three sides (added) from those of the
// reconstructed face states
DG_HD void reconstruct_adj(const Side& m, const Side& c, const Side& p, double xm, double xc, double xp, double x_l,
                           double x_r, double h_dry, const SideBar& lb, const SideBar& rb, SideBar& mb, SideBar& cb,
                           SideBar& pb) {
    const double eta0 = m.h + m.b, eta1 = c.h + c.b, eta2 = p.h + p.b;
    double ws = -INFINITY;
    if ((m.h >= h_dry) && (eta0 > ws)) ws = eta0;
    if ((c.h >= h_dry) && (eta1 > ws)) ws = eta1;
    if ((p.h >= h_dry) && (eta2 > ws)) ws = eta2;
    const bool flooding =
        ((m.h < h_dry) && (eta0 < ws)) || ((c.h < h_dry) && (eta1 < ws)) || ((p.h < h_dry) && (eta2 < ws));
    const Inv3 inv = slope_inv(xm, xc, xp, x_l, x_r);
    const double s_h = lim_slope(m.h, c.h, p.h, inv);
    double s_hb = 0.0, s_etab = 0.0, s_ub = 0.0, s_vb = 0.0, eta1b = 0.0;
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const SideBar& bb = sd == 0 ? lb : rb;
        const double d = (sd == 0 ? x_l : x_r) - xc;
        const double x = c.h + s_h * d;
        const double hhb = bb.h - bb.b;
        cb.u += bb.u; s_ub += bb.u * d;
        cb.v += bb.v; s_vb += bb.v * d;
        eta1b += bb.b; s_etab += bb.b * d;
        if (takes_first_max(x, 0.0)) { cb.h += hhb; s_hb += hhb * d; }
    }
    lim_slope_adj(m.h, c.h, p.h, inv, s_hb, mb.h, cb.h, pb.h);
    double e0b = 0.0, e2b = 0.0;
    if (!flooding) lim_slope_adj(eta0, eta1, eta2, inv, s_etab, e0b, eta1b, e2b);
    mb.h += e0b; mb.b += e0b;
    cb.h += eta1b; cb.b += eta1b;
    pb.h += e2b; pb.b += e2b;
    lim_slope_adj(m.u, c.u, p.u, inv, s_ub, mb.u, cb.u, pb.u);
    lim_slope_adj(m.v, c.v, p.v, inv, s_vb, mb.v, cb.v, pb.v);
}

}  // namespace dgswe

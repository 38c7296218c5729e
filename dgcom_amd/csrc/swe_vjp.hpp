// Discrete adjoint of the fused stage: the reverse half of the stage kernel's
// VJP instantiation. stage_block<N, SUB, true> recomputes the element's forward
// (node states, metric, face fluxes, volume or subcell term, RHS) with the very
// arithmetic of the forward kernel, then hands its registers and LDS to
// vjp_tail, which recomputes the post-processing (combination, damping, wet/dry
// correction) and runs the element backwards:
//   out cotangent -> correct_node -> theta limiter -> damping -> combination
//   -> Coriolis, 1/J -> volume (or subcell) term and surface term -> node states.
// Cotangents of this element's own nodes are added to the input's cotangent and
// to the bed's, except those through the face Riemann fluxes F*: the element
// writes the cotangent of F* per face node, edge_vjp_kernel runs each Riemann
// solve backwards (one thread per edge node, which keeps the element kernels'
// registers low) and writes the cotangents of both sides per face node, and
// vjp_gather_kernel adds them, with the neighbour's inner node that a subcell
// element reconstructs from, to the nodes afterwards. No two threads write one
// address and the sums do not depend on scheduling. Included by
// swe_kernels.hpp before stage_block.
#pragma once

#include <vector>

#include "swe_adjoint.hpp"

// Three wavefronts per SIMD for the edge kernel (it needs a few registers more)
#ifdef DGSWE_EMULATE
#define DG_WAVES_PER_EU(n)
#else
#define DG_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif

namespace dgswe {

// The reverse stage's per-node reads, issued with the forward's loads at the
// start of the stage kernel: the saved RHS, the base state, Manning n of the
// element, the output cotangent, and the accumulators it adds to (the base's
// and the bed's cotangents)
struct VjpPre { double rh, rhu, rhv, bh, bhu, bhv, mn, wh, whu, whv, gbh, gbhu, gbhv, gbed; };

struct VjpParams {
    StageParams s;                 // the forward stage's inputs (its outputs are not written)
    const double *wh, *whu, *whv;  // cotangent of the stage output (of the RHS when s.rhs_only)
    double *gbh, *gbhu, *gbhv;     // += a * (cotangent of the combination); null: not wanted
    double *gih, *gihu, *gihv;     // += the element's own part of the input's cotangent (= with input_write)
    int input_write;               // the input's cotangent is a fresh buffer: write it instead of adding
    double* gbed;                  // += the element's own part of the bed's cotangent
    double* gman;                  // += per element Manning n cotangent (null: not wanted)
    double* facecot;               // [K][4(N+1)][3]: cotangent of F* at each face node (element kernels)
    double* facebuf;               // [K][4(N+1)][4]: cotangent of the neighbour's face node (h, hu, hv, B)
    double* innerbuf;              // [K][4(N+1)][4]: of the neighbour's inner node (subcell elements)
    uint8_t* kind;                 // [K]: 1 where the element takes the subcell volume term
};

DG_HD void nbar_add(NodeBar& a, const NodeBar& b) {
    a.h += b.h; a.hu += b.hu; a.hv += b.hv; a.u += b.u; a.v += b.v; a.b += b.b;
}
DG_HD void nbar_add_side(NodeBar& a, const SideBar& b) { a.h += b.h; a.u += b.u; a.v += b.v; a.b += b.b; }
DG_HD SideBar sbar_zero() { return {0.0, 0.0, 0.0, 0.0}; }

// Index of the inner node behind face node fi of face f of an element (the
// node one step into the element along the GLL line through the face node)
template <int N> __device__ __forceinline__ int inner_node(int f, int fi) {
    constexpr int N1 = N + 1;
    const int fn = face_node<N>(f, fi);
    return f == 0 ? fn + N1 : f == 1 ? fn - 1 : f == 2 ? fn - N1 : fn + 1;
}

// Cotangent of a neighbour's inner node read by outer_node for face node fi
// of face f: the (h, hu, hv, B) cotangents, written to the inner-node buffer
template <int N>
__device__ __forceinline__ void outer_node_adj(const VjpParams& P, int64_t k, int f, int fi, const SideBar& ob) {
    constexpr int N1 = N + 1, NN = N1 * N1, NF = 4 * N1;
    const StageParams& p = P.s;
    const uint32_t l = p.links[4 * k + f];
    const int64_t nb = (int64_t)(l >> 4);
    const int nbf = (int)(l & 3u);
    const size_t m = (size_t)nb * NN + inner_node<N>(nbf, N1 - 1 - fi);
    double hb = 0.0, hub = 0.0, hvb = 0.0, bb = 0.0;
    const NodeBar q = {ob.h, 0.0, 0.0, ob.u, ob.v, ob.b};
    make_node_adj(p.ih[m], p.ihu[m], p.ihv[m], p.h_min, q, hb, hub, hvb, bb);
    double* o = P.innerbuf + ((size_t)k * NF + f * N1 + fi) * 4;
    o[0] = hb; o[1] = hub; o[2] = hvb; o[3] = bb;
}

template <int N, bool SUB>
__device__ __forceinline__ void vjp_tail(const VjpParams& P, int64_t k, bool active, const Quad& c, bool affine,
                                         double det, double det_inv, double hin, double huin, double hvin, double bin,
                                         const Cons& rhs, const VjpPre& pre, const double* sh,
                                         const double* shu, const double* shv, const double* su, const double* sv,
                                         const double* sb, const double* srx, const double* sry, const double* ssx,
                                         const double* ssy, const SideRows<SUB ? 64 : 1>* rl,
                                         const SideRows<SUB ? 64 : 1>* rr) {
    const StageParams& p = P.s;
    constexpr int N1 = N + 1, NN = N1 * N1, EPB = 64 / NN, NF = 4 * N1;
    constexpr int NT = (NF + NN - 1) / NN;
    const GllTables& T = c_gll[N];
    // One LDS region for the phases' own arrays (separated by lds_sync): the
    // wet/dry recompute (xh..) and its reverse (gy.., gtb), pw through both, the
    // bed term's line cotangents (sdb), the face nodes' own cotangents (tcot),
    // the Manning reduction (sman). Less LDS per workgroup, more workgroups per CU.
    constexpr int NREG = EPB * NF * 6 > 6 * 64 ? EPB * NF * 6 : 6 * 64;
    __shared__ double region[NREG];
    double *xh = region, *xhu = region + 64, *xhv = region + 128, *pw = region + 256;
    double *gyh = region, *gyhu = region + 64, *gyhv = region + 128, *gtb = region + 192;
    double(*sdb)[3][64] = reinterpret_cast<double(*)[3][64]>(region);
    NodeBar* tcot = reinterpret_cast<NodeBar*>(region);
    double* sman = region;
    __shared__ double red[EPB * 8];
    __shared__ int wfl[EPB];
    __shared__ double swh[64], swhu[64], swhv[64];
    __shared__ SideBar rlb[2][SUB ? 64 : 1], rrb[2][SUB ? 64 : 1], rlx[2][SUB ? 64 : 1];
    // the interface fluxes' cotangents (ifab, ifbb: until the interface HLLs'
    // reverse) and the reconstruction's cotangents for the line neighbours
    // (tol, tor: from the reconstruction's reverse on) share one array
    constexpr int SL = SUB ? 64 : 1;
    __shared__ double line_ovl[2 * 2 * SL * (sizeof(SideBar) / sizeof(double))];
    SideBar(*tol)[SL] = reinterpret_cast<SideBar(*)[SL]>(line_ovl);
    SideBar(*tor)[SL] = reinterpret_cast<SideBar(*)[SL]>(line_ovl + 2 * SL * (sizeof(SideBar) / sizeof(double)));
    Cons(*ifab)[SL] = reinterpret_cast<Cons(*)[SL]>(line_ovl);
    Cons(*ifbb)[SL] = reinterpret_cast<Cons(*)[SL]>(line_ovl + 2 * SL * (sizeof(Cons) / sizeof(double)));

    const int t = threadIdx.x;
    const int e = t / NN;
    const int i = t - e * NN;
    const int a = i % N1, j = i / N1;
    const int eb = e * NN;
    const size_t n = (size_t)k * NN + i;
    const bool post = !p.rhs_only;
    const bool friction = (p.flags & DG_FRICTION) != 0, wet_dry = (p.flags & DG_WET_DRY) != 0;

    // ---- the post-processing, recomputed as the forward stage does it ----
    double vh = 0.0, vhu = 0.0, vhv = 0.0;  // the combination a*base + b*input + c*rhs
    double fac = 1.0, rate = 0.0;
    bool zeroed = false, damped = false, fric = false, thin = false;
    double uf = 0.0, vf = 0.0, speed = 0.0, mn = 0.0, hh = 0.0, cbr = 0.0;
    if (active && post) {
        if (p.a != 0.0) { vh = p.a * pre.bh; vhu = p.a * pre.bhu; vhv = p.a * pre.bhv; }
        if (p.b != 0.0) { vh += p.b * hin; vhu += p.b * huin; vhv += p.b * hvin; }
        if (p.c != 0.0) { vh += p.c * rhs.h; vhu += p.c * rhs.hu; vhv += p.c * rhs.hv; }
        if (friction || wet_dry) {
            if (vh <= 0.0) zeroed = true;
            else {
                if (friction) {
                    if (wet_dry) desing_velocity(hin, huin, hvin, p.h_dry, uf, vf);
                    else if (hin > p.h_min) { double inv = 1.0 / hin; uf = huin * inv; vf = hvin * inv; }
                    speed = sqrt_rn(uf * uf + vf * vf);
                    if (speed > 0.0) {
                        fric = true;
                        mn = pre.mn;
                        double gn2 = p.g * mn * mn;
                        hh = fmax_(vh, p.h_min);
                        cbr = manning_cbrt(hh);
                        rate += gn2 * speed / (cbr * hh);
                    }
                }
                if (wet_dry && vh < p.h_dry) {
                    thin = true;
                    double ex = p.h_dry / vh - 1.0;
                    rate += ex * ex / p.tau;
                }
                fac = 1.0 / (1.0 + p.c * rate);
                damped = true;
            }
        }
    }
    // after damping
    const double dh = vh;
    const double dhu = zeroed ? 0.0 : (damped ? vhu * fac : vhu);
    const double dhv = zeroed ? 0.0 : (damped ? vhv * fac : vhv);

    // wet/dry correction: the theta limiter over the element, then correct_node
    double yh = dh, yhu = dhu, yhv = dhv, yph = dh;  // after the limiter (yph: h before its clamp at 0)
    bool needs = false;
    double ah = 0.0, ahu = 0.0, ahv = 0.0, hmin = 0.0, theta = 1.0, inv_total = 0.0;
    int imin = 0;
    if (post && wet_dry) {
        if (t < EPB) wfl[t] = 0;
        lds_sync();
        if (active) {
            pw[t] = T.w2[i] * det;
            xh[t] = dh; xhu[t] = dhu; xhv[t] = dhv;
            if (dh < p.h_dry) wfl[e] = 1;
        }
        lds_sync();
        needs = active && wfl[e] != 0;
        if (needs && i == 0) {
            double area = 0.0;
            for (int q = 0; q < NN; ++q) area += pw[eb + q];
            double it = 1.0 / area;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, mnh = INFINITY;
            int im = 0;
            for (int q = 0; q < NN; ++q) {
                double w = pw[eb + q];
                s0 += w * xh[eb + q];
                s1 += w * xhu[eb + q];
                s2 += w * xhv[eb + q];
            }
            for (int q = 0; q < NN; ++q)
                if (xh[eb + q] < mnh) { mnh = xh[eb + q]; im = q; }
            red[e * 8 + 0] = s0 * it;
            red[e * 8 + 1] = s1 * it;
            red[e * 8 + 2] = s2 * it;
            red[e * 8 + 3] = mnh;
            red[e * 8 + 4] = it;
            red[e * 8 + 5] = (double)im;
        }
        lds_sync();
        if (needs) {
            ah = red[e * 8 + 0]; ahu = red[e * 8 + 1]; ahv = red[e * 8 + 2];
            hmin = red[e * 8 + 3]; inv_total = red[e * 8 + 4]; imin = (int)red[e * 8 + 5];
            if (ah <= 0.0) { yh = 0.0; yhu = 0.0; yhv = 0.0; }
            else {
                if (hmin < 0.0) {
                    theta = fmax_(0.0, fmin_(1.0, ah / (ah - hmin)));
                    yph = theta * (dh - ah) + ah;
                    yhu = theta * (dhu - ahu) + ahu;
                    yhv = theta * (dhv - ahv) + ahv;
                    yh = fmax_(yph, 0.0);
                }
                if (ah < p.h_dry) { yhu = 0.0; yhv = 0.0; }
            }
        }
    }

    // ---- reverse ----
    double gh = 0.0, ghu = 0.0, ghv = 0.0;
    if (active) { gh = pre.wh; ghu = pre.whu; ghv = pre.whv; }
    double gi_h = 0.0, gi_hu = 0.0, gi_hv = 0.0;  // input cotangent, this node
    double gman = 0.0;
    Cons rb = {0.0, 0.0, 0.0};                     // cotangent of the RHS
    if (post) {
        if (wet_dry) {
            if (active) correct_node_adj(yh, yhu, yhv, p.h_dry, p.u_max, gh, ghu, ghv);
            // theta limiter: y' = theta (x - mean) + mean, h clamped at 0, means over the element
            bool lim = false;
            if (needs) {
                if (ah <= 0.0) { gh = 0.0; ghu = 0.0; ghv = 0.0; }
                else {
                    if (ah < p.h_dry) { ghu = 0.0; ghv = 0.0; }
                    if (hmin < 0.0) {
                        lim = true;
                        if (!takes_first_max(yph, 0.0)) gh = 0.0;
                    }
                }
            }
            gyh[t] = gh; gyhu[t] = ghu; gyhv[t] = ghv;
            gtb[t] = lim ? gh * (dh - ah) + ghu * (dhu - ahu) + ghv * (dhv - ahv) : 0.0;
            lds_sync();
            if (needs && i == 0 && ah > 0.0) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, tb = 0.0, hminb = 0.0;
                for (int q = 0; q < NN; ++q) { s0 += gyh[eb + q]; s1 += gyhu[eb + q]; s2 += gyhv[eb + q]; tb += gtb[eb + q]; }
                double a0 = 0.0, a1 = 0.0, a2 = 0.0;
                if (hmin < 0.0) {
                    // y' = theta x + (1 - theta) mean
                    a0 = (1.0 - theta) * s0; a1 = (1.0 - theta) * s1; a2 = (1.0 - theta) * s2;
                    const double r = ah / (ah - hmin);
                    const double inner = fmin_(1.0, r);
                    if (!takes_first_min(1.0, r) && !takes_first_max(0.0, inner)) {
                        const double den = ah - hmin;
                        const double denb = -tb * r / den;
                        a0 += tb / den + denb;
                        hminb = -denb;
                    }
                } else { a0 = s0; a1 = s1; a2 = s2; }
                if (hmin < 0.0) {
                    red[e * 8 + 0] = a0 * inv_total; red[e * 8 + 1] = a1 * inv_total; red[e * 8 + 2] = a2 * inv_total;
                } else {
                    // no limiting: the element passes through (the means do not enter)
                    red[e * 8 + 0] = 0.0; red[e * 8 + 1] = 0.0; red[e * 8 + 2] = 0.0;
                }
                red[e * 8 + 3] = hminb;
            }
            lds_sync();
            if (needs && ah > 0.0 && hmin < 0.0) {
                gh = theta * gh + pw[t] * red[e * 8 + 0];
                ghu = theta * ghu + pw[t] * red[e * 8 + 1];
                ghv = theta * ghv + pw[t] * red[e * 8 + 2];
                if (i == imin) gh += red[e * 8 + 3];
            }
        }
        // damping: (hu, hv) / (1 + c rate), rate from the input's velocity and the combination's depth
        if (active && (friction || wet_dry)) {
            if (zeroed) { ghu = 0.0; ghv = 0.0; }
            else if (damped) {
                const double facb = ghu * vhu + ghv * vhv;
                ghu *= fac;
                ghv *= fac;
                const double rateb = -facb * fac * fac * p.c;
                if (thin) {
                    const double q = p.h_dry / vh;
                    const double ex = q - 1.0;
                    const double exb = rateb * 2.0 * ex / p.tau;
                    gh += -exb * q / vh;
                }
                if (fric) {
                    const double gn2 = p.g * mn * mn;
                    const double num = gn2 * speed, den = cbr * hh;
                    const double iden = 1.0 / den;
                    const double term = num * iden;
                    const double numb = rateb * iden, denb = -rateb * term * iden;
                    const double gn2b = numb * speed, speedb = numb * gn2;
                    const double cbb = denb * hh;
                    // the Halley iterate's derivative is that of the cube root, 1/(3 y^2),
                    // to rounding (the iteration converges cubically)
                    const double hhb = denb * cbr + cbb / (3.0 * cbr * cbr);
                    if (takes_first_max(vh, p.h_min)) gh += hhb;
                    gman += gn2b * (p.g * mn + p.g * mn);
                    const double isp = 1.0 / speed;
                    const double ub = speedb * uf * isp, vb = speedb * vf * isp;
                    if (wet_dry) desing_velocity_adj(hin, huin, hvin, p.h_dry, ub, vb, gi_h, gi_hu, gi_hv);
                    else if (hin > p.h_min) {
                        const double inv = 1.0 / hin;
                        gi_hu += ub * inv;
                        gi_hv += vb * inv;
                        gi_h += -(ub * huin + vb * hvin) * inv * inv;
                    }
                }
            }
        }
        // combination
        if (active) {
            if (p.a != 0.0 && P.gbh) {
                P.gbh[n] = pre.gbh + p.a * gh; P.gbhu[n] = pre.gbhu + p.a * ghu; P.gbhv[n] = pre.gbhv + p.a * ghv;
            }
            if (p.b != 0.0) { gi_h += p.b * gh; gi_hu += p.b * ghu; gi_hv += p.b * ghv; }
            if (p.c != 0.0) rb = {p.c * gh, p.c * ghu, p.c * ghv};
        }
    } else rb = {gh, ghu, ghv};

    // Coriolis and 1/J
    if (active && p.f0 != 0.0) {
        gi_hv += p.f0 * rb.hu;
        gi_hu += (-p.f0) * rb.hv;
    }
    const Cons W = {det_inv * rb.h, det_inv * rb.hu, det_inv * rb.hv};
    swh[t] = W.h; swhu[t] = W.hu; swhv[t] = W.hv;
    lds_sync();

    NodeBar nb = nodebar_zero();  // cotangent of this node's state (h, hu, hv, u, v, B)
    const int rlb0 = eb + j * N1, slb0 = eb + a;
    if (!SUB) {
        // flux-differencing volume term with the bed slope, per line
        if (active) {
            const Node me = {sh[t], shu[t], shv[t], su[t], sv[t], sb[t]};
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int lb = dir == 0 ? rlb0 : slb0, ls = dir == 0 ? 1 : N1, pp = dir == 0 ? a : j;
                const double* smx = dir == 0 ? srx : ssx;
                const double* smy = dir == 0 ? sry : ssy;
                const double mpx = smx[t], mpy = smy[t];
                const Cons Wp = {swh[t], swhu[t], swhv[t]};
                double db = 0.0, dbx = 0.0, dby = 0.0;
#pragma unroll 1
                for (int q = 0; q < N1; ++q) {
                    const int iq = lb + q * ls;
                    const Node o = {sh[iq], shu[iq], shv[iq], su[iq], sv[iq], sb[iq]};
                    const double dpq = T.d1[pp * N1 + q], dqp = T.d1[q * N1 + pp];
                    const double mqx = smx[iq], mqy = smy[iq];
                    // The pair's flux is F#(q_q, q_p): its cotangent w collects both
                    // ends' terms, and this node takes the right argument's part.
                    // The diagonal pair (q = p) enters both arguments with the same
                    // state; its two equal parts are one part of twice its w, which
                    // is what the off-diagonal formula gives there (D_pp twice, W_p
                    // twice). The metric average at q = p and on an affine element
                    // (one metric at every node) is the node's metric exactly. So
                    // every pair takes one code path: a branch on q = p would
                    // diverge within the wavefront and run both.
                    const double mx = 0.5 * (mqx + mpx);
                    const double my = 0.5 * (mqy + mpy);
                    const Cons Wq = {swh[iq], swhu[iq], swhv[iq]};
                    const Cons w = cadd(cscale(-2.0 * dpq, Wp), cscale(-2.0 * dqp, Wq));
                    NodeBar other = nodebar_zero();
                    ec_flux_adj(o, me, mx, my, p.g, w, other, nb);
                    const double d_b = dpq * o.b;
                    db += d_b;
                    dbx += d_b * mqx;
                    dby += d_b * mqy;
                }
                double bx, by;
                if (affine) { bx = db * mpx; by = db * mpy; }
                else { bx = 0.5 * (db * mpx + dbx); by = 0.5 * (db * mpy + dby); }
                nb.h += -p.g * (bx * Wp.hu + by * Wp.hv);
                const double bxb = -p.g * me.h * Wp.hu, byb = -p.g * me.h * Wp.hv;
                if (affine) {
                    sdb[dir][0][t] = bxb * mpx + byb * mpy;
                    sdb[dir][1][t] = 0.0;
                    sdb[dir][2][t] = 0.0;
                } else {
                    sdb[dir][0][t] = 0.5 * (bxb * mpx + byb * mpy);
                    sdb[dir][1][t] = 0.5 * bxb;
                    sdb[dir][2][t] = 0.5 * byb;
                }
            }
        }
        lds_sync();
        if (active) {
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int lb = dir == 0 ? rlb0 : slb0, ls = dir == 0 ? 1 : N1, pp = dir == 0 ? a : j;
                const double mqx = (dir == 0 ? srx : ssx)[t], mqy = (dir == 0 ? sry : ssy)[t];
#pragma unroll
                for (int q = 0; q < N1; ++q) {
                    const int iq = lb + q * ls;
                    nb.b += T.d1[q * N1 + pp] * (sdb[dir][0][iq] + sdb[dir][1][iq] * mqx + sdb[dir][2][iq] * mqy);
                }
            }
        }
    } else {
        // subcell finite volumes per line: end fluxes, interface HLL, bed and metric-balance terms
        if (active) {
            const Node me = {sh[t], shu[t], shv[t], su[t], sv[t], sb[t]};
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int lb = dir == 0 ? rlb0 : slb0, ls = dir == 0 ? 1 : N1, pp = dir == 0 ? a : j;
                const double* smx = dir == 0 ? srx : ssx;
                const double* smy = dir == 0 ? sry : ssy;
                const Cons Wp = {swh[t], swhu[t], swhv[t]};
                const double iw = 1.0 / T.w1[pp];
                if (pp == 0) ec_flux_adj(me, me, smx[t], smy[t], p.g, cscale(1.0 / T.w1[0], Wp), nb, nb);
                if (pp == N1 - 1) ec_flux_adj(me, me, smx[t], smy[t], p.g, cscale(-(1.0 / T.w1[N1 - 1]), Wp), nb, nb);
                if (pp > 0) ifbb[dir][t - ls] = cscale(iw, Wp);
                ifab[dir][t] = pp < N1 - 1 ? cscale(-iw, Wp) : Cons{0.0, 0.0, 0.0};
                double mlx, mly, mrx, mry;
                interface_pair<N>(pp, lb, ls, smx, smy, affine, mlx, mly, mrx, mry);
                const Side L = rl[dir].get(t), R = rr[dir].get(t);
                double bedb, metb;
                if (affine) { bedb = Wp.hu * smx[lb] + Wp.hv * smy[lb]; metb = 0.0; }
                else {
                    bedb = Wp.hu * 0.5 * (mlx + mrx) + Wp.hv * 0.5 * (mly + mry);
                    metb = Wp.hu * (mrx - mlx) + Wp.hv * (mry - mly);
                }
                // bed = K1 (L.h + R.h)(R.b - L.b), metric = K2 ((R.h + h)(b - R.b) + (L.h + h)(b - L.b))
                const double K1 = -0.5 * p.g * iw, K2 = 0.25 * p.g * iw;
                SideBar Lb = sbar_zero(), Rb = sbar_zero();
                const double k1 = bedb * K1;
                Lb.h += k1 * (R.b - L.b); Rb.h += k1 * (R.b - L.b);
                Rb.b += k1 * (L.h + R.h); Lb.b -= k1 * (L.h + R.h);
                const double k2 = metb * K2;
                Rb.h += k2 * (me.b - R.b);
                Lb.h += k2 * (me.b - L.b);
                nb.h += k2 * ((me.b - R.b) + (me.b - L.b));
                nb.b += k2 * ((R.h + me.h) + (L.h + me.h));
                Rb.b -= k2 * (R.h + me.h);
                Lb.b -= k2 * (L.h + me.h);
                rlb[dir][t] = Lb;
                rrb[dir][t] = Rb;
            }
        }
        lds_sync();
        if (active) {
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int lb = dir == 0 ? rlb0 : slb0, ls = dir == 0 ? 1 : N1, pp = dir == 0 ? a : j;
                if (pp < N1 - 1) {
                    double mlx, mly, mx, my;  // im[pp + 1]
                    interface_pair<N>(pp, lb, ls, dir == 0 ? srx : ssx, dir == 0 ? sry : ssy, affine, mlx, mly, mx, my);
                    SideBar sab = sbar_zero(), sbb = sbar_zero();
                    hydro_hll_adj(rr[dir].get(t), rl[dir].get(t + ls), mx, my, p.g, p.h_min, ifab[dir][t], ifbb[dir][t], sab,
                                  sbb);
                    SideBar& own = rrb[dir][t];
                    own.h += sab.h; own.u += sab.u; own.v += sab.v; own.b += sab.b;
                    rlx[dir][t + ls] = sbb;
                }
            }
        }
        lds_sync();
        if (active) {
            const Side me = {sh[t], su[t], sv[t], sb[t]};
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int ls = dir == 0 ? 1 : N1, pp = dir == 0 ? a : j;
                SideBar Lb = rlb[dir][t];
                if (pp > 0) { const SideBar x = rlx[dir][t]; Lb.h += x.h; Lb.u += x.u; Lb.v += x.v; Lb.b += x.b; }
                const SideBar Rb = rrb[dir][t];
                // the neighbours of the reconstruction, as the forward takes them
                Side lft = me, rgt = me;
                double xlf = 0.0, xrt = 0.0;
                bool has_l = true, has_r = true;
                const int fl = dir == 0 ? 3 : 0, fil = dir == 0 ? N1 - 1 - j : a;
                const int fr = dir == 0 ? 1 : 2, fir = dir == 0 ? j : N1 - 1 - a;
                if (pp > 0) { int q = t - ls; lft = {sh[q], su[q], sv[q], sb[q]}; xlf = T.x1[pp - 1]; }
                if (pp < N1 - 1) { int q = t + ls; rgt = {sh[q], su[q], sv[q], sb[q]}; xrt = T.x1[pp + 1]; }
                // one outer node per direction at most (as the forward)
                const bool at_l = pp == 0;
                if (at_l || pp == N1 - 1) {
                    Side o = me;
                    double d = 0.0;
                    const bool has = outer_node<N>(p, k, at_l ? fl : fr, at_l ? fil : fir, o, d);
                    if (at_l) { lft = o; xlf = -1.0 - d; has_l = has; }
                    else { rgt = o; xrt = 1.0 + d; has_r = has; }
                }
                SideBar lftb = sbar_zero(), rgtb = sbar_zero(), meb = sbar_zero();
                if (has_l && has_r)
                    reconstruct_adj(lft, me, rgt, xlf, T.x1[pp], xrt, T.xl[pp], T.xr[pp], p.h_dry, Lb, Rb, lftb, meb, rgtb);
                else {
                    meb = {Lb.h + Rb.h, Lb.u + Rb.u, Lb.v + Rb.v, Lb.b + Rb.b};
                }
                nbar_add_side(nb, meb);
                tol[dir][t] = sbar_zero();
                tor[dir][t] = sbar_zero();
                if (has_l && has_r) {
                    if (pp > 0) tol[dir][t] = lftb;
                    else outer_node_adj<N>(P, k, fl, fil, lftb);
                    if (pp < N1 - 1) tor[dir][t] = rgtb;
                    else outer_node_adj<N>(P, k, fr, fir, rgtb);
                } else {
                    // no reconstruction at a boundary line end: nothing reaches the neighbour's inner nodes
                    if (pp == 0 && has_l) outer_node_adj<N>(P, k, fl, fil, sbar_zero());
                    if (pp == N1 - 1 && has_r) outer_node_adj<N>(P, k, fr, fir, sbar_zero());
                }
            }
        }
        lds_sync();
        if (active) {
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int ls = dir == 0 ? 1 : N1, pp = dir == 0 ? a : j;
                if (pp > 0) nbar_add_side(nb, tor[dir][t - ls]);
                if (pp < N1 - 1) nbar_add_side(nb, tol[dir][t + ls]);
            }
        }
    }

    lds_sync();  // the region's next use (tcot) overlays the line cotangents
    // Faces: rhs += sJ LIFT (F(q).n - F*) at the element's face nodes, spread
    // over its lanes. The F(q).n part is this node's own; the cotangent of F*
    // goes to the face buffer for edge_vjp_kernel, which runs the Riemann
    // solve backwards for both sides of the face.
    if (active) {
#pragma unroll
        for (int r = 0; r < NT; ++r) {
            const int task = i + r * NN;
            if (task >= NF) break;
            const int f = task / N1, fi = task - f * N1;
            const int li = eb + face_node<N>(f, fi);
            const Node q = {sh[li], shu[li], shv[li], su[li], sv[li], sb[li]};
            const size_t fk = (size_t)k * 4 + f, nf4 = (size_t)p.n_elements * 4;
            const double nx = p.fgeo[fk], ny = p.fgeo[nf4 + fk], sj = p.fgeo[2 * nf4 + fk];
            const Cons sfdb = cscale(sj * T.lift[f][fi], Cons{swh[li], swhu[li], swhv[li]});
            NodeBar qb = nodebar_zero();
            ec_flux_adj(q, q, nx, ny, p.g, sfdb, qb, qb);
            double* o = P.facecot + ((size_t)k * NF + task) * 3;
            o[0] = -sfdb.h; o[1] = -sfdb.hu; o[2] = -sfdb.hv;
            tcot[e * NF + task] = qb;
        }
    }
    lds_sync();
    if (active) {
        if (j == 0) nbar_add(nb, tcot[e * NF + 0 * N1 + a]);
        if (a == N1 - 1) nbar_add(nb, tcot[e * NF + 1 * N1 + j]);
        if (j == N1 - 1) nbar_add(nb, tcot[e * NF + 2 * N1 + (N1 - 1 - a)]);
        if (a == 0) nbar_add(nb, tcot[e * NF + 3 * N1 + (N1 - 1 - j)]);
        double gb = 0.0;
        make_node_adj(hin, huin, hvin, p.h_min, nb, gi_h, gi_hu, gi_hv, gb);
        if (P.input_write) {
            P.gih[n] = gi_h;
            P.gihu[n] = gi_hu;
            P.gihv[n] = gi_hv;
        } else {
            P.gih[n] += gi_h;
            P.gihu[n] += gi_hu;
            P.gihv[n] += gi_hv;
        }
        P.gbed[n] = pre.gbed + gb;
    }
    if (P.gman) {
        lds_sync();  // sman overlays tcot
        sman[t] = gman;
        lds_sync();
        if (active && i == 0) {
            double s = 0.0;
            for (int q = 0; q < NN; ++q) s += sman[eb + q];
            P.gman[k] += s;
        }
    }
    (void)bin;
}

// The faces whose Riemann solves the edge kernels run backwards, as k << 2 | f,
// in element order. `edges`: every boundary face, and every interior edge once,
// from its left side (the side whose link has the left bit), where the normal
// its right side solved with equals the left's own (the right's own reversed on
// a shared face: equal to the left's up to the sign of a zero component on a
// conforming edge; the left's own on a periodic pair). `split`: the edges whose
// two normals differ, from their left side (for edge_vjp_split_kernel). nx, ny:
// the faces' unit normals [K][4].
inline void vjp_edge_lists(const uint32_t* links, const double* nx, const double* ny, int64_t K,
                           std::vector<uint32_t>& edges, std::vector<uint32_t>& split) {
    edges.clear();
    split.clear();
    edges.reserve((size_t)K * 2 + 16);
    for (int64_t k = 0; k < K; ++k)
        for (int f = 0; f < 4; ++f) {
            const size_t fk = (size_t)k * 4 + f;
            const uint32_t l = links[fk];
            const uint32_t own = (uint32_t)k << 2 | (uint32_t)f;
            if (l == LINK_BOUNDARY) { edges.push_back(own); continue; }
            if (!(l & 4u)) continue;  // listed from its left side
            const size_t fo = (size_t)(l >> 4) * 4 + (l & 3u);
            const bool agree = !(links[fo] & 8u) || (-nx[fo] == nx[fk] && -ny[fo] == ny[fk]);
            (agree ? edges : split).push_back(own);
        }
}

// The Riemann solves backwards, one lane per node of a listed face, from the
// cotangents of F* that the element kernels wrote per face node. The forward
// solved an interior edge once in each of its elements, both times from the
// left side with the same states and (for a listed edge) the same normal, and
// took F*_a on the left and -F*_b on the right: one reverse solve with both
// cotangents (that of F*_a from the left, minus the right's of F*_b) carries
// the edge, and writes the cotangents of both face nodes; a boundary face (wall
// ghost, open boundary) is its element's alone. The (h, hu, hv, B) cotangent
// of every element face node goes to that node's own slot of facebuf: each
// slot is written by one lane (an edge whose normals differ is
// edge_vjp_split_kernel's).
template <int N>
__global__ __launch_bounds__(64) DG_WAVES_PER_EU(3) void edge_vjp_kernel(VjpParams P, const uint32_t* edges,
                                                                         uint32_t n_edges) {
    constexpr int N1 = N + 1, NN = N1 * N1, NF = 4 * N1;
    const StageParams& p = P.s;
    const size_t gid = (size_t)blockIdx.x * 64 + threadIdx.x;
    if (gid >= (size_t)n_edges * N1) return;
    const uint32_t ed = edges[gid / N1];
    const int fi = (int)(gid % N1);
    const int64_t k = (int64_t)(ed >> 2);
    const int f = (int)(ed & 3u);
    const size_t n = (size_t)k * NN + face_node<N>(f, fi);
    const double h = p.ih[n], hu = p.ihu[n], hv = p.ihv[n], bd = p.bed[n];
    const Node q = make_node(h, hu, hv, bd, p.h_min);
    const size_t fk = (size_t)k * 4 + f, nf4 = (size_t)p.n_elements * 4;
    const double nx = p.fgeo[fk], ny = p.fgeo[nf4 + fk];
    const uint32_t l = p.links[fk];
    const double* fc = P.facecot + ((size_t)k * NF + f * N1 + fi) * 3;
    const Cons fsb = {fc[0], fc[1], fc[2]};
    NodeBar qb = nodebar_zero();
    if (l == LINK_BOUNDARY && p.tags[fk] == DG_BC_TIDE) {
        const bool incoming = (p.flags & DG_INCOMING_WAVE) != 0;
        const Cons qbd = characteristic(q.h, q.hu, q.hv, q.b, p.eta_e, nx, ny, p.g, p.h_min, incoming);
        const Node nbn = make_node(qbd.h, qbd.hu, qbd.hv, q.b, p.h_min);
        NodeBar nbb = nodebar_zero();
        ec_flux_adj(nbn, nbn, nx, ny, p.g, fsb, nbb, nbb);
        double bh = 0.0, bhu = 0.0, bhv = 0.0, bbb = 0.0;
        make_node_adj(qbd.h, qbd.hu, qbd.hv, p.h_min, nbb, bh, bhu, bhv, bbb);
        qb.b += bbb;
        characteristic_adj(q.h, q.hu, q.hv, q.b, p.eta_e, nx, ny, p.g, p.h_min, incoming, Cons{bh, bhu, bhv}, qb.h,
                           qb.hu, qb.hv, qb.b);
    } else {
        // this lane's face is the edge's left side (or a wall's, the ghost on the right)
        const bool wall = l == LINK_BOUNDARY;
        size_t slot_o = 0;
        double xh = 0.0, xhu = 0.0, xhv = 0.0, xb = 0.0;
        Cons fso = {0.0, 0.0, 0.0};
        if (wall) {
            const Cons gs = wall_ghost(q.h, q.hu, q.hv, nx, ny, p.h_min);
            xh = gs.h; xhu = gs.hu; xhv = gs.hv; xb = q.b;
        } else {
            const int fo = (int)(l & 3u), fio = N1 - 1 - fi;
            const size_t m = (size_t)(l >> 4) * NN + face_node<N>(fo, fio);
            xh = p.ih[m]; xhu = p.ihu[m]; xhv = p.ihv[m]; xb = p.bed[m];
            slot_o = (size_t)(l >> 4) * NF + fo * N1 + fio;
            const double* fco = P.facecot + slot_o * 3;
            fso = {fco[0], fco[1], fco[2]};
        }
        const Node qo = make_node(xh, xhu, xhv, xb, p.h_min);
        const Cons fbb = {-fso.h, -fso.hu, -fso.hv};
        SideBar sab = sbar_zero(), sbb = sbar_zero();
        hydro_hll_adj(side(q), side(qo), nx, ny, p.g, p.h_min, fsb, fbb, sab, sbb);
        nbar_add_side(qb, sab);
        const NodeBar ob = {sbb.h, 0.0, 0.0, sbb.u, sbb.v, sbb.b};
        double oh = 0.0, ohu = 0.0, ohv = 0.0, obb = 0.0;
        make_node_adj(xh, xhu, xhv, p.h_min, ob, oh, ohu, ohv, obb);
        if (wall) {
            qb.b += obb;
            wall_ghost_adj(q.h, q.hu, q.hv, nx, ny, p.h_min, Cons{oh, ohu, ohv}, qb.h, qb.hu, qb.hv);
        } else {
            double* o = P.facebuf + slot_o * 4;
            o[0] = oh; o[1] = ohu; o[2] = ohv; o[3] = obb;
        }
    }
    double gh = 0.0, ghu = 0.0, ghv = 0.0, gb = 0.0;
    make_node_adj(h, hu, hv, p.h_min, qb, gh, ghu, ghv, gb);
    double* o = P.facebuf + ((size_t)k * NF + f * N1 + fi) * 4;
    o[0] = gh; o[1] = ghu; o[2] = ghv; o[3] = gb;
}

// The edges whose two normals differ (each element solved the edge with its
// own: the left side with its normal for F*_a, the right side with its own
// reversed for F*_b): both solves backwards, one lane per edge node, each with
// its own element's cotangent, the two sides' parts summed, then both face
// nodes' slots written. Not on a conforming mesh.
template <int N>
__global__ __launch_bounds__(64) void edge_vjp_split_kernel(VjpParams P, const uint32_t* edges, uint32_t n_edges) {
    constexpr int N1 = N + 1, NN = N1 * N1, NF = 4 * N1;
    const StageParams& p = P.s;
    const size_t gid = (size_t)blockIdx.x * 64 + threadIdx.x;
    if (gid >= (size_t)n_edges * N1) return;
    const uint32_t ed = edges[gid / N1];
    const int fi = (int)(gid % N1);
    const int64_t k = (int64_t)(ed >> 2);
    const int f = (int)(ed & 3u);
    const uint32_t l = p.links[(size_t)k * 4 + f];
    const int64_t kr = (int64_t)(l >> 4);
    const int fr = (int)(l & 3u), fir = N1 - 1 - fi;
    const size_t nl = (size_t)k * NN + face_node<N>(f, fi), nr = (size_t)kr * NN + face_node<N>(fr, fir);
    const size_t sl = (size_t)k * NF + f * N1 + fi, sr = (size_t)kr * NF + fr * N1 + fir;
    const size_t nf4 = (size_t)p.n_elements * 4, fl = (size_t)k * 4 + f, fq = (size_t)kr * 4 + fr;
    const double lh = p.ih[nl], lhu = p.ihu[nl], lhv = p.ihv[nl], rh = p.ih[nr], rhu = p.ihu[nr], rhv = p.ihv[nr];
    const Side sa = side(make_node(lh, lhu, lhv, p.bed[nl], p.h_min));
    const Side sb = side(make_node(rh, rhu, rhv, p.bed[nr], p.h_min));
    const double* cl = P.facecot + sl * 3;
    const double* cr = P.facecot + sr * 3;
    const Cons zero = {0.0, 0.0, 0.0};
    SideBar a1 = sbar_zero(), b1 = sbar_zero(), a2 = sbar_zero(), b2 = sbar_zero();
    hydro_hll_adj(sa, sb, p.fgeo[fl], p.fgeo[nf4 + fl], p.g, p.h_min, Cons{cl[0], cl[1], cl[2]}, zero, a1, b1);
    hydro_hll_adj(sa, sb, -p.fgeo[fq], -p.fgeo[nf4 + fq], p.g, p.h_min, zero, Cons{-cr[0], -cr[1], -cr[2]}, a2, b2);
    const NodeBar ql = {a1.h + a2.h, 0.0, 0.0, a1.u + a2.u, a1.v + a2.v, a1.b + a2.b};
    const NodeBar qr = {b1.h + b2.h, 0.0, 0.0, b1.u + b2.u, b1.v + b2.v, b1.b + b2.b};
    double gh = 0.0, ghu = 0.0, ghv = 0.0, gb = 0.0;
    make_node_adj(lh, lhu, lhv, p.h_min, ql, gh, ghu, ghv, gb);
    double* o = P.facebuf + sl * 4;
    o[0] = gh; o[1] = ghu; o[2] = ghv; o[3] = gb;
    gh = ghu = ghv = gb = 0.0;
    make_node_adj(rh, rhu, rhv, p.h_min, qr, gh, ghu, ghv, gb);
    o = P.facebuf + sr * 4;
    o[0] = gh; o[1] = ghu; o[2] = ghv; o[3] = gb;
}

// The face parts of the input's and the bed's cotangents: every node adds,
// face by face in face order, its own slot of facebuf on that face (the edge
// kernel's cotangent of this face node), and what a neighbour that took the
// subcell volume term wrote for it as the inner node behind that neighbour's
// face. One thread per node.
template <int N>
__global__ __launch_bounds__(256) void vjp_gather_kernel(const uint32_t* links, const double* facebuf,
                                                         const double* innerbuf,
                                                         const uint8_t* kind, int64_t K, double* gh, double* ghu,
                                                         double* ghv, double* gbed) {
    constexpr int N1 = N + 1, NN = N1 * N1, NF = 4 * N1;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)K * NN) return;
    const int64_t k = (int64_t)(gid / NN);
    const int i = (int)(gid - (size_t)k * NN);
    const int a = i % N1, j = i / N1;
    double sh = 0.0, shu = 0.0, shv = 0.0, sb = 0.0;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const uint32_t l = links[4 * k + f];
        // position of this node on face f, and as the inner node behind face f
        const int fi = f == 0 ? (j == 0 ? a : -1)
                     : f == 1 ? (a == N1 - 1 ? j : -1)
                     : f == 2 ? (j == N1 - 1 ? N1 - 1 - a : -1)
                              : (a == 0 ? N1 - 1 - j : -1);
        const int gi = f == 0 ? (j == 1 ? a : -1)
                     : f == 1 ? (a == N1 - 2 ? j : -1)
                     : f == 2 ? (j == N1 - 2 ? N1 - 1 - a : -1)
                              : (a == 1 ? N1 - 1 - j : -1);
        if (fi >= 0) {
            const double* o = facebuf + ((size_t)k * NF + f * N1 + fi) * 4;
            sh += o[0]; shu += o[1]; shv += o[2]; sb += o[3];
        }
        if (l == LINK_BOUNDARY) continue;
        const int64_t nb = (int64_t)(l >> 4);
        const int nbf = (int)(l & 3u);
        if (gi >= 0 && kind[nb]) {
            const double* o = innerbuf + ((size_t)nb * NF + nbf * N1 + (N1 - 1 - gi)) * 4;
            sh += o[0]; shu += o[1]; shv += o[2]; sb += o[3];
        }
    }
    gh[gid] += sh;
    ghu[gid] += shu;
    ghv[gid] += shv;
    gbed[gid] += sb;
}

}  // namespace dgswe

// Device code of the 2D shallow-water stage: the fused SSP stage kernel, the
// time-step kernel and the state utilities. Included by dgswe.hip (and by a
// host-thread emulation of workgroups used to check the kernels on a CPU).
#pragma once

#include <cmath>
#include <cstring>

#include "../../include/dgswe.h"
#include "swe_physics.hpp"

namespace dgswe {

// ---------------------------------------------------------------------------
// GLL tables per order (set once from the host)
// ---------------------------------------------------------------------------
struct GllTables {
    double x1[4], w1[4], d1[16], w2[16];
    double lift[4][4];       // LIFT value at face node fi of face f (its node's entry)
    double xl[4], xr[4];     // subcell a spans [xl[a], xr[a]]
    // 2 D split by position for the volume term of node p: vsub[p][q] = 2 D_pq
    // where q < p (terms subtracted from the node), vacc[p][q] = 2 D_pq where
    // q >= p (terms of the node's own sum); zero elsewhere
    double vsub[16], vacc[16];
};
__constant__ GllTables c_gll[4];

constexpr uint32_t LINK_BOUNDARY = 0xFFFFFFFFu;

// Node of face node fi of face f (faces 2 and 3 list their nodes in reverse):
// a start node and a step of 1 (faces 0, 2) or N+1 (faces 1, 3), backwards on
// faces 2 and 3; integer selects, no branch on a neighbour's face
template <int N> __device__ __forceinline__ int face_node(int f, int fi) {
    constexpr int N1 = N + 1, NN = N1 * N1;
    const int step = (f & 1) ? fi * N1 : fi;
    const int start = f == 0 ? 0 : f == 1 ? N : f == 2 ? NN - 1 : N * N1;
    return (f & 2) ? start - step : start + step;
}

// The four corners of a quadrilateral, counter-clockwise from (r, s) = (-1, -1),
// held as named scalars (kept in registers)
struct Quad { double x0, y0, x1, y1, x2, y2, x3, y3; };
__device__ __forceinline__ Quad load_quad(const double* corners, int64_t k) {
    const double* c = corners + 8 * k;
    return {c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]};
}

// Element data at a 32-bit byte offset from a kernel-argument base: one SGPR base
// and one VGPR offset per access, no 64-bit address arithmetic per lane. Every
// per-node array of a context is below 4 GiB (dg_upload_form refuses more).
template <class V> __device__ __forceinline__ V ld32(const V* base, uint32_t byte_off) {
    return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <class V> __device__ __forceinline__ void st32(V* base, uint32_t byte_off, V v) {
    *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + byte_off) = v;
}
__device__ __forceinline__ Quad load_quad32(const double* corners, uint32_t k) {
    const uint32_t o = k * 64u;
    return {ld32(corners, o), ld32(corners, o + 8u), ld32(corners, o + 16u), ld32(corners, o + 24u),
            ld32(corners, o + 32u), ld32(corners, o + 40u), ld32(corners, o + 48u), ld32(corners, o + 56u)};
}

// a / b from y, the correctly rounded 1/b (Markstein's theorem: q0 = a y, the
// remainder r = a - b q0 is exact, and q0 + r y rounded once is the correctly
// rounded quotient unless it underflows): the IEEE quotient bit for bit for the
// metric's normal-range values, with one division shared by four. A zero
// remainder keeps q0, so a zero numerator keeps its sign.
__device__ __forceinline__ double div_by(double a, double b, double y) {
    const double q0 = a * y;
    const double r = fma(-b, q0, a);
    return r == 0.0 ? q0 : fma(r, y, q0);
}

// Bilinear map derivatives at (r, s) (operators/geometric.rs:438-460)
struct Metric { double x_r, x_s, y_r, y_s, det; };
__device__ __forceinline__ Metric bilinear(const Quad& c, double r, double s) {
    Metric m;
    m.x_r = 0.25 * ((1.0 - s) * (c.x1 - c.x0) + (1.0 + s) * (c.x2 - c.x3));
    m.y_r = 0.25 * ((1.0 - s) * (c.y1 - c.y0) + (1.0 + s) * (c.y2 - c.y3));
    m.x_s = 0.25 * ((1.0 - r) * (c.x3 - c.x0) + (1.0 + r) * (c.x2 - c.x1));
    m.y_s = 0.25 * ((1.0 - r) * (c.y3 - c.y0) + (1.0 + r) * (c.y2 - c.y1));
    m.det = m.x_r * m.y_s - m.x_s * m.y_r;
    return m;
}

// sJ n of face f: the outward contravariant vector of the bilinear map at its
// face nodes (geometric.rs:188-205). On the face the map's derivative along it
// is half the edge vector from vertex f to vertex f+1, exactly (the other
// corner pair enters with a factor 0 and the rest are powers of two), so the
// vector is that half edge turned clockwise, bit for bit the reference's at
// every node of the face.
template <int N> __device__ __forceinline__ void face_vector(const Quad& c, int f, double& vx, double& vy) {
    const double v0x = 0.5 * (c.y1 - c.y0), v0y = -0.5 * (c.x1 - c.x0);
    const double v1x = 0.5 * (c.y2 - c.y1), v1y = -0.5 * (c.x2 - c.x1);
    const double v2x = 0.5 * (c.y3 - c.y2), v2y = -0.5 * (c.x3 - c.x2);
    const double v3x = 0.5 * (c.y0 - c.y3), v3y = -0.5 * (c.x0 - c.x3);
    vx = f == 0 ? v0x : f == 1 ? v1x : f == 2 ? v2x : v3x;
    vy = f == 0 ? v0y : f == 1 ? v1y : f == 2 ? v2y : v3y;
}

// The per-node geometry cache of geometry_kernel: J of the node's own map, for
// the wet/dry correction's masses (read only for the elements it corrects). The
// stage and time-step kernels evaluate their metric, bed-slope factors and
// grad r, grad s from the element's corners (64 B per element) with
// GeometricFactors2D::compute's expressions, so the same bits.
enum {
    GEO_DET,
    GEO_COUNT
};

struct StageParams {
    const double *bh, *bhu, *bhv;
    const double *ih, *ihu, *ihv;
    double *oh, *ohu, *ohv;
    const double* bed;
    const double* corners;
    const double* manning;
    const uint32_t* links;
    const uint8_t* tags;
    const uint8_t* affine;
    int64_t n_elements;
    double a, b, c;
    double eta_e;
    double g, h_min, h_dry, u_max, tau, f0;
    double sqrt_g;       // sqrt(g), correctly rounded on the host
    int rhs_only;
    int flags;
    unsigned long long* clips;
    double* fstar_out;   // optional [K][4][N+1][3]: F* at every face node, outward (this element's side)
    double *rh, *rhu, *rhv;  // optional: the stage's RHS, written by the forward, read by the VJP
    // Static geometry cached per node, SoA [GEO_COUNT][K*NN] (geometry_kernel): J; per face
    // [3][K*4]: the unit normal and sJ
    const double* geo;
    const double* fgeo;
    uint32_t* queue;     // [blocks][EPB]: the block's elements that take the subcell volume term
    uint8_t* queue_count;  // [blocks]: how many
    unsigned int xcd_stride;  // workgroups per XCD share of the wet kernel's grid (0: no remap)
    // Local time stepping (the kernels' LTS instantiation; multirate.rs:169-200
    // element_stage): the pass's elements elist[n_list]; an element on level l
    // takes c = lvl_c[l] and the tide elevation lvl_eta[l] of its own stage time;
    // with acc (the scheme's last stage) its value is summed, acc <- w out + keep
    // acc with w = 2^(level - rate) and keep = (level < depth), instead of
    // post-processed and stored
    const uint32_t* elist;
    int64_t n_list;
    const uint8_t* level;
    const uint8_t* lrate;  // the last stage's rates
    int depth;
    double *ah, *ahu, *ahv;
    double lvl_c[DG_MAX_LEVELS + 1], lvl_eta[DG_MAX_LEVELS + 1];
    // Boundary tides per face node (BoundaryTides, tidal_atlas.rs:858-1140): the
    // slot of face node (k, f, fi) (UINT32_MAX: the uniform tide), per slot the
    // constants [n_c][6] (C, S of eta, u, v), the mean level (0 unless used) and
    // the source model's depth; the stage time's ramp and (f, cos, sin) per constituent
    const uint32_t* tide_slot;
    const double *tide_coef, *tide_mean, *tide_src;
    int atlas_nc, atlas_velocity;
    double atlas_max_ratio;
    double atlas_moment[1 + 3 * ATLAS_MAX_CONSTITUENTS];
    // The nesting band (NestingRelaxation2D, ocean_nesting.rs:890-1060): every
    // node's slot (UINT32_MAX: none), per slot its snapshots' (zeta, u, v) [n_times][3]
    // (f32), band weight and parent depth; the stage time's snapshot stencil and ramp
    const uint32_t* nest_slot;
    const float* nest_series;
    const double *nest_weight, *nest_depth;
    int nest_times, nest_len, nest_velocity, nest_momentum_only;
    int nest_idx[4];
    double nest_w[4];
    double nest_rate, nest_wet_depth, nest_max_ratio, nest_reference, nest_ramp;
    int nest_boundary;   // the open boundary's external state is the parent's at its face nodes' slots
    // the gridded weather (dg_set_weather): per snapshot, field-major f32 [n_times][5][K*NN] of u, v,
    // dp/dx, dp/dy, p - p_ref; the stage's snapshot stencil, the ramped switches and the constants
    const float* wx_series;
    int64_t wx_nodes;
    int wx_len, wx_drag;
    int wx_idx[4];
    double wx_w[4];
    double wx_wind, wx_pres, wx_cd, wx_rho_air, wx_rho_water, wx_h_min;
    double wx_raise;   // the open boundary's level raised by this times the inverse barometer (0: not)
    // local time stepping with forcing: per level a row of LVL_MOMENT values, the moments at the level's stage time
    const double* lvl_moment;
    // Manning n per node [K*NN] (SpatiallyVaryingManning2D, friction.rs:463-553) and a beta-plane's f per node
    // [K*NN] (CoriolisSource2D::beta_plane, coriolis.rs:101-121), or NULL: per element n, the f-plane f0
    const double *manning_node, *coriolis_node;
};

// A level's row of the forcing moments: the atlas moment (25 values), then the nesting band's and the weather's
// snapshot moments at LM_NEST and LM_WX, each [ramp or wind, 0 or pressure, len, idx x 4, w x 4]
constexpr int LM_NEST = 1 + 3 * ATLAS_MAX_CONSTITUENTS, LM_WX = LM_NEST + 11, LVL_MOMENT = LM_WX + 11;

struct SnapMoment { int len; int idx[4]; double w[4]; double a, b; };

// The nesting band's (WX false: a = ramp) or the weather's (WX true: a, b = the ramped switches) snapshot
// moment at the element's stage time: the kernel's arguments, or with local time stepping the level's row
template <bool LTS, bool WX>
__device__ __forceinline__ SnapMoment snap_moment(const StageParams& p, int lv) {
    SnapMoment m;
    if constexpr (LTS) {
        const double* r = p.lvl_moment + (size_t)lv * LVL_MOMENT + (WX ? LM_WX : LM_NEST);
        m.a = r[0]; m.b = r[1]; m.len = (int)r[2];
#pragma unroll
        for (int q = 0; q < 4; ++q) { m.idx[q] = (int)r[3 + q]; m.w[q] = r[7 + q]; }
    } else if constexpr (WX) {
        m.a = p.wx_wind; m.b = p.wx_pres; m.len = p.wx_len;
#pragma unroll
        for (int q = 0; q < 4; ++q) { m.idx[q] = p.wx_idx[q]; m.w[q] = p.wx_w[q]; }
    } else {
        m.a = p.nest_ramp; m.b = 0.0; m.len = p.nest_len;
#pragma unroll
        for (int q = 0; q < 4; ++q) { m.idx[q] = p.nest_idx[q]; m.w[q] = p.nest_w[q]; }
    }
    return m;
}

// LDS exchange among the lanes of a workgroup that is one wavefront: the
// wavefront's LDS operations complete in issue order, so only the compiler's
// ordering needs fencing (no s_barrier, and no wait on outstanding global loads).
__device__ __forceinline__ void lds_sync() {
#ifdef DGSWE_EMULATE
    __syncthreads();
#else
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#endif
}

// Workgroup b of a grid of 8*stride runs on XCD b % 8; give each XCD a contiguous
// eighth of the element blocks, so that neighbour gathers hit its own L2.
__device__ __forceinline__ unsigned int xcd_block(unsigned int b, unsigned int stride) {
    return stride ? (b & 7u) * stride + (b >> 3) : b;
}

// One line's flux-differencing volume term with the bed slope, for the node at
// position p (swe_2d_split_form.rs:843-894). Every pair flux of the node is
// formed first (independent, so they overlap), then they are summed in the
// reference's order: the pairs whose first node is an earlier one (q < p)
// straight into the node, then the node's own sum from the diagonal on. The
// two-point flux and the metric average are bitwise symmetric in the pair.
template <int N>
__device__ __forceinline__ void volume_line(Cons& rhs, int p, int lb, int ls, const double* sh, const double* shu,
                                            const double* shv, const double* su, const double* sv,
                                            const double* smx, const double* smy, double g, double bx, double by) {
    constexpr int N1 = N + 1;
    const GllTables& T = c_gll[N];
    const int ip = lb + p * ls;
    const Node me = {sh[ip], shu[ip], shv[ip], su[ip], sv[ip], 0.0};
    const double mpx = smx[ip], mpy = smy[ip];
    // The metric average of the pair: at the node itself (q = p) and on an
    // affine element (whose cached metric is the same at every node) the
    // average of two equal values is that value exactly, so it is formed
    // for every pair without a select.
    Cons f[N1];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
        const int iq = lb + q * ls;
        const Node o = {sh[iq], shu[iq], shv[iq], su[iq], sv[iq], 0.0};
        const double mx = 0.5 * (smx[iq] + mpx);
        const double my = 0.5 * (smy[iq] + mpy);
        f[q] = ec_flux(o, me, mx, my, g);
    }
    // The reference's order with the lane's position p as data instead of
    // control (the coefficient tables hold 2 D_pq or zero by position):
    // subtracting a zero term from the node (q >= p) or adding one to the
    // node's own sum before its first term (q < p) leaves the value as it is
    // (up to the sign of a zero result). The node's own sum starts from its first
    // term, not from a zero it is added to.
    Cons acc = cscale(T.vacc[p * N1], f[0]);
    rhs = csub(rhs, cscale(T.vsub[p * N1], f[0]));
#pragma unroll
    for (int q = 1; q < N1; ++q) {
        rhs = csub(rhs, cscale(T.vsub[p * N1 + q], f[q]));
        acc = cadd(acc, cscale(T.vacc[p * N1 + q], f[q]));
    }
    // the bed slope's factors (bx, by): bed_line below, before the passes
    const double force = g * me.h;
    acc = cadd(acc, Cons{0.0, force * bx, force * by});
    rhs = csub(rhs, acc);
}

// The volume bed-slope factors of the node at position p of a line
// (swe_2d_split_form.rs:872-890): g h_p (bx, by) is the bed term, with
// bx = 1/2 (Ja_p (D B)_p + (D (Ja B))_p) (Ja_p (D B)_p when the element is affine)
template <int N>
__device__ __forceinline__ void bed_line(int p, int lb, int ls, const double* sb, const double* smx, const double* smy,
                                         bool affine, double& bx, double& by) {
    constexpr int N1 = N + 1;
    const GllTables& T = c_gll[N];
    const int ip = lb + p * ls;
    const double mpx = smx[ip], mpy = smy[ip];
    double db = 0.0, dbx = 0.0, dby = 0.0;
#pragma unroll
    for (int q = 0; q < N1; ++q) {
        const int iq = lb + q * ls;
        double d_b = T.d1[p * N1 + q] * sb[iq];
        db += d_b;
        dbx += d_b * smx[iq];
        dby += d_b * smy[iq];
    }
    if (affine) { bx = db * mpx; by = db * mpy; }
    else { bx = 0.5 * (db * mpx + dbx); by = 0.5 * (db * mpy + dby); }
}

// Subcell interface directions of a line (subcells.rs:63-90): im[0] and im[N1]
// are the end nodes' metric, im[a+1] the running sum of the telescoped metric
// over subcells 0..a. Only the pair around position p is kept, as the sums
// run (ml = im[p], mr = im[p+1]), so the whole table is never held.
template <int N>
__device__ __forceinline__ void interface_pair(int p, int lb, int ls, const double* smx, const double* smy, bool affine,
                                               double& mlx, double& mly, double& mrx, double& mry) {
    constexpr int N1 = N + 1;
    const GllTables& T = c_gll[N];
    if (affine) {
        mlx = mrx = smx[lb];
        mly = mry = smy[lb];
        return;
    }
    mlx = smx[lb];
    mly = smy[lb];
    double m0 = 0.0, m1 = 0.0;
#pragma unroll
    for (int a = 0; a < N1 - 1; ++a) {
#pragma unroll
        for (int c = 0; c < N1; ++c) {
            if (c == a) continue;
            double q2 = 2.0 * T.w1[a] * T.d1[a * N1 + c];
            m0 += q2 * 0.5 * (smx[lb + a * ls] + smx[lb + c * ls]);
            m1 += q2 * 0.5 * (smy[lb + a * ls] + smy[lb + c * ls]);
        }
        if (a + 1 == p) { mlx = m0; mly = m1; }
        if (a == p) { mrx = m0; mry = m1; }
    }
    if (p == N1 - 1) { mrx = smx[lb + (N1 - 1) * ls]; mry = smy[lb + (N1 - 1) * ls]; }
}

// The subcell update's contributions to the node at position p of a line
// (swe_2d_split_form.rs:938-1035), from the reconstructed face states of the
// line's subcells (rl, rr) and its interface fluxes (fa, fb) in LDS.
template <int N, int M>
__device__ __forceinline__ void subcell_line(Cons& rhs, int p, int lb, int ls, const double* sh, const double* shu,
                                             const double* shv, const double* su, const double* sv, const double* sb,
                                             const double* smx, const double* smy, bool affine, double g,
                                             const SideRows<M>& rl, const SideRows<M>& rr, const ConsRows<M>& ifa,
                                             const ConsRows<M>& ifb) {
    constexpr int N1 = N + 1;
    const GllTables& T = c_gll[N];
    const int ip = lb + p * ls;
    Node me = {sh[ip], shu[ip], shv[ip], su[ip], sv[ip], sb[ip]};
    if (p == 0) rhs = cadd(rhs, cscale(1.0 / T.w1[0], ec_flux(me, me, smx[ip], smy[ip], g)));
    if (p == N1 - 1) rhs = csub(rhs, cscale(1.0 / T.w1[N1 - 1], ec_flux(me, me, smx[ip], smy[ip], g)));
    if (p > 0) rhs = cadd(rhs, cscale(1.0 / T.w1[p], ifb.get(ip - ls)));
    if (p < N1 - 1) rhs = csub(rhs, cscale(1.0 / T.w1[p], ifa.get(ip)));
    double mlx, mly, mrx, mry;
    interface_pair<N>(p, lb, ls, smx, smy, affine, mlx, mly, mrx, mry);
    const Side left = rl.get(ip);
    const Side right = rr.get(ip);
    double bed = -0.5 * g * (left.h + right.h) * (right.b - left.b) / T.w1[p];
    double fx, fy;
    if (affine) { fx = bed * smx[lb]; fy = bed * smy[lb]; }
    else {
        double metric = 0.25 * g * ((right.h + me.h) * (me.b - right.b) + (left.h + me.h) * (me.b - left.b)) / T.w1[p];
        fx = bed * 0.5 * (mlx + mrx) + metric * (mrx - mlx);
        fy = bed * 0.5 * (mly + mry) + metric * (mry - mly);
    }
    rhs = cadd(rhs, Cons{0.0, fx, fy});
}

// Element height normal to face f at face node fi (J / sJ), affine: node-0 values
template <int N>
__device__ __forceinline__ double face_height(const Quad& c, bool affine, int f, int fi) {
    const GllTables& T = c_gll[N];
    constexpr int N1 = N + 1;
    double vx, vy;
    face_vector<N>(c, f, vx, vy);
    double sj = sqrt_rn(vx * vx + vy * vy);
    int node = affine ? 0 : face_node<N>(f, fi);
    Metric m = bilinear(c, T.x1[node % N1], T.x1[node / N1]);
    return m.det / sj;
}

// The outer node of face node fi of face f (swe_2d_split_form.rs:1038-1085): the
// neighbour's first interior node on the GLL line, and its reference distance
// (static: geometry_kernel caches it per face node with outer_distance below)
template <int N>
__device__ __forceinline__ bool outer_node(const StageParams& p, int64_t k, int f, int fi, Side& out, double& dist) {
    constexpr int N1 = N + 1, NN = N1 * N1;
    uint32_t l = p.links[4 * k + f];
    if (l == LINK_BOUNDARY) return false;
    int64_t nb = (int64_t)(l >> 4);
    int nbf = (int)(l & 3u);
    int nb_node = face_node<N>(nbf, N1 - 1 - fi);
    int inner = nbf == 0 ? nb_node + N1 : nbf == 1 ? nb_node - 1 : nbf == 2 ? nb_node - N1 : nb_node + 1;
    dist = p.fgeo[(size_t)12 * p.n_elements + ((size_t)k * 4 + f) * N1 + fi];
    size_t n = (size_t)nb * NN + inner;
    Node q = make_node(p.ih[n], p.ihu[n], p.ihv[n], p.bed[n], p.h_min);
    out = side(q);
    return true;
}

// The reference distance of the outer node behind face node fi of face f:
// the neighbour's first GLL spacing scaled by the ratio of the element
// heights normal to the face (swe_2d_split_form.rs:1061-1080)
template <int N>
__device__ __forceinline__ double outer_distance(const double* corners, const uint8_t* affine_flags, const Quad& c,
                                                 bool affine, uint32_t l, int f, int fi) {
    constexpr int N1 = N + 1;
    const GllTables& T = c_gll[N];
    const int64_t nb = (int64_t)(l >> 4);
    const int nbf = (int)(l & 3u);
    const Quad cn = load_quad(corners, nb);
    const bool nb_affine = affine_flags[nb] != 0;
    return (T.x1[1] - T.x1[0]) * face_height<N>(cn, nb_affine, nbf, N1 - 1 - fi) / face_height<N>(c, affine, f, fi);
}

// The wet/dry correction of one element held one node per lane: if any node is
// below h_dry, Zhang-Shu scaling towards the mass-weighted mean until h >= 0,
// momentum zeroed if the mean is below h_dry, a negative mean emptied and
// counted; then the desingularised velocity and the speed cap at every node
// (wetting_drying.rs:194-213; limiters/swe_2d.rs:104-197). wflag[] must be
// zero on entry; sh, shu, shv, pw and savg are overwritten.
template <int N>
__device__ __forceinline__ void wet_dry_correct(double h_dry, double u_max, const double* geo, int64_t K,
                                                unsigned long long* clips, bool active, size_t n, double& vh,
                                                double& vhu, double& vhv, double* sh, double* shu, double* shv,
                                                double* pw, int* wflag, double* savg) {
    constexpr int N1 = N + 1, NN = N1 * N1;
    const GllTables& T = c_gll[N];
    const int t = threadIdx.x, e = t / NN, i = t - e * NN, eb = e * NN;
    if (active) {
        sh[t] = vh; shu[t] = vhu; shv[t] = vhv;
        if (vh < h_dry) wflag[e] = 1;
    }
    lds_sync();
    const bool needs = active && wflag[e] != 0;
    if (needs) {
        // the node masses of an element that is corrected (J read here only)
        pw[t] = T.w2[i] * geo[GEO_DET * (size_t)K * NN + n];
    }
    lds_sync();
    if (needs && i == 0) {
        double area = 0.0;
        for (int q = 0; q < NN; ++q) area += pw[eb + q];
        double inv_total = 1.0 / area;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, hmin = INFINITY;
        for (int q = 0; q < NN; ++q) {
            double w = pw[eb + q];
            s0 += w * sh[eb + q];
            s1 += w * shu[eb + q];
            s2 += w * shv[eb + q];
        }
        for (int q = 0; q < NN; ++q) hmin = fmin_(hmin, sh[eb + q]);
        savg[e * 4 + 0] = s0 * inv_total;
        savg[e * 4 + 1] = s1 * inv_total;
        savg[e * 4 + 2] = s2 * inv_total;
        savg[e * 4 + 3] = hmin;
    }
    lds_sync();
    if (needs) {
        double ah = savg[e * 4 + 0], ahu = savg[e * 4 + 1], ahv = savg[e * 4 + 2], hmin = savg[e * 4 + 3];
        if (ah <= 0.0) {
            vh = 0.0; vhu = 0.0; vhv = 0.0;
            if (i == 0 && ah < 0.0) atomicAdd(clips, 1ull);
        } else {
            if (hmin < 0.0) {
                double theta = fmax_(0.0, fmin_(1.0, ah / (ah - hmin)));
                vh = theta * (vh - ah) + ah;
                vhu = theta * (vhu - ahu) + ahu;
                vhv = theta * (vhv - ahv) + ahv;
                vh = fmax_(vh, 0.0);
            }
            if (ah < h_dry) { vhu = 0.0; vhv = 0.0; }
        }
    }
    if (active) correct_node(vh, vhu, vhv, h_dry, u_max);
}

// One workgroup-wavefront of EPB elements through a fused stage: the element
// RHS (volume term, Riemann fluxes at the 4(N+1) face nodes from the left side
// of each edge, LIFT, 1/J, Coriolis), then out = a*base + b*input + c*RHS, the
// point-implicit damping and the wet/dry correction (builder.rs:652-737).
// SUB: the subcell volume term (elements with a node below h_dry); otherwise
// the flux-differencing one, and an element with a shallow node is handed to
// the subcell queue instead.
}  // namespace dgswe

#include "swe_vjp.hpp"

namespace dgswe {

// VJP: the same forward up to the element RHS, then the reverse (vjp_tail)
// instead of the post-processing and the stores.
// LTS: k is an element of the pass's list p.elist, lts_blk the list block.
// k is a valid element on every lane (an inactive lane's is not used).
template <int N, bool SUB, bool VJP = false, bool LTS = false, bool FORCE = false>
__device__ __forceinline__ void stage_block(const StageParams& p, int64_t k, bool active,
                                            const VjpParams* vp = nullptr, int64_t lts_blk = 0) {
    constexpr int N1 = N + 1, NN = N1 * N1, EPB = 64 / NN, NF = 4 * N1;
    constexpr int NT = (NF + NN - 1) / NN;  // face tasks per lane
    const GllTables& T = c_gll[N];
    __shared__ double sh[64], shu[64], shv[64], su[64], sv[64], sb[64];
    __shared__ double srx[64], sry[64], ssx[64], ssy[64];
    __shared__ ConsRows<EPB * NF> sfd;
    __shared__ double ssj[EPB * 4];
    __shared__ double pw[64];
    __shared__ int sflag[EPB], wflag[EPB];
    __shared__ double savg[EPB * 4];
    __shared__ SideRows<SUB ? 64 : 1> rl[2], rr[2];
    __shared__ ConsRows<SUB ? 64 : 1> ifa[2], ifb[2];
    __shared__ int64_t skid[LTS && !SUB ? EPB : 1];

    const int t = threadIdx.x;
    const int e = t / NN;
    const int i = t - e * NN;
    const int a = i % N1, j = i / N1;
    const size_t n = (size_t)k * NN + i;
    const uint32_t k32 = (uint32_t)k, nb = ((uint32_t)k * NN + i) * 8u;  // the node's byte offset in a state array
    const int eb = e * NN;

    // Every global load is issued here, before the first LDS exchange (whose
    // fences the compiler does not move loads across): the face tasks' links,
    // this node, the corners, the neighbours' face nodes, the cached geometry,
    // the base state and Manning n; one wait for all of them.
    uint32_t lk[NT];
    uint8_t tg[NT];
    double nh[NT], nhu[NT], nhv[NT], nbd[NT];
    double fnx[NT], fny[NT], fsj[NT];
    Quad c = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double hin = 0.0, huin = 0.0, hvin = 0.0, bin = 0.0;
    double g_rx = 0.0, g_ry = 0.0, g_sx = 0.0, g_sy = 0.0, det = 1.0, det_inv = 1.0;
    double bxr = 0.0, byr = 0.0, bxs = 0.0, bys = 0.0;
    double bh0 = 0.0, bhu0 = 0.0, bhv0 = 0.0, mann = 0.0;
    VjpPre vpre = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // SUB: the outer nodes behind this node's line ends (r-line, s-line), read with the other loads
    Side o_side[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    double o_dist[2] = {0.0, 0.0};
    bool o_has[2] = {false, false};
    bool affine = false;
    int lv = 0;  // LTS: the element's level
    if (active) {
        affine = ld32(p.affine, k32) != 0;
        if constexpr (LTS) lv = p.level[k];
#pragma unroll
        for (int r = 0; r < NT; ++r) {
            int task = i + r * NN;
            lk[r] = LINK_BOUNDARY;
            tg[r] = 0;
            if (task < NF) {
                lk[r] = ld32(p.links, (4u * k32 + (uint32_t)(task / N1)) * 4u);
                tg[r] = ld32(p.tags, 4u * k32 + (uint32_t)(task / N1));
            }
        }
        hin = ld32(p.ih, nb); huin = ld32(p.ihu, nb); hvin = ld32(p.ihv, nb); bin = ld32(p.bed, nb);
#pragma unroll
        for (int r = 0; r < NT; ++r) {
            int task = i + r * NN;
            nh[r] = nhu[r] = nhv[r] = nbd[r] = 0.0;
            if (task < NF) {
                // the neighbour's face node (at a boundary this node, unused)
                int fi = task - (task / N1) * N1;
                uint32_t l = lk[r];
                const uint32_t mb = l == LINK_BOUNDARY ? nb
                                    : ((l >> 4) * NN + (uint32_t)face_node<N>((int)(l & 3u), N1 - 1 - fi)) * 8u;
                nh[r] = ld32(p.ih, mb); nhu[r] = ld32(p.ihu, mb); nhv[r] = ld32(p.ihv, mb); nbd[r] = ld32(p.bed, mb);
            }
        }
        // the static metric of this node from the element's corners (evaluated
        // below); the faces' unit normals and sJ from the cache
        const size_t nt = (size_t)p.n_elements * NN;
        c = load_quad32(p.corners, k32);
        if (VJP) det = p.geo[GEO_DET * nt + n];  // the forward reads J only where the wet/dry correction needs it
        if (!VJP) {
            const size_t nf4 = (size_t)p.n_elements * 4;
            // (with two face tasks per lane, N < 3, they are loaded at use:
            // held from here they would not fit the registers)
#pragma unroll
            for (int r = 0; r < (NT == 1 ? 1 : 0); ++r) {
                const int task = i + r * NN;
                fnx[r] = fny[r] = fsj[r] = 0.0;
                if (task < NF) {
                    const uint32_t fb = (4u * k32 + (uint32_t)(task / N1)) * 8u;
                    fnx[r] = ld32(p.fgeo, fb); fny[r] = ld32(p.fgeo + nf4, fb); fsj[r] = ld32(p.fgeo + 2 * nf4, fb);
                }
            }
            if (NT == 1 && !p.rhs_only && p.a != 0.0) { bh0 = ld32(p.bh, nb); bhu0 = ld32(p.bhu, nb); bhv0 = ld32(p.bhv, nb); }
            if (NT == 1 && (p.flags & DG_FRICTION)) mann = ld32(p.manning, k32 * 8u);
        } else {
            // the reverse's own reads: the saved RHS, the base state, Manning n
            // and the output cotangent
            if (p.rh) { vpre.rh = p.rh[n]; vpre.rhu = p.rhu[n]; vpre.rhv = p.rhv[n]; }
            if (!p.rhs_only && p.a != 0.0) { vpre.bh = p.bh[n]; vpre.bhu = p.bhu[n]; vpre.bhv = p.bhv[n]; }
            if (p.flags & DG_FRICTION) vpre.mn = p.manning[k];
            vpre.wh = vp->wh[n]; vpre.whu = vp->whu[n]; vpre.whv = vp->whv[n];
            if (!p.rhs_only && p.a != 0.0 && vp->gbh) { vpre.gbh = vp->gbh[n]; vpre.gbhu = vp->gbhu[n]; vpre.gbhv = vp->gbhv[n]; }
            vpre.gbed = vp->gbed[n];
        }
        if constexpr (SUB && !VJP) {
            // a node is at one end of its line at most (N >= 1): one outer node per direction
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int pos = dir == 0 ? a : j;
                const bool at_l = pos == 0;
                if (at_l || pos == N1 - 1) {
                    const int f = dir == 0 ? (at_l ? 3 : 1) : (at_l ? 0 : 2);
                    const int fi = dir == 0 ? (at_l ? N1 - 1 - j : j) : (at_l ? a : N1 - 1 - a);
                    o_has[dir] = outer_node<N>(p, k, f, fi, o_side[dir], o_dist[dir]);
                }
            }
        }
    }
    if (active) {
        // GeometricFactors2D::compute's expressions (GeometricFactors2D::compute, geometric.rs:150-175): the
        // contravariant metric J grad r, J grad s and 1/J of the node's map (a parallelogram's at node 0)
        const double r = affine ? T.x1[0] : T.x1[a], s = affine ? T.x1[0] : T.x1[j];
        const Metric mm = bilinear(c, r, s);
        det_inv = 1.0 / mm.det;
        g_rx = mm.det * div_by(mm.y_s, mm.det, det_inv);
        g_ry = mm.det * div_by(-mm.x_s, mm.det, det_inv);
        g_sx = mm.det * div_by(-mm.y_r, mm.det, det_inv);
        g_sy = mm.det * div_by(mm.x_r, mm.det, det_inv);
    }
    if (t < EPB) { sflag[t] = 0; wflag[t] = 0; }
    if constexpr (LTS && !SUB) {
        if (i == 0 && e < EPB) skid[e] = active ? k : -1;
    }
    lds_sync();
    const bool subcells_on = (p.flags & DG_SUBCELLS) != 0;
    if (!SUB && active && subcells_on && hin < p.h_dry) sflag[e] = 1;
    lds_sync();
    if (!SUB) {
        // Elements with a shallow node take the subcell volume term: listed per
        // block (in element order) for the subcell kernel
        if (t == 0) {
            int cnt = 0;
            const int64_t qb = LTS ? lts_blk : k / EPB;
            for (int q = 0; q < EPB; ++q) {
                const int64_t kq = LTS ? skid[LTS ? q : 0] : k - e + q;
                if ((LTS ? kq >= 0 : kq < p.n_elements) && sflag[q] != 0) p.queue[(size_t)qb * EPB + cnt++] = (uint32_t)kq;
            }
            p.queue_count[qb] = (uint8_t)cnt;
        }
        if constexpr (VJP) {
            if (active && i == 0) vp->kind[k] = (uint8_t)(sflag[e] != 0);
        }
        if (active && sflag[e] != 0) active = false;
    }

    if (active) {
        Node q = make_node(hin, huin, hvin, bin, p.h_min);
        sh[t] = q.h; shu[t] = q.hu; shv[t] = q.hv; su[t] = q.u; sv[t] = q.v; sb[t] = q.b;
        srx[t] = g_rx; sry[t] = g_ry; ssx[t] = g_sx; ssy[t] = g_sy;
    }
    lds_sync();

    // Lines through this node: the r-line (lanes eb + j*N1 + q) and the s-line (eb + q*N1 + a)
    const int rlb = eb + j * N1, slb = eb + a;
    if (!SUB && !VJP && active) {
        // the volume bed-slope factors of the node's r- and s-line
        bed_line<N>(a, rlb, 1, sb, srx, sry, affine, bxr, byr);
        bed_line<N>(j, slb, N1, sb, ssx, ssy, affine, bxs, bys);
    }

    // Faces: F* at the element's face nodes, spread over its lanes (the VJP
    // reads the RHS that its forward saved instead)
    if (active && !VJP) {
#pragma unroll
        for (int r = 0; r < NT; ++r) {
            const int task = i + r * NN;
            if (task >= NF) break;
            const int f = task / N1, fi = task - f * N1;
            const int li = eb + face_node<N>(f, fi);
            Node q = {sh[li], shu[li], shv[li], su[li], sv[li], sb[li]};
            const size_t nf4 = (size_t)p.n_elements * 4, fk = (size_t)k * 4 + f;
            const double nx = NT == 1 ? fnx[r] : p.fgeo[fk], ny = NT == 1 ? fny[r] : p.fgeo[nf4 + fk],
                         sj = NT == 1 ? fsj[r] : p.fgeo[2 * nf4 + fk];
            const uint32_t l = lk[r];
            Cons fstar;
            if (l == LINK_BOUNDARY && tg[r] == DG_BC_TIDE) {
                const uint32_t slot = (FORCE && p.tide_slot) ? p.tide_slot[((size_t)k * 4 + f) * N1 + fi] : 0xFFFFFFFFu;
                double raise = 0.0;
                if (FORCE && p.wx_raise != 0.0) {
                    // InverseBarometer over the weather's level at the face node (characteristic.rs:324-328,
                    // inverse_barometer_level, gridded_atmosphere.rs:198-200): -ramp p / (rho g), p summed
                    // over the stage's snapshots at the node
                    const size_t nf = (size_t)k * NN + face_node<N>(f, fi);
                    const SnapMoment m = snap_moment<LTS, true>(p, lv);
                    double sp = 0.0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (q >= m.len) break;
                        sp += m.w[q] * (double)p.wx_series[((size_t)m.idx[q] * 5 + 4) * p.wx_nodes + nf];
                    }
                    raise = p.wx_raise * (-m.b * sp / (p.wx_rho_water * 9.81));
                }
                const uint32_t pslot = (FORCE && p.nest_boundary) ? p.nest_slot[(size_t)k * NN + face_node<N>(f, fi)]
                                                                  : 0xFFFFFFFFu;
                Cons qb;
                if (FORCE && pslot != 0xFFFFFFFFu) {
                    // the nesting parent's state at the face node's slot (OceanModelState::external_state,
                    // ocean_nesting.rs:823-826), raised by the inverse barometer
                    const SnapMoment m = snap_moment<LTS, false>(p, lv);
                    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                    for (int qq = 0; qq < 4; ++qq) {
                        if (qq >= m.len) break;
                        const float* v = p.nest_series + ((size_t)pslot * p.nest_times + m.idx[qq]) * 3;
                        s0 += m.w[qq] * (double)v[0];
                        s1 += m.w[qq] * (double)v[1];
                        s2 += m.w[qq] * (double)v[2];
                    }
                    double eta, pu, pv;
                    bool hv_;
                    parent_state(s0, s1, s2, m.a, p.nest_reference, p.nest_velocity != 0, p.nest_max_ratio,
                                 p.nest_depth[pslot], q.b, eta, hv_, pu, pv);
                    if (p.wx_raise != 0.0) eta += raise;
                    qb = hv_ ? characteristic_velocity(q.h, q.hu, q.hv, q.b, eta, pu, pv, nx, ny, p.g, p.h_min)
                             : characteristic(q.h, q.hu, q.hv, q.b, eta, nx, ny, p.g, p.h_min,
                                              (p.flags & DG_INCOMING_WAVE) != 0);
                } else if (FORCE && slot != 0xFFFFFFFFu) {
                    // the node's own atlas tide at the element's stage time
                    double eta, ue, ve;
                    if constexpr (LTS)
                        atlas_state(p.tide_coef + (size_t)slot * 6 * p.atlas_nc, p.atlas_nc,
                                    p.lvl_moment + (size_t)lv * LVL_MOMENT, p.tide_mean[slot], p.atlas_velocity != 0,
                                    p.atlas_max_ratio, p.tide_src[slot], q.b, eta, ue, ve);
                    else
                        atlas_state(p.tide_coef + (size_t)slot * 6 * p.atlas_nc, p.atlas_nc, p.atlas_moment,
                                    p.tide_mean[slot], p.atlas_velocity != 0, p.atlas_max_ratio, p.tide_src[slot], q.b,
                                    eta, ue, ve);
                    if (p.wx_raise != 0.0) eta += raise;
                    qb = p.atlas_velocity ? characteristic_velocity(q.h, q.hu, q.hv, q.b, eta, ue, ve, nx, ny, p.g, p.h_min)
                                          : characteristic(q.h, q.hu, q.hv, q.b, eta, nx, ny, p.g, p.h_min,
                                                           (p.flags & DG_INCOMING_WAVE) != 0);
                } else {
                    // the tide at the element's own stage time (its level's)
                    double seta = p.eta_e;
                    if constexpr (LTS) {
#pragma unroll
                        for (int l = 0; l <= DG_MAX_LEVELS; ++l) seta = lv == l ? p.lvl_eta[l] : seta;
                    }
                    if (FORCE && p.wx_raise != 0.0) seta += raise;
                    qb = characteristic(q.h, q.hu, q.hv, q.b, seta, nx, ny, p.g, p.h_min, (p.flags & DG_INCOMING_WAVE) != 0);
                }
                Node nb = make_node(qb.h, qb.hu, qb.hv, q.b, p.h_min);
                fstar = ec_flux(nb, nb, nx, ny, p.g);
            } else {
                // One Riemann solve per face node, as the left side of the edge
                // sees it: an interior face's left side (this element, or the
                // neighbour with this face's normal reversed), or a wall's
                // reflected ghost on the right
                const bool wall = l == LINK_BOUNDARY;
                const bool left = wall || (l & 4u);
                // the outside state: the neighbour's face node, or at a wall the
                // reflected ghost (formed only where a lane has a wall face)
                double oh = nh[r], ohu = nhu[r], ohv = nhv[r], ob = nbd[r];
                if (wall) {
                    const Cons gs = wall_ghost(q.h, q.hu, q.hv, nx, ny, p.h_min);
                    oh = gs.h; ohu = gs.hu; ohv = gs.hv; ob = q.b;
                }
                Node qo = make_node(oh, ohu, ohv, ob, p.h_min);
                double lx = nx, ly = ny;
                if (!left) {
                    lx = -nx; ly = -ny;
                    if (!(l & 8u)) {
                        // a periodic pair: the left side's own normal
                        const size_t fo = (size_t)(l >> 4) * 4 + (l & 3u);
                        lx = p.fgeo[fo]; ly = p.fgeo[nf4 + fo];
                    }
                }
                if (p.flags & DG_ENTROPY_STABLE) {
                    // EntropyStable faces (swe_2d_split_form.rs:611-651): the EC flux and
                    // its dissipation from the left side, plus each side's bed term
                    const Node na = left ? q : qo, nb = left ? qo : q;
                    const Cons shared = cadd(ec_flux(na, nb, lx, ly, p.g), es_dissipation(na, nb, lx, ly, p.g));
                    fstar = left ? cadd(shared, bed_interface(q.h, q.b, qo.b, lx, ly, p.g))
                                 : cadd(cneg(shared), bed_interface(q.h, q.b, qo.b, -lx, -ly, p.g));
                } else {
                    const Side sa = left ? side(q) : side(qo), sbb = left ? side(qo) : side(q);
                    Cons fa, fb;
                    hydro_hll(sa, sbb, lx, ly, p.g, p.sqrt_g, p.h_min, fa, fb);
                    fstar = left ? fa : cneg(fb);
                }
            }
            if (p.fstar_out) {
                double* o = p.fstar_out + ((size_t)k * NF + task) * 3;
                o[0] = fstar.h; o[1] = fstar.hu; o[2] = fstar.hv;
            }
            sfd.put(e * NF + task, csub(ec_flux(q, q, nx, ny, p.g), fstar));
            if (fi == 0) ssj[e * 4 + f] = sj;
        }
    }
    // (the subcell path after the faces: the faces' registers are free by then)
    if (SUB) {
        // Reconstructed face states of this node's subcell on its r- and s-line
        if (active) {
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int lb = dir == 0 ? rlb : slb, ls = dir == 0 ? 1 : N1;
                int pos = dir == 0 ? a : j;
                Side me = {sh[t], su[t], sv[t], sb[t]};
                Side lft = me, rgt = me;
                double xlf = 0.0, xrt = 0.0;
                bool has_l = true, has_r = true;
                if (pos > 0) { int q = t - ls; lft = {sh[q], su[q], sv[q], sb[q]}; xlf = T.x1[pos - 1]; }
                if (pos < N1 - 1) { int q = t + ls; rgt = {sh[q], su[q], sv[q], sb[q]}; xrt = T.x1[pos + 1]; }
                // A line end's neighbour is the outer node behind the face there;
                // a node is at one end of its line at most (N >= 1), so one read
                // per direction with the face as data (not a branch per end)
                const bool at_l = pos == 0;
                if (at_l || pos == N1 - 1) {
                    const int f = dir == 0 ? (at_l ? 3 : 1) : (at_l ? 0 : 2);
                    const int fi = dir == 0 ? (at_l ? N1 - 1 - j : j) : (at_l ? a : N1 - 1 - a);
                    Side o = me;
                    double d = 0.0;
                    bool has;
                    if constexpr (!VJP) { o = o_side[dir]; d = o_dist[dir]; has = o_has[dir]; }
                    else has = outer_node<N>(p, k, f, fi, o, d);
                    if (at_l) { lft = o; xlf = -1.0 - d; has_l = has; }
                    else { rgt = o; xrt = 1.0 + d; has_r = has; }
                }
                Side sl = me, sr = me;
                if (has_l && has_r)
                    reconstruct(lft, me, rgt, xlf, T.x1[pos], xrt, T.xl[pos], T.xr[pos], p.h_dry, sl, sr);
                rl[dir].put(t, sl);
                rr[dir].put(t, sr);
            }
        }
        lds_sync();
        if (active && !VJP) {
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                const int lb = dir == 0 ? rlb : slb, ls = dir == 0 ? 1 : N1;
                int pos = dir == 0 ? a : j;
                if (pos < N1 - 1) {
                    double mlx, mly, mx, my;  // the interface to the right of this subcell: im[pos + 1]
                    interface_pair<N>(pos, lb, ls, dir == 0 ? srx : ssx, dir == 0 ? sry : ssy, affine, mlx, mly, mx, my);
                    Cons fa, fb;
                    hydro_hll(rr[dir].get(t), rl[dir].get(t + ls), mx, my, p.g, p.sqrt_g, p.h_min, fa, fb);
                    ifa[dir].put(t, fa);
                    ifb[dir].put(t, fb);
                }
            }
        }
    }

    lds_sync();

    Cons rhs = {0.0, 0.0, 0.0};
    if (active && !VJP) {
        const bool r_first = j <= a;
        auto line_pass = [&](int pass) {
            // The line of this pass as data (a branch on it would diverge
            // within the wavefront and run both lines' code on every lane)
            const bool rline = (pass == 0) == r_first;
            const int lp = rline ? a : j, lb = rline ? rlb : slb, ls = rline ? 1 : N1, dir = rline ? 0 : 1;
            const double* smx = rline ? srx : ssx;
            const double* smy = rline ? sry : ssy;
            if (SUB)
                subcell_line<N>(rhs, lp, lb, ls, sh, shu, shv, su, sv, sb, smx, smy, affine, p.g, rl[dir], rr[dir],
                                ifa[dir], ifb[dir]);
            else
                volume_line<N>(rhs, lp, lb, ls, sh, shu, shv, su, sv, smx, smy, p.g, rline ? bxr : bxs,
                               rline ? byr : bys);
        };
        if constexpr (SUB) {
            // one copy of the subcell line for both passes (two copies
            // scheduled together would not fit two wavefronts per SIMD)
#pragma unroll 1
            for (int pass = 0; pass < 2; ++pass) line_pass(pass);
        } else {
#pragma unroll
            for (int pass = 0; pass < 2; ++pass) line_pass(pass);
        }
        // Surface terms in face order: LIFT sJ (F(q).n - F*)
        if (j == 0) rhs = cadd(rhs, cscale(ssj[e * 4 + 0] * T.lift[0][a], sfd.get(e * NF + 0 * N1 + a)));
        if (a == N1 - 1) rhs = cadd(rhs, cscale(ssj[e * 4 + 1] * T.lift[1][j], sfd.get(e * NF + 1 * N1 + j)));
        if (j == N1 - 1) rhs = cadd(rhs, cscale(ssj[e * 4 + 2] * T.lift[2][N1 - 1 - a], sfd.get(e * NF + 2 * N1 + (N1 - 1 - a))));
        if (a == 0) rhs = cadd(rhs, cscale(ssj[e * 4 + 3] * T.lift[3][N1 - 1 - j], sfd.get(e * NF + 3 * N1 + (N1 - 1 - j))));
        rhs = {det_inv * rhs.h, det_inv * rhs.hu, det_inv * rhs.hv};
        if (FORCE && p.coriolis_node) {
            // a beta-plane: f0 + beta (y - y0) at the node (CoriolisSource2D::add_element, coriolis.rs:157-169)
            const double f = p.coriolis_node[n];
            rhs.hu += f * hvin;
            rhs.hv += (-f) * huin;
        } else if (p.f0 != 0.0) {
            rhs.hu += p.f0 * hvin;
            rhs.hv += (-p.f0) * huin;
        }
        if (FORCE && p.nest_slot) {
            // the nesting band's relaxation, after Coriolis (swe_2d.rs:929-973)
            const uint32_t slot = p.nest_slot[n];
            if (slot != 0xFFFFFFFFu) {
                const SnapMoment m = snap_moment<LTS, false>(p, lv);
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (q >= m.len) break;
                    const float* v = p.nest_series + ((size_t)slot * p.nest_times + m.idx[q]) * 3;
                    s0 += m.w[q] * (double)v[0];
                    s1 += m.w[q] * (double)v[1];
                    s2 += m.w[q] * (double)v[2];
                }
                double eta, pu, pv;
                bool hv_;
                parent_state(s0, s1, s2, m.a, p.nest_reference, p.nest_velocity != 0, p.nest_max_ratio,
                             p.nest_depth[slot], bin, eta, hv_, pu, pv);
                rhs = cadd(rhs, relaxation_source(p.nest_weight[slot], p.nest_rate, p.nest_wet_depth,
                                                  p.nest_momentum_only != 0, hin, huin, hvin, bin, eta, hv_, pu, pv));
            }
        }
        if (FORCE && p.wx_series && hin >= p.wx_h_min) {
            // the weather's wind stress and pressure gradient, after the band (GriddedAtmosphere2D::add_element,
            // gridded_atmosphere.rs:703-731): the snapshots summed in the stencil's order (add_snapshot)
            const SnapMoment m = snap_moment<LTS, true>(p, lv);
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q >= m.len) break;
                const float* v = p.wx_series + (size_t)m.idx[q] * 5 * p.wx_nodes + n;
                s0 += m.w[q] * (double)v[0];
                s1 += m.w[q] * (double)v[p.wx_nodes];
                s2 += m.w[q] * (double)v[2 * p.wx_nodes];
                s3 += m.w[q] * (double)v[3 * p.wx_nodes];
            }
            double su, sv;
            atmosphere_source(hin, m.a * s0, m.a * s1, m.b * s2, m.b * s3, p.wx_drag, p.wx_cd,
                              p.wx_rho_air, p.wx_rho_water, p.wx_h_min, su, sv);
            rhs.hu += su;
            rhs.hv += sv;
        }
        if (p.rh) { p.rh[n] = rhs.h; p.rhu[n] = rhs.hu; p.rhv[n] = rhs.hv; }
    }
    if (VJP && active && p.rh) rhs = {vpre.rh, vpre.rhu, vpre.rhv};
    if constexpr (VJP) {
        vjp_tail<N, SUB>(*vp, k, active, c, affine, det, det_inv, hin, huin, hvin, bin, rhs, vpre, sh, shu, shv, su, sv,
                         sb, srx, sry, ssx, ssy, rl, rr);
        return;
    }
    if (p.rhs_only) {
        if (active) { p.oh[n] = rhs.h; p.ohu[n] = rhs.hu; p.ohv[n] = rhs.hv; }
        return;
    }

    // Shu-Osher combination and point-implicit damping (builder.rs:668-690; wetting_drying.rs:324-363),
    // c the element's own (its level's)
    double sc = p.c;
    if constexpr (LTS) {
#pragma unroll
        for (int l = 0; l <= DG_MAX_LEVELS; ++l) sc = lv == l ? p.lvl_c[l] : sc;
    }
    double vh = 0.0, vhu = 0.0, vhv = 0.0;
    const bool friction = (p.flags & DG_FRICTION) != 0, wet_dry = (p.flags & DG_WET_DRY) != 0;
    if (active) {
        if (p.a != 0.0) {
            if (NT != 1) { bh0 = ld32(p.bh, nb); bhu0 = ld32(p.bhu, nb); bhv0 = ld32(p.bhv, nb); }
            vh = p.a * bh0; vhu = p.a * bhu0; vhv = p.a * bhv0;
        }
        if (p.b != 0.0) { vh += p.b * hin; vhu += p.b * huin; vhv += p.b * hvin; }
        if (sc != 0.0) { vh += sc * rhs.h; vhu += sc * rhs.hu; vhv += sc * rhs.hv; }
        if (friction || wet_dry) {
            if (vh <= 0.0) { vhu = 0.0; vhv = 0.0; }
            else {
                double rate = 0.0;
                if (friction) {
                    double u, v;
                    if (wet_dry && !SUB && hin >= p.h_dry) {
                        // desingularized_velocity of a node at or above h_dry (> h_min) is make_node's
                        // hu (1/h), already in LDS
                        u = su[t]; v = sv[t];
                    } else if (wet_dry) desing_velocity(hin, huin, hvin, p.h_dry, u, v);
                    else { u = 0.0; v = 0.0; if (hin > p.h_min) { double inv = 1.0 / hin; u = huin * inv; v = hvin * inv; } }
                    double speed = sqrt_rn(u * u + v * v);
                    if (speed > 0.0) {
                        double mn = NT == 1 ? mann : ld32(p.manning, k32 * 8u);
                        if (FORCE && p.manning_node) mn = ld32(p.manning_node, nb);  // the node's own n
                        double gn2 = p.g * mn * mn;
                        double hh = fmax_(vh, p.h_min);
                        rate += gn2 * speed / (manning_cbrt(hh) * hh);
                    }
                }
                if (wet_dry && vh < p.h_dry) {
                    double ex = p.h_dry / vh - 1.0;
                    rate += ex * ex / p.tau;
                }
                double factor = 1.0 / (1.0 + sc * rate);
                vhu *= factor;
                vhv *= factor;
            }
        }
    }
    if constexpr (LTS) {
        if (p.ah) {
            // The scheme's last stage: summed into the substep's result, which
            // finish_kernel post-processes (builder.rs:716-728)
            if (active) {
                const double w = ldexp(1.0, lv - (int)p.lrate[k]);  // 2^(level - last-stage rate)
                if (lv >= p.depth) { p.ah[n] = w * vh; p.ahu[n] = w * vhu; p.ahv[n] = w * vhv; }
                else { p.ah[n] = w * vh + p.ah[n]; p.ahu[n] = w * vhu + p.ahu[n]; p.ahv[n] = w * vhv + p.ahv[n]; }
            }
            return;
        }
    }
    // Wet/dry correction (wetting_drying.rs:194-213; limiters/swe_2d.rs:104-197)
    if (wet_dry) {
        lds_sync();  // the node states in LDS are reused here
        wet_dry_correct<N>(p.h_dry, p.u_max, p.geo, p.n_elements, p.clips, active, n, vh, vhu, vhv, sh, shu, shv, pw,
                           wflag, savg);
    }
    if (active) { st32(p.oh, nb, vh); st32(p.ohu, nb, vhu); st32(p.ohv, nb, vhv); }
}

// The static geometry cache, with GeometricFactors2D::compute's expressions (so the
// cached values are its bits): per node J and grad r, grad s of the node's own
// map; per face the unit normal and sJ; per face node the subcell reconstruction's
// outer-node distance. Run once per form.
template <int N>
__global__ __launch_bounds__(64) void geometry_kernel(const double* corners, const uint8_t* affine_flags,
                                                      const uint32_t* links, int64_t K, double* geo, double* fgeo,
                                                      double* emass) {
    constexpr int N1 = N + 1, NN = N1 * N1, EPB = 64 / NN;
    const GllTables& T = c_gll[N];
    __shared__ double smass[64];
    const int t = threadIdx.x, e = t / NN, i = t - e * NN, a = i % N1, j = i / N1;
    const int64_t k = (int64_t)blockIdx.x * EPB + e;
    const bool active = e < EPB && k < K;
    const size_t n = (size_t)k * NN + i, nt = (size_t)K * NN;
    if (active) smass[t] = T.w2[i] * bilinear(load_quad(corners, k), T.x1[a], T.x1[j]).det;
    lds_sync();
    if (active && i == 0) {
        // the positivity bound's static ratio (swe_2d.rs:2165-2203): the node masses w_i J_i
        // summed over the element and over its boundary nodes, in node order, and their quotient
        double weight = 0.0, weight_b = 0.0;
        for (int q = 0; q < NN; ++q) {
            const int qa = q % N1, qj = q / N1;
            weight += smass[t + q];
            if (qa == 0 || qa == N1 - 1 || qj == 0 || qj == N1 - 1) weight_b += smass[t + q];
        }
        emass[k] = weight / weight_b;
    }
    if (active) {
        const bool affine = affine_flags[k] != 0;
        const Quad c = load_quad(corners, k);
        const Metric m = bilinear(c, T.x1[a], T.x1[j]);
        geo[GEO_DET * nt + n] = m.det;
        if (i < 4) {
            double vx, vy;
            face_vector<N>(c, i, vx, vy);
            const double sj = sqrt_rn(vx * vx + vy * vy);
            const size_t fk = (size_t)k * 4 + i, nf4 = (size_t)K * 4;
            fgeo[fk] = vx / sj;
            fgeo[nf4 + fk] = vy / sj;
            fgeo[2 * nf4 + fk] = sj;
        }
        // per face node, after the faces: the outer node's reference distance
        constexpr int NF = 4 * N1;
        for (int task = i; task < NF; task += NN) {
            const int f = task / N1, fi = task - f * N1;
            const uint32_t l = links[4 * k + f];
            fgeo[(size_t)12 * K + (size_t)k * NF + task] =
                l == LINK_BOUNDARY ? 0.0 : outer_distance<N>(corners, affine_flags, c, affine, l, f, fi);
        }
    }
}

// f = f0 + beta (y - y0) at every node, y by the element's bilinear map (Mesh2D::bilinear, mesh2d.rs:578-595)
template <int N>
__global__ __launch_bounds__(256) void coriolis_node_kernel(const double* corners, int64_t K, double f0, double beta,
                                                            double y0, double* f) {
    constexpr int N1 = N + 1, NN = N1 * N1;
    const GllTables& T = c_gll[N];
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= K * NN) return;
    const int64_t k = n / NN;
    const int i = (int)(n - k * NN);
    const double r = T.x1[i % N1], s = T.x1[i / N1];
    const double* v = corners + 8 * k;
    const double n0 = (1.0 - r) * (1.0 - s) / 4.0, n1 = (1.0 + r) * (1.0 - s) / 4.0;
    const double n2 = (1.0 + r) * (1.0 + s) / 4.0, n3 = (1.0 - r) * (1.0 + s) / 4.0;
    const double y = n0 * v[1] + n1 * v[3] + n2 * v[5] + n3 * v[7];
    f[n] = f0 + beta * (y - y0);
}

// Elements with every node at or above h_dry, in mesh order (XCD-aware block
// mapping); the others are listed for the subcell kernel.
// Four wavefronts per SIMD: N = 2 needs two registers fewer than it would take
// by itself (measured 3 % faster); N = 1 and 3 fit already.
#ifdef DGSWE_EMULATE
#define DG_WET_OCCUPANCY
#else
#define DG_WET_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4)))
#endif

// FORCE: the forcing given per node (tides per boundary node, the nesting band, the weather) compiled in;
// runs without it do not carry its registers
template <int N, bool FORCE = false>
__global__ __launch_bounds__(64) DG_WET_OCCUPANCY void wet_stage_kernel(StageParams p) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    const int64_t blk = xcd_block(blockIdx.x, p.xcd_stride);
    if (blk * EPB >= p.n_elements) return;
    const int e = threadIdx.x / NN;
    const int64_t k = blk * EPB + (e < EPB ? e : 0);
    // (a lane past the mesh's end works on the last element and writes nothing)
    stage_block<N, false, false, false, FORCE>(p, k < p.n_elements ? k : p.n_elements - 1,
                                               e < EPB && blk * EPB + e < p.n_elements);
}

// A pass of local time stepping: the listed elements (p.elist[0..n_list)),
// EPB per workgroup in list order (LocalTimeStepping::stage_where,
// builder.rs:652-737); those with a shallow node are queued per block for
// subcell_stage_kernel<N, true>.
template <int N, bool FORCE = false>
__global__ __launch_bounds__(64) DG_WET_OCCUPANCY void wet_lts_kernel(StageParams p) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    const int64_t blk = xcd_block(blockIdx.x, p.xcd_stride);
    if (blk * EPB >= p.n_list) return;
    const int e = threadIdx.x / NN;
    const bool active = e < EPB && blk * EPB + e < p.n_list;
    const int64_t i = active ? blk * EPB + e : blk * EPB;
    const int64_t k = p.elist ? (int64_t)p.elist[i] : i;   // no list: every element, in mesh order
    stage_block<N, false, false, true, FORCE>(p, k, active, nullptr, blk);
}

// Compaction of the per-block subcell lists into one list in element order:
// chunk_count_kernel sums the counts of each chunk of 1024 blocks, then
// compact_kernel scans them and scatters the element indices.
__global__ __launch_bounds__(256) void chunk_count_kernel(const uint8_t* counts, unsigned int n_blocks,
                                                          unsigned int* chunk_total) {
    __shared__ unsigned int red[256];
    const unsigned int b0 = blockIdx.x * 1024u + threadIdx.x * 4u;
    unsigned int s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) s += b0 + q < n_blocks ? counts[b0 + q] : 0u;
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_total[blockIdx.x] = red[0];
}

// Exclusive prefix sum over the 256 threads of a workgroup (Hillis-Steele in
// LDS) and the sum of chunk_total[0..blockIdx.x), each thread a share: the
// list offsets of a compaction chunk (integer sums: any order is exact). The
// first chunk reads no totals: a one-chunk compaction needs no count launch.
__device__ __forceinline__ void chunk_offsets(const unsigned int* chunk_total, unsigned int s, unsigned int& base,
                                              unsigned int& excl, unsigned int& chunk_sum) {
    __shared__ unsigned int scan[256], part[256];
    unsigned int b = 0;
    for (unsigned int c = threadIdx.x; c < blockIdx.x; c += 256) b += chunk_total[c];
    scan[threadIdx.x] = s;
    part[threadIdx.x] = b;
    __syncthreads();
    for (unsigned int w = 1; w < 256; w <<= 1) {
        const unsigned int v = threadIdx.x >= w ? scan[threadIdx.x - w] : 0u;
        const unsigned int q = threadIdx.x + w < 256 ? part[threadIdx.x + w] : 0u;
        __syncthreads();
        scan[threadIdx.x] += v;
        if ((threadIdx.x & (2 * w - 1)) == 0) part[threadIdx.x] += q;
        __syncthreads();
    }
    base = part[0];
    excl = scan[threadIdx.x] - s;
    chunk_sum = scan[255];
}

__global__ __launch_bounds__(256) void compact_kernel(const uint8_t* counts, const uint32_t* queue, unsigned int n_blocks,
                                                      const unsigned int* chunk_total, int epb, uint32_t* list,
                                                      unsigned int* list_count) {
    const unsigned int b0 = blockIdx.x * 1024u + threadIdx.x * 4u;
    unsigned int cnt[4], s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) { cnt[q] = b0 + q < n_blocks ? counts[b0 + q] : 0u; s += cnt[q]; }
    unsigned int base, excl, total;
    chunk_offsets(chunk_total, s, base, excl, total);
    unsigned int off = base + excl;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for (unsigned int r = 0; r < cnt[q]; ++r) list[off++] = queue[(size_t)(b0 + q) * epb + r];
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *list_count = base + total;
}

// The listed elements with a shallow node (subcell volume term), EPB per
// workgroup pass, grid-stride over the compacted list. Two wavefronts per SIMD
// (the register budget of 256 holds it without spills; its LDS allows 7
// workgroups per CU).
#ifdef DGSWE_EMULATE
#define DG_SUB_OCCUPANCY
#else
#define DG_SUB_OCCUPANCY __attribute__((amdgpu_waves_per_eu(2)))
#endif
template <int N, bool LTS = false, bool FORCE = false>
__global__ __launch_bounds__(64) DG_SUB_OCCUPANCY void subcell_stage_kernel(StageParams p, const uint32_t* list, const unsigned int* count_ptr) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    const unsigned int count = *count_ptr;
    const int e = threadIdx.x / NN;
    for (size_t first = (size_t)blockIdx.x * EPB; first < count; first += (size_t)gridDim.x * EPB) {
        const bool active = e < EPB && first + e < count;
        stage_block<N, true, false, LTS, FORCE>(p, active ? (int64_t)list[first + e] : 0, active);
        lds_sync();
    }
}

// LocalTimeStepping::finish_where (builder.rs:739-782): the listed elements'
// state from the substep's sum, then the wet/dry correction; one lane per node.
struct FinishParams {
    const uint32_t* elist;
    int64_t n_list;
    const double *ah, *ahu, *ahv;
    double *h, *hu, *hv;
    const double* geo;
    int64_t n_elements;
    double h_dry, u_max;
    int wet_dry;
    unsigned long long* clips;
};

template <int N>
__global__ __launch_bounds__(64) void finish_kernel(FinishParams p) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    __shared__ double sh[64], shu[64], shv[64], pw[64], savg[EPB * 4];
    __shared__ int wflag[EPB];
    const int64_t blk = blockIdx.x;
    const int t = threadIdx.x, e = t / NN, i = t - e * NN;
    const bool active = e < EPB && blk * EPB + e < p.n_list;
    const int64_t k = !active ? 0 : p.elist ? (int64_t)p.elist[blk * EPB + e] : blk * EPB + e;
    const size_t n = (size_t)k * NN + i;
    double vh = 0.0, vhu = 0.0, vhv = 0.0;
    if (active) { vh = p.ah[n]; vhu = p.ahu[n]; vhv = p.ahv[n]; }
    if (t < EPB) wflag[t] = 0;
    if (p.wet_dry) {
        lds_sync();
        wet_dry_correct<N>(p.h_dry, p.u_max, p.geo, p.n_elements, p.clips, active, n, vh, vhu, vhv, sh, shu, shv, pw,
                           wflag, savg);
    }
    if (active) { p.h[n] = vh; p.hu[n] = vhu; p.hv[n] = vhv; }
}

// The VJP of a stage: wet elements first (the others are listed as in the
// forward), then the listed elements, grid-stride.
template <int N>
__global__ __launch_bounds__(64) DG_WAVES_PER_EU(N == 3 ? 3 : 4) void wet_vjp_kernel(VjpParams P) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    const int64_t blk = xcd_block(blockIdx.x, P.s.xcd_stride);
    if (blk * EPB >= P.s.n_elements) return;
    const int e = threadIdx.x / NN;
    const int64_t k = blk * EPB + (e < EPB ? e : 0);
    stage_block<N, false, true>(P.s, k < P.s.n_elements ? k : P.s.n_elements - 1,
                                e < EPB && blk * EPB + e < P.s.n_elements, &P);
}

template <int N>
__global__ __launch_bounds__(64) void subcell_vjp_kernel(VjpParams P, const uint32_t* list, const unsigned int* count_ptr) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    const unsigned int count = *count_ptr;
    const int e = threadIdx.x / NN;
    for (size_t first = (size_t)blockIdx.x * EPB; first < count; first += (size_t)gridDim.x * EPB) {
        const bool active = e < EPB && first + e < count;
        stage_block<N, true, true>(P.s, active ? (int64_t)list[first + e] : 0, active, &P);
        lds_sync();
    }
}

// ---------------------------------------------------------------------------
// Time step: per element bound and the global minimum (swe_2d.rs:2072-2258),
// one lane per node
// ---------------------------------------------------------------------------
constexpr int DT_SLOTS = 256;           // partial minima of the time step (atomic, one per 256th of the grid)
constexpr unsigned long long DT_SLOT_EMPTY = 0x7FF0000000000000ull;  // +infinity

struct DtParams {
    const double *h, *hu, *hv;
    const double* corners;              // [K][4][2]: grad r, grad s and J of each node from its element's map
    const double* emass;                // [K]: the element's node masses summed, over those of its boundary nodes
    const uint32_t* links;
    const uint8_t* affine;
    int64_t n_elements;
    double cfl, g, h_min, h_dry;
    double pcfl, max_cfl, max_cfl_sub;
    int positivity;
    double* element_dt;                 // optional
    unsigned long long* slots;          // [DT_SLOTS]: running minima, as the bits of non-negative doubles
    unsigned int xcd_stride;
};

// node_reference_rate (swe_2d.rs:1346-1367) with the metric's norms |grad r|, |grad s| given
// (nr, ns: sqrt(rx^2 + ry^2), sqrt(sx^2 + sy^2) of the same metric, formed once per node)
__device__ __forceinline__ double node_rate(double h, double hu, double hv, double rx, double ry, double sx, double sy,
                                            double nr, double ns, double g, double h_min) {
    double u = 0.0, v = 0.0;
    if (h > h_min) { double inv = 1.0 / h; u = hu * inv; v = hv * inv; }
    double cc = sqrt_rn(g * fmax_(h, 0.0));
    double lr = fabs(u * rx + v * ry) + cc * nr;
    double ls = fabs(u * sx + v * sy) + cc * ns;
    return 0.25 * (lr + ls);
}

template <int N>
__device__ __forceinline__ double dt_block(const DtParams& p, int64_t blk) {
    constexpr int N1 = N + 1, NN = N1 * N1, EPB = 64 / NN, NF = 4 * N1;
    constexpr int NT = (NF + NN - 1) / NN;
    const GllTables& T = c_gll[N];
    __shared__ double mrx[64], mry[64], msx[64], msy[64], mnr[64], mns[64], sr[64], swater[64];
    __shared__ double sdt[EPB];
    __shared__ int sshallow[EPB];
    const int t = threadIdx.x, e = t / NN, i = t - e * NN, eb = e * NN;
    const int64_t k = blk * EPB + e;
    const bool active = e < EPB && k < p.n_elements;
    const int a = i % N1, j = i / N1;
    const uint32_t k32 = (uint32_t)k, nb = ((uint32_t)k * NN + i) * 8u;  // the node's byte offset in a state array
    // every load first: links, this node, the element's corners, the
    // neighbours' face nodes
    uint32_t lk[NT];
    double nh[NT], nhu[NT], nhv[NT];
    double h = 0.0, hu = 0.0, hv = 0.0, rx = 0.0, ry = 0.0, sx = 0.0, sy = 0.0, det = 0.0;
    Quad c = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    bool affine = false;
    if (active) {
#pragma unroll
        for (int r = 0; r < NT; ++r) {
            int task = i + r * NN;
            lk[r] = task < NF ? ld32(p.links, (4u * k32 + (uint32_t)(task / N1)) * 4u) : LINK_BOUNDARY;
        }
        h = ld32(p.h, nb); hu = ld32(p.hu, nb); hv = ld32(p.hv, nb);
        c = load_quad32(p.corners, k32);
        affine = ld32(p.affine, k32) != 0;
#pragma unroll
        for (int r = 0; r < NT; ++r) {
            int task = i + r * NN;
            nh[r] = nhu[r] = nhv[r] = 0.0;
            if (task < NF && lk[r] != LINK_BOUNDARY) {
                int fi = task - (task / N1) * N1;
                const uint32_t mb = ((lk[r] >> 4) * NN + (uint32_t)face_node<N>((int)(lk[r] & 3u), N1 - 1 - fi)) * 8u;
                nh[r] = ld32(p.h, mb); nhu[r] = ld32(p.hu, mb); nhv[r] = ld32(p.hv, mb);
            }
        }
    }
    if (active) {
        // grad r, grad s and J of the node's own map, GeometricFactors2D::compute's quotients
        // (geometric.rs:150-175): y_s / J, -x_s / J, -y_r / J, x_r / J, each the IEEE quotient
        const Metric m = bilinear(c, T.x1[a], T.x1[j]);
        const double y = 1.0 / m.det;
        det = m.det;
        rx = div_by(m.y_s, m.det, y); ry = div_by(-m.x_s, m.det, y);
        sx = div_by(-m.y_r, m.det, y); sy = div_by(m.x_r, m.det, y);
        mrx[t] = rx; mry[t] = ry; msx[t] = sx; msy[t] = sy;
        // the metric's norms, once per node (the node's own rate and a neighbour's face node read them)
        mnr[t] = sqrt_rn(rx * rx + ry * ry);
        mns[t] = sqrt_rn(sx * sx + sy * sy);
        swater[t] = (T.w2[i] * det) * fmax_(h, 0.0);
    }
    if (i == 0 && e < EPB) sshallow[e] = 0;
    lds_sync();
    double rate = 0.0;
    if (active) {
        // an affine element's own nodes take its node-0 metric (the map's constant one)
        if (h > p.h_min) {
            const int l0 = affine ? eb : t;
            rate = node_rate(h, hu, hv, mrx[l0], mry[l0], msx[l0], msy[l0], mnr[l0], mns[l0], p.g, p.h_min);
        }
#pragma unroll
        for (int r = 0; r < NT; ++r) {
            int task = i + r * NN;
            if (task >= NF || lk[r] == LINK_BOUNDARY || nh[r] <= p.h_min) continue;
            int f = task / N1, fi = task - f * N1;
            int li = eb + face_node<N>(f, fi);
            rate = fmax_(rate, node_rate(nh[r], nhu[r], nhv[r], mrx[li], mry[li], msx[li], msy[li], mnr[li], mns[li],
                                         p.g, p.h_min));
        }
        sr[t] = rate;
        // a node at or below the shallow depth (an order-free "any")
        if (h <= 100.0 * p.h_dry) sshallow[e] = 1;
    }
    lds_sync();
    // the element's largest rate: a tree over its lanes (a maximum does not depend on the order)
    constexpr int P = NN <= 4 ? 4 : NN <= 16 ? 16 : 32;  // a power of two >= NN
#pragma unroll
    for (int w = P / 2; w > 0; w >>= 1) {
        if (active && i < w && i + w < NN) sr[t] = fmax_(sr[t], sr[t + w]);
        lds_sync();
    }
    if (active && i == 0) {
        const double rmax = sr[t];
        double cfl = p.cfl;
        if (p.positivity) {
            // the water over the element and over its boundary nodes, summed in node order; the
            // masses' sums are static (geometry_kernel)
            double water = 0.0, water_b = 0.0;
#pragma unroll
            for (int q = 0; q < NN; ++q) {
                const int qa = q % N1, qj = q / N1;
                const double wi = swater[eb + q];
                water += wi;
                if (qa == 0 || qa == N1 - 1 || qj == 0 || qj == N1 - 1) water_b += wi;
            }
            const bool shallow = sshallow[e] != 0;
            double ratio = fmin_(water / water_b, ld32(p.emass, k32 * 8u));
            double rho = fmax_(0.9 * ratio, 1.0);
            const double cap_sub = p.max_cfl_sub, cap_lin = p.max_cfl;
            double cap = shallow ? cap_sub : cap_lin;  // values: selecting a field's address would put p in memory
            cfl = fmin_(cfl, fmin_(p.pcfl * rho, cap));
        }
        double dt = rmax < 1e-14 ? INFINITY : cfl / ((2.0 * N + 1.0) * rmax);
        if (p.element_dt) p.element_dt[k] = dt;
        sdt[e] = dt;
    }
    if (!active && i == 0 && e < EPB) sdt[e] = INFINITY;
    lds_sync();
    double m = INFINITY;
    if (t == 0)
        for (int q = 0; q < EPB; ++q) m = fmin_(m, sdt[q]);
    lds_sync();
    return m;
}

// One element block per workgroup (XCD-aware mapping); the block's minimum is
// folded into one of DT_SLOTS running minima (an atomic min on the bits, which
// order like the values for non-negative doubles; the minimum does not depend
// on the order of the folds).
template <int N>
__global__ __launch_bounds__(64) void dt_kernel(DtParams p) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    const int64_t nblk = (p.n_elements + EPB - 1) / EPB;
    const unsigned int blk = xcd_block(blockIdx.x, p.xcd_stride);
    if (blk >= nblk) return;
    const double m = dt_block<N>(p, blk);
    if (threadIdx.x == 0) atomicMin(p.slots + (blockIdx.x & (DT_SLOTS - 1)), (unsigned long long)__double_as_longlong(m));
}

// The minimum of the slots into *out; the slots are left empty for the next
// time step (so no clearing launch or memset is needed, in a graph either)
__global__ __launch_bounds__(DT_SLOTS) void slots_min_kernel(unsigned long long* slots, double* out) {
    __shared__ double red[DT_SLOTS];
    red[threadIdx.x] = __longlong_as_double((long long)slots[threadIdx.x]);
    slots[threadIdx.x] = DT_SLOT_EMPTY;
    __syncthreads();
    for (int w = DT_SLOTS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmin_(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// The minimum of n values (one workgroup of 1024 lanes), into *out
__global__ __launch_bounds__(1024) void min_kernel(const double* v, size_t n, double* out) {
    __shared__ double red[1024];
    double m = INFINITY;
    for (size_t q = threadIdx.x; q < n; q += 1024) m = fmin_(m, v[q]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmin_(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ---------------------------------------------------------------------------
// Utilities
// ---------------------------------------------------------------------------
__global__ void fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }
__global__ void fill_u32(unsigned int* p, unsigned int v) { *p = v; }

__global__ void finite_kernel(const double* x, size_t n, int* bad) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    int local = 0;
    for (; i < n; i += stride) local |= !isfinite(x[i]);
    if (local) atomicOr(bad, 1);
}

template <int N>
__global__ void masks_kernel(const double* h, int64_t K, double h_dry, int subcells, uint8_t* node_dry, uint8_t* elem_sub) {
    constexpr int NN = (N + 1) * (N + 1);
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    uint8_t any = 0;
    for (int i = 0; i < NN; ++i) {
        uint8_t d = h[(size_t)k * NN + i] < h_dry ? 1 : 0;
        node_dry[(size_t)k * NN + i] = d;
        any |= d;
    }
    elem_sub[k] = subcells ? any : 0;
}

// Bit j of element k: every node of k deeper than d[j] (the line-local elements
// and the coarse-step bound of local time stepping, builder.rs:569-637)
template <int N>
__global__ void depth_flags_kernel(const double* h, int64_t K, double d0, double d1, double d2, uint8_t* out) {
    constexpr int NN = (N + 1) * (N + 1);
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    bool a0 = true, a1 = true, a2 = true;
    for (int i = 0; i < NN; ++i) {
        const double x = h[(size_t)k * NN + i];
        a0 = a0 && x > d0; a1 = a1 && x > d1; a2 = a2 && x > d2;
    }
    out[k] = (uint8_t)((a0 ? 1 : 0) | (a1 ? 2 : 0) | (a2 ? 4 : 0));
}

__global__ void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

__global__ void zero_kernel(double* x, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) x[i] = 0.0;
}

// Water level at stations (Probe2D, src/solver/probe.rs:1-30): eta_s = sum_i w_si (h + B)
// over the nodes of the station's element, one thread per station
__global__ void station_sample_kernel(const double* h, const double* bed, const int64_t* elem, const double* w,
                                      int n_st, int nn, double* eta) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_st) return;
    const size_t n0 = (size_t)elem[s] * nn;
    double v = 0.0;
    for (int i = 0; i < nn; ++i) v += w[(size_t)s * nn + i] * (h[n0 + i] + bed[n0 + i]);
    eta[s] = v;
}

// The misfit's cotangent at a sample: 2 scale (eta_s - obs_s) w_si into the depth and
// the bed of each station's nodes, stations in order by one thread (stations may
// share an element)
__global__ void station_seed_kernel(const int64_t* elem, const double* w, int n_st, int nn, const double* eta,
                                    const double* obs, double scale, double* gh, double* gbed) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int s = 0; s < n_st; ++s) {
        const double r = 2.0 * scale * (eta[s] - obs[s]);
        const size_t n0 = (size_t)elem[s] * nn;
        for (int i = 0; i < nn; ++i) {
            const double v = r * w[(size_t)s * nn + i];
            gh[n0 + i] += v;
            gbed[n0 + i] += v;
        }
    }
}


// Host: the tables of order n
inline void legendre(int n, double x, double& p_n, double& dp_n) {
    // three-term recurrence and P'_N from P_N, P_{N-1} (polynomial/legendre.rs:62-93)
    if (n == 1) { p_n = x; dp_n = 1.0; return; }
    double pp = 1.0, pc = x;
    for (int k = 1; k < n; ++k) {
        double nx = ((double)(2 * k + 1) * x * pc - (double)k * pp) / (double)(k + 1);
        pp = pc;
        pc = nx;
    }
    p_n = pc;
    if (std::fabs(x - 1.0) < 1e-14) dp_n = (double)(n * (n + 1)) / 2.0;
    else if (std::fabs(x + 1.0) < 1e-14) dp_n = ((n % 2 == 0) ? -1.0 : 1.0) * (double)(n * (n + 1)) / 2.0;
    else dp_n = (double)n * (x * pc - pp) / (x * x - 1.0);
}

inline GllTables make_tables(int n) {
    // GLL points by Newton's method on (1 - x^2) P_N' from the Chebyshev-Lobatto
    // points (polynomial/nodes.rs:17-107), weights 2 / (N(N+1) P_N^2), and the
    // Lagrange differentiation matrix on them.
    GllTables t;
    std::memset(&t, 0, sizeof(t));
    const int n1 = n + 1;
    for (int j = 0; j <= n; ++j) t.x1[j] = -std::cos(M_PI * (double)j / (double)n);
    t.x1[0] = -1.0;
    t.x1[n] = 1.0;
    for (int j = 1; j < n; ++j) {
        double x = t.x1[j];
        for (int it = 0; it < 100; ++it) {
            double p, dp;
            legendre(n, x, p, dp);
            double upd = (1.0 - x * x) * dp / ((double)n * (double)(n + 1) * p);
            if (std::fabs(upd) < 1e-15) break;
            x += upd;
        }
        t.x1[j] = x;
    }
    double pn[4];
    for (int j = 0; j < n1; ++j) {
        double dp;
        legendre(n, t.x1[j], pn[j], dp);
        t.w1[j] = 2.0 / ((double)(n * (n + 1)) * pn[j] * pn[j]);
    }
    for (int r = 0; r < n1; ++r)
        for (int c = 0; c < n1; ++c) t.d1[r * n1 + c] = r == c ? 0.0 : pn[r] / (pn[c] * (t.x1[r] - t.x1[c]));
    t.d1[0] = -(double)(n * (n + 1)) / 4.0;
    t.d1[n1 * n1 - 1] = (double)(n * (n + 1)) / 4.0;
    for (int j = 0; j < n1; ++j)
        for (int a = 0; a < n1; ++a) t.w2[j * n1 + a] = t.w1[a] * t.w1[j];
    for (int r = 0; r < n1; ++r)
        for (int c = 0; c < n1; ++c) {
            t.vsub[r * n1 + c] = c < r ? 2.0 * t.d1[r * n1 + c] : 0.0;
            t.vacc[r * n1 + c] = c < r ? 0.0 : 2.0 * t.d1[r * n1 + c];
        }
    auto fnode = [&](int f, int fi) {
        return f == 0 ? fi : f == 1 ? fi * n1 + n1 - 1 : f == 2 ? (n1 - 1) * n1 + (n1 - 1 - fi) : (n1 - 1 - fi) * n1;
    };
    for (int f = 0; f < 4; ++f)
        for (int fi = 0; fi < n1; ++fi) t.lift[f][fi] = (1.0 / t.w2[fnode(f, fi)]) * (1.0 * t.w1[fi]);
    double xl = -1.0;
    for (int a = 0; a < n1; ++a) {
        double xr = xl + t.w1[a];
        t.xl[a] = xl;
        t.xr[a] = xr;
        xl = xr;
    }
    return t;
}

}  // namespace dgswe

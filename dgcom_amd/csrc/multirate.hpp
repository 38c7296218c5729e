// Multirate local time stepping, the host side (MultirateStepper,
// src/time/multirate.rs:713-1330, with SWEPhysics2D's line-local flags and
// coarse-step bound, src/physics/builder.rs:563-637): levels from the element
// time steps, the level-jump limit, the stage rates (node by node over
// line-local elements, else one face per stage), then a coarse step's passes
// over the device's element-subset stages. Included by dgswe.hip (host code).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace dgswe {
namespace multirate {

constexpr int MAX_LEVELS = 16;                    // multirate.rs:455
constexpr double LANDING_SLACK = 1e-9;            // multirate.rs:588
constexpr double SUBCELL_CAP_DEPTH_FACTOR = 100.0;  // swe_2d.rs:2185
constexpr double LINE_LOCAL_HEADROOM = 1.0;       // builder.rs:148

struct Stage { double start, input, beta, c; };   // ShuOsherStage (integrator.rs:239-263)
inline int stages(int scheme, Stage* out) {
    static const Stage rk3[3] = {{1.0, 0.0, 1.0, 0.0}, {0.75, 0.25, 0.25, 1.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 0.5}};
    static const Stage rk43[4] = {{1.0, 0.0, 0.5, 0.0}, {0.0, 1.0, 0.5, 0.5}, {2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 1.0},
                                  {0.0, 1.0, 0.5, 0.5}};
    const Stage* s = scheme == 0 ? rk3 : rk43;
    const int n = scheme == 0 ? 3 : 4;
    for (int i = 0; i < n; ++i) out[i] = s[i];
    return n;
}

// assign_levels (multirate.rs:526-568): returns the coarse step, `finest` and
// the levels; (inf, 0) when no element limits the step
inline double assign_levels(const std::vector<double>& edt, int max_levels, double cap, int& finest,
                            std::vector<uint8_t>& level) {
    const size_t K = edt.size();
    level.assign(K, 0);
    double lo = std::numeric_limits<double>::infinity(), hi = 0.0;
    for (double d : edt)
        if (std::isfinite(d)) { lo = std::fmin(lo, d); hi = std::fmax(hi, d); }
    finest = 0;
    if (!std::isfinite(lo)) return std::numeric_limits<double>::infinity();
    while (finest < max_levels) {
        const double coarser = lo * (double)(2ull << finest);
        if (coarser > hi || coarser > cap) break;
        ++finest;
    }
    const double dt = lo * (double)(1ull << finest);
    for (size_t k = 0; k < K; ++k) {
        int l = 0;
        double dl = dt;
        while (l < finest && dl > edt[k]) { dl *= 0.5; ++l; }
        level[k] = (uint8_t)l;
    }
    return dt;
}

// levels_for_step (multirate.rs:574-586)
inline int levels_for_step(const std::vector<double>& edt, double dt, int max_levels, std::vector<uint8_t>& level) {
    int finest = 0;
    for (size_t k = 0; k < edt.size(); ++k) {
        int l = 0;
        double dl = dt;
        while (l < max_levels && dl > edt[k]) { dl *= 0.5; ++l; }
        level[k] = (uint8_t)l;
        finest = finest > l ? finest : l;
    }
    return finest;
}

// spread_to_neighbours (multirate.rs:592-613): to[k] = max(from[k], from[j] - drop)
inline bool spread_to_neighbours(const std::vector<int64_t>& nb, const std::vector<uint8_t>& from,
                                 std::vector<uint8_t>& to, int drop) {
    const size_t K = from.size();
    to.resize(K);
    bool changed = false;
    for (size_t k = 0; k < K; ++k) {
        int m = from[k];
        for (int f = 0; f < 4; ++f) {
            const int64_t j = nb[4 * k + f];
            if (j >= 0) { const int v = (int)from[(size_t)j] - drop; m = v > m ? v : m; }
        }
        to[k] = (uint8_t)m;
        changed = changed || to[k] != from[k];
    }
    return changed;
}

// The node of face f at position fi (node j*n1 + i on r-line j, s-line i),
// counter-clockwise: node fi of a face is node n1 - 1 - fi of the neighbour's
inline int face_node(int n1, int f, int fi) {
    return f == 0 ? fi : f == 1 ? fi * n1 + n1 - 1 : f == 2 ? (n1 - 1) * n1 + (n1 - 1 - fi) : (n1 - 1 - fi) * n1;
}
inline int inner_offset(int n1, int f) { return f == 0 ? n1 : f == 1 ? -1 : f == 2 ? -n1 : 1; }

// spread_over_lines (multirate.rs:705-767): node rates of a stage's value from
// those of its input, and each element's RHS rate
inline void spread_over_lines(const std::vector<int64_t>& nb, const std::vector<uint32_t>& links, int n1,
                              const std::vector<uint8_t>& line_local, const std::vector<uint8_t>& from,
                              std::vector<uint8_t>& to, std::vector<uint8_t>& rate) {
    const size_t nn = (size_t)n1 * n1, K = line_local.size();
    to.resize(K * nn);
    rate.resize(K);
    for (size_t k = 0; k < K; ++k) {
        const uint8_t* own = from.data() + k * nn;
        uint8_t* t = to.data() + k * nn;
        if (!line_local[k]) {
            uint8_t m = 0;
            for (size_t i = 0; i < nn; ++i) m = own[i] > m ? own[i] : m;
            for (int f = 0; f < 4; ++f) {
                const int64_t j = nb[4 * k + f];
                if (j < 0) continue;
                const int g = (int)(links[4 * k + f] & 3u);
                const uint8_t* other = from.data() + (size_t)j * nn;
                for (int fi = 0; fi < n1; ++fi) {
                    const int node = face_node(n1, g, fi);
                    m = other[node] > m ? other[node] : m;
                    const uint8_t in = other[node + inner_offset(n1, g)];
                    m = in > m ? in : m;
                }
            }
            std::memset(t, m, nn);
            rate[k] = m;
            continue;
        }
        for (int j = 0; j < n1; ++j) {
            uint8_t m = 0;
            for (int i = 0; i < n1; ++i) m = own[j * n1 + i] > m ? own[j * n1 + i] : m;
            for (int i = 0; i < n1; ++i) t[j * n1 + i] = m;
        }
        for (int i = 0; i < n1; ++i) {
            uint8_t m = 0;
            for (int j = 0; j < n1; ++j) m = own[j * n1 + i] > m ? own[j * n1 + i] : m;
            for (int j = 0; j < n1; ++j) t[j * n1 + i] = t[j * n1 + i] > m ? t[j * n1 + i] : m;
        }
        for (int f = 0; f < 4; ++f) {
            const int64_t j = nb[4 * k + f];
            if (j < 0) continue;
            const int g = (int)(links[4 * k + f] & 3u);
            const uint8_t* other = from.data() + (size_t)j * nn;
            for (int fi = 0; fi < n1; ++fi) {
                const int node = face_node(n1, f, fi);
                const uint8_t v = other[face_node(n1, g, n1 - 1 - fi)];
                t[node] = t[node] > v ? t[node] : v;
            }
        }
        uint8_t m = 0;
        for (size_t i = 0; i < nn; ++i) m = t[i] > m ? t[i] : m;
        rate[k] = m;
    }
}

// The levels, the limit and the stage rates of a coarse step from the element
// time steps and the depth flags (bit 0: every node deeper than
// SUBCELL_CAP_DEPTH_FACTOR h_dry, bit 1: than h_dry, bit 2: than the line-local
// depth h_dry + LINE_LOCAL_HEADROOM): MultirateStepper::assign_levels and
// ::assign (multirate.rs:1014-1168) with SWEPhysics2D::coarse_step_bound and
// ::line_local (builder.rs:563-637)
struct Assigner {
    int max_levels = 0, n_stages = 3, n1 = 3;
    bool line_rates = true;
    std::vector<int64_t> nb;    // [K][4], -1 on a boundary
    std::vector<uint32_t> links;
    // the last assignment
    int finest = 0;
    double stable = 0.0;
    std::vector<uint8_t> level, rates;  // [K], [n_stages][K]
    int64_t evals_per_step = 0;
    bool changed = true;  // the level limit or the rates differ from the last assignment's
    // kept while their inputs are (multirate.rs:1114-1121)
    bool memo = false;
    int memo_finest = -1;
    std::vector<uint8_t> memo_given, memo_ll, memo_limited;
    std::vector<uint8_t> ll, scratch, node_from, node_to, rate_tmp;

    void init(const uint32_t* face_links, size_t K) {
        links.assign(face_links, face_links + 4 * K);
        nb.resize(4 * K);
        for (size_t i = 0; i < 4 * K; ++i) nb[i] = links[i] == 0xFFFFFFFFu ? -1 : (int64_t)(links[i] >> 4);
    }

    // Returns the coarse step fitted to `landing` (<= 0: none); `stable` the one
    // the elements allow. dt_cap <= 0: none.
    double assign(const double* edt_p, size_t K, const uint8_t* flags, bool wet_dry, double dt_cap, double landing) {
        edt.assign(edt_p, edt_p + K);
        return assign_kept(flags, wet_dry, dt_cap, landing);
    }
    std::vector<double> edt;

  private:
    double assign_kept(const uint8_t* flags, bool wet_dry, double dt_cap, double landing) {
        const size_t K = edt.size();
        double bound = std::numeric_limits<double>::infinity();
        if (wet_dry) {
            for (int bit : {1, 2}) {
                double longest = 0.0;
                for (size_t k = 0; k < K; ++k)
                    if (std::isfinite(edt[k]) && (flags[k] & bit)) longest = std::fmax(longest, edt[k]);
                if (longest > 0.0) { bound = longest; break; }
            }
        }
        const double cap = std::fmin(dt_cap > 0.0 ? dt_cap : std::numeric_limits<double>::infinity(), bound);
        int fin = 0;
        double dt = assign_levels(edt, max_levels, cap, fin, level);
        stable = dt;
        if (landing > 0.0 && fin > 0 && std::isfinite(dt)) {
            const double steps = std::fmax(std::ceil(landing / (dt * (1.0 + LANDING_SLACK))), 1.0);
            const double fitted = landing / steps;
            if (fitted < dt) {
                dt = fitted;
                fin = levels_for_step(edt, dt, max_levels, level);
            }
        }
        finest = fin;
        const bool lines = line_rates && n1 >= 2;
        ll.assign(K, 0);
        if (lines)
            for (size_t k = 0; k < K; ++k) ll[k] = (flags[k] & 4) ? 1 : 0;
        if (memo && memo_finest == fin && memo_given == level && memo_ll == ll) {
            level = memo_limited;
            return dt;
        }
        changed = true;
        memo_given = level;
        while (spread_to_neighbours(nb, level, scratch, 1)) level.swap(scratch);
        rates.resize((size_t)n_stages * K);
        if (lines) {
            const size_t nn = (size_t)n1 * n1;
            node_from.resize(K * nn);
            for (size_t k = 0; k < K; ++k) std::memset(node_from.data() + k * nn, level[k], nn);
            for (int q = 0; q < n_stages; ++q) {
                spread_over_lines(nb, links, n1, ll, node_from, node_to, rate_tmp);
                std::memcpy(rates.data() + (size_t)q * K, rate_tmp.data(), K);
                node_from.swap(node_to);
            }
        } else {
            std::vector<uint8_t> from = level, to;
            for (int q = 0; q < n_stages; ++q) {
                spread_to_neighbours(nb, from, to, 0);
                std::memcpy(rates.data() + (size_t)q * K, to.data(), K);
                from.swap(to);
            }
        }
        evals_per_step = 0;
        for (uint8_t r : rates) evals_per_step += int64_t(1) << r;
        memo = true;
        memo_finest = fin;
        memo_ll = ll;
        memo_limited = level;
        return dt;
    }
};

// The coarsest level whose substeps start at finest substep p (multirate.rs:842-849)
inline int depth(int p, int finest) {
    if (p == 0 || p == (1 << finest)) return 0;
    int tz = 0;
    while (!((p >> tz) & 1)) ++tz;
    return finest - tz;
}

}  // namespace multirate
}  // namespace dgswe

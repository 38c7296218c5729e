// Multirate local time stepping, the host side (MultirateStepper,
// src/time/multirate.rs:713-1330, with SWEPhysics2D's line-local flags and
// coarse-step bound, src/physics/builder.rs:563-637): levels from the element
// time steps, the level-jump limit, the stage rates (node by node over
// line-local elements, else one face per stage), then a coarse step's passes
// over the device's element-subset stages. Included by dgswe.hip (host code).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

// The coarse step's plan and an element's level are evaluated on the device
// too (the same IEEE operations, so the same bits)
#if defined(__HIPCC__) || defined(DGSWE_EMULATE)
#define MR_HD __host__ __device__ inline
#else
#define MR_HD inline
#endif

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

// The coarse step from the extremes of the finite element time steps: lo, hi
// the least and the largest (lo = inf: none is finite), long1, long2 the largest
// of an element whose depth flags have bit 0, bit 1 (0: none) for the module's
// bound (builder.rs:614-637); assign_levels' coarsening (multirate.rs:526-568)
// and the fit to the landing (:1014-1043). The levels are element_level(edt, dt,
// bound) for every element; with `fitted` the finest level is their largest
// (levels_for_step, :574-586), else `finest`.
struct Plan {
    double dt, stable;  // the coarse step (fitted to the landing), and the one the elements allow
    int finest, bound;  // assign_levels' finest level; the level loop's limit
    int fitted;
};
MR_HD Plan plan_coarse_step(double lo, double hi, double long1, double long2, bool wet_dry, double dt_cap,
                            double landing, int max_levels) {
    const double inf = __builtin_inf();
    double bound = inf;
    if (wet_dry) {
        if (long1 > 0.0) bound = long1;
        else if (long2 > 0.0) bound = long2;
    }
    const double cap = fmin(dt_cap > 0.0 ? dt_cap : inf, bound);
    Plan p = {inf, inf, 0, 0, 0};
    if (!(lo < inf)) return p;
    int finest = 0;
    while (finest < max_levels) {
        const double coarser = lo * (double)(2ull << finest);
        if (coarser > hi || coarser > cap) break;
        ++finest;
    }
    const double dt = lo * (double)(1ull << finest);
    p.dt = p.stable = dt;
    p.finest = p.bound = finest;
    if (landing > 0.0 && finest > 0 && dt < inf) {
        const double steps = fmax(ceil(landing / (dt * (1.0 + LANDING_SLACK))), 1.0);
        const double fitted = landing / steps;
        if (fitted < dt) {
            p.dt = fitted;
            p.bound = max_levels;
            p.fitted = 1;
        }
    }
    return p;
}

// An element's level: halve dt until it is at most the element's step, at most `bound` times
MR_HD int element_level(double edt, double dt, int bound) {
    int l = 0;
    double dl = dt;
    while (l < bound && dl > edt) { dl *= 0.5; ++l; }
    return l;
}

// The host loops over elements on threads: host_parts(n) parts of [0, n) (up to
// 16, one per 32768 items, at most the hardware's threads), and host_parallel
// runs body(begin, end, part) on each contiguous part, one thread per part.
// Every loop run this way writes each element's own results only.
inline int host_parts(size_t n) {
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t by_size = n / 32768 + 1;
    size_t t = std::min<size_t>(std::min<size_t>(hw ? hw : 1, 16), by_size);
    return (int)(t < 1 ? 1 : t);
}
template <class F>
inline void host_parallel(size_t n, int parts, F body) {
    if (parts <= 1) { body((size_t)0, n, 0); return; }
    std::vector<std::thread> th;
    th.reserve((size_t)parts);
    for (int q = 0; q < parts; ++q) {
        const size_t b = n * (size_t)q / (size_t)parts, e = n * (size_t)(q + 1) / (size_t)parts;
        th.emplace_back([&body, b, e, q] { body(b, e, q); });
    }
    for (auto& t : th) t.join();
}

// spread_to_neighbours (multirate.rs:592-613): to[k] = max(from[k], from[j] - drop)
inline bool spread_to_neighbours(const std::vector<int64_t>& nb, const std::vector<uint8_t>& from,
                                 std::vector<uint8_t>& to, int drop) {
    const size_t K = from.size();
    to.resize(K);
    const int parts = host_parts(K);
    std::vector<uint8_t> changed((size_t)parts, 0);
    host_parallel(K, parts, [&](size_t b, size_t e, int q) {
        bool c = false;
        for (size_t k = b; k < e; ++k) {
            int m = from[k];
            for (int f = 0; f < 4; ++f) {
                const int64_t j = nb[4 * k + f];
                if (j >= 0) { const int v = (int)from[(size_t)j] - drop; m = v > m ? v : m; }
            }
            to[k] = (uint8_t)m;
            c = c || to[k] != from[k];
        }
        changed[(size_t)q] = c ? 1 : 0;
    });
    for (uint8_t c : changed)
        if (c) return true;
    return false;
}

// The node of face f at position fi (node j*n1 + i on r-line j, s-line i),
// counter-clockwise: node fi of a face is node n1 - 1 - fi of the neighbour's
inline int face_node(int n1, int f, int fi) {
    return f == 0 ? fi : f == 1 ? fi * n1 + n1 - 1 : f == 2 ? (n1 - 1) * n1 + (n1 - 1 - fi) : (n1 - 1 - fi) * n1;
}
inline int inner_offset(int n1, int f) { return f == 0 ? n1 : f == 1 ? -1 : f == 2 ? -n1 : 1; }

// spread_over_lines (multirate.rs:705-767): node rates of a stage's value from
// those of its input, and each element's RHS rate
template <int n1>
inline void spread_over_lines_n(const std::vector<int64_t>& nb, const std::vector<uint32_t>& links,
                                const std::vector<uint8_t>& line_local, const std::vector<uint8_t>& from,
                                std::vector<uint8_t>& to, std::vector<uint8_t>& rate) {
    constexpr size_t nn = (size_t)n1 * n1;
    const size_t K = line_local.size();
    to.resize(K * nn);
    rate.resize(K);
    host_parallel(K, host_parts(K), [&](size_t kb, size_t ke, int) {
        for (size_t k = kb; k < ke; ++k) {
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
    });
}
inline void spread_over_lines(const std::vector<int64_t>& nb, const std::vector<uint32_t>& links, int n1,
                              const std::vector<uint8_t>& line_local, const std::vector<uint8_t>& from,
                              std::vector<uint8_t>& to, std::vector<uint8_t>& rate) {
    switch (n1) {
        case 2: spread_over_lines_n<2>(nb, links, line_local, from, to, rate); break;
        case 3: spread_over_lines_n<3>(nb, links, line_local, from, to, rate); break;
        default: spread_over_lines_n<4>(nb, links, line_local, from, to, rate); break;
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

    // The levels of an element step sequence under its plan (the host's own
    // pass; the library evaluates the extremes and the levels on the device)
    double assign_kept(const uint8_t* flags, bool wet_dry, double dt_cap, double landing) {
        const size_t K = edt.size();
        const double inf = std::numeric_limits<double>::infinity();
        double lo = inf, hi = 0.0, long1 = 0.0, long2 = 0.0;
        for (size_t k = 0; k < K; ++k)
            if (std::isfinite(edt[k])) {
                lo = std::fmin(lo, edt[k]);
                hi = std::fmax(hi, edt[k]);
                if (flags[k] & 1) long1 = std::fmax(long1, edt[k]);
                if (flags[k] & 2) long2 = std::fmax(long2, edt[k]);
            }
        const Plan p = plan_coarse_step(lo, hi, long1, long2, wet_dry, dt_cap, landing, max_levels);
        level.resize(K);
        int top = 0;
        for (size_t k = 0; k < K; ++k) {
            const int l = element_level(edt[k], p.dt, p.bound);
            level[k] = (uint8_t)l;
            top = top > l ? top : l;
        }
        const bool lines = line_rates && n1 >= 2;
        ll.assign(K, 0);
        if (lines)
            for (size_t k = 0; k < K; ++k) ll[k] = (flags[k] & 4) ? 1 : 0;
        const int fin = p.fitted ? top : p.finest;
        stable = p.stable;
        if (kept(fin, level, ll)) return p.dt;
        apply(fin);
        return p.dt;
    }

    // After the device half of an assignment (the plan, whether some element's
    // level or line-local flag differs from the kept assignment's inputs, the
    // largest level): the finest level, and whether the kept assignment holds
    // (then `level` is its limited levels already); else the caller brings the
    // levels and flags back into `level` and `ll` and calls apply(finest)
    bool device_kept(const Plan& p, int differs, int top) {
        stable = p.stable;
        finest = p.fitted ? top : p.finest;
        return memo && memo_finest == finest && !differs;
    }

    // The last assignment holds when its finest level and its levels before the
    // limit and line-local flags are the same (multirate.rs:1114-1121)
    bool kept(int fin, const std::vector<uint8_t>& given, const std::vector<uint8_t>& given_ll) {
        finest = fin;
        if (memo && memo_finest == fin && memo_given == given && memo_ll == given_ll) {
            level = memo_limited;
            return true;
        }
        return false;
    }
    // A new assignment from `level` (before the limit) and `ll`: the level-jump
    // limit, the stage rates, kept for the next coarse steps
    void apply(int fin) {
        const size_t K = level.size();
        finest = fin;
        changed = true;
        memo_given = level;
        while (spread_to_neighbours(nb, level, scratch, 1)) level.swap(scratch);
        rates.resize((size_t)n_stages * K);
        if (line_rates && n1 >= 2) {
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
    }
};

#if defined(__HIPCC__) || defined(DGSWE_EMULATE)
// The device half of an assignment: the extremes of the element time steps,
// the plan, every element's level and line-local flag, and whether they differ
// from the kept assignment's (so that a coarse step whose levels are the last
// one's brings back a few bytes instead of the element time steps).

// slots: the bits of lo, hi, long1, long2 (non-negative doubles order like
// their bits); out: [0] some element differs from the kept assignment, [1]
// the largest level
__global__ void mr_clear_kernel(unsigned long long* slots, int* out) {
    if (threadIdx.x == 0) {
        slots[0] = 0x7FF0000000000000ull;  // +inf
        slots[1] = slots[2] = slots[3] = 0ull;
        out[0] = out[1] = 0;
    }
}

__global__ __launch_bounds__(256) void mr_extremes_kernel(const double* edt, const uint8_t* flags, int64_t K,
                                                          unsigned long long* slots) {
    __shared__ double red[4][256];
    const int t = threadIdx.x;
    const double inf = __builtin_inf();
    double lo = inf, hi = 0.0, l1 = 0.0, l2 = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + t; k < K; k += (int64_t)gridDim.x * 256) {
        const double d = edt[k];
        if (!(d < inf)) continue;  // not finite (the steps are non-negative)
        const uint8_t f = flags[k];
        lo = d < lo ? d : lo;
        hi = d > hi ? d : hi;
        if ((f & 1) && d > l1) l1 = d;
        if ((f & 2) && d > l2) l2 = d;
    }
    red[0][t] = lo; red[1][t] = hi; red[2][t] = l1; red[3][t] = l2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] = red[0][t + w] < red[0][t] ? red[0][t + w] : red[0][t];
            for (int q = 1; q < 4; ++q) red[q][t] = red[q][t + w] > red[q][t] ? red[q][t + w] : red[q][t];
        }
        __syncthreads();
    }
    if (t == 0) {
        atomicMin(slots + 0, (unsigned long long)__double_as_longlong(red[0][0]));
        for (int q = 1; q < 4; ++q) atomicMax(slots + q, (unsigned long long)__double_as_longlong(red[q][0]));
    }
}

__global__ void mr_plan_kernel(const unsigned long long* slots, int wet_dry, double dt_cap, double landing,
                               int max_levels, Plan* plan) {
    if (threadIdx.x != 0) return;
    *plan = plan_coarse_step(__longlong_as_double((long long)slots[0]), __longlong_as_double((long long)slots[1]),
                             __longlong_as_double((long long)slots[2]), __longlong_as_double((long long)slots[3]),
                             wet_dry != 0, dt_cap, landing, max_levels);
}

// memo_given, memo_ll: the kept assignment's levels before the limit and its
// line-local flags (null: none kept)
__global__ __launch_bounds__(256) void mr_levels_kernel(const double* edt, const uint8_t* flags, int64_t K,
                                                        const Plan* plan, int lines, const uint8_t* memo_given,
                                                        const uint8_t* memo_ll, uint8_t* level, uint8_t* ll, int* out) {
    __shared__ int red[2][256];
    const int t = threadIdx.x;
    const double dt = plan->dt;
    const int bound = plan->bound;
    int diff = memo_given ? 0 : 1, top = 0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + t; k < K; k += (int64_t)gridDim.x * 256) {
        const int l = element_level(edt[k], dt, bound);
        const uint8_t b = lines && (flags[k] & 4) ? 1 : 0;
        level[k] = (uint8_t)l;
        ll[k] = b;
        if (memo_given && (memo_given[k] != (uint8_t)l || memo_ll[k] != b)) diff = 1;
        top = l > top ? l : top;
    }
    red[0][t] = diff;
    red[1][t] = top;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] |= red[0][t + w];
            red[1][t] = red[1][t + w] > red[1][t] ? red[1][t + w] : red[1][t];
        }
        __syncthreads();
    }
    if (t == 0) {
        if (red[0][0]) atomicOr(out, 1);
        atomicMax(out + 1, red[1][0]);
    }
}
#endif

// The coarsest level whose substeps start at finest substep p (multirate.rs:842-849)
inline int depth(int p, int finest) {
    if (p == 0 || p == (1 << finest)) return 0;
    int tz = 0;
    while (!((p >> tz) & 1)) ++tz;
    return finest - tz;
}

}  // namespace multirate
}  // namespace dgswe

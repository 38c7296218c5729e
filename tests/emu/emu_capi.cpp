// C ABI over the emulated product kernels (tests only).
#include "hip_emu.hpp"

#include <algorithm>

#include "../../dgcom_amd/csrc/swe_kernels.hpp"
#include "../../dgcom_amd/csrc/multirate.hpp"

using namespace dgswe;

extern "C" {

void emu_init(void) {
    for (int n = 1; n <= 3; ++n) c_gll[n] = make_tables(n);
}

// The static geometry cache (geometry_kernel) of a form
static std::vector<double> g_emass;  // the positivity bound's static mass ratios of the last form
static void run_geometry(int order, int64_t K, StageParams& p, std::vector<double>& geo, std::vector<double>& fgeo) {
    g_emass.assign((size_t)K, 0.0);
    size_t nn = (size_t)(order + 1) * (order + 1);
    geo.assign((size_t)GEO_COUNT * K * nn, 0.0);
    fgeo.assign((size_t)(3 * 4 + 4 * (order + 1)) * K, 0.0);
    int epb = 64 / (int)nn;
    unsigned grid = (unsigned)((K + epb - 1) / epb);
    switch (order) {
        case 1: emu_launch(grid, 64, [&] { geometry_kernel<1>(p.corners, p.affine, p.links, K, geo.data(), fgeo.data(), g_emass.data()); }); break;
        case 2: emu_launch(grid, 64, [&] { geometry_kernel<2>(p.corners, p.affine, p.links, K, geo.data(), fgeo.data(), g_emass.data()); }); break;
        default: emu_launch(grid, 64, [&] { geometry_kernel<3>(p.corners, p.affine, p.links, K, geo.data(), fgeo.data(), g_emass.data()); }); break;
    }
    p.geo = geo.data();
    p.fgeo = fgeo.data();
}

// The forward stage's launch sequence (wet kernel, list compaction, subcell kernel)
static void run_stage(int order, int64_t K, StageParams& p) {
    size_t nn = (size_t)(order + 1) * (order + 1);
    int epb = 64 / (int)nn;
    unsigned nblk = (unsigned)((K + epb - 1) / epb);
    unsigned nchunk = (nblk + 1023) / 1024;
    std::vector<uint32_t> queue((size_t)nblk * epb), list((size_t)K);
    std::vector<uint8_t> qcount(nblk);
    std::vector<unsigned int> ctot(nchunk, ~0u);  // unread by a one-chunk compaction
    unsigned int lcount = 0;
    p.queue = queue.data();
    p.queue_count = qcount.data();
    p.xcd_stride = (nblk + 7) / 8;
    unsigned grid = 8 * p.xcd_stride, sub_grid = std::min(nblk, 5u);
    switch (order) {
        case 1: emu_launch(grid, 64, [&] { wet_stage_kernel<1>(p); }); break;
        case 2: emu_launch(grid, 64, [&] { wet_stage_kernel<2>(p); }); break;
        default: emu_launch(grid, 64, [&] { wet_stage_kernel<3>(p); }); break;
    }
    if (nchunk > 1) emu_launch(nchunk, 256, [&] { chunk_count_kernel(qcount.data(), nblk, ctot.data()); });
    emu_launch(nchunk, 256, [&] { compact_kernel(qcount.data(), queue.data(), nblk, ctot.data(), epb, list.data(), &lcount); });
    switch (order) {
        case 1: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<1>(p, list.data(), &lcount); }); break;
        case 2: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<2>(p, list.data(), &lcount); }); break;
        default: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<3>(p, list.data(), &lcount); }); break;
    }
    p.queue = nullptr;
    p.queue_count = nullptr;
}

unsigned long long emu_stage(int order, int64_t K, const double* corners, const uint32_t* links, const uint8_t* tags,
                             const uint8_t* affine, const double* bed, const double* manning, const double* base,
                             const double* input, double* out, double a, double b, double c, double eta_e, double g,
                             double h_min, double h_dry, double u_max, double tau, double f0, int rhs_only, int flags,
                             double* fstar_out) {
    size_t nn = (size_t)(order + 1) * (order + 1), nt = (size_t)K * nn;
    unsigned long long clips = 0;
    StageParams p{};
    p.bh = base; p.bhu = base + nt; p.bhv = base + 2 * nt;
    p.ih = input; p.ihu = input + nt; p.ihv = input + 2 * nt;
    p.oh = out; p.ohu = out + nt; p.ohv = out + 2 * nt;
    p.bed = bed; p.corners = corners; p.manning = manning; p.links = links; p.tags = tags; p.affine = affine;
    p.n_elements = K; p.a = a; p.b = b; p.c = c; p.eta_e = eta_e;
    p.g = g; p.h_min = h_min; p.h_dry = h_dry; p.u_max = u_max; p.tau = tau; p.f0 = f0; p.sqrt_g = std::sqrt(g);
    p.rhs_only = rhs_only; p.flags = flags; p.clips = &clips; p.fstar_out = fstar_out;
    p.rh = p.rhu = p.rhv = nullptr;
    std::vector<double> geo, fgeo;
    run_geometry(order, K, p, geo, fgeo);
    run_stage(order, K, p);
    return clips;
}

// A pass of local time stepping over the listed elements (wet_lts_kernel, the
// list compaction, the subcell kernel's LTS instantiation), each element with its
// level's c and tide elevation; acc (if not null) the last stage's sum.
unsigned long long emu_stage_where(int order, int64_t K, const double* corners, const uint32_t* links,
                                   const uint8_t* tags, const uint8_t* affine, const double* bed,
                                   const double* manning, const double* base, const double* input, double* out,
                                   double* acc, const uint32_t* elist, int64_t n_list, const uint8_t* level,
                                   const uint8_t* lrate, int depth, double a, double b, const double* lvl_c,
                                   const double* lvl_eta, double g, double h_min, double h_dry, double u_max,
                                   double tau, double f0, int flags) {
    size_t nn = (size_t)(order + 1) * (order + 1), nt = (size_t)K * nn;
    unsigned long long clips = 0;
    StageParams p{};
    std::memset(&p, 0, sizeof(p));
    p.bh = base; p.bhu = base + nt; p.bhv = base + 2 * nt;
    p.ih = input; p.ihu = input + nt; p.ihv = input + 2 * nt;
    p.oh = out; p.ohu = out + nt; p.ohv = out + 2 * nt;
    p.bed = bed; p.corners = corners; p.manning = manning; p.links = links; p.tags = tags; p.affine = affine;
    p.n_elements = K; p.a = a; p.b = b;
    p.g = g; p.h_min = h_min; p.h_dry = h_dry; p.u_max = u_max; p.tau = tau; p.f0 = f0; p.sqrt_g = std::sqrt(g);
    p.flags = flags; p.clips = &clips;
    p.elist = elist; p.n_list = n_list; p.level = level; p.lrate = lrate; p.depth = depth;
    if (acc) { p.ah = acc; p.ahu = acc + nt; p.ahv = acc + 2 * nt; }
    for (int l = 0; l <= DG_MAX_LEVELS; ++l) { p.lvl_c[l] = lvl_c[l]; p.lvl_eta[l] = lvl_eta[l]; }
    std::vector<double> geo, fgeo;
    run_geometry(order, K, p, geo, fgeo);
    int epb = 64 / (int)nn;
    unsigned nblk = (unsigned)((n_list + epb - 1) / epb);
    unsigned nchunk = (nblk + 1023) / 1024;
    std::vector<uint32_t> queue((size_t)nblk * epb), list((size_t)K);
    std::vector<uint8_t> qcount(nblk);
    std::vector<unsigned int> ctot(nchunk, ~0u);  // unread by a one-chunk compaction
    unsigned int lcount = 0;
    p.queue = queue.data();
    p.queue_count = qcount.data();
    p.xcd_stride = (nblk + 7) / 8;
    unsigned grid = 8 * p.xcd_stride, sub_grid = std::min(nblk, 5u);
    switch (order) {
        case 1: emu_launch(grid, 64, [&] { wet_lts_kernel<1>(p); }); break;
        case 2: emu_launch(grid, 64, [&] { wet_lts_kernel<2>(p); }); break;
        default: emu_launch(grid, 64, [&] { wet_lts_kernel<3>(p); }); break;
    }
    if (nchunk > 1) emu_launch(nchunk, 256, [&] { chunk_count_kernel(qcount.data(), nblk, ctot.data()); });
    emu_launch(nchunk, 256, [&] { compact_kernel(qcount.data(), queue.data(), nblk, ctot.data(), epb, list.data(), &lcount); });
    switch (order) {
        case 1: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<1, true>(p, list.data(), &lcount); }); break;
        case 2: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<2, true>(p, list.data(), &lcount); }); break;
        default: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<3, true>(p, list.data(), &lcount); }); break;
    }
    return clips;
}

// The same pass with the forcing given per node compiled in (the kernels' FORCE
// instantiations): tides per boundary face node (slot [K][4][N+1], per slot the
// constants [n_c][6], mean level and source depth) with each level's moment row
// (rows [DG_MAX_LEVELS + 1][LVL_MOMENT]), Manning n per node and Coriolis f per node.
unsigned long long emu_stage_where_forced(int order, int64_t K, const double* corners, const uint32_t* links,
                                          const uint8_t* tags, const uint8_t* affine, const double* bed,
                                          const double* manning, const double* base, const double* input, double* out,
                                          double* acc, const uint32_t* elist, int64_t n_list, const uint8_t* level,
                                          const uint8_t* lrate, int depth, double a, double b, const double* lvl_c,
                                          const double* lvl_eta, double g, double h_min, double h_dry, double u_max,
                                          double tau, double f0, int flags, const uint32_t* tide_slot,
                                          const double* tide_coef, const double* tide_mean, const double* tide_src,
                                          int atlas_nc, int atlas_velocity, double atlas_max_ratio,
                                          const double* lvl_moment, const double* manning_node,
                                          const double* coriolis_node) {
    size_t nn = (size_t)(order + 1) * (order + 1), nt = (size_t)K * nn;
    unsigned long long clips = 0;
    StageParams p{};
    std::memset(&p, 0, sizeof(p));
    p.bh = base; p.bhu = base + nt; p.bhv = base + 2 * nt;
    p.ih = input; p.ihu = input + nt; p.ihv = input + 2 * nt;
    p.oh = out; p.ohu = out + nt; p.ohv = out + 2 * nt;
    p.bed = bed; p.corners = corners; p.manning = manning; p.links = links; p.tags = tags; p.affine = affine;
    p.n_elements = K; p.a = a; p.b = b;
    p.g = g; p.h_min = h_min; p.h_dry = h_dry; p.u_max = u_max; p.tau = tau; p.f0 = f0; p.sqrt_g = std::sqrt(g);
    p.flags = flags; p.clips = &clips;
    p.elist = elist; p.n_list = n_list; p.level = level; p.lrate = lrate; p.depth = depth;
    if (acc) { p.ah = acc; p.ahu = acc + nt; p.ahv = acc + 2 * nt; }
    for (int l = 0; l <= DG_MAX_LEVELS; ++l) { p.lvl_c[l] = lvl_c[l]; p.lvl_eta[l] = lvl_eta[l]; }
    p.tide_slot = tide_slot; p.tide_coef = tide_coef; p.tide_mean = tide_mean; p.tide_src = tide_src;
    p.atlas_nc = atlas_nc; p.atlas_velocity = atlas_velocity; p.atlas_max_ratio = atlas_max_ratio;
    p.lvl_moment = lvl_moment; p.manning_node = manning_node; p.coriolis_node = coriolis_node;
    std::vector<double> geo, fgeo;
    run_geometry(order, K, p, geo, fgeo);
    int epb = 64 / (int)nn;
    unsigned nblk = (unsigned)((n_list + epb - 1) / epb);
    unsigned nchunk = (nblk + 1023) / 1024;
    std::vector<uint32_t> queue((size_t)nblk * epb), list((size_t)K);
    std::vector<uint8_t> qcount(nblk);
    std::vector<unsigned int> ctot(nchunk, ~0u);  // unread by a one-chunk compaction
    unsigned int lcount = 0;
    p.queue = queue.data();
    p.queue_count = qcount.data();
    p.xcd_stride = (nblk + 7) / 8;
    unsigned grid = 8 * p.xcd_stride, sub_grid = std::min(nblk, 5u);
    switch (order) {
        case 1: emu_launch(grid, 64, [&] { wet_lts_kernel<1, true>(p); }); break;
        case 2: emu_launch(grid, 64, [&] { wet_lts_kernel<2, true>(p); }); break;
        default: emu_launch(grid, 64, [&] { wet_lts_kernel<3, true>(p); }); break;
    }
    if (nchunk > 1) emu_launch(nchunk, 256, [&] { chunk_count_kernel(qcount.data(), nblk, ctot.data()); });
    emu_launch(nchunk, 256, [&] { compact_kernel(qcount.data(), queue.data(), nblk, ctot.data(), epb, list.data(), &lcount); });
    switch (order) {
        case 1: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<1, true, true>(p, list.data(), &lcount); }); break;
        case 2: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<2, true, true>(p, list.data(), &lcount); }); break;
        default: emu_launch(sub_grid, 64, [&] { subcell_stage_kernel<3, true, true>(p, list.data(), &lcount); }); break;
    }
    return clips;
}

// f0 + beta (y - y0) at every node (coriolis_node_kernel)
void emu_coriolis_nodes(int order, int64_t K, const double* corners, double f0, double beta, double y0, double* f) {
    size_t nt = (size_t)K * (order + 1) * (order + 1);
    unsigned grid = (unsigned)((nt + 255) / 256);
    switch (order) {
        case 1: emu_launch(grid, 256, [&] { coriolis_node_kernel<1>(corners, K, f0, beta, y0, f); }); break;
        case 2: emu_launch(grid, 256, [&] { coriolis_node_kernel<2>(corners, K, f0, beta, y0, f); }); break;
        default: emu_launch(grid, 256, [&] { coriolis_node_kernel<3>(corners, K, f0, beta, y0, f); }); break;
    }
}

int emu_lvl_moment(void) { return LVL_MOMENT; }

unsigned long long emu_finish_where(int order, int64_t K, const double* corners, const uint8_t* affine,
                                    double* state, const double* acc, const uint32_t* elist, int64_t n_list,
                                    double h_dry, double u_max, int wet_dry) {
    size_t nn = (size_t)(order + 1) * (order + 1), nt = (size_t)K * nn;
    unsigned long long clips = 0;
    StageParams sp{};
    std::memset(&sp, 0, sizeof(sp));
    std::vector<double> bed(nt, 0.0), geo, fgeo;
    std::vector<uint32_t> links((size_t)4 * K, 0xFFFFFFFFu);
    sp.corners = corners; sp.bed = bed.data(); sp.affine = affine; sp.links = links.data();
    run_geometry(order, K, sp, geo, fgeo);
    FinishParams p;
    p.elist = elist; p.n_list = n_list;
    p.ah = acc; p.ahu = acc + nt; p.ahv = acc + 2 * nt;
    p.h = state; p.hu = state + nt; p.hv = state + 2 * nt;
    p.geo = geo.data(); p.n_elements = K; p.h_dry = h_dry; p.u_max = u_max; p.wet_dry = wet_dry; p.clips = &clips;
    int epb = 64 / (int)nn;
    unsigned grid = (unsigned)((n_list + epb - 1) / epb);
    switch (order) {
        case 1: emu_launch(grid, 64, [&] { finish_kernel<1>(p); }); break;
        case 2: emu_launch(grid, 64, [&] { finish_kernel<2>(p); }); break;
        default: emu_launch(grid, 64, [&] { finish_kernel<3>(p); }); break;
    }
    return clips;
}

// The library's multirate level assignment (multirate::Assigner, host code) over
// a sequence of element time steps and depth flags, one assigner throughout
// (its kept rates included)
void emu_multirate_assign(int64_t K, int n1, const uint32_t* links, int max_levels, int n_stages, int line_rates,
                          int calls, const double* edt, const uint8_t* flags, int wet_dry, const double* dt_cap,
                          const double* landing, double* dt_out, double* stable_out, int* finest_out,
                          uint8_t* level_out, uint8_t* rates_out) {
    dgswe::multirate::Assigner as;
    as.max_levels = max_levels;
    as.n_stages = n_stages;
    as.n1 = n1;
    as.line_rates = line_rates != 0;
    as.init(links, (size_t)K);
    for (int q = 0; q < calls; ++q) {
        dt_out[q] = as.assign(edt + (size_t)q * K, (size_t)K, flags + (size_t)q * K, wet_dry != 0, dt_cap[q], landing[q]);
        stable_out[q] = as.stable;
        finest_out[q] = as.finest;
        std::memcpy(level_out + (size_t)q * K, as.level.data(), (size_t)K);
        std::memcpy(rates_out + (size_t)q * n_stages * K, as.rates.data(), (size_t)n_stages * K);
    }
}

// The stage VJP: gbase (if not null) and ginput are added to; gbed and gman too.
// The forward stage runs first and saves its RHS, as the library does.
void emu_stage_vjp(int order, int64_t K, const double* corners, const uint32_t* links, const uint8_t* tags,
                   const uint8_t* affine, const double* bed, const double* manning, const double* base,
                   const double* input, const double* w, double* gbase, double* ginput, double* gbed, double* gman,
                   double a, double b, double c, double eta_e, double g, double h_min, double h_dry, double u_max,
                   double tau, double f0, int rhs_only, int flags) {
    size_t nn = (size_t)(order + 1) * (order + 1), nt = (size_t)K * nn, nf = 4 * (size_t)(order + 1);
    unsigned long long clips = 0;
    VjpParams P{};
    StageParams& p = P.s;
    p.bh = base; p.bhu = base + nt; p.bhv = base + 2 * nt;
    p.ih = input; p.ihu = input + nt; p.ihv = input + 2 * nt;
    p.oh = p.ohu = p.ohv = nullptr;
    p.bed = bed; p.corners = corners; p.manning = manning; p.links = links; p.tags = tags; p.affine = affine;
    p.n_elements = K; p.a = a; p.b = b; p.c = c; p.eta_e = eta_e;
    p.g = g; p.h_min = h_min; p.h_dry = h_dry; p.u_max = u_max; p.tau = tau; p.f0 = f0; p.sqrt_g = std::sqrt(g);
    p.rhs_only = rhs_only; p.flags = flags; p.clips = &clips; p.fstar_out = nullptr;
    std::vector<double> rhs(3 * nt, 0.0), fout(3 * nt, 0.0), geo, fgeo;
    run_geometry(order, K, p, geo, fgeo);
    if (!rhs_only) {
        p.oh = fout.data(); p.ohu = p.oh + nt; p.ohv = p.oh + 2 * nt;
        p.rh = rhs.data(); p.rhu = p.rh + nt; p.rhv = p.rh + 2 * nt;
        run_stage(order, K, p);
    } else {
        p.rh = p.rhu = p.rhv = nullptr;
    }
    P.wh = w; P.whu = w + nt; P.whv = w + 2 * nt;
    P.gbh = gbase; P.gbhu = gbase ? gbase + nt : nullptr; P.gbhv = gbase ? gbase + 2 * nt : nullptr;
    P.gih = ginput; P.gihu = ginput + nt; P.gihv = ginput + 2 * nt;
    P.gbed = gbed; P.gman = gman;
    std::vector<double> fb((size_t)K * nf * 4, 0.0), ib((size_t)K * nf * 4, 0.0), fc((size_t)K * nf * 3, 0.0);
    std::vector<uint8_t> kind((size_t)K, 0);
    P.facebuf = fb.data(); P.innerbuf = ib.data(); P.kind = kind.data(); P.facecot = fc.data();
    P.input_write = 0;
    int epb = 64 / (int)nn;
    unsigned nblk = (unsigned)((K + epb - 1) / epb);
    unsigned nchunk = (nblk + 1023) / 1024;
    std::vector<uint32_t> queue((size_t)nblk * epb), list((size_t)K);
    std::vector<uint8_t> qcount(nblk);
    std::vector<unsigned int> ctot(nchunk, ~0u);  // unread by a one-chunk compaction
    unsigned int lcount = 0;
    p.queue = queue.data();
    p.queue_count = qcount.data();
    p.xcd_stride = (nblk + 7) / 8;
    unsigned grid = 8 * p.xcd_stride, sub_grid = std::min(nblk, 5u);
    switch (order) {
        case 1: emu_launch(grid, 64, [&] { wet_vjp_kernel<1>(P); }); break;
        case 2: emu_launch(grid, 64, [&] { wet_vjp_kernel<2>(P); }); break;
        default: emu_launch(grid, 64, [&] { wet_vjp_kernel<3>(P); }); break;
    }
    if (nchunk > 1) emu_launch(nchunk, 256, [&] { chunk_count_kernel(qcount.data(), nblk, ctot.data()); });
    emu_launch(nchunk, 256, [&] { compact_kernel(qcount.data(), queue.data(), nblk, ctot.data(), epb, list.data(), &lcount); });
    switch (order) {
        case 1: emu_launch(sub_grid, 64, [&] { subcell_vjp_kernel<1>(P, list.data(), &lcount); }); break;
        case 2: emu_launch(sub_grid, 64, [&] { subcell_vjp_kernel<2>(P, list.data(), &lcount); }); break;
        default: emu_launch(sub_grid, 64, [&] { subcell_vjp_kernel<3>(P, list.data(), &lcount); }); break;
    }
    std::vector<uint32_t> edges, split;
    vjp_edge_lists(links, fgeo.data(), fgeo.data() + (size_t)K * 4, K, edges, split);
    const uint32_t ne = (uint32_t)edges.size(), ns = (uint32_t)split.size();
    unsigned fgrid = (unsigned)(((size_t)ne * (order + 1) + 63) / 64);
    switch (order) {
        case 1: emu_launch(fgrid, 64, [&] { edge_vjp_kernel<1>(P, edges.data(), ne); }); break;
        case 2: emu_launch(fgrid, 64, [&] { edge_vjp_kernel<2>(P, edges.data(), ne); }); break;
        default: emu_launch(fgrid, 64, [&] { edge_vjp_kernel<3>(P, edges.data(), ne); }); break;
    }
    if (ns) {
        unsigned sgrid = (unsigned)(((size_t)ns * (order + 1) + 63) / 64);
        switch (order) {
            case 1: emu_launch(sgrid, 64, [&] { edge_vjp_split_kernel<1>(P, split.data(), ns); }); break;
            case 2: emu_launch(sgrid, 64, [&] { edge_vjp_split_kernel<2>(P, split.data(), ns); }); break;
            default: emu_launch(sgrid, 64, [&] { edge_vjp_split_kernel<3>(P, split.data(), ns); }); break;
        }
    }
    unsigned ggrid = (unsigned)((nt + 255) / 256);
    switch (order) {
        case 1: emu_launch(ggrid, 256, [&] { vjp_gather_kernel<1>(links, fb.data(), ib.data(), kind.data(), K, ginput, ginput + nt, ginput + 2 * nt, gbed); }); break;
        case 2: emu_launch(ggrid, 256, [&] { vjp_gather_kernel<2>(links, fb.data(), ib.data(), kind.data(), K, ginput, ginput + nt, ginput + 2 * nt, gbed); }); break;
        default: emu_launch(ggrid, 256, [&] { vjp_gather_kernel<3>(links, fb.data(), ib.data(), kind.data(), K, ginput, ginput + nt, ginput + 2 * nt, gbed); }); break;
    }
}

// The same sequence through the library's device half (the mr_*_kernel
// launches of dg_multirate_assign_levels, emulated) and its host remainder
void emu_multirate_assign_device(int64_t K, int n1, const uint32_t* links, int max_levels, int n_stages,
                                 int line_rates, int calls, const double* edt, const uint8_t* flags, int wet_dry,
                                 const double* dt_cap, const double* landing, double* dt_out, double* stable_out,
                                 int* finest_out, uint8_t* level_out, uint8_t* rates_out) {
    namespace mr = dgswe::multirate;
    mr::Assigner as;
    as.max_levels = max_levels;
    as.n_stages = n_stages;
    as.n1 = n1;
    as.line_rates = line_rates != 0;
    as.init(links, (size_t)K);
    std::vector<unsigned long long> slots(4);
    std::vector<uint8_t> dlevel((size_t)K), dll((size_t)K), memo_given((size_t)K), memo_ll((size_t)K);
    mr::Plan plan{};
    int out[2] = {0, 0};
    const unsigned egrid = (unsigned)std::min<int64_t>((K + 4095) / 4096, 128);
    const unsigned rgrid = (unsigned)std::min<int64_t>((K + 255) / 256, 1024);
    const bool lines = as.line_rates && n1 >= 2;
    for (int q = 0; q < calls; ++q) {
        const double* e = edt + (size_t)q * K;
        const uint8_t* f = flags + (size_t)q * K;
        emu_launch(1, 64, [&] { mr::mr_clear_kernel(slots.data(), out); });
        emu_launch(egrid, 256, [&] { mr::mr_extremes_kernel(e, f, K, slots.data()); });
        emu_launch(1, 64, [&] { mr::mr_plan_kernel(slots.data(), wet_dry, dt_cap[q], landing[q], max_levels, &plan); });
        const bool kept = as.memo;
        emu_launch(rgrid, 256, [&] {
            mr::mr_levels_kernel(e, f, K, &plan, lines ? 1 : 0, kept ? memo_given.data() : nullptr,
                                 kept ? memo_ll.data() : nullptr, dlevel.data(), dll.data(), out);
        });
        if (!as.device_kept(plan, out[0], out[1])) {
            as.level = dlevel;
            as.ll = dll;
            as.apply(as.finest);
            memo_given = as.memo_given;
            memo_ll = as.memo_ll;
        }
        dt_out[q] = plan.dt;
        stable_out[q] = plan.stable;
        finest_out[q] = as.finest;
        std::memcpy(level_out + (size_t)q * K, as.level.data(), (size_t)K);
        std::memcpy(rates_out + (size_t)q * n_stages * K, as.rates.data(), (size_t)n_stages * K);
    }
}

double emu_dt(int order, int64_t K, const double* state, const double* corners, const uint32_t* links,
              const uint8_t* affine, double cfl, double g, double h_min, double h_dry, double pcfl, double max_cfl,
              double max_cfl_sub, int positivity, double* element_dt) {
    size_t nn = (size_t)(order + 1) * (order + 1), nt = (size_t)K * nn;
    int epb = 64 / (int)nn;
    unsigned nblk = (unsigned)((K + epb - 1) / epb);
    // the form's static geometry (the time step reads the mass sums)
    StageParams sp{};
    std::memset(&sp, 0, sizeof(sp));
    std::vector<double> bed(nt, 0.0), geo, fgeo;
    sp.corners = corners; sp.bed = bed.data(); sp.affine = affine; sp.links = links;
    run_geometry(order, K, sp, geo, fgeo);
    std::vector<unsigned long long> slots(DT_SLOTS, DT_SLOT_EMPTY);
    DtParams p;
    p.h = state; p.hu = state + nt; p.hv = state + 2 * nt;
    p.corners = corners; p.emass = g_emass.data(); p.links = links; p.affine = affine; p.n_elements = K;
    p.cfl = cfl; p.g = g; p.h_min = h_min; p.h_dry = h_dry; p.pcfl = pcfl; p.max_cfl = max_cfl;
    p.max_cfl_sub = max_cfl_sub; p.positivity = positivity; p.element_dt = element_dt; p.slots = slots.data();
    p.xcd_stride = (nblk + 7) / 8;
    unsigned grid = 8 * p.xcd_stride;
    switch (order) {
        case 1: emu_launch(grid, 64, [&] { dt_kernel<1>(p); }); break;
        case 2: emu_launch(grid, 64, [&] { dt_kernel<2>(p); }); break;
        default: emu_launch(grid, 64, [&] { dt_kernel<3>(p); }); break;
    }
    double out = 0.0;
    emu_launch(1, DT_SLOTS, [&] { slots_min_kernel(slots.data(), &out); });
    for (unsigned long long v : slots)
        if (v != DT_SLOT_EMPTY) return -1.0;  // the reduction leaves the slots empty
    return out;
}

}  // extern "C"

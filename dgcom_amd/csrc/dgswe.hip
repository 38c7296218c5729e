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
#include <string>
#include <vector>

#include "../../include/dgswe.h"
#include "swe_kernels.hpp"

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

// Workgroups of the time-step kernel (a multiple of 8: XCDs)
constexpr unsigned int DT_GROUPS = 8 * 4096;
// Workgroups of the subcell kernel (a grid-stride loop over its list)
constexpr unsigned int SUB_GROUPS = 16384;

struct dg_state {
    double* d = nullptr;  // 3*K*NN: h, hu, hv
};

struct dg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int order = 0, nn = 0;
    int64_t K = 0;
    int flags = 0;
    double g = 9.81, h_min = 1e-6, h_dry = 1e-3, tau = 1.0, u_max = 20.0, f0 = 0.0;
    dg_tide tide{};
    double *corners = nullptr, *bed = nullptr, *manning = nullptr, *element_dt = nullptr;
    uint32_t* links = nullptr;
    uint8_t *tags = nullptr, *affine = nullptr;
    unsigned long long* clips = nullptr;
    uint32_t* queue = nullptr;
    uint8_t* queue_count = nullptr;
    double *block_min = nullptr, *dt_value = nullptr;
    uint32_t* sub_list = nullptr;
    unsigned int *chunk_total = nullptr, *sub_count = nullptr;
    int* flag = nullptr;
    dg_state* work[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double stage_ms = 0.0, dt_ms = 0.0;
    int64_t stage_n = 0, dt_n = 0;
};

#define DG_TRY(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            g_err = std::string(#expr) + ": " + hipGetErrorString(e_);                            \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

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

static double tide_elevation(const dg_tide& t, double time) {
    double tide = 0.0;
    for (int i = 0; i < t.n; ++i) tide += 1.0 * t.amplitude[i] * std::cos(t.omega[i] * time + t.phase[i] + 0.0);
    double rf = 1.0;
    if (t.ramp > 0.0) {
        double tau = std::min(std::max(time / t.ramp, 0.0), 1.0);
        rf = tau * tau * (3.0 - 2.0 * tau);
    }
    return t.mean + rf * tide;
}

static int free_statics(dg_ctx* c) {
    void* ptrs[] = {c->corners, c->bed, c->manning, c->element_dt, c->links, c->tags, c->affine, c->queue,
                    c->queue_count, c->block_min, c->sub_list, c->chunk_total};
    for (void* p : ptrs)
        if (p) DG_TRY(hipFree(p));
    c->corners = c->bed = c->manning = c->element_dt = nullptr;
    c->links = nullptr;
    c->tags = c->affine = nullptr;
    c->queue = nullptr;
    c->queue_count = nullptr;
    c->block_min = nullptr;
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

static int launch_stage(dg_ctx* c, const double* base, const double* input, double* out, double a, double b, double cc,
                        double time, int rhs_only, double* fstar_out = nullptr) {
    size_t nt = (size_t)c->K * c->nn;
    StageParams p;
    p.bh = base; p.bhu = base + nt; p.bhv = base + 2 * nt;
    p.ih = input; p.ihu = input + nt; p.ihv = input + 2 * nt;
    p.oh = out; p.ohu = out + nt; p.ohv = out + 2 * nt;
    p.bed = c->bed; p.corners = c->corners; p.manning = c->manning;
    p.links = c->links; p.tags = c->tags; p.affine = c->affine;
    p.n_elements = c->K;
    p.a = a; p.b = b; p.c = cc;
    p.eta_e = tide_elevation(c->tide, time);
    p.g = c->g; p.h_min = c->h_min; p.h_dry = c->h_dry; p.u_max = c->u_max; p.tau = c->tau; p.f0 = c->f0;
    p.rhs_only = rhs_only;
    p.flags = c->flags;
    p.clips = c->clips;
    p.fstar_out = fstar_out;
    p.queue = c->queue;
    p.queue_count = c->queue_count;
    const int epb = 64 / c->nn;
    const unsigned int nblk = (unsigned int)((c->K + epb - 1) / epb);
    const unsigned int nchunk = (nblk + 1023) / 1024;
    p.xcd_stride = (nblk + 7) / 8;
    dim3 grid(8 * p.xcd_stride), sub_grid(std::min(nblk, SUB_GROUPS));
    if (c->timing) DG_TRY(hipEventRecord(c->ev[0], c->stream));
    switch (c->order) {
        case 1: hipLaunchKernelGGL(wet_stage_kernel<1>, grid, dim3(64), 0, c->stream, p); break;
        case 2: hipLaunchKernelGGL(wet_stage_kernel<2>, grid, dim3(64), 0, c->stream, p); break;
        default: hipLaunchKernelGGL(wet_stage_kernel<3>, grid, dim3(64), 0, c->stream, p); break;
    }
    hipLaunchKernelGGL(chunk_count_kernel, dim3(nchunk), dim3(256), 0, c->stream, c->queue_count, nblk, c->chunk_total);
    hipLaunchKernelGGL(compact_kernel, dim3(nchunk), dim3(256), 0, c->stream, c->queue_count, c->queue, nblk,
                       c->chunk_total, epb, c->sub_list, c->sub_count);
    switch (c->order) {
        case 1: hipLaunchKernelGGL(subcell_stage_kernel<1>, sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
        case 2: hipLaunchKernelGGL(subcell_stage_kernel<2>, sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
        default: hipLaunchKernelGGL(subcell_stage_kernel<3>, sub_grid, dim3(64), 0, c->stream, p, c->sub_list, c->sub_count); break;
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

static int launch_dt(dg_ctx* c, const double* state, double cfl, int scheme, bool element_out, double* dt) {
    size_t nt = (size_t)c->K * c->nn;
    DtParams p;
    p.h = state; p.hu = state + nt; p.hv = state + 2 * nt;
    p.corners = c->corners; p.links = c->links; p.affine = c->affine;
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
    p.block_min = c->block_min;
    const unsigned int groups = DT_GROUPS;
    p.xcd_stride = 0;
    dim3 grid(groups);
    if (c->timing) DG_TRY(hipEventRecord(c->ev[0], c->stream));
    switch (c->order) {
        case 1: hipLaunchKernelGGL(dt_kernel<1>, grid, dim3(64), 0, c->stream, p); break;
        case 2: hipLaunchKernelGGL(dt_kernel<2>, grid, dim3(64), 0, c->stream, p); break;
        default: hipLaunchKernelGGL(dt_kernel<3>, grid, dim3(64), 0, c->stream, p); break;
    }
    hipLaunchKernelGGL(min_kernel, dim3(1), dim3(1024), 0, c->stream, c->block_min, (size_t)groups, c->dt_value);
    DG_TRY(hipGetLastError());
    if (c->timing) DG_TRY(hipEventRecord(c->ev[1], c->stream));
    double value = 0.0;
    DG_TRY(hipMemcpyAsync(&value, c->dt_value, 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
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

static int step_impl(dg_ctx* c, dg_state* s, double t, double dt, int scheme) {
    // Shu-Osher stages (integrator.rs:239-263); stage values in their own buffers,
    // the last one the new state (multirate.rs:1285-1311 with one level)
    static const double rk3[3][4] = {{1.0, 0.0, 1.0, 0.0}, {0.75, 0.25, 0.25, 1.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 0.5}};
    static const double rk43[4][4] = {{1.0, 0.0, 0.5, 0.0}, {0.0, 1.0, 0.5, 0.5}, {2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 1.0},
                                      {0.0, 1.0, 0.5, 0.5}};
    if (int r = ensure_work(c)) return r;
    int ns = scheme == DG_SSP_RK3 ? 3 : 4;
    const double(*st)[4] = scheme == DG_SSP_RK3 ? rk3 : rk43;
    const double* input = s->d;
    for (int q = 0; q < ns; ++q) {
        double* out = c->work[q]->d;
        if (int r = launch_stage(c, s->d, input, out, st[q][0], st[q][1], st[q][2] * dt, t + st[q][3] * dt, 0)) return r;
        input = out;
    }
    std::swap(s->d, c->work[ns - 1]->d);
    return 0;
}

extern "C" {

const char* dg_last_error(void) { return g_err.c_str(); }

int dg_create(int device, dg_ctx** out) {
    DG_REQUIRE(out != nullptr, "dg_create: null output");
    DG_TRY(hipSetDevice(device));
    dg_ctx* c = new dg_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; g_err = hipGetErrorString(e); return 1; }
    static bool tables = false;
    if (!tables) {
        GllTables t[4];
        std::memset(&t[0], 0, sizeof(GllTables));
        for (int n = 1; n <= 3; ++n) t[n] = make_tables(n);
        DG_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_gll), t, sizeof(t)));
        tables = true;
    }
    DG_TRY(hipMalloc(&c->clips, sizeof(unsigned long long)));
    DG_TRY(hipMalloc(&c->dt_value, sizeof(double)));
    DG_TRY(hipMalloc(&c->sub_count, sizeof(unsigned int)));
    DG_TRY(hipMalloc(&c->flag, sizeof(int)));
    DG_TRY(hipMemset(c->clips, 0, sizeof(unsigned long long)));
    DG_TRY(hipEventCreate(&c->ev[0]));
    DG_TRY(hipEventCreate(&c->ev[1]));
    *out = c;
    return 0;
}

int dg_destroy(dg_ctx* c) {
    if (!c) return 0;
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    if (int r = free_statics(c)) return r;
    DG_TRY(hipFree(c->clips));
    DG_TRY(hipFree(c->dt_value));
    DG_TRY(hipFree(c->sub_count));
    DG_TRY(hipFree(c->flag));
    DG_TRY(hipEventDestroy(c->ev[0]));
    DG_TRY(hipEventDestroy(c->ev[1]));
    DG_TRY(hipStreamDestroy(c->stream));
    delete c;
    return 0;
}

int dg_upload_form(dg_ctx* c, const dg_form* f) {
    DG_REQUIRE(c && f, "dg_upload_form: null argument");
    DG_REQUIRE(f->order >= 1 && f->order <= 3, "dg_upload_form: order must be 1, 2 or 3");
    DG_REQUIRE(f->n_elements > 0 && f->n_elements < (int64_t(1) << 28) - 1, "dg_upload_form: element count out of range");
    DG_REQUIRE(f->corners && f->face_links && f->face_tags && f->bed, "dg_upload_form: missing mesh, tags or bed");
    DG_TRY(hipSetDevice(c->device));
    if (int r = free_statics(c)) return r;
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
        DG_TRY(hipMalloc(&c->block_min, DT_GROUPS * sizeof(double)));
        DG_TRY(hipMalloc(&c->sub_list, K * sizeof(uint32_t)));
        DG_TRY(hipMalloc(&c->chunk_total, ((nblk + 1023) / 1024) * sizeof(unsigned int)));
    }
    DG_TRY(hipMemcpy(c->corners, f->corners, 8 * K * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->bed, f->bed, (size_t)K * nn * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->manning, man.data(), K * sizeof(double), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->links, f->face_links, 4 * K * sizeof(uint32_t), hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->tags, f->face_tags, 4 * K, hipMemcpyHostToDevice));
    DG_TRY(hipMemcpy(c->affine, aff.data(), K, hipMemcpyHostToDevice));
    return 0;
}

int dg_set_tide(dg_ctx* c, const dg_tide* t) {
    DG_REQUIRE(c && t, "dg_set_tide: null argument");
    DG_REQUIRE(t->n >= 0 && t->n <= 8, "dg_set_tide: at most 8 constituents");
    c->tide = *t;
    return 0;
}

int dg_state_alloc(dg_ctx* c, dg_state** s) {
    DG_REQUIRE(c && s && c->K > 0, "dg_state_alloc: upload a form first");
    DG_TRY(hipSetDevice(c->device));
    dg_state* st = new dg_state;
    hipError_t e = hipMalloc(&st->d, (size_t)3 * c->K * c->nn * sizeof(double));
    if (e != hipSuccess) { delete st; g_err = hipGetErrorString(e); return 1; }
    *s = st;
    return 0;
}

int dg_state_free(dg_ctx* c, dg_state* s) {
    if (!s) return 0;
    DG_TRY(hipSetDevice(c->device));
    DG_TRY(hipStreamSynchronize(c->stream));
    DG_TRY(hipFree(s->d));
    delete s;
    return 0;
}

int dg_state_upload(dg_ctx* c, dg_state* s, const double* h, const double* hu, const double* hv) {
    DG_REQUIRE(c && s && h && hu && hv, "dg_state_upload: null argument");
    size_t nt = (size_t)c->K * c->nn;
    DG_TRY(hipMemcpyAsync(s->d, h, nt * 8, hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipMemcpyAsync(s->d + nt, hu, nt * 8, hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipMemcpyAsync(s->d + 2 * nt, hv, nt * 8, hipMemcpyHostToDevice, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_state_download(dg_ctx* c, const dg_state* s, double* h, double* hu, double* hv) {
    DG_REQUIRE(c && s && h && hu && hv, "dg_state_download: null argument");
    size_t nt = (size_t)c->K * c->nn;
    DG_TRY(hipMemcpyAsync(h, s->d, nt * 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipMemcpyAsync(hu, s->d + nt, nt * 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipMemcpyAsync(hv, s->d + 2 * nt, nt * 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_state_copy(dg_ctx* c, dg_state* dst, const dg_state* src) {
    DG_REQUIRE(c && dst && src, "dg_state_copy: null argument");
    DG_TRY(hipMemcpyAsync(dst->d, src->d, (size_t)3 * c->K * c->nn * 8, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

int dg_rhs(dg_ctx* c, const dg_state* in, double t, dg_state* out) {
    DG_REQUIRE(c && in && out && c->K > 0, "dg_rhs: null argument or no form");
    DG_REQUIRE(in != out, "dg_rhs: in and out must differ");
    DG_TRY(hipSetDevice(c->device));
    if (int r = launch_stage(c, in->d, in->d, out->d, 0.0, 0.0, 0.0, t, 1)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_face_fluxes(dg_ctx* c, const dg_state* in, double t, double* out) {
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
    DG_REQUIRE(c && base && input && out && st && c->K > 0, "dg_stage: null argument or no form");
    DG_REQUIRE(out != input && out != base, "dg_stage: out must differ from base and input");
    DG_TRY(hipSetDevice(c->device));
    if (int r = launch_stage(c, base->d, input->d, out->d, st->a, st->b, st->c, st->time, 0)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_element_dt(dg_ctx* c, const dg_state* s, double cfl, int scheme, double* out) {
    DG_REQUIRE(c && s && out && c->K > 0, "dg_element_dt: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    double dt;
    if (int r = launch_dt(c, s->d, cfl, scheme, true, &dt)) return r;
    DG_TRY(hipMemcpy(out, c->element_dt, c->K * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int dg_dt_ssp(dg_ctx* c, const dg_state* s, double cfl, int scheme, double* dt) {
    DG_REQUIRE(c && s && dt && c->K > 0, "dg_dt_ssp: null argument or no form");
    DG_TRY(hipSetDevice(c->device));
    return launch_dt(c, s->d, cfl, scheme, false, dt);
}

int dg_step(dg_ctx* c, dg_state* s, double t, double dt, int scheme) {
    DG_REQUIRE(c && s && c->K > 0, "dg_step: null argument or no form");
    DG_REQUIRE(scheme == DG_SSP_RK3 || scheme == DG_SSP_RK43, "dg_step: unknown scheme");
    DG_TRY(hipSetDevice(c->device));
    if (int r = step_impl(c, s, t, dt, scheme)) return r;
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_run(dg_ctx* c, dg_state* s, double t, int n_steps, double cfl, int scheme, double* t_end, double* dts) {
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

int dg_is_finite(dg_ctx* c, const dg_state* s, int* finite) {
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
    DG_REQUIRE(c && clips, "dg_clip_count: null argument");
    unsigned long long v = 0;
    DG_TRY(hipMemcpyAsync(&v, c->clips, 8, hipMemcpyDeviceToHost, c->stream));
    DG_TRY(hipStreamSynchronize(c->stream));
    *clips = v;
    return 0;
}

int dg_masks(dg_ctx* c, const dg_state* s, uint8_t* node_dry, uint8_t* elem_subcell) {
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

int dg_timing_enable(dg_ctx* c, int on) {
    DG_REQUIRE(c, "dg_timing_enable: null context");
    c->timing = on != 0;
    c->stage_ms = c->dt_ms = 0.0;
    c->stage_n = c->dt_n = 0;
    return 0;
}

int dg_timing_read(dg_ctx* c, double* stage_ms, int64_t* stage_n, double* dt_ms, int64_t* dt_n) {
    DG_REQUIRE(c, "dg_timing_read: null context");
    if (stage_ms) *stage_ms = c->stage_ms;
    if (stage_n) *stage_n = c->stage_n;
    if (dt_ms) *dt_ms = c->dt_ms;
    if (dt_n) *dt_n = c->dt_n;
    return 0;
}

int dg_sync(dg_ctx* c) {
    DG_REQUIRE(c, "dg_sync: null context");
    DG_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_copy_bench(dg_ctx* c, size_t bytes, int reps, double* ms_per_copy) {
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

}  // extern "C"

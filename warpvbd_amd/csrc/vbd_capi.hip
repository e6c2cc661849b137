// vbd_capi.hip -- extern "C" entry points declared in include/warpvbd.h.
#include <algorithm>
#include <cstring>
#include <string>

#include "vbd_internal.h"

namespace vbd {
namespace {
thread_local std::string g_err;
thread_local const char *g_err_literal = nullptr;
}
void set_error(const std::string &msg)
{
    g_err = msg;
    g_err_literal = nullptr;
}
void set_error_literal(const char *msg) { g_err_literal = msg; }
}  // namespace vbd

using namespace vbd;

#define CHECK_ARG(cond, msg)                                                                   \
    do {                                                                                       \
        if (!(cond)) {                                                                         \
            set_error(msg);                                                                    \
            return VBD_EINVAL;                                                                 \
        }                                                                                      \
    } while (0)

extern "C" {

const char *vbd_last_error(void) { return g_err_literal ? g_err_literal : g_err.c_str(); }

void vbd_destroy(vbd_solver_t *s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    release_batch(s);
    void *dev[] = {s->perm, s->iperm, s->X,      s->u,       s->uprev, s->vel,  s->tgt,    s->mass,
                   s->active, s->slot_off, s->slot_ent, s->slots, s->slots2, s->colors, s->inc_off, s->inc, s->part, s->ctrl,
                   s->xio,  s->d_col_start, s->d_ranges, s->d_grav, s->d_body_out, s->body_new, s->tets_dev, s->dict,
                   s->st_nb, s->st_rec, s->st_vrec, s->small_rec, s->u_last, s->d_dx_max,
                   s->tail_v, s->tail_rec, s->tail_nb, s->st_dl, s->dict_ext};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    for (double4 *p : s->snap)
        if (p) (void)hipFree(p);
    if (s->ctrl_host) (void)hipHostFree(s->ctrl_host);
    if (s->xio_host) (void)hipHostFree(s->xio_host);
    if (s->body_out_host) (void)hipHostFree(s->body_out_host);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    if (s->ev_in) (void)hipEventDestroy(s->ev_in);
    if (s->ev_out) (void)hipEventDestroy(s->ev_out);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int vbd_create(const double *rest_xyz, int32_t n_vertices, const int32_t *tets, int32_t n_tets,
               const double *density, const double *youngs, const double *poisson, const uint8_t *active,
               const double *gravity, double damping, int32_t material, vbd_solver_t **out)
{
    CHECK_ARG(out, "out is NULL");
    *out = nullptr;
    CHECK_ARG(n_vertices > 0 && n_tets >= 0, "n_vertices must be > 0 and n_tets >= 0");
    CHECK_ARG(rest_xyz && gravity, "rest_xyz and gravity are required");
    CHECK_ARG(n_tets == 0 || (tets && density && youngs && poisson), "element arrays are required");
    CHECK_ARG(material == VBD_MATERIAL_SNH || material == VBD_MATERIAL_STVK, "unknown material");
    CHECK_ARG((int64_t)n_tets * 4 < (int64_t)1 << 31, "too many elements");
    for (int64_t i = 0; i < (int64_t)4 * n_tets; ++i) {
        if (tets[i] < 0 || tets[i] >= n_vertices) {
            set_error("element " + std::to_string(i / 4) + " references vertex " + std::to_string(tets[i]) +
                      " outside [0, " + std::to_string(n_vertices) + ")");
            return VBD_EMESH;
        }
    }
    vbd_solver *s = new vbd_solver();
    s->n = n_vertices;
    s->m = n_tets;
    s->material = material;
    s->damping = damping;
    for (int k = 0; k < 3; ++k) s->grav[k] = gravity[k];
    int r = 0;
    if (hipGetDevice(&s->device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("no HIP device available");
        delete s;
        return VBD_EHIP;
    }
    s->uniform_E = n_tets ? youngs[0] : 0.0;
    s->uniform_nu = n_tets ? poisson[0] : 0.0;
    for (int e = 1; e < n_tets; ++e)
        if (youngs[e] != youngs[0] || poisson[e] != poisson[0]) {
            s->uniform_E = s->uniform_nu = NAN;
            break;
        }
    r = setup_mesh(s, rest_xyz, tets, density, youngs, poisson, active, nullptr, 1, nullptr);
    if (r == 0) r = build_staging(s);
    if (r == 0) plan_segments(s);  // the big colours' launches take their lane layouts
    if (r != 0) {
        vbd_destroy(s);
        return r;
    }
    *out = s;
    return 0;
}

int vbd_batch_create(const double *rest_xyz, int32_t n_vertices, const int32_t *tets, int32_t n_tets,
                     const int32_t *body_vertex_offsets, int32_t n_bodies, const double *density,
                     const double *youngs, const double *poisson, const uint8_t *active, const double *gravity,
                     double damping, int32_t material, vbd_solver_t **out)
{
    CHECK_ARG(out, "out is NULL");
    *out = nullptr;
    CHECK_ARG(n_vertices > 0 && n_tets >= 0 && n_bodies >= 1, "bad sizes");
    CHECK_ARG(rest_xyz && gravity && body_vertex_offsets, "rest_xyz, gravity and body offsets are required");
    CHECK_ARG(n_tets == 0 || (tets && density && youngs && poisson), "element arrays are required");
    CHECK_ARG(material == VBD_MATERIAL_SNH || material == VBD_MATERIAL_STVK, "unknown material");
    CHECK_ARG((int64_t)(kMaxColors + 1) * n_bodies < ((int64_t)1 << 31), "too many bodies");
    CHECK_ARG(body_vertex_offsets[0] == 0 && body_vertex_offsets[n_bodies] == n_vertices,
              "body offsets must start at 0 and end at n_vertices");
    std::vector<int32_t> body_of(n_vertices);
    for (int b = 0; b < n_bodies; ++b) {
        CHECK_ARG(body_vertex_offsets[b] <= body_vertex_offsets[b + 1], "body offsets must not decrease");
        for (int v = body_vertex_offsets[b]; v < body_vertex_offsets[b + 1]; ++v) body_of[v] = b;
    }
    for (int64_t e = 0; e < n_tets; ++e) {
        int b0 = -1;
        for (int k = 0; k < 4; ++k) {
            const int v = tets[4 * e + k];
            if (v < 0 || v >= n_vertices) {
                set_error("element " + std::to_string(e) + " references a vertex outside [0, n_vertices)");
                return VBD_EMESH;
            }
            if (k == 0) b0 = body_of[v];
            else if (body_of[v] != b0) {
                set_error("element " + std::to_string(e) + " joins vertices of two bodies");
                return VBD_EMESH;
            }
        }
    }
    vbd_solver *s = new vbd_solver();
    s->n = n_vertices;
    s->m = n_tets;
    s->material = material;
    s->damping = damping;
    s->batched = true;
    for (int k = 0; k < 3; ++k) s->grav[k] = gravity[k];
    if (hipGetDevice(&s->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("no HIP device available");
        delete s;
        return VBD_EHIP;
    }
    int r = setup_mesh(s, rest_xyz, tets, density, youngs, poisson, active, body_vertex_offsets, n_bodies, gravity);
    if (r == 0) r = build_staging(s);
    if (r != 0) {
        vbd_destroy(s);
        return r;
    }
    *out = s;
    return 0;
}

int vbd_batch_step(vbd_solver_t *s, double dt, double tol, int32_t max_iter, int32_t *iters_out, double *metric_out)
{
    CHECK_ARG(s && s->batched, "not a batched handle");
    CHECK_ARG(dt > 0.0, "dt must be > 0");
    CHECK_ARG(max_iter >= 1, "max_iter must be >= 1");
    VBD_HIP(hipSetDevice(s->device));
    VBD_TRY(step_bodies(s, dt, tol, max_iter));
    for (int b = 0; b < s->n_bodies; ++b) {
        if (iters_out) iters_out[b] = s->body_out_host[b].x;
        if (metric_out) metric_out[b] = (double)__builtin_bit_cast(float, (uint32_t)s->body_out_host[b].y);
    }
    return 0;
}

int vbd_batch_results(vbd_solver_t *s, int32_t *iters_out, double *metric_out)
{
    CHECK_ARG(s && s->batched, "not a batched handle");
    for (int b = 0; b < s->n_bodies; ++b) {
        if (iters_out) iters_out[b] = s->body_out_host[b].x;
        if (metric_out) metric_out[b] = (double)__builtin_bit_cast(float, (uint32_t)s->body_out_host[b].y);
    }
    return 0;
}

int vbd_batch_set_mode(vbd_solver_t *s, int32_t mode, int32_t lanes, int32_t block, int32_t use_graph)
{
    CHECK_ARG(s && s->batched, "not a batched handle");
    CHECK_ARG(mode == VBD_BATCH_LAUNCHES || mode == VBD_BATCH_RESIDENT || mode == VBD_BATCH_LAUNCHES_FULL,
              "unknown batch mode");
    CHECK_ARG(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32,
              "lanes_per_vertex must be 1, 2, 4, 8, 16 or 32");
    CHECK_ARG(block >= 64 && block <= 1024 && block % 64 == 0, "block_size must be a multiple of 64 in [64, 1024]");
    s->batch_mode = mode == VBD_BATCH_RESIDENT ? kBatchResident : kBatchLaunches;
    s->launch.compact = mode != VBD_BATCH_LAUNCHES_FULL;
    s->launch.lanes = lanes;
    if (mode == VBD_BATCH_RESIDENT) s->launch.resident_block = block;
    else s->launch.block = block;
    s->launch.graph = use_graph != 0;
    return 0;
}

int vbd_batch_get_mode(vbd_solver_t *s, int32_t *mode, int32_t *lanes, int32_t *block, int32_t *use_graph)
{
    CHECK_ARG(s && s->batched, "not a batched handle");
    if (mode)
        *mode = s->batch_mode == kBatchResident ? VBD_BATCH_RESIDENT
                : s->launch.compact                ? VBD_BATCH_LAUNCHES
                                                   : VBD_BATCH_LAUNCHES_FULL;
    if (lanes) *lanes = s->launch.lanes;
    if (block) *block = s->batch_mode == VBD_BATCH_RESIDENT ? s->launch.resident_block : s->launch.block;
    if (use_graph) *use_graph = s->launch.graph ? 1 : 0;
    return 0;
}

int vbd_batch_autotune(vbd_solver_t *s, double dt, double tol, int32_t max_iter, int32_t trial_steps,
                       double *best_ms_per_step)
{
    CHECK_ARG(s && s->batched, "not a batched handle");
    CHECK_ARG(dt > 0.0 && max_iter >= 1 && trial_steps >= 1, "bad arguments");
    VBD_HIP(hipSetDevice(s->device));
    return autotune_bodies(s, dt, tol, max_iter, trial_steps, best_ms_per_step);
}

int vbd_set_uniform_material(vbd_solver_t *s, double youngs, double poisson)
{
    CHECK_ARG(s, "null handle");
    CHECK_ARG(youngs > 0.0 && poisson > -1.0 && poisson < 0.5, "need E > 0 and -1 < nu < 0.5");
    CHECK_ARG(s->uniform_E == s->uniform_E, "the solver's elements do not share one material");
    VBD_HIP(hipSetDevice(s->device));
    VBD_TRY(rescale_material(s, s->uniform_E, s->uniform_nu, youngs, poisson));
    s->uniform_E = youngs;
    s->uniform_nu = poisson;
    ++s->edits;
    return 0;
}

int vbd_set_gravity(vbd_solver_t *s, const double *gravity)
{
    CHECK_ARG(s && gravity, "null argument");
    CHECK_ARG(!s->batched, "batched handles take per-body gravity at create");
    VBD_HIP(hipSetDevice(s->device));
    for (int k = 0; k < 3; ++k) s->grav[k] = gravity[k];
    const float4 g = make_float4((float)gravity[0], (float)gravity[1], (float)gravity[2], 0.f);
    VBD_HIP(hipMemcpyAsync(s->d_grav, &g, sizeof(float4), hipMemcpyHostToDevice, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    ++s->edits;
    return 0;
}

int vbd_reset(vbd_solver_t *s)
{
    CHECK_ARG(s, "null handle");
    VBD_HIP(hipSetDevice(s->device));
    VBD_HIP(hipMemsetAsync(s->uprev, 0, sizeof(double4) * s->n, s->stream));
    VBD_HIP(hipMemsetAsync(s->vel, 0, sizeof(float4) * s->n, s->stream));
    VBD_HIP(hipMemsetAsync(s->u, 0, sizeof(float4) * s->n, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    s->last_iters = s->prev_iters = 0;
    s->hist_valid = false;
    s->step_serial += 2;
    ++s->edits;
    return 0;
}

int vbd_step(vbd_solver_t *s, const double *positions, double dt, double tol, int32_t max_iter, double *out_xyz,
             int32_t *iters_out, double *metric_out)
{
    CHECK_ARG(s, "null handle");
    CHECK_ARG(dt > 0.0, "dt must be > 0");
    CHECK_ARG(max_iter >= 1, "max_iter must be >= 1");
    VBD_HIP(hipSetDevice(s->device));
    if (positions) {
        VBD_TRY(import_positions(s, positions, s->uprev));
        ++s->edits;
    }
    int it = 0;
    double met = 0.0;
    if (s->batched) {
        VBD_TRY(step_bodies(s, dt, tol, max_iter));
        float mx = 0.f;
        for (int b = 0; b < s->n_bodies; ++b) {
            it = std::max(it, s->body_out_host[b].x);
            const float m = __builtin_bit_cast(float, (uint32_t)s->body_out_host[b].y);
            if (!(m <= mx)) mx = m;  // NaN wins
        }
        met = mx;
    } else {
        VBD_TRY(step_device(s, dt, tol, max_iter, &it, &met));
    }
    if (out_xyz) VBD_TRY(export_positions(s, s->u, out_xyz));
    if (iters_out) *iters_out = it;
    if (metric_out) *metric_out = met;
    return 0;
}

int vbd_step_dev(vbd_solver_t *s, const double *positions_dev, double dt, double tol, int32_t max_iter,
                 double *out_dev, void *stream, int32_t *iters_out, double *metric_out)
{
    CHECK_ARG(s, "null handle");
    CHECK_ARG(dt > 0.0, "dt must be > 0");
    CHECK_ARG(max_iter >= 1, "max_iter must be >= 1");
    VBD_HIP(hipSetDevice(s->device));
    hipStream_t cs = (hipStream_t)stream;
    if (!s->ev_in) VBD_HIP(hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming));
    if (!s->ev_out) VBD_HIP(hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming));
    // the step runs after what the caller queued on its stream ...
    VBD_HIP(hipEventRecord(s->ev_in, cs));
    VBD_HIP(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    if (positions_dev) {
        VBD_TRY(import_positions_dev(s, positions_dev, s->uprev));
        ++s->edits;
    }
    int it = 0;
    double met = 0.0;
    if (s->batched) {
        VBD_TRY(step_bodies(s, dt, tol, max_iter));
        float mx = 0.f;
        for (int b = 0; b < s->n_bodies; ++b) {
            it = std::max(it, s->body_out_host[b].x);
            const float m = __builtin_bit_cast(float, (uint32_t)s->body_out_host[b].y);
            if (!(m <= mx)) mx = m;
        }
        met = mx;
    } else {
        VBD_TRY(step_device(s, dt, tol, max_iter, &it, &met));
    }
    if (out_dev) VBD_TRY(export_positions_dev(s, s->u, out_dev));
    // ... and what the caller queues next runs after the result is written
    VBD_HIP(hipEventRecord(s->ev_out, s->stream));
    VBD_HIP(hipStreamWaitEvent(cs, s->ev_out, 0));
    if (iters_out) *iters_out = it;
    if (metric_out) *metric_out = met;
    return 0;
}

int vbd_get_state(vbd_solver_t *s, double *x_prev, double *v_prev)
{
    CHECK_ARG(s, "null handle");
    VBD_HIP(hipSetDevice(s->device));
    if (x_prev) VBD_TRY(export_positions(s, s->uprev, x_prev));
    if (v_prev) VBD_TRY(export_velocities(s, v_prev));
    return 0;
}

int vbd_set_state(vbd_solver_t *s, const double *x_prev, const double *v_prev)
{
    CHECK_ARG(s, "null handle");
    VBD_HIP(hipSetDevice(s->device));
    if (x_prev) VBD_TRY(import_positions(s, x_prev, s->uprev));
    if (v_prev) VBD_TRY(import_velocities(s, v_prev));
    if (x_prev || v_prev) ++s->edits;
    return 0;
}

int vbd_snapshot(vbd_solver_t *s, int32_t slot)
{
    CHECK_ARG(s, "null handle");
    CHECK_ARG(slot >= 0 && slot < VBD_MAX_SNAPSHOTS, "snapshot slot out of range");
    VBD_HIP(hipSetDevice(s->device));
    if (!s->snap[slot]) VBD_HIP(hipMalloc((void **)&s->snap[slot], sizeof(double4) * s->n));
    VBD_HIP(hipMemcpyAsync(s->snap[slot], s->uprev, sizeof(double4) * s->n, hipMemcpyDeviceToDevice, s->stream));
    return 0;
}

int vbd_read_snapshot(vbd_solver_t *s, int32_t slot, double *out_xyz)
{
    CHECK_ARG(s && out_xyz, "null argument");
    CHECK_ARG(slot >= 0 && slot < VBD_MAX_SNAPSHOTS && s->snap[slot], "empty snapshot slot");
    VBD_HIP(hipSetDevice(s->device));
    return export_positions(s, s->snap[slot], out_xyz);
}

int vbd_load_snapshot(vbd_solver_t *s, int32_t slot)
{
    CHECK_ARG(s, "null handle");
    CHECK_ARG(slot >= 0 && slot < VBD_MAX_SNAPSHOTS && s->snap[slot], "empty snapshot slot");
    VBD_HIP(hipSetDevice(s->device));
    VBD_HIP(hipMemcpyAsync(s->uprev, s->snap[slot], sizeof(double4) * s->n, hipMemcpyDeviceToDevice, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    s->hist_valid = false;
    s->step_serial += 2;
    ++s->edits;
    return 0;
}

int vbd_info(vbd_solver_t *s, int32_t *n_vertices, int32_t *n_tets, int32_t *n_colors, int32_t *max_incident,
             int32_t *max_degree, int64_t *n_slots)
{
    CHECK_ARG(s, "null handle");
    if (n_vertices) *n_vertices = s->n;
    if (n_tets) *n_tets = s->m;
    if (n_colors) *n_colors = s->ncolors;
    if (max_incident) *max_incident = s->max_incident;
    if (max_degree) *max_degree = s->max_degree;
    if (n_slots) *n_slots = s->n_slots;
    return 0;
}

int vbd_get_colors(vbd_solver_t *s, int32_t *colors)
{
    CHECK_ARG(s && colors, "null argument");
    VBD_HIP(hipSetDevice(s->device));
    VBD_HIP(hipMemcpyAsync(colors, s->colors, sizeof(int32_t) * s->n, hipMemcpyDeviceToHost, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

int vbd_get_color_sizes(vbd_solver_t *s, int32_t *sizes)
{
    CHECK_ARG(s && sizes, "null argument");
    for (int c = 0; c < s->ncolors; ++c) sizes[c] = s->col_size[c];
    return 0;
}

int vbd_get_incidence(vbd_solver_t *s, int32_t *adj)
{
    CHECK_ARG(s && adj, "null argument");
    VBD_HIP(hipSetDevice(s->device));
    std::vector<int32_t> off(s->n + 1), inc(std::max(1, 4 * s->m));
    VBD_HIP(hipMemcpyAsync(off.data(), s->inc_off, sizeof(int32_t) * (s->n + 1), hipMemcpyDeviceToHost, s->stream));
    if (s->m)
        VBD_HIP(hipMemcpyAsync(inc.data(), s->inc, sizeof(int32_t) * 4 * s->m, hipMemcpyDeviceToHost, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    const int W = s->max_incident;
    for (int v = 0; v < s->n; ++v) {
        int k = 0;
        for (int q = off[v]; q < off[v + 1]; ++q, ++k) adj[(size_t)v * W + k] = inc[q] >> 2;
        for (; k < W; ++k) adj[(size_t)v * W + k] = -1;
    }
    return 0;
}

int vbd_color(const int32_t *tets, int32_t n_tets, int32_t n_vertices, int32_t *colors, int32_t *n_colors_out,
              int32_t *max_degree_out)
{
    CHECK_ARG(tets && colors && n_vertices > 0 && n_tets >= 0, "bad arguments");
    for (int64_t i = 0; i < (int64_t)4 * n_tets; ++i) {
        if (tets[i] < 0 || tets[i] >= n_vertices) {
            set_error("element index outside [0, n_vertices)");
            return VBD_EMESH;
        }
    }
    hipStream_t st;
    VBD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int32_t *d_tets = nullptr, *d_col = nullptr, *d_off = nullptr, *d_inc = nullptr;
    int rc = 0, nc = 0, md = 0;
    auto cleanup = [&]() {
        if (d_tets) (void)hipFree(d_tets);
        if (d_col) (void)hipFree(d_col);
        if (d_off) (void)hipFree(d_off);
        if (d_inc) (void)hipFree(d_inc);
        (void)hipStreamDestroy(st);
    };
    if (hipMalloc((void **)&d_tets, sizeof(int32_t) * std::max(1, 4 * n_tets)) != hipSuccess ||
        hipMalloc((void **)&d_col, sizeof(int32_t) * n_vertices) != hipSuccess ||
        hipMalloc((void **)&d_off, sizeof(int32_t) * (n_vertices + 1)) != hipSuccess ||
        hipMalloc((void **)&d_inc, sizeof(int32_t) * std::max(1, 4 * n_tets)) != hipSuccess) {
        cleanup();
        set_error("hipMalloc failed");
        return VBD_ENOMEM;
    }
    if (n_tets && hipMemcpyAsync(d_tets, tets, sizeof(int32_t) * 4 * n_tets, hipMemcpyHostToDevice, st) != hipSuccess) {
        cleanup();
        set_error("hipMemcpy failed");
        return VBD_EHIP;
    }
    rc = color_tets_device(d_tets, n_tets, n_vertices, d_col, d_off, d_inc, &nc, &md, st);
    if (rc == 0) {
        if (hipMemcpyAsync(colors, d_col, sizeof(int32_t) * n_vertices, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("hipMemcpy failed");
            rc = VBD_EHIP;
        }
    }
    cleanup();
    if (rc == 0) {
        if (n_colors_out) *n_colors_out = nc;
        if (max_degree_out) *max_degree_out = md;
    }
    return rc;
}

int vbd_color_noalloc(const int32_t *tets_dev, int32_t n_tets, int32_t n_vertices, int32_t *colors_dev,
                      int32_t *n_colors_out, int32_t *max_degree_out, void *stream)
{
    if (!tets_dev || !colors_dev || n_vertices <= 0 || n_tets < 0 || n_tets > (1 << 29)) {
        set_error_literal("bad arguments");
        return VBD_EINVAL;
    }
    int nc = 0, md = 0;
    const int rc = color_tets_noalloc(tets_dev, n_tets, n_vertices, colors_dev, &nc, &md, (hipStream_t)stream);
    if (rc == 0) {
        if (n_colors_out) *n_colors_out = nc;
        if (max_degree_out) *max_degree_out = md;
    }
    return rc;
}

int vbd_color_graph(const int32_t *adj_off, const int32_t *adj, int32_t n_vertices, int32_t *colors,
                    int32_t *n_colors_out)
{
    CHECK_ARG(adj_off && colors && n_vertices > 0, "bad arguments");
    const int64_t ne = adj_off[n_vertices];
    CHECK_ARG(ne >= 0 && (ne == 0 || adj), "bad adjacency offsets");
    for (int v = 0; v < n_vertices; ++v) CHECK_ARG(adj_off[v] <= adj_off[v + 1], "adjacency offsets must not decrease");
    hipStream_t st;
    VBD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int32_t *d_off = nullptr, *d_adj = nullptr, *d_col = nullptr;
    auto cleanup = [&]() {
        if (d_off) (void)hipFree(d_off);
        if (d_adj) (void)hipFree(d_adj);
        if (d_col) (void)hipFree(d_col);
        (void)hipStreamDestroy(st);
    };
    if (hipMalloc((void **)&d_off, sizeof(int32_t) * (n_vertices + 1)) != hipSuccess ||
        hipMalloc((void **)&d_adj, sizeof(int32_t) * std::max<int64_t>(1, ne)) != hipSuccess ||
        hipMalloc((void **)&d_col, sizeof(int32_t) * n_vertices) != hipSuccess) {
        cleanup();
        set_error("hipMalloc failed");
        return VBD_ENOMEM;
    }
    int rc = 0, nc = 0;
    if (hipMemcpyAsync(d_off, adj_off, sizeof(int32_t) * (n_vertices + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        (ne && hipMemcpyAsync(d_adj, adj, sizeof(int32_t) * ne, hipMemcpyHostToDevice, st) != hipSuccess)) {
        cleanup();
        set_error("hipMemcpy failed");
        return VBD_EHIP;
    }
    rc = color_graph_device(d_off, d_adj, n_vertices, d_col, &nc, st);
    if (rc == 0 && (hipMemcpyAsync(colors, d_col, sizeof(int32_t) * n_vertices, hipMemcpyDeviceToHost, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess)) {
        set_error("hipMemcpy failed");
        rc = VBD_EHIP;
    }
    cleanup();
    if (rc == 0 && n_colors_out) *n_colors_out = nc;
    return rc;
}

int vbd_use_device(int32_t device)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device available");
        return VBD_EHIP;
    }
    CHECK_ARG(device >= 0 && device < count, "device index out of range");
    VBD_HIP(hipSetDevice(device));
    return 0;
}

int vbd_set_graph(vbd_solver_t *s, int32_t use_graph)
{
    CHECK_ARG(s, "null handle");
    s->launch.graph = use_graph != 0;
    return 0;
}

int vbd_profile_enable(vbd_solver_t *s, int32_t enable, int32_t min_vertices)
{
    CHECK_ARG(s, "null handle");
    s->prof = enable != 0;
    s->prof_min = min_vertices;
    s->prof_ms = s->prof_bytes = 0.0;
    s->prof_launches = 0;
    return 0;
}

int vbd_last_dx_max(vbd_solver_t *s, double *dx_max)
{
    CHECK_ARG(s && dx_max, "null argument");
    *dx_max = s->last_dx_max;
    return 0;
}

int vbd_profile_read(vbd_solver_t *s, double *ms, int64_t *launches, double *bytes)
{
    CHECK_ARG(s, "null handle");
    if (ms) *ms = s->prof_ms;
    if (launches) *launches = s->prof_launches;
    if (bytes) *bytes = s->prof_bytes;
    s->prof_ms = s->prof_bytes = 0.0;
    s->prof_launches = 0;
    return 0;
}

void vbd_use_record_dictionary(int32_t enable) { g_record_dictionary = enable != 0; }
void vbd_use_neighbour_staging(int32_t enable) { g_staging = enable != 0; }
void vbd_use_tail_sweep(int32_t enable) { g_tail = enable != 0; }
void vbd_use_edge_form(int32_t enable) { g_edge_form = enable != 0; }

int vbd_edge_form(vbd_solver_t *s, int32_t *on)
{
    CHECK_ARG(s && on, "null argument");
    *on = s->dict_ext != nullptr;
    return 0;
}

int vbd_tail_sweep(vbd_solver_t *s, int32_t *n_colors, int32_t *n_clusters)
{
    CHECK_ARG(s && n_colors && n_clusters, "null argument");
    const bool on = !s->segments.empty() && s->segments.back().tail;
    *n_colors = on ? s->tail_k : 0;
    *n_clusters = on ? s->tail_clusters : 0;
    return 0;
}

int vbd_color_layout_bytes(vbd_solver_t *s, double *bytes)
{
    CHECK_ARG(s && bytes, "null argument");
#if VBD_TET_RECORDS
    const double per_slot = 20.0;
#else
    // one entry per slot: 48 bytes of constants (the rest-edge constants go
    // with shared entries only)
    const double per_slot = s->dict_shared ? 16.0 : 64.0;
#endif
    for (int c = 0; c < s->ncolors; ++c) {
        const double nv = (double)(s->col_start[c + 1] - s->col_start[c]);
        // packed position 16 + FP64 target 32 + slot offset 4 + write-back 16
        bytes[c] = 68.0 * nv + per_slot * s->col_slots[c] + 16.0 * s->col_boundary[c];
        // big-colour sweep with staging lists (the default launch): per vertex
        // an 8-byte word (first entry and slot count, record sequence) in place
        // of the slot offsets, the 4-byte records of the distinct sequences the
        // colour uses, and the blocks' lists of 4-byte entries (one per vertex
        // and neighbour)
        const LaneLayout *l = c < (int)s->st_col.size() ? &s->st_col[c] : nullptr;
        if (s->st_ok && l && l->ids >= 0 && l->b1 > 0 && s->launch.lanes == 4 && s->launch.block == 256)
            bytes[c] = 72.0 * nv + 4.0 * s->st_prec[c] + 16.0 * s->col_boundary[c] +
                       4.0 * (double)stage_list_offset(l->b1, l->b4);
    }
    return 0;
}

int vbd_staged_colors(vbd_solver_t *s, int32_t *n_colors)
{
    CHECK_ARG(s && n_colors, "null argument");
    int k = 0;
    if (s->st_ok)
        for (const LaneLayout &l : s->st_col) k += l.ids >= 0 && l.b1 > 0;
    *n_colors = k;
    return 0;
}

int vbd_records(vbd_solver_t *s, int32_t *n_dictionary, int32_t *bytes_per_slot)
{
    CHECK_ARG(s, "null handle");
#if VBD_TET_RECORDS
    if (n_dictionary) *n_dictionary = 0;
    if (bytes_per_slot) *bytes_per_slot = 20;  // 64-byte tet record shared by 4 slots + 4-byte entry
#else
    if (n_dictionary) *n_dictionary = s->n_dict;
    if (bytes_per_slot) *bytes_per_slot = s->dict_shared ? 16 : 64;
#endif
    return 0;
}

int vbd_color_bytes(vbd_solver_t *s, double *bytes)
{
    CHECK_ARG(s && bytes, "null argument");
    for (int c = 0; c < s->ncolors; ++c) bytes[c] = s->col_bytes[c];
    return 0;
}

int vbd_set_launch(vbd_solver_t *s, int32_t lanes, int32_t block, int32_t small_color, int32_t chunk)
{
    CHECK_ARG(s, "null handle");
    CHECK_ARG(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32,
              "lanes_per_vertex must be 1, 2, 4, 8, 16 or 32");
    CHECK_ARG(block >= 64 && block <= 1024 && block % 64 == 0, "block_size must be a multiple of 64 in [64, 1024]");
    CHECK_ARG(small_color >= 0 && chunk >= 0, "small_color and chunk must be >= 0");
    s->launch.lanes = lanes;
    s->launch.block = block;
    s->launch.small_color = small_color;
    s->launch.chunk = chunk;
    plan_segments(s);
    VBD_TRY(build_small_records(s));
    for (auto &g : s->graphs) {  // captured with the previous plan
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    s->graphs.clear();
    return 0;
}

}  // extern "C"

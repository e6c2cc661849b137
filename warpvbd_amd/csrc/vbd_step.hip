// vbd_step.hip -- the VBD time step on the device (src/vbd_solver.py:509-612).
//
//   k_step_init    x <- x_prev + dt v_prev + dt^2 g for active vertices (:277-292);
//                  also the inertia target and M_v/dt^2 used by every sweep
//   k_sweep        one colour: per vertex, a group of L lanes walks its slots,
//                  each lane evaluating one incident tet per visit; the group
//                  reduces gradient and Hessian by butterfly shuffles, adds
//                  inertia (and damping), solves the 3x3 block in registers and
//                  writes x += dx (:149-274 fused, no per-slot scratch)
//   k_sweep_fused  a run of consecutive small colours in one workgroup with a
//                  barrier between colours; optionally closes the iteration
//   k_check        max|dx_x+dx_y+dx_z| < tol  (:578) on the device
//   k_step_finish  v = (x - x_prev)/dt, x_prev = x (:609-610)
//
// The iteration index and the converged flag live in device memory (Ctrl), so
// the host enqueues chunks of iterations and waits once per chunk; launches
// after convergence return at their first instruction.
#include <algorithm>
#include <cmath>

#include "vbd_device.h"
#include "vbd_internal.h"

namespace vbd {

namespace {

struct SweepArgs {
    float4 *__restrict__ u;
    const float4 *__restrict__ tgt;
    const float4 *__restrict__ uprev;
    const int32_t *__restrict__ slot_off;
    const float4 *__restrict__ slots;
    uint32_t *part;  // this launch's per-block metric maxima
    Ctrl *ctrl;
    float kd_dt;  // damping_coefficient / dt
};

// Max over the block of the metric bits; thread 0 gets the result.  Avoids a
// device-scope atomic per wave: those serialise on one address across XCDs.
__device__ __forceinline__ uint32_t block_max_u32(uint32_t x)
{
    __shared__ uint32_t red[16];
    x = wave_max_u32(x);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = x;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < nw; ++i) x = x > red[i] ? x : red[i];
    return x;
}

// Max over np per-block partial maxima with the whole (<= 1024-thread) block;
// thread 0 gets the result.
__device__ __forceinline__ uint32_t reduce_partials(const uint32_t *part, int np, uint32_t x)
{
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        const uint32_t y = part[i];
        x = x > y ? x : y;
    }
    return block_max_u32(x);
}

__device__ __forceinline__ void close_iteration(Ctrl *c, int it, uint32_t bits, float tol)
{
    c->metric = bits;
    c->iters = it + 1;
    if (__uint_as_float(bits) < tol) c->converged = 1;
    c->it = it + 1;
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD one
// contiguous share of the colour so its L2 sees a compact neighbourhood.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int per = nb >> 3, rem = nb & 7;
    const int x = b & 7, i = b >> 3;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// Evaluate and apply one vertex update for a lane group.  Returns the metric
// bits (0 for lanes without a vertex).
template <int L, int B, int MAT>
__device__ __forceinline__ uint32_t update_vertex(const SweepArgs &a, int v, bool valid, int lane)
{
    Acc acc;
    acc_zero(acc);
    const float4 uv4 = a.u[v];
    const float3 us = make_float3(uv4.x, uv4.y, uv4.z);
    if (valid) {
        const int s0 = a.slot_off[v], s1 = a.slot_off[v + 1];
        // B slots per lane per batch: all records, then all gathers, then the
        // arithmetic, so each batch costs two memory latencies, not 2B.
        for (int s = s0 + lane; s < s1; s += B * L) {
            float4 q[B][4];
            float4 x[B][3];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int sj = s + j * L;
                if (sj < s1) {
                    const float4 *r = a.slots + 4 * (size_t)sj;
                    q[j][0] = r[0];
                    q[j][1] = r[1];
                    q[j][2] = r[2];
                    q[j][3] = r[3];
                }
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (s + j * L < s1) {
                    x[j][0] = a.u[__float_as_int(q[j][0].x)];
                    x[j][1] = a.u[__float_as_int(q[j][0].y)];
                    x[j][2] = a.u[__float_as_int(q[j][0].z)];
                }
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (s + j * L < s1) {
                    const float4 *qq = q[j];
                    const float4 *xx = x[j];
                    slot_eval<MAT>(qq[0].w, qq[1].w, qq[2].w, make_float3(qq[1].x, qq[1].y, qq[1].z),
                                   make_float3(qq[2].x, qq[2].y, qq[2].z), make_float3(qq[3].x, qq[3].y, qq[3].z),
                                   make_float3(xx[0].x - us.x, xx[0].y - us.y, xx[0].z - us.z),
                                   make_float3(xx[1].x - us.x, xx[1].y - us.y, xx[1].z - us.z),
                                   make_float3(xx[2].x - us.x, xx[2].y - us.y, xx[2].z - us.z), acc);
                }
            }
        }
    }
    acc_group_sum<L>(acc);
    if (!valid) return 0u;
    if (a.kd_dt != 0.f) {
        // damping (src/vbd_solver.py:222-227): g += kd/dt H_el (x - x_prev), H += kd/dt H_el
        const float4 up = a.uprev[v];
        const float e0 = us.x - up.x, e1 = us.y - up.y, e2 = us.z - up.z;
        const float k = a.kd_dt;
        acc.g0 += k * (acc.h00 * e0 + acc.h01 * e1 + acc.h02 * e2);
        acc.g1 += k * (acc.h01 * e0 + acc.h11 * e1 + acc.h12 * e2);
        acc.g2 += k * (acc.h02 * e0 + acc.h12 * e1 + acc.h22 * e2);
        const float f = 1.f + k;
        acc.h00 *= f; acc.h01 *= f; acc.h02 *= f; acc.h11 *= f; acc.h12 *= f; acc.h22 *= f;
    }
    // inertia + gravity (src/vbd_solver.py:15-47, summed over slots): M/dt^2 (x - t)
    const float4 tg = a.tgt[v];
    const float w = tg.w;
    acc.g0 += w * (us.x - tg.x);
    acc.g1 += w * (us.y - tg.y);
    acc.g2 += w * (us.z - tg.z);
    acc.h00 += w;
    acc.h11 += w;
    acc.h22 += w;
    const float3 dx = solve3_guarded(acc);
    if (lane == 0) a.u[v] = make_float4(us.x + dx.x, us.y + dx.y, us.z + dx.z, 0.f);
    return metric_bits(dx);
}

template <int L, int B, int MAT>
__global__ __launch_bounds__(1024) void k_sweep(SweepArgs a, int v0, int nv)
{
    const Ctrl *c = a.ctrl;
    if (c->converged) return;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int t = b * blockDim.x + threadIdx.x;
    const int vi = t / L, lane = t % L;
    const bool valid = vi < nv;
    const int v = v0 + (valid ? vi : 0);
    const uint32_t mb = block_max_u32(update_vertex<L, B, MAT>(a, v, valid, lane));
    if (threadIdx.x == 0) a.part[blockIdx.x] = mb;
}

// Colours [c0, c1) in one workgroup; col_start is the colour-major range table.
// close != 0: this launch ends the iteration -- it folds the np partial maxima
// of the iteration's earlier launches (at a.part - np_before) with its own and
// decides convergence (src/vbd_solver.py:578).
template <int L, int B, int MAT>
__global__ __launch_bounds__(1024) void k_sweep_fused(SweepArgs a, const int32_t *col_start, int c0, int c1,
                                                      int close, int np_before, float tol)
{
    Ctrl *c = a.ctrl;
    if (c->converged) return;
    const int it = c->it;
    const int groups = blockDim.x / L;
    const int gi = threadIdx.x / L, lane = threadIdx.x % L;
    uint32_t mb = 0;
    for (int col = c0; col < c1; ++col) {
        const int v0 = col_start[col], v1 = col_start[col + 1];
        for (int base = v0; base < v1; base += groups) {
            const int v = base + gi;
            const bool valid = v < v1;
            const uint32_t m = update_vertex<L, B, MAT>(a, valid ? v : v0, valid, lane);
            mb = mb > m ? mb : m;
        }
        __syncthreads();
    }
    if (close) {
        const uint32_t all = reduce_partials(a.part - np_before, np_before, mb);
        if (threadIdx.x == 0) close_iteration(c, it, all, tol);
    } else {
        const uint32_t all = block_max_u32(mb);
        if (threadIdx.x == 0) a.part[0] = all;
    }
}

__global__ __launch_bounds__(1024) void k_check(Ctrl *c, const uint32_t *part, int np, float tol)
{
    if (c->converged) return;
    const int it = c->it;
    const uint32_t all = reduce_partials(part, np, 0u);
    if (threadIdx.x == 0) close_iteration(c, it, all, tol);
}

__global__ void k_step_init(int n, float4 *u, const float4 *uprev, const float4 *vel, float4 *tgt, const float *mass,
                            const uint8_t *active, float dt, float gx, float gy, float gz, Ctrl *ctrl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctrl->it = 0;
        ctrl->converged = 0;
        ctrl->iters = 0;
        ctrl->metric = 0u;
    }
    if (i >= n) return;
    const float4 p = uprev[i];
    if (active[i]) {
        const float4 v = vel[i];
        const float dt2 = dt * dt;
        const float4 t = make_float4(p.x + dt * v.x + dt2 * gx, p.y + dt * v.y + dt2 * gy, p.z + dt * v.z + dt2 * gz,
                                     mass[i] / dt2);
        u[i] = make_float4(t.x, t.y, t.z, 0.f);
        tgt[i] = t;
    } else {
        u[i] = p;
        tgt[i] = make_float4(p.x, p.y, p.z, 0.f);
    }
}

__global__ void k_step_finish(int n, const float4 *u, float4 *uprev, float4 *vel, float inv_dt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = u[i], p = uprev[i];
    vel[i] = make_float4((x.x - p.x) * inv_dt, (x.y - p.y) * inv_dt, (x.z - p.z) * inv_dt, 0.f);
    uprev[i] = x;
}

// x (original order, f64) = X + u[iperm]
__global__ void k_export(int n, const float4 *src, const int32_t *iperm, const double *X, double *out)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float4 d = src[iperm[v]];
    out[3 * v] = X[3 * v] + (double)d.x;
    out[3 * v + 1] = X[3 * v + 1] + (double)d.y;
    out[3 * v + 2] = X[3 * v + 2] + (double)d.z;
}

__global__ void k_import(int n, const double *in, const int32_t *perm, const double *X, float4 *dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = perm[i];
    dst[i] = make_float4((float)(in[3 * v] - X[3 * v]), (float)(in[3 * v + 1] - X[3 * v + 1]),
                         (float)(in[3 * v + 2] - X[3 * v + 2]), 0.f);
}

__global__ void k_export_vec(int n, const float4 *src, const int32_t *iperm, double *out)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float4 d = src[iperm[v]];
    out[3 * v] = d.x;
    out[3 * v + 1] = d.y;
    out[3 * v + 2] = d.z;
}

__global__ void k_import_vec(int n, const double *in, const int32_t *perm, float4 *dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = perm[i];
    dst[i] = make_float4((float)in[3 * v], (float)in[3 * v + 1], (float)in[3 * v + 2], 0.f);
}

inline int nblk(int64_t n, int tb) { return (int)((n + tb - 1) / tb); }

// slots per lane per load batch, by lanes per vertex
template <int L>
struct Batch {
    static constexpr int value = L == 1 ? 4 : L == 2 ? 4 : L == 4 ? 3 : L == 8 ? 2 : 1;
};

template <int L, int MAT>
void launch_sweep(const SweepArgs &a, int v0, int nv, int block, hipStream_t st)
{
    const int64_t threads = (int64_t)nv * L;
    k_sweep<L, Batch<L>::value, MAT><<<nblk(threads, block), block, 0, st>>>(a, v0, nv);
}

template <int L, int MAT>
void launch_fused(const SweepArgs &a, const int32_t *col_start, int c0, int c1, int close, int np_before, float tol,
                  hipStream_t st)
{
    k_sweep_fused<L, Batch<L>::value, MAT><<<1, 1024, 0, st>>>(a, col_start, c0, c1, close, np_before, tol);
}

template <int MAT>
void dispatch_sweep(int L, const SweepArgs &a, int v0, int nv, int block, hipStream_t st)
{
    switch (L) {
    case 1: launch_sweep<1, MAT>(a, v0, nv, block, st); break;
    case 2: launch_sweep<2, MAT>(a, v0, nv, block, st); break;
    case 8: launch_sweep<8, MAT>(a, v0, nv, block, st); break;
    case 16: launch_sweep<16, MAT>(a, v0, nv, block, st); break;
    case 32: launch_sweep<32, MAT>(a, v0, nv, block, st); break;
    default: launch_sweep<4, MAT>(a, v0, nv, block, st); break;
    }
}

template <int MAT>
void dispatch_fused(int L, const SweepArgs &a, const int32_t *cs, int c0, int c1, int close, int np_before, float tol,
                    hipStream_t st)
{
    switch (L) {
    case 1: launch_fused<1, MAT>(a, cs, c0, c1, close, np_before, tol, st); break;
    case 2: launch_fused<2, MAT>(a, cs, c0, c1, close, np_before, tol, st); break;
    case 8: launch_fused<8, MAT>(a, cs, c0, c1, close, np_before, tol, st); break;
    case 16: launch_fused<16, MAT>(a, cs, c0, c1, close, np_before, tol, st); break;
    case 32: launch_fused<32, MAT>(a, cs, c0, c1, close, np_before, tol, st); break;
    default: launch_fused<4, MAT>(a, cs, c0, c1, close, np_before, tol, st); break;
    }
}

}  // namespace

void plan_segments(vbd_solver *s)
{
    s->segments.clear();
    int c = 0;
    while (c < s->ncolors) {
        if (s->col_size[c] == 0) {
            ++c;
            continue;
        }
        if (s->col_size[c] < s->launch.small_color) {
            int c1 = c + 1;
            while (c1 < s->ncolors && s->col_size[c1] < s->launch.small_color) ++c1;
            s->segments.push_back({c, c1, true, 0, 1});
            c = c1;
        } else {
            const int64_t threads = (int64_t)s->col_size[c] * s->launch.lanes;
            s->segments.push_back({c, c + 1, false, 0, nblk(threads, s->launch.block)});
            ++c;
        }
    }
    int off = 0;
    for (auto &g : s->segments) {
        g.part_off = off;
        off += g.nparts;
    }
    s->n_parts = off;
}

static int ensure_partials(vbd_solver *s)
{
    const int need = std::max(1, s->n_parts);
    if (s->part && s->part_cap >= need) return 0;
    if (s->part) VBD_HIP(hipFree(s->part));
    s->part = nullptr;
    VBD_HIP(hipMalloc((void **)&s->part, sizeof(uint32_t) * need));
    s->part_cap = need;
    return 0;
}

static int record_event(vbd_solver *s, int color, int iter, hipEvent_t *ev)
{
    const size_t need = 2 * (s->ev_pending.size() + 1);
    while (s->ev_pool.size() < need) {
        hipEvent_t e;
        VBD_HIP(hipEventCreate(&e));
        s->ev_pool.push_back(e);
    }
    const int idx = (int)s->ev_pending.size();
    s->ev_pending.push_back({idx, color | (iter << 9)});
    ev[0] = s->ev_pool[2 * idx];
    ev[1] = s->ev_pool[2 * idx + 1];
    return 0;
}

static int enqueue_iteration(vbd_solver *s, int iter, float tol, bool profile)
{
    SweepArgs a;
    a.u = s->u;
    a.tgt = s->tgt;
    a.uprev = s->uprev;
    a.slot_off = s->slot_off;
    a.slots = s->slots;
    a.ctrl = s->ctrl;
    a.kd_dt = s->kd_dt;
    const int32_t *cs = s->d_col_start;
    const int L = s->launch.lanes, block = s->launch.block;
    bool closed = false;
    for (size_t k = 0; k < s->segments.size(); ++k) {
        const Segment &g = s->segments[k];
        a.part = s->part + g.part_off;
        if (g.fused) {
            const int close = (k + 1 == s->segments.size()) ? 1 : 0;
            if (s->material == kSNH) dispatch_fused<kSNH>(L, a, cs, g.c0, g.c1, close, g.part_off, tol, s->stream);
            else dispatch_fused<kStVK>(L, a, cs, g.c0, g.c1, close, g.part_off, tol, s->stream);
            closed = close;
        } else {
            const int v0 = s->col_start[g.c0], nv = s->col_size[g.c0];
            hipEvent_t ev[2] = {nullptr, nullptr};
            const bool timed = profile && nv >= s->prof_min;
            if (timed) {
                VBD_TRY(record_event(s, g.c0, iter, ev));
                VBD_HIP(hipEventRecord(ev[0], s->stream));
            }
            if (s->material == kSNH) dispatch_sweep<kSNH>(L, a, v0, nv, block, s->stream);
            else dispatch_sweep<kStVK>(L, a, v0, nv, block, s->stream);
            if (timed) VBD_HIP(hipEventRecord(ev[1], s->stream));
        }
    }
    if (!closed) k_check<<<1, 1024, 0, s->stream>>>(s->ctrl, s->part, s->n_parts, tol);
    VBD_HIP(hipGetLastError());
    return 0;
}

int step_device(vbd_solver *s, double dt, double tol, int max_iter, int *iters, double *metric)
{
    VBD_TRY(ensure_partials(s));
    const int n = s->n;
    const float fdt = (float)dt;
    s->kd_dt = (float)(s->damping / dt);
    k_step_init<<<nblk(std::max(n, 1), 256), 256, 0, s->stream>>>(
        n, s->u, s->uprev, s->vel, s->tgt, s->mass, s->active, fdt, (float)s->grav[0], (float)s->grav[1],
        (float)s->grav[2], s->ctrl);
    VBD_HIP(hipGetLastError());
    // A tolerance is compared in FP32 on the device; tol = 0 never converges.
    const float ftol = (float)tol;
    int launched = 0;
    int chunk = s->launch.chunk > 0 ? s->launch.chunk : std::max(1, s->last_iters);
    bool conv = false;
    while (launched < max_iter && !conv) {
        const int k = std::min(chunk, max_iter - launched);
        for (int j = 0; j < k; ++j) VBD_TRY(enqueue_iteration(s, launched + j, ftol, s->prof));
        launched += k;
        VBD_HIP(hipMemcpyAsync(s->ctrl_host, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
        VBD_HIP(hipStreamSynchronize(s->stream));
        conv = s->ctrl_host->converged != 0;
        chunk = s->launch.chunk > 0 ? s->launch.chunk : 2;
    }
    k_step_finish<<<nblk(n, 256), 256, 0, s->stream>>>(n, s->u, s->uprev, s->vel, (float)(1.0 / dt));
    VBD_HIP(hipGetLastError());
    const int it = s->ctrl_host->iters;
    s->last_iters = it;
    if (iters) *iters = it;
    if (metric) *metric = (double)__builtin_bit_cast(float, s->ctrl_host->metric);
    if (s->prof) {
        VBD_HIP(hipStreamSynchronize(s->stream));
        for (auto &pe : s->ev_pending) {
            const int color = pe.second & 511, iter = pe.second >> 9;
            if (iter >= it) continue;  // launch after convergence: returned at once
            float ms = 0.f;
            VBD_HIP(hipEventElapsedTime(&ms, s->ev_pool[2 * pe.first], s->ev_pool[2 * pe.first + 1]));
            s->prof_ms += ms;
            s->prof_bytes += s->col_bytes[color];
            s->prof_launches += 1;
        }
        s->ev_pending.clear();
    }
    return 0;
}

int export_positions(vbd_solver *s, const float4 *src, double *host_out)
{
    const int n = s->n;
    k_export<<<nblk(n, 256), 256, 0, s->stream>>>(n, src, s->iperm, s->X, s->xio);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipMemcpyAsync(s->xio_host, s->xio, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    std::copy(s->xio_host, s->xio_host + 3 * (size_t)n, host_out);
    return 0;
}

int import_positions(vbd_solver *s, const double *host_in, float4 *dst)
{
    const int n = s->n;
    std::copy(host_in, host_in + 3 * (size_t)n, s->xio_host);
    VBD_HIP(hipMemcpyAsync(s->xio, s->xio_host, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
    k_import<<<nblk(n, 256), 256, 0, s->stream>>>(n, s->xio, s->perm, s->X, dst);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

int export_velocities(vbd_solver *s, double *host_out)
{
    const int n = s->n;
    k_export_vec<<<nblk(n, 256), 256, 0, s->stream>>>(n, s->vel, s->iperm, s->xio);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipMemcpyAsync(s->xio_host, s->xio, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    std::copy(s->xio_host, s->xio_host + 3 * (size_t)n, host_out);
    return 0;
}

int import_velocities(vbd_solver *s, const double *host_in)
{
    const int n = s->n;
    std::copy(host_in, host_in + 3 * (size_t)n, s->xio_host);
    VBD_HIP(hipMemcpyAsync(s->xio, s->xio_host, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
    k_import_vec<<<nblk(n, 256), 256, 0, s->stream>>>(n, s->xio, s->perm, s->vel);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

}  // namespace vbd

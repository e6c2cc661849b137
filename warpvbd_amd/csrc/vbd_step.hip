// vbd_step.hip -- the VBD time step on the device (src/vbd_solver.py:509-612).
//
//   k_step_init    x <- x_prev + dt v_prev + dt^2 g for active vertices (:277-292);
//                  also the inertia target and M_v/dt^2 used by every sweep
//   k_sweep        one colour: per vertex, a group of L lanes walks its slots,
//                  each lane evaluating one incident tet per visit; the group
//                  reduces gradient and Hessian by butterfly shuffles, adds
//                  inertia (and damping), solves the 3x3 block in registers and
//                  writes x += dx (:149-274 fused, no per-slot scratch)
//                  the last block of the iteration's last launch folds the
//                  per-block maxima: max|dx_x+dx_y+dx_z| < tol (:578) on the
//                  device
//   k_step_resident (vbd_sweep.h) the whole step in one workgroup when every
//                  colour is small
//   k_step_finish_if  v = (x - x_prev)/dt, x_prev = x (:609-610), behind
//                  every chunk, applied by the chunk that completes the step
//
// The iteration index and the converged flag live in device memory (Ctrl), so
// the host enqueues chunks of iterations and waits once per chunk; launches
// after convergence return at their first instruction.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <hip/hip_ext.h>

#include "vbd_device.h"
#include "vbd_internal.h"
#include "vbd_mpnn.h"
#include "vbd_sweep.h"

namespace vbd {

// The one-workgroup step serves a body whose largest colour has at most this
// many vertices: one workgroup sweeps a colour in passes of 1024 threads, so
// its iteration grows with the colour while a launch per colour stays near
// the launch latency.  Measured on the mesh-size sweep's cantilevers (wall s
// per simulated s, tol 1e-6, resident against launches): 176 vertices 0.21
// resident; 1,029: 0.57 against 0.43; 3,100: 1.96 against 0.47; 6,929: 3.5
// against 0.53.
constexpr int kResidentMaxColor = 128;

namespace {

// Max over the np per-block partial maxima (this thread's share; entry skip
// reads as zero).  Eight loads are requested before any is used, so the
// fold waits for about np / (8 blockDim.x) memory latencies instead of
// np / blockDim.x.  ATOMIC: relaxed agent-scope loads, for partials that
// other blocks of the running launch wrote (read after their counter).
template <bool ATOMIC>
__device__ __forceinline__ uint32_t load_partial(const uint32_t *p)
{
    if constexpr (ATOMIC) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}

template <bool ATOMIC>
__device__ __forceinline__ uint32_t fold_partials(const uint32_t *part, int np, int skip)
{
    constexpr int U = 8;
    uint32_t x = 0;
    int i = threadIdx.x;
    const int step = blockDim.x;
    for (; i + (U - 1) * step < np; i += U * step) {
        uint32_t y[U];
#pragma unroll
        for (int k = 0; k < U; ++k)
            y[k] = load_partial<ATOMIC>(&part[i + k * step]);
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const uint32_t v = i + k * step == skip ? 0u : y[k];
            x = x > v ? x : v;
        }
    }
    for (; i < np; i += step) {
        const uint32_t v = i == skip ? 0u : load_partial<ATOMIC>(&part[i]);
        x = x > v ? x : v;
    }
    return x;
}

// A launch's end: the block's metric maximum mb (thread 0) into
// a.part[blockIdx.x]; with close, the iteration's last launch folds every
// partial maximum and decides convergence.
__device__ __forceinline__ void finish_sweep(const SweepArgs &a, Ctrl *c, uint32_t mb, int close,
                                             const uint32_t *part_all, int np_all, float tol)
{
    __shared__ int s_last;
    if (threadIdx.x == 0) a.part[blockIdx.x] = mb;
    if (!close) return;
    if (gridDim.x == 1) {
        // The only block of the iteration's last launch: the other launches'
        // partial maxima were written by kernels that finished before this one
        // started, and this block's own is mb (thread 0), so no device-scope
        // fence or block counter is needed.  Its own slot still holds an older
        // value for the other threads, so it is skipped.
        const int own = (int)(a.part - part_all);
        const uint32_t x0 = fold_partials<false>(part_all, np_all, own);
        __syncthreads();  // block_max_u32's staging is free again
        const uint32_t x = block_max_u32(x0);
        if (threadIdx.x == 0) close_iteration(c, c->it, x > mb ? x : mb, tol);
        return;
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t prev = atomicAdd(&c->blocks_done, 1u);
        s_last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    const uint32_t x = block_max_u32(fold_partials<true>(part_all, np_all, -1));
    if (threadIdx.x == 0) {
        const int it = c->it;
        c->blocks_done = 0u;
        close_iteration(c, it, x, tol);
    }
}

// One colour: a group of L lanes per vertex.  Each block leaves its metric
// maximum in a.part[blockIdx.x].  close != 0 on the iteration's last launch:
// the block that finishes last (device counter in Ctrl) folds all np_all
// partial maxima of the iteration and decides convergence
// (src/vbd_solver.py:578), so no separate check launch is needed.
template <int L, int B, int MAT>
__global__ __launch_bounds__(1024) void k_sweep(
    SweepArgs a, int v0, int nv, int close, const uint32_t *part_all, int np_all, float tol)
{
    Ctrl *c = a.ctrl;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int t = b * blockDim.x + threadIdx.x;
    const int vi = t / L, lane = t % L;
    const bool valid = vi < nv;
    const int v = v0 + (valid ? vi : 0);
    // The vertex's own loads and the converged flag are requested together: a
    // launch after convergence sweeps nothing (every lane without a vertex)
    // and stores nothing, instead of branching on the flag before its loads.
    const VertexPre pre = vertex_pre(a, v, valid);
    const bool live = !c->converged;
    const uint32_t mb = block_max_u32(update_vertex<L, B, MAT>(a, v, valid && live, lane, pre));
    if (!live) return;
    finish_sweep(a, c, mb, close, part_all, np_all, tol);
}

// A single body's big colour (at least kStageMinColor vertices at the
// default launch): 256-thread blocks over the colour's lanes as lay lays them
// out (4, 2 or 1 lanes per vertex by slot count, so that no lane walks more
// than kLaneSlots slots of a voxel mesh); blocks below lay.b1 stage their
// neighbours in LDS (staged_sweep in vbd_sweep.h), the others read them from
// global memory (lanes_sweep).  The block's metric maximum goes to a.part, and
// the iteration's last launch folds them all.
template <int MAT, bool WIDE>
__global__ __launch_bounds__(256) void k_sweep_staged(SweepArgs a, int v0, int nv, LaneLayout lay, int close,
                                                      const uint32_t *part_all, int np_all, float tol,
                                                      StageTables st, int n_dict)
{
    const int b = lane_block(blockIdx.x, lay);
    uint32_t m;
    if (b >= lay.b1) {
        if (!lanes_sweep<MAT>(a, v0, nv, lay, b, m)) return;
    } else {
        int body;
        if (!staged_sweep<MAT, false, WIDE, false>(a, v0, nv, lay, b, st, n_dict, nullptr, nullptr, nullptr, body, m))
            return;
    }
    const uint32_t mb = block_max_u32(m);
    finish_sweep(a, a.ctrl, mb, close, part_all, np_all, tol);
}

// A colour below small_color vertices: 32 lanes per vertex, one slot per
// lane, as k_sweep<32, 1, MAT> (the same arithmetic, so the same results bit
// for bit), but each lane reads its compact record from a table padded to 32
// records per vertex (entry -1 past the vertex's slots), so a lane needs no
// slot offset: two memory rounds (the vertex's data and its record; then the
// record's constants and the three neighbour positions) instead of three.
// These colours are latency bound.  In the iteration's last launch with a
// single block the fold of the other launches' maxima (final: written by
// launches that have finished) is requested with the first round, so its
// latency overlaps the vertex's.
template <int MAT>
__global__ __launch_bounds__(1024) void k_sweep_small(SweepArgs a, int v0, int nv, int close, const uint32_t *part_all,
                                                      int np_all, float tol, const float4 *prec)
{
    constexpr int L = 32;
    constexpr int U = 16;  // fold loads per thread requested with the first round
    Ctrl *c = a.ctrl;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int vi = t / L, lane = t % L;
    const bool valid = vi < nv;
    const int vr = valid ? vi : 0;
    const int v = v0 + vr;
    const bool early = close && gridDim.x == 1;
    const int own = (int)(a.part - part_all);
    uint32_t fl[U];
    if (early) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int i = threadIdx.x + k * blockDim.x;
            fl[k] = i < np_all && i != own ? part_all[i] : 0u;
        }
    }
    const float4 uv = a.u[v];
    const double4 tg = a.tgt[v];
    const float4 r = prec[(size_t)vr * L + lane];
    const bool live = !c->converged;
    VertexPre pre;
    pre.uv = uv;
    pre.tg = tg;
    pre.s0 = pre.s1 = 0;
    Acc acc;
    acc_zero(acc);
    VertexLocal ql = vertex_local(pre);
    const int e = __float_as_int(r.w);
    if (valid && live && e >= 0) {
        const float4 *d = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(a.dict) + (uint32_t)e * 48u);
        const float4 d0 = d[0], d1 = d[1], d2 = d[2];
        const bool edge = a.dict_ext != nullptr;
        const EdgeK ek = edge_k<MAT>(a.dict_ext, (uint32_t)e, edge);
        const float4 x0 = *at(a.u, __float_as_uint(r.x));
        const float4 x1 = *at(a.u, __float_as_uint(r.y));
        const float4 x2 = *at(a.u, __float_as_uint(r.z));
        const SlotRel sr = slot_rel(x0, x1, x2, ql.hi, ql.lo);
        slot_eval<MAT>(d2.z, d0.w, slot_kc<MAT>(d2.z), d2.w, make_float3(d0.x, d0.y, d0.z),
                       make_float3(d1.x, d1.y, d1.z), make_float3(d1.w, d2.x, d2.y), ek, edge, sr.e0, sr.e1, sr.d2,
                       acc);
    }
    double3 us;
    const float3 dx = vertex_finish<L>(a, v, valid && live, acc, ql, us);
    uint32_t m = 0u;
    if (valid && live) {
        if (lane == 0) a.u[v] = pack_disp(us.x + dx.x, us.y + dx.y, us.z + dx.z);
        m = metric_bits(dx);
    }
    const uint32_t mb = block_max_u32(m);
    if (!live) return;
    if (!early) {
        finish_sweep(a, c, mb, close, part_all, np_all, tol);
        return;
    }
    // the rest of the fold (partials beyond the first U per thread), then the
    // iteration's close
    uint32_t x = 0u;
#pragma unroll
    for (int k = 0; k < U; ++k) x = x > fl[k] ? x : fl[k];
    for (int i = threadIdx.x + U * blockDim.x; i < np_all; i += blockDim.x) {
        const uint32_t y = i == own ? 0u : part_all[i];
        x = x > y ? x : y;
    }
    __syncthreads();  // block_max_u32's staging is free again
    const uint32_t all = block_max_u32(x);
    if (threadIdx.x == 0) {
        a.part[0] = mb;
        close_iteration(c, c->it, all > mb ? all : mb, tol);
    }
}

// The tail sweep's tables for one launch (the tail_* fields of the solver).
// Row j of colour k in cluster c is row c * rows_total + row_off[k] + j of v
// and of rec (32 lanes per row).
struct TailArgs {
    const int2 *v;        // per row: vertex (-1: none), its index in the cluster's table (-1: none)
    const uint2 *rec;     // per row and lane: n0 | n1 << 10 | n2 << 20, dictionary entry (-1: none)
    const int32_t *nb;    // per cluster: nb_stride neighbour ids, -1 past its count
    int nb_stride;
    int rows_total;
    int rows[kMaxTail];   // rows of colour k per cluster
    int row_off[kMaxTail];
    int n_dict;
};

// Lanes per vertex of the tail sweep: each lane evaluates 32 / kTailLanes
// consecutive slot positions of its vertex.
#ifndef VBD_TAIL_LANES
#define VBD_TAIL_LANES 16
#endif
constexpr int kTailLanes = VBD_TAIL_LANES;
static_assert(kTailLanes == 8 || kTailLanes == 16 || kTailLanes == 32, "8, 16 or 32 lanes per tail vertex");

// One slot of a tail vertex from zero (record r, entry -1: none).
template <int MAT>
__device__ __forceinline__ Acc tail_slot(uint2 r, const float4 *s_x, const float4 *s_dict, const float4 *s_ext,
                                         const VertexLocal &ql)
{
    Acc p;
    acc_zero(p);
    const int e = (int)r.y;
    if (e >= 0) {
        const float4 x0 = s_x[r.x & 1023u], x1 = s_x[(r.x >> 10) & 1023u], x2 = s_x[(r.x >> 20) & 1023u];
        const float4 d0 = s_dict[3 * e], d1 = s_dict[3 * e + 1], d2 = s_dict[3 * e + 2];
        const EdgeK ek = edge_k<MAT>(s_ext, (uint32_t)e);
        const SlotRel sr = slot_rel(x0, x1, x2, ql.hi, ql.lo);
        slot_eval<MAT>(d2.z, d0.w, slot_kc<MAT>(d2.z), d2.w, make_float3(d0.x, d0.y, d0.z),
                       make_float3(d1.x, d1.y, d1.z), make_float3(d1.w, d2.x, d2.y), ek, true, sr.e0, sr.e1, sr.d2,
                       p);
    }
    return p;
}

// One lane's share of a tail vertex's slots: its S = 32 / kTailLanes
// consecutive slot positions (records q), each evaluated from zero and summed
// as a balanced tree (s0 + s1, (s0 + s1) + (s2 + s3)).  With
// group_sum<kTailLanes> over the lanes after it, a vertex's sum is the
// balanced tree over its 32 slot positions that group_sum<32> forms over
// k_sweep_small's one slot per lane, so the two give the same result bit for
// bit.
template <int MAT, int S>
__device__ __forceinline__ Acc tail_slots(const uint2 *q, const float4 *s_x, const float4 *s_dict,
                                          const float4 *s_ext, const VertexLocal &ql)
{
    if constexpr (S == 1) return tail_slot<MAT>(q[0], s_x, s_dict, s_ext, ql);
    else
        return acc_add(tail_slots<MAT, S / 2>(q, s_x, s_dict, s_ext, ql),
                       tail_slots<MAT, S / 2>(q + S / 2, s_x, s_dict, s_ext, ql));
}

// The tail colours (K trailing colours of 32 lanes per vertex) of a single
// body in one launch: one block per cluster, 32 rows of a colour x
// kTailLanes lanes, lane l evaluating the row vertex's slot positions
// S l .. S l + S - 1 (tail_slots).  No edge joins two clusters, so a
// cluster's vertices depend only on each other and on the colours before the
// tail, which this launch does not write.  Two memory rounds: (1) every row's
// vertex and every lane's records of the K colours, the cluster's neighbour
// ids, the dictionary, the iteration index, and this block's share of the
// other launches' block maxima (final: those launches have finished); (2) the
// rows' own data and the neighbours' packed positions, into LDS.  Then colour
// by colour from LDS, a block barrier between colours: a row's update goes
// to memory and to its entry in the table.  The results are k_sweep_small's
// bit for bit.  Each block adds its metric maximum into Ctrl.launch_max and
// counts itself done (release); the last one reads the maximum (acquire) and
// closes the iteration (src/vbd_solver.py:578).
template <int MAT, int K>
__global__ __launch_bounds__(kTailRows * kTailLanes) void k_sweep_tail(SweepArgs a, TailArgs t,
                                                                       const uint32_t *part_all, int np_before,
                                                                       float tol)
{
    constexpr int L = kTailLanes, S = 32 / L, kThreads = kTailRows * L, kNbPer = kTailNb / kThreads;
    static_assert(kTailNb % kThreads == 0, "the neighbour table splits evenly over the block");
    __shared__ float4 s_x[kTailNb];
    __shared__ float4 s_dict[3 * kStageDict];
    __shared__ float4 s_ext[MAT == kSNH ? 3 * kStageDict : 1];  // the entries' rest-edge constants
    static_assert(3 * kStageDict <= kThreads, "a dictionary row per thread");
    Ctrl *c = a.ctrl;
    const int row = (int)threadIdx.x / L, lane = (int)threadIdx.x % L;
    const int64_t cl = blockIdx.x;
    // round 1
    int2 vo[K];
    uint2 rc[K][S];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        vo[k] = make_int2(-1, -1);
#pragma unroll
        for (int j = 0; j < S; ++j) rc[k][j] = make_uint2(0u, 0xFFFFFFFFu);
        if (row < t.rows[k]) {
            const int64_t r = cl * t.rows_total + t.row_off[k] + row;
            vo[k] = t.v[r];
#pragma unroll
            for (int j = 0; j < S; ++j) rc[k][j] = t.rec[r * 32 + S * lane + j];
        }
    }
    int id[kNbPer];
#pragma unroll
    for (int j = 0; j < kNbPer; ++j) {
        const int i = (int)threadIdx.x + kThreads * j;
        id[j] = i < t.nb_stride ? t.nb[cl * t.nb_stride + i] : -1;
    }
    const bool has_dict = (int)threadIdx.x < 3 * t.n_dict;
    const float4 drow = has_dict ? a.dict[threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 xrow = MAT == kSNH && has_dict ? a.dict_ext[threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t m = 0u;
    for (int64_t i = cl * blockDim.x + threadIdx.x; i < np_before; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t y = part_all[i];
        m = m > y ? m : y;
    }
    const int it = c->it;
    // the same for every block: a launch after convergence sweeps nothing
    if (c->converged) return;
    // round 2
    float4 px[kNbPer];
#pragma unroll
    for (int j = 0; j < kNbPer; ++j)
        px[j] = id[j] >= 0 ? *at(a.u, (uint32_t)id[j]) : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 uv[K];
    double4 tg[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        uv[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        tg[k] = make_double4(0.0, 0.0, 0.0, 0.0);
        if (vo[k].x >= 0) {
            uv[k] = a.u[vo[k].x];
            tg[k] = a.tgt[vo[k].x];
        }
    }
#pragma unroll
    for (int j = 0; j < kNbPer; ++j) s_x[threadIdx.x + kThreads * j] = px[j];
    if (has_dict) {
        s_dict[threadIdx.x] = drow;
        if constexpr (MAT == kSNH) s_ext[threadIdx.x] = xrow;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int v = vo[k].x;
        const bool valid = v >= 0;
        VertexPre pre;
        pre.uv = uv[k];
        pre.tg = tg[k];
        pre.s0 = pre.s1 = 0;
        Acc acc;
        acc_zero(acc);
        VertexLocal ql = vertex_local(pre);
        if (valid) acc = tail_slots<MAT, S>(rc[k], s_x, s_dict, s_ext, ql);
        double3 us;
        const float3 dx = vertex_finish<L>(a, valid ? v : 0, valid, acc, ql, us);
        if (valid) {
            if (lane == 0) {
                const float4 nu = pack_disp(us.x + dx.x, us.y + dx.y, us.z + dx.z);
                a.u[v] = nu;
                if (vo[k].y >= 0) s_x[vo[k].y] = nu;
            }
            const uint32_t mk = metric_bits(dx);
            m = m > mk ? m : mk;
        }
        if (k + 1 < K) __syncthreads();  // the next colour reads this one's updates
    }
    const uint32_t mb = block_max_u32(m);
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_max(&c->launch_max, mb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t prev = __hip_atomic_fetch_add(&c->blocks_done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == gridDim.x - 1) {
            // every block's maximum is in launch_max (their increments came first)
            const uint32_t x = __hip_atomic_load(&c->launch_max, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            c->blocks_done = 0u;
            c->launch_max = 0u;
            close_iteration(c, it, x, tol);
        }
    }
}

// The padded record table of one colour: row (v - v0) holds the vertex's
// compact records in its 32 lanes, entry -1 past its last slot.
__global__ void k_pad_records(const float4 *slots, const int32_t *slot_off, int v0, int nv, float4 *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nv * 32) return;
    const int v = v0 + (int)(t >> 5), lane = (int)(t & 31);
    const int s0 = slot_off[v], s1 = slot_off[v + 1];
    out[t] = s0 + lane < s1 ? slots[s0 + lane] : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
}

// Iteration close when no colour has a vertex to sweep.
__global__ void k_check_empty(Ctrl *c, float tol)
{
    if (c->converged) return;
    close_iteration(c, c->it, 0u, tol);
}

__global__ void k_step_init(int n, float4 *u, const double4 *uprev, const float4 *vel, double4 *tgt, const float *mass,
                            const uint8_t *active, float dt, float4 g, Ctrl *ctrl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctrl->it = 0;
        ctrl->converged = 0;
        ctrl->iters = 0;
        ctrl->metric = 0u;
        ctrl->blocks_done = 0u;
        ctrl->launch_max = 0u;
    }
    if (i < n) init_vertex(i, u, uprev, vel, tgt, mass, active, dt, g);
}

// The largest |dx| component of the last iteration over the swept vertices:
// u against u_last, the iterate before it (one workgroup; used only by steps
// that ran their last allowed iteration).
__global__ __launch_bounds__(1024) void k_dx_max(int n, const float4 *u, const float4 *u_last, double *out)
{
    __shared__ double red[16];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double4 a = unpack_disp4(u[i]), b = unpack_disp4(u_last[i]);
        m = fmax(m, fmax(fabs(a.x - b.x), fmax(fabs(a.y - b.y), fabs(a.z - b.z))));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
        *out = m;
    }
}

// v = (x - x_prev)/dt, x_prev = x once the step is complete (converged, or
// its last allowed iteration ran), else nothing: enqueued behind every chunk
// of iterations, so the velocity update follows the last chunk without a wait
// on the host.
__global__ void k_step_finish_if(int n, const float4 *u, double4 *uprev, float4 *vel, float dt, const Ctrl *ctrl,
                                 int max_iter)
{
    if (!ctrl->converged && ctrl->iters < max_iter) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 x = unpack_disp4(u[i]), p = uprev[i];
    vel[i] = step_velocity(x, p, dt);
    uprev[i] = x;
}

__device__ __forceinline__ double4 load_disp(const double4 *p, int i) { return p[i]; }
__device__ __forceinline__ double4 load_disp(const float4 *p, int i) { return unpack_disp4(p[i]); }

// x (original order, f64) = X + u[iperm] (u in FP64 or packed)
template <typename T>
__global__ void k_export(int n, const T *src, const int32_t *iperm, const double *X, double *out)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const double4 d = load_disp(src, iperm[v]);
    out[3 * v] = X[3 * v] + d.x;
    out[3 * v + 1] = X[3 * v + 1] + d.y;
    out[3 * v + 2] = X[3 * v + 2] + d.z;
}

__global__ void k_import(int n, const double *in, const int32_t *perm, const double *X, double4 *dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = perm[i];
    dst[i] = make_double4(in[3 * v] - X[3 * v], in[3 * v + 1] - X[3 * v + 1], in[3 * v + 2] - X[3 * v + 2], 0.0);
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

// hipExtLaunchKernelGGL: the optional events take the dispatch's own start and
// end timestamps, so a timed launch measures the kernel alone.
template <int L, int MAT>
void launch_sweep(const SweepArgs &a, int v0, int nv, int block, int close, const uint32_t *part_all, int np_all,
                  float tol, hipEvent_t e0, hipEvent_t e1, hipStream_t st)
{
    const int64_t threads = (int64_t)nv * L;
    hipExtLaunchKernelGGL(k_sweep<L, Batch<L>::value, MAT>, dim3(nblk(threads, block)), dim3(block), 0, st, e0, e1, 0,
                          a, v0, nv, close, part_all, np_all, tol);
}

void launch_staged(int material, const SweepArgs &a, int v0, int nv, const LaneLayout &lay, int close,
                   const uint32_t *part_all, int np_all, float tol, const StageTables &tables, int n_dict,
                   hipEvent_t e0, hipEvent_t e1, hipStream_t st)
{
    const dim3 grid(lay.nblocks);
    // the larger lists (and LDS) only when a staged block holds 2-lane vertices
    const bool wide = lay.b4 < lay.b1;
    auto kern = material == kSNH ? (wide ? k_sweep_staged<kSNH, true> : k_sweep_staged<kSNH, false>)
                                 : (wide ? k_sweep_staged<kStVK, true> : k_sweep_staged<kStVK, false>);
    hipExtLaunchKernelGGL(kern, grid, dim3(256), 0, st, e0, e1, 0, a, v0, nv, lay, close, part_all, np_all, tol, tables,
                          n_dict);
}

void launch_small(int material, const SweepArgs &a, int v0, int nv, int block, int close, const uint32_t *part_all,
                  int np_all, float tol, const float4 *prec, hipEvent_t e0, hipEvent_t e1, hipStream_t st)
{
    const dim3 grid(nblk((int64_t)nv * 32, block));
    if (material == kSNH)
        hipExtLaunchKernelGGL(k_sweep_small<kSNH>, grid, dim3(block), 0, st, e0, e1, 0, a, v0, nv, close, part_all,
                              np_all, tol, prec);
    else
        hipExtLaunchKernelGGL(k_sweep_small<kStVK>, grid, dim3(block), 0, st, e0, e1, 0, a, v0, nv, close, part_all,
                              np_all, tol, prec);
}

template <int MAT>
void launch_tail_k(int K, const SweepArgs &a, const TailArgs &t, int clusters, const uint32_t *part_all,
                   int np_before, float tol, hipEvent_t e0, hipEvent_t e1, hipStream_t st)
{
    auto kern = K == 2 ? k_sweep_tail<MAT, 2> : K == 3 ? k_sweep_tail<MAT, 3> : k_sweep_tail<MAT, 4>;
    hipExtLaunchKernelGGL(kern, dim3(clusters), dim3(kTailRows * kTailLanes), 0, st, e0, e1, 0, a, t, part_all,
                          np_before, tol);
}

void launch_tail(const vbd_solver *s, const SweepArgs &a, int np_before, float tol, hipEvent_t e0, hipEvent_t e1)
{
    TailArgs t;
    t.v = s->tail_v;
    t.rec = s->tail_rec;
    t.nb = s->tail_nb;
    t.nb_stride = s->tail_nb_stride;
    t.rows_total = 0;
    for (int k = 0; k < kMaxTail; ++k) {
        t.rows[k] = k < s->tail_k ? s->tail_rows[k] : 0;
        t.row_off[k] = t.rows_total;
        t.rows_total += t.rows[k];
    }
    t.n_dict = s->n_dict;
    if (s->material == kSNH)
        launch_tail_k<kSNH>(s->tail_k, a, t, s->tail_clusters, s->part, np_before, tol, e0, e1, s->stream);
    else
        launch_tail_k<kStVK>(s->tail_k, a, t, s->tail_clusters, s->part, np_before, tol, e0, e1, s->stream);
}

template <int MAT>
void dispatch_sweep(int L, const SweepArgs &a, int v0, int nv, int block, int close, const uint32_t *part_all,
                    int np_all, float tol, hipEvent_t e0, hipEvent_t e1, hipStream_t st)
{
    switch (L) {
    case 1: launch_sweep<1, MAT>(a, v0, nv, block, close, part_all, np_all, tol, e0, e1, st); break;
    case 2: launch_sweep<2, MAT>(a, v0, nv, block, close, part_all, np_all, tol, e0, e1, st); break;
    case 8: launch_sweep<8, MAT>(a, v0, nv, block, close, part_all, np_all, tol, e0, e1, st); break;
    case 16: launch_sweep<16, MAT>(a, v0, nv, block, close, part_all, np_all, tol, e0, e1, st); break;
    case 32: launch_sweep<32, MAT>(a, v0, nv, block, close, part_all, np_all, tol, e0, e1, st); break;
    default: launch_sweep<4, MAT>(a, v0, nv, block, close, part_all, np_all, tol, e0, e1, st); break;
    }
}

}  // namespace

const LaneLayout *big_color_layout(const vbd_solver *s, int c)
{
    if (s->batched || s->launch.lanes != 4 || s->launch.block != 256 || c >= (int)s->st_col.size()) return nullptr;
    const LaneLayout &l = s->st_col[c];
    if (l.nblocks == 0 || s->col_size[c] < s->launch.small_color) return nullptr;
    return &l;
}

// One launch per non-empty colour.  Colours below small_color vertices are
// latency bound: they get 32 lanes per vertex (one slot per lane) so a vertex
// finishes in one load round.  Big colours of a single body take their lane
// layout by slot count (k_sweep_staged).  When every colour is small the
// whole step runs in one workgroup instead (k_step_resident).
void plan_segments(vbd_solver *s)
{
    s->segments.clear();
    int mx = 0;
    for (int c = 0; c < s->ncolors; ++c) {
        const int nv = s->col_size[c];
        mx = std::max(mx, nv);
        if (nv == 0) continue;
        const int lanes = nv < s->launch.small_color ? 32 : s->launch.lanes;
        const int64_t threads = (int64_t)nv * lanes;
        const LaneLayout *lay = big_color_layout(s, c);
        s->segments.push_back({c, c + 1, 0, lay ? lay->nblocks : nblk(threads, s->launch.block), lanes,
                               s->launch.block});
    }
    // the tail sweep's colours, when its tables were built for this plan's
    // trailing small colours, in one launch (k_sweep_tail)
    if (s->tail_ok && (int)s->segments.size() >= s->tail_k) {
        const size_t first = s->segments.size() - s->tail_k;
        bool match = true;
        for (int k = 0; k < s->tail_k; ++k) {
            const Segment &g = s->segments[first + k];
            match = match && g.c0 == s->tail_col[k] && g.lanes == 32 && s->col_size[g.c0] < s->launch.small_color;
        }
        if (match) {
            s->segments.resize(first);
            s->segments.push_back({s->tail_col[0], s->tail_col[s->tail_k - 1] + 1, 0, s->tail_clusters, kTailLanes,
                                   kTailRows * kTailLanes, 1});
        }
    }
    // the iteration's last launch: one block of up to 1024 threads when its
    // colour fits (the padded small-colour kernel then folds the other
    // launches' maxima alongside its own vertices)
    if (!s->segments.empty() && !s->segments.back().tail) {
        Segment &g = s->segments.back();
        const int64_t threads = (int64_t)s->col_size[g.c0] * g.lanes;
        if (g.lanes == 32 && threads <= 1024) {
            g.block = 1024;
            g.nparts = 1;
        }
    }
    int off = 0;
    for (auto &g : s->segments) {
        g.part_off = off;
        off += g.nparts;
    }
    s->n_parts = off;
    s->resident = s->launch.resident && mx < s->launch.small_color && mx <= kResidentMaxColor;
    // lanes of the resident form: as many as keep the largest colour within one
    // pass of the 1024-thread workgroup
    int lanes = 1;
    while (lanes < 32 && (int64_t)std::max(mx, 1) * lanes * 2 <= 1024) lanes *= 2;
    s->resident_lanes = lanes;
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

// The tail sweep's tables (tail_*): the trailing colours that the plan sweeps
// with 32 lanes per vertex (below launch.small_color; the last kMaxTail of
// them, at least 2), when their vertices split into at most
// kTailMaxClusters clusters joined by no edge with at most kTailRows vertices
// of a colour and kTailNb distinct neighbours each (fewer of the trailing
// colours are tried when they do not);
// else the solver keeps a launch per colour.  A shared dictionary of at most
// kStageDict entries is needed (it is staged in LDS).  Built on the host from
// the tail vertices' slot records (the tail is the end of the swept range).
static int build_tail(vbd_solver *s)
{
    s->tail_ok = false;
    for (void **p : {(void **)&s->tail_v, (void **)&s->tail_rec, (void **)&s->tail_nb}) {
        if (*p) VBD_HIP(hipFree(*p));
        *p = nullptr;
    }
#if VBD_TET_RECORDS
    return 0;
#else
    if (!g_tail || s->n_bodies > 1 || s->batched || !s->slots || !s->dict_shared || !s->dict_ext ||
        s->n_dict > kStageDict ||
        s->n_dict <= 0)
        return 0;
    std::vector<int> cand;  // the trailing small colours, in order
    for (int k = (int)s->segments.size() - 1; k >= 0; --k) {
        const Segment &g = s->segments[k];
        if (g.tail || g.c1 != g.c0 + 1 || g.lanes != 32 || s->col_size[g.c0] >= s->launch.small_color) break;
        cand.insert(cand.begin(), g.c0);
    }
    if (cand.size() < 2) return 0;
    if ((int)cand.size() > kMaxTail) cand.erase(cand.begin(), cand.end() - kMaxTail);
    const int va = s->col_start[cand[0]], vb = s->col_start[cand.back()] + s->col_size[cand.back()];
    std::vector<int32_t> off(vb - va + 1);
    VBD_HIP(hipMemcpyAsync(off.data(), s->slot_off + va, sizeof(int32_t) * off.size(), hipMemcpyDeviceToHost,
                           s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    std::vector<float4> rec(off.back() - off[0]);
    if (!rec.empty())
        VBD_HIP(hipMemcpyAsync(rec.data(), s->slots + off[0], sizeof(float4) * rec.size(), hipMemcpyDeviceToHost,
                               s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    auto ids = [&](int v, int64_t q, uint32_t w[4]) { std::memcpy(w, &rec[off[v - va] - off[0] + q], 16); };
    std::vector<int32_t> stamp(s->n, -1), local(s->n, 0);
    for (int K = (int)cand.size(); K >= 2; --K) {
        const int c0 = cand[cand.size() - K];
        const int v0 = s->col_start[c0], nt = vb - v0;
        // clusters: components of the tail vertices joined by a slot
        std::vector<int> par(nt);
        for (int i = 0; i < nt; ++i) par[i] = i;
        auto find = [&](int i) {
            while (par[i] != i) i = par[i] = par[par[i]];
            return i;
        };
        for (int v = v0; v < vb; ++v)
            for (int64_t q = 0; q < off[v + 1 - va] - off[v - va]; ++q) {
                uint32_t w[4];
                ids(v, q, w);
                for (int j = 0; j < 3; ++j)
                    if ((int)w[j] >= v0 && (int)w[j] < vb) par[find(v - v0)] = find((int)w[j] - v0);
            }
        std::vector<int> cl_of(nt, -1), root_cl(nt, -1);
        int ncl = 0;
        for (int i = 0; i < nt; ++i) {
            const int r = find(i);
            if (root_cl[r] < 0) root_cl[r] = ncl++;
            cl_of[i] = root_cl[r];
        }
        if (ncl > kTailMaxClusters) continue;
        // rows: per (cluster, colour) its vertices in order
        std::vector<int> cnt((size_t)ncl * K, 0), rowk(nt);
        int rows[kMaxTail] = {};
        bool ok = true;
        for (int k = 0; k < K && ok; ++k) {
            const int c = cand[cand.size() - K + k];
            for (int v = s->col_start[c]; v < s->col_start[c] + s->col_size[c]; ++v) {
                int &n = cnt[(size_t)cl_of[v - v0] * K + k];
                rowk[v - v0] = n++;
                rows[k] = std::max(rows[k], n);
                ok = ok && n <= kTailRows && off[v + 1 - va] - off[v - va] <= 32;
            }
        }
        if (!ok) continue;
        int rt = 0, roff[kMaxTail] = {};
        for (int k = 0; k < K; ++k) {
            roff[k] = rt;
            rt += rows[k];
        }
        // neighbour tables: per cluster its distinct neighbours in order of
        // first appearance over its vertices' slots
        std::vector<std::vector<int>> members(ncl);
        for (int i = 0; i < nt; ++i) members[cl_of[i]].push_back(v0 + i);
        std::vector<int32_t> nb;
        int stride = 0;
        std::vector<int> ncount(ncl, 0);
        for (int cl = 0; cl < ncl && ok; ++cl)
            for (int v : members[cl])
                for (int64_t q = 0; q < off[v + 1 - va] - off[v - va] && ok; ++q) {
                    uint32_t w[4];
                    ids(v, q, w);
                    for (int j = 0; j < 3; ++j)
                        if (stamp[w[j]] != cl + K * s->n) {  // stamps differ between attempts
                            stamp[w[j]] = cl + K * s->n;
                            ++ncount[cl];
                        }
                    ok = ncount[cl] <= kTailNb;
                }
        if (!ok) continue;
        for (int cl = 0; cl < ncl; ++cl) stride = std::max(stride, ncount[cl]);
        stride = (stride + 63) & ~63;
        // at most 64 MB of tables
        if ((double)ncl * rt * 32 * sizeof(uint2) + (double)ncl * stride * 4 > 64.0 * (1 << 20)) continue;
        nb.assign((size_t)ncl * stride, -1);
        std::vector<int2> tv((size_t)ncl * rt, make_int2(-1, -1));
        std::vector<uint2> tr((size_t)ncl * rt * 32, make_uint2(0u, 0xFFFFFFFFu));
        for (int cl = 0; cl < ncl; ++cl) {
            int n = 0;
            const int tag = -2 - cl;  // this pass's stamps (the counting pass left nonnegative ones)
            for (int v : members[cl])
                for (int64_t q = 0; q < off[v + 1 - va] - off[v - va]; ++q) {
                    uint32_t w[4];
                    ids(v, q, w);
                    for (int j = 0; j < 3; ++j)
                        if (stamp[w[j]] != tag) {
                            stamp[w[j]] = tag;
                            local[w[j]] = n;
                            nb[(size_t)cl * stride + n++] = (int32_t)w[j];
                        }
                }
            for (int v : members[cl]) {
                int k = K - 1;
                while (v < s->col_start[cand[cand.size() - K + k]]) --k;
                const size_t r = (size_t)cl * rt + roff[k] + rowk[v - v0];
                tv[r] = make_int2(v, stamp[v] == tag ? local[v] : -1);
                for (int64_t q = 0; q < off[v + 1 - va] - off[v - va]; ++q) {
                    uint32_t w[4];
                    ids(v, q, w);
                    tr[r * 32 + q] = make_uint2((uint32_t)local[w[0]] | (uint32_t)local[w[1]] << 10 |
                                                    (uint32_t)local[w[2]] << 20,
                                                w[3]);
                }
            }
        }
        // every index the kernel takes from the tables stays in its buffer
        for (int cl = 0; cl < ncl && ok; ++cl)
            for (int r = 0; r < rt && ok; ++r) {
                const int2 q = tv[(size_t)cl * rt + r];
                ok = q.x < vb && q.y < ncount[cl];
                for (int l = 0; l < 32 && ok; ++l) {
                    const uint2 e = tr[((size_t)cl * rt + r) * 32 + l];
                    if ((int)e.y < 0) continue;
                    ok = q.x >= 0 && (int)e.y < s->n_dict && (int)(e.x & 1023u) < ncount[cl] &&
                         (int)((e.x >> 10) & 1023u) < ncount[cl] && (int)(e.x >> 20) < ncount[cl];
                }
            }
        if (!ok) return 0;
        VBD_HIP(hipMalloc((void **)&s->tail_v, sizeof(int2) * tv.size()));
        VBD_HIP(hipMalloc((void **)&s->tail_rec, sizeof(uint2) * tr.size()));
        VBD_HIP(hipMalloc((void **)&s->tail_nb, sizeof(int32_t) * nb.size()));
        VBD_HIP(hipMemcpyAsync(s->tail_v, tv.data(), sizeof(int2) * tv.size(), hipMemcpyHostToDevice, s->stream));
        VBD_HIP(hipMemcpyAsync(s->tail_rec, tr.data(), sizeof(uint2) * tr.size(), hipMemcpyHostToDevice,
                               s->stream));
        VBD_HIP(hipMemcpyAsync(s->tail_nb, nb.data(), sizeof(int32_t) * nb.size(), hipMemcpyHostToDevice,
                               s->stream));
        VBD_HIP(hipStreamSynchronize(s->stream));
        s->tail_k = K;
        s->tail_clusters = ncl;
        s->tail_nb_stride = stride;
        for (int k = 0; k < kMaxTail; ++k) {
            s->tail_col[k] = k < K ? cand[cand.size() - K + k] : 0;
            s->tail_rows[k] = k < K ? rows[k] : 0;
        }
        s->tail_ok = true;
        return 0;
    }
    return 0;
#endif
}

int build_small_records(vbd_solver *s)
{
    // a launch per colour while the tables are built
    s->tail_ok = false;
    plan_segments(s);
    if (s->small_rec) VBD_HIP(hipFree(s->small_rec));
    s->small_rec = nullptr;
    s->small_off.assign(s->ncolors, -1);
#if VBD_TET_RECORDS
    return 0;
#else
    if (s->n_bodies > 1 || !s->slots) return 0;
    int64_t rows = 0;
    for (const Segment &g : s->segments) {
        if (g.lanes != 32) continue;
        s->small_off[g.c0] = rows;
        rows += s->col_size[g.c0];
    }
    if (rows == 0) return 0;
    VBD_HIP(hipMalloc((void **)&s->small_rec, sizeof(float4) * 32 * rows));
    for (const Segment &g : s->segments) {
        if (g.lanes != 32) continue;
        const int nv = s->col_size[g.c0];
        k_pad_records<<<nblk((int64_t)nv * 32, 256), 256, 0, s->stream>>>(
            s->slots, s->slot_off, s->col_start[g.c0], nv, s->small_rec + 32 * s->small_off[g.c0]);
    }
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipStreamSynchronize(s->stream));
    VBD_TRY(build_tail(s));
    plan_segments(s);
    return 0;
#endif
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
    a.dict_ext = s->dict_ext;
    a.dict = s->dict;
    a.ent = s->slot_ent;
    a.ns = std::max<int64_t>(s->n_slots, 1);
    a.ctrl = s->ctrl;
    a.kd_dt = s->kd_dt;
    for (size_t k = 0; k < s->segments.size(); ++k) {
        const Segment &g = s->segments[k];
        a.part = s->part + g.part_off;
        const int close = (k + 1 == s->segments.size()) ? 1 : 0;
        const int v0 = s->col_start[g.c0], nv = s->col_size[g.c0];
        hipEvent_t ev[2] = {nullptr, nullptr};
        if (profile && nv >= s->prof_min) VBD_TRY(record_event(s, g.c0, iter, ev));
        // big colours at the default launch (4 lanes, 256 threads): the lane
        // layout by slot count, with neighbour staging where it was built
        const LaneLayout *lay = big_color_layout(s, g.c0);
        const float4 *prec = s->small_rec && g.lanes == 32 && s->small_off[g.c0] >= 0
                                 ? s->small_rec + 32 * s->small_off[g.c0]
                                 : nullptr;
        if (g.tail)
            launch_tail(s, a, g.part_off, tol, ev[0], ev[1]);
        else if (lay)
            launch_staged(s->material, a, v0, nv, *lay, close, s->part, s->n_parts, tol,
                          StageTables{s->st_nb, s->st_rec, s->st_vrec},
                          s->n_dict, ev[0], ev[1], s->stream);
        else if (prec)
            launch_small(s->material, a, v0, nv, g.block, close, s->part, s->n_parts, tol, prec, ev[0], ev[1],
                         s->stream);
        else if (s->material == kSNH)
            dispatch_sweep<kSNH>(g.lanes, a, v0, nv, g.block, close, s->part, s->n_parts, tol, ev[0], ev[1],
                                 s->stream);
        else
            dispatch_sweep<kStVK>(g.lanes, a, v0, nv, g.block, close, s->part, s->n_parts, tol, ev[0], ev[1],
                                  s->stream);
    }
    if (s->segments.empty()) k_check_empty<<<1, 1, 0, s->stream>>>(s->ctrl, tol);
    VBD_HIP(hipGetLastError());
    return 0;
}

// A hipGraph of k iterations (all colour launches) for the current launch
// plan, tolerance and damping; kernels take every per-step value from device
// memory, so one graph serves every step.
static int single_graph(vbd_solver *s, int k, float tol, hipGraphExec_t *out)
{
    for (auto &g : s->graphs)
        if (g.k == k && g.tol == tol && g.kd_dt == s->kd_dt && g.lanes == s->launch.lanes &&
            g.block == s->launch.block && g.max_iter == s->launch.small_color) {
            graph_touch(s, g);
            *out = g.exec;
            return 0;
        }
    graph_make_room(s);
    CachedGraph g;
    g.k = k;
    g.tol = tol;
    g.kd_dt = s->kd_dt;
    g.lanes = s->launch.lanes;
    g.block = s->launch.block;
    g.max_iter = s->launch.small_color;
    VBD_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int j = 0; j < k && rc == 0; ++j) rc = enqueue_iteration(s, j, tol, false);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s->stream, &graph);
    if (rc != 0) return rc;
    if (e != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return VBD_EHIP;
    }
    g.graph = graph;
    VBD_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    graph_touch(s, g);
    s->graphs.push_back(g);
    *out = g.exec;
    return 0;
}

static int ensure_u_last(vbd_solver *s)
{
    if (s->u_last) return 0;
    VBD_HIP(hipMalloc((void **)&s->u_last, sizeof(float4) * std::max(s->n_swept, 1)));
    VBD_HIP(hipMalloc((void **)&s->d_dx_max, sizeof(double)));
    return 0;
}

// After a step of `it` iterations: the last iteration's largest |dx| component
// when it was the last allowed one (the reference prints it then,
// src/vbd_solver.py:581-582), else NaN.
static int report_dx_max(vbd_solver *s, int it, int max_iter)
{
    s->last_dx_max = std::nan("");
    if (it != max_iter || !s->u_last || s->n_swept == 0) return 0;
    k_dx_max<<<1, 1024, 0, s->stream>>>(s->n_swept, s->u, s->u_last, s->d_dx_max);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipMemcpyAsync(&s->last_dx_max, s->d_dx_max, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

static int step_resident(vbd_solver *s, double dt, double tol, int max_iter, int *iters, double *metric)
{
    SweepArgs a;
    a.u = s->u;
    a.tgt = s->tgt;
    a.uprev = s->uprev;
    a.slot_off = s->slot_off;
    a.slots = s->slots;
    a.dict_ext = s->dict_ext;
    a.dict = s->dict;
    a.ent = s->slot_ent;
    a.ns = std::max<int64_t>(s->n_slots, 1);
    a.part = s->part;
    a.ctrl = s->ctrl;
    a.kd_dt = (float)(s->damping / dt);
    StepParams p;
    p.dt = (float)dt;
    p.inv_dt = (float)(1.0 / dt);
    p.tol = (float)tol;
    p.kd_dt = a.kd_dt;
    p.max_iter = max_iter;
    p.skip_init = 0;
    VBD_TRY(ensure_u_last(s));
    p.u_last = s->u_last;
    if (s->material == kSNH)
        dispatch_resident<kSNH>(s->resident_lanes, a, s->d_ranges, s->ncolors, s->d_grav, s->mass, s->active,
                                s->uprev, s->vel, p, s->d_body_out, 1, 1024, s->stream);
    else
        dispatch_resident<kStVK>(s->resident_lanes, a, s->d_ranges, s->ncolors, s->d_grav, s->mass, s->active,
                                 s->uprev, s->vel, p, s->d_body_out, 1, 1024, s->stream);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipMemcpyAsync(s->body_out_host, s->d_body_out, sizeof(int2), hipMemcpyDeviceToHost, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    const int it = s->body_out_host[0].x;
    s->prev_iters = s->last_iters;
    s->last_iters = it;
    if (iters) *iters = it;
    if (metric) *metric = (double)__builtin_bit_cast(float, (uint32_t)s->body_out_host[0].y);
    return report_dx_max(s, it, max_iter);
}

namespace {
int step_device_run(vbd_solver *s, double dt, double tol, int max_iter, int *iters, double *metric)
{
    if (s->resident && !s->warm) return step_resident(s, dt, tol, max_iter, iters, metric);
    VBD_TRY(ensure_partials(s));
    const int n = s->n;
    s->kd_dt = (float)(s->damping / dt);
    const float4 g = make_float4((float)s->grav[0], (float)s->grav[1], (float)s->grav[2], 0.f);
    bool fused = false;
    if (s->warm)
        VBD_TRY(mpnn_init(s->warm, n, s->u, s->uprev, s->vel, s->tgt, s->mass, s->active, (float)dt, g, s->ctrl,
                          &fused));
    if (!fused)
        k_step_init<<<nblk(std::max(n, 1), 256), 256, 0, s->stream>>>(n, s->u, s->uprev, s->vel, s->tgt, s->mass,
                                                                       s->active, (float)dt, g, s->ctrl);
    VBD_HIP(hipGetLastError());
    if (s->warm) VBD_TRY(mpnn_apply_guess(s->warm, dt));
    // A tolerance is compared in FP32 on the device; tol = 0 never converges.
    const float ftol = (float)tol;
    int launched = 0;
    // first chunk: the larger of the last two steps' iteration counts, so a
    // step one iteration longer than the last needs no second host round trip
    int chunk = s->launch.chunk > 0 ? s->launch.chunk : std::max(1, std::max(s->last_iters, s->prev_iters));
    bool conv = false;
    // k iterations from iteration `from`: as graphs of 2^j iterations, largest
    // first (a few graphs per launch plan serve every chunk length, so a step
    // whose count has not been seen before captures nothing), or launch by launch
    auto enqueue = [&](int from, int k) -> int {
        if (s->launch.graph && !s->prof) {
            for (int rest = k; rest > 0;) {
                const int p = graph_piece(rest);
                hipGraphExec_t ex;
                VBD_TRY(single_graph(s, p, ftol, &ex));
                VBD_HIP(hipGraphLaunch(ex, s->stream));
                rest -= p;
            }
        } else {
            for (int j = 0; j < k; ++j) VBD_TRY(enqueue_iteration(s, from + j, ftol, s->prof));
        }
        return 0;
    };
    while (launched < max_iter && !conv) {
        const int k = std::min(chunk, max_iter - launched);
        if (launched + k == max_iter) {
            // the chunk runs the last allowed iteration: keep the iterate from
            // before it, for the largest |dx| of that iteration
            VBD_TRY(ensure_u_last(s));
            VBD_TRY(enqueue(launched, k - 1));
            if (s->n_swept > 0)
                VBD_HIP(hipMemcpyAsync(s->u_last, s->u, sizeof(float4) * s->n_swept, hipMemcpyDeviceToDevice,
                                       s->stream));
            VBD_TRY(enqueue(launched + k - 1, 1));
        } else {
            VBD_TRY(enqueue(launched, k));
        }
        launched += k;
        // the velocity update if this chunk completed the step, then one wait
        k_step_finish_if<<<nblk(n, 256), 256, 0, s->stream>>>(n, s->u, s->uprev, s->vel, (float)dt, s->ctrl,
                                                               max_iter);
        VBD_HIP(hipGetLastError());
        VBD_HIP(hipMemcpyAsync(s->ctrl_host, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
        VBD_HIP(hipStreamSynchronize(s->stream));
        conv = s->ctrl_host->converged != 0;
        chunk = s->launch.chunk > 0 ? s->launch.chunk : 2;
    }
    const int it = s->ctrl_host->iters;
    s->prev_iters = s->last_iters;
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
    return report_dx_max(s, it, max_iter);
}
}  // namespace

int step_device(vbd_solver *s, double dt, double tol, int max_iter, int *iters, double *metric)
{
    // the correction pass is deferred to the init (mpnn_init), which may fold it in
    if (s->warm) VBD_TRY(mpnn_before_init(s->warm, dt, true));
    const int rc = step_device_run(s, dt, tol, max_iter, iters, metric);
    s->hist_valid = rc == 0;
    s->hist_dt = dt;
    if (rc == 0) ++s->step_serial;
    return rc;
}

template <typename T>
static int export_positions_t(vbd_solver *s, const T *src, double *host_out)
{
    const int n = s->n;
    k_export<<<nblk(n, 256), 256, 0, s->stream>>>(n, src, s->iperm, s->X, s->xio);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipMemcpyAsync(s->xio_host, s->xio, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    std::copy(s->xio_host, s->xio_host + 3 * (size_t)n, host_out);
    return 0;
}

int import_positions(vbd_solver *s, const double *host_in, double4 *dst)
{
    const int n = s->n;
    std::copy(host_in, host_in + 3 * (size_t)n, s->xio_host);
    VBD_HIP(hipMemcpyAsync(s->xio, s->xio_host, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
    k_import<<<nblk(n, 256), 256, 0, s->stream>>>(n, s->xio, s->perm, s->X, dst);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

template <typename T>
static int export_positions_dev_t(vbd_solver *s, const T *src, double *dev_out)
{
    const int n = s->n;
    k_export<<<nblk(n, 256), 256, 0, s->stream>>>(n, src, s->iperm, s->X, dev_out);
    VBD_HIP(hipGetLastError());
    return 0;
}

int export_positions(vbd_solver *s, const double4 *src, double *host_out) { return export_positions_t(s, src, host_out); }
int export_positions(vbd_solver *s, const float4 *src, double *host_out) { return export_positions_t(s, src, host_out); }
int export_positions_dev(vbd_solver *s, const double4 *src, double *dev_out)
{
    return export_positions_dev_t(s, src, dev_out);
}
int export_positions_dev(vbd_solver *s, const float4 *src, double *dev_out)
{
    return export_positions_dev_t(s, src, dev_out);
}

int import_positions_dev(vbd_solver *s, const double *dev_in, double4 *dst)
{
    const int n = s->n;
    k_import<<<nblk(n, 256), 256, 0, s->stream>>>(n, dev_in, s->perm, s->X, dst);
    VBD_HIP(hipGetLastError());
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

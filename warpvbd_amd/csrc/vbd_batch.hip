// vbd_batch.hip -- many bodies of uneven size in one solver.
//
// The reference runs parallel environments by concatenating them into one
// mesh with one global convergence test (examples/run_mass_spring_parallel.py:
// 40-58,107-119).  Here the vertices are ordered colour-major, body-minor (a
// colour-major offset table: colour c of body 0, colour c of body 1, ...), each
// body keeps its own convergence state, and a body that has converged drops
// out of the remaining sweeps of the step.  Two execution forms:
//
//   colour launches  one launch per colour across every body; the first
//                    launch of an iteration also retires the bodies that
//                    converged in the iteration before (each vertex takes the
//                    same decision from the previous iteration's per-body
//                    maxima; extra waves, one per body, write the new state),
//                    and a run of iterations ends with one closing launch.
//                    Runs are captured in hipGraphs and replayed in chunks,
//                    the host waiting once per chunk on a device count of
//                    live bodies.  Rows of kStageRows vertices of a colour
//                    whose bodies have all converged are skipped by the
//                    launches (a live count per row, kept by whichever
//                    launch retires a body).
//   resident         one workgroup per body runs its whole step
//                    (k_step_resident), one launch per step
//
// The per-body result equals that body's single-body solve: colouring a
// disjoint union in body-major vertex order gives each body its own first-fit
// colouring, and the body's iterations stop on its own max|dx_x+dx_y+dx_z|.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "vbd_device.h"
#include "vbd_internal.h"
#include "vbd_mpnn.h"
#include "vbd_sweep.h"

namespace vbd {

namespace {

#ifndef VBD_FULL_CHUNK_MARGIN
#define VBD_FULL_CHUNK_MARGIN 1
#endif
// Iterations the first chunk of an uncompacted step adds to the last step's
// largest body count: an iteration after every body has converged costs only
// launches whose blocks leave at once, a chunk too short costs a wait on the
// host and another submission.
constexpr int kFullChunkMargin = VBD_FULL_CHUNK_MARGIN;

// The per-body bookkeeping of an iteration, done by extra waves (one per
// body) in the iteration's first colour launch: in an iteration after the
// first of a run (ph.fold) the body's new state from the previous
// iteration's state and maxima (the same decision its vertices take in that
// launch; a body already done is copied), and for a body that retires the
// live counts of its rows and the live-body count; in every iteration the
// maxima array of the iteration after next is cleared.  The launch's first
// nfold blocks do this (dispatched first, they run beside the colour's own
// blocks instead of after them).
struct BodyFold {
    BodyState *next = nullptr;
    uint32_t *mclear = nullptr;
    int32_t *n_live = nullptr;
    const int32_t *row_off = nullptr, *rows = nullptr;
    int32_t *row_live = nullptr;
    int nb = 0;
    int nfold = 0;
};

// A body's state after an iteration with maxima m (s: before it).
__device__ __forceinline__ BodyState body_after(const BodyState &s, uint32_t m, float tol, int max_iter, bool &retire)
{
    retire = false;
    if (s.done) return s;
    retire = body_retires(s, m, tol, max_iter);
    return BodyState{retire ? 1 : 0, s.iters + 1, m, 0};
}

__device__ __forceinline__ void retire_rows(int b, int lane, const int32_t *row_off, const int32_t *rows,
                                            int32_t *row_live)
{
    for (int i = row_off[b] + lane; i < row_off[b + 1]; i += 64) atomicSub(&row_live[rows[i]], 1);
}

__device__ void fold_bodies(const BodyPhase &ph, const BodyFold &f)
{
    const int b = (int)(((int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x) >> 6);
    const int lane = (int)(threadIdx.x & 63);
    if (b >= f.nb) return;
    if (lane == 0) f.mclear[b] = 0u;
    if (!ph.fold) return;
    // every lane reads the same state; lane 0 writes the new one
    bool retire;
    const BodyState o = body_after(ph.prev[b], ph.mprev[b], ph.tol, ph.max_iter, retire);
    if (lane == 0) {
        f.next[b] = o;
        if (retire) atomicSub(f.n_live, 1);
    }
    if (retire) retire_rows(b, lane, f.row_off, f.rows, f.row_live);
}

template <int L, int B, int MAT>
__global__ __launch_bounds__(1024) void k_sweep_bodies(SweepArgs a, const int32_t *body_new, BodyPhase ph,
                                                       BodyFold f, const int32_t *rows, int v0, int nv)
{
    if ((int)blockIdx.x < f.nfold) {
        fold_bodies(ph, f);
        return;
    }
    const int b = xcd_remap((int)blockIdx.x - f.nfold, (int)gridDim.x - f.nfold);
    const int t = b * blockDim.x + threadIdx.x;
    const int vi = t / L, lane = t % L;
    const bool in = vi < nv;
    const int v = v0 + (in ? vi : 0);
    // the vertex's own loads go out with the body lookup, not after it
    const VertexPre pre = vertex_pre(a, v, in);
    const int body = body_new[v];
    // A wave's 64 / L vertices lie in one row of kStageRows (rows start at
    // multiples of kStageRows in the colour), so its live count is a scalar,
    // and a wave of a row without live bodies leaves here.  The loads above
    // stay ahead of the test (sunk past it, they would wait for the count).
    const int live = rows[__builtin_amdgcn_readfirstlane(min(vi, nv - 1) / kStageRows)];
    asm volatile("" ::"v"(body), "s"(live) : "memory");
    if (live == 0) return;
    // read by every lane (v is a swept vertex for all), so no branch around the loads
    const bool valid = in && !body_done(ph, body);
    const uint32_t m = update_vertex<L, B, MAT>(a, v, valid, lane, pre);
    body_metric_max(valid ? body : -1, m, ph.macc);
}

// Colour sweep over a task list: block b updates the vertices
// [tasks[b].x, tasks[b].x + tasks[b].y) (at most blockDim.x / L of them) of
// bodies still live.  Used once most bodies have converged, so a launch
// covers only their ranges (one iteration per run: ph.fold is 0).
template <int L, int B, int MAT>
__global__ __launch_bounds__(1024) void k_sweep_tasks(SweepArgs a, const int32_t *body_new, BodyPhase ph,
                                                      const int2 *tasks)
{
    const int2 t = tasks[xcd_remap(blockIdx.x, gridDim.x)];
    const int vi = threadIdx.x / L, lane = threadIdx.x % L;
    const bool in = vi < t.y;
    const int v = t.x + (in ? vi : 0);
    // the vertex's own loads go out with the body lookup, not after it
    const VertexPre pre = vertex_pre(a, v, in);
    const int body = body_new[v];
    const bool valid = in && !body_done(ph, body);
    const uint32_t m = update_vertex<L, B, MAT>(a, v, valid, lane, pre);
    body_metric_max(valid ? body : -1, m, ph.macc);
}

// A staged colour (staged_sweep in vbd_sweep.h) across every body: 256
// threads per block of kStageRows consecutive vertices of the colour; block
// blist[i].x for the i-th block when a list is given (the blocks that still
// hold a live body's vertex), else all of them.  DL: per-block dictionary
// lists (staged_sweep).
template <int MAT, bool DL>
__global__ __launch_bounds__(256) void k_sweep_staged_bodies(SweepArgs a, int v0, int nv, LaneLayout lay,
                                                             StageTables st, int n_dict,
                                                             const int32_t *body_new, BodyPhase ph, BodyFold f,
                                                             const int32_t *rows, const int2 *blist, const int32_t *dl)
{
    if ((int)blockIdx.x < f.nfold) {
        fold_bodies(ph, f);
        return;
    }
    const int i = (int)blockIdx.x - f.nfold;
    const int b = blist ? blist[i].x : xcd_remap(i, (int)gridDim.x - f.nfold);
    uint32_t m;
    int body;
    // The block's maxima of its first kFoldBodies bodies gather in LDS and go
    // out as one device atomic each at the end: a body's vertices fill
    // consecutive blocks that finish together, and one atomic per wave on
    // the same word costs the launch about 4% (cleared here, before
    // staged_sweep's barrier).
    constexpr int kFoldBodies = 8;
    __shared__ uint32_t s_bm[kFoldBodies];
    if (threadIdx.x < kFoldBodies) s_bm[threadIdx.x] = 0u;
    const int b0 = body_new[v0 + b * kStageRows];
    // (the whole block returns together)
    if (!staged_sweep<MAT, true, false, DL>(a, v0, nv, lay, b, st, n_dict, body_new, &ph, rows, body, m, dl))
        return;
    body_metric_max(body, m, ph.macc, s_bm, b0, kFoldBodies);
    __syncthreads();
    if (threadIdx.x < kFoldBodies && s_bm[threadIdx.x]) atomicMax(&ph.macc[b0 + threadIdx.x], s_bm[threadIdx.x]);
}

// After a run of iterations (one wave per body): the bodies' state after its
// last iteration from the state before it (prev) and its maxima (mprev), into
// out, the state the next run starts from; the maxima array of the next run's
// first iteration (mclear) is cleared.  prev and out, mprev and mclear may be
// the same arrays (each body's entry is read before it is written).
__global__ __launch_bounds__(256) void k_close_bodies(int nb, const BodyState *prev, const uint32_t *mprev,
                                                      BodyState *out, uint32_t *mclear, float tol, int max_iter,
                                                      int32_t *n_live, const int32_t *row_off,
                                                      const int32_t *rows, int32_t *row_live)
{
    const int b = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = (int)(threadIdx.x & 63);
    if (b >= nb) return;
    bool retire;
    const BodyState o = body_after(prev[b], mprev[b], tol, max_iter, retire);
    if (lane == 0) {
        out[b] = o;
        mclear[b] = 0u;
        if (retire) atomicSub(n_live, 1);
    }
    if (retire) retire_rows(b, lane, row_off, rows, row_live);
}

__global__ void k_init_bodies(int n, float4 *u, const double4 *uprev, const float4 *vel, double4 *tgt, const float *mass,
                              const uint8_t *active, const int32_t *body_new, const float4 *grav, float dt, int nb,
                              BodyState *bs, uint32_t *bmet, int32_t *n_live, int nrows, const int32_t *row_live0,
                              int32_t *row_live)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) bs[i] = BodyState{0, 0, 0u, 0};
    if (i < 3 * nb) bmet[i] = 0u;
    if (i < nrows) row_live[i] = row_live0[i];
    if (i == 0) *n_live = nb;
    if (i < n) init_vertex(i, u, uprev, vel, tgt, mass, active, dt, grav[body_new[i]]);
}

// The velocity update once every body is done (else nothing): enqueued behind
// every chunk of iterations, so it follows the last chunk without a wait on
// the host.
__global__ void k_finish_all_if(int n, const float4 *u, double4 *uprev, float4 *vel, float dt, const int32_t *n_live)
{
    if (*n_live != 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 x = unpack_disp4(u[i]), p = uprev[i];
    vel[i] = step_velocity(x, p, dt);
    uprev[i] = x;
}

template <int L, int MAT>
void launch_sweep_bodies(const SweepArgs &a, const int32_t *body_new, const BodyPhase &ph, BodyFold f,
                         const int32_t *rows, int v0, int nv, int block, hipStream_t st)
{
    const int nmain = nblk((int64_t)nv * L, block);
    f.nfold = f.nb > 0 ? nblk((int64_t)f.nb * 64, block) : 0;
    k_sweep_bodies<L, Batch<L>::value, MAT><<<nmain + f.nfold, block, 0, st>>>(a, body_new, ph, f, rows, v0, nv);
}

template <int MAT>
void dispatch_sweep_bodies(int L, const SweepArgs &a, const int32_t *body_new, const BodyPhase &ph,
                           const BodyFold &f, const int32_t *rows, int v0, int nv, int block, hipStream_t st)
{
    switch (L) {
    case 1: launch_sweep_bodies<1, MAT>(a, body_new, ph, f, rows, v0, nv, block, st); break;
    case 2: launch_sweep_bodies<2, MAT>(a, body_new, ph, f, rows, v0, nv, block, st); break;
    case 8: launch_sweep_bodies<8, MAT>(a, body_new, ph, f, rows, v0, nv, block, st); break;
    case 16: launch_sweep_bodies<16, MAT>(a, body_new, ph, f, rows, v0, nv, block, st); break;
    case 32: launch_sweep_bodies<32, MAT>(a, body_new, ph, f, rows, v0, nv, block, st); break;
    default: launch_sweep_bodies<4, MAT>(a, body_new, ph, f, rows, v0, nv, block, st); break;
    }
}

template <int L, int MAT>
void launch_sweep_tasks(const SweepArgs &a, const int32_t *body_new, const BodyPhase &ph, const int2 *tasks,
                        int ntasks, int block, hipStream_t st)
{
    k_sweep_tasks<L, Batch<L>::value, MAT><<<ntasks, block, 0, st>>>(a, body_new, ph, tasks);
}

template <int MAT>
void dispatch_sweep_tasks(int L, const SweepArgs &a, const int32_t *body_new, const BodyPhase &ph,
                          const int2 *tasks, int ntasks, int block, hipStream_t st)
{
    switch (L) {
    case 1: launch_sweep_tasks<1, MAT>(a, body_new, ph, tasks, ntasks, block, st); break;
    case 2: launch_sweep_tasks<2, MAT>(a, body_new, ph, tasks, ntasks, block, st); break;
    case 8: launch_sweep_tasks<8, MAT>(a, body_new, ph, tasks, ntasks, block, st); break;
    case 16: launch_sweep_tasks<16, MAT>(a, body_new, ph, tasks, ntasks, block, st); break;
    case 32: launch_sweep_tasks<32, MAT>(a, body_new, ph, tasks, ntasks, block, st); break;
    default: launch_sweep_tasks<4, MAT>(a, body_new, ph, tasks, ntasks, block, st); break;
    }
}

// Lanes per vertex of colour c's launch: colours below launch.small_color
// vertices are latency-bound and get 32 (one slot per lane), the others the
// tuned count.
int color_lanes(const vbd_solver *s, int c)
{
    return s->col_size[c] < s->launch.small_color ? 32 : s->launch.lanes;
}

// The staged kernel for colour c when the solver has staging tables for it.
bool staged_color(const vbd_solver *s, int c)
{
    return s->st_ok && c < (int)s->st_col.size() && s->st_col[c].ids >= 0 && s->st_col[c].b1 > 0;
}

void launch_staged_bodies(vbd_solver *s, const SweepArgs &a, int c, const int2 *blist, int nblocks,
                          const BodyPhase &ph, BodyFold f)
{
    const int v0 = s->col_start[c], nv = s->col_size[c];
    const int nmain = blist ? nblocks : nblk(nv, kStageRows);
    f.nfold = f.nb > 0 ? nblk((int64_t)f.nb * 64, 256) : 0;
    if (nmain + f.nfold == 0) return;
    const dim3 grid(nmain + f.nfold);
    const int32_t *rows = s->row_live + s->row_base[c];
    const StageTables st{s->st_nb, s->st_rec, s->st_vrec};
    const bool dl = s->st_dl != nullptr;
    auto kern = s->material == kSNH ? (dl ? k_sweep_staged_bodies<kSNH, true> : k_sweep_staged_bodies<kSNH, false>)
                                    : (dl ? k_sweep_staged_bodies<kStVK, true> : k_sweep_staged_bodies<kStVK, false>);
    kern<<<grid, 256, 0, s->stream>>>(a, v0, nv, s->st_col[c], st, s->n_dict, s->body_new, ph, f, rows, blist,
                                      s->st_dl);
}

// Per-body state and maxima of iteration k of a run: two state arrays
// (iteration k reads array k % 2) and three maxima arrays (k % 3), so that
// an iteration's first launch can retire bodies from the previous
// iteration's arrays while writing the next ones.  Array 0 holds the state a
// run starts from and the host reads.
BodyState *body_state(vbd_solver *s, int k) { return s->bstate + (size_t)(k & 1) * s->n_bodies; }
uint32_t *body_maxima(vbd_solver *s, int k) { return s->dxm + (size_t)(k % 3) * s->n_bodies; }

BodyPhase body_phase(vbd_solver *s, int k, float tol, int max_iter)
{
    BodyPhase ph;
    ph.cur = body_state(s, k);
    ph.macc = body_maxima(s, k);
    ph.tol = tol;
    ph.max_iter = max_iter;
    if (k > 0) {
        ph.prev = body_state(s, k - 1);
        ph.mprev = body_maxima(s, k - 1);
        ph.fold = 1;
    }
    return ph;
}

BodyFold body_fold(vbd_solver *s, int k)
{
    BodyFold f;
    f.next = body_state(s, k);
    f.mclear = body_maxima(s, k + 1);
    f.n_live = s->n_live;
    f.row_off = s->body_row_off;
    f.rows = s->body_rows;
    f.row_live = s->row_live;
    f.nb = s->n_bodies;
    return f;
}

// Close a run of k iterations: the state after its last one into array 0.
void close_run(vbd_solver *s, int k, float tol, int max_iter)
{
    k_close_bodies<<<nblk((int64_t)s->n_bodies * 64, 256), 256, 0, s->stream>>>(
        s->n_bodies, body_state(s, k - 1), body_maxima(s, k - 1), body_state(s, 0), body_maxima(s, 0), tol, max_iter,
        s->n_live, s->body_row_off, s->body_rows, s->row_live);
}

SweepArgs sweep_args(vbd_solver *s, float kd_dt)
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
    a.kd_dt = kd_dt;
    return a;
}

int ensure_batch_buffers(vbd_solver *s)
{
    if (s->bstate) return 0;
    VBD_HIP(hipMalloc((void **)&s->bstate, sizeof(BodyState) * 2 * s->n_bodies));
    VBD_HIP(hipMalloc((void **)&s->dxm, sizeof(uint32_t) * 3 * s->n_bodies));
    VBD_HIP(hipMemset(s->dxm, 0, sizeof(uint32_t) * 3 * s->n_bodies));
    // rows of kStageRows vertices per colour, and per body the rows that hold
    // its vertices (a body's colour range is contiguous)
    const int C = s->ncolors, nb = s->n_bodies;
    s->row_base.assign(C + 1, 0);
    for (int c = 0; c < C; ++c) s->row_base[c + 1] = s->row_base[c] + nblk(s->col_size[c], kStageRows);
    s->n_rows = s->row_base[C];
    std::vector<int32_t> live0(std::max(s->n_rows, 1), 0), off(nb + 1, 0), rows;
    for (int b = 0; b < nb; ++b) {
        for (int c = 0; c < C; ++c) {
            const int2 r = s->ranges_host[(size_t)b * (C + 1) + c];
            if (r.y <= 0) continue;
            const int j0 = (r.x - s->col_start[c]) / kStageRows, j1 = (r.x + r.y - 1 - s->col_start[c]) / kStageRows;
            for (int j = j0; j <= j1; ++j) {
                rows.push_back(s->row_base[c] + j);
                ++live0[s->row_base[c] + j];
            }
        }
        off[b + 1] = (int32_t)rows.size();
    }
    rows.resize(std::max<size_t>(rows.size(), 1), 0);
    VBD_HIP(hipMalloc((void **)&s->row_live, sizeof(int32_t) * live0.size()));
    VBD_HIP(hipMalloc((void **)&s->row_live0, sizeof(int32_t) * live0.size()));
    VBD_HIP(hipMalloc((void **)&s->body_row_off, sizeof(int32_t) * off.size()));
    VBD_HIP(hipMalloc((void **)&s->body_rows, sizeof(int32_t) * rows.size()));
    VBD_HIP(hipMemcpy(s->row_live0, live0.data(), sizeof(int32_t) * live0.size(), hipMemcpyHostToDevice));
    VBD_HIP(hipMemcpy(s->row_live, live0.data(), sizeof(int32_t) * live0.size(), hipMemcpyHostToDevice));
    VBD_HIP(hipMemcpy(s->body_row_off, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice));
    VBD_HIP(hipMemcpy(s->body_rows, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice));

    VBD_HIP(hipMalloc((void **)&s->n_live, sizeof(int32_t)));
    VBD_HIP(hipHostMalloc((void **)&s->n_live_host, sizeof(int32_t), hipHostMallocDefault));
    VBD_HIP(hipHostMalloc((void **)&s->bstate_host, sizeof(BodyState) * s->n_bodies, hipHostMallocDefault));
    return 0;
}

// A run of k iterations of every live body: per iteration one launch per
// colour, the first carrying the per-body bookkeeping of the iteration before
// (extra waves), then one closing launch for the run.
int enqueue_body_run(vbd_solver *s, int k, float kd_dt, float tol, int max_iter)
{
    const SweepArgs a = sweep_args(s, kd_dt);
    const int block = s->launch.block;
    for (int it = 0; it < k; ++it) {
        // the first launch decides the retirements (fold), the others read
        // the state its extra waves wrote
        const BodyPhase ph0 = body_phase(s, it, tol, max_iter);
        BodyPhase rest = ph0;
        rest.fold = 0;
        bool first = true;
        for (int c = 0; c < s->ncolors; ++c) {
            const int nv = s->col_size[c];
            if (!nv) continue;
            const BodyFold f = first ? body_fold(s, it) : BodyFold{};
            const BodyPhase &ph = first ? ph0 : rest;
            first = false;
            const int L = color_lanes(s, c);
            const int32_t *rows = s->row_live + s->row_base[c];
            if (staged_color(s, c))
                launch_staged_bodies(s, a, c, nullptr, 0, ph, f);
            else if (s->material == kSNH)
                dispatch_sweep_bodies<kSNH>(L, a, s->body_new, ph, f, rows, s->col_start[c], nv, block, s->stream);
            else
                dispatch_sweep_bodies<kStVK>(L, a, s->body_new, ph, f, rows, s->col_start[c], nv, block,
                                             s->stream);
        }
    }
    close_run(s, k, tol, max_iter);
    VBD_HIP(hipGetLastError());
    return 0;
}

// Task lists for the bodies not yet done (bstate_host must be current): per
// colour, the bodies' colour ranges with adjacent ranges merged, cut into
// blocks of block / L vertices (L the colour's lanes per vertex).
int build_tasks(vbd_solver *s)
{
    const int C = s->ncolors, nb = s->n_bodies;
    auto vpb_of = [&](int c) { return std::max(1, s->launch.block / color_lanes(s, c)); };
    s->tasks_host.clear();
    s->task_off.assign(C + 1, 0);
    s->task_vertices = 0;
    for (int c = 0; c < C; ++c) {
        s->task_off[c] = (int)s->tasks_host.size();
        if (staged_color(s, c)) {
            // the colour's staging blocks that hold a live body's vertex, as
            // (block index, 0) entries
            const int c0 = s->col_start[c], nblocks = nblk(s->col_size[c], kStageRows);
            int last = -1;
            for (int b = 0; b < nb; ++b) {
                if (s->bstate_host[b].done) continue;
                const int2 r = s->ranges_host[(size_t)b * (C + 1) + c];
                if (r.y <= 0) continue;
                const int j0 = std::max(last + 1, (r.x - c0) / kStageRows);
                const int j1 = std::min(nblocks - 1, (r.x + r.y - 1 - c0) / kStageRows);
                for (int j = j0; j <= j1; ++j) s->tasks_host.push_back(make_int2(j, 0));
                last = std::max(last, j1);
                s->task_vertices += r.y;
            }
            continue;
        }
        const int vpb = vpb_of(c);
        int run0 = -1, run1 = -1;
        auto flush = [&]() {
            for (int v = run0; v < run1; v += vpb) s->tasks_host.push_back(make_int2(v, std::min(vpb, run1 - v)));
            s->task_vertices += run1 - run0;
        };
        for (int b = 0; b < nb; ++b) {
            if (s->bstate_host[b].done) continue;
            const int2 r = s->ranges_host[(size_t)b * (C + 1) + c];
            if (r.y <= 0) continue;
            if (r.x == run1) {
                run1 += r.y;
            } else {
                if (run1 > run0) flush();
                run0 = r.x;
                run1 = r.x + r.y;
            }
        }
        if (run1 > run0) flush();
    }
    s->task_off[C] = (int)s->tasks_host.size();
    // sized once for any list (a reallocation in the middle of a run stalls
    // it: hipFree waits for the device): a colour's list has at most one
    // entry per vpb of its vertices plus one per body, or one per staging block
    size_t nt = std::max<size_t>(1, s->tasks_host.size());
    if (nt > s->task_cap) {
        size_t cap = 16;
        for (int c = 0; c < C; ++c) cap += (size_t)s->col_size[c] / vpb_of(c) + nb + 1;
        nt = std::max(nt, cap);
        if (s->d_tasks) VBD_HIP(hipFree(s->d_tasks));
        s->d_tasks = nullptr;
        VBD_HIP(hipMalloc((void **)&s->d_tasks, sizeof(int2) * nt));
        s->task_cap = nt;
    }
    if (!s->tasks_host.empty())
        VBD_HIP(hipMemcpyAsync(s->d_tasks, s->tasks_host.data(), sizeof(int2) * s->tasks_host.size(),
                               hipMemcpyHostToDevice, s->stream));
    return 0;
}

// One iteration of the live bodies through the current task lists (a run
// of one iteration).
int enqueue_task_iteration(vbd_solver *s, float kd_dt, float tol, int max_iter)
{
    const SweepArgs a = sweep_args(s, kd_dt);
    const int block = s->launch.block;
    const BodyPhase ph = body_phase(s, 0, tol, max_iter);
    for (int c = 0; c < s->ncolors; ++c) {
        const int t0 = s->task_off[c], nt = s->task_off[c + 1] - t0;
        if (!nt) continue;
        const int L = color_lanes(s, c);
        if (staged_color(s, c))
            launch_staged_bodies(s, a, c, s->d_tasks + t0, nt, ph, BodyFold{});
        else if (s->material == kSNH)
            dispatch_sweep_tasks<kSNH>(L, a, s->body_new, ph, s->d_tasks + t0, nt, block, s->stream);
        else
            dispatch_sweep_tasks<kStVK>(L, a, s->body_new, ph, s->d_tasks + t0, nt, block, s->stream);
    }
    close_run(s, 1, tol, max_iter);
    VBD_HIP(hipGetLastError());
    return 0;
}

void drop_graphs(vbd_solver *s)
{
    for (auto &g : s->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    s->graphs.clear();
}

// A graph of k iterations for the current (lanes, block, tol, max_iter, kd).
int get_graph(vbd_solver *s, int k, float kd_dt, float tol, int max_iter, hipGraphExec_t *out)
{
    for (auto &g : s->graphs)
        if (g.k == k && g.tol == tol && g.max_iter == max_iter && g.kd_dt == kd_dt && g.lanes == s->launch.lanes &&
            g.block == s->launch.block) {
            graph_touch(s, g);
            *out = g.exec;
            return 0;
        }
    graph_make_room(s);
    CachedGraph g;
    g.k = k;
    g.tol = tol;
    g.max_iter = max_iter;
    g.kd_dt = kd_dt;
    g.lanes = s->launch.lanes;
    g.block = s->launch.block;
    VBD_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_body_run(s, k, kd_dt, tol, max_iter);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &graph);
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

int step_bodies_launches(vbd_solver *s, double dt, double tol, int max_iter)
{
    VBD_TRY(ensure_batch_buffers(s));
    const float kd_dt = (float)(s->damping / dt);
    const float ftol = (float)tol;
    k_init_bodies<<<nblk(std::max({s->n, 3 * s->n_bodies, s->n_rows}), 256), 256, 0, s->stream>>>(
        s->n, s->u, s->uprev, s->vel, s->tgt, s->mass, s->active, s->body_new, s->d_grav, (float)dt, s->n_bodies,
        s->bstate, s->dxm, s->n_live, s->n_rows, s->row_live0, s->row_live);
    VBD_HIP(hipGetLastError());
    if (s->warm) VBD_TRY(mpnn_apply_guess(s->warm, dt));
    // First chunk: every body, replayed from a graph (rows whose bodies have
    // all converged are skipped on the device).  Without compaction its length
    // is the previous step's largest body iteration count plus
    // kFullChunkMargin, so that a step seldom waits on the host more than
    // once; with compaction it is the median count, after which about half the
    // bodies are done and the remaining chunks launch only the live bodies'
    // ranges (lists rebuilt whenever the live vertices fall below half of what
    // the current list covers).
    const bool compact = s->launch.compact && s->n_bodies > 1;
    const int prev_max = s->last_iters;
    int launched = 0;
    int chunk = std::max(1, compact ? s->median_iters : s->last_iters + kFullChunkMargin);
    bool tasks = false;
    while (launched < max_iter) {
        const int k = std::min(chunk, max_iter - launched);
        if (tasks) {
            for (int j = 0; j < k; ++j) VBD_TRY(enqueue_task_iteration(s, kd_dt, ftol, max_iter));
        } else if (s->launch.graph) {
            // as graphs of 2^j iterations, largest first (no capture for a new count)
            for (int rest = k; rest > 0;) {
                const int p = graph_piece(rest);
                hipGraphExec_t ex;
                VBD_TRY(get_graph(s, p, kd_dt, ftol, max_iter, &ex));
                VBD_HIP(hipGraphLaunch(ex, s->stream));
                rest -= p;
            }
        } else {
            VBD_TRY(enqueue_body_run(s, k, kd_dt, ftol, max_iter));
        }
        launched += k;
        // the velocity update if every body is done, then the live count and
        // every body's state after the chunk, in one wait (after the last
        // chunk that state is the step's result)
        k_finish_all_if<<<nblk(s->n, 256), 256, 0, s->stream>>>(s->n, s->u, s->uprev, s->vel, (float)dt,
                                                                 s->n_live);
        VBD_HIP(hipGetLastError());
        VBD_HIP(hipMemcpyAsync(s->n_live_host, s->n_live, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
        VBD_HIP(hipMemcpyAsync(s->bstate_host, s->bstate, sizeof(BodyState) * s->n_bodies, hipMemcpyDeviceToHost,
                               s->stream));
        VBD_HIP(hipStreamSynchronize(s->stream));
        if (*s->n_live_host == 0) break;
        if (compact) {
            int64_t live_v = 0;
            for (int b = 0; b < s->n_bodies; ++b)
                if (!s->bstate_host[b].done) live_v += s->body_voff_host[b + 1] - s->body_voff_host[b];
            const int64_t covered = tasks ? s->task_vertices : (int64_t)s->n;
            if (2 * live_v < covered) {
                VBD_TRY(build_tasks(s));
                tasks = true;
            }
        }
        // the rest of the step in one more chunk if it needs as many
        // iterations as the last step's slowest body did
        chunk = s->launch.chunk > 0 ? s->launch.chunk : tasks ? std::max(2, prev_max - launched) : 2;
    }
    int mx = 0;
    std::vector<int> its(s->n_bodies);
    for (int b = 0; b < s->n_bodies; ++b) {
        s->body_out_host[b] = make_int2(s->bstate_host[b].iters, (int)s->bstate_host[b].metric);
        mx = std::max(mx, s->bstate_host[b].iters);
        its[b] = s->bstate_host[b].iters;
    }
    s->last_iters = mx;
    if (!its.empty()) {
        std::nth_element(its.begin(), its.begin() + its.size() / 2, its.end());
        s->median_iters = its[its.size() / 2];
    }
    return 0;
}

int step_bodies_resident(vbd_solver *s, double dt, double tol, int max_iter)
{
    SweepArgs a = sweep_args(s, (float)(s->damping / dt));
    StepParams p;
    p.dt = (float)dt;
    p.inv_dt = (float)(1.0 / dt);
    p.tol = (float)tol;
    p.kd_dt = a.kd_dt;
    p.max_iter = max_iter;
    p.skip_init = 0;
    if (s->warm) {
        // initial guess outside the resident kernel: plain guess, then the network's where it wins
        VBD_TRY(ensure_batch_buffers(s));
        k_init_bodies<<<nblk(std::max({s->n, 3 * s->n_bodies, s->n_rows}), 256), 256, 0, s->stream>>>(
            s->n, s->u, s->uprev, s->vel, s->tgt, s->mass, s->active, s->body_new, s->d_grav, (float)dt,
            s->n_bodies, s->bstate, s->dxm, s->n_live, s->n_rows, s->row_live0, s->row_live);
        VBD_HIP(hipGetLastError());
        VBD_TRY(mpnn_apply_guess(s->warm, dt));
        p.skip_init = 1;
    }
    const int L = s->launch.lanes, block = s->launch.resident_block;
    if (s->material == kSNH)
        dispatch_resident<kSNH>(L, a, s->d_ranges, s->ncolors, s->d_grav, s->mass, s->active, s->uprev, s->vel, p,
                                s->d_body_out, s->n_bodies, block, s->stream);
    else
        dispatch_resident<kStVK>(L, a, s->d_ranges, s->ncolors, s->d_grav, s->mass, s->active, s->uprev, s->vel, p,
                                 s->d_body_out, s->n_bodies, block, s->stream);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipMemcpyAsync(s->body_out_host, s->d_body_out, sizeof(int2) * s->n_bodies, hipMemcpyDeviceToHost,
                           s->stream));
    VBD_HIP(hipStreamSynchronize(s->stream));
    int mx = 0;
    for (int b = 0; b < s->n_bodies; ++b) mx = std::max(mx, s->body_out_host[b].x);
    s->last_iters = mx;
    return 0;
}

}  // namespace

int step_bodies(vbd_solver *s, double dt, double tol, int max_iter)
{
    if (s->warm) VBD_TRY(mpnn_before_init(s->warm, dt));
    const int rc = s->batch_mode == kBatchResident ? step_bodies_resident(s, dt, tol, max_iter)
                                                   : step_bodies_launches(s, dt, tol, max_iter);
    s->hist_valid = rc == 0;
    s->hist_dt = dt;
    if (rc == 0) ++s->step_serial;
    return rc;
}

void release_batch(vbd_solver *s)
{
    drop_graphs(s);
    if (s->d_tasks) (void)hipFree(s->d_tasks);
    s->d_tasks = nullptr;
    s->task_cap = 0;
    if (s->bstate) (void)hipFree(s->bstate);
    if (s->dxm) (void)hipFree(s->dxm);

    for (int32_t **p : {&s->row_live, &s->row_live0, &s->body_row_off, &s->body_rows}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    s->n_rows = 0;
    if (s->n_live) (void)hipFree(s->n_live);
    if (s->n_live_host) (void)hipHostFree(s->n_live_host);
    if (s->bstate_host) (void)hipHostFree(s->bstate_host);
    s->bstate = nullptr;
    s->dxm = nullptr;
    s->n_live = nullptr;
    s->n_live_host = nullptr;
    s->bstate_host = nullptr;
}

// Time each candidate execution form on the solver's own state (saved and
// restored around the trials) and keep the fastest.
int autotune_bodies(vbd_solver *s, double dt, double tol, int max_iter, int trial_steps, double *best_ms)
{
    const int n = s->n;
    double4 *save_u = nullptr, *save_t = nullptr;
    float4 *save_v = nullptr;
    VBD_HIP(hipMalloc((void **)&save_u, sizeof(double4) * n));
    VBD_HIP(hipMalloc((void **)&save_v, sizeof(float4) * n));
    if (hipMalloc((void **)&save_t, sizeof(double4) * n) != hipSuccess) {
        (void)hipFree(save_u);
        (void)hipFree(save_v);
        set_error("hipMalloc failed (autotune)");
        return VBD_ENOMEM;
    }
    VBD_HIP(hipMemcpyAsync(save_u, s->uprev, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
    VBD_HIP(hipMemcpyAsync(save_v, s->vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, s->stream));
    VBD_HIP(hipMemcpyAsync(save_t, s->tgt, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
    const bool keep_hist = s->hist_valid;
    const double keep_hist_dt = s->hist_dt;
    const int64_t keep_steps = s->step_serial;
    // the last step's iteration counts size the first chunk of the next
    const int keep_last = s->last_iters, keep_median = s->median_iters;
    struct Cand {
        int mode, lanes, block;
        bool compact;
    };
    std::vector<Cand> cands;
    // (1 and 16 lanes per vertex never won on the voxel bodies measured)
    for (bool compact : {true, false})
        for (int L : {2, 4, 8})
            for (int blk : {128, 256, 512}) cands.push_back({kBatchLaunches, L, blk, compact});
    for (int L : {2, 4, 8})
        for (int blk : {512, 1024}) cands.push_back({kBatchResident, L, blk, true});
    double best = 1e300;
    Cand bc = {s->batch_mode, s->launch.lanes, s->launch.block, s->launch.compact};
    const Launch keep = s->launch;
    const int keep_mode = s->batch_mode;
    int rc = 0;
    for (const Cand &c : cands) {
        s->batch_mode = c.mode;
        s->launch.lanes = c.lanes;
        s->launch.compact = c.compact;
        if (c.mode == kBatchResident) s->launch.resident_block = c.block;
        else s->launch.block = c.block;
        // restore the start state and run the trial steps twice, timing the
        // second pass: the steps repeat exactly, so every graph of iterations
        // the timed pass replays was captured in the first
        for (int rep = 0; rep < 2 && rc == 0; ++rep) {
            (void)hipMemcpyAsync(s->uprev, save_u, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream);
            (void)hipMemcpyAsync(s->vel, save_v, sizeof(float4) * n, hipMemcpyDeviceToDevice, s->stream);
            s->last_iters = keep_last;
            s->median_iters = keep_median;
            const int steps = trial_steps;
            auto t0 = std::chrono::steady_clock::now();
            for (int k = 0; k < steps && rc == 0; ++k) rc = step_bodies(s, dt, tol, max_iter);
            auto t1 = std::chrono::steady_clock::now();
            const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
            if (rep == 1 && rc == 0 && ms < best) {
                best = ms;
                bc = c;
            }
        }
        if (rc != 0) break;
    }
    // the trials' graphs of the other forms would crowd the cache
    for (auto it = s->graphs.begin(); it != s->graphs.end();) {
        if (bc.mode == kBatchLaunches && it->lanes == bc.lanes && it->block == bc.block) {
            ++it;
            continue;
        }
        (void)hipGraphExecDestroy(it->exec);
        (void)hipGraphDestroy(it->graph);
        it = s->graphs.erase(it);
    }
    s->launch = keep;
    s->batch_mode = bc.mode;
    s->launch.lanes = bc.lanes;
    s->launch.compact = bc.compact;
    if (bc.mode == kBatchResident) s->launch.resident_block = bc.block;
    else s->launch.block = bc.block;
    if (best >= 1e300) {
        s->batch_mode = keep_mode;
        s->launch = keep;
    }
    (void)hipMemcpyAsync(s->uprev, save_u, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream);
    (void)hipMemcpyAsync(s->vel, save_v, sizeof(float4) * n, hipMemcpyDeviceToDevice, s->stream);
    (void)hipMemcpyAsync(s->tgt, save_t, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream);
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(save_u);
    (void)hipFree(save_v);
    (void)hipFree(save_t);
    s->hist_valid = keep_hist;
    s->hist_dt = keep_hist_dt;
    s->step_serial = keep_steps;
    s->last_iters = keep_last;
    s->median_iters = keep_median;
    if (best_ms) *best_ms = best;
    return rc;
}

}  // namespace vbd

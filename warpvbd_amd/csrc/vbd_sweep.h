// vbd_sweep.h -- vertex update shared by the per-colour, fused, resident and
// batched kernels (included by vbd_step.hip and vbd_batch.hip).
#pragma once
#include "vbd_device.h"
#include "vbd_internal.h"

namespace vbd {
namespace {

inline int nblk(int64_t n, int tb) { return (int)((n + tb - 1) / tb); }

#ifndef VBD_BATCH_L4
#define VBD_BATCH_L4 1
#endif
#ifndef VBD_BATCH_L8
#define VBD_BATCH_L8 1
#endif

struct SweepArgs {
    float4 *__restrict__ u;  // packed displacements (pack_disp)
    const double4 *__restrict__ tgt;
    const double4 *__restrict__ uprev;
    const int32_t *__restrict__ slot_off;
    const float4 *__restrict__ slots;   // compact slot records (ids, entry), or tet records (VBD_TET_RECORDS)
    const float4 *__restrict__ dict_ext;  // rest-edge constants of the dictionary's entries
    const float4 *__restrict__ dict;    // record dictionary
    const int32_t *__restrict__ ent;    // per-slot tet << 2 | local index (VBD_TET_RECORDS)
    int64_t ns;
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

// The block of a lane layout that hardware block i sweeps.  Blocks are
// dispatched round-robin over the 8 XCDs (hardware block i on XCD i % 8, as
// its (i / 8)-th).  Each XCD takes one contiguous share of the blocks holding
// 1- or 2-lane vertices (from lay.b4 on: heavier staging lists, or unstaged),
// as even as the count allows, and sweeps it first, then one contiguous share
// of the 4-lane blocks, the rest of its blocks: every XCD gets its part of
// the heavier blocks, and its L2 sees compact neighbourhoods.
__device__ __forceinline__ int lane_block(int i, const LaneLayout &lay)
{
    const int nb = lay.nblocks, nh = nb - lay.b4;
    const int x = i & 7, j = i >> 3;
    const int hq = nh >> 3, hr = nh & 7, cq = nb >> 3, cr = nb & 7;
    const int hx = hq + (x < hr);               // this XCD's share of the heavier blocks
    const int h0 = x * hq + min(x, hr);         // ... starting here
    if (j < hx) return lay.b4 + h0 + j;
    const int f0 = x * cq + min(x, cr) - h0;    // its 4-lane share starts after the other XCDs' shares
    return f0 + (j - hx);
}

// &base[i] for a 32-bit unsigned element index (arrays here hold fewer than
// 2^26 float4: create rejects larger meshes), formed as base + 16 i in 32 bits
// so that a load can take the base from scalar registers and the offset from
// one vector register.
__device__ __forceinline__ const float4 *at(const float4 *base, uint32_t i)
{
    return reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + (uint32_t)(i << 4));
}

// Packed displacement, one float4 per vertex (16 bytes, one load per gather):
//   xyz  the FP32 values nearest to the displacement's components
//   w    bits 0-7, 8-15, 16-23: the remainders x - xyz of x, y, z as integers
//        in [-128, 127] (stored + 128) times a quantum q = 2^(E-31), E the
//        exponent of the largest |xyz|; bits 24-31: q's biased FP32 exponent
//        (0: q = 0, used when that largest |xyz| is below 2^-95)
// so the displacement is held to 2^-8 of an FP32 ulp of its largest component
// (32 significant bits).  The remainder of each component is at most half its
// own ulp <= 128 q, so only an exact half-ulp remainder of the largest one is
// clamped (by q).  All-zero bits are the zero displacement.
__device__ __forceinline__ float4 pack_disp(double x, double y, double z)
{
    const float hx = (float)x, hy = (float)y, hz = (float)z;
    const float m = fmaxf(fabsf(hx), fmaxf(fabsf(hy), fabsf(hz)));
    const int em = (int)(__float_as_uint(m) >> 23);  // biased exponent of m
    const int eb = em > 31 ? em - 31 : 0;            // biased exponent of q
    uint32_t w = 0x808080u;
    if (eb) {
        // 1/q = 2^(127 - eb), as an FP64 power of two
        const double iq = __longlong_as_double((long long)(1023 + 127 - eb) << 52);
        const int qx = min(127, max(-128, __double2int_rn((x - (double)hx) * iq)));
        const int qy = min(127, max(-128, __double2int_rn((y - (double)hy) * iq)));
        const int qz = min(127, max(-128, __double2int_rn((z - (double)hz) * iq)));
        w = (uint32_t)(qx + 128) | (uint32_t)(qy + 128) << 8 | (uint32_t)(qz + 128) << 16;
    }
    return make_float4(hx, hy, hz, __uint_as_float(w | (uint32_t)eb << 24));
}

// the remainders of a packed displacement (exact in FP32)
__device__ __forceinline__ float3 disp_lo(const float4 &p)
{
    const uint32_t w = __float_as_uint(p.w);
    const float q = __uint_as_float((w >> 24) << 23), nq = -128.f * q;
    return make_float3(fmaf((float)(w & 0xFFu), q, nq), fmaf((float)((w >> 8) & 0xFFu), q, nq),
                       fmaf((float)((w >> 16) & 0xFFu), q, nq));
}

// the displacement in FP64 (exact)
__device__ __forceinline__ double3 unpack_disp(const float4 &p)
{
    const float3 lo = disp_lo(p);
    return make_double3((double)p.x + (double)lo.x, (double)p.y + (double)lo.y, (double)p.z + (double)lo.z);
}

__device__ __forceinline__ double4 unpack_disp4(const float4 &p)
{
    const double3 d = unpack_disp(p);
    return make_double4(d.x, d.y, d.z, 0.0);
}

// The differences slot_eval takes, from the packed displacements x0, x1, x2
// of a slot's other vertices and the swept vertex's FP32 parts hi and
// remainders lo: with d_k = u_ok - u_self (the FP32 parts and the remainders
// differenced separately, so d_k carries the rounding of one FP32 operation
// on the difference, not that of the stored values), e0 = d_0 - d_2,
// e1 = d_1 - d_2 and d2 = d_2.  A vertex's d_k depends only on the vertex
// and the neighbour, so the staged sweep forms it once per (vertex,
// neighbour) pair (neighbour_diff) and every slot reads it.
struct SlotRel {
    float3 e0, e1, d2;
};

__device__ __forceinline__ float3 neighbour_diff(const float4 &x, const float3 &lx, const float3 &hi, const float3 &lo)
{
    return make_float3((x.x - hi.x) + (lx.x - lo.x), (x.y - hi.y) + (lx.y - lo.y), (x.z - hi.z) + (lx.z - lo.z));
}

__device__ __forceinline__ SlotRel slot_rel_d(const float3 &d0, const float3 &d1, const float3 &d2)
{
    SlotRel r;
    r.e0 = make_float3(d0.x - d2.x, d0.y - d2.y, d0.z - d2.z);
    r.e1 = make_float3(d1.x - d2.x, d1.y - d2.y, d1.z - d2.z);
    r.d2 = d2;
    return r;
}

__device__ __forceinline__ SlotRel slot_rel(const float4 &x0, const float4 &x1, const float4 &x2, const float3 &hi,
                                            const float3 &lo)
{
    return slot_rel_d(neighbour_diff(x0, disp_lo(x0), hi, lo), neighbour_diff(x1, disp_lo(x1), hi, lo),
                      neighbour_diff(x2, disp_lo(x2), hi, lo));
}

// A vertex's own data, loaded before anything that depends on other loads so
// that a kernel can request it ahead of its early-exit test.
struct VertexPre {
    int s0, s1;  // slot range (empty for lanes without a vertex)
    float4 uv;   // current displacement (packed)
    double4 tg;  // inertia target, M/dt^2
};

__device__ __forceinline__ VertexPre vertex_pre(const SweepArgs &a, int v, bool valid)
{
    VertexPre p;
    p.uv = a.u[v];
    p.tg = a.tgt[v];
    // v is a swept vertex for every lane (lanes without one use the range's first)
    p.s0 = a.slot_off[v];
    const int e = a.slot_off[v + 1];
    p.s1 = valid ? e : p.s0;
    return p;
}

// The swept vertex's own terms, formed before its slot loop: its packed
// displacement split into FP32 parts (hi) and remainders (lo), and the
// inertia term's M/dt^2 (w) and x - t (t0..t2), formed in FP64 and rounded
// once, so the loop carries four FP32 values instead of the FP64 target.
struct VertexLocal {
    float3 hi, lo;
    float w, t0, t1, t2;
};

__device__ __forceinline__ VertexLocal vertex_local(const VertexPre &pre)
{
    VertexLocal q;
    const float4 uv4 = pre.uv;
    q.hi = make_float3(uv4.x, uv4.y, uv4.z);
    q.lo = disp_lo(uv4);
    const double4 tg = pre.tg;
    const double3 u0 = unpack_disp(uv4);
    q.w = (float)tg.w;
    q.t0 = (float)(u0.x - tg.x);
    q.t1 = (float)(u0.y - tg.y);
    q.t2 = (float)(u0.z - tg.z);
    // an empty asm that reads and writes them keeps the compiler from sinking
    // the subtraction (and the FP64 target with it) past the loop
    asm volatile("" : "+v"(q.t0), "+v"(q.t1), "+v"(q.t2), "+v"(q.w));
    return q;
}

// The vertex's current displacement in FP64 from the parts the loop holds
// (the empty asm keeps the compiler from carrying the FP64 displacement
// through the loop instead of rebuilding it here).
__device__ __forceinline__ double3 vertex_disp(VertexLocal &q)
{
    asm volatile("" : "+v"(q.hi.x), "+v"(q.hi.y), "+v"(q.hi.z), "+v"(q.lo.x), "+v"(q.lo.y), "+v"(q.lo.z));
    return make_double3((double)q.hi.x + (double)q.lo.x, (double)q.hi.y + (double)q.lo.y,
                        (double)q.hi.z + (double)q.lo.z);
}

// After the group sum: damping and inertia, and the guarded 3x3 solve (us:
// the vertex's current displacement).  Lanes without a vertex get dx = 0.
__device__ __forceinline__ float3 vertex_close(const SweepArgs &a, int v, bool valid, Acc &acc, const VertexLocal &q,
                                               const double3 &us)
{
    if (!valid) return make_float3(0.f, 0.f, 0.f);
    if (a.kd_dt != 0.f) {
        // damping (src/vbd_solver.py:222-227): g += kd/dt H_el (x - x_prev), H += kd/dt H_el
        const double4 up = a.uprev[v];
        const float e0 = (float)(us.x - up.x), e1 = (float)(us.y - up.y), e2 = (float)(us.z - up.z);
        const float k = a.kd_dt;
        acc.g0 += k * (acc.h00 * e0 + acc.h01 * e1 + acc.h02 * e2);
        acc.g1 += k * (acc.h01 * e0 + acc.h11 * e1 + acc.h12 * e2);
        acc.g2 += k * (acc.h02 * e0 + acc.h12 * e1 + acc.h22 * e2);
        const float f = 1.f + k;
        acc.h00 *= f; acc.h01 *= f; acc.h02 *= f; acc.h11 *= f; acc.h12 *= f; acc.h22 *= f;
    }
    // inertia + gravity (src/vbd_solver.py:15-47, summed over slots): M/dt^2 (x - t)
    acc.g0 += q.w * q.t0;
    acc.g1 += q.w * q.t1;
    acc.g2 += q.w * q.t2;
    acc.h00 += q.w;
    acc.h11 += q.w;
    acc.h22 += q.w;
    return solve3_guarded(acc);
}

// After the slot loop: the group sum of gradient and Hessian, damping and
// inertia, and the guarded 3x3 solve.  us receives the vertex's current
// displacement in FP64.  Lanes without a vertex get dx = 0.
template <int L>
__device__ __forceinline__ float3 vertex_finish(const SweepArgs &a, int v, bool valid, Acc &acc, VertexLocal &q,
                                                double3 &us)
{
    us = vertex_disp(q);
    acc_group_sum<L>(acc);
    return vertex_close(a, v, valid, acc, q, us);
}

// vertex_finish for groups of 1 << lg lanes that vary from lane to lane
// (lg <= 2): a wave whose groups all have 4 lanes takes the fixed form; the
// per-lane form adds the same values in the same order.
__device__ __forceinline__ float3 vertex_finish_lanes(const SweepArgs &a, int v, bool valid, Acc &acc,
                                                      VertexLocal &q, double3 &us, int lg)
{
    us = vertex_disp(q);
    if (__all(lg == 2)) acc_group_sum<4>(acc);
    else acc_group_sum_lanes(acc, lg);
    return vertex_close(a, v, valid, acc, q, us);
}

// A lane's slots s, s + step, ... below s1, one per batch, software-pipelined:
// the next compact record is requested after the current slot's constants
// and gathers, so each batch waits for one memory latency, not two.
template <int MAT>
__device__ __forceinline__ void slot_loop(const SweepArgs &a, int s, int s1, int step, const float3 &hi,
                                          const float3 &lo, Acc &acc)
{
    float4 c = s < s1 ? a.slots[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool edge = a.dict_ext != nullptr;  // the same for every thread
    for (; s < s1; s += step) {
        // 32-bit unsigned byte offsets from the arrays' (uniform) bases:
        // address = base + offset, no 64-bit index arithmetic per load
        // entry * 48 bytes in 32 bits (create admits fewer than 2^26 slots)
        const float4 *d = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(a.dict) +
                                                           __float_as_uint(c.w) * 48u);
        const float4 d0 = d[0], d1 = d[1], d2 = d[2];
        const EdgeK ek = edge_k<MAT>(a.dict_ext, __float_as_uint(c.w), edge);
        const float4 x0 = *at(a.u, __float_as_uint(c.x));
        const float4 x1 = *at(a.u, __float_as_uint(c.y));
        const float4 x2 = *at(a.u, __float_as_uint(c.z));
        // keep the request for the next record behind this slot's loads
        __builtin_amdgcn_sched_barrier(0);
        const float4 cn = *at(a.slots, (uint32_t)min(s + step, s1 - 1));  // the last one reloads a record in cache
        __builtin_amdgcn_sched_barrier(0);
        const SlotRel sr = slot_rel(x0, x1, x2, hi, lo);
        slot_eval<MAT>(d2.z, d0.w, slot_kc<MAT>(d2.z), d2.w, make_float3(d0.x, d0.y, d0.z),
                       make_float3(d1.x, d1.y, d1.z), make_float3(d1.w, d2.x, d2.y), ek, edge, sr.e0, sr.e1, sr.d2,
                       acc);
        c = cn;
    }
}

// Gradient and Hessian of one vertex's local energy (all incident tets plus
// inertia and damping) accumulated by a group of L lanes, and the block solve:
// the update dx (zero for lanes without a vertex).  us receives the vertex's
// current displacement.
template <int L, int B, int MAT>
__device__ __forceinline__ float3 vertex_dx(const SweepArgs &a, int v, bool valid, int lane, double3 &us,
                                            const VertexPre &pre)
{
    Acc acc;
    acc_zero(acc);
    VertexLocal ql = vertex_local(pre);
    const float3 hi = ql.hi, lo = ql.lo;
    if (valid) {
        const int s0 = pre.s0, s1 = pre.s1;
#if VBD_TET_RECORDS
        // B slots per lane per batch: entries, tet records, gathers, arithmetic
        for (int s = s0 + lane; s < s1; s += B * L) {
            int e[B];
            float4 q[B][4];
            float4 x[B][3];
#pragma unroll
            for (int j = 0; j < B; ++j)
                if (s + j * L < s1) e[j] = a.ent[s + j * L];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (s + j * L < s1) {
                    const float4 *r = a.slots + 4 * (size_t)(e[j] >> 2);
                    q[j][0] = r[0];
                    q[j][1] = r[1];
                    q[j][2] = r[2];
                    q[j][3] = r[3];
                }
            }
            int o[B][3];
            float3 bs[B], b0[B], b1[B];
            float ka[B], kb[B], kc[B];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (s + j * L < s1) {
                    tet_slot(q[j], e[j] & 3, o[j][0], o[j][1], o[j][2], bs[j], b0[j], b1[j], ka[j], kb[j], kc[j]);
                    x[j][0] = a.u[o[j][0]];
                    x[j][1] = a.u[o[j][1]];
                    x[j][2] = a.u[o[j][2]];
                }
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (s + j * L < s1) {
                    const SlotRel r = slot_rel(x[j][0], x[j][1], x[j][2], hi, lo);
                    const EdgeK ek = edge_k<MAT>(nullptr, 0u, false);
                    slot_eval<MAT>(ka[j], kb[j], kc[j], slot_bb(bs[j]), bs[j], b0[j], b1[j], ek, false, r.e0, r.e1,
                                   r.d2, acc);
                }
            }
        }
#else
        if constexpr (B == 1) {
            slot_loop<MAT>(a, s0 + lane, s1, L, hi, lo, acc);
        } else
        // B slots per lane per batch: all records, then all gathers, then the
        // arithmetic, so each batch costs two memory latencies, not 2B.
        for (int s = s0 + lane; s < s1; s += B * L) {
            float4 q[B][3];
            float2 qd[B];
            float qbb[B];
            EdgeK qk[B];
            float4 x[B][3];
            // compact record (ids, entry), then the entry's constants
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int sj = s + j * L;
                if (sj < s1) {
                    const float4 c = a.slots[sj];
                    const float4 *d = a.dict + 3 * (size_t)__float_as_int(c.w);
                    const float4 d2 = d[2];
                    q[j][0] = make_float4(c.x, c.y, c.z, d2.z);
                    q[j][1] = d[0];
                    q[j][2] = d[1];
                    qd[j] = make_float2(d2.x, d2.y);
                    qbb[j] = d2.w;
                    qk[j] = edge_k<MAT>(a.dict_ext, (uint32_t)__float_as_int(c.w), a.dict_ext != nullptr);
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
                    const SlotRel r = slot_rel(x[j][0], x[j][1], x[j][2], hi, lo);
                    slot_eval<MAT>(qq[0].w, qq[1].w, slot_kc<MAT>(qq[0].w), qbb[j], make_float3(qq[1].x, qq[1].y, qq[1].z),
                                   make_float3(qq[2].x, qq[2].y, qq[2].z), make_float3(qq[2].w, qd[j].x, qd[j].y),
                                   qk[j], a.dict_ext != nullptr, r.e0, r.e1, r.d2, acc);
                }
            }
        }
#endif
    }
    return vertex_finish<L>(a, v, valid, acc, ql, us);
}


// Evaluate and apply one vertex update for a lane group.  Returns the metric
// bits (0 for lanes without a vertex).
template <int L, int B, int MAT>
__device__ __forceinline__ float3 vertex_dx(const SweepArgs &a, int v, bool valid, int lane, double3 &us)
{
    return vertex_dx<L, B, MAT>(a, v, valid, lane, us, vertex_pre(a, v, valid));
}

template <int L, int B, int MAT>
__device__ __forceinline__ uint32_t update_vertex(const SweepArgs &a, int v, bool valid, int lane,
                                                  const VertexPre &pre)
{
    double3 us;
    const float3 dx = vertex_dx<L, B, MAT>(a, v, valid, lane, us, pre);
    if (!valid) return 0u;
    if (lane == 0) a.u[v] = pack_disp(us.x + dx.x, us.y + dx.y, us.z + dx.z);
    return metric_bits(dx);
}

template <int L, int B, int MAT>
__device__ __forceinline__ uint32_t update_vertex(const SweepArgs &a, int v, bool valid, int lane)
{
    double3 us;
    const float3 dx = vertex_dx<L, B, MAT>(a, v, valid, lane, us);
    if (!valid) return 0u;
    if (lane == 0) a.u[v] = pack_disp(us.x + dx.x, us.y + dx.y, us.z + dx.z);
    return metric_bits(dx);
}

// The update a Jacobi pass from the positions in a.u would take (no write).
template <int L, int B, int MAT>
__device__ __forceinline__ float3 eval_vertex(const SweepArgs &a, int v, bool valid, int lane)
{
    double3 us;
    return vertex_dx<L, B, MAT>(a, v, valid, lane, us);
}

// Slot records a lane of the staged sweep holds: 32 incident tets / 4 lanes.
constexpr int kStagedSlots = kMaxElementsPerVertex / 4;

// Per-body convergence state seen by one iteration of a batched solver.  An
// iteration reads its bodies' state from cur, except the first colour launch
// of an iteration after the first of a run (fold), which decides each body's
// retirement itself from the previous iteration's state and per-body maxima
// (prev, mprev): every vertex of a body recomputes the same decision, and
// extra waves of that launch write the new state (fold_bodies).  macc
// receives this iteration's per-body maxima.
struct BodyPhase {
    const BodyState *cur = nullptr;
    const BodyState *prev = nullptr;
    const uint32_t *mprev = nullptr;
    uint32_t *macc = nullptr;
    float tol = 0.f;
    int max_iter = 0;
    int fold = 0;
};

// A live body retires after an iteration whose max |dx_x+dx_y+dx_z| fell
// below tol, or that used its last allowed iteration.
__device__ __forceinline__ bool body_retires(const BodyState &s, uint32_t m, float tol, int max_iter)
{
    return (__uint_as_float(m) < tol) | (s.iters + 1 >= max_iter);
}

// Whether body's vertices sit out this iteration (both loads are issued
// without a branch between them).
__device__ __forceinline__ bool body_done(const BodyPhase &ph, int body)
{
    if (!ph.fold) return ph.cur[body].done != 0;
    const BodyState s = ph.prev[body];
    const uint32_t m = ph.mprev[body];
    return (s.done != 0) | body_retires(s, m, ph.tol, ph.max_iter);
}


// The staging tables of a solver (build_staging): per block a list of
// entries, vertex by vertex each vertex's distinct neighbours (neighbour id |
// owner << 25, the owner the vertex's index in the block; -1 past the list);
// per vertex its first entry; per slot n0 | n1 << 5 | n2 << 10 | entry << 15,
// n_k the index of its k-th other vertex in the vertex's entries and entry
// its record dictionary entry.
struct StageTables {
    const int32_t *nb;
    const uint32_t *rec;
    const int2 *vrec;
};

// One block of the staged colour sweep: 256 threads over lanes
// [256 b, 256 b + 256) of a colour of nv vertices at v0 laid out by lay
// (lane_vertex: 4, 2 or 1 lanes per vertex by slot count; a batch's colours
// have 4 lanes for every vertex, blocks of kStageRows).  Two memory rounds,
// then the neighbour differences, then arithmetic only:
//   1. the vertex's own data and its word in st.vrec (first entry, slot
//      count, record sequence), the block's entries (at lay.ids +
//      stage_list_offset: kStageMax per block below lay.b4, kStageMaxWide
//      from there) and the record dictionary;
//   2. the entries' neighbour positions and every 4-byte slot record of the
//      lane from the vertex's sequence in st.rec (the table of distinct
//      sequences ends with 32 zero records, so a lane reads its 8 without a
//      bound).
// Each vertex's own displacement goes to LDS, then each entry's difference
// d = u_neighbour - u_vertex (neighbour_diff) is formed once and stored in
// LDS, and each slot reads its three differences and its dictionary entry
// from LDS (slot_rel_d).  The neighbours belong to other colours, which the
// launch does not write, so staging them first is exact, and the arithmetic
// is slot_rel's, so the results are the same bit for bit as lanes_sweep's.
// BODIES: a batched solver's launch -- a vertex of a body that has converged
// is not updated (its lanes get body -1), instead of the step's convergence
// flag being tested, and rows[b] counts the live bodies with a vertex in the
// block.  Returns false when the (single-body) step has converged already, or
// no body of the block is live (the whole block then does nothing); m
// receives the lane's metric bits (0 for lanes without an updated vertex),
// body the vertex's body.
// WIDE: the launch has blocks of 2-lane vertices, whose lists hold up to
// kStageMaxWide entries (more LDS per block); otherwise kStageMax.
// DL: each block of a batch has its own dictionary list (dl, kStageDict
// entries per block, -1 padded, the colour's first block's at lay.ids /
// kStageMax: a batch's blocks all take that stride; read in round 1, the
// rows in round 2); otherwise every block stages the shared dictionary.
template <int MAT, bool BODIES, bool WIDE, bool DL>
__device__ __forceinline__ bool staged_sweep(const SweepArgs &a, int v0, int nv, const LaneLayout &lay, int b,
                                             const StageTables &st, int n_dict, const int32_t *body_new,
                                             const BodyPhase *ph, const int32_t *rows, int &body_out, uint32_t &m,
                                             const int32_t *dl = nullptr)
{
    static_assert(!(BODIES && WIDE), "a batch's blocks hold 4-lane vertices only");
    constexpr int kCap = WIDE ? kStageMaxWide : kStageMax;
    constexpr int kStageUnroll = (kCap + 255) / 256;
    constexpr int kOwners = WIDE ? 128 : 64;  // vertices of a staged block
    static_assert(kStageMaxWide >= kStageMax, "a wide list holds a narrow one");
    static_assert(3 * kStageDict <= 256, "one dictionary row per thread");
    __shared__ float4 s_d[kCap];           // per entry: d.xyz
    __shared__ float4 s_own[kOwners];      // per vertex: FP32 parts .xyz, remainder x in .w
    __shared__ float2 s_own2[kOwners];     // remainders y, z
    __shared__ float4 s_dict[3 * kStageDict];
    __shared__ float4 s_ext[MAT == kSNH ? 3 * kStageDict : 1];  // the entries' rest-edge constants
    const int t = b * 256 + threadIdx.x;
    int g, lg, vi, vfirst;
    if constexpr (BODIES) {
        vi = t >> 2;
        g = t & 3;
        lg = 2;
        vfirst = b * 64;
    } else {
        vi = lane_vertex(t, lay.n4, lay.n2, g, lg);
        int g0, lg0;
        vfirst = lane_vertex(b * 256, lay.n4, lay.n2, g0, lg0);  // the block's first vertex
    }
    const bool in = vi < nv;
    const int v = v0 + (in ? vi : 0);
    // round 1: the vertex's own data, its first entry, slot count and record
    // sequence (st.vrec), the block's entries and the dictionary
    VertexPre pre;
    pre.uv = a.u[v];
    pre.tg = a.tgt[v];
    const int2 vr = st.vrec[v];
    const int ov = vr.x & 0xFFFF;
    pre.s0 = vr.y;
    pre.s1 = in ? vr.y + (vr.x >> 16) : vr.y;
    const bool wide = b >= lay.b4;  // the same for every thread
    const int cap = wide ? kStageMaxWide : kStageMax;
    const int32_t *ids = st.nb + lay.ids + stage_list_offset(b, lay.b4);
    int id[kStageUnroll];
#pragma unroll
    for (int q = 0; q < kStageUnroll; ++q) {
        const int k = (int)threadIdx.x + 256 * q;
        id[q] = k < kCap && k < cap ? ids[k] : -1;
    }
    // the dictionary rows: the shared dictionary's, or (DL) the entry of the
    // block's list this thread's row belongs to, its row read in round 2
    const bool has_dict = (int)threadIdx.x < 3 * (DL ? kStageDict : n_dict);
    float4 drow = !DL && has_dict ? a.dict[threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 xrow = MAT == kSNH && !DL && has_dict ? a.dict_ext[threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    int de = -1;
    static_assert(!DL || BODIES, "per-block dictionaries are a batch's");
    if constexpr (DL)
        if (has_dict) de = dl[(lay.ids / kStageMax + b) * kStageDict + (int)threadIdx.x / 3];
    int body = 0;
    if constexpr (BODIES) {
        body = body_new[v];
        // this block's row: the live bodies with a vertex in it (same for every thread)
        const int live = rows[b];
        // Keep the round's loads (the body's among them) ahead of the test
        // below: sunk past it, they would start only after the row's count
        // arrived, a memory latency later.
        asm volatile("" ::"v"(body), "s"(live) : "memory");
        if (live == 0) {
            m = 0u;
            body_out = -1;
            return false;
        }
    } else {
        // the flag was written by an earlier launch: the same for every thread
        if (a.ctrl->converged) {
            m = 0u;
            return false;
        }
    }
    // round 2 (the body's flag is read by every thread, so that no branch
    // makes the neighbour loads wait for it)
    int done = 0;
    if constexpr (BODIES) done = body_done(*ph, body);
    float4 p[kStageUnroll];
#pragma unroll
    for (int q = 0; q < kStageUnroll; ++q)
        p[q] = id[q] != -1 ? a.u[(uint32_t)id[q] & 0x1FFFFFFu] : make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t rec[kStagedSlots];
    const uint32_t *rl = st.rec + pre.s0 + g;
#pragma unroll
    for (int k = 0; k < kStagedSlots; ++k) rec[k] = rl[k << lg];
    if constexpr (DL) {
        if (de >= 0) {
            drow = a.dict[3 * de + (int)threadIdx.x % 3];
            if constexpr (MAT == kSNH) xrow = a.dict_ext[3 * de + (int)threadIdx.x % 3];
        }
    }
    if (has_dict) {
        s_dict[threadIdx.x] = drow;
        if constexpr (MAT == kSNH) s_ext[threadIdx.x] = xrow;
    }
    VertexLocal ql = vertex_local(pre);
    const float3 hi = ql.hi, lo = ql.lo;
    if (in && g == 0) {
        s_own[vi - vfirst] = make_float4(hi.x, hi.y, hi.z, lo.x);
        s_own2[vi - vfirst] = make_float2(lo.y, lo.z);
    }
    __syncthreads();
    // each entry's difference to its vertex
#pragma unroll
    for (int q = 0; q < kStageUnroll; ++q) {
        if (id[q] != -1) {
            const int o = (int)((uint32_t)id[q] >> 25);
            const float4 oh = s_own[o];
            const float2 ol = s_own2[o];
            const float3 d = neighbour_diff(p[q], disp_lo(p[q]), make_float3(oh.x, oh.y, oh.z),
                                            make_float3(oh.w, ol.x, ol.y));
            s_d[threadIdx.x + 256 * q] = make_float4(d.x, d.y, d.z, 0.f);
        }
    }
    __syncthreads();
    const bool valid = in && !done;
    Acc acc;
    acc_zero(acc);
    // this lane's slots: records pre.s0 + g, + (1 << lg), ... below pre.s1 of
    // the vertex's sequence
    const int nk = valid && pre.s1 > pre.s0 + g ? (pre.s1 - pre.s0 - g + (1 << lg) - 1) >> lg : 0;
    constexpr uint32_t kM = (1u << kStageEntryBits) - 1u;
#pragma unroll
    for (int k = 0; k < kStagedSlots; ++k) {
        if (k < nk) {
            const uint32_t r = rec[k];
            const uint32_t e = r >> 15, n0 = r & kM, n1 = (r >> 5) & kM, n2 = (r >> 10) & kM;
            const float4 d0 = s_dict[3 * e], d1 = s_dict[3 * e + 1], d2 = s_dict[3 * e + 2];
            const EdgeK ek = edge_k<MAT>(s_ext, e);
            const float4 q0 = s_d[ov + n0], q1 = s_d[ov + n1], q2 = s_d[ov + n2];
            const SlotRel sr = slot_rel_d(make_float3(q0.x, q0.y, q0.z), make_float3(q1.x, q1.y, q1.z),
                                          make_float3(q2.x, q2.y, q2.z));
            slot_eval<MAT>(d2.z, d0.w, slot_kc<MAT>(d2.z), d2.w, make_float3(d0.x, d0.y, d0.z),
                           make_float3(d1.x, d1.y, d1.z), make_float3(d1.w, d2.x, d2.y), ek, true, sr.e0, sr.e1,
                           sr.d2, acc);
        }
    }
    double3 us;
    float3 dx;
    if constexpr (BODIES) dx = vertex_finish<4>(a, v, valid, acc, ql, us);
    else dx = vertex_finish_lanes(a, v, valid, acc, ql, us, lg);
    m = 0u;
    body_out = valid ? body : -1;
    if (valid) {
        if (g == 0) a.u[v] = pack_disp(us.x + dx.x, us.y + dx.y, us.z + dx.z);
        m = metric_bits(dx);
    }
    return true;
}

// One block of a single body's big-colour sweep without staging lists (the
// blocks from lay.b1 on): the lanes of lay, each walking its slots by the
// compact records (slot_loop).  Same arithmetic and lane groups as
// staged_sweep.  Returns false when the step has converged already.
template <int MAT>
__device__ __forceinline__ bool lanes_sweep(const SweepArgs &a, int v0, int nv, const LaneLayout &lay, int b,
                                            uint32_t &m)
{
    const int t = b * 256 + threadIdx.x;
    int g, lg;
    const int vi = lane_vertex(t, lay.n4, lay.n2, g, lg);
    const bool in = vi < nv;
    const int v = v0 + (in ? vi : 0);
    const VertexPre pre = vertex_pre(a, v, in);
    if (a.ctrl->converged) {
        m = 0u;
        return false;
    }
    Acc acc;
    acc_zero(acc);
    VertexLocal ql = vertex_local(pre);
    const float3 hi = ql.hi, lo = ql.lo;
    if (in) slot_loop<MAT>(a, pre.s0 + g, pre.s1, 1 << lg, hi, lo, acc);
    double3 us;
    const float3 dx = vertex_finish_lanes(a, v, in, acc, ql, us, lg);
    m = 0u;
    if (in) {
        if (g == 0) a.u[v] = pack_disp(us.x + dx.x, us.y + dx.y, us.z + dx.z);
        m = metric_bits(dx);
    }
    return true;
}

// A batched launch's metrics into the per-body maxima bmet (bits of
// max |dx_x + dx_y + dx_z| of the iteration so far): per wave, one
// device-scope atomic max per body among its lanes (a wave's vertices are
// consecutive in colour-major, body-minor order, so mostly one body).  Lanes
// with body < 0 take no part.
__device__ __forceinline__ void body_metric_max(int body, uint32_t m, uint32_t *bmet, uint32_t *lds = nullptr,
                                                int b0 = 0, int nlds = 0)
{
    const int lid = (int)(threadIdx.x & 63);
    bool pending = body >= 0;
    while (true) {
        const uint64_t mask = __ballot(pending);
        if (!mask) break;
        const int lead = __builtin_ctzll(mask);
        const int lb = __builtin_amdgcn_readlane(body, lead);
        const bool mine = pending && body == lb;
        const uint32_t x = wave_max_u32(mine ? m : 0u);
        if (lid == lead) {
            // bodies b0 .. b0 + nlds - 1 into the block's LDS maxima instead
            const unsigned r = (unsigned)(lb - b0);
            if (r < (unsigned)nlds) atomicMax(&lds[r], x);
            else atomicMax(&bmet[lb], x);
        }
        pending = pending && !mine;
    }
}

// slots per lane per load batch, by lanes per vertex
template <int L>
struct Batch {
    static constexpr int value = L == 1 ? 4 : L == 2 ? 4 : L == 4 ? VBD_BATCH_L4 : L == 8 ? VBD_BATCH_L8 : L == 16 ? 2 : 1;
};

struct StepParams {
    float dt, inv_dt, tol, kd_dt;
    int max_iter;
    int skip_init;            // the initial guess is already in u (warm start)
    float4 *u_last = nullptr;  // if set: u before the last allowed iteration
};

// inertia target x_prev + dt v + dt^2 g of an active vertex (xyz) and M/dt^2 (w)
__device__ __forceinline__ double4 inertia_target(const double4 p, const float4 v, float mass, float dt, float4 g)
{
    const double h = dt, h2 = h * h;
    // explicit fused forms: every kernel that inlines this rounds alike
    return make_double4(fma(h2, (double)g.x, fma(h, (double)v.x, p.x)), fma(h2, (double)g.y, fma(h, (double)v.y, p.y)),
                        fma(h2, (double)g.z, fma(h, (double)v.z, p.z)), (double)mass / h2);
}

__device__ __forceinline__ void init_vertex(int i, float4 *u, const double4 *uprev, const float4 *vel, double4 *tgt,
                                            const float *mass, const uint8_t *active, float dt, float4 g)
{
    const double4 p = uprev[i];
    if (active[i]) {
        const double4 t = inertia_target(p, vel[i], mass[i], dt, g);
        u[i] = pack_disp(t.x, t.y, t.z);
        tgt[i] = t;
    } else {
        u[i] = pack_disp(p.x, p.y, p.z);
        tgt[i] = make_double4(p.x, p.y, p.z, 0.0);
    }
}

// v = (x - x_prev) / dt from the FP64 displacements, stored in FP32
__device__ __forceinline__ float4 step_velocity(const double4 x, const double4 q, float dt)
{
    const double r = 1.0 / (double)dt;
    return make_float4((float)((x.x - q.x) * r), (float)((x.y - q.y) * r), (float)((x.z - q.z) * r), 0.f);
}

// A whole time step of one body in one workgroup: initial guess, colour
// sweeps with a barrier between colours until the body's own
// max|dx_x+dx_y+dx_z| < tol (or max_iter), then the velocity update.  Used
// when every colour is small enough that a per-colour launch would be
// launch-latency bound; blockIdx.x selects the body (ranges: per body, C
// colour ranges then the range of its other vertices, as (start, count)).
template <int L, int B, int MAT>
__global__ __launch_bounds__(1024) void k_step_resident(SweepArgs a, const int2 *ranges, int C, const float4 *grav,
                                                        const float *mass, const uint8_t *active, double4 *uprev,
                                                        float4 *vel, StepParams p, int2 *out)
{
    __shared__ uint32_t s_bits;
    const int2 *r = ranges + (size_t)blockIdx.x * (C + 1);
    const float4 g = grav[blockIdx.x];
    if (!p.skip_init) {
        for (int c = 0; c <= C; ++c) {
            const int2 rc = r[c];
            for (int i = rc.x + threadIdx.x; i < rc.x + rc.y; i += blockDim.x)
                init_vertex(i, a.u, uprev, vel, const_cast<double4 *>(a.tgt), mass, active, p.dt, g);
        }
        __syncthreads();
    }
    const int groups = blockDim.x / L;
    const int gi = threadIdx.x / L, lane = threadIdx.x % L;
    int it = 0;
    uint32_t bits = 0;
    for (; it < p.max_iter; ++it) {
        if (p.u_last && it == p.max_iter - 1) {
            for (int c = 0; c < C; ++c)
                for (int i = r[c].x + threadIdx.x; i < r[c].x + r[c].y; i += blockDim.x) p.u_last[i] = a.u[i];
            __syncthreads();
        }
        uint32_t mb = 0;
        for (int c = 0; c < C; ++c) {
            const int2 rc = r[c];
            for (int base = rc.x; base < rc.x + rc.y; base += groups) {
                const int v = base + gi;
                const bool valid = v < rc.x + rc.y;
                const uint32_t m = update_vertex<L, B, MAT>(a, valid ? v : rc.x, valid, lane);
                mb = mb > m ? mb : m;
            }
            __syncthreads();
        }
        const uint32_t all = block_max_u32(mb);
        if (threadIdx.x == 0) s_bits = all;
        __syncthreads();
        bits = s_bits;
        if (__uint_as_float(bits) < p.tol) break;
    }
    if (it == p.max_iter) it = p.max_iter - 1;
    for (int c = 0; c <= C; ++c) {
        const int2 rc = r[c];
        for (int i = rc.x + threadIdx.x; i < rc.x + rc.y; i += blockDim.x) {
            const double4 x = unpack_disp4(a.u[i]), q = uprev[i];
            vel[i] = step_velocity(x, q, p.dt);
            uprev[i] = x;
        }
    }
    if (threadIdx.x == 0) out[blockIdx.x] = make_int2(it + 1, (int)bits);
}

template <int L, int MAT>
void launch_resident(const SweepArgs &a, const int2 *ranges, int C, const float4 *grav, const float *mass,
                     const uint8_t *active, double4 *uprev, float4 *vel, StepParams p, int2 *out, int n_bodies,
                     int block, hipStream_t st)
{
    k_step_resident<L, Batch<L>::value, MAT><<<n_bodies, block, 0, st>>>(a, ranges, C, grav, mass, active, uprev, vel,
                                                                         p, out);
}

template <int MAT>
void dispatch_resident(int L, const SweepArgs &a, const int2 *ranges, int C, const float4 *grav, const float *mass,
                       const uint8_t *active, double4 *uprev, float4 *vel, StepParams p, int2 *out, int n_bodies,
                       int block, hipStream_t st)
{
    switch (L) {
    case 1: launch_resident<1, MAT>(a, ranges, C, grav, mass, active, uprev, vel, p, out, n_bodies, block, st); break;
    case 2: launch_resident<2, MAT>(a, ranges, C, grav, mass, active, uprev, vel, p, out, n_bodies, block, st); break;
    case 8: launch_resident<8, MAT>(a, ranges, C, grav, mass, active, uprev, vel, p, out, n_bodies, block, st); break;
    case 16: launch_resident<16, MAT>(a, ranges, C, grav, mass, active, uprev, vel, p, out, n_bodies, block, st); break;
    case 32: launch_resident<32, MAT>(a, ranges, C, grav, mass, active, uprev, vel, p, out, n_bodies, block, st); break;
    default: launch_resident<4, MAT>(a, ranges, C, grav, mass, active, uprev, vel, p, out, n_bodies, block, st); break;
    }
}

}  // namespace
}  // namespace vbd

// vbd_device.h -- per-slot VBD physics, 3x3 solve and lane-group reductions (gfx950).
//
// A "slot" is one (vertex, incident tet) pair.  Its constants are
//     ka                                  material constant (below)
//     qb = (bs.xyz, kb)                   shape gradient b of the slot's own vertex
//     qc = (b0.xyz, b1.x)                 b of vertex o0, first component of b of o1
//     qd = (b1.y, b1.z)                   ; b of o2 = -(bs+b0+b1)
// with o0, o1, o2 the tet's other three vertices.  Slots of a vertex are stored
// contiguously in colour-major vertex order as compact 16-byte records
//     slots  (1 float4 per slot)  c = (o0, o1, o2, entry as int bits)
// and the constants live in a record dictionary
//     dict   (3 float4 per entry) d0 = qb, d1 = qc, d2 = (qd.x, qd.y, ka, |bs|^2).
// Meshes built from voxels have a handful of distinct rest shapes (5 for the
// 5-tet split of a uniform grid), so their slots share few entries (20 per
// material) and a sweep reads 16 bytes per slot from HBM, the dictionary from
// cache.  Otherwise every slot has its own entry, in slot order (64 bytes per
// slot, read as two streams).
//
// b is the row of invDm belonging to the vertex's local index (minus the row sum
// for local index 3), so dF/dx_{v,i} = e_i b^T (src/vbd_solver.py:86-94,450-465).
// Material constants, volume folded in:
//   SNH : ka = V mu', kb = V lambda', kc = 0.75 ka  (= V lambda'(alpha-1))
//   StVK: ka = V mu,  kb = V lambda,  kc = 0
// kc follows from ka and the material, so it is not stored.
//
// Solvers whose dictionary is read from LDS (at most 64 entries, or a batch's
// per-block lists) also give each entry rest-edge constants (3 float4 in
// dict_ext, read by SNH only; null otherwise, where three more scattered loads
// per slot would cost more than the arithmetic they save):
//     x0 = (P.xyz, k)   x1 = (Q.xyz, R.x)   x2 = (R.y, R.z)
// with E0 = X_o0 - X_o2, E1 = X_o1 - X_o2, R = X_self - X_o2 the rest edges
// (the rows of [b0 b1 bs]^-1), k = det[b0 b1 bs], P = k E0 and Q = k E1; so
// bs = k E0 x E1 and bs . R = 1.
//
// Positions are carried as displacements u = x - X from the rest shape, so
// F = I + D with D = sum_i (u_i - u_self) b_i^T is formed from small
// differences; FP32 absolute coordinates cannot resolve a 1e-7 m tolerance on a
// 2.56 m beam.  The iterate u is packed to 32 significant bits (pack_disp in
// vbd_sweep.h): FP32 displacements alone (a 1e-9 m ulp at 1 cm) move
// tolerance-terminated trajectories off the FP64 ones by iteration-count flips.
// Each difference u_i - u_self reaches the slot arithmetic rounded once to FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// VBD_TET_RECORDS = 1 replaces the per-slot records by one 64-byte record per
// tet, shared by its four slots, and a 4-byte entry per slot (tet << 2 | local
// index, the incidence CSR entry itself):
//
//   r0 = (id0, id1, id2, id3 as int bits)    colour-major vertex ids
//   r1 = (b0.xyz, ka)   r2 = (b1.xyz, kb)   r3 = (b2.xyz, kc)
//
// with b0..b2 the rows of invDm and b3 = -(b0+b1+b2).  The tet records of a
// mesh are a quarter of the per-slot records' size (84 MB for 1.31 M tets),
// small enough to stay in the 256 MB Infinity Cache between the four colour
// sweeps that read each tet in an iteration.
#ifndef VBD_TET_RECORDS
#define VBD_TET_RECORDS 0
#endif

namespace vbd {

enum Material : int { kSNH = 0, kStVK = 1 };

// kc of a slot (see the record layout above)
template <int MAT>
__device__ __forceinline__ float slot_kc(float ka)
{
    return MAT == kSNH ? 0.75f * ka : 0.f;
}

// A slot's view of its tet record (VBD_TET_RECORDS layout): the other three
// vertex ids (ascending local index) and the b vectors of the vertex itself
// and of the first two others; the third other's b is -(bs + b0 + b1).
__device__ __forceinline__ void tet_slot(const float4 r[4], int loc, int &oa, int &ob, int &oc, float3 &bs,
                                         float3 &b0, float3 &b1, float &ka, float &kb, float &kc)
{
    const int i0 = __float_as_int(r[0].x), i1 = __float_as_int(r[0].y), i2 = __float_as_int(r[0].z),
              i3 = __float_as_int(r[0].w);
    oa = loc == 0 ? i1 : i0;
    ob = loc <= 1 ? i2 : i1;
    oc = loc <= 2 ? i3 : i2;
    const float3 c0 = make_float3(r[1].x, r[1].y, r[1].z), c1 = make_float3(r[2].x, r[2].y, r[2].z),
                 c2 = make_float3(r[3].x, r[3].y, r[3].z);
    const float3 c3 = make_float3(-(c0.x + c1.x + c2.x), -(c0.y + c1.y + c2.y), -(c0.z + c1.z + c2.z));
    bs = loc == 0 ? c0 : loc == 1 ? c1 : loc == 2 ? c2 : c3;
    b0 = loc == 0 ? c1 : c0;
    b1 = loc <= 1 ? c2 : c1;
    ka = r[1].w;
    kb = r[2].w;
    kc = r[3].w;
}

struct Acc {
    float g0, g1, g2;
    float h00, h01, h02, h11, h12, h22;
};

__device__ __forceinline__ void acc_zero(Acc &a)
{
    a.g0 = a.g1 = a.g2 = 0.f;
    a.h00 = a.h01 = a.h02 = a.h11 = a.h12 = a.h22 = 0.f;
}

// Elastic gradient and Hessian block of one slot, added into a.
//   e0, e1, d2: u_o0 - u_o2, u_o1 - u_o2, u_o2 - u_self, so that with
//   b2 = -(bs + b0 + b1) the displacement gradient
//   D = sum_k (u_ok - u_self) b_k^T = e0 b0^T + e1 b1^T - d2 bs^T.
// SNH (src/vbd_solver.py:101-124) in the equivalent compact form
//   g = V P b,  P = mu'(1 - 1/(Ic+1)) F + lambda'(J - alpha) cof(F)
//   H = V[2mu'/(Ic+1)^2 (Fb)(Fb)^T + mu'(1 - 1/(Ic+1))|b|^2 I + lambda'(cof(F) b)(cof(F) b)^T]
// StVK (src/vbd_solver.py:127-144):
//   g = V F S b, S = lambda tr(E) I + 2 mu E
//   H = V[(lambda+mu)(Fb)(Fb)^T + mu |b|^2 F F^T + (lambda tr(E)|b|^2 + 2 mu b^T E b) I]
// |bs|^2 of a slot, as the record dictionary stores it (d2.w): every path
// that evaluates a slot uses this one rounding
__device__ __forceinline__ float slot_bb(const float3 bs) { return fmaf(bs.z, bs.z, fmaf(bs.y, bs.y, bs.x * bs.x)); }

// The rest-edge constants of a slot (x0, x1, x2 above) from its b vectors,
// in FP64; the dictionary's are made by the same function.
__device__ __forceinline__ void edge_consts(const float3 bs, const float3 b0, const float3 b1, float4 &x0, float4 &x1,
                                            float2 &x2)
{
    // M = [b0 b1 bs] (columns) and its adjugate A; M^-1 = A / det M
    const double m00 = b0.x, m01 = b1.x, m02 = bs.x, m10 = b0.y, m11 = b1.y, m12 = bs.y, m20 = b0.z, m21 = b1.z,
                 m22 = bs.z;
    const double a00 = m11 * m22 - m12 * m21, a01 = m02 * m21 - m01 * m22, a02 = m01 * m12 - m02 * m11;
    const double a10 = m12 * m20 - m10 * m22, a11 = m00 * m22 - m02 * m20, a12 = m02 * m10 - m00 * m12;
    const double a20 = m10 * m21 - m11 * m20, a21 = m01 * m20 - m00 * m21, a22 = m00 * m11 - m01 * m10;
    const double det = m00 * a00 + m01 * a10 + m02 * a20;
    const double id = det != 0.0 ? 1.0 / det : 0.0;
    x0 = make_float4((float)a00, (float)a01, (float)a02, (float)det);
    x1 = make_float4((float)a10, (float)a11, (float)a12, (float)(a20 * id));
    x2 = make_float2((float)(a21 * id), (float)(a22 * id));
}

// The rest-edge constants of entry e of a table laid out as dict_ext (global
// memory or LDS); zero, and not read, for StVK, which does not use them, and
// when edge is false (a solver without the table).
struct EdgeK {
    float4 x0, x1;
    float2 x2;
};

template <int MAT>
__device__ __forceinline__ EdgeK edge_k(const float4 *ext, uint32_t e, bool edge = true)
{
    EdgeK k;
    if (MAT == kSNH && edge) {
        const float4 *p = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(ext) + e * 48u);
        k.x0 = p[0];
        k.x1 = p[1];
        k.x2 = *reinterpret_cast<const float2 *>(p + 2);
    } else {
        k.x0 = k.x1 = make_float4(0.f, 0.f, 0.f, 0.f);
        k.x2 = make_float2(0.f, 0.f);
    }
    return k;
}

template <int MAT>
__device__ __forceinline__ void slot_eval(float ka, float kb, float kc, float bb, const float3 bs, const float3 b0,
                                          const float3 b1, const EdgeK &ek, bool edge, const float3 e0,
                                          const float3 e1, const float3 d2, Acc &a)
{
    const float4 x0 = ek.x0, x1 = ek.x1;
    const float2 x2 = ek.x2;
    // D (row r, column c)
    const float D00 = e0.x * b0.x + e1.x * b1.x - d2.x * bs.x;
    const float D01 = e0.x * b0.y + e1.x * b1.y - d2.x * bs.y;
    const float D02 = e0.x * b0.z + e1.x * b1.z - d2.x * bs.z;
    const float D10 = e0.y * b0.x + e1.y * b1.x - d2.y * bs.x;
    const float D11 = e0.y * b0.y + e1.y * b1.y - d2.y * bs.y;
    const float D12 = e0.y * b0.z + e1.y * b1.z - d2.y * bs.z;
    const float D20 = e0.z * b0.x + e1.z * b1.x - d2.z * bs.x;
    const float D21 = e0.z * b0.y + e1.z * b1.y - d2.z * bs.y;
    const float D22 = e0.z * b0.z + e1.z * b1.z - d2.z * bs.z;
    const float F00 = 1.f + D00, F11 = 1.f + D11, F22 = 1.f + D22;
    const float F01 = D01, F02 = D02, F10 = D10, F12 = D12, F20 = D20, F21 = D21;
    const float Fb0 = F00 * bs.x + F01 * bs.y + F02 * bs.z;
    const float Fb1 = F10 * bs.x + F11 * bs.y + F12 * bs.z;
    const float Fb2 = F20 * bs.x + F21 * bs.y + F22 * bs.z;
    if constexpr (MAT == kSNH) {
        const float Ic = F00 * F00 + F01 * F01 + F02 * F02 + F10 * F10 + F11 * F11 + F12 * F12 + F20 * F20 +
                         F21 * F21 + F22 * F22;
        float cb0, cb1, cb2, Jm1;
        if (edge) {
            // cof(F) b = dJ/dx_self = k A x C with A = E0 + e0 and C = E1 + e1
            // the deformed edges, so cb = bs + c with the change
            //   c = k (A x C - E0 x E1) = P x e1 + e0 x (Q + k e1),
            // formed from the small differences alone.  J is affine in x_self
            // and zero on the face (o0, o1, o2): J = cb . (R - d2), and with
            // bs . R = 1
            //   J - 1 = c . R - cb . d2,
            // accurate for small D.  15 VALU instructions fewer than below.
            const float w0 = fmaf(x0.w, e1.x, x1.x), w1 = fmaf(x0.w, e1.y, x1.y), w2 = fmaf(x0.w, e1.z, x1.z);
            const float c0 = x0.y * e1.z - x0.z * e1.y + e0.y * w2 - e0.z * w1;
            const float c1 = x0.z * e1.x - x0.x * e1.z + e0.z * w0 - e0.x * w2;
            const float c2 = x0.x * e1.y - x0.y * e1.x + e0.x * w1 - e0.y * w0;
            cb0 = bs.x + c0;
            cb1 = bs.y + c1;
            cb2 = bs.z + c2;
            Jm1 = fmaf(-cb2, d2.z, fmaf(-cb1, d2.y, fmaf(-cb0, d2.x, c0 * x1.w + c1 * x2.x + c2 * x2.y)));
        } else {
            // J - 1 = tr D + (principal 2x2 minors of D) + det D, accurate for small D
            const float m2 = (D00 * D11 - D01 * D10) + (D00 * D22 - D02 * D20) + (D11 * D22 - D12 * D21);
            const float detD = D00 * (D11 * D22 - D12 * D21) - D01 * (D10 * D22 - D12 * D20) +
                               D02 * (D10 * D21 - D11 * D20);
            Jm1 = (D00 + D11 + D22) + m2 + detD;
            // cof(F) b.  Row i of cof(F) = J F^-T is F_j x F_k (rows of F,
            // (i,j,k) cyclic).  With p = F_1 x b the scalar triple products give
            //   cb0 = (F_1 x F_2) . b = -F_2 . p,   cb2 = (F_0 x F_1) . b = F_0 . p,
            // and cb1 = (F_2 x F_0) . b directly.
            const float p0 = F11 * bs.z - F12 * bs.y, p1 = F12 * bs.x - F10 * bs.z, p2 = F10 * bs.y - F11 * bs.x;
            cb0 = -(F20 * p0 + F21 * p1 + F22 * p2);
            cb2 = F00 * p0 + F01 * p1 + F02 * p2;
            cb1 = (F21 * F02 - F22 * F01) * bs.x + (F22 * F00 - F20 * F02) * bs.y + (F20 * F01 - F21 * F00) * bs.z;
        }
        // hardware reciprocal (1 ulp): the SNH weights need no correctly rounded divide
        const float r = __builtin_amdgcn_rcpf(Ic + 1.f);
        const float s1 = ka * (Ic * r);        // V mu'(1 - 1/(Ic+1))
        const float s2 = 2.f * ka * r * r;     // V 2mu'/(Ic+1)^2
        const float sj = kb * Jm1 - kc;        // V lambda'(J - alpha)
        // accumulated term by term, each product fused into the running sum
        a.g0 = fmaf(sj, cb0, fmaf(s1, Fb0, a.g0));
        a.g1 = fmaf(sj, cb1, fmaf(s1, Fb1, a.g1));
        a.g2 = fmaf(sj, cb2, fmaf(s1, Fb2, a.g2));
        const float sd = s1 * bb;
        const float t0 = s2 * Fb0, t1 = s2 * Fb1, t2 = s2 * Fb2;
        const float q0 = kb * cb0, q1 = kb * cb1, q2 = kb * cb2;
        a.h00 = fmaf(q0, cb0, fmaf(t0, Fb0, a.h00 + sd));
        a.h01 = fmaf(q0, cb1, fmaf(t0, Fb1, a.h01));
        a.h02 = fmaf(q0, cb2, fmaf(t0, Fb2, a.h02));
        a.h11 = fmaf(q1, cb1, fmaf(t1, Fb1, a.h11 + sd));
        a.h12 = fmaf(q1, cb2, fmaf(t1, Fb2, a.h12));
        a.h22 = fmaf(q2, cb2, fmaf(t2, Fb2, a.h22 + sd));
    } else {
        // E = (D + D^T + D^T D) / 2
        const float E00 = D00 + 0.5f * (D00 * D00 + D10 * D10 + D20 * D20);
        const float E11 = D11 + 0.5f * (D01 * D01 + D11 * D11 + D21 * D21);
        const float E22 = D22 + 0.5f * (D02 * D02 + D12 * D12 + D22 * D22);
        const float E01 = 0.5f * (D01 + D10 + D00 * D01 + D10 * D11 + D20 * D21);
        const float E02 = 0.5f * (D02 + D20 + D00 * D02 + D10 * D12 + D20 * D22);
        const float E12 = 0.5f * (D12 + D21 + D01 * D02 + D11 * D12 + D21 * D22);
        const float trE = E00 + E11 + E22;
        const float Eb0 = E00 * bs.x + E01 * bs.y + E02 * bs.z;
        const float Eb1 = E01 * bs.x + E11 * bs.y + E12 * bs.z;
        const float Eb2 = E02 * bs.x + E12 * bs.y + E22 * bs.z;
        const float bEb = bs.x * Eb0 + bs.y * Eb1 + bs.z * Eb2;
        // w = S b (V folded)
        const float w0 = kb * trE * bs.x + 2.f * ka * Eb0;
        const float w1 = kb * trE * bs.y + 2.f * ka * Eb1;
        const float w2 = kb * trE * bs.z + 2.f * ka * Eb2;
        a.g0 = fmaf(F02, w2, fmaf(F01, w1, fmaf(F00, w0, a.g0)));
        a.g1 = fmaf(F12, w2, fmaf(F11, w1, fmaf(F10, w0, a.g1)));
        a.g2 = fmaf(F22, w2, fmaf(F21, w1, fmaf(F20, w0, a.g2)));
        const float kab = ka + kb, kbb = ka * bb;
        const float sd = kb * trE * bb + 2.f * ka * bEb;
        const float u0 = kab * Fb0, u1 = kab * Fb1, u2 = kab * Fb2;
        a.h00 = fmaf(u0, Fb0, fmaf(kbb, F00 * F00 + F01 * F01 + F02 * F02, a.h00 + sd));
        a.h01 = fmaf(u0, Fb1, fmaf(kbb, F00 * F10 + F01 * F11 + F02 * F12, a.h01));
        a.h02 = fmaf(u0, Fb2, fmaf(kbb, F00 * F20 + F01 * F21 + F02 * F22, a.h02));
        a.h11 = fmaf(u1, Fb1, fmaf(kbb, F10 * F10 + F11 * F11 + F12 * F12, a.h11 + sd));
        a.h12 = fmaf(u1, Fb2, fmaf(kbb, F10 * F20 + F11 * F21 + F12 * F22, a.h12));
        a.h22 = fmaf(u2, Fb2, fmaf(kbb, F20 * F20 + F21 * F21 + F22 * F22, a.h22 + sd));
    }
}

// a + b, component by component
__device__ __forceinline__ Acc acc_add(const Acc &a, const Acc &b)
{
    Acc r;
    r.g0 = a.g0 + b.g0;
    r.g1 = a.g1 + b.g1;
    r.g2 = a.g2 + b.g2;
    r.h00 = a.h00 + b.h00;
    r.h01 = a.h01 + b.h01;
    r.h02 = a.h02 + b.h02;
    r.h11 = a.h11 + b.h11;
    r.h12 = a.h12 + b.h12;
    r.h22 = a.h22 + b.h22;
    return r;
}

// x from another lane of the same 16-lane row by a DPP lane pattern
// (quad_perm, row mirror), without a trip through LDS.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF,
                                                                 false));
}
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // lane i -> 7-i within each 8 lanes
constexpr int kDppMirror = 0x140;      // lane i -> 15-i within each 16 lanes

// Sum over aligned groups of L lanes (L a power of two <= 64), the same value
// in every lane of a group: pairs within quads, then quads within 8 and 8s
// within 16 by mirroring, then shuffles across rows.
template <int L>
__device__ __forceinline__ float group_sum(float x)
{
    if constexpr (L >= 2) x += dpp_f32<kDppXor1>(x);
    if constexpr (L >= 4) x += dpp_f32<kDppXor2>(x);
    if constexpr (L >= 8) x += dpp_f32<kDppHalfMirror>(x);
    if constexpr (L >= 16) x += dpp_f32<kDppMirror>(x);
    if constexpr (L >= 32) x += __shfl_xor(x, 16, L);
    if constexpr (L >= 64) x += __shfl_xor(x, 32, 64);
    return x;
}

template <int L>
__device__ __forceinline__ void acc_group_sum(Acc &a)
{
    if constexpr (L > 1) {
        a.g0 = group_sum<L>(a.g0);
        a.g1 = group_sum<L>(a.g1);
        a.g2 = group_sum<L>(a.g2);
        a.h00 = group_sum<L>(a.h00);
        a.h01 = group_sum<L>(a.h01);
        a.h02 = group_sum<L>(a.h02);
        a.h11 = group_sum<L>(a.h11);
        a.h12 = group_sum<L>(a.h12);
        a.h22 = group_sum<L>(a.h22);
    }
}

// group_sum for aligned groups of 1 << lg lanes, lg in {0, 1, 2} per lane:
// group_sum<4>'s two DPP stages, each kept only by the lanes whose group
// spans it (same additions in the same order as group_sum<2> / <4>).
__device__ __forceinline__ float group_sum_lanes(float x, int lg)
{
    const float y = dpp_f32<kDppXor1>(x);
    x = lg >= 1 ? x + y : x;
    const float z = dpp_f32<kDppXor2>(x);
    return lg >= 2 ? x + z : x;
}

__device__ __forceinline__ void acc_group_sum_lanes(Acc &a, int lg)
{
    a.g0 = group_sum_lanes(a.g0, lg);
    a.g1 = group_sum_lanes(a.g1, lg);
    a.g2 = group_sum_lanes(a.g2, lg);
    a.h00 = group_sum_lanes(a.h00, lg);
    a.h01 = group_sum_lanes(a.h01, lg);
    a.h02 = group_sum_lanes(a.h02, lg);
    a.h11 = group_sum_lanes(a.h11, lg);
    a.h12 = group_sum_lanes(a.h12, lg);
    a.h22 = group_sum_lanes(a.h22, lg);
}

// Solve H dx = -G for a symmetric 3x3 block by the adjugate.  A block whose
// determinant is finite but tiny relative to its diagonal product is skipped
// (dx = 0): this is the guard for vertices without incident tets and for
// indefinite StVK blocks.  NaN or Inf in G or H propagates into dx.
__device__ __forceinline__ float3 solve3_guarded(const Acc &a)
{
    const float c00 = a.h11 * a.h22 - a.h12 * a.h12;
    const float c01 = a.h02 * a.h12 - a.h01 * a.h22;
    const float c02 = a.h01 * a.h12 - a.h02 * a.h11;
    const float c11 = a.h00 * a.h22 - a.h02 * a.h02;
    const float c12 = a.h01 * a.h02 - a.h00 * a.h12;
    const float c22 = a.h00 * a.h11 - a.h01 * a.h01;
    const float det = a.h00 * c00 + a.h01 * c01 + a.h02 * c02;
    const float scale = fabsf(a.h00 * a.h11 * a.h22);
    if (fabsf(det) <= 1e-6f * scale && det == det && fabsf(det) != INFINITY &&
        a.g0 == a.g0 && a.g1 == a.g1 && a.g2 == a.g2)
        return make_float3(0.f, 0.f, 0.f);
    // hardware reciprocal (1 ulp): the solve needs no correctly rounded divide
    const float id = -__builtin_amdgcn_rcpf(det);
    return make_float3(id * (c00 * a.g0 + c01 * a.g1 + c02 * a.g2), id * (c01 * a.g0 + c11 * a.g1 + c12 * a.g2),
                       id * (c02 * a.g0 + c12 * a.g1 + c22 * a.g2));
}

__device__ __forceinline__ float wave_max(float x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmaxf(x, __shfl_xor(x, off, 64));
    return x;
}

// |dx_x + dx_y + dx_z| as the ordered bit pattern of a non-negative float (NaN
// sorts above +Inf, so an unsigned max keeps it).  src/vbd_solver.py:578.
__device__ __forceinline__ uint32_t metric_bits(float3 dx)
{
    float s = fabsf(dx.x + dx.y + dx.z);
    return __float_as_uint(s);
}

// Max over the wave (every lane active); lane 0 gets the result.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x)
{
    x = max(x, dpp_u32<kDppXor1>(x));
    x = max(x, dpp_u32<kDppXor2>(x));
    x = max(x, dpp_u32<kDppHalfMirror>(x));
    x = max(x, dpp_u32<kDppMirror>(x));
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)x, 0);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)x, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)x, 32);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)x, 48);
    return max(max(r0, r1), max(r2, r3));
}

} // namespace vbd

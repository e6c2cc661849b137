// vbd_setup.hip -- device-side mesh preprocessing for the VBD solver.
//
// Replaces VBDSolver.__init__ (src/vbd_solver.py:367-499) and the colouring of
// src/coloring.py:24-72.  Everything below runs on the GPU:
//   1. per-tet Lame parameters, rest shape inverse, volume, mass (FP64)
//        src/vbd_solver.py:295-362
//   2. vertex -> incident tet CSR, ascending tet order per vertex
//        src/vbd_solver.py:485-498
//   3. first-fit colouring, bit-exact with the sequential reference: the
//      recurrence colour[v] = mex{colour[u] : u ~ v, u < v} is iterated in
//      parallel until a round changes nothing; its fixed point is unique (by
//      induction on v) and equals the sequential result.
//   4. colour-major renumbering (active vertices of colour 0 ascending, then
//      colour 1, ..., then inactive/uncoloured vertices), so each colour sweep
//      is one contiguous range
//   5. 64-byte slot records in colour-major order (layout in vbd_device.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>

#include <rocprim/rocprim.hpp>

#include "vbd_device.h"
#include "vbd_internal.h"

namespace vbd {

namespace {

constexpr int kTB = 256;

inline int nblk(int64_t n, int tb = kTB) { return (int)((n + tb - 1) / tb); }

template <typename T>
int dmalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return VBD_ENOMEM;
    }
    return 0;
}

struct Scratch {
    std::vector<void *> ptrs;
    ~Scratch()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    int get(T **p, size_t count)
    {
        int r = dmalloc(p, count);
        if (r == 0) ptrs.push_back((void *)*p);
        return r;
    }
};

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// incidence counts and the largest referenced vertex
__global__ void k_count_incidence(const int32_t *tets, int m, int32_t *cnt, int32_t *maxref)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    int mx = -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int v = tets[4 * e + k];
        atomicAdd(&cnt[v], 1);
        mx = max(mx, v);
    }
    atomicMax(maxref, mx);
}

__global__ void k_fill_incidence(const int32_t *tets, int m, int32_t *cursor, int32_t *inc)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int v = tets[4 * e + k];
        int pos = atomicAdd(&cursor[v], 1);
        inc[pos] = (e << 2) | k;
    }
}

// Sort each vertex's incidence list by tet id (the reference lists incident
// elements in ascending index, src/vbd_solver.py:488-491) and record the
// largest list.
__global__ void k_sort_incidence(const int32_t *off, int32_t *inc, int n, int32_t *maxinc)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int a = off[v], b = off[v + 1];
    for (int i = a + 1; i < b; ++i) {
        int x = inc[i];
        int j = i - 1;
        while (j >= a && inc[j] > x) {
            inc[j + 1] = inc[j];
            --j;
        }
        inc[j + 1] = x;
    }
    atomicMax(maxinc, b - a);
}

// Neighbour-set size of each vertex (compute_adjacency_dict's degree,
// src/coloring.py:24-38, including the self entry a degenerate element adds).
// A neighbour is counted at its first occurrence in the vertex's incidence list.
__global__ void k_degree(const int32_t *tets, const int32_t *off, const int32_t *inc, int n, int32_t *maxdeg)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int a = off[v], b = off[v + 1];
    int cnt = 0;
    bool self = false;
    for (int s = a; s < b; ++s) {
        int e = inc[s] >> 2;
        int rep = 0;
        for (int k = 0; k < 4; ++k) {
            int u = tets[4 * e + k];
            if (u == v) {
                ++rep;
                continue;
            }
            bool seen = false;
            for (int k2 = 0; k2 < k && !seen; ++k2) seen = tets[4 * e + k2] == u;
            for (int s2 = a; s2 < s && !seen; ++s2) {
                int e2 = inc[s2] >> 2;
                for (int k2 = 0; k2 < 4; ++k2) seen |= tets[4 * e2 + k2] == u;
            }
            if (!seen) ++cnt;
        }
        self |= rep > 1;
    }
    atomicMax(maxdeg, cnt + (self ? 1 : 0));
}

// Distinct lower-indexed neighbours of each vertex (the only ones first-fit
// consults): counted in the first pass, written in the second.
template <bool FILL>
__global__ void k_lower_neighbours(const int32_t *tets, const int32_t *off, const int32_t *inc, int n, int32_t *cnt,
                                   const int32_t *lo_off, int32_t *lo)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int a = off[v], b = off[v + 1];
    int k = 0;
    int w = FILL ? lo_off[v] : 0;
    for (int s = a; s < b; ++s) {
        int e = inc[s] >> 2;
        for (int q = 0; q < 4; ++q) {
            int u = tets[4 * e + q];
            if (u >= v) continue;
            bool seen = false;
            for (int q2 = 0; q2 < q && !seen; ++q2) seen = tets[4 * e + q2] == u;
            for (int s2 = a; s2 < s && !seen; ++s2) {
                int e2 = inc[s2] >> 2;
                for (int q2 = 0; q2 < 4; ++q2) seen |= tets[4 * e2 + q2] == u;
            }
            if (seen) continue;
            if (FILL) lo[w + k] = u;
            ++k;
        }
    }
    if (!FILL) cnt[v] = k;
}

__global__ void k_color_init(int32_t *colors, int n, int nref)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) colors[v] = v < nref ? 0 : -1;
}

// One parallel round of colour[v] = mex{colour[u] : u in adj(v), u < v} over an
// adjacency list (the mesh's lower-neighbour lists, or graph_coloring's dict
// input, src/coloring.py:40-72): only neighbours with a smaller index are
// already coloured when sequential first-fit visits v.
__global__ void k_color_round_graph(const int32_t *off, const int32_t *adj, int n, int32_t *colors,
                                    int32_t *changed, int32_t *overflow)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t used[kMaxColors / 32];
#pragma unroll
    for (int i = 0; i < kMaxColors / 32; ++i) used[i] = 0;
    for (int s = off[v]; s < off[v + 1]; ++s) {
        int u = adj[s];
        if (u >= 0 && u < v) {
            int c = colors[u];
            if (c >= 0 && c < kMaxColors) used[c >> 5] |= 1u << (c & 31);
        }
    }
    int c = kMaxColors;
    for (int w = 0; w < kMaxColors / 32; ++w) {
        uint32_t freebits = ~used[w];
        if (freebits) {
            c = w * 32 + __builtin_ctz(freebits);
            break;
        }
    }
    if (c == kMaxColors) {
        atomicExch(overflow, 1);
        c = kMaxColors - 1;
    }
    const bool ch = c != colors[v];
    if (ch) __atomic_store_n(&colors[v], c, __ATOMIC_RELAXED);
    // one plain store per wave that changed something: a device atomic per
    // vertex serialises on the flag's address
    if (__any(ch) && (threadIdx.x & 63) == 0) *(volatile int32_t *)changed = 1;
}

// Per-tet rest shape in FP64: Dm = [X0-X3 | X1-X3 | X2-X3], invDm, V = |det|/6,
// m = rho V (src/vbd_solver.py:317-334); Lame mu, lambda (:356-362).
__global__ void k_tet_setup(const double *X, const int32_t *tets, int m, const double *rho, const double *E,
                            const double *nu, double *invDm, double *vol, double *mass, double *mu, double *lam,
                            int32_t *singular)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    int t0 = tets[4 * e], t1 = tets[4 * e + 1], t2 = tets[4 * e + 2], t3 = tets[4 * e + 3];
    double D[3][3];
    for (int r = 0; r < 3; ++r) {
        double x3 = X[3 * t3 + r];
        D[r][0] = X[3 * t0 + r] - x3;
        D[r][1] = X[3 * t1 + r] - x3;
        D[r][2] = X[3 * t2 + r] - x3;
    }
    double c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
    double c01 = D[0][2] * D[2][1] - D[0][1] * D[2][2];
    double c02 = D[0][1] * D[1][2] - D[0][2] * D[1][1];
    double c10 = D[1][2] * D[2][0] - D[1][0] * D[2][2];
    double c11 = D[0][0] * D[2][2] - D[0][2] * D[2][0];
    double c12 = D[0][2] * D[1][0] - D[0][0] * D[1][2];
    double c20 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
    double c21 = D[0][1] * D[2][0] - D[0][0] * D[2][1];
    double c22 = D[0][0] * D[1][1] - D[0][1] * D[1][0];
    double det = D[0][0] * c00 + D[1][0] * c01 + D[2][0] * c02;
    if (!(det != 0.0) || !isfinite(det)) atomicExch(singular, 1);
    double s = det != 0.0 ? 1.0 / det : 0.0;
    double *o = invDm + 9 * (size_t)e;
    o[0] = c00 * s; o[1] = c01 * s; o[2] = c02 * s;
    o[3] = c10 * s; o[4] = c11 * s; o[5] = c12 * s;
    o[6] = c20 * s; o[7] = c21 * s; o[8] = c22 * s;
    double V = fabs(det) / 6.0;
    vol[e] = V;
    mass[e] = rho[e] * V;
    mu[e] = E[e] / (2.0 * (1.0 + nu[e]));
    lam[e] = (E[e] * nu[e]) / ((1.0 + nu[e]) * (1.0 - 2.0 * nu[e]));
}

// Sort key of the colour-major order: colour-major, body-minor, then (with
// ncls = 3) the lane class of the vertex's slot count; vertices that are not
// swept (pinned or uncoloured) go after every colour.  Vertices of one colour
// are not neighbours, so their order within a colour does not change a sweep.
__global__ void k_sort_keys(const int32_t *colors, const uint8_t *active, const int32_t *body_of, int nb, int n,
                            const int32_t *inc_off, int ncls, uint32_t *keys, int32_t *vals)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int c = colors[v];
    const bool swept = active[v] && c >= 0;
    uint32_t ck = swept ? (uint32_t)c : (uint32_t)kMaxColors;
    const int cls = ncls > 1 && swept ? lane_class(inc_off[v + 1] - inc_off[v]) : 0;
    keys[v] = (ck * (uint32_t)nb + (uint32_t)body_of[v]) * (uint32_t)ncls + (uint32_t)cls;
    vals[v] = v;
}

__global__ void k_gather_i32(const int32_t *src, const int32_t *idx, int n, int32_t *dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_inverse_perm(const int32_t *perm, int n, int32_t *iperm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = i;
}

// Run lengths of the sorted keys: each run's first element records its start,
// its last its end (no atomics; absent keys keep 0).
__global__ void k_key_runs(const uint32_t *keys_sorted, int n, int32_t *first, int32_t *end)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys_sorted[i];
    if (i == 0 || keys_sorted[i - 1] != k) first[k] = i;
    if (i == n - 1 || keys_sorted[i + 1] != k) end[k] = i + 1;
}

__global__ void k_slot_counts(const int32_t *perm, const int32_t *inc_off, int nswept, int32_t *cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nswept) return;
    if (i == nswept) {
        cnt[i] = 0;
        return;
    }
    int v = perm[i];
    cnt[i] = inc_off[v + 1] - inc_off[v];
}

// Lumped vertex mass sum_e m_e/4 in ascending tet order (src/vbd_solver.py:32)
// for every vertex, in colour-major order.
__global__ void k_vertex_mass(const int32_t *perm, const int32_t *inc_off, const int32_t *inc, const double *mass_e,
                              int n, float *mass_v, const uint8_t *active_orig, uint8_t *active_new)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = perm[i];
    double M = 0.0;
    for (int s = inc_off[v]; s < inc_off[v + 1]; ++s) M += mass_e[inc[s] >> 2] / 4.0;
    mass_v[i] = (float)M;
    active_new[i] = active_orig[v];
}

// Material constants of one tet: SNH ka = V mu', kb = V lambda', kc = V
// lambda'(alpha - 1) = 0.75 ka; StVK ka = V mu, kb = V lambda, kc = 0.
__device__ __forceinline__ void tet_material(int material, double V, double mu, double lam, float &ka, float &kb,
                                             float &kc)
{
    if (material == kSNH) {
        const double mup = 4.0 * mu / 3.0;       // src/vbd_solver.py:108
        const double lamp = lam + 5.0 * mu / 6.0;  // :109
        ka = (float)(V * mup);
        kb = (float)(V * lamp);
        kc = 0.75f * ka;  // alpha from :110
    } else {
        ka = (float)(V * mu);
        kb = (float)(V * lam);
        kc = 0.f;
    }
}

__device__ __forceinline__ void brow(const double *iD, int k, double b[3])
{
    if (k < 3) {
        b[0] = iD[3 * k];
        b[1] = iD[3 * k + 1];
        b[2] = iD[3 * k + 2];
    } else {
        b[0] = -(iD[0] + iD[3] + iD[6]);
        b[1] = -(iD[1] + iD[4] + iD[7]);
        b[2] = -(iD[2] + iD[5] + iD[8]);
    }
}

// 56-byte slot records (layout documented in vbd_device.h): qa|qb|qc in
// slots, qd in slots2.
__global__ void k_build_slots(const int32_t *perm, const int32_t *iperm, const int32_t *inc_off, const int32_t *inc,
                              const int32_t *tets, const double *invDm, const double *vol, const double *mu,
                              const double *lam, int material, int nswept, const int32_t *slot_off, float4 *slots,
                              float2 *slots2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nswept) return;
    int v = perm[i];
    int s = slot_off[i];
    for (int q = inc_off[v]; q < inc_off[v + 1]; ++q, ++s) {
        int e = inc[q] >> 2, loc = inc[q] & 3;
        const double *iD = invDm + 9 * (size_t)e;
        int other[3], n_o = 0;
        for (int k = 0; k < 4; ++k)
            if (k != loc) other[n_o++] = k;
        double bs[3], b0[3], b1[3];
        brow(iD, loc, bs);
        brow(iD, other[0], b0);
        brow(iD, other[1], b1);
        float ka, kb, kc;
        tet_material(material, vol[e], mu[e], lam[e], ka, kb, kc);
        int o0 = iperm[tets[4 * e + other[0]]];
        int o1 = iperm[tets[4 * e + other[1]]];
        int o2 = iperm[tets[4 * e + other[2]]];
        (void)kc;  // follows from ka and the material (vbd_device.h)
        slots[3 * (size_t)s] = make_float4(__int_as_float(o0), __int_as_float(o1), __int_as_float(o2), ka);
        slots[3 * (size_t)s + 1] = make_float4((float)bs[0], (float)bs[1], (float)bs[2], kb);
        slots[3 * (size_t)s + 2] = make_float4((float)b0[0], (float)b0[1], (float)b0[2], (float)b1[0]);
        slots2[s] = make_float2((float)b1[1], (float)b1[2]);
    }
}

// 64-byte tet records (VBD_TET_RECORDS layout, documented in vbd_device.h).
__global__ void k_build_tet_records(int m, const int32_t *tets, const int32_t *iperm, const double *invDm,
                                    const double *vol, const double *mu, const double *lam, int material,
                                    float4 *rec)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const double *iD = invDm + 9 * (size_t)e;
    float ka, kb, kc;
    tet_material(material, vol[e], mu[e], lam[e], ka, kb, kc);
    rec[4 * (size_t)e] = make_float4(__int_as_float(iperm[tets[4 * e]]), __int_as_float(iperm[tets[4 * e + 1]]),
                                     __int_as_float(iperm[tets[4 * e + 2]]), __int_as_float(iperm[tets[4 * e + 3]]));
    rec[4 * (size_t)e + 1] = make_float4((float)iD[0], (float)iD[1], (float)iD[2], ka);
    rec[4 * (size_t)e + 2] = make_float4((float)iD[3], (float)iD[4], (float)iD[5], kb);
    rec[4 * (size_t)e + 3] = make_float4((float)iD[6], (float)iD[7], (float)iD[8], kc);
}

// Per-slot entries in colour-major vertex order: the incidence CSR entries
// (tet << 2 | local index, ascending tet) of each swept vertex.
__global__ void k_build_slot_entries(const int32_t *perm, const int32_t *inc_off, const int32_t *inc, int nswept,
                                     const int32_t *slot_off, int32_t *ent)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nswept) return;
    const int v = perm[i];
    int s = slot_off[i];
    for (int q = inc_off[v]; q < inc_off[v + 1]; ++q, ++s) ent[s] = inc[q];
}

// ----------------------------------------------------------------- record dictionary
// 64-bit hash of a slot record's 11 constant words (ka, qb, qc, qd).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xbf58476d1ce4e5b9ULL;
    z ^= z >> 27;
    z *= 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ void record_words(const float4 *slots, const float2 *slots2, int64_t s, uint32_t w[11])
{
    const float4 qa = slots[3 * s], qb = slots[3 * s + 1], qc = slots[3 * s + 2];
    const float2 qd = slots2[s];
    const float f[11] = {qa.w, qb.x, qb.y, qb.z, qb.w, qc.x, qc.y, qc.z, qc.w, qd.x, qd.y};
#pragma unroll
    for (int k = 0; k < 11; ++k) w[k] = __float_as_uint(f[k]);
}

__global__ void k_record_hash(int64_t ns, const float4 *slots, const float2 *slots2, uint64_t *h, int32_t *idx)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    uint32_t w[11];
    record_words(slots, slots2, s, w);
    uint64_t x = 0x9e3779b97f4a7c15ULL;
#pragma unroll
    for (int k = 0; k < 11; ++k) x = mix64(x ^ ((uint64_t)w[k] + 0x632be59bd9b4e019ULL * (uint64_t)(k + 1)));
    h[s] = x;
    idx[s] = (int32_t)s;
}

// In hash order: head[i] = 1 where a new hash starts; a record equal in hash
// but not in content to its predecessor sets *collide.
__global__ void k_record_heads(int64_t ns, const uint64_t *hs, const int32_t *is, const float4 *slots,
                               const float2 *slots2, int32_t *head, int32_t *collide)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const bool h = i == 0 || hs[i] != hs[i - 1];
    head[i] = h ? 1 : 0;
    if (!h) {
        uint32_t a[11], b[11];
        record_words(slots, slots2, is[i], a);
        record_words(slots, slots2, is[i - 1], b);
        bool same = true;
#pragma unroll
        for (int k = 0; k < 11; ++k) same = same && a[k] == b[k];
        if (!same) atomicExch(collide, 1);
    }
}

// Compact records and shared dictionary entries (entry of hash position i:
// pos[i] + head[i] - 1, pos the exclusive scan of head).
__global__ void k_record_dict(int64_t ns, const int32_t *is, const int32_t *head, const int32_t *pos,
                              const float4 *slots, const float2 *slots2, float4 *compact, float4 *dict)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int64_t s = is[i];
    const int k = pos[i] + head[i] - 1;
    const float4 qa = slots[3 * s];
    compact[s] = make_float4(qa.x, qa.y, qa.z, __int_as_float(k));
    if (head[i]) {
        const float2 qd = slots2[s];
        dict[3 * (size_t)k] = slots[3 * s + 1];
        dict[3 * (size_t)k + 1] = slots[3 * s + 2];
        const float4 qb = slots[3 * s + 1];
        dict[3 * (size_t)k + 2] = make_float4(qd.x, qd.y, qa.w, slot_bb(make_float3(qb.x, qb.y, qb.z)));
    }
}

// Compact records with one dictionary entry per slot, in slot order.
__global__ void k_record_identity(int64_t ns, const float4 *slots, const float2 *slots2, float4 *compact,
                                  float4 *dict)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const float4 qa = slots[3 * s];
    const float2 qd = slots2[s];
    compact[s] = make_float4(qa.x, qa.y, qa.z, __int_as_float((int)s));
    dict[3 * s] = slots[3 * s + 1];
    dict[3 * s + 1] = slots[3 * s + 2];
    const float4 qb = slots[3 * s + 1];
    dict[3 * s + 2] = make_float4(qd.x, qd.y, qa.w, slot_bb(make_float3(qb.x, qb.y, qb.z)));
}

// Rest-edge constants of each dictionary entry (layout and derivation in
// vbd_device.h), from the entry's b vectors.
__global__ void k_dict_ext(int64_t nd, const float4 *dict, float4 *ext)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nd) return;
    const float4 d0 = dict[3 * k], d1 = dict[3 * k + 1], d2 = dict[3 * k + 2];
    float4 x0, x1;
    float2 x2;
    edge_consts(make_float3(d0.x, d0.y, d0.z), make_float3(d1.x, d1.y, d1.z), make_float3(d1.w, d2.x, d2.y), x0, x1,
                x2);
    ext[3 * k] = x0;
    ext[3 * k + 1] = x1;
    ext[3 * k + 2] = make_float4(x2.x, x2.y, 0.f, 0.f);
}

// Distinct neighbour vertices gathered by one colour's sweep (for the byte model).
__global__ void k_mark_boundary(const float4 *slots, const int32_t *ent, const int32_t *slot_off, int v0, int v1,
                                int tag, int32_t *mark, int32_t *count)
{
    int i = v0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v1) return;
    int local = 0;
    for (int s = slot_off[i]; s < slot_off[i + 1]; ++s) {
        int o[3];
#if VBD_TET_RECORDS
        const float4 r0 = slots[4 * (size_t)(ent[s] >> 2)];
        const int loc = ent[s] & 3;
        const int ids[4] = {__float_as_int(r0.x), __float_as_int(r0.y), __float_as_int(r0.z), __float_as_int(r0.w)};
        for (int k = 0, n_o = 0; k < 4; ++k)
            if (k != loc) o[n_o++] = ids[k];
#else
        (void)ent;
        const float4 q0 = slots[3 * (size_t)s];
        o[0] = __float_as_int(q0.x);
        o[1] = __float_as_int(q0.y);
        o[2] = __float_as_int(q0.z);
#endif
        for (int k = 0; k < 3; ++k)
            if (atomicExch(&mark[o[k]], tag) != tag) ++local;
    }
    if (local) atomicAdd(count, local);
}

int exclusive_scan_i32(const int32_t *in, int32_t *out, int n, hipStream_t st, Scratch &scr);

// Turn the full slot records (3 float4 + 1 float2 per slot, as built) into
// compact records and a record dictionary: shared entries when the distinct
// constant sets are at most a quarter of the slots, no hash collides and
// `share` is set; one entry per slot otherwise.
int build_record_dictionary(vbd_solver *s, bool share, hipStream_t st, Scratch &scr)
{
    const int64_t ns = s->n_slots;
    float4 *compact = nullptr, *dict = nullptr;
    auto install = [&](int64_t nd) -> int {
        (void)hipFree(s->slots);
        (void)hipFree(s->slots2);
        s->slots = compact;
        s->slots2 = nullptr;
        s->dict = dict;
        s->n_dict = (int32_t)nd;
        s->dict_shared = nd < ns;
        // rest-edge constants where the sweeps read the dictionary from LDS
        // (the staged and tail sweeps need them there)
        (void)hipFree(s->dict_ext);
        s->dict_ext = nullptr;
        if (!g_edge_form || !s->dict_shared || (nd > kStageDict && !s->batched)) return 0;
        VBD_TRY(dmalloc(&s->dict_ext, 3 * (size_t)nd));
        k_dict_ext<<<nblk(nd), kTB, 0, st>>>(nd, s->dict, s->dict_ext);
        VBD_HIP(hipGetLastError());
        VBD_HIP(hipStreamSynchronize(st));
        return 0;
    };
    auto identity = [&]() -> int {
        VBD_TRY(dmalloc(&compact, (size_t)std::max<int64_t>(ns, 1)));
        if (dmalloc(&dict, 3 * (size_t)std::max<int64_t>(ns, 1))) {
            (void)hipFree(compact);
            return VBD_ENOMEM;
        }
        if (ns > 0) k_record_identity<<<nblk(ns), kTB, 0, st>>>(ns, s->slots, s->slots2, compact, dict);
        VBD_HIP(hipGetLastError());
        VBD_HIP(hipStreamSynchronize(st));
        return install(ns);
    };
    if (!share || ns < 4) return identity();
    uint64_t *h, *hs;
    int32_t *is0, *is, *head, *pos, *collide;
    VBD_TRY(scr.get(&h, ns));
    VBD_TRY(scr.get(&hs, ns));
    VBD_TRY(scr.get(&is0, ns));
    VBD_TRY(scr.get(&is, ns));
    VBD_TRY(scr.get(&head, ns));
    VBD_TRY(scr.get(&pos, ns));
    VBD_TRY(scr.get(&collide, 1));
    k_fill_i32<<<1, kTB, 0, st>>>(collide, 1, 0);
    k_record_hash<<<nblk(ns), kTB, 0, st>>>(ns, s->slots, s->slots2, h, is0);
    {
        size_t tb = 0;
        VBD_HIP(rocprim::radix_sort_pairs(nullptr, tb, h, hs, is0, is, (size_t)ns, 0, 64, st));
        void *tmp;
        VBD_TRY(scr.get((char **)&tmp, tb));
        VBD_HIP(rocprim::radix_sort_pairs(tmp, tb, h, hs, is0, is, (size_t)ns, 0, 64, st));
    }
    k_record_heads<<<nblk(ns), kTB, 0, st>>>(ns, hs, is, s->slots, s->slots2, head, collide);
    VBD_TRY(exclusive_scan_i32(head, pos, (int)ns, st, scr));
    int32_t last[2] = {0, 0}, col = 0;
    VBD_HIP(hipMemcpyAsync(&last[0], pos + ns - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipMemcpyAsync(&last[1], head + ns - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipMemcpyAsync(&col, collide, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    const int64_t nd = (int64_t)last[0] + last[1];
    if (col || nd < 1 || 4 * nd > ns) return identity();
    VBD_TRY(dmalloc(&compact, (size_t)ns));
    if (dmalloc(&dict, 3 * (size_t)nd)) {
        (void)hipFree(compact);
        return VBD_ENOMEM;
    }
    k_record_dict<<<nblk(ns), kTB, 0, st>>>(ns, is, head, pos, s->slots, s->slots2, compact, dict);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipStreamSynchronize(st));
    return install(nd);
}

int exclusive_scan_i32(const int32_t *in, int32_t *out, int n, hipStream_t st, Scratch &scr)
{
    size_t tb = 0;
    VBD_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), st));
    void *tmp = nullptr;
    VBD_TRY(scr.get((char **)&tmp, tb));
    VBD_HIP(rocprim::exclusive_scan(tmp, tb, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), st));
    return 0;
}

// Incidence CSR (ascending tet ids) + first-fit colouring on the device.
int build_incidence_and_colors(const int32_t *d_tets, int m, int n, int32_t *d_inc_off, int32_t *d_inc,
                               int32_t *d_colors, int *ncolors, int *max_incident, int *max_degree, int *nref,
                               bool enforce_valence, hipStream_t st, Scratch &scr)
{
    int32_t *cnt, *cursor, *scal;  // scal: maxref, maxinc, maxdeg, overflow
    VBD_TRY(scr.get(&cnt, (size_t)n + 1));
    VBD_TRY(scr.get(&cursor, (size_t)n + 1));
    VBD_TRY(scr.get(&scal, 4));
    k_fill_i32<<<nblk(n + 1), kTB, 0, st>>>(cnt, n + 1, 0);
    k_fill_i32<<<1, kTB, 0, st>>>(scal, 4, 0);
    k_fill_i32<<<1, 1, 0, st>>>(scal, 1, -1);
    if (m > 0) k_count_incidence<<<nblk(m), kTB, 0, st>>>(d_tets, m, cnt, scal);
    VBD_TRY(exclusive_scan_i32(cnt, d_inc_off, n + 1, st, scr));
    VBD_HIP(hipMemcpyAsync(cursor, d_inc_off, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToDevice, st));
    if (m > 0) k_fill_incidence<<<nblk(m), kTB, 0, st>>>(d_tets, m, cursor, d_inc);
    k_sort_incidence<<<nblk(n), kTB, 0, st>>>(d_inc_off, d_inc, n, scal + 1);
    int32_t h[4];
    VBD_HIP(hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    *nref = h[0] + 1;
    *max_incident = h[1];
    if (enforce_valence && h[1] > kMaxElementsPerVertex) {
        set_error("a vertex has " + std::to_string(h[1]) + " incident elements; at most 32 are supported "
                  "(MAX_ELEMENTS_PER_VERTEX)");
        return VBD_EVALENCE;
    }
    k_degree<<<nblk(n), kTB, 0, st>>>(d_tets, d_inc_off, d_inc, n, scal + 2);
    // lower-neighbour CSR, then parallel first-fit rounds over it
    int32_t *lo_cnt, *lo_off, *lo;
    VBD_TRY(scr.get(&lo_cnt, (size_t)n + 1));
    VBD_TRY(scr.get(&lo_off, (size_t)n + 1));
    k_fill_i32<<<nblk(n + 1), kTB, 0, st>>>(lo_cnt, n + 1, 0);
    k_lower_neighbours<false><<<nblk(n), kTB, 0, st>>>(d_tets, d_inc_off, d_inc, n, lo_cnt, nullptr, nullptr);
    VBD_TRY(exclusive_scan_i32(lo_cnt, lo_off, n + 1, st, scr));
    int32_t n_lo = 0;
    VBD_HIP(hipMemcpyAsync(&n_lo, lo_off + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    VBD_TRY(scr.get(&lo, (size_t)std::max(n_lo, 1)));
    k_lower_neighbours<true><<<nblk(n), kTB, 0, st>>>(d_tets, d_inc_off, d_inc, n, nullptr, lo_off, lo);
    k_color_init<<<nblk(n), kTB, 0, st>>>(d_colors, n, *nref);
    // rounds in batches; each round has its own change flag
    constexpr int kBatch = 32;
    int32_t *changed;
    VBD_TRY(scr.get(&changed, kBatch));
    int32_t hc[kBatch];
    const int max_rounds = std::max(64, 4 * (*nref) + 8);
    bool done = false;
    for (int r = 0; r < max_rounds && !done; r += kBatch) {
        k_fill_i32<<<1, kBatch, 0, st>>>(changed, kBatch, 0);
        for (int b = 0; b < kBatch; ++b)
            k_color_round_graph<<<nblk(*nref), kTB, 0, st>>>(lo_off, lo, *nref, d_colors, changed + b, scal + 3);
        VBD_HIP(hipMemcpyAsync(hc, changed, sizeof(hc), hipMemcpyDeviceToHost, st));
        VBD_HIP(hipStreamSynchronize(st));
        for (int b = 0; b < kBatch; ++b) done |= (hc[b] == 0);
    }
    // the overflow flag of the rounds above may come from a transient colour;
    // one more round at the fixed point decides it
    k_fill_i32<<<1, 1, 0, st>>>(scal + 3, 1, 0);
    k_color_round_graph<<<nblk(*nref), kTB, 0, st>>>(lo_off, lo, *nref, d_colors, changed, scal + 3);
    VBD_HIP(hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    if (!done) {
        set_error("colouring did not reach its fixed point");
        return VBD_EHIP;
    }
    if (h[3]) {
        set_error("a vertex needs more than 256 colours");
        return VBD_ECOLORS;
    }
    *max_degree = h[2];
    // number of colours = max colour + 1
    std::vector<int32_t> col(n);
    VBD_HIP(hipMemcpyAsync(col.data(), d_colors, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    int mx = -1;
    for (int v = 0; v < n; ++v) mx = std::max(mx, col[v]);
    *ncolors = mx + 1;
    return 0;
}

#if VBD_TET_RECORDS
// Scale every tet record's material constants from one uniform material to
// another.
__global__ void k_rescale_material(int64_t ns, float4 *slots, float fa, float fb, int material)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    float4 *r = slots + 4 * s;
    r[1].w *= fa;
    r[2].w *= fb;
    r[3].w = material == kSNH ? 0.75f * r[1].w : 0.f;
}
#endif

// Scale the material constants of dictionary entries: ka in d2.z, kb in d0.w.
__global__ void k_rescale_dict(int64_t nd, float4 *dict, float fa, float fb)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nd) return;
    dict[3 * k + 2].z *= fa;
    dict[3 * k].w *= fb;
}

}  // namespace

bool g_record_dictionary = true;  // vbd_use_record_dictionary
bool g_staging = true;            // vbd_use_neighbour_staging
bool g_tail = true;               // vbd_use_tail_sweep
bool g_edge_form = true;          // vbd_use_edge_form

int rescale_material(vbd_solver *s, double E0, double nu0, double E1, double nu1)
{
    auto lame = [](double E, double nu, double &mu, double &lam) {
        mu = E / (2.0 * (1.0 + nu));
        lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    };
    double mu0, lam0, mu1, lam1;
    lame(E0, nu0, mu0, lam0);
    lame(E1, nu1, mu1, lam1);
    // new / old, 1 when both are zero (the constants stay zero)
    auto ratio = [](double a1, double a0) { return a0 == 0.0 && a1 == 0.0 ? 1.0 : a1 / a0; };
    double fa, fb;
    if (s->material == kSNH) {
        fa = ratio(mu1, mu0);  // mu' = 4mu/3
        fb = ratio(lam1 + 5.0 * mu1 / 6.0, lam0 + 5.0 * mu0 / 6.0);
    } else {
        fa = ratio(mu1, mu0);
        fb = ratio(lam1, lam0);
    }
    // The constants are rescaled in place, which needs the old parameters to
    // be nonzero (a solver created with nu = 0 has lambda = 0, and StVK scales
    // by lambda1 / lambda0); such a solver keeps its material.
    if (!std::isfinite(fa) || !std::isfinite(fb)) {
        set_error("set_material: the solver's current material has a zero Lame parameter (E = 0, or nu = 0 "
                  "with StVK), so its constants cannot be rescaled; create a solver with the new material");
        return VBD_EINVAL;
    }
#if VBD_TET_RECORDS
    const int64_t nrec = s->m;
#else
    const int64_t nrec = s->n_slots;
#endif
#if VBD_TET_RECORDS
    if (nrec > 0)
        k_rescale_material<<<nblk(nrec), kTB, 0, s->stream>>>(nrec, s->slots, (float)fa, (float)fb, s->material);
#else
    (void)nrec;
    if (s->n_dict > 0)
        k_rescale_dict<<<nblk(s->n_dict), kTB, 0, s->stream>>>(s->n_dict, s->dict, (float)fa, (float)fb);
#endif
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

hipError_t rocprim_exclusive_scan_i32(void *tmp, size_t &bytes, const int32_t *in, int32_t *out, int n,
                                      hipStream_t st)
{
    return rocprim::exclusive_scan(tmp, bytes, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), st);
}

int color_graph_device(const int32_t *d_off, const int32_t *d_adj, int n, int32_t *d_colors, int *ncolors,
                       hipStream_t st)
{
    Scratch scr;
    constexpr int kBatch = 32;
    int32_t *changed, *overflow;
    VBD_TRY(scr.get(&changed, kBatch));
    VBD_TRY(scr.get(&overflow, 1));
    k_fill_i32<<<nblk(n), kTB, 0, st>>>(d_colors, n, 0);
    k_fill_i32<<<1, 1, 0, st>>>(overflow, 1, 0);
    int32_t hc[kBatch];
    bool done = false;
    for (int r = 0; r < n + 2 * kBatch && !done; r += kBatch) {
        k_fill_i32<<<1, kBatch, 0, st>>>(changed, kBatch, 0);
        for (int b = 0; b < kBatch; ++b)
            k_color_round_graph<<<nblk(n), kTB, 0, st>>>(d_off, d_adj, n, d_colors, changed + b, overflow);
        VBD_HIP(hipMemcpyAsync(hc, changed, sizeof(hc), hipMemcpyDeviceToHost, st));
        VBD_HIP(hipStreamSynchronize(st));
        for (int b = 0; b < kBatch; ++b) done |= (hc[b] == 0);
    }
    // the overflow flag of the rounds above may come from a transient colour;
    // one more round at the fixed point decides it
    k_fill_i32<<<1, 1, 0, st>>>(overflow, 1, 0);
    k_color_round_graph<<<nblk(n), kTB, 0, st>>>(d_off, d_adj, n, d_colors, changed, overflow);
    int32_t of = 0;
    VBD_HIP(hipMemcpyAsync(&of, overflow, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    std::vector<int32_t> col(n);
    VBD_HIP(hipMemcpyAsync(col.data(), d_colors, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    if (!done) {
        set_error("colouring did not reach its fixed point");
        return VBD_EHIP;
    }
    if (of) {
        set_error("a vertex needs more than 256 colours");
        return VBD_ECOLORS;
    }
    int mx = -1;
    for (int v = 0; v < n; ++v) mx = std::max(mx, col[v]);
    *ncolors = mx + 1;
    return 0;
}

int color_tets_device(const int32_t *d_tets, int m, int n, int32_t *d_colors, int32_t *d_inc_off, int32_t *d_inc,
                      int *ncolors, int *max_degree, hipStream_t stream)
{
    Scratch scr;
    int maxinc = 0, nref = 0;
    return build_incidence_and_colors(d_tets, m, n, d_inc_off, d_inc, d_colors, ncolors, &maxinc, max_degree, &nref,
                                      false, stream, scr);
}

int setup_mesh(vbd_solver *s, const double *rest_xyz, const int32_t *tets, const double *density,
               const double *youngs, const double *poisson, const uint8_t *active, const int32_t *body_voff,
               int n_bodies, const double *body_grav)
{
    const int n = s->n, m = s->m;
    const int nb = std::max(1, n_bodies);
    s->n_bodies = nb;
    s->body_voff_host.assign(nb + 1, 0);
    for (int b = 0; b <= nb; ++b) s->body_voff_host[b] = body_voff ? body_voff[b] : (b == 0 ? 0 : n);
    hipStream_t st = s->stream;
    Scratch scr;
    int32_t *d_tets;
    double *d_rho, *d_E, *d_nu, *d_invDm, *d_vol, *d_mass, *d_mu, *d_lam;
    uint8_t *d_act_orig;
    VBD_TRY(dmalloc(&s->X, 3 * (size_t)n));
    VBD_TRY(dmalloc(&s->tets_dev, 4 * (size_t)m));
    d_tets = s->tets_dev;
    VBD_TRY(scr.get(&d_rho, m));
    VBD_TRY(scr.get(&d_E, m));
    VBD_TRY(scr.get(&d_nu, m));
    VBD_TRY(scr.get(&d_invDm, 9 * (size_t)m));
    VBD_TRY(scr.get(&d_vol, m));
    VBD_TRY(scr.get(&d_mass, m));
    VBD_TRY(scr.get(&d_mu, m));
    VBD_TRY(scr.get(&d_lam, m));
    VBD_TRY(scr.get(&d_act_orig, n));
    VBD_HIP(hipMemcpyAsync(s->X, rest_xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
    if (m) {
        VBD_HIP(hipMemcpyAsync(d_tets, tets, sizeof(int32_t) * 4 * m, hipMemcpyHostToDevice, st));
        VBD_HIP(hipMemcpyAsync(d_rho, density, sizeof(double) * m, hipMemcpyHostToDevice, st));
        VBD_HIP(hipMemcpyAsync(d_E, youngs, sizeof(double) * m, hipMemcpyHostToDevice, st));
        VBD_HIP(hipMemcpyAsync(d_nu, poisson, sizeof(double) * m, hipMemcpyHostToDevice, st));
    }
    std::vector<uint8_t> act(n, 1);
    if (active)
        for (int i = 0; i < n; ++i) act[i] = active[i] ? 1 : 0;
    VBD_HIP(hipMemcpyAsync(d_act_orig, act.data(), n, hipMemcpyHostToDevice, st));

    // 1. rest shape + Lame
    int32_t *d_singular;
    VBD_TRY(scr.get(&d_singular, 1));
    k_fill_i32<<<1, 1, 0, st>>>(d_singular, 1, 0);
    if (m)
        k_tet_setup<<<nblk(m), kTB, 0, st>>>(s->X, d_tets, m, d_rho, d_E, d_nu, d_invDm, d_vol, d_mass, d_mu, d_lam,
                                             d_singular);
    int32_t singular = 0;
    VBD_HIP(hipMemcpyAsync(&singular, d_singular, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    if (m) {
        std::vector<double> vol(m);
        VBD_HIP(hipMemcpy(vol.data(), d_vol, sizeof(double) * m, hipMemcpyDeviceToHost));
        double sv = 0.0;
        for (double x : vol) sv += x;
        s->edge_length = std::cbrt(6.0 * sv / m);
    }
    if (singular) {
        set_error("an element has a singular rest shape matrix");
        return VBD_EDEGENERATE;
    }

    // 2-3. incidence + colouring
    VBD_TRY(dmalloc(&s->inc_off, (size_t)n + 1));
    VBD_TRY(dmalloc(&s->inc, 4 * (size_t)m));
    VBD_TRY(dmalloc(&s->colors, n));
    VBD_TRY(build_incidence_and_colors(d_tets, m, n, s->inc_off, s->inc, s->colors, &s->ncolors, &s->max_incident,
                                       &s->max_degree, &s->n_ref, true, st, scr));

    // 4. colour-major renumbering (stable radix sort on colour, then body, then
    // for a single body the lane class)
    std::vector<int32_t> body_of(n, 0);
    if (body_voff)
        for (int b = 0; b < nb; ++b)
            for (int v = body_voff[b]; v < body_voff[b + 1]; ++v) body_of[v] = b;
    int32_t *d_body_of;
    VBD_TRY(scr.get(&d_body_of, n));
    VBD_HIP(hipMemcpyAsync(d_body_of, body_of.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    int ncls = 1;
    if (!s->batched) {
        // lane classes only for interior-dominated big colours (kLaneClassMinShare)
        std::vector<int32_t> hc(n), ho(n + 1);
        VBD_HIP(hipMemcpy(hc.data(), s->colors, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        VBD_HIP(hipMemcpy(ho.data(), s->inc_off, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
        std::vector<int64_t> nvc(std::max(s->ncolors, 1), 0), n4c(std::max(s->ncolors, 1), 0);
        for (int v = 0; v < n; ++v)
            if (act[v] && hc[v] >= 0 && hc[v] < s->ncolors) {
                ++nvc[hc[v]];
                n4c[hc[v]] += lane_class(ho[v + 1] - ho[v]) == 0;
            }
        int64_t big = 0, big4 = 0;
        for (int c = 0; c < s->ncolors; ++c)
            if (nvc[c] >= kStageMinColor) {
                big += nvc[c];
                big4 += n4c[c];
            }
        if (big > 0 && (double)big4 >= kLaneClassMinShare * (double)big) ncls = 3;
    }
    const int64_t nkeys = (int64_t)(kMaxColors + 1) * nb * ncls;
    unsigned end_bit = 1;
    while (((int64_t)1 << end_bit) < nkeys) ++end_bit;
    uint32_t *keys, *keys_sorted;
    int32_t *vals;
    VBD_TRY(scr.get(&keys, n));
    VBD_TRY(scr.get(&keys_sorted, n));
    VBD_TRY(scr.get(&vals, n));
    VBD_TRY(dmalloc(&s->perm, n));
    VBD_TRY(dmalloc(&s->iperm, n));
    VBD_TRY(dmalloc(&s->body_new, n));
    k_sort_keys<<<nblk(n), kTB, 0, st>>>(s->colors, d_act_orig, d_body_of, nb, n, s->inc_off, ncls, keys, vals);
    {
        size_t tb = 0;
        VBD_HIP(rocprim::radix_sort_pairs(nullptr, tb, keys, keys_sorted, vals, s->perm, (size_t)n, 0, end_bit, st));
        void *tmp;
        VBD_TRY(scr.get((char **)&tmp, tb));
        VBD_HIP(rocprim::radix_sort_pairs(tmp, tb, keys, keys_sorted, vals, s->perm, (size_t)n, 0, end_bit, st));
    }
    k_inverse_perm<<<nblk(n), kTB, 0, st>>>(s->perm, n, s->iperm);
    k_gather_i32<<<nblk(n), kTB, 0, st>>>(d_body_of, s->perm, n, s->body_new);
    int32_t *kfirst, *kend;
    VBD_TRY(scr.get(&kfirst, nkeys));
    VBD_TRY(scr.get(&kend, nkeys));
    k_fill_i32<<<nblk(nkeys), kTB, 0, st>>>(kfirst, nkeys, 0);
    k_fill_i32<<<nblk(nkeys), kTB, 0, st>>>(kend, nkeys, 0);
    k_key_runs<<<nblk(n), kTB, 0, st>>>(keys_sorted, n, kfirst, kend);
    std::vector<int32_t> hf(nkeys), he(nkeys), hh(nkeys);
    VBD_HIP(hipMemcpyAsync(hf.data(), kfirst, sizeof(int32_t) * nkeys, hipMemcpyDeviceToHost, st));
    VBD_HIP(hipMemcpyAsync(he.data(), kend, sizeof(int32_t) * nkeys, hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    // per (colour, body): the vertex count over the lane classes
    std::vector<int32_t> hcb((size_t)(kMaxColors + 1) * nb, 0);
    for (int64_t k = 0; k < nkeys; ++k) {
        hh[k] = he[k] - hf[k];
        hcb[k / ncls] += hh[k];
    }
    s->col_start.assign(s->ncolors + 1, 0);
    s->col_size.assign(s->ncolors, 0);
    s->col_n42.assign(s->ncolors, make_int2(0, 0));
    for (int c = 0; c < s->ncolors; ++c) {
        int tot = 0;
        for (int b = 0; b < nb; ++b) tot += hcb[(size_t)c * nb + b];
        s->col_size[c] = tot;
        s->col_start[c + 1] = s->col_start[c] + tot;
        s->col_n42[c] = ncls == 3 ? make_int2(hh[3 * (size_t)c], hh[3 * (size_t)c + 1]) : make_int2(tot, 0);
    }
    s->n_swept = s->col_start[s->ncolors];
    // per body: (start, count) of each colour, then of its other vertices
    s->ranges_host.assign((size_t)nb * (s->ncolors + 1), make_int2(0, 0));
    {
        int64_t pos = 0;
        for (int c = 0; c <= kMaxColors; ++c) {
            const int cc = c < s->ncolors ? c : (c == kMaxColors ? s->ncolors : -1);
            for (int b = 0; b < nb; ++b) {
                const int cnt = hcb[(size_t)c * nb + b];
                if (cc >= 0) s->ranges_host[(size_t)b * (s->ncolors + 1) + cc] = make_int2((int)pos, cnt);
                pos += cnt;
            }
        }
    }
    VBD_TRY(dmalloc(&s->d_col_start, (size_t)s->ncolors + 1));
    VBD_HIP(hipMemcpyAsync(s->d_col_start, s->col_start.data(), sizeof(int32_t) * (s->ncolors + 1),
                           hipMemcpyHostToDevice, st));

    // 5. slot records
    int32_t *kcnt;
    VBD_TRY(scr.get(&kcnt, (size_t)s->n_swept + 1));
    VBD_TRY(dmalloc(&s->slot_off, (size_t)s->n_swept + 1));
    k_slot_counts<<<nblk(s->n_swept + 1), kTB, 0, st>>>(s->perm, s->inc_off, s->n_swept, kcnt);
    VBD_TRY(exclusive_scan_i32(kcnt, s->slot_off, s->n_swept + 1, st, scr));
    int32_t ns = 0;
    VBD_HIP(hipMemcpyAsync(&ns, s->slot_off + s->n_swept, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    s->n_slots = ns;
    if (ns >= (1 << 26) || n >= (1 << 26)) {
        // the sweep addresses records, dictionary entries and vertices with
        // 32-bit byte offsets
        set_error("mesh too large: at most 2^26 - 1 vertices and vertex-tet incidences (16.7 M tets)");
        return VBD_EMESH;
    }
    VBD_TRY(dmalloc(&s->mass, n));
    VBD_TRY(dmalloc(&s->active, n));
    k_vertex_mass<<<nblk(n), kTB, 0, st>>>(s->perm, s->inc_off, s->inc, d_mass, n, s->mass, d_act_orig, s->active);
#if VBD_TET_RECORDS
    VBD_TRY(dmalloc(&s->slots, 4 * (size_t)std::max(m, 1)));
    VBD_TRY(dmalloc(&s->slot_ent, (size_t)std::max<int64_t>(ns, 1)));
    if (m)
        k_build_tet_records<<<nblk(m), kTB, 0, st>>>(m, d_tets, s->iperm, d_invDm, d_vol, d_mu, d_lam, s->material,
                                                     s->slots);
    if (s->n_swept)
        k_build_slot_entries<<<nblk(s->n_swept), kTB, 0, st>>>(s->perm, s->inc_off, s->inc, s->n_swept, s->slot_off,
                                                              s->slot_ent);
#else
    VBD_TRY(dmalloc(&s->slots, 3 * (size_t)std::max<int64_t>(ns, 1)));
    VBD_TRY(dmalloc(&s->slots2, (size_t)std::max<int64_t>(ns, 1)));
    if (s->n_swept)
        k_build_slots<<<nblk(s->n_swept), kTB, 0, st>>>(s->perm, s->iperm, s->inc_off, s->inc, d_tets, d_invDm, d_vol,
                                                       d_mu, d_lam, s->material, s->n_swept, s->slot_off, s->slots,
                                                       s->slots2);
#endif

    // algorithmic bytes per colour sweep: sum_v (40 + 68 k_v) + 12 |boundary|
    {
        int32_t *mark, *count;
        VBD_TRY(scr.get(&mark, n));
        VBD_TRY(scr.get(&count, s->ncolors > 0 ? s->ncolors : 1));
        k_fill_i32<<<nblk(n), kTB, 0, st>>>(mark, n, -1);
        k_fill_i32<<<1, kTB, 0, st>>>(count, std::max(1, s->ncolors), 0);
        for (int c = 0; c < s->ncolors; ++c) {
            int v0 = s->col_start[c], v1 = s->col_start[c + 1];
            if (v1 > v0)
                k_mark_boundary<<<nblk(v1 - v0), kTB, 0, st>>>(s->slots, s->slot_ent, s->slot_off, v0, v1, c, mark,
                                                               count + c);
        }
        std::vector<int32_t> hcnt(std::max(1, s->ncolors));
        std::vector<int32_t> hoff(s->n_swept + 1);
        VBD_HIP(hipMemcpyAsync(hcnt.data(), count, sizeof(int32_t) * hcnt.size(), hipMemcpyDeviceToHost, st));
        VBD_HIP(hipMemcpyAsync(hoff.data(), s->slot_off, sizeof(int32_t) * hoff.size(), hipMemcpyDeviceToHost, st));
        VBD_HIP(hipStreamSynchronize(st));
        s->col_bytes.assign(s->ncolors, 0.0);
        for (int c = 0; c < s->ncolors; ++c) {
            int v0 = s->col_start[c], v1 = s->col_start[c + 1];
            double k = (double)(hoff[v1] - hoff[v0]);
            s->col_bytes[c] = 40.0 * (v1 - v0) + 68.0 * k + 12.0 * hcnt[c];
            s->col_slots.push_back(k);
            s->col_boundary.push_back((double)hcnt[c]);
        }
    }

#if !VBD_TET_RECORDS
    VBD_TRY(build_record_dictionary(s, g_record_dictionary, st, scr));
#endif

    // state buffers
    VBD_TRY(dmalloc(&s->u, n));
    VBD_TRY(dmalloc(&s->uprev, n));
    VBD_TRY(dmalloc(&s->vel, n));
    VBD_TRY(dmalloc(&s->tgt, n));
    VBD_TRY(dmalloc(&s->xio, 3 * (size_t)n));
    VBD_TRY(dmalloc(&s->ctrl, 1));
    VBD_HIP(hipHostMalloc((void **)&s->ctrl_host, sizeof(Ctrl), hipHostMallocDefault));
    VBD_HIP(hipHostMalloc((void **)&s->xio_host, sizeof(double) * 3 * std::max(n, 1), hipHostMallocDefault));
    VBD_HIP(hipMemsetAsync(s->u, 0, sizeof(float4) * n, st));
    VBD_HIP(hipMemsetAsync(s->uprev, 0, sizeof(double4) * n, st));
    VBD_HIP(hipMemsetAsync(s->vel, 0, sizeof(float4) * n, st));
    VBD_HIP(hipMemsetAsync(s->tgt, 0, sizeof(double4) * n, st));
    // per-body range table, gravity and step results
    {
        VBD_TRY(dmalloc(&s->d_ranges, s->ranges_host.size()));
        VBD_HIP(hipMemcpyAsync(s->d_ranges, s->ranges_host.data(), sizeof(int2) * s->ranges_host.size(),
                               hipMemcpyHostToDevice, st));
        std::vector<float4> g(nb);
        for (int b = 0; b < nb; ++b) {
            const double *gb = body_grav ? body_grav + 3 * b : s->grav;
            g[b] = make_float4((float)gb[0], (float)gb[1], (float)gb[2], 0.f);
        }
        VBD_TRY(dmalloc(&s->d_grav, nb));
        VBD_HIP(hipMemcpyAsync(s->d_grav, g.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, st));
        VBD_TRY(dmalloc(&s->d_body_out, nb));
        VBD_HIP(hipHostMalloc((void **)&s->body_out_host, sizeof(int2) * nb, hipHostMallocDefault));
    }
    VBD_HIP(hipStreamSynchronize(st));
    plan_segments(s);
    VBD_TRY(build_small_records(s));
    return 0;
}

// Lane layouts and neighbour staging tables (the st_* fields).  Colours of at
// least kStageMinColor vertices get their lane layout (lane_layout: a single
// body's vertices take 4, 2 or 1 lanes by slot count, a batch's 4) and, with
// a shared dictionary, staging lists: each 256-thread block below the
// layout's b1 lists, vertex by vertex, each vertex's distinct neighbours in
// order of first appearance in its slots (entries id | owner << 25, owner the
// vertex's index in the block; kStageMax entries per block of 4-lane
// vertices, kStageMaxWide per block holding 2-lane ones), and each of a
// vertex's slots gets the record n0 | n1 << 5 | n2 << 10 | entry << 15 with
// n_k the index of the slot's k-th other vertex in the vertex's own list.
// Vertices with the same incidence pattern (a voxel mesh's interior vertices,
// each face's, edge's ...) have the same record sequence, so st_rec holds
// each distinct sequence once and each vertex gets one 8-byte word (st_vrec):
// its first entry | its slot count << 16, and its sequence's first record.  A batch whose dictionary holds more
// than kStageDict entries (bodies of many materials) also lists per block the
// dictionary entries its slots use, in order of first use (st_dl, kStageDict
// per block), and a slot's entry field is its index in that list (a single
// body's solver stages only a shared dictionary).  A block
// with more entries than its lists hold, a vertex with more than 31
// neighbours, or a mesh of 2^25 - 1 vertices or more turns staging off for
// the solver (the layouts stay, and every block reads its neighbours from
// global memory).
int build_staging(vbd_solver *s)
{
    s->st_ok = false;
    s->st_col.assign(s->ncolors, LaneLayout());
    s->st_ids.assign(s->ncolors, 0);
    const int min_color = std::max(s->launch.small_color, kStageMinColor);
    for (int c = 0; c < s->ncolors; ++c) {
        const int nv = s->col_size[c];
        if (nv >= min_color) s->st_col[c] = lane_layout(s->col_n42[c].x, s->col_n42[c].y, nv, false, -1);
    }
    // (ids below 2^25 - 1, so that no entry reads as the -1 that ends a list)
    const bool per_block = s->n_dict > kStageDict;
    if (!g_staging || !s->dict_shared || !s->dict_ext || (per_block && !s->batched) || s->n_slots <= 0 ||
        s->n >= (1 << 25) - 1)
        return 0;
    hipStream_t st = s->stream;
    const int nsw = s->n_swept;
    std::vector<int32_t> off(nsw + 1);
    std::vector<float4> rec(s->n_slots);
    VBD_HIP(hipMemcpyAsync(off.data(), s->slot_off, sizeof(int32_t) * (nsw + 1), hipMemcpyDeviceToHost, st));
    VBD_HIP(hipMemcpyAsync(rec.data(), s->slots, sizeof(float4) * s->n_slots, hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    constexpr uint32_t kNbMask = (1u << kStageEntryBits) - 1u;
    std::vector<int32_t> nb;  // the blocks' lists, -1 past each block's entries
    std::vector<uint32_t> srec;  // the distinct record sequences
    std::unordered_map<std::string, std::pair<uint32_t, int>> seqs;  // sequence -> (first record, last colour)
    std::vector<int2> vrec(std::max(nsw, 1), make_int2(0, 0));
    std::vector<double> prec(s->ncolors, 0.0);
    std::vector<uint32_t> seq;
    std::vector<int32_t> stamp(s->n, -1), local(s->n, 0);
    std::vector<int32_t> dl;  // per-block dictionary lists
    std::vector<int32_t> estamp(per_block ? s->n_dict : 0, -1), elocal(per_block ? s->n_dict : 0, 0);
    std::vector<LaneLayout> col(s->ncolors);
    std::vector<int64_t> ids(s->ncolors, 0);
    int nblocks = 0;
    for (int c = 0; c < s->ncolors; ++c) {
        const int nv = s->col_size[c];
        if (nv < min_color) continue;
        const int n4 = s->col_n42[c].x, n2 = s->col_n42[c].y;
        const LaneLayout l = lane_layout(n4, n2, nv, true, (int64_t)nb.size());
        // a batch's per-block lists are found at lay.ids / kStageMax
        if (per_block && (l.b4 != l.nblocks || l.ids != (int64_t)nblocks * kStageMax)) return 0;
        const int64_t lanes = 4 * (int64_t)n4 + 2 * (int64_t)n2 + l.n1;
        const int c0 = s->col_start[c];
        nb.resize(nb.size() + stage_list_offset(l.b1, l.b4), -1);
        for (int b = 0; b < l.b1; ++b, ++nblocks) {
            int g, lg;
            const int va = c0 + lane_vertex(256 * (int64_t)b, n4, n2, g, lg);
            const int vb = c0 + lane_vertex(std::min<int64_t>(256 * (int64_t)b + 256, lanes) - 1, n4, n2, g, lg) + 1;
            if (vb - va > 128) return 0;  // the owner field holds 7 bits
            const size_t base = (size_t)(l.ids + stage_list_offset(b, l.b4));
            const int cap = b < l.b4 ? kStageMax : kStageMaxWide;
            int cnt = 0, ne = 0;
            if (per_block) dl.resize((size_t)(nblocks + 1) * kStageDict, -1);
            for (int v = va; v < vb; ++v) {
                if (cnt > 0xFFFF || off[v + 1] - off[v] > kMaxElementsPerVertex) return 0;
                const int first = cnt;
                seq.clear();
                int nd = 0;  // the vertex's distinct neighbours so far
                for (int64_t q = off[v]; q < off[v + 1]; ++q) {
                    uint32_t w[4];
                    std::memcpy(w, &rec[q], sizeof(w));
                    uint32_t lk[3];
                    for (int k = 0; k < 3; ++k) {
                        const int id = (int)w[k];
                        if (stamp[id] != v) {
                            if (cnt == cap || nd == (int)kNbMask + 1) return 0;
                            stamp[id] = v;
                            local[id] = nd++;
                            nb[base + cnt++] = (int32_t)((uint32_t)id | (uint32_t)(v - va) << 25);
                        }
                        lk[k] = (uint32_t)local[id];
                    }
                    uint32_t e = w[3];
                    if (per_block) {
                        if (e >= (uint32_t)s->n_dict) return 0;
                        if (estamp[e] != nblocks) {
                            if (ne == kStageDict) return 0;
                            estamp[e] = nblocks;
                            elocal[e] = ne;
                            dl[(size_t)nblocks * kStageDict + ne++] = (int32_t)e;
                        }
                        e = (uint32_t)elocal[e];
                    }
                    if (e >= (1u << 17)) return 0;
                    seq.push_back(lk[0] | lk[1] << 5 | lk[2] << 10 | e << 15);
                }
                const std::string key(reinterpret_cast<const char *>(seq.data()), seq.size() * sizeof(uint32_t));
                auto it = seqs.find(key);
                if (it == seqs.end()) {
                    if (srec.size() + seq.size() >= (size_t)INT32_MAX) return 0;
                    it = seqs.emplace(key, std::make_pair((uint32_t)srec.size(), -1)).first;
                    srec.insert(srec.end(), seq.begin(), seq.end());
                }
                if (it->second.second != c) {  // records the colour reads, once per sequence
                    it->second.second = c;
                    prec[c] += (double)seq.size();
                }
                vrec[v] = make_int2(first | (int)seq.size() << 16, (int)it->second.first);
            }
            ids[c] += cnt;
        }
        col[c] = l;
    }
    if (nblocks == 0) return 0;
    // zero records past the last slot: the sweep loads a lane's records
    // without a bound (the ones past its vertex's last slot are not used)
    srec.resize(srec.size() + kMaxElementsPerVertex, 0u);
    VBD_TRY(dmalloc(&s->st_nb, nb.size()));
    VBD_TRY(dmalloc(&s->st_rec, srec.size()));
    VBD_TRY(dmalloc(&s->st_vrec, vrec.size()));
    VBD_HIP(hipMemcpyAsync(s->st_nb, nb.data(), sizeof(int32_t) * nb.size(), hipMemcpyHostToDevice, st));
    VBD_HIP(hipMemcpyAsync(s->st_rec, srec.data(), sizeof(uint32_t) * srec.size(), hipMemcpyHostToDevice, st));
    VBD_HIP(hipMemcpyAsync(s->st_vrec, vrec.data(), sizeof(int2) * vrec.size(), hipMemcpyHostToDevice, st));
    if (per_block) {
        VBD_TRY(dmalloc(&s->st_dl, dl.size()));
        VBD_HIP(hipMemcpyAsync(s->st_dl, dl.data(), sizeof(int32_t) * dl.size(), hipMemcpyHostToDevice, st));
    }
    VBD_HIP(hipStreamSynchronize(st));
    for (int c = 0; c < s->ncolors; ++c)
        if (col[c].nblocks > 0) s->st_col[c] = col[c];
    s->st_ids = ids;
    s->st_prec = prec;
    s->st_ok = true;
    return 0;
}


// ------------------------------------------------------------------ colouring without allocation
//
// color_tets_noalloc: color_tets_device on device buffers that the caller
// owns, with every intermediate array (incidence CSR, counters, lower-neighbour
// CSR, scan temporaries, round flags) carved out of one file-scope device
// array.  Each GPU holds its own copy of the array for the life of the
// process; its address on a device comes from hipGetSymbolAddress, taken
// once per device with that device current.  All calls on one GPU share the
// array, so calls that overlap on one GPU race.
namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// incidence counts with a range check (the tets are the caller's device data)
__global__ void k_count_incidence_checked(const int32_t *tets, int m, int n, int32_t *cnt, int32_t *scal)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    int mx = -1;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int v = tets[4 * e + k];
        if (v < 0 || v >= n) {
            bad = true;
            continue;
        }
        atomicAdd(&cnt[v], 1);
        mx = max(mx, v);
    }
    atomicMax(&scal[0], mx);
    if (bad) atomicOr(&scal[4], 1);
}

// largest colour (scal[5]); colours start at -1 for unreferenced vertices
__global__ void k_max_color(const int32_t *colors, int n, int32_t *scal)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int c = v < n ? colors[v] : -1;
    c = max(c, __shfl_xor(c, 1));
    c = max(c, __shfl_xor(c, 2));
    c = max(c, __shfl_xor(c, 4));
    c = max(c, __shfl_xor(c, 8));
    c = max(c, __shfl_xor(c, 16));
    c = max(c, __shfl_xor(c, 32));
    if ((threadIdx.x & 63) == 0) atomicMax(&scal[5], c);
}

// Bump allocator over the scratch array: 256-byte aligned pieces.
struct Carve {
    char *base;
    size_t used, cap;
    template <typename T>
    T *take(size_t count)
    {
        const size_t off = (used + 255) & ~(size_t)255;
        const size_t bytes = count * sizeof(T);
        if (off > cap || bytes > cap - off) return nullptr;
        used = off + bytes;
        return reinterpret_cast<T *>(base + off);
    }
};

}  // namespace

int scratch_address(int device, void **out)
{
    if (device < 0 || device >= kMaxDevices) {
        set_error_literal("device index out of range for the allocation-free entry points");
        return VBD_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            set_error_literal("hipGetSymbolAddress(scratch) failed");
            return VBD_EHIP;
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

#define VBD_NA(call)                                                                           \
    do {                                                                                       \
        if ((call) != hipSuccess) {                                                            \
            set_error_literal("a HIP call of vbd_color_noalloc failed: " #call);               \
            return VBD_EHIP;                                                                   \
        }                                                                                      \
    } while (0)

int color_tets_noalloc(const int32_t *d_tets, int m, int n, int32_t *d_colors, int *ncolors, int *max_degree,
                       hipStream_t st)
{
    int dev = 0;
    VBD_NA(hipGetDevice(&dev));
    void *base = nullptr;
    VBD_TRY(scratch_address(dev, &base));
    Carve cv{static_cast<char *>(base), 0, kNoAllocScratchBytes};
    // the fixed part: 5 (n + 1) + 4 m counters and lists, 64 scalars and flags,
    // the scan's temporary storage; the lower-neighbour list comes last
    if ((int64_t)20 * ((int64_t)n + 1) + (int64_t)16 * m > (int64_t)kNoAllocScratchBytes) {
        set_error_literal("mesh too large for vbd_color_noalloc's scratch (20 bytes per vertex + 16 per tet + the "
                          "lower-neighbour lists must fit in 64 MiB)");
        return VBD_EINVAL;
    }
    size_t tb = 0;
    VBD_NA(rocprim::exclusive_scan(nullptr, tb, (const int32_t *)nullptr, (int32_t *)nullptr, 0, (size_t)n + 1,
                                   rocprim::plus<int32_t>(), st));
    int32_t *inc_off = cv.take<int32_t>((size_t)n + 1), *inc = cv.take<int32_t>(std::max<size_t>(1, 4 * (size_t)m));
    int32_t *cnt = cv.take<int32_t>((size_t)n + 1), *cursor = cv.take<int32_t>((size_t)n + 1);
    int32_t *lo_cnt = cv.take<int32_t>((size_t)n + 1), *lo_off = cv.take<int32_t>((size_t)n + 1);
    int32_t *scal = cv.take<int32_t>(8), *changed = cv.take<int32_t>(32);
    void *tmp = cv.take<char>(std::max<size_t>(tb, 1));
    if (!inc_off || !inc || !cnt || !cursor || !lo_cnt || !lo_off || !scal || !changed || !tmp) {
        set_error_literal("mesh too large for vbd_color_noalloc's scratch (64 MiB)");
        return VBD_EINVAL;
    }
    // scal: 0 largest referenced vertex, 1 largest incidence, 2 largest degree,
    // 3 colour overflow, 4 index out of range, 5 largest colour
    k_fill_i32<<<nblk(n + 1), kTB, 0, st>>>(cnt, n + 1, 0);
    k_fill_i32<<<1, kTB, 0, st>>>(scal, 8, 0);
    k_fill_i32<<<1, 1, 0, st>>>(scal, 1, -1);
    k_fill_i32<<<1, 1, 0, st>>>(scal + 5, 1, -1);
    if (m > 0) k_count_incidence_checked<<<nblk(m), kTB, 0, st>>>(d_tets, m, n, cnt, scal);
    int32_t h[8];
    VBD_NA(hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, st));
    VBD_NA(hipStreamSynchronize(st));
    if (h[4]) {
        set_error_literal("element index outside [0, n_vertices)");
        return VBD_EMESH;
    }
    const int nref = h[0] + 1;
    VBD_NA(rocprim::exclusive_scan(tmp, tb, cnt, inc_off, 0, (size_t)n + 1, rocprim::plus<int32_t>(), st));
    VBD_NA(hipMemcpyAsync(cursor, inc_off, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToDevice, st));
    if (m > 0) k_fill_incidence<<<nblk(m), kTB, 0, st>>>(d_tets, m, cursor, inc);
    k_sort_incidence<<<nblk(n), kTB, 0, st>>>(inc_off, inc, n, scal + 1);
    k_degree<<<nblk(n), kTB, 0, st>>>(d_tets, inc_off, inc, n, scal + 2);
    k_fill_i32<<<nblk(n + 1), kTB, 0, st>>>(lo_cnt, n + 1, 0);
    k_lower_neighbours<false><<<nblk(n), kTB, 0, st>>>(d_tets, inc_off, inc, n, lo_cnt, nullptr, nullptr);
    VBD_NA(rocprim::exclusive_scan(tmp, tb, lo_cnt, lo_off, 0, (size_t)n + 1, rocprim::plus<int32_t>(), st));
    int32_t n_lo = 0;
    VBD_NA(hipMemcpyAsync(&n_lo, lo_off + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VBD_NA(hipStreamSynchronize(st));
    int32_t *lo = cv.take<int32_t>((size_t)std::max(n_lo, 1));
    if (!lo) {
        set_error_literal("mesh too large for vbd_color_noalloc's scratch (64 MiB): the lower-neighbour lists "
                          "do not fit");
        return VBD_EINVAL;
    }
    k_lower_neighbours<true><<<nblk(n), kTB, 0, st>>>(d_tets, inc_off, inc, n, nullptr, lo_off, lo);
    k_color_init<<<nblk(n), kTB, 0, st>>>(d_colors, n, nref);
    constexpr int kBatch = 32;
    int32_t hc[kBatch];
    const int max_rounds = std::max(64, 4 * nref + 8);
    bool done = false;
    for (int r = 0; r < max_rounds && !done; r += kBatch) {
        k_fill_i32<<<1, kBatch, 0, st>>>(changed, kBatch, 0);
        for (int b = 0; b < kBatch; ++b)
            k_color_round_graph<<<nblk(std::max(nref, 1)), kTB, 0, st>>>(lo_off, lo, nref, d_colors, changed + b,
                                                                          scal + 3);
        VBD_NA(hipMemcpyAsync(hc, changed, sizeof(hc), hipMemcpyDeviceToHost, st));
        VBD_NA(hipStreamSynchronize(st));
        for (int b = 0; b < kBatch; ++b) done |= (hc[b] == 0);
    }
    // the overflow flag of the rounds above may come from a transient colour;
    // one more round at the fixed point decides it
    k_fill_i32<<<1, 1, 0, st>>>(scal + 3, 1, 0);
    k_color_round_graph<<<nblk(std::max(nref, 1)), kTB, 0, st>>>(lo_off, lo, nref, d_colors, changed, scal + 3);
    k_max_color<<<nblk(n), kTB, 0, st>>>(d_colors, n, scal);
    VBD_NA(hipGetLastError());
    VBD_NA(hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, st));
    VBD_NA(hipStreamSynchronize(st));
    if (!done) {
        set_error_literal("colouring did not reach its fixed point");
        return VBD_EHIP;
    }
    if (h[3]) {
        set_error_literal("a vertex needs more than 256 colours");
        return VBD_ECOLORS;
    }
    *max_degree = h[2];
    *ncolors = h[5] + 1;
    return 0;
}

}  // namespace vbd

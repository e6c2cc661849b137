// vbd_mpnn.hip -- learned warm start for the VBD step (no counterpart in WarpVBD).
//
// A 4-layer message-passing network on the mesh's vertex graph predicts each
// step's converged correction c = x_{n+1} - x_inertia from the state at the
// start of the step.  VBD then iterates from x_inertia + c to the same
// tolerance; the plain inertia guess is kept whenever the network guess has a
// larger first-sweep residual (Jacobi evaluation of both guesses).
//
// Network (same definition as the FP64 reference in oracle/mpnn_ref.py):
//   h0      = relu(bf(f) bf(W_in)^T + b_in)                      f: 32 features
//   a_l     = mean_{j in N(i)} h_l[j]
//   h_{l+1} = h_l + relu(bf([h_l, a_l]) bf(W_l)^T + b_l)         l = 0..3, width 64
//   out     = bf(h_4) bf(W_out)^T + b_out                        3 outputs
// The width-64 products run on MFMA (v_mfma_f32_16x16x16_bf16, BF16 operands,
// FP32 accumulation); weights are kept in FP32 for Adam and copied to BF16
// after every update.  Weight gradients accumulate in FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include <hip/hip_ext.h>

#include "vbd_device.h"
#include "vbd_internal.h"
#include "vbd_mpnn.h"
#include "vbd_sweep.h"

namespace vbd {

namespace {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4v_t __attribute__((ext_vector_type(4)));

constexpr int kFin = 32, kW = 64, kLayers = 4, kFout = 3;
constexpr int kOffWin = 0;
constexpr int kOffBin = kOffWin + kW * kFin;
constexpr int kLayerStride = kW * 2 * kW + kW;
__host__ __device__ constexpr int off_W(int l) { return kOffBin + kW + l * kLayerStride; }
__host__ __device__ constexpr int off_b(int l) { return off_W(l) + kW * 2 * kW; }
constexpr int kOffWout = off_W(kLayers);
constexpr int kOffBout = kOffWout + kFout * kW;
constexpr int kNumParams = kOffBout + kFout;

__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float bfr(float x) { return bf2f(f2bf(x)); }

// ----------------------------------------------------------------- MFMA product
// Y[M, 64*gridDim.y] = X[M, K] W^T + bias with X = [X1 | X2] (row-major FP32,
// converted to BF16 on load) and W row-major BF16 [Nout, K].  A block of 256
// threads computes 64 rows x 64 columns: wave w owns rows 16w..16w+15 and four
// 16x16 MFMA tiles.  Fragment layout of v_mfma_f32_16x16x16_bf16: lane l holds
// A[row l%16][k 4(l/16)..+3], B[k 4(l/16)..+3][col l%16] and D[rows 4(l/16)..+3][col l%16].
// MODE 0: z = Y, h = relu(Y)            (input layer; writes z and h)
// MODE 1: z = Y, h = hin + relu(Y)      (message-passing layer)
// MODE 2: z = Y                         (plain product, no bias)
template <int K, int MODE>
__global__ __launch_bounds__(256) void k_mfma_linear(int M, const float *__restrict__ X1, int K1,
                                                     const float *__restrict__ X2, const uint16_t *__restrict__ Wb,
                                                     const float *__restrict__ bias, float *__restrict__ z,
                                                     const float *__restrict__ hin, float *__restrict__ h, int ldo)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = blockIdx.x * 64 + wave * 16;
    const int col0 = blockIdx.y * 64;
    const int r = row0 + (lane & 15);
    const int kq = 4 * (lane >> 4);
    f32x4v_t acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4v_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int k = k0 + kq;
        bf16x4_t a;
        if (r < M) {
            const float *src = k < K1 ? X1 + (size_t)r * K1 + k : X2 + (size_t)r * (K - K1) + (k - K1);
            const float4 v = *reinterpret_cast<const float4 *>(src);
            a = bf16x4_t{(short)f2bf(v.x), (short)f2bf(v.y), (short)f2bf(v.z), (short)f2bf(v.w)};
        } else {
            a = bf16x4_t{0, 0, 0, 0};
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int o = col0 + 16 * t + (lane & 15);
            const bf16x4_t b = *reinterpret_cast<const bf16x4_t *>(Wb + (size_t)o * K + k);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int o = col0 + 16 * t + (lane & 15);
        const float bo = MODE == 2 ? 0.f : bias[o];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int rr = row0 + 4 * (lane >> 4) + i;
            if (rr >= M) continue;
            const float y = acc[t][i] + bo;
            const size_t idx = (size_t)rr * ldo + o;
            z[idx] = y;
            if (MODE == 0) h[idx] = fmaxf(y, 0.f);
            if (MODE == 1) h[idx] = hin[idx] + fmaxf(y, 0.f);
        }
    }
}

// a[i] = mean of h over the graph neighbours of i (0 without neighbours)
__global__ void k_neighbour_mean(int n, const int32_t *nb_off, const int32_t *nb, const float *h, float *a)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * kW) return;
    const int i = (int)(t / kW), f = (int)(t % kW);
    const int b = nb_off[i], e = nb_off[i + 1];
    float s = 0.f;
    for (int q = b; q < e; ++q) s += h[(size_t)nb[q] * kW + f];
    a[t] = e > b ? s / (float)(e - b) : 0.f;
}

// dh[i] += dx[i, :64] + sum_{j in N(i)} dx[j, 64:] / deg(j)   (adjoint of the mean)
__global__ void k_backprop_mean(int n, const int32_t *nb_off, const int32_t *nb, const float *dx, float *dh)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * kW) return;
    const int i = (int)(t / kW), f = (int)(t % kW);
    float s = dh[t] + dx[(size_t)i * 2 * kW + f];
    for (int q = nb_off[i]; q < nb_off[i + 1]; ++q) {
        const int j = nb[q];
        s += dx[(size_t)j * 2 * kW + kW + f] / (float)(nb_off[j + 1] - nb_off[j]);
    }
    dh[t] = s;
}

__global__ void k_relu_grad(int64_t n, const float *dh, const float *z, float *dz)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dz[t] = z[t] > 0.f ? dh[t] : 0.f;
}

// Output layer, loss and its gradient; one thread per vertex.
//   out = bf(h4) bf(W_out)^T + b_out;  loss += sum m |out - t|^2 / denom
//   dout = 2 m (out - t) / denom;  dh4 = bf(dout) bf(W_out);  gW_out += dout bf(h4)^T
__global__ __launch_bounds__(256) void k_output(int n, const float *h4, const uint16_t *Wb, const float *P,
                                                const float *target, const uint8_t *mask, float inv_denom,
                                                int with_grad, float *out, float *dh4, float *G, float *loss)
{
    __shared__ float s_g[kFout * kW + kFout];
    __shared__ float s_loss;
    for (int q = threadIdx.x; q < kFout * kW + kFout; q += blockDim.x) s_g[q] = 0.f;
    if (threadIdx.x == 0) s_loss = 0.f;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float *hi = h4 + (size_t)i * kW;
        float o[kFout];
        for (int c = 0; c < kFout; ++c) {
            float s = P[kOffBout + c];
            for (int k = 0; k < kW; ++k) s += bfr(hi[k]) * bf2f(Wb[kOffWout + c * kW + k]);
            o[c] = s;
            out[(size_t)i * kFout + c] = s;
        }
        if (with_grad) {
            const float m = mask[i] ? 1.f : 0.f;
            float d[kFout], l = 0.f;
            for (int c = 0; c < kFout; ++c) {
                const float df = m * (o[c] - target[(size_t)i * kFout + c]);
                l += df * df;
                d[c] = 2.f * df * inv_denom;
            }
            atomicAdd(&s_loss, l * inv_denom);
            for (int k = 0; k < kW; ++k) {
                float s = 0.f;
                for (int c = 0; c < kFout; ++c) s += bfr(d[c]) * bf2f(Wb[kOffWout + c * kW + k]);
                dh4[(size_t)i * kW + k] = s;
            }
            for (int c = 0; c < kFout; ++c) {
                if (d[c] == 0.f) continue;
                for (int k = 0; k < kW; ++k) atomicAdd(&s_g[c * kW + k], d[c] * bfr(hi[k]));
                atomicAdd(&s_g[kFout * kW + c], d[c]);
            }
        }
    }
    __syncthreads();
    if (!with_grad) return;
    for (int q = threadIdx.x; q < kFout * kW + kFout; q += blockDim.x)
        if (s_g[q] != 0.f) atomicAdd(&G[kOffWout + q], s_g[q]);
    if (threadIdx.x == 0) atomicAdd(loss, s_loss);
}

// Weight gradient of one linear layer: gW[o][k] += sum_i dz[i][o] bf(x[i][k]),
// gb[o] += sum_i dz[i][o], x = [X1 | X2] with K columns.  Each block folds a
// strided set of 32-row chunks (staged in LDS) into private sums and writes
// one partial row; k_sum_partials adds the partials into G.
template <int K>
__global__ __launch_bounds__(256) void k_weight_grad(int n, const float *dz, const float *X1, int K1, const float *X2,
                                                     float *partial)
{
    constexpr int RB = 32;
    constexpr int PER = kW * K / 256;  // outputs per thread
    __shared__ float s_x[RB][K + 1];
    __shared__ float s_dz[RB][kW];
    const int o = threadIdx.x / (256 / kW);
    const int kb = (threadIdx.x % (256 / kW)) * PER;
    float acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = 0.f;
    float accb = 0.f;
    for (int r0 = blockIdx.x * RB; r0 < n; r0 += gridDim.x * RB) {
        for (int q = threadIdx.x; q < RB * K; q += 256) {
            const int rr = q / K, k = q % K, i = r0 + rr;
            float v = 0.f;
            if (i < n) v = k < K1 ? X1[(size_t)i * K1 + k] : X2[(size_t)i * (K - K1) + (k - K1)];
            s_x[rr][k] = bfr(v);
        }
        for (int q = threadIdx.x; q < RB * kW; q += 256) {
            const int rr = q / kW, oo = q % kW, i = r0 + rr;
            s_dz[rr][oo] = i < n ? dz[(size_t)i * kW + oo] : 0.f;
        }
        __syncthreads();
        for (int rr = 0; rr < RB; ++rr) {
            const float d = s_dz[rr][o];
#pragma unroll
            for (int q = 0; q < PER; ++q) acc[q] += d * s_x[rr][kb + q];
            accb += d;
        }
        __syncthreads();
    }
    float *p = partial + (size_t)blockIdx.x * (kW * K + kW);
#pragma unroll
    for (int q = 0; q < PER; ++q) p[o * K + kb + q] = acc[q];
    if (kb == 0) p[kW * K + o] = accb;
}

__global__ void k_sum_partials(int nparts, int len, const float *partial, float *G, int goff_w, int goff_b, int nw)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len) return;
    float s = 0.f;
    for (int b = 0; b < nparts; ++b) s += partial[(size_t)b * len + t];
    G[t < nw ? goff_w + t : goff_b + (t - nw)] += s;
}

// Adam on the FP32 master weights, then the BF16 copies (and the transposed
// copies of the layer weights used by the data-gradient products); clears G.
__global__ void k_adam(int np, float *P, float *G, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2, uint16_t *Pb, uint16_t *PbT)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const float g = G[t];
    const float mm = b1 * m[t] + (1.f - b1) * g;
    const float vv = b2 * v[t] + (1.f - b2) * g * g;
    m[t] = mm;
    v[t] = vv;
    const float p = P[t] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    P[t] = p;
    G[t] = 0.f;
    const uint16_t b = f2bf(p);
    Pb[t] = b;
    for (int l = 0; l < kLayers; ++l) {
        const int w0 = off_W(l);
        if (t >= w0 && t < w0 + kW * 2 * kW) {
            const int o = (t - w0) / (2 * kW), k = (t - w0) % (2 * kW);
            PbT[(size_t)l * 2 * kW * kW + (size_t)k * kW + o] = b;
        }
    }
}

__global__ void k_refresh_bf16(int np, const float *P, uint16_t *Pb, uint16_t *PbT)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const uint16_t b = f2bf(P[t]);
    Pb[t] = b;
    for (int l = 0; l < kLayers; ++l) {
        const int w0 = off_W(l);
        if (t >= w0 && t < w0 + kW * 2 * kW) {
            const int o = (t - w0) / (2 * kW), k = (t - w0) % (2 * kW);
            PbT[(size_t)l * 2 * kW * kW + (size_t)k * kW + o] = b;
        }
    }
}

// ----------------------------------------------------------------- features and targets
struct FeatParams {
    float dt, s;             // step, target scale dt^2 |g|
    float inv_h;             // 1 / edge length
};

// Per body: geo = (centre xyz, 1 / bounding-box diagonal), mat = (log10 E, nu).
struct BodyFeat {
    const int32_t *body;     // body of each vertex (colour-major)
    const float4 *geo;
    const float2 *mat;
    const float4 *grav;
};

__device__ __forceinline__ float3 inertia_disp(int i, const float4 *uprev, const float4 *vel, const uint8_t *active,
                                               float dt, float4 g)
{
    const float4 u = uprev[i];
    if (!active[i]) return make_float3(u.x, u.y, u.z);
    const float4 v = vel[i];
    const float dt2 = dt * dt;
    return make_float3(u.x + dt * v.x + dt2 * g.x, u.y + dt * v.y + dt2 * g.y, u.z + dt * v.z + dt2 * g.z);
}

// Per-vertex input features (colour-major order); the network is defined in
// oracle/mpnn_ref.py.  From the solver state at the start of a step: rest
// position relative to its body, inertia-prediction displacement, its
// increment, velocity, a local deformation estimate from the neighbours'
// inertia predictions, pin flag, the body's material and gravity direction,
// and the update a Jacobi pass from the inertia prediction would take (dxj,
// swept vertices [0, nsw) only).
__global__ void k_features(int n, const float4 *Xr, const float4 *uprev, const float4 *vel, const uint8_t *active,
                           const int32_t *nb_off, const int32_t *nb, FeatParams p, BodyFeat bf, const float4 *dxj,
                           int nsw, float *f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *o = f + (size_t)i * kFin;
    const int b = bf.body[i];
    const float4 geo = bf.geo[b];
    const float2 mat = bf.mat[b];
    const float4 g = bf.grav[b];
    const float4 X = Xr[i];
    const float3 uy = inertia_disp(i, uprev, vel, active, p.dt, g);
    const float4 up = uprev[i], v = vel[i];
    o[0] = (X.x - geo.x) * geo.w;
    o[1] = (X.y - geo.y) * geo.w;
    o[2] = (X.z - geo.z) * geo.w;
    o[3] = uy.x * p.inv_h;
    o[4] = uy.y * p.inv_h;
    o[5] = uy.z * p.inv_h;
    o[6] = (uy.x - up.x) / p.s;
    o[7] = (uy.y - up.y) / p.s;
    o[8] = (uy.z - up.z) / p.s;
    o[9] = v.x * p.dt * p.inv_h;
    o[10] = v.y * p.dt * p.inv_h;
    o[11] = v.z * p.dt * p.inv_h;
    float D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int q0 = nb_off[i], q1 = nb_off[i + 1];
    for (int q = q0; q < q1; ++q) {
        const int j = nb[q];
        const float4 Xj = Xr[j];
        const float3 uj = inertia_disp(j, uprev, vel, active, p.dt, g);
        const float dX[3] = {Xj.x - X.x, Xj.y - X.y, Xj.z - X.z};
        const float du[3] = {uj.x - uy.x, uj.y - uy.y, uj.z - uy.z};
        const float w = 1.f / (dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2]);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) D[3 * r + c] += du[r] * dX[c] * w;
    }
    const float sc = q1 > q0 ? 1.f / (float)(q1 - q0) : 0.f;
    for (int q = 0; q < 9; ++q) o[12 + q] = D[q] * sc * 100.f;
    o[21] = active[i] ? 0.f : 1.f;
    o[22] = mat.x - 5.f;
    o[23] = (mat.y - 0.375f) * 10.f;
    o[24] = 1.f;
    const float gn = sqrtf(g.x * g.x + g.y * g.y + g.z * g.z) + 1e-30f;
    o[25] = g.x / gn;
    o[26] = g.y / gn;
    o[27] = g.z / gn;
    const float4 d = i < nsw ? dxj[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    o[28] = d.x / p.s;
    o[29] = d.y / p.s;
    o[30] = d.z / p.s;
    o[31] = 0.f;
}

// The plain initial guess (inertia prediction, per-body gravity) in u and the
// inertia target in tgt, as the step's own init writes them.
__global__ void k_plain_guess(int n, float4 *u, const float4 *uprev, const float4 *vel, float4 *tgt,
                              const float *mass, const uint8_t *active, const int32_t *body, const float4 *grav,
                              float dt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) init_vertex(i, u, uprev, vel, tgt, mass, active, dt, grav[body[i]]);
}

// training target of the step just taken: (x_{n+1} - x_inertia) / s
__global__ void k_target(int n, const float4 *u, const float4 *tgt, const uint8_t *active, float inv_s, float *t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = u[i], y = tgt[i];
    const bool a = active[i];
    t[(size_t)i * 3] = a ? (x.x - y.x) * inv_s : 0.f;
    t[(size_t)i * 3 + 1] = a ? (x.y - y.y) * inv_s : 0.f;
    t[(size_t)i * 3 + 2] = a ? (x.z - y.z) * inv_s : 0.f;
}

// network guess: x_inertia + s out (active vertices)
__global__ void k_guess(int n, const float4 *tgt, const uint8_t *active, const float *out, float s, float4 *dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 y = tgt[i];
    if (active[i])
        dst[i] = make_float4(y.x + s * out[(size_t)i * 3], y.y + s * out[(size_t)i * 3 + 1],
                             y.z + s * out[(size_t)i * 3 + 2], 0.f);
    else
        dst[i] = make_float4(y.x, y.y, y.z, 0.f);
}

// The update a Jacobi pass from the positions in a.u would take at each swept
// vertex (no update is applied): dx (nullable) and the first-sweep residual
// bits |dx_x+dx_y+dx_z| per vertex in m.
template <int L, int B, int MAT>
__global__ __launch_bounds__(256) void k_jacobi(SweepArgs a, int nv, float4 *dx, uint32_t *m)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int vi = t / L, lane = t % L;
    const bool valid = vi < nv;
    const int v = valid ? vi : 0;
    const float3 d = eval_vertex<L, B, MAT>(a, v, valid, lane);
    if (valid && lane == 0) {
        if (dx) dx[v] = make_float4(d.x, d.y, d.z, 0.f);
        m[v] = metric_bits(d);
    }
}

// One workgroup per body: keep the guess with the smaller first-sweep
// residual (max over the body's vertices) and write it to u; force = 1 takes
// the network guess without the comparison.
__global__ __launch_bounds__(256) void k_choose_guess(const int2 *ranges, int C, const uint32_t *m0,
                                                      const uint32_t *m1, const float4 *alt, float4 *u,
                                                      int32_t *use_alt, int force)
{
    __shared__ int s_use;
    const int2 *r = ranges + (size_t)blockIdx.x * (C + 1);
    uint32_t a0 = 0, a1 = 0;
    if (!force) {
        for (int c = 0; c < C; ++c) {
            const int2 rc = r[c];
            for (int i = rc.x + threadIdx.x; i < rc.x + rc.y; i += blockDim.x) {
                a0 = max(a0, m0[i]);
                a1 = max(a1, m1[i]);
            }
        }
        a0 = block_max_u32(a0);
        __syncthreads();
        a1 = block_max_u32(a1);
    }
    if (threadIdx.x == 0) {
        s_use = force || __uint_as_float(a1) < __uint_as_float(a0) ? 1 : 0;
        use_alt[blockIdx.x] = s_use;
    }
    __syncthreads();
    if (!s_use) return;
    for (int c = 0; c < C; ++c) {
        const int2 rc = r[c];
        for (int i = rc.x + threadIdx.x; i < rc.x + rc.y; i += blockDim.x) u[i] = alt[i];
    }
}

// full neighbour lists (distinct, without self), colour-major ids
template <bool FILL>
__global__ void k_graph(const int32_t *perm, const int32_t *iperm, const int32_t *tets, const int32_t *off,
                        const int32_t *inc, int n, int32_t *cnt, const int32_t *g_off, int32_t *g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = perm[i];
    const int a = off[v], b = off[v + 1];
    int k = 0;
    const int w = FILL ? g_off[i] : 0;
    for (int s = a; s < b; ++s) {
        const int e = inc[s] >> 2;
        for (int q = 0; q < 4; ++q) {
            const int u = tets[4 * e + q];
            if (u == v) continue;
            bool seen = false;
            for (int q2 = 0; q2 < q && !seen; ++q2) seen = tets[4 * e + q2] == u;
            for (int s2 = a; s2 < s && !seen; ++s2) {
                const int e2 = inc[s2] >> 2;
                for (int q2 = 0; q2 < 4; ++q2) seen |= tets[4 * e2 + q2] == u;
            }
            if (seen) continue;
            if (FILL) g[w + k] = iperm[u];
            ++k;
        }
    }
    if (!FILL) cnt[i] = k;
}

__global__ void k_rest_new(int n, const int32_t *perm, const double *X, float4 *Xr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = perm[i];
    Xr[i] = make_float4((float)X[3 * v], (float)X[3 * v + 1], (float)X[3 * v + 2], 0.f);
}

template <int L, int MAT>
void launch_jacobi(const SweepArgs &a, int nv, float4 *dx, uint32_t *m, hipStream_t st)
{
    k_jacobi<L, Batch<L>::value, MAT><<<nblk((int64_t)std::max(nv, 1) * L, 256), 256, 0, st>>>(a, nv, dx, m);
}

}  // namespace
}  // namespace vbd

using namespace vbd;

struct vbd_mpnn {
    vbd_solver *s = nullptr;
    int n = 0, n_edges = 0;
    int32_t *nb_off = nullptr, *nb = nullptr;
    float4 *Xr = nullptr;
    float *P = nullptr, *G = nullptr, *Am = nullptr, *Av = nullptr;
    uint16_t *Pb = nullptr, *PbT = nullptr;
    float *f = nullptr, *z0 = nullptr, *h[kLayers + 1] = {}, *a[kLayers] = {}, *z[kLayers] = {};
    float *out = nullptr, *dh = nullptr, *dx = nullptr, *dz = nullptr, *partial = nullptr, *loss = nullptr;
    int wg_blocks = 0;
    // sample store
    int cap = 0;
    float *F = nullptr, *T = nullptr;
    float4 *guess = nullptr;
    uint32_t *rm0 = nullptr, *rm1 = nullptr;  // per-vertex first-sweep residuals of the two guesses
    float4 *dxj = nullptr;                     // Jacobi update from the plain guess (a feature)
    int policy = VBD_WARM_CHOOSE;
    int32_t *use_alt = nullptr, *use_alt_host = nullptr;  // per body
    float *loss_host = nullptr;
    int adam_t = 0;
    FeatParams fp{};
    float4 *body_geo = nullptr;
    float2 *body_mat = nullptr;
    std::vector<float2> body_mat_host;
    int n_active = 0;
};

namespace {

int mpnn_free(vbd_mpnn *m)
{
    void *ps[] = {m->nb_off, m->nb, m->Xr, m->P, m->G, m->Am, m->Av, m->Pb, m->PbT, m->f, m->z0, m->out, m->dh,
                  m->dx, m->dz, m->partial, m->loss, m->F, m->T, m->guess, m->rm0, m->rm1, m->use_alt,
                  m->body_geo, m->body_mat, m->dxj};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    for (auto p : m->h)
        if (p) (void)hipFree(p);
    for (auto p : m->a)
        if (p) (void)hipFree(p);
    for (auto p : m->z)
        if (p) (void)hipFree(p);
    if (m->use_alt_host) (void)hipHostFree(m->use_alt_host);
    if (m->loss_host) (void)hipHostFree(m->loss_host);
    delete m;
    return 0;
}

template <typename T>
int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
        set_error("hipMalloc failed (warm-start network)");
        return VBD_ENOMEM;
    }
    return 0;
}

// forward of sample features f (m->f); with_grad: loss + backward into m->G
int mpnn_forward_backward(vbd_mpnn *m, const float *target, bool with_grad, float inv_denom)
{
    hipStream_t st = m->s->stream;
    const int n = m->n;
    const dim3 g1(nblk(n, 64), 1), g2(nblk(n, 64), 2);
    k_mfma_linear<kFin, 0><<<g1, 256, 0, st>>>(n, m->f, kFin, nullptr, m->Pb + kOffWin, m->P + kOffBin, m->z0,
                                                nullptr, m->h[0], kW);
    for (int l = 0; l < kLayers; ++l) {
        k_neighbour_mean<<<nblk((int64_t)n * kW, 256), 256, 0, st>>>(n, m->nb_off, m->nb, m->h[l], m->a[l]);
        k_mfma_linear<2 * kW, 1><<<g1, 256, 0, st>>>(n, m->h[l], kW, m->a[l], m->Pb + off_W(l), m->P + off_b(l),
                                                      m->z[l], m->h[l], m->h[l + 1], kW);
    }
    k_output<<<nblk(n, 256), 256, 0, st>>>(n, m->h[kLayers], m->Pb, m->P, target, m->s->active, inv_denom,
                                            with_grad ? 1 : 0, m->out, m->dh, m->G, m->loss);
    VBD_HIP(hipGetLastError());
    if (!with_grad) return 0;
    const int64_t nw = (int64_t)n * kW;
    for (int l = kLayers - 1; l >= 0; --l) {
        k_relu_grad<<<nblk(nw, 256), 256, 0, st>>>(nw, m->dh, m->z[l], m->dz);
        k_weight_grad<2 * kW><<<m->wg_blocks, 256, 0, st>>>(n, m->dz, m->h[l], kW, m->a[l], m->partial);
        const int len = kW * 2 * kW + kW;
        k_sum_partials<<<nblk(len, 256), 256, 0, st>>>(m->wg_blocks, len, m->partial, m->G, off_W(l), off_b(l),
                                                        kW * 2 * kW);
        // dx = bf(dz) bf(W_l): [n, 128]; the transposed BF16 copy is [128, 64]
        k_mfma_linear<kW, 2><<<g2, 256, 0, st>>>(n, m->dz, kW, nullptr, m->PbT + (size_t)l * 2 * kW * kW, nullptr,
                                                  m->dx, nullptr, nullptr, 2 * kW);
        k_backprop_mean<<<nblk(nw, 256), 256, 0, st>>>(n, m->nb_off, m->nb, m->dx, m->dh);
    }
    k_relu_grad<<<nblk(nw, 256), 256, 0, st>>>(nw, m->dh, m->z0, m->dz);
    k_weight_grad<kFin><<<m->wg_blocks, 256, 0, st>>>(n, m->dz, m->f, kFin, nullptr, m->partial);
    const int len = kW * kFin + kW;
    k_sum_partials<<<nblk(len, 256), 256, 0, st>>>(m->wg_blocks, len, m->partial, m->G, kOffWin, kOffBin, kW * kFin);
    VBD_HIP(hipGetLastError());
    return 0;
}

void set_feature_params(vbd_mpnn *m, double dt)
{
    vbd_solver *s = m->s;
    m->fp.dt = (float)dt;
    const double gn = std::sqrt(s->grav[0] * s->grav[0] + s->grav[1] * s->grav[1] + s->grav[2] * s->grav[2]);
    m->fp.s = (float)(dt * dt * std::max(gn, 1e-3));
}

BodyFeat body_feat(vbd_mpnn *m)
{
    return BodyFeat{m->s->body_new, m->body_geo, m->body_mat, m->s->d_grav};
}

SweepArgs sweep_args_at(vbd_solver *s, float4 *u, double dt)
{
    SweepArgs a;
    a.u = u;
    a.tgt = s->tgt;
    a.uprev = s->uprev;
    a.slot_off = s->slot_off;
    a.slots = s->slots;
    a.ent = s->slot_ent;
    a.ns = std::max<int64_t>(s->n_slots, 1);
    a.part = nullptr;
    a.ctrl = s->ctrl;
    a.kd_dt = (float)(s->damping / dt);
    return a;
}

// Jacobi update from the guess in u (and its residual bits into res).
void jacobi(vbd_mpnn *m, float4 *u, double dt, float4 *dx, uint32_t *res)
{
    vbd_solver *s = m->s;
    constexpr int L = 8;
    const SweepArgs a = sweep_args_at(s, u, dt);
    if (s->material == kSNH) launch_jacobi<L, kSNH>(a, s->n_swept, dx, res, s->stream);
    else launch_jacobi<L, kStVK>(a, s->n_swept, dx, res, s->stream);
}

// Features of the solver's state at the start of a step of size dt into f.
// plain_in_u: s->u and s->tgt already hold the plain guess and the inertia
// target (the step's init ran); otherwise the plain guess is formed in
// m->guess and the solver's positions are left alone.
void features(vbd_mpnn *m, double dt, bool plain_in_u, float *f)
{
    vbd_solver *s = m->s;
    hipStream_t st = s->stream;
    set_feature_params(m, dt);
    float4 *u = s->u;
    if (!plain_in_u) {
        u = m->guess;
        k_plain_guess<<<nblk(m->n, 256), 256, 0, st>>>(m->n, u, s->uprev, s->vel, s->tgt, s->mass, s->active,
                                                       s->body_new, s->d_grav, (float)dt);
    }
    jacobi(m, u, dt, m->dxj, m->rm0);
    k_features<<<nblk(m->n, 256), 256, 0, st>>>(m->n, m->Xr, s->uprev, s->vel, s->active, m->nb_off, m->nb, m->fp,
                                                 body_feat(m), m->dxj, s->n_swept, f);
}

}  // namespace

// ----------------------------------------------------------------- host hooks used by vbd_step.hip
namespace vbd {

// Replace the plain initial guess in s->u by the network guess when its
// first-sweep residual is smaller (per body; the policy may force either
// guess).  Called after the step's init.
int mpnn_apply_guess(vbd_mpnn *m, double dt)
{
    vbd_solver *s = m->s;
    hipStream_t st = s->stream;
    if (m->policy == VBD_WARM_PLAIN) {
        VBD_HIP(hipMemsetAsync(m->use_alt, 0, sizeof(int32_t) * s->n_bodies, st));
        return 0;
    }
    features(m, dt, true, m->f);
    VBD_TRY(mpnn_forward_backward(m, nullptr, false, 0.f));
    k_guess<<<nblk(m->n, 256), 256, 0, st>>>(m->n, s->tgt, s->active, m->out, m->fp.s, m->guess);
    const int force = m->policy == VBD_WARM_NETWORK ? 1 : 0;
    if (!force) jacobi(m, m->guess, dt, nullptr, m->rm1);
    k_choose_guess<<<s->n_bodies, 256, 0, st>>>(s->d_ranges, s->ncolors, m->rm0, m->rm1, m->guess, s->u, m->use_alt,
                                                force);
    VBD_HIP(hipGetLastError());
    return 0;
}

int mpnn_used_guess(vbd_mpnn *m, int *used)
{
    const int nb = m->s->n_bodies;
    VBD_HIP(hipMemcpyAsync(m->use_alt_host, m->use_alt, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, m->s->stream));
    VBD_HIP(hipStreamSynchronize(m->s->stream));
    int c = 0;
    for (int b = 0; b < nb; ++b) c += m->use_alt_host[b] ? 1 : 0;
    *used = c;
    return 0;
}

}  // namespace vbd

// ----------------------------------------------------------------- C ABI
extern "C" {

int vbd_mpnn_create(vbd_solver_t *s, int32_t sample_capacity, vbd_mpnn_t **out)
{
    if (!s || !out || sample_capacity < 0) {
        set_error("vbd_mpnn_create: needs a solver handle and sample_capacity >= 0");
        return VBD_EINVAL;
    }
    *out = nullptr;
    VBD_HIP(hipSetDevice(s->device));
    vbd_mpnn *m = new vbd_mpnn();
    m->s = s;
    m->n = s->n;
    const int n = s->n;
    hipStream_t st = s->stream;
    int rc = 0;
    auto fail = [&](int r) {
        mpnn_free(m);
        return r;
    };
    // graph
    int32_t *cnt = nullptr;
    if ((rc = dalloc(&cnt, (size_t)n + 1)) || (rc = dalloc(&m->nb_off, (size_t)n + 1))) return fail(rc);
    (void)hipMemsetAsync(cnt, 0, sizeof(int32_t) * (n + 1), st);
    k_graph<false><<<nblk(n, 256), 256, 0, st>>>(s->perm, s->iperm, s->tets_dev, s->inc_off, s->inc, n, cnt, nullptr,
                                                 nullptr);
    size_t tb = 0;
    (void)rocprim_exclusive_scan_i32(nullptr, tb, cnt, m->nb_off, n + 1, st);
    void *tmp = nullptr;
    if (hipMalloc(&tmp, std::max<size_t>(tb, 1)) != hipSuccess) {
        (void)hipFree(cnt);
        return fail(VBD_ENOMEM);
    }
    (void)rocprim_exclusive_scan_i32(tmp, tb, cnt, m->nb_off, n + 1, st);
    int32_t ne = 0;
    (void)hipMemcpyAsync(&ne, m->nb_off + n, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(tmp);
    (void)hipFree(cnt);
    m->n_edges = ne;
    if ((rc = dalloc(&m->nb, ne))) return fail(rc);
    k_graph<true><<<nblk(n, 256), 256, 0, st>>>(s->perm, s->iperm, s->tets_dev, s->inc_off, s->inc, n, nullptr,
                                                m->nb_off, m->nb);
    if ((rc = dalloc(&m->Xr, n))) return fail(rc);
    k_rest_new<<<nblk(n, 256), 256, 0, st>>>(n, s->perm, s->X, m->Xr);
    // parameters
    if ((rc = dalloc(&m->P, kNumParams)) || (rc = dalloc(&m->G, kNumParams)) || (rc = dalloc(&m->Am, kNumParams)) ||
        (rc = dalloc(&m->Av, kNumParams)) || (rc = dalloc(&m->Pb, kNumParams)) ||
        (rc = dalloc(&m->PbT, (size_t)kLayers * 2 * kW * kW)))
        return fail(rc);
    (void)hipMemsetAsync(m->P, 0, sizeof(float) * kNumParams, st);
    (void)hipMemsetAsync(m->G, 0, sizeof(float) * kNumParams, st);
    (void)hipMemsetAsync(m->Am, 0, sizeof(float) * kNumParams, st);
    (void)hipMemsetAsync(m->Av, 0, sizeof(float) * kNumParams, st);
    // activations
    const size_t nw = (size_t)n * kW;
    if ((rc = dalloc(&m->f, (size_t)n * kFin)) || (rc = dalloc(&m->z0, nw)) || (rc = dalloc(&m->out, (size_t)n * 3)) ||
        (rc = dalloc(&m->dh, nw)) || (rc = dalloc(&m->dx, 2 * nw)) || (rc = dalloc(&m->dz, nw)) ||
        (rc = dalloc(&m->loss, 1)) || (rc = dalloc(&m->guess, n)) || (rc = dalloc(&m->use_alt, s->n_bodies)) ||
        (rc = dalloc(&m->rm0, std::max(1, s->n_swept))) || (rc = dalloc(&m->rm1, std::max(1, s->n_swept))) ||
        (rc = dalloc(&m->dxj, std::max(1, s->n_swept))) ||
        (rc = dalloc(&m->body_geo, s->n_bodies)) || (rc = dalloc(&m->body_mat, s->n_bodies)))
        return fail(rc);
    for (int l = 0; l <= kLayers; ++l)
        if ((rc = dalloc(&m->h[l], nw))) return fail(rc);
    for (int l = 0; l < kLayers; ++l)
        if ((rc = dalloc(&m->a[l], nw)) || (rc = dalloc(&m->z[l], nw))) return fail(rc);
    m->wg_blocks = std::max(1, std::min(512, nblk(n, 32)));
    if ((rc = dalloc(&m->partial, (size_t)m->wg_blocks * (kW * 2 * kW + kW)))) return fail(rc);
    if (hipHostMalloc((void **)&m->use_alt_host, sizeof(int32_t) * s->n_bodies, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&m->loss_host, sizeof(float), hipHostMallocDefault) != hipSuccess)
        return fail(VBD_ENOMEM);
    // samples
    m->cap = sample_capacity;
    if (m->cap > 0) {
        if ((rc = dalloc(&m->F, (size_t)m->cap * n * kFin)) || (rc = dalloc(&m->T, (size_t)m->cap * n * 3)))
            return fail(rc);
    }
    // per-body feature normalisation from the rest shape
    std::vector<double> X(3 * (size_t)n);
    if (hipMemcpy(X.data(), s->X, sizeof(double) * 3 * n, hipMemcpyDeviceToHost) != hipSuccess) return fail(VBD_EHIP);
    {
        const int nb = s->n_bodies;
        std::vector<float4> geo(nb);
        for (int b = 0; b < nb; ++b) {
            const int v0 = s->body_voff_host[b], v1 = s->body_voff_host[b + 1];
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int v = v0; v < v1; ++v)
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], X[3 * v + k]);
                    hi[k] = std::max(hi[k], X[3 * v + k]);
                }
            if (v1 <= v0) lo[0] = lo[1] = lo[2] = hi[0] = hi[1] = hi[2] = 0.0;
            const double L = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                       (hi[2] - lo[2]) * (hi[2] - lo[2]));
            geo[b] = make_float4((float)(0.5 * (lo[0] + hi[0])), (float)(0.5 * (lo[1] + hi[1])),
                                 (float)(0.5 * (lo[2] + hi[2])), (float)(1.0 / std::max(L, 1e-12)));
        }
        m->body_mat_host.assign(nb, make_float2(5.f, 0.4f));
        (void)hipMemcpy(m->body_geo, geo.data(), sizeof(float4) * nb, hipMemcpyHostToDevice);
        (void)hipMemcpy(m->body_mat, m->body_mat_host.data(), sizeof(float2) * nb, hipMemcpyHostToDevice);
    }
    // edge length: cube root of the mean tet volume x 6
    m->fp.inv_h = (float)(1.0 / std::max(s->edge_length, 1e-12));
    std::vector<uint8_t> act(n);
    (void)hipMemcpy(act.data(), s->active, n, hipMemcpyDeviceToHost);
    for (uint8_t x : act) m->n_active += x ? 1 : 0;
    VBD_HIP(hipStreamSynchronize(st));
    *out = m;
    return 0;
}

void vbd_mpnn_destroy(vbd_mpnn_t *m)
{
    if (m) mpnn_free(m);
}

int vbd_mpnn_num_params(void) { return kNumParams; }

int vbd_mpnn_set_policy(vbd_mpnn_t *m, int32_t policy)
{
    if (!m || policy < VBD_WARM_CHOOSE || policy > VBD_WARM_PLAIN) {
        set_error("vbd_mpnn_set_policy: policy must be VBD_WARM_CHOOSE, VBD_WARM_NETWORK or VBD_WARM_PLAIN");
        return VBD_EINVAL;
    }
    m->policy = policy;
    return 0;
}

int vbd_mpnn_set_params(vbd_mpnn_t *m, const float *params)
{
    if (!m || !params) {
        set_error("null argument");
        return VBD_EINVAL;
    }
    hipStream_t st = m->s->stream;
    VBD_HIP(hipMemcpyAsync(m->P, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, st));
    k_refresh_bf16<<<nblk(kNumParams, 256), 256, 0, st>>>(kNumParams, m->P, m->Pb, m->PbT);
    VBD_HIP(hipMemsetAsync(m->Am, 0, sizeof(float) * kNumParams, st));
    VBD_HIP(hipMemsetAsync(m->Av, 0, sizeof(float) * kNumParams, st));
    VBD_HIP(hipMemsetAsync(m->G, 0, sizeof(float) * kNumParams, st));
    VBD_HIP(hipStreamSynchronize(st));
    m->adam_t = 0;
    return 0;
}

int vbd_mpnn_get_params(vbd_mpnn_t *m, float *params)
{
    if (!m || !params) {
        set_error("null argument");
        return VBD_EINVAL;
    }
    VBD_HIP(hipMemcpyAsync(params, m->P, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, m->s->stream));
    VBD_HIP(hipStreamSynchronize(m->s->stream));
    return 0;
}

int vbd_mpnn_set_material(vbd_mpnn_t *m, double youngs, double poisson)
{
    if (!m || !(youngs > 0.0)) {
        set_error("bad arguments");
        return VBD_EINVAL;
    }
    for (auto &x : m->body_mat_host) x = make_float2((float)std::log10(youngs), (float)poisson);
    VBD_HIP(hipMemcpy(m->body_mat, m->body_mat_host.data(), sizeof(float2) * m->body_mat_host.size(),
                      hipMemcpyHostToDevice));
    return 0;
}

int vbd_mpnn_set_body_materials(vbd_mpnn_t *m, const double *youngs, const double *poisson)
{
    if (!m || !youngs || !poisson) {
        set_error("null argument");
        return VBD_EINVAL;
    }
    for (size_t b = 0; b < m->body_mat_host.size(); ++b) {
        if (!(youngs[b] > 0.0)) {
            set_error("Young's modulus must be > 0");
            return VBD_EINVAL;
        }
        m->body_mat_host[b] = make_float2((float)std::log10(youngs[b]), (float)poisson[b]);
    }
    VBD_HIP(hipMemcpy(m->body_mat, m->body_mat_host.data(), sizeof(float2) * m->body_mat_host.size(),
                      hipMemcpyHostToDevice));
    return 0;
}

// One plain step of the solver that also stores a training sample: the
// features of the state before the step and the converged correction.
int vbd_mpnn_collect(vbd_mpnn_t *m, int32_t slot, double dt, double tol, int32_t max_iter, int32_t *iters_out)
{
    if (!m || slot < 0 || slot >= m->cap || !(dt > 0.0) || max_iter < 1) {
        set_error("vbd_mpnn_collect: bad arguments");
        return VBD_EINVAL;
    }
    vbd_solver *s = m->s;
    hipStream_t st = s->stream;
    VBD_HIP(hipSetDevice(s->device));
    features(m, dt, false, m->F + (size_t)slot * m->n * kFin);
    int it = 0;
    double met = 0.0;
    if (s->batched) {
        VBD_TRY(step_bodies(s, dt, tol, max_iter));
        it = s->last_iters;
    } else {
        VBD_TRY(step_device(s, dt, tol, max_iter, &it, &met));
    }
    k_target<<<nblk(m->n, 256), 256, 0, st>>>(m->n, s->uprev, s->tgt, s->active, 1.f / m->fp.s,
                                              m->T + (size_t)slot * m->n * 3);
    VBD_HIP(hipGetLastError());
    VBD_HIP(hipStreamSynchronize(st));
    if (iters_out) *iters_out = it;
    return 0;
}

// Loss (and optionally gradients, accumulated into the gradient buffer) of one
// stored sample.  grads (nullable): receives the accumulated gradient and
// clears it.
int vbd_mpnn_loss_grad(vbd_mpnn_t *m, int32_t slot, float *loss_out, float *grads)
{
    if (!m || slot < 0 || slot >= m->cap) {
        set_error("vbd_mpnn_loss_grad: bad arguments");
        return VBD_EINVAL;
    }
    hipStream_t st = m->s->stream;
    VBD_HIP(hipMemcpyAsync(m->f, m->F + (size_t)slot * m->n * kFin, sizeof(float) * m->n * kFin,
                           hipMemcpyDeviceToDevice, st));
    VBD_HIP(hipMemsetAsync(m->loss, 0, sizeof(float), st));
    const float inv_denom = 1.f / (3.f * (float)std::max(1, m->n_active));
    VBD_TRY(mpnn_forward_backward(m, m->T + (size_t)slot * m->n * 3, true, inv_denom));
    VBD_HIP(hipMemcpyAsync(m->loss_host, m->loss, sizeof(float), hipMemcpyDeviceToHost, st));
    if (grads) {
        VBD_HIP(hipMemcpyAsync(grads, m->G, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, st));
        VBD_HIP(hipMemsetAsync(m->G, 0, sizeof(float) * kNumParams, st));
    }
    VBD_HIP(hipStreamSynchronize(st));
    if (loss_out) *loss_out = *m->loss_host;
    return 0;
}

// Adam over the stored samples [0, n_samples): `epochs` passes in a shuffled
// order, `batch` samples per update.  loss_per_epoch (nullable, epochs
// entries): mean sample loss of each epoch.
int vbd_mpnn_train(vbd_mpnn_t *m, int32_t n_samples, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                   float *loss_per_epoch)
{
    if (!m || n_samples < 1 || n_samples > m->cap || epochs < 1 || batch < 1 || !(lr > 0.f)) {
        set_error("vbd_mpnn_train: bad arguments");
        return VBD_EINVAL;
    }
    hipStream_t st = m->s->stream;
    VBD_HIP(hipSetDevice(m->s->device));
    std::mt19937_64 rng(seed);
    std::vector<int> order(n_samples);
    for (int i = 0; i < n_samples; ++i) order[i] = i;
    const float inv_denom = 1.f / (3.f * (float)std::max(1, m->n_active) * (float)batch);
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        VBD_HIP(hipMemsetAsync(m->loss, 0, sizeof(float), st));
        int nb = 0;
        for (int b0 = 0; b0 < n_samples; b0 += batch) {
            const int b1 = std::min(n_samples, b0 + batch);
            for (int q = b0; q < b1; ++q) {
                const int slot = order[q];
                VBD_HIP(hipMemcpyAsync(m->f, m->F + (size_t)slot * m->n * kFin, sizeof(float) * m->n * kFin,
                                       hipMemcpyDeviceToDevice, st));
                VBD_TRY(mpnn_forward_backward(m, m->T + (size_t)slot * m->n * 3, true, inv_denom));
            }
            ++m->adam_t;
            const float b1c = 0.9f, b2c = 0.999f;
            const float c1 = 1.f - std::pow(b1c, (float)m->adam_t), c2 = 1.f - std::pow(b2c, (float)m->adam_t);
            k_adam<<<nblk(kNumParams, 256), 256, 0, st>>>(kNumParams, m->P, m->G, m->Am, m->Av, lr, b1c, b2c, 1e-8f,
                                                          c1, c2, m->Pb, m->PbT);
            ++nb;
        }
        VBD_HIP(hipGetLastError());
        VBD_HIP(hipMemcpyAsync(m->loss_host, m->loss, sizeof(float), hipMemcpyDeviceToHost, st));
        VBD_HIP(hipStreamSynchronize(st));
        if (loss_per_epoch) loss_per_epoch[ep] = *m->loss_host * (float)batch / (float)n_samples;
    }
    return 0;
}

// Network prediction (out, [n, 3] in the original vertex order, in units of
// dt^2 |g|) for the features of stored sample `slot` (slot < 0: of the
// solver's current state).
int vbd_mpnn_predict(vbd_mpnn_t *m, int32_t slot, double dt, float *out_host)
{
    if (!m || !out_host || slot >= m->cap) {
        set_error("vbd_mpnn_predict: bad arguments");
        return VBD_EINVAL;
    }
    vbd_solver *s = m->s;
    hipStream_t st = s->stream;
    set_feature_params(m, dt);
    if (slot >= 0)
        VBD_HIP(hipMemcpyAsync(m->f, m->F + (size_t)slot * m->n * kFin, sizeof(float) * m->n * kFin,
                               hipMemcpyDeviceToDevice, st));
    else
        features(m, dt, false, m->f);
    VBD_TRY(mpnn_forward_backward(m, nullptr, false, 0.f));
    std::vector<float> tmp((size_t)m->n * 3);
    std::vector<int32_t> perm(m->n);
    VBD_HIP(hipMemcpyAsync(tmp.data(), m->out, sizeof(float) * 3 * m->n, hipMemcpyDeviceToHost, st));
    VBD_HIP(hipMemcpyAsync(perm.data(), s->perm, sizeof(int32_t) * m->n, hipMemcpyDeviceToHost, st));
    VBD_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < m->n; ++i)
        for (int k = 0; k < 3; ++k) out_host[(size_t)perm[i] * 3 + k] = tmp[(size_t)i * 3 + k];
    return 0;
}

// Copy stored sample `slot` to the host in the original vertex order:
// features [n, 32], targets [n, 3] (either may be NULL); plus the graph in
// original ids: nb_off [n+1], nb [n_edges] (nullable).
int vbd_mpnn_read_sample(vbd_mpnn_t *m, int32_t slot, float *features, float *targets, int32_t *nb_off,
                         int32_t *nb, int32_t *n_edges)
{
    if (!m || slot < 0 || slot >= m->cap) {
        set_error("vbd_mpnn_read_sample: bad arguments");
        return VBD_EINVAL;
    }
    vbd_solver *s = m->s;
    const int n = m->n;
    std::vector<int32_t> perm(n), off(n + 1), g(std::max(1, m->n_edges));
    std::vector<float> F((size_t)n * kFin), T((size_t)n * 3);
    VBD_HIP(hipMemcpy(perm.data(), s->perm, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    VBD_HIP(hipMemcpy(off.data(), m->nb_off, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
    if (m->n_edges) VBD_HIP(hipMemcpy(g.data(), m->nb, sizeof(int32_t) * m->n_edges, hipMemcpyDeviceToHost));
    VBD_HIP(hipMemcpy(F.data(), m->F + (size_t)slot * n * kFin, sizeof(float) * n * kFin, hipMemcpyDeviceToHost));
    VBD_HIP(hipMemcpy(T.data(), m->T + (size_t)slot * n * 3, sizeof(float) * n * 3, hipMemcpyDeviceToHost));
    if (n_edges) *n_edges = m->n_edges;
    if (features)
        for (int i = 0; i < n; ++i) std::memcpy(features + (size_t)perm[i] * kFin, &F[(size_t)i * kFin], sizeof(float) * kFin);
    if (targets)
        for (int i = 0; i < n; ++i) std::memcpy(targets + (size_t)perm[i] * 3, &T[(size_t)i * 3], sizeof(float) * 3);
    if (nb_off && nb) {
        // original order: vertex perm[i] owns colour-major row i
        std::vector<int32_t> deg(n);
        for (int i = 0; i < n; ++i) deg[perm[i]] = off[i + 1] - off[i];
        nb_off[0] = 0;
        for (int v = 0; v < n; ++v) nb_off[v + 1] = nb_off[v] + deg[v];
        for (int i = 0; i < n; ++i) {
            const int v = perm[i];
            for (int q = off[i]; q < off[i + 1]; ++q) nb[nb_off[v] + (q - off[i])] = perm[g[q]];
        }
    }
    return 0;
}

// A time step that starts from the network's guess where it beats the plain
// guess on the first-sweep residual (per body).  used_network_out: the number
// of bodies that started from the network guess.
int vbd_step_warm(vbd_solver_t *s, vbd_mpnn_t *m, double dt, double tol, int32_t max_iter, int32_t *iters_out,
                  double *metric_out, int32_t *used_network_out)
{
    if (!s || !m || m->s != s || !(dt > 0.0) || max_iter < 1) {
        set_error("vbd_step_warm: bad arguments");
        return VBD_EINVAL;
    }
    VBD_HIP(hipSetDevice(s->device));
    s->warm = m;
    int it = 0;
    double met = 0.0;
    int rc = 0;
    if (s->batched) {
        rc = step_bodies(s, dt, tol, max_iter);
        it = s->last_iters;
        float mx = 0.f;
        for (int b = 0; b < s->n_bodies && rc == 0; ++b) {
            const float x = __builtin_bit_cast(float, (uint32_t)s->body_out_host[b].y);
            if (!(x <= mx)) mx = x;
        }
        met = mx;
    } else {
        rc = step_device(s, dt, tol, max_iter, &it, &met);
    }
    s->warm = nullptr;
    if (rc) return rc;
    int used = 0;
    VBD_TRY(mpnn_used_guess(m, &used));
    if (iters_out) *iters_out = it;
    if (metric_out) *metric_out = met;
    if (used_network_out) *used_network_out = used;
    return 0;
}

}  // extern "C"

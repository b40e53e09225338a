// Path-guiding network kernels: weight packing, inference, training
// (forward + back-propagation on the matrix cores, deterministic gradient
// reduction) and Adam.  No reference counterpart (kazen_proto has no path
// guiding); the methods are named in kp_guide.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/kp_hip.h"
#include "kp_ctx.h"
#include "kp_guide.h"
#include "kp_guide_host.h"

namespace kp {

// ------------------------------------------------------------ packing ----
// Converts the FP32 masters into the per-lane bf16 matrix-core fragments.
__global__ void k_guide_pack(const float* __restrict__ P, u32x4* wf, u32x4* wb) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int total_f = KP_G_WF_FRAGS * 64, total_b = KP_G_WB_FRAGS * 64;
    if (tid < total_f) {
        const int frag = tid / 64, lane = tid % 64, g = lane >> 4, r = lane & 15;
        int l = frag < 32 ? frag / 8 : 4;
        const int local = frag - g_wf_base(l);
        const int t = local / 2, c = local % 2;
        const int din = g_dim_in(l), dout = g_dim_out(l);
        const float* W = P + g_layer_offset(l);
        float v[8];
        for (int j = 0; j < 8; ++j) {
            const int k = 32 * c + 8 * g + j, n = 16 * t + r;
            v[j] = k < din ? W[k * dout + n] : 0.0f;
        }
        wf[tid] = u32x4{pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3]), pack_bf2(v[4], v[5]), pack_bf2(v[6], v[7])};
    } else if (tid < total_f + total_b) {
        const int id = tid - total_f;
        const int frag = id / 64, lane = id % 64, g = lane >> 4, r = lane & 15;
        const int l = 1 + frag / 8;
        const int local = frag - g_wb_base(l);
        const int nchunks = g_dim_out(l) / 32;
        const int kt = local / nchunks, c = local % nchunks;
        const int dout = g_dim_out(l);
        const float* W = P + g_layer_offset(l);
        float v[8];
        for (int j = 0; j < 8; ++j) v[j] = W[(16 * kt + r) * dout + 32 * c + 8 * g + j];
        wb[id] = u32x4{pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3]), pack_bf2(v[4], v[5]), pack_bf2(v[6], v[7])};
    }
}

// Writes the 48 encoded features (bf16) of a sample into an LDS row; columns
// 48..63 are zero.
__device__ __forceinline__ void g_write_features(uint32_t* row, const float* in6, bool valid) {
    float f[48];
    if (valid) g_encode(in6[0], in6[1], in6[2], in6[3], in6[4], in6[5], f);
#pragma unroll
    for (int w = 0; w < 24; ++w) row[w] = valid ? pack_bf2(f[2 * w], f[2 * w + 1]) : 0u;
#pragma unroll
    for (int w = 24; w < 32; ++w) row[w] = 0u;
}

// ---------------------------------------------------------- inference ----
// mode 0: outputs; 1: mixture pdf of dirs; 2: mixture sample from u3.
__global__ __launch_bounds__(256) void k_guide_eval(GuideNet net, const float* __restrict__ in6, uint32_t n,
                                                    float* out32, const float* dirs, const float* u3,
                                                    float* dirs_out, float* pdf_out, int mode) {
    // mode 0: 64 feature rows + two 16-row activation buffers per wave
    __shared__ __align__(16) uint32_t lds[4][96 * KP_G_ROW];   // 16-byte row fragments
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    uint32_t* feat = lds[wave];
    uint32_t* ping = feat + 64 * KP_G_ROW;
    uint32_t* pong = ping + 16 * KP_G_ROW;
    const uint32_t s = (blockIdx.x * 4 + wave) * 64 + lane;
    const bool valid = s < n;
    if (mode == 0) {
        g_write_features(feat + lane * KP_G_ROW, in6 + 6 * (size_t)(valid ? s : 0), valid);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t* act[4] = {ping, pong, ping, pong};
        for (int q = 0; q < 4; ++q) {
            uint32_t* rows = feat + 16 * q * KP_G_ROW;
            g_forward_chunk(net, rows, act, (float*)rows, KP_G_ROW);
        }
        const float* o = (const float*)(feat + lane * KP_G_ROW);
        if (valid)
            for (int i = 0; i < KP_G_OUT; ++i) out32[(size_t)s * KP_G_OUT + i] = o[i];
        return;
    }
    // the wave-level query routine the guided shade kernel uses
    const size_t si = valid ? s : 0;
    GuideQuery q;
    q.px = in6[6 * si]; q.py = in6[6 * si + 1]; q.pz = in6[6 * si + 2];
    q.nx = in6[6 * si + 3]; q.ny = in6[6 * si + 4]; q.nz = in6[6 * si + 5];
    q.sample = mode == 2;
    if (mode == 1) {
        q.ax = dirs[3 * si]; q.ay = dirs[3 * si + 1]; q.az = dirs[3 * si + 2];
        q.u0 = q.u1 = q.u2 = 0.0f;
    } else {
        q.ax = q.ay = q.az = 0.0f;
        q.u0 = u3[3 * si]; q.u1 = u3[3 * si + 1]; q.u2 = u3[3 * si + 2];
    }
    q.bx = q.ax; q.by = q.ay; q.bz = q.az;
    GuideAnswer a;
    g_wave_guide(net, feat, q, a);
    if (!valid) return;
    if (mode == 1) {
        pdf_out[s] = a.pdf_a;
    } else {
        dirs_out[3 * (size_t)s] = a.dx;
        dirs_out[3 * (size_t)s + 1] = a.dy;
        dirs_out[3 * (size_t)s + 2] = a.dz;
        pdf_out[s] = a.pdf_b;
    }
}

// ----------------------------------------------------------- training ----
// d(-scale * w * log q)/d outputs, same formulas as the oracle, split over
// the four consecutive lanes of a sample: lane
// part (0..3) takes lobes part and part + 4 and writes their 8 outputs to
// g8 (lobe part, then lobe part + 4); the mixture density q is summed over
// the four lanes.  The mixture weights use FP32 exponentials and an FP64
// normalisation.  All four lanes must call it.
__device__ inline void g_output_grad_part(const float* o, int part, float wx, float wy, float wz, float weight,
                                          float* g8) {
    float mx = o[0];
#pragma unroll
    for (int i = 1; i < KP_G_LOBES; ++i) mx = fmaxf(mx, o[4 * i]);
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < KP_G_LOBES; ++i) sum += (double)expf(o[4 * i] - mx);
    const double inv_sum = 1.0 / sum;
    float v[2], pi[2], k[2], c[2], mxv[2], myv[2], mzv[2];
    float qp = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float* o4 = o + 4 * (part + 4 * h);
        pi[h] = (float)((double)expf(o4[0] - mx) * inv_sum);
        const float kk = fminf(g_softplus(o4[1]) + 0.01f, 1.0e4f);
        const float norm = kk / (2.0f * 3.14159265358979f * (-expm1f(-2.0f * kk)));
        const float z = tanhf(o4[2]);
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sp, cp;
        sincosf(3.14159265358979f * tanhf(o4[3]), &sp, &cp);
        mxv[h] = rr * cp;
        myv[h] = z;
        mzv[h] = rr * sp;
        k[h] = kk;
        c[h] = mxv[h] * wx + myv[h] * wy + mzv[h] * wz;
        v[h] = norm * expf(kk * (c[h] - 1.0f));
        qp += pi[h] * v[h];
    }
    qp += __shfl_xor(qp, 1);
    qp += __shfl_xor(qp, 2);
    const float q = fmaxf(qp, 1e-30f);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float* o4 = o + 4 * (part + 4 * h);
        float* g = g8 + 4 * h;
        const float kk = k[h];
        g[0] = -weight * pi[h] * (v[h] - q) / q;
        const float dlogc = 1.0f / kk - 2.0f / expm1f(2.0f * kk);
        const float dv_dk = v[h] * (dlogc + c[h] - 1.0f);
        const float b = o4[1];
        const bool clamped = g_softplus(b) + 0.01f > 1.0e4f;
        const float sig = 1.0f / (1.0f + expf(-b));
        g[1] = clamped ? 0.0f : -weight * pi[h] * dv_dk / q * sig;
        const float sc = -weight * pi[h] * v[h] * kk / q;
        const float z = myv[h];
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        const float td = tanhf(o4[3]);
        const float dphi = 3.14159265358979f * (1.0f - td * td);
        const float cp = rr > 0.0f ? mxv[h] / rr : 1.0f, sp = rr > 0.0f ? mzv[h] / rr : 0.0f;
        g[2] = sc * (-z * rr * cp * wx + (1.0f - z * z) * wy - z * rr * sp * wz);
        g[3] = sc * (-rr * sp * dphi * wx + rr * cp * dphi * wz);
    }
}

// Weight-gradient tiles (16 x 16) of the network in a fixed enumeration:
// layer 0 has 3 k-tiles (48 inputs) x 4 n-tiles, layers 1-3 4 x 4, layer 4
// 4 x 2; 68 tiles, tile id % 4 = the wave that owns it, id / 4 = its slot in
// that wave's accumulators.
__host__ __device__ constexpr int g_ktiles(int l) { return l == 0 ? 3 : 4; }
__host__ __device__ constexpr int g_tile_base(int l) {
    return l == 0 ? 0 : g_tile_base(l - 1) + g_ktiles(l - 1) * (g_dim_out(l - 1) / 16);
}
#define KP_G_TILES 68
#define KP_G_WTILES 17

// Eight consecutive samples s0 .. s0 + 7 of column f of a sample-major bf16
// block (row stride KP_G_ROW words), packed two per word in sample order: the
// matrix-core operand with samples along k, gathered by 16-bit LDS reads.
__device__ __forceinline__ u32x4 g_column8(const uint32_t* rows, int s0, int f) {
    const uint16_t* p = reinterpret_cast<const uint16_t*>(rows + s0 * KP_G_ROW) + f;
    constexpr int st = 2 * KP_G_ROW;   // row stride in halves
    u32x4 v;
    v[0] = (uint32_t)p[0] | ((uint32_t)p[st] << 16);
    v[1] = (uint32_t)p[2 * st] | ((uint32_t)p[3 * st] << 16);
    v[2] = (uint32_t)p[4 * st] | ((uint32_t)p[5 * st] << 16);
    v[3] = (uint32_t)p[6 * st] | ((uint32_t)p[7 * st] << 16);
    return v;
}

// dW[k][n] += sum over the group's 64 samples of X[s][k] dY[s][n]: A = X^T
// (16 features x 32 samples), B = dY (32 samples x 16 outputs), both gathered
// from the sample-major blocks (no transposed copies: they would double the
// block's LDS and keep it at one block per CU).
template <int L>
__device__ __forceinline__ void g_dw_tiles(const uint32_t* X, const uint32_t* dY, int wave, int lane,
                                           f32x4* acc) {
    constexpr int ntn = g_dim_out(L) / 16;
    const int g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int kt = 0; kt < g_ktiles(L); ++kt)
#pragma unroll
        for (int nt = 0; nt < ntn; ++nt) {
            const int id = g_tile_base(L) + kt * ntn + nt;
            if ((id & 3) != wave) continue;
            f32x4 a = acc[id >> 2];
#pragma unroll
            for (int sc = 0; sc < 2; ++sc) {
                const u32x4 av = g_column8(X, 32 * sc + 8 * g, 16 * kt + r);
                const u32x4 bv = g_column8(dY, 32 * sc + 8 * g, 16 * nt + r);
                a = mfma16(av, bv, a);
            }
            acc[id >> 2] = a;
        }
}

template <int L>
__device__ __forceinline__ void g_dw_store(float* part, int wave, int lane, const f32x4* acc) {
    constexpr int ntn = g_dim_out(L) / 16, dout = g_dim_out(L);
    const int g = lane >> 4, r = lane & 15;
    float* gW = part + g_layer_offset(L);
#pragma unroll
    for (int kt = 0; kt < g_ktiles(L); ++kt)
#pragma unroll
        for (int nt = 0; nt < ntn; ++nt) {
            const int id = g_tile_base(L) + kt * ntn + nt;
            if ((id & 3) != wave) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) gW[(16 * kt + 4 * g + j) * dout + 16 * nt + r] = acc[id >> 2][j];
        }
}

// Back-propagation through layer L for one 64-sample group: accumulate its
// weight and bias gradients, then (L > 0) write dX = (W dY) masked by the
// ReLU of the layer input, sample-major, for this wave's 16 samples.
template <int L>
__device__ __forceinline__ void g_backward_layer(const GuideNet& net, uint32_t (*act)[64 * KP_G_ROW],
                                                 const uint32_t* dY, uint32_t* dX, int tid, int wave, int lane,
                                                 f32x4* acc, float* bacc) {
    constexpr int dout = g_dim_out(L);
    g_dw_tiles<L>(act[L], dY, wave, lane, acc);
    if (tid < dout) {
        // bias gradient: column tid of dY, summed in sample order
        float sb = bacc[L];
        const uint16_t* col = reinterpret_cast<const uint16_t*>(dY) + tid;
#pragma unroll 8
        for (int s = 0; s < 64; ++s) sb += __uint_as_float((uint32_t)col[s * 2 * KP_G_ROW] << 16);
        bacc[L] = sb;
    }
    if constexpr (L > 0) {
        const int g = lane >> 4, r = lane & 15;
        constexpr int nch = dout / 32;
        const uint32_t* dyrow = dY + (16 * wave + r) * KP_G_ROW;
        const uint32_t* xrow = act[L] + (16 * wave + r) * KP_G_ROW;
        uint32_t* dxrow = dX + (16 * wave + r) * KP_G_ROW;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int c = 0; c < nch; ++c) {
                const u32x4 b = *(const u32x4*)(dyrow + 16 * c + 4 * g);
                d = mfma16(net.wb[(g_wb_base(L) + kt * nch + c) * 64 + lane], b, d);
            }
            const int k0 = 16 * kt + 4 * g;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t xw = xrow[(k0 + j) >> 1];
                const float xv = ((k0 + j) & 1) ? bf_hi(xw) : bf_lo(xw);
                v[j] = xv > 0.0f ? d[j] : 0.0f;
            }
            dxrow[k0 / 2] = pack_bf2(v[0], v[1]);
            dxrow[k0 / 2 + 1] = pack_bf2(v[2], v[3]);
        }
    }
}

// Persistent training blocks: block b runs the 64-sample groups b, b + grid,
// ... (4 waves x 16 samples each), keeps its weight-gradient tiles in the
// matrix-core accumulators and its bias gradients in registers across groups,
// and writes one partial gradient (canonical layout) to partial[b].
__global__ __launch_bounds__(256, 2) void k_guide_train(GuideNet net, const float* __restrict__ in6,
                                                        const float* __restrict__ dirs,
                                                        const float* __restrict__ weights,
                                                        const uint32_t* n_dev, const float* scale_dev,
                                                        float* __restrict__ partial) {
    // 63 KB: two blocks per CU
    __shared__ __align__(16) uint32_t act[KP_G_LAYERS][64 * KP_G_ROW];    // layer inputs, sample-major (bf16)
    __shared__ __align__(16) uint32_t dyb[2][64 * KP_G_ROW];              // gradients, sample-major
    // network outputs (f32): read by the output gradient, before the first
    // backward layer writes dyb[1]
    float* outs = reinterpret_cast<float*>(dyb[1]);
    const uint32_t n = *n_dev;
    const uint32_t groups = (n + 63) / 64;
    const int tid = threadIdx.x, wave = tid >> 6, lane = __lane_id();
    const float scale = *scale_dev;
    f32x4 acc[KP_G_WTILES];
#pragma unroll
    for (int i = 0; i < KP_G_WTILES; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bacc[KP_G_LAYERS] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};

    for (uint32_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const uint32_t base = grp * 64;
        {
            // four lanes per sample: lanes 0-2 encode position x / y / z (6
            // frequencies each), lane 3 the normal (2 frequencies per axis);
            // the same features as g_write_features
            const int smp = tid >> 2, part = tid & 3;
            const bool valid = base + smp < n;
            const float* src = in6 + 6 * (size_t)(valid ? base + smp : 0);
            uint32_t wd[6];
            if (part < 3) {
                float sn[6], cs[6];
                g_octaves<6>(src[part], sn, cs);
#pragma unroll
                for (int k = 0; k < 6; ++k) wd[k] = valid ? pack_bf2(sn[k], cs[k]) : 0u;
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float sn[2], cs[2];
                    g_octaves<2>(src[3 + c], sn, cs);
#pragma unroll
                    for (int k = 0; k < 2; ++k) wd[2 * c + k] = valid ? pack_bf2(sn[k], cs[k]) : 0u;
                }
            }
            uint32_t* row = act[0] + smp * KP_G_ROW;
#pragma unroll
            for (int j = 0; j < 6; ++j) row[6 * part + j] = wd[j];
            row[24 + 2 * part] = 0u;
            row[25 + 2 * part] = 0u;
        }
        __syncthreads();
        {
            uint32_t* a[4] = {act[1] + 16 * wave * KP_G_ROW, act[2] + 16 * wave * KP_G_ROW,
                              act[3] + 16 * wave * KP_G_ROW, act[4] + 16 * wave * KP_G_ROW};
            g_forward_chunk(net, act[0] + 16 * wave * KP_G_ROW, a, outs + 16 * wave * KP_G_ROW, KP_G_ROW);
        }
        __syncthreads();
        {
            const int smp = tid >> 2, part = tid & 3;
            const uint32_t sg = base + smp;
            const bool valid = sg < n;
            const size_t si = valid ? sg : 0;
            float g8[8];
            g_output_grad_part(outs + smp * KP_G_ROW, part, dirs[3 * si], dirs[3 * si + 1], dirs[3 * si + 2],
                               valid ? weights[si] * scale : 0.0f, g8);
            uint32_t* row = dyb[0] + smp * KP_G_ROW;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int lobe = part + 4 * h;
                const float* g = g8 + 4 * h;
                row[2 * lobe] = valid ? pack_bf2(g[0], g[1]) : 0u;
                row[2 * lobe + 1] = valid ? pack_bf2(g[2], g[3]) : 0u;
            }
#pragma unroll
            for (int w = 16 + 4 * part; w < 20 + 4 * part; ++w) row[w] = 0u;
        }
        __syncthreads();
        g_backward_layer<4>(net, act, dyb[0], dyb[1], tid, wave, lane, acc, bacc);
        __syncthreads();
        g_backward_layer<3>(net, act, dyb[1], dyb[0], tid, wave, lane, acc, bacc);
        __syncthreads();
        g_backward_layer<2>(net, act, dyb[0], dyb[1], tid, wave, lane, acc, bacc);
        __syncthreads();
        g_backward_layer<1>(net, act, dyb[1], dyb[0], tid, wave, lane, acc, bacc);
        __syncthreads();
        g_backward_layer<0>(net, act, dyb[0], nullptr, tid, wave, lane, acc, bacc);
        __syncthreads();
    }
    float* part = partial + (size_t)blockIdx.x * KP_G_PARAMS;
    g_dw_store<0>(part, wave, lane, acc);
    g_dw_store<1>(part, wave, lane, acc);
    g_dw_store<2>(part, wave, lane, acc);
    g_dw_store<3>(part, wave, lane, acc);
    g_dw_store<4>(part, wave, lane, acc);
#pragma unroll
    for (int l = 0; l < KP_G_LAYERS; ++l)
        if (tid < g_dim_out(l)) part[g_layer_offset(l) + g_dim_in(l) * g_dim_out(l) + tid] = bacc[l];
}

// First stage of the gradient reduction: seg[s][p] = sum over partials
// s*per .. (s+1)*per - 1 (fixed order).
__global__ __launch_bounds__(256) void k_guide_reduce_seg(const float* __restrict__ partial, uint32_t nparts,
                                                          uint32_t per, float* __restrict__ seg) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= KP_G_PARAMS) return;
    const uint32_t b0 = blockIdx.y * per, b1 = b0 + per < nparts ? b0 + per : nparts;
    float sum = 0.0f;
    for (uint32_t b = b0; b < b1; ++b) sum += partial[(size_t)b * KP_G_PARAMS + p];
    seg[(size_t)blockIdx.y * KP_G_PARAMS + p] = sum;
}

// One block: sums weights[0..n) in a fixed order (FP64), then writes the
// batch count, the loss scale 1 / sum and the number of positive weights.
// Sum and count of the positive record weights (one block); with clamp > 0,
// weights above clamp x (mean positive weight) are first capped in place.
__global__ __launch_bounds__(1024) void k_guide_normalise(float* __restrict__ weights, uint32_t n, float clamp,
                                                          uint32_t* count, float* scale, uint32_t* valid) {
    __shared__ double ssum[1024];
    __shared__ uint32_t scnt[1024];
    const int t = threadIdx.x;
    auto reduce = [&](double acc, uint32_t nv) {
        ssum[t] = acc;
        scnt[t] = nv;
        __syncthreads();
        for (int h = 512; h > 0; h >>= 1) {
            if (t < h) {
                ssum[t] += ssum[t + h];
                scnt[t] += scnt[t + h];
            }
            __syncthreads();
        }
    };
    double acc = 0.0;
    uint32_t nv = 0;
    for (uint32_t i = t; i < n; i += 1024) {
        const float w = weights[i];
        acc += (double)w;
        nv += w > 0.0f ? 1u : 0u;
    }
    reduce(acc, nv);
    const double total = ssum[0];
    const uint32_t nvalid = scnt[0];
    __syncthreads();
    double kept = total;
    if (clamp > 0.0f && nvalid > 0u) {
        const float cap = (float)(clamp * total / (double)nvalid);
        acc = 0.0;
        for (uint32_t i = t; i < n; i += 1024) {
            const float w = fminf(weights[i], cap);
            weights[i] = w;
            acc += (double)w;
        }
        reduce(acc, 0u);
        kept = ssum[0];
    }
    if (t == 0) {
        *count = n;
        *scale = kept > 0.0 ? (float)(1.0 / kept) : 0.0f;
        *valid = nvalid;
    }
}

// grad = sum of the nb segment sums (fixed order); optional Adam step.
__global__ void k_guide_reduce_adam(const float* __restrict__ partial, uint32_t nb, float* grad,
                                    float* params, float* m, float* v, int step, float lr, float beta1,
                                    float beta2, float eps, int apply) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= KP_G_PARAMS) return;
    float gsum = 0.0f;
    for (uint32_t b = 0; b < nb; ++b) gsum += partial[(size_t)b * KP_G_PARAMS + p];
    grad[p] = gsum;
    if (!apply) return;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    const float mm = beta1 * m[p] + (1.0f - beta1) * gsum;
    const float vv = beta2 * v[p] + (1.0f - beta2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

}  // namespace kp

// =================================================================== host ==
namespace kp {

void guide_state_free(kp_ctx* c) {
    GuideState* g = c->guide;
    if (!g) return;
    for (void* p : {(void*)g->params, (void*)g->m, (void*)g->v, (void*)g->grad, (void*)g->wf, (void*)g->wb,
                    (void*)g->partial, (void*)g->in6, (void*)g->dirs, (void*)g->weights, (void*)g->count,
                    (void*)g->scale, (void*)g->rec_aux, (void*)g->rec_valid, (void*)g->seg})
        if (p) (void)hipFree(p);
    delete g;
    c->guide = nullptr;
}

GuideNet guide_net(kp_ctx* c) {
    GuideNet n;
    n.wf = c->guide->wf;
    n.wb = c->guide->wb;
    n.params = c->guide->params;
    n.bmin = c->guide->bmin;
    n.binv = c->guide->binv;
    return n;
}

static int guide_pack(kp_ctx* c) {
    const int total = (KP_G_WF_FRAGS + KP_G_WB_FRAGS) * 64;
    hipLaunchKernelGGL(k_guide_pack, dim3((total + 255) / 256), dim3(256), 0, c->stream, c->guide->params,
                       c->guide->wf, c->guide->wb);
    KP_HIP_TRY(c, hipGetLastError());
    return KP_OK;
}

static int guide_alloc(kp_ctx* c) {
    if (c->guide) return KP_OK;
    GuideState* g = new GuideState();
    c->guide = g;
    const size_t P = KP_G_PARAMS * sizeof(float);
    KP_HIP_TRY(c, hipMalloc((void**)&g->params, P));
    KP_HIP_TRY(c, hipMalloc((void**)&g->m, P));
    KP_HIP_TRY(c, hipMalloc((void**)&g->v, P));
    KP_HIP_TRY(c, hipMalloc((void**)&g->grad, P));
    KP_HIP_TRY(c, hipMalloc((void**)&g->wf, KP_G_WF_FRAGS * 64 * sizeof(u32x4)));
    KP_HIP_TRY(c, hipMalloc((void**)&g->wb, KP_G_WB_FRAGS * 64 * sizeof(u32x4)));
    // persistent training grid: every block resident (two per CU at 63 KB of LDS)
    int per_cu = 0;
    KP_HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_guide_train, 256, 0));
    g->train_blocks = (uint32_t)((c->cus > 0 ? c->cus : 256) * (per_cu > 0 ? per_cu : 1));
    KP_HIP_TRY(c, hipMalloc((void**)&g->partial, (size_t)g->train_blocks * P));
    KP_HIP_TRY(c, hipMalloc((void**)&g->seg, (size_t)KP_G_SEGS * P));
    KP_HIP_TRY(c, hipMalloc((void**)&g->in6, (size_t)g->max_batch * 6 * sizeof(float)));
    KP_HIP_TRY(c, hipMalloc((void**)&g->dirs, (size_t)g->max_batch * 3 * sizeof(float)));
    KP_HIP_TRY(c, hipMalloc((void**)&g->weights, (size_t)g->max_batch * sizeof(float)));
    KP_HIP_TRY(c, hipMalloc((void**)&g->count, 64));
    KP_HIP_TRY(c, hipMalloc((void**)&g->scale, 64));
    KP_HIP_TRY(c, hipMalloc((void**)&g->rec_aux, (size_t)g->max_batch * 2 * sizeof(float4)));
    KP_HIP_TRY(c, hipMalloc((void**)&g->rec_valid, 64));
    return KP_OK;
}

static int guide_set_params(kp_ctx* c, const float* host) {
    GuideState* g = c->guide;
    const size_t P = KP_G_PARAMS * sizeof(float);
    KP_HIP_TRY(c, hipMemcpyAsync(g->params, host, P, hipMemcpyHostToDevice, c->stream));
    KP_HIP_TRY(c, hipMemsetAsync(g->m, 0, P, c->stream));
    KP_HIP_TRY(c, hipMemsetAsync(g->v, 0, P, c->stream));
    g->step = 0;
    g->train_passes = 0;
    int rc = guide_pack(c);
    if (rc) return rc;
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    return KP_OK;
}

// One training step on the batch staged in g->in6/dirs/weights, with the
// sample count and loss scale read from device memory.
int guide_train_device(kp_ctx* c, float lr, bool apply) {
    GuideState* g = c->guide;
    GuideNet net = guide_net(c);
    const uint32_t blocks = g->train_blocks;
    hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(256), 0, c->stream, net, g->in6, g->dirs, g->weights,
                       g->count, g->scale, g->partial);
    KP_HIP_TRY(c, hipGetLastError());
    const uint32_t per = (blocks + KP_G_SEGS - 1) / KP_G_SEGS;
    hipLaunchKernelGGL(k_guide_reduce_seg, dim3((KP_G_PARAMS + 255) / 256, KP_G_SEGS), dim3(256), 0, c->stream,
                       g->partial, blocks, per, g->seg);
    KP_HIP_TRY(c, hipGetLastError());
    if (apply) ++g->step;
    hipLaunchKernelGGL(k_guide_reduce_adam, dim3((KP_G_PARAMS + 255) / 256), dim3(256), 0, c->stream,
                       g->seg, (uint32_t)KP_G_SEGS, g->grad, g->params, g->m, g->v, g->step > 0 ? g->step : 1, lr,
                       g->beta1, g->beta2, g->eps, apply ? 1 : 0);
    KP_HIP_TRY(c, hipGetLastError());
    if (apply) return guide_pack(c);
    return KP_OK;
}

int guide_normalise_records(kp_ctx* c, uint32_t n) {
    GuideState* g = c->guide;
    hipLaunchKernelGGL(k_guide_normalise, dim3(1), dim3(1024), 0, c->stream, g->weights, n, g->weight_clamp,
                       g->count, g->scale, g->rec_valid);
    KP_HIP_TRY(c, hipGetLastError());
    return KP_OK;
}

void guide_set_box(kp_ctx* c, const float* lo, const float* hi) {
    GuideState* g = c->guide;
    float inv[3];
    for (int i = 0; i < 3; ++i) {
        const float ext = hi[i] - lo[i];
        inv[i] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    g->bmin = make_float3(lo[0], lo[1], lo[2]);
    g->binv = make_float3(inv[0], inv[1], inv[2]);
}

static int guide_ready(kp_ctx* c) {
    if (!c->guide) return fail(c, KP_ERR_STATE, "guiding network not initialised (kp_guiding_init)");
    return KP_OK;
}

static int guide_eval_common(kp_ctx* c, const float* in6, uint32_t n, float* out32, const float* dirs,
                             const float* u3, float* dirs_out, float* pdf_out, int mode) {
    int rc = guide_ready(c);
    if (rc) return rc;
    if (n == 0) return KP_OK;
    KP_USE_DEVICE(c);
    float *d_in = nullptr, *d_out = nullptr, *d_dirs = nullptr, *d_u = nullptr, *d_do = nullptr, *d_pdf = nullptr;
    auto cleanup = [&]() {
        for (float* p : {d_in, d_out, d_dirs, d_u, d_do, d_pdf})
            if (p) (void)hipFree(p);
    };
    auto up = [&](float** d, const float* h, size_t count) -> hipError_t {
        hipError_t e = hipMalloc((void**)d, count * sizeof(float));
        if (e == hipSuccess && h) e = hipMemcpy(*d, h, count * sizeof(float), hipMemcpyHostToDevice);
        return e;
    };
    hipError_t e = up(&d_in, in6, 6 * (size_t)n);
    if (e == hipSuccess && mode == 0) e = up(&d_out, nullptr, 32 * (size_t)n);
    if (e == hipSuccess && mode == 1) e = up(&d_dirs, dirs, 3 * (size_t)n);
    if (e == hipSuccess && mode == 2) e = up(&d_u, u3, 3 * (size_t)n);
    if (e == hipSuccess && mode == 2) e = up(&d_do, nullptr, 3 * (size_t)n);
    if (e == hipSuccess && mode >= 1) e = up(&d_pdf, nullptr, (size_t)n);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_guide_eval, dim3((n + 255) / 256), dim3(256), 0, c->stream, guide_net(c), d_in, n,
                           d_out, d_dirs, d_u, d_do, d_pdf, mode);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && mode == 0) e = hipMemcpy(out32, d_out, 32 * (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && mode == 2) e = hipMemcpy(dirs_out, d_do, 3 * (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && mode >= 1) e = hipMemcpy(pdf_out, d_pdf, (size_t)n * 4, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return hip_fail(c, e, "guiding evaluation");
    return KP_OK;
}

}  // namespace kp

extern "C" {

int kp_guiding_init(kp_ctx* c, uint64_t seed) {
    if (!c) return KP_ERR_INVALID;
    KP_USE_DEVICE(c);
    int rc = kp::guide_alloc(c);
    if (rc) return rc;
    // Glorot-uniform weights from a PCG32 stream, zero biases; the output
    // layer starts 10x smaller so the first mixture is nearly uniform.
    std::vector<float> P(KP_G_PARAMS, 0.0f);
    uint64_t state = 0, inc = (seed << 1u) | 1u;
    auto next = [&]() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    };
    next();
    state += 0x853c49e6748fea9bULL;
    next();
    for (int l = 0; l < KP_G_LAYERS; ++l) {
        const int din = kp::g_dim_in(l), dout = kp::g_dim_out(l);
        const float lim = std::sqrt(6.0f / (float)(din + dout)) * (l == KP_G_LAYERS - 1 ? 0.1f : 1.0f);
        float* W = P.data() + kp::g_layer_offset(l);
        for (int i = 0; i < din * dout; ++i) W[i] = lim * (2.0f * ((float)(next() >> 8) * (1.0f / 16777216.0f)) - 1.0f);
    }
    return kp::guide_set_params(c, P.data());
}

int kp_guiding_set_params(kp_ctx* c, const float* params, uint32_t n) {
    if (!c || !params) return KP_ERR_INVALID;
    if (n != KP_G_PARAMS) return kp::fail(c, KP_ERR_INVALID, "guiding parameter count must be 17696");
    KP_USE_DEVICE(c);
    int rc = kp::guide_alloc(c);
    if (rc) return rc;
    return kp::guide_set_params(c, params);
}

int kp_guiding_get_params(kp_ctx* c, float* params, uint32_t n) {
    if (!c || !params) return KP_ERR_INVALID;
    int rc = kp::guide_ready(c);
    if (rc) return rc;
    if (n != KP_G_PARAMS) return kp::fail(c, KP_ERR_INVALID, "guiding parameter count must be 17696");
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    KP_HIP_TRY(c, hipMemcpy(params, c->guide->params, KP_G_PARAMS * 4, hipMemcpyDeviceToHost));
    return KP_OK;
}

// Checkpoint: a 64-byte header {magic, version, parameter count, Adam step,
// training passes, beta1, beta2, eps, box origin[3], box inverse extent[3],
// reserved}, then the FP32 masters, the first and the second Adam moments.
namespace {
constexpr uint32_t kCkMagic = 0x4347504Bu;   // "KPGC"
constexpr uint32_t kCkVersion = 1u;
struct CkHeader {
    uint32_t magic, version, n_params;
    int32_t step;
    uint32_t train_passes;
    float beta1, beta2, eps;
    float bmin[3], binv[3];
    uint32_t reserved[2];
};
static_assert(sizeof(CkHeader) == 64, "checkpoint header");
}  // namespace

int kp_guiding_save(kp_ctx* c, void* buf, uint64_t bytes, uint64_t* written) {
    if (!c) return KP_ERR_INVALID;
    if (written) *written = KP_GUIDE_CHECKPOINT_BYTES;
    if (!buf) return KP_OK;   // size query
    if (bytes < KP_GUIDE_CHECKPOINT_BYTES)
        return kp::fail(c, KP_ERR_INVALID, "kp_guiding_save: buffer smaller than KP_GUIDE_CHECKPOINT_BYTES");
    int rc = kp::guide_ready(c);
    if (rc) return rc;
    KP_USE_DEVICE(c);
    GuideState* g = c->guide;
    CkHeader h{};
    h.magic = kCkMagic;
    h.version = kCkVersion;
    h.n_params = KP_G_PARAMS;
    h.step = g->step;
    h.train_passes = g->train_passes;
    h.beta1 = g->beta1;
    h.beta2 = g->beta2;
    h.eps = g->eps;
    h.bmin[0] = g->bmin.x; h.bmin[1] = g->bmin.y; h.bmin[2] = g->bmin.z;
    h.binv[0] = g->binv.x; h.binv[1] = g->binv.y; h.binv[2] = g->binv.z;
    unsigned char* out = (unsigned char*)buf;
    std::memcpy(out, &h, sizeof(h));
    const size_t P = KP_G_PARAMS * sizeof(float);
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    KP_HIP_TRY(c, hipMemcpy(out + sizeof(h), g->params, P, hipMemcpyDeviceToHost));
    KP_HIP_TRY(c, hipMemcpy(out + sizeof(h) + P, g->m, P, hipMemcpyDeviceToHost));
    KP_HIP_TRY(c, hipMemcpy(out + sizeof(h) + 2 * P, g->v, P, hipMemcpyDeviceToHost));
    return KP_OK;
}

int kp_guiding_load(kp_ctx* c, const void* buf, uint64_t bytes) {
    if (!c || !buf) return KP_ERR_INVALID;
    CkHeader h;
    if (bytes < KP_GUIDE_CHECKPOINT_BYTES)
        return kp::fail(c, KP_ERR_INVALID, "kp_guiding_load: buffer smaller than KP_GUIDE_CHECKPOINT_BYTES");
    std::memcpy(&h, buf, sizeof(h));
    if (h.magic != kCkMagic || h.version != kCkVersion || h.n_params != KP_G_PARAMS)
        return kp::fail(c, KP_ERR_INVALID, "kp_guiding_load: not a guiding checkpoint of this network");
    if (h.step < 0) return kp::fail(c, KP_ERR_INVALID, "kp_guiding_load: negative Adam step");
    KP_USE_DEVICE(c);
    int rc = kp::guide_alloc(c);
    if (rc) return rc;
    GuideState* g = c->guide;
    const unsigned char* in = (const unsigned char*)buf;
    const size_t P = KP_G_PARAMS * sizeof(float);
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    KP_HIP_TRY(c, hipMemcpy(g->params, in + sizeof(h), P, hipMemcpyHostToDevice));
    KP_HIP_TRY(c, hipMemcpy(g->m, in + sizeof(h) + P, P, hipMemcpyHostToDevice));
    KP_HIP_TRY(c, hipMemcpy(g->v, in + sizeof(h) + 2 * P, P, hipMemcpyHostToDevice));
    g->step = h.step;
    g->train_passes = h.train_passes;
    g->beta1 = h.beta1;
    g->beta2 = h.beta2;
    g->eps = h.eps;
    g->bmin = make_float3(h.bmin[0], h.bmin[1], h.bmin[2]);
    g->binv = make_float3(h.binv[0], h.binv[1], h.binv[2]);
    if ((rc = kp::guide_pack(c))) return rc;
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    return KP_OK;
}

int kp_guiding_eval(kp_ctx* c, const float* in6, uint32_t n, float* out32) {
    if (!c || (n && (!in6 || !out32))) return KP_ERR_INVALID;
    return kp::guide_eval_common(c, in6, n, out32, nullptr, nullptr, nullptr, nullptr, 0);
}

int kp_guiding_pdf(kp_ctx* c, const float* in6, const float* dirs, uint32_t n, float* pdf) {
    if (!c || (n && (!in6 || !dirs || !pdf))) return KP_ERR_INVALID;
    return kp::guide_eval_common(c, in6, n, nullptr, dirs, nullptr, nullptr, pdf, 1);
}

int kp_guiding_sample(kp_ctx* c, const float* in6, const float* u3, uint32_t n, float* dirs, float* pdf) {
    if (!c || (n && (!in6 || !u3 || !dirs || !pdf))) return KP_ERR_INVALID;
    return kp::guide_eval_common(c, in6, n, nullptr, nullptr, u3, dirs, pdf, 2);
}

int kp_guiding_train_step(kp_ctx* c, const float* in6, const float* dirs, const float* weights, uint32_t n,
                          float lr, int apply, float* grad_out) {
    if (!c || (n && (!in6 || !dirs || !weights))) return KP_ERR_INVALID;
    int rc = kp::guide_ready(c);
    if (rc) return rc;
    GuideState* g = c->guide;
    if (n == 0 || n > g->max_batch) return kp::fail(c, KP_ERR_INVALID, "batch must hold 1..262144 samples");
    KP_USE_DEVICE(c);
    KP_HIP_TRY(c, hipMemcpyAsync(g->in6, in6, (size_t)n * 24, hipMemcpyHostToDevice, c->stream));
    KP_HIP_TRY(c, hipMemcpyAsync(g->dirs, dirs, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
    KP_HIP_TRY(c, hipMemcpyAsync(g->weights, weights, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    const float scale = 1.0f / (float)n;   // mean over the batch
    KP_HIP_TRY(c, hipMemcpyAsync(g->count, &n, 4, hipMemcpyHostToDevice, c->stream));
    KP_HIP_TRY(c, hipMemcpyAsync(g->scale, &scale, 4, hipMemcpyHostToDevice, c->stream));
    if ((rc = kp::guide_train_device(c, lr, apply != 0))) return rc;
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (grad_out) KP_HIP_TRY(c, hipMemcpy(grad_out, g->grad, KP_G_PARAMS * 4, hipMemcpyDeviceToHost));
    return KP_OK;
}

int kp_guiding_configure(kp_ctx* c, const kp_guiding_desc* d) {
    if (!c || !d) return KP_ERR_INVALID;
    int rc = kp::guide_ready(c);
    if (rc) return rc;
    if (!(d->bsdf_fraction > 0.0f && d->bsdf_fraction <= 1.0f))
        return kp::fail(c, KP_ERR_INVALID, "kp_guiding_configure: bsdf_fraction must be in (0, 1]");
    if (!(d->min_alpha >= 0.0f) || !(d->lr >= 0.0f) || !(d->weight_clamp >= 0.0f))
        return kp::fail(c, KP_ERR_INVALID, "kp_guiding_configure: min_alpha, lr and weight_clamp must be >= 0");
    if (d->max_records == 0 || d->max_records > c->guide->max_batch || d->pass_paths == 0)
        return kp::fail(c, KP_ERR_INVALID, "kp_guiding_configure: max_records must be 1..262144, pass_paths >= 1");
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(d->bmin[i]) || !std::isfinite(d->bmax[i]) || d->bmax[i] < d->bmin[i])
            return kp::fail(c, KP_ERR_INVALID, "kp_guiding_configure: bad box");
    GuideState* g = c->guide;
    g->bsdf_fraction = d->bsdf_fraction;
    g->min_alpha = d->min_alpha;
    g->weight_clamp = d->weight_clamp;
    g->relative_weights = d->relative_weights != 0;
    g->lr = d->lr;
    g->train_steps = d->train_steps;
    g->guide_after = d->guide_after;
    g->max_records = d->max_records;
    g->pass_paths = d->pass_paths;
    g->max_train_passes = d->train_passes;
    g->late_train_steps = d->late_train_steps;
    g->max_guided_depth = d->max_guided_depth;
    g->auto_box = d->bmin[0] == d->bmax[0] && d->bmin[1] == d->bmax[1] && d->bmin[2] == d->bmax[2];
    for (int i = 0; i < 3; ++i) {
        g->box_min[i] = d->bmin[i];
        g->box_max[i] = d->bmax[i];
    }
    if (!g->auto_box) kp::guide_set_box(c, d->bmin, d->bmax);
    return KP_OK;
}

}  // extern "C"

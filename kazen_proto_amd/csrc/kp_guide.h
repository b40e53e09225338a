// Path-guiding network: device layout shared by the guiding kernels and the
// guided shade kernel.
//
// No reference counterpart (kazen_proto has no path guiding).  Network:
// frequency-encoded position (3 -> 36) and normal (3 -> 12), four hidden
// ReLU layers of 64, 32 outputs = 8 vMF lobes x (logit, concentration, z, phi).
// Canonical parameter order (FP32 masters): W0[48x64] b0 W1[64x64] b1 W2 b2 W3
// b3 W4[64x32] b4, row-major input-major, 17,696 values.
//
// Matrix-core operands (v_mfma_f32_16x16x32_bf16; lane i, g = i/16, r = i%16):
//   A (16x32): lane holds row r, k = 8g..8g+7;  B (32x16): col r, k = 8g..8g+7;
//   D (16x16 f32): lane holds rows 4g..4g+3 of column r.
// Forward layers compute Y^T = W^T X^T, so A = a 16-neuron x 32-input slice of
// W^T (packed per lane, `wf`), B = 32 inputs of 16 samples (from LDS), and the
// result is 4 neurons x 1 sample per lane.  Back-propagation computes
// dX^T = W dY^T with A = a 16-input x 32-output slice of W (`wb`).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kp {

#define KP_G_LAYERS 5
#define KP_G_IN 48
#define KP_G_HID 64
#define KP_G_OUT 32
#define KP_G_LOBES 8
#define KP_G_PARAMS 17696
// bf16 fragments: forward 18 (n-tile, k-chunk) pairs, backward 28 pairs,
// 64 lanes x 16 bytes each
#define KP_G_WF_FRAGS 36
#define KP_G_WB_FRAGS 28

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int g_dim_in(int l) { return l == 0 ? KP_G_IN : KP_G_HID; }
__host__ __device__ constexpr int g_dim_out(int l) { return l == KP_G_LAYERS - 1 ? KP_G_OUT : KP_G_HID; }
__host__ __device__ constexpr int g_layer_offset(int l) {
    return l == 0 ? 0 : g_layer_offset(l - 1) + g_dim_in(l - 1) * g_dim_out(l - 1) + g_dim_out(l - 1);
}
// first forward fragment of layer l: layers 0-3 have 4 n-tiles x 2 k-chunks
__host__ __device__ constexpr int g_wf_base(int l) { return l * 8; }
// first backward fragment of layer l (l >= 1): 4 k-tiles x (out/32) chunks
__host__ __device__ constexpr int g_wb_base(int l) { return (l - 1) * 8; }

struct GuideNet {
    const u32x4* wf;       // [KP_G_WF_FRAGS][64]
    const u32x4* wb;       // [KP_G_WB_FRAGS][64]
    const float* params;   // canonical FP32 masters (biases read from here)
    float3 bmin, binv;     // scene box: position -> [0,1]^3
};

__device__ __forceinline__ uint32_t f2bf_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return u >> 16;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ uint32_t pack_bf2(float lo, float hi) {
    return f2bf_bits(lo) | (f2bf_bits(hi) << 16);
}
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

__device__ __forceinline__ f32x4 mfma16(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Frequency encoding of one sample into 48 features (see the file comment).
__device__ __forceinline__ void g_encode(float px, float py, float pz, float nx, float ny, float nz,
                                         float* f) {
    const float pc[3] = {px, py, pz};
    const float nc[3] = {nx, ny, nz};
    int o = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            float s, co;
            sincospif(pc[c] * (float)(1 << k), &s, &co);
            f[o++] = s;
            f[o++] = co;
        }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            float s, co;
            sincospif(nc[c] * (float)(1 << k), &s, &co);
            f[o++] = s;
            f[o++] = co;
        }
}

// LDS row stride (in 32-bit words) of a 64-wide bf16 activation row: 32 words
// + 4 padding words, which keeps 16-byte fragment loads aligned and spreads
// the 16 rows a fragment reads over the banks (at most 2-way conflicts).
#define KP_G_ROW 36

// One 16-sample chunk through the network.  `xin` holds 16 rows of 64 bf16
// input features (row stride KP_G_ROW words); hidden layer l (0..3) writes its
// bf16 activations to act[l] (same shape); the 32 outputs of sample r are
// written to out[r*out_stride + o] (f32).  Every lane of the wave must call it.
__device__ inline void g_forward_chunk(const GuideNet& net, const uint32_t* xin, uint32_t* const* act,
                                       float* out, int out_stride) {
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
    const uint32_t* src = xin;
#pragma unroll
    for (int l = 0; l < KP_G_LAYERS; ++l) {
        const int ntiles = g_dim_out(l) / 16;
        const float* bias = net.params + g_layer_offset(l) + g_dim_in(l) * g_dim_out(l);
        // B fragments of this sample (column r): inputs 32c + 8g .. +7
        const u32x4 b0 = *(const u32x4*)(src + r * KP_G_ROW + 4 * g);
        const u32x4 b1 = *(const u32x4*)(src + r * KP_G_ROW + 16 + 4 * g);
        uint32_t* dst = l < KP_G_LAYERS - 1 ? act[l] : nullptr;
        for (int t = 0; t < ntiles; ++t) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma16(net.wf[(g_wf_base(l) + 2 * t) * 64 + lane], b0, acc);
            acc = mfma16(net.wf[(g_wf_base(l) + 2 * t + 1) * 64 + lane], b1, acc);
            const int n0 = 16 * t + 4 * g;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = acc[j] + bias[n0 + j];
            if (l < KP_G_LAYERS - 1) {
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = fmaxf(v[j], 0.0f);
                dst[r * KP_G_ROW + n0 / 2] = pack_bf2(v[0], v[1]);
                dst[r * KP_G_ROW + n0 / 2 + 1] = pack_bf2(v[2], v[3]);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) out[r * out_stride + n0 + j] = v[j];
            }
        }
        src = dst;
    }
}

// ------------------------------------------------------------ mixture ----
struct Lobe {
    float pi, kappa, norm;
    float mx, my, mz;
};

__device__ __forceinline__ float g_softplus(float x) { return x > 20.0f ? x : log1pf(expf(x)); }

// Mixture from the 32 outputs of a sample; the softmax normalisation is done
// in FP64.
__device__ inline void g_mixture(const float* o, Lobe* L) {
    float mx = o[0];
#pragma unroll
    for (int i = 1; i < KP_G_LOBES; ++i) mx = fmaxf(mx, o[4 * i]);
    double e[KP_G_LOBES], sum = 0.0;
#pragma unroll
    for (int i = 0; i < KP_G_LOBES; ++i) {
        e[i] = exp((double)(o[4 * i] - mx));
        sum += e[i];
    }
#pragma unroll
    for (int i = 0; i < KP_G_LOBES; ++i) {
        L[i].pi = (float)(e[i] / sum);
        const float k = fminf(g_softplus(o[4 * i + 1]) + 0.01f, 1.0e4f);
        L[i].kappa = k;
        L[i].norm = k / (2.0f * 3.14159265358979f * (-expm1f(-2.0f * k)));
        const float z = tanhf(o[4 * i + 2]);
        const float phi = 3.14159265358979f * tanhf(o[4 * i + 3]);
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sp, cp;
        sincosf(phi, &sp, &cp);
        L[i].mx = rr * cp;
        L[i].my = z;
        L[i].mz = rr * sp;
    }
}

__device__ __forceinline__ float g_pdf(const Lobe* L, float wx, float wy, float wz) {
    float q = 0.0f;
#pragma unroll
    for (int i = 0; i < KP_G_LOBES; ++i) {
        const float c = L[i].mx * wx + L[i].my * wy + L[i].mz * wz;
        q += L[i].pi * L[i].norm * expf(L[i].kappa * (c - 1.0f));
    }
    return q;
}

// Lobe by inverse CDF of pi, then the vMF sample (W. Jakob 2012).
__device__ inline void g_sample(const Lobe* L, float u0, float u1, float u2, float& wx, float& wy,
                                float& wz) {
    int lobe = KP_G_LOBES - 1;
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < KP_G_LOBES; ++i) {
        acc += L[i].pi;
        if (u0 < acc && lobe == KP_G_LOBES - 1 && i < KP_G_LOBES - 1) lobe = i;
    }
    // (the loop keeps the first i with u0 < acc; lobe 7 otherwise)
    Lobe l = L[0];
#pragma unroll
    for (int i = 1; i < KP_G_LOBES; ++i)
        if (i == lobe) l = L[i];
    const float k = l.kappa;
    const float w = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * k)) / k;
    const float sr = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    float sp, cp;
    sincosf(2.0f * 3.14159265358979f * u2, &sp, &cp);
    const float lx = sr * cp, ly = sr * sp;
    float tx, ty, tz;
    if (fabsf(l.mx) > fabsf(l.my)) {
        const float inv = 1.0f / sqrtf(l.mx * l.mx + l.mz * l.mz);
        tx = l.mz * inv; ty = 0.0f; tz = -l.mx * inv;
    } else {
        const float inv = 1.0f / sqrtf(l.my * l.my + l.mz * l.mz);
        tx = 0.0f; ty = l.mz * inv; tz = -l.my * inv;
    }
    const float bx = ty * l.mz - tz * l.my, by = tz * l.mx - tx * l.mz, bz = tx * l.my - ty * l.mx;
    float dx = lx * tx + ly * bx + w * l.mx;
    float dy = lx * ty + ly * by + w * l.my;
    float dz = lx * tz + ly * bz + w * l.mz;
    const float len = sqrtf(dx * dx + dy * dy + dz * dz);
    wx = dx / len;
    wy = dy / len;
    wz = dz / len;
}

}  // namespace kp

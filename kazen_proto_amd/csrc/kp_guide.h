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

// FP32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32)
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf2(float lo, float hi) {
    const bf16x2 v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ uint32_t f2bf_bits(float f) { return pack_bf2(f, 0.0f) & 0xffffu; }
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

__device__ __forceinline__ f32x4 mfma16(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// sin / cos of pi x 2^k for k = 0 .. N-1: k = 0 from sincospif, each higher
// octave by the double-angle formulas sin 2a = 2 sin a cos a and
// cos 2a = (cos a - sin a)(cos a + sin a).  The absolute error about doubles
// per octave (about 2e-6 at k = 5), far below the bf16 spacing (2^-8 near 1)
// the features are rounded to; 6 sincospif per sample instead of 24.
template <int N>
__device__ __forceinline__ void g_octaves(float x, float* s, float* c) {
    sincospif(x, &s[0], &c[0]);
#pragma unroll
    for (int k = 1; k < N; ++k) {
        s[k] = 2.0f * s[k - 1] * c[k - 1];
        c[k] = (c[k - 1] - s[k - 1]) * (c[k - 1] + s[k - 1]);
    }
}

// Frequency encoding of one sample into 48 features (see the file comment).
__device__ __forceinline__ void g_encode(float px, float py, float pz, float nx, float ny, float nz,
                                         float* f) {
    const float pc[3] = {px, py, pz};
    const float nc[3] = {nx, ny, nz};
    int o = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float s[6], co[6];
        g_octaves<6>(pc[c], s, co);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            f[o++] = s[k];
            f[o++] = co[k];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float s[2], co[2];
        g_octaves<2>(nc[c], s, co);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            f[o++] = s[k];
            f[o++] = co[k];
        }
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
                                       float* out, int out_stride, float* last8 = nullptr) {
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
    const uint32_t* src = xin;
    // Launder the fragment pointer so the weight loads are not hoisted out of
    // a caller's chunk loop (they would pin 144 registers per lane).
    const u32x4* wf = net.wf;
    asm volatile("" : "+s"(wf));
#pragma unroll
    for (int l = 0; l < KP_G_LAYERS; ++l) {
        const int ntiles = g_dim_out(l) / 16;
        const float* bias = net.params + g_layer_offset(l) + g_dim_in(l) * g_dim_out(l);
        // B fragments of this sample (column r): inputs 32c + 8g .. +7
        const u32x4 b0 = *(const u32x4*)(src + r * KP_G_ROW + 4 * g);
        const u32x4 b1 = *(const u32x4*)(src + r * KP_G_ROW + 16 + 4 * g);
        uint32_t* dst = l < KP_G_LAYERS - 1 ? act[l] : nullptr;
#pragma unroll
        for (int t = 0; t < ntiles; ++t) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma16(wf[(g_wf_base(l) + 2 * t) * 64 + lane], b0, acc);
            acc = mfma16(wf[(g_wf_base(l) + 2 * t + 1) * 64 + lane], b1, acc);
            const int n0 = 16 * t + 4 * g;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = acc[j] + bias[n0 + j];
            if (l < KP_G_LAYERS - 1) {
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = fmaxf(v[j], 0.0f);
                dst[r * KP_G_ROW + n0 / 2] = pack_bf2(v[0], v[1]);
                dst[r * KP_G_ROW + n0 / 2 + 1] = pack_bf2(v[2], v[3]);
            } else if (out) {
#pragma unroll
                for (int j = 0; j < 4; ++j) out[r * out_stride + n0 + j] = v[j];
            } else {
                // outputs 16t + 4g + j: this lane holds lobes g (t = 0) and g + 4 (t = 1)
#pragma unroll
                for (int j = 0; j < 4; ++j) last8[4 * t + j] = v[j];
            }
        }
        // the next layer reads rows written by other lanes of the wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        src = dst;
    }
}

// Rows of LDS (KP_G_ROW words each) one wave needs for g_wave_guide.
#define KP_G_WAVE_ROWS 64

// Forward pass of a wave's 64 samples in place, layer-major: each layer's
// weight fragments and biases are read from memory once and applied to the
// four 16-sample chunks (the per-chunk order reads them four times and waits
// on the memory system 20 times instead of 5).  x holds the 64 bf16 input
// rows (KP_G_ROW words each); a chunk's rows are read into registers before
// its outputs overwrite them.  On return row s holds sample s's 32 outputs
// as f32 (words 0..31).  Every lane of the wave must call it.
__device__ inline void g_wave_forward(const GuideNet& net, uint32_t* x) {
    const int lane = __lane_id();
    const int g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int l = 0; l < KP_G_LAYERS; ++l) {
        const int ntiles = g_dim_out(l) / 16;
        const float* bias = net.params + g_layer_offset(l) + g_dim_in(l) * g_dim_out(l);
        // launder the fragment pointer per layer: loads of later layers are
        // not hoisted here (they would pin 32 registers per lane and layer)
        const u32x4* wf = net.wf;
        asm volatile("" : "+s"(wf));
        u32x4 w[8];
#pragma unroll
        for (int i = 0; i < 2 * ntiles; ++i) w[i] = wf[(g_wf_base(l) + i) * 64 + lane];
        float bv[16];
#pragma unroll
        for (int t = 0; t < ntiles; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[4 * t + j] = bias[16 * t + 4 * g + j];
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
            const int row = 16 * c + r;
            const u32x4 b0 = *(const u32x4*)(x + row * KP_G_ROW + 4 * g);
            const u32x4 b1 = *(const u32x4*)(x + row * KP_G_ROW + 16 + 4 * g);
#pragma unroll
            for (int t = 0; t < ntiles; ++t) {
                // the bias is the accumulator's initial value
                f32x4 acc = {bv[4 * t], bv[4 * t + 1], bv[4 * t + 2], bv[4 * t + 3]};
                acc = mfma16(w[2 * t], b0, acc);
                acc = mfma16(w[2 * t + 1], b1, acc);
                const int n0 = 16 * t + 4 * g;
                float v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = acc[j];
                if (l < KP_G_LAYERS - 1) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = fmaxf(v[j], 0.0f);
                    *reinterpret_cast<uint2*>(x + row * KP_G_ROW + n0 / 2) =
                        make_uint2(pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3]));
                } else {
                    *reinterpret_cast<float4*>(x + row * KP_G_ROW + n0) = make_float4(v[0], v[1], v[2], v[3]);
                }
            }
        }
        // the next layer reads rows written by other lanes of the wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ------------------------------------------------------------ mixture ----
// The mixture is evaluated straight from the 32 network outputs `o`, one
// lobe at a time, so no per-lane lobe array is needed.
// Lobe i: pi_i = softmax(o[4i]) (normalised in FP64), kappa_i =
// min(softplus(o[4i+1]) + 0.01, 1e4), mu_i = (r cos phi, z, r sin phi) with
// z = tanh(o[4i+2]), phi = pi tanh(o[4i+3]).

__device__ __forceinline__ float g_softplus(float x) { return x > 20.0f ? x : log1pf(expf(x)); }

// ---------------------------------------------- wave-level guide queries --
// Features of one sample written straight into an LDS row as bf16 pairs
// (sin, cos), columns 48..63 zero.
__device__ __forceinline__ void g_encode_row(uint32_t* row, float px, float py, float pz, float nx, float ny,
                                             float nz) {
    const float pc[3] = {px, py, pz};
    const float nc[3] = {nx, ny, nz};
    uint32_t wd[32];
    int w = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float s[6], co[6];
        g_octaves<6>(pc[c], s, co);
#pragma unroll
        for (int k = 0; k < 6; ++k) wd[w++] = pack_bf2(s[k], co[k]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float s[2], co[2];
        g_octaves<2>(nc[c], s, co);
#pragma unroll
        for (int k = 0; k < 2; ++k) wd[w++] = pack_bf2(s[k], co[k]);
    }
#pragma unroll
    for (; w < 32; ++w) wd[w] = 0u;
    // 16-byte stores: a lane's row is 8 stores, not 32
#pragma unroll
    for (int i = 0; i < 8; ++i)
        reinterpret_cast<u32x4*>(row)[i] = u32x4{wd[4 * i], wd[4 * i + 1], wd[4 * i + 2], wd[4 * i + 3]};
}

// Per-lane query of the guiding distribution at one shading point.
struct GuideQuery {
    float px, py, pz, nx, ny, nz;   // position in [0,1]^3, unit normal
    float ax, ay, az;               // direction A: density wanted (zero = none)
    float bx, by, bz;               // direction B: density wanted, unless sampling
    bool sample;                    // draw a direction (u0: lobe, u1, u2) instead of B
    float u0, u1, u2;
};

struct GuideAnswer {
    float pdf_a, pdf_b;             // densities at A and at B (or at the sample)
    float dx, dy, dz;               // the sampled direction
};

// Hardware transcendentals for the guided shade's mixture queries
// (v_exp_f32, v_log_f32, v_rcp_f32, v_rsq_f32, v_sin_f32 / v_cos_f32 on
// revolutions): about 1e-6 relative, far inside the 2e-3 at which mixture
// densities are compared with the oracle.  Training keeps the libm forms.
__device__ __forceinline__ float g_fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float g_flog(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
__device__ __forceinline__ float g_frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float g_ftanh(float x) {
    const float t = g_fexp(-2.0f * fabsf(x));
    return copysignf((1.0f - t) * g_frcp(1.0f + t), x);
}

// One lobe (raw outputs o4, mixture weight pi) set up for density queries:
// concentration, pi x vMF normalisation, mean direction.
struct LobeQ {
    float k, pn, mx, my, mz;
};
__device__ __forceinline__ LobeQ g_lobe_q(const float* o4, float pi) {
    LobeQ L;
    const float sp_in = o4[1];
    const float soft = sp_in > 20.0f ? sp_in : g_flog(1.0f + g_fexp(sp_in));
    const float k = fminf(soft + 0.01f, 1.0e4f);
    const float norm = k * g_frcp(2.0f * 3.14159265358979f * (1.0f - g_fexp(-2.0f * k)));
    const float z = g_ftanh(o4[2]);
    const float rr = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - z * z));
    // phi = pi tanh(o4[3]): half a revolution per unit
    const float rev = 0.5f * g_ftanh(o4[3]);
    L.k = k;
    L.pn = pi * norm;
    L.mx = rr * __builtin_amdgcn_cosf(rev);
    L.my = z;
    L.mz = rr * __builtin_amdgcn_sinf(rev);
    return L;
}
// Density contribution of the lobe for the unit direction w.
__device__ __forceinline__ float g_lobe_eval(const LobeQ& L, float wx, float wy, float wz) {
    const float c = L.mx * wx + L.my * wy + L.mz * wz;
    return L.pn * g_fexp(L.k * (c - 1.0f));
}

// Answers every lane's query.  The wave's 64 samples go through the network
// together (g_wave_forward), which leaves sample s's 32 outputs in row s;
// then each lane evaluates its own sample's mixture: mixture weights (FP32
// exponentials, FP64 normalisation), the sampled direction if asked (lobe by
// inverse CDF of the weights, then the vMF sample), and the densities at A
// and at B or the sample, one lobe at a time.  `lds` is this wave's
// KP_G_WAVE_ROWS rows (KP_G_ROW words each).  Every lane must call it.
__device__ inline void g_wave_guide(const GuideNet& net, uint32_t* lds, const GuideQuery& q, GuideAnswer& ans) {
    const int lane = __lane_id();
    uint32_t* row = lds + lane * KP_G_ROW;
    g_encode_row(row, q.px, q.py, q.pz, q.nx, q.ny, q.nz);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    g_wave_forward(net, lds);
    float o[KP_G_OUT];
#pragma unroll
    for (int i = 0; i < KP_G_OUT / 4; ++i) {
        const float4 v = reinterpret_cast<const float4*>(row)[i];
        o[4 * i] = v.x;
        o[4 * i + 1] = v.y;
        o[4 * i + 2] = v.z;
        o[4 * i + 3] = v.w;
    }
    float mx = o[0];
#pragma unroll
    for (int i = 1; i < KP_G_LOBES; ++i) mx = fmaxf(mx, o[4 * i]);
    float e[KP_G_LOBES];
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < KP_G_LOBES; ++i) {
        e[i] = g_fexp(o[4 * i] - mx);
        sum += (double)e[i];
    }
    const double inv_sum = 1.0 / sum;
    float pi[KP_G_LOBES];
#pragma unroll
    for (int i = 0; i < KP_G_LOBES; ++i) pi[i] = (float)((double)e[i] * inv_sum);
    float sx = q.bx, sy = q.by, sz = q.bz;
    if (q.sample) {
        int sel = KP_G_LOBES - 1;
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < KP_G_LOBES - 1; ++i) {
            acc += pi[i];
            if (sel == KP_G_LOBES - 1 && q.u0 < acc) sel = i;
        }
        // the selected lobe's raw outputs, from LDS (a dynamic index into o
        // would put o in scratch memory)
        const float4 ov = reinterpret_cast<const float4*>(row)[sel];
        const float o4[4] = {ov.x, ov.y, ov.z, ov.w};
        const LobeQ L = g_lobe_q(o4, 1.0f);
        const float k = L.k, mxx = L.mx, myy = L.my, mzz = L.mz;
        // cosine to the mean (W. Jakob 2012), clamped to -1 where exp(-2k)
        // underflows and u1 = 0
        const float u1 = q.u1, u2 = q.u2;
        const float w = fmaxf(1.0f + g_flog(u1 + (1.0f - u1) * g_fexp(-2.0f * k)) * g_frcp(k), -1.0f);
        const float srr = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - w * w));
        const float lx = srr * __builtin_amdgcn_cosf(u2), ly = srr * __builtin_amdgcn_sinf(u2);
        float tx, ty, tz;
        if (fabsf(mxx) > fabsf(myy)) {
            const float inv = __builtin_amdgcn_rsqf(mxx * mxx + mzz * mzz);
            tx = mzz * inv; ty = 0.0f; tz = -mxx * inv;
        } else {
            const float inv = __builtin_amdgcn_rsqf(myy * myy + mzz * mzz);
            tx = 0.0f; ty = mzz * inv; tz = -myy * inv;
        }
        const float bxx = ty * mzz - tz * myy, byy = tz * mxx - tx * mzz, bzz = tx * myy - ty * mxx;
        const float dx = lx * tx + ly * bxx + w * mxx;
        const float dy = lx * ty + ly * byy + w * myy;
        const float dz = lx * tz + ly * bzz + w * mzz;
        const float il = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
        sx = dx * il;
        sy = dy * il;
        sz = dz * il;
    }
    float pa = 0.0f, pb = 0.0f;
#pragma unroll
    for (int i = 0; i < KP_G_LOBES; ++i) {
        const LobeQ L = g_lobe_q(o + 4 * i, pi[i]);
        pa += g_lobe_eval(L, q.ax, q.ay, q.az);
        pb += g_lobe_eval(L, sx, sy, sz);
    }
    ans.pdf_a = pa;
    ans.pdf_b = pb;
    ans.dx = sx;
    ans.dy = sy;
    ans.dz = sz;
}

}  // namespace kp

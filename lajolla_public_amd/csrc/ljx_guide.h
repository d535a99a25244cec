// Learned path guiding: a small MLP maps (position, normal, outgoing
// direction) at a surface vertex to a mixture of 8 von Mises-Fisher lobes
// over the sphere of incident directions.  The mixture is combined with BSDF
// sampling by one-sample MIS, so the estimator stays unbiased whatever the
// network outputs.
//
// Network: one-blob encoded inputs (position 3 x 32 bins, normal 2 x 16,
// outgoing direction 2 x 16 = 160), 4 hidden layers of 64 ReLU units, 40
// outputs (8 lobes x (3 mean-direction components, log kappa, weight logit)).
// Weights and activations are BF16 operands of v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation; biases and the mixture arithmetic are FP32.
//
// Matrix layout.  Each layer is computed transposed, H_out^T = W^T H_in^T:
// the A operand is a 16 x 32 block of W^T (rows: output neurons, K: input
// neurons), the B operand a 32 x 16 block of H_in^T (K: input neurons,
// columns: 16 samples).  Lane l = j + 16 g of the accumulator holds rows
// 4g..4g+3 of column j, so two 16-row output tiles of one layer are exactly
// the 8 K-elements lane l needs as the next layer's B operand, provided the
// weights are stored with the same K order (kslot_hidden).  Activations never
// move between lanes; the weight fragments are pre-permuted instead.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ljx {
namespace guide {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFeat = 8;        // 7 features in [0, 1] + pad
constexpr int kIn = 160;
constexpr int kHid = 64;
constexpr int kOut = 40;
constexpr int kOutRows = 48;    // output rows padded to 3 tiles of 16
constexpr int kLobes = 8;
constexpr int kLayers = 5;
constexpr int kNumParams = 25384;
constexpr int kFwdFrags = 50;   // 20 + 8 + 8 + 8 + 6
constexpr int kBwdFrags = 32;   // 4 layers x 4 row tiles x 2 K blocks
constexpr int kFragElems = 512; // 64 lanes x 8 bf16
constexpr int kBiasFwd = 304;   // 4 x 64 + 48
constexpr float kLogKappaLo = -3.0f, kLogKappaHi = 8.0f;
constexpr float kTwoPi = 6.283185307179586f;

__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? kIn : kHid; }
__host__ __device__ constexpr int layer_out(int l) { return l == 4 ? kOut : kHid; }
__host__ __device__ constexpr int layer_rows(int l) { return l == 4 ? kOutRows : kHid; }
__host__ __device__ constexpr int layer_tiles(int l) { return layer_rows(l) / 16; }
__host__ __device__ constexpr int layer_kblocks(int l) { return layer_in(l) / 32; }
// canonical parameter layout: per layer W[in][out] row-major, then b[out]
__host__ __device__ constexpr int param_w(int l) {
    return l == 0 ? 0 : param_w(l - 1) + layer_in(l - 1) * layer_out(l - 1) + layer_out(l - 1);
}
__host__ __device__ constexpr int param_b(int l) { return param_w(l) + layer_in(l) * layer_out(l); }
__host__ __device__ constexpr int fwd_frag_base(int l) {
    return l == 0 ? 0 : fwd_frag_base(l - 1) + layer_tiles(l - 1) * layer_kblocks(l - 1);
}
__host__ __device__ constexpr int bias_base(int l) { return l * kHid; }
static_assert(param_b(4) + kOut == kNumParams, "parameter count");
static_assert(fwd_frag_base(4) + 6 == kFwdFrags, "fragment count");

// K order of hidden-layer operands: element e of lane group g in K block kb
// is the neuron held by accumulator tile 2 kb + e / 4, row 4 g + e % 4.
__host__ __device__ constexpr int kslot_hidden(int kb, int g, int e) { return 32 * kb + 16 * (e >> 2) + 4 * g + (e & 3); }

// Output row r (0..47) <-> canonical output o = 5 * lobe + param.  Lane group
// g owns lobes 2g and 2g+1: slots 4t + e (t: tile, e: element) 0..4 are the
// first lobe, 5..9 the second, 10 and 11 are padding.
__host__ __device__ constexpr int out_of_row(int r) {
    return ((4 * (r >> 4) + (r & 3)) >= 10) ? -1
                                            : 5 * (2 * ((r & 15) >> 2) + (4 * (r >> 4) + (r & 3)) / 5) +
                                                  (4 * (r >> 4) + (r & 3)) % 5;
}
__host__ __device__ constexpr int row_of_out(int o) {
    return 16 * ((5 * ((o / 5) & 1) + o % 5) >> 2) + 4 * ((o / 5) >> 1) + ((5 * ((o / 5) & 1) + o % 5) & 3);
}

// One-blob encoding (Mueller et al. 2019): feature s in [0, 1] over nb bins,
// bin b = exp(-((b + 1/2) - s nb)^2 / 2), a Gaussian one bin wide.
// Input order: px 32 | py 32 | pz 32 | n_theta 16 | n_phi 16 | wo_theta 16 | wo_phi 16.
__host__ __device__ inline void input_group(int G, int &feature, int &nbins, int &bin0) {
    // G: group of 8 consecutive inputs (0..19)
    if (G < 12) {
        feature = G >> 2;
        nbins = 32;
        bin0 = 8 * (G & 3);
    } else {
        feature = 3 + ((G - 12) >> 1);
        nbins = 16;
        bin0 = 8 * ((G - 12) & 1);
    }
}

// ------------------------------------------------------------ device side
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 zero4() { return f32x4{0.f, 0.f, 0.f, 0.f}; }

// Bin e of the group: exp(-(e - c)^2 / 2) = 2^(e c L - c^2 L / 2 - e^2 L / 2)
// with L = log2(e): one add to the running exponent and one exp2 per bin.
__device__ __forceinline__ bf16x8 encode8(float s, int nbins, int bin0) {
    constexpr float L = 1.4426950408889634f;
    const float c = s * float(nbins) - float(bin0) - 0.5f;
    const float cl = c * L;
    float q = -0.5f * L * c * c;  // exponent of bin 0
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        r[e] = (__bf16)__builtin_amdgcn_exp2f(q);
        q += cl - L * (float(e) + 0.5f);  // exponent of bin e + 1
    }
    return r;
}

// Forward pass of NS sample tiles (16 samples each) held by one wave.
//   feat:  LDS, kFeat floats per sample, sample 16 s + j at feat[(16 s + j) * kFeat]
//   frag:  LDS, kFwdFrags fragments of kFragElems bf16
//   bias:  LDS, kBiasFwd floats (row order of the accumulator tiles)
// out[s][t][e] (t < 3): output row 16 t + 4 g + e of sample 16 s + j.
// If H is non-null, the ReLU outputs of hidden layer k (0..3) are kept as
// B operands in H[k][s][kb].
template <int NS>
__device__ __forceinline__ void mlp_forward(const float *feat, const bf16x8 *frag, const float *bias,
                                            f32x4 (&out)[NS][3], bf16x8 (*H)[NS][2]) {
    const int lane = __lane_id();
    const int j = lane & 15, g = lane >> 4;
    f32x4 acc[NS][4];
    bf16x8 hb[NS][2];
    // every layer's accumulators start from its bias (rows 16 t + 4 g + e)
    auto bias4 = [&](int l, int t) {
        const float *b = bias + bias_base(l) + 16 * t + 4 * g;
        return f32x4{b[0], b[1], b[2], b[3]};
    };
    // layer 0: one-blob input computed per K block
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const f32x4 b = bias4(0, t);
#pragma unroll
        for (int s = 0; s < NS; s++) acc[s][t] = b;
    }
#pragma unroll
    for (int kb = 0; kb < 5; kb++) {
        int f, nb, b0;
        input_group(4 * kb + g, f, nb, b0);
        bf16x8 xb[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) xb[s] = encode8(feat[(16 * s + j) * kFeat + f], nb, b0);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            bf16x8 a = frag[(fwd_frag_base(0) + t * 5 + kb) * 64 + lane];
#pragma unroll
            for (int s = 0; s < NS; s++) acc[s][t] = mfma(a, xb[s], acc[s][t]);
        }
    }
    // hidden layers 1..3 and the output layer
#pragma unroll
    for (int l = 1; l <= 4; l++) {
        // ReLU of layer l-1 (its bias is in the accumulators), packed as
        // layer l's B operand
#pragma unroll
        for (int s = 0; s < NS; s++) {
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
#pragma unroll
                for (int e = 0; e < 8; e++) hb[s][kb][e] = (__bf16)fmaxf(acc[s][2 * kb + (e >> 2)][e & 3], 0.f);
                if (H) H[l - 1][s][kb] = hb[s][kb];
            }
        }
        const int T = l == 4 ? 3 : 4;
#pragma unroll
        for (int t = 0; t < T; t++) {
            const f32x4 b = bias4(l, t);
#pragma unroll
            for (int s = 0; s < NS; s++) acc[s][t] = b;
        }
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
#pragma unroll
            for (int t = 0; t < T; t++) {
                bf16x8 a = frag[(fwd_frag_base(l) + t * 2 + kb) * 64 + lane];
#pragma unroll
                for (int s = 0; s < NS; s++) acc[s][t] = mfma(a, hb[s][kb], acc[s][t]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int t = 0; t < 3; t++) out[s][t] = acc[s][t];
}

// Raw parameters of the lane's two lobes from its 12 output slots.
struct Lobe {
    float rx, ry, rz;  // mean direction before normalisation
    float rk;          // log kappa before clamping
    float logit;
};
__device__ __forceinline__ void lane_lobes(const f32x4 (&o)[3], Lobe &a, Lobe &b) {
    auto slot = [&](int sl) { return o[sl >> 2][sl & 3]; };
    a = Lobe{slot(0), slot(1), slot(2), slot(3), slot(4)};
    b = Lobe{slot(5), slot(6), slot(7), slot(8), slot(9)};
}

// Reductions over the 4 lanes (j, j+16, j+32, j+48) that hold one sample.
__device__ __forceinline__ float quad_max(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}
__device__ __forceinline__ float quad_sum(float v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

// Evaluated form of one lobe: pdf_j(w) = exp(dot(km, w) + c).
struct LobeEval {
    float mx, my, mz, kappa, logw;  // unit mean, kappa, log mixture weight
    float kmx, kmy, kmz, c;
};
__device__ __forceinline__ float log_vmf_norm(float kappa) {
    // log(kappa / (2 pi (1 - exp(-2 kappa))))
    return __logf(kappa) - __logf(kTwoPi) - __logf(-expm1f(-2.f * kappa));
}
__device__ __forceinline__ LobeEval eval_lobe(const Lobe &L, float log_z) {
    LobeEval e;
    const float il = rsqrtf(fmaxf(L.rx * L.rx + L.ry * L.ry + L.rz * L.rz, 1e-12f));
    e.mx = L.rx * il;
    e.my = L.ry * il;
    e.mz = L.rz * il;
    e.kappa = __expf(fminf(fmaxf(L.rk, kLogKappaLo), kLogKappaHi));
    e.logw = L.logit - log_z;
    e.kmx = e.kappa * e.mx;
    e.kmy = e.kappa * e.my;
    e.kmz = e.kappa * e.mz;
    e.c = e.logw + log_vmf_norm(e.kappa) - e.kappa;
    return e;
}

// Mixture pdf of a stored lobe set (8 float4: kappa * mean, log constant).
__device__ __forceinline__ float mixture_pdf(const float4 *lobes, float wx, float wy, float wz) {
    float p = 0.f;
#pragma unroll
    for (int k = 0; k < kLobes; k++) {
        float4 L = lobes[k];
        p += __expf(L.x * wx + L.y * wy + L.z * wz + L.w);
    }
    return p;
}

// Samples a vMF lobe (Wood 1994 / Jakob 2012 inversion of the cosine CDF).
__device__ __forceinline__ void sample_vmf(float mx, float my, float mz, float kappa, float u1, float u2, float &wx,
                                           float &wy, float &wz) {
    float w = 1.f + __logf(u1 + (1.f - u1) * __expf(-2.f * kappa)) / kappa;
    w = fminf(fmaxf(w, -1.f), 1.f);
    float st = sqrtf(fmaxf(0.f, 1.f - w * w));
    float sp, cp;
    __sincosf(kTwoPi * u2, &sp, &cp);
    // orthonormal basis around the mean (Duff et al. 2017)
    float sg = copysignf(1.f, mz);
    float a = -1.f / (sg + mz);
    float b = mx * my * a;
    float t1x = 1.f + sg * mx * mx * a, t1y = sg * b, t1z = -sg * mx;
    float t2x = b, t2y = sg + my * my * a, t2z = -my;
    wx = w * mx + st * (cp * t1x + sp * t2x);
    wy = w * my + st * (cp * t1y + sp * t2y);
    wz = w * mz + st * (cp * t1z + sp * t2z);
}

// Loads a fragment table (bf16) and a bias table into LDS; all threads of the
// block take part.  Returns after a block barrier.
__device__ __forceinline__ void load_lds(bf16x8 *dst, const bf16x8 *src, int nvec, float *bdst, const float *bsrc,
                                         int nb) {
    for (int i = threadIdx.x; i < nvec; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) bdst[i] = bsrc[i];
    __syncthreads();
}

// Gradient of -w log q(dir) with respect to the lane's two lobes' raw
// outputs, and the sample's loss; every lane of the quad returns the loss.
__device__ __forceinline__ float lobe_grads(const f32x4 (&o)[3], float dx, float dy, float dz, float w,
                                            f32x4 (&grad)[3]) {
    Lobe L[2];
    lane_lobes(o, L[0], L[1]);
    float lm = quad_max(fmaxf(L[0].logit, L[1].logit));
    float zs = quad_sum(__expf(L[0].logit - lm) + __expf(L[1].logit - lm));
    float log_z = lm + __logf(zs);
    float a[2], mw[2], kap[2], len[2], mx[2], my[2], mz[2];
    bool inside[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        float l2 = fmaxf(L[k].rx * L[k].rx + L[k].ry * L[k].ry + L[k].rz * L[k].rz, 1e-12f);
        len[k] = sqrtf(l2);
        mx[k] = L[k].rx / len[k];
        my[k] = L[k].ry / len[k];
        mz[k] = L[k].rz / len[k];
        inside[k] = L[k].rk > kLogKappaLo && L[k].rk < kLogKappaHi;
        kap[k] = __expf(fminf(fmaxf(L[k].rk, kLogKappaLo), kLogKappaHi));
        mw[k] = mx[k] * dx + my[k] * dy + mz[k] * dz;
        a[k] = (L[k].logit - log_z) + log_vmf_norm(kap[k]) + kap[k] * (mw[k] - 1.f);
    }
    float am = quad_max(fmaxf(a[0], a[1]));
    float ss = quad_sum(__expf(a[0] - am) + __expf(a[1] - am));
    float log_q = am + __logf(ss);
    float g[2][5];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        float gam = __expf(a[k] - log_q);
        float pi = __expf(L[k].logit - log_z);
        // d log q / d logit
        g[k][4] = gam - pi;
        // d log q / d rk = gamma (1 - 2 kappa / expm1(2 kappa) + kappa (mu.w - 1))
        float kk = kap[k];
        float dk = 1.f - 2.f * kk / expm1f(2.f * kk) + kk * (mw[k] - 1.f);
        g[k][3] = inside[k] ? gam * dk : 0.f;
        // d log q / d r = gamma kappa (w - (mu.w) mu) / |r|
        float sc = gam * kk / len[k];
        g[k][0] = sc * (dx - mw[k] * mx[k]);
        g[k][1] = sc * (dy - mw[k] * my[k]);
        g[k][2] = sc * (dz - mw[k] * mz[k]);
    }
#pragma unroll
    for (int sl = 0; sl < 12; sl++) {
        float v = sl < 10 ? -w * g[sl / 5][sl % 5] : 0.f;
        grad[sl >> 2][sl & 3] = v;
    }
    return -w * log_q;
}

}  // namespace guide
}  // namespace ljx

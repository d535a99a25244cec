// Guiding network: parameter storage, BF16 fragment packing, the MFMA
// forward / backward / weight-gradient kernels, Adam, and the network's C
// entry points (include/lajolla_amd.h, "learned path guiding").
#include "ljx_guide_net.h"
#include "ljx_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace ljx {
namespace guide {

// Activation tensors of one batch, in groups of 32 samples; inside a group a
// tensor is [row][32 samples] BF16, so that a 16-row x 32-sample block is one
// contiguous 1 KB MFMA operand with the samples along K.
//   0..3  H_k   (ReLU output of hidden layer k, 64 rows)
//   4..7  dZ_k  (loss gradient at hidden layer k's pre-activation, 64 rows)
//   8     dZ_4  (loss gradient at the 48 padded output rows)
__host__ __device__ inline size_t act_offset(int which, int groups) { return size_t(which) * groups * 64 * 32; }
__host__ __device__ inline int act_rows(int which) { return which == 8 ? kOutRows : 64; }
inline size_t act_elems(int groups) { return act_offset(8, groups) + size_t(groups) * kOutRows * 32; }

__device__ __forceinline__ uint16_t bf16_bits(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
// Bits of element e of a BF16 vector, through its 32-bit words (a direct
// bit_cast of the extracted __bf16 element was seen to return element e & ~3).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint16_t bf16_elem_bits(bf16x8 v, int e) {
    u32x4 u = __builtin_bit_cast(u32x4, v);
    return uint16_t(u[e >> 1] >> (16 * (e & 1)));
}

__global__ void k_pack(const float *master, const int *fmap, __bf16 *ffrag, int nf, const int *bmap, __bf16 *bfrag,
                       int nb, const int *biasmap, float *fbias, int nbias) {
    int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf + nb + nbias; i += stride) {
        if (i < nf) {
            int p = fmap[i];
            ffrag[i] = (__bf16)(p >= 0 ? master[p] : 0.f);
        } else if (i < nf + nb) {
            int k = i - nf, p = bmap[k];
            bfrag[k] = (__bf16)(p >= 0 ? master[p] : 0.f);
        } else {
            int k = i - nf - nb, p = biasmap[k];
            fbias[k] = p >= 0 ? master[p] : 0.f;
        }
    }
}

__global__ void k_zero_f32(float *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Raw network outputs for n feature rows (for tests and tools).
__global__ __launch_bounds__(256) void k_forward_raw(const float4 *feats, int n, const bf16x8 *ffrag,
                                                    const float *fbias, float *out) {
    __shared__ bf16x8 sfrag[kFwdFrags * 64];
    __shared__ float sbias[kBiasFwd];
    __shared__ float sfeat[4][64 * kFeat];
    load_lds(sfrag, ffrag, kFwdFrags * 64, sbias, fbias, kBiasFwd);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        int i = base + w * 64 + lane;
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
        if (i < n) {
            a = feats[2 * i];
            b = feats[2 * i + 1];
        }
        float *f = &sfeat[w][lane * kFeat];
        f[0] = a.x, f[1] = a.y, f[2] = a.z, f[3] = a.w, f[4] = b.x, f[5] = b.y, f[6] = b.z, f[7] = b.w;
        __syncthreads();
        f32x4 o[4][3];
        mlp_forward<4>(sfeat[w], sfrag, sbias, o, nullptr);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            int sample = base + w * 64 + 16 * s + j;
            if (sample >= n) continue;
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    int oo = out_of_row(16 * t + 4 * g + e);
                    if (oo >= 0) out[size_t(sample) * kOut + oo] = o[s][t][e];
                }
        }
        __syncthreads();
    }
}

// Forward and backward of 32-sample groups: stores every hidden activation
// and every pre-activation gradient in the group layout above, and adds the
// batch's weighted loss and weight sum as per-wave partials (k_red_sum adds
// them into red[0], red[1]).
// Most blocks of k_fwdbwd (each leaves 8 loss partials for k_red_sum).
constexpr int kMaxFwdGrid = 1024;

__global__ __launch_bounds__(512) void k_fwdbwd(const float4 *feats, const float4 *dirs, int n, const bf16x8 *ffrag,
                                               const float *fbias, const bf16x8 *bfrag, uint16_t *acts, int groups,
                                               float *red) {
    __shared__ bf16x8 sfrag[kFwdFrags * 64];
    __shared__ bf16x8 sbfrag[kBwdFrags * 64];
    __shared__ float sbias[kBiasFwd];
    __shared__ float sfeat[8][32 * kFeat];
    __shared__ float4 sdir[8][32];
    for (int i = threadIdx.x; i < kBwdFrags * 64; i += blockDim.x) sbfrag[i] = bfrag[i];
    load_lds(sfrag, ffrag, kFwdFrags * 64, sbias, fbias, kBiasFwd);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const int ngroups = (n + 31) / 32;
    float loss = 0.f, wsum = 0.f;
    for (int gb = blockIdx.x * 8; gb < ngroups; gb += gridDim.x * 8) {
        const int grp = gb + w;
        if (lane < 32) {
            int i = grp * 32 + lane;
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, d = a;
            if (grp < ngroups && i < n) {
                a = feats[2 * i];
                b = feats[2 * i + 1];
                d = dirs[i];
            }
            float *f = &sfeat[w][lane * kFeat];
            f[0] = a.x, f[1] = a.y, f[2] = a.z, f[3] = a.w, f[4] = b.x, f[5] = b.y, f[6] = b.z, f[7] = b.w;
            sdir[w][lane] = d;
        }
        __syncthreads();
        if (grp < ngroups) {
            f32x4 out[2][3];
            bf16x8 H[4][2][2];
            mlp_forward<2>(sfeat[w], sfrag, sbias, out, H);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint16_t *dst = acts + act_offset(k, groups) + size_t(grp) * 64 * 32;
#pragma unroll
                for (int s = 0; s < 2; s++)
#pragma unroll
                    for (int kb = 0; kb < 2; kb++)
#pragma unroll
                        for (int e = 0; e < 8; e++)
                            dst[kslot_hidden(kb, g, e) * 32 + 16 * s + j] = bf16_elem_bits(H[k][s][kb], e);
            }
            f32x4 gr[2][3];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                float4 d = sdir[w][16 * s + j];
                float l = lobe_grads(out[s], d.x, d.y, d.z, d.w, gr[s]);
                if (g == 0) {
                    loss += l;
                    wsum += d.w;
                }
            }
            bf16x8 dz[2][2];
            {
                uint16_t *dst = acts + act_offset(8, groups) + size_t(grp) * kOutRows * 32;
#pragma unroll
                for (int s = 0; s < 2; s++) {
#pragma unroll
                    for (int t = 0; t < 3; t++)
#pragma unroll
                        for (int e = 0; e < 4; e++) dst[(16 * t + 4 * g + e) * 32 + 16 * s + j] = bf16_bits(gr[s][t][e]);
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        dz[s][0][e] = (__bf16)gr[s][e >> 2][e & 3];
                        dz[s][1][e] = (__bf16)(e < 4 ? gr[s][2][e] : 0.f);
                    }
                }
            }
#pragma unroll
            for (int l = 4; l >= 1; l--) {
                f32x4 acc[2][4];
#pragma unroll
                for (int s = 0; s < 2; s++)
#pragma unroll
                    for (int t = 0; t < 4; t++) acc[s][t] = zero4();
#pragma unroll
                for (int kb = 0; kb < 2; kb++)
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        bf16x8 a = sbfrag[(8 * (l - 1) + 2 * t + kb) * 64 + lane];
#pragma unroll
                        for (int s = 0; s < 2; s++) acc[s][t] = mfma(a, dz[s][kb], acc[s][t]);
                    }
                uint16_t *dst = acts + act_offset(4 + (l - 1), groups) + size_t(grp) * 64 * 32;
#pragma unroll
                for (int s = 0; s < 2; s++)
#pragma unroll
                    for (int t = 0; t < 4; t++)
#pragma unroll
                        for (int e = 0; e < 4; e++) {
                            float v = acc[s][t][e];
                            if (!(float(H[l - 1][s][t >> 1][(t & 1) * 4 + e]) > 0.f)) v = 0.f;
                            acc[s][t][e] = v;
                            dst[(16 * t + 4 * g + e) * 32 + 16 * s + j] = bf16_bits(v);
                        }
                if (l > 1) {
#pragma unroll
                    for (int s = 0; s < 2; s++)
#pragma unroll
                        for (int kb = 0; kb < 2; kb++)
#pragma unroll
                            for (int e = 0; e < 8; e++) dz[s][kb][e] = (__bf16)acc[s][2 * kb + (e >> 2)][e & 3];
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        loss += __shfl_xor(loss, off);
        wsum += __shfl_xor(wsum, off);
    }
    // one partial per wave; k_red_sum adds them in a fixed order
    if (lane == 0) {
        red[4 + 2 * (blockIdx.x * 8 + w)] = loss;
        red[5 + 2 * (blockIdx.x * 8 + w)] = wsum;
    }
}

// Sums the per-wave loss and weight partials of k_fwdbwd into red[0],
// red[1] in a fixed order, so the normalisation of every gradient is the
// same on every run: thread t adds partials t, t + 256, ... in turn, each
// wave adds its threads' sums by a butterfly, and thread 0 adds the four
// waves' sums.
__global__ __launch_bounds__(256) void k_red_sum(float *red, int nparts) {
    float l = 0.f, ws = 0.f;
    for (int p = threadIdx.x; p < nparts; p += 256) {
        l += red[4 + 2 * p];
        ws += red[5 + 2 * p];
    }
    for (int off = 32; off > 0; off >>= 1) {
        l += __shfl_xor(l, off);
        ws += __shfl_xor(ws, off);
    }
    __shared__ float sl[4], sw[4];
    if ((threadIdx.x & 63) == 0) {
        sl[threadIdx.x >> 6] = l;
        sw[threadIdx.x >> 6] = ws;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        red[0] = (sl[0] + sl[1]) + (sl[2] + sl[3]);
        red[1] = (sw[0] + sw[1]) + (sw[2] + sw[3]);
    }
}

// Weight gradient of layer L over a range of groups: dW^T = sum dZ^T H_in,
// samples along K.  Wave w owns output rows 16w..16w+15 and every input
// column tile; layer 0's input (the one-blob code) is recomputed from the
// features.  Writes this block's partial gradient of layer L's weights and
// biases into partial[blockIdx.x].
template <int L>
__global__ __launch_bounds__(256) void k_wgrad(const float4 *feats, int n, const uint16_t *acts, int groups,
                                              int ngroups, int gpb, float *partial) {
    constexpr int TC = layer_in(L) / 16;
    constexpr int ROWS = layer_rows(L);
    __shared__ float sfeat[32 * kFeat];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const bool active = w < ROWS / 16;
    const int g0 = blockIdx.x * gpb, g1 = min(g0 + gpb, ngroups);
    const uint16_t *dZ = acts + act_offset(L == 4 ? 8 : 4 + L, groups);
    const uint16_t *Hin = L > 0 ? acts + act_offset(L - 1, groups) : nullptr;
    f32x4 acc[TC];
#pragma unroll
    for (int c = 0; c < TC; c++) acc[c] = zero4();
    float bsum = 0.f;
    for (int grp = g0; grp < g1; grp++) {
        if (L == 0) {
            __syncthreads();
            if (threadIdx.x < 64) {
                // 32 samples x 2 float4; rows past n are zero (their deltas are zero too)
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (grp * 32 + int(threadIdx.x >> 1) < n) v = feats[size_t(grp) * 64 + threadIdx.x];
                float *f = &sfeat[threadIdx.x * 4];
                f[0] = v.x, f[1] = v.y, f[2] = v.z, f[3] = v.w;
            }
            __syncthreads();
        }
        if (active) {
            bf16x8 a = *reinterpret_cast<const bf16x8 *>(dZ + (size_t(grp) * ROWS + 16 * w + j) * 32 + 8 * g);
#pragma unroll
            for (int e = 0; e < 8; e++) bsum += float(a[e]);
#pragma unroll
            for (int c = 0; c < TC; c++) {
                bf16x8 b;
                if (L == 0) {
                    int col = 16 * c + j, f, nb, b0;
                    input_group(col >> 3, f, nb, b0);
                    float cen = float(b0 + (col & 7)) + 0.5f;
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        float d = cen - sfeat[(8 * g + e) * kFeat + f] * float(nb);
                        b[e] = (__bf16)__expf(-0.5f * d * d);
                    }
                } else {
                    b = *reinterpret_cast<const bf16x8 *>(Hin + (size_t(grp) * 64 + 16 * c + j) * 32 + 8 * g);
                }
                acc[c] = mfma(a, b, acc[c]);
            }
        }
    }
    if (!active) return;
    float *dst = partial + size_t(blockIdx.x) * kNumParams;
#pragma unroll
    for (int c = 0; c < TC; c++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int row = 16 * w + 4 * g + e;
            int o = L == 4 ? out_of_row(row) : row;
            if (o >= 0) dst[param_w(L) + (16 * c + j) * layer_out(L) + o] = acc[c][e];
        }
    bsum = quad_sum(bsum);
    if (g == 0) {
        int row = 16 * w + j;
        int o = L == 4 ? out_of_row(row) : row;
        if (o >= 0) dst[param_b(L) + o] = bsum;
    }
}

// First stage of the partial-gradient sum: slice blockIdx.y of kReduceSlices
// adds its share of the nblk per-block rows (in a fixed order).
constexpr int kReduceSlices = 32;
__global__ void k_reduce_partials(const float *partial, int nblk, float *out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNumParams) return;
    float s = 0.f;
    for (int b = blockIdx.y; b < nblk; b += kReduceSlices) s += partial[size_t(b) * kNumParams + p];
    out[size_t(blockIdx.y) * kNumParams + p] = s;
}

// Sums the sliced partial gradients, normalises by the batch's weight sum
// and, with update set, takes one Adam step (Kingma and Ba 2015,
// bias-corrected).
__global__ void k_adam(const float *sliced, const float *red, float *master, float *m, float *v, float *grad,
                       float lr, float b1, float b2, float eps, float bc1, float bc2, int update) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNumParams) return;
    float s = 0.f;
    for (int b = 0; b < kReduceSlices; b++) s += sliced[size_t(b) * kNumParams + p];
    float ws = red[1];
    float gp = ws > 0.f ? s / ws : 0.f;
    grad[p] = gp;
    if (update) {
        float mm = b1 * m[p] + (1.f - b1) * gp;
        float vv = b2 * v[p] + (1.f - b2) * gp * gp;
        m[p] = mm;
        v[p] = vv;
        master[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    }
}

namespace {

void build_maps(std::vector<int> &fm, std::vector<int> &bm, std::vector<int> &biasm) {
    fm.assign(size_t(kFwdFrags) * kFragElems, -1);
    for (int l = 0; l < kLayers; l++)
        for (int t = 0; t < layer_tiles(l); t++)
            for (int kb = 0; kb < layer_kblocks(l); kb++) {
                int f = fwd_frag_base(l) + t * layer_kblocks(l) + kb;
                for (int lane = 0; lane < 64; lane++)
                    for (int e = 0; e < 8; e++) {
                        int i = lane & 15, g = lane >> 4;
                        int row = 16 * t + i;
                        int o = l == 4 ? out_of_row(row) : row;
                        if (o < 0) continue;
                        int k = l == 0 ? 32 * kb + 8 * g + e : kslot_hidden(kb, g, e);
                        fm[(size_t(f) * 64 + lane) * 8 + e] = param_w(l) + k * layer_out(l) + o;
                    }
            }
    bm.assign(size_t(kBwdFrags) * kFragElems, -1);
    for (int l = 1; l < kLayers; l++)
        for (int t = 0; t < 4; t++)
            for (int kb = 0; kb < 2; kb++) {
                int f = 8 * (l - 1) + 2 * t + kb;
                for (int lane = 0; lane < 64; lane++)
                    for (int e = 0; e < 8; e++) {
                        int i = lane & 15, g = lane >> 4;
                        int in = 16 * t + i;
                        int k = kslot_hidden(kb, g, e);
                        int o = l == 4 ? (k < kOutRows ? out_of_row(k) : -1) : k;
                        if (o < 0) continue;
                        bm[(size_t(f) * 64 + lane) * 8 + e] = param_w(l) + in * layer_out(l) + o;
                    }
            }
    biasm.assign(kBiasFwd, -1);
    for (int l = 0; l < kLayers; l++)
        for (int r = 0; r < layer_rows(l); r++) {
            int o = l == 4 ? out_of_row(r) : r;
            if (o >= 0) biasm[bias_base(l) + r] = param_b(l) + o;
        }
}

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// He-uniform hidden layers; a small output layer whose biases spread the 8
// lobe means over the cube diagonals with kappa = 3 and equal weights.
std::vector<float> init_params(uint64_t seed) {
    std::vector<float> p(kNumParams, 0.f);
    uint64_t s = seed;
    auto uni = [&]() { return double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); };
    for (int l = 0; l < kLayers; l++) {
        double a = l < 4 ? std::sqrt(6.0 / layer_in(l)) : 0.1 * std::sqrt(6.0 / layer_in(l));
        for (int k = 0; k < layer_in(l) * layer_out(l); k++) p[param_w(l) + k] = float((2 * uni() - 1) * a);
    }
    for (int lobe = 0; lobe < kLobes; lobe++) {
        float sx = (lobe & 1) ? 1.f : -1.f, sy = (lobe & 2) ? 1.f : -1.f, sz = (lobe & 4) ? 1.f : -1.f;
        float inv = 1.f / std::sqrt(3.f);
        p[param_b(4) + 5 * lobe + 0] = sx * inv;
        p[param_b(4) + 5 * lobe + 1] = sy * inv;
        p[param_b(4) + 5 * lobe + 2] = sz * inv;
        p[param_b(4) + 5 * lobe + 3] = std::log(3.f);
        p[param_b(4) + 5 * lobe + 4] = 0.f;
    }
    return p;
}

template <class T>
void dalloc(T *&p, size_t count) {
    HIPX(hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)));
}

void repack(Net &n, hipStream_t s) {
    int total = kFwdFrags * kFragElems + kBwdFrags * kFragElems + kBiasFwd;
    hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), dim3(256), 0, s, n.master, n.fmap,
                       reinterpret_cast<__bf16 *>(n.ffrag), kFwdFrags * kFragElems, n.bmap,
                       reinterpret_cast<__bf16 *>(n.bfrag), kBwdFrags * kFragElems, n.biasmap, n.fbias, kBiasFwd);
    HIPX(hipGetLastError());
}

}  // namespace

Net *net_create(int device, const ljx_guide_config &cfg) {
    int count = 0;
    HIPX(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) fail(LJX_ERR_INVALID_ARGUMENT, "no such device");
    HIPX(hipSetDevice(device));
    Net *n = new Net();
    try {
        n->device = device;
        n->cfg = cfg;
        hipDeviceProp_t prop;
        HIPX(hipGetDeviceProperties(&prop, device));
        n->num_cu = prop.multiProcessorCount;
        HIPX(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
        dalloc(n->master, kNumParams);
        dalloc(n->adam_m, kNumParams);
        dalloc(n->adam_v, kNumParams);
        dalloc(n->grad, kNumParams);
        dalloc(n->red, 4 + 2 * 8 * kMaxFwdGrid);
        dalloc(n->ffrag, size_t(kFwdFrags) * 64);
        dalloc(n->bfrag, size_t(kBwdFrags) * 64);
        dalloc(n->fbias, kBiasFwd);
        std::vector<int> fm, bm, biasm;
        build_maps(fm, bm, biasm);
        dalloc(n->fmap, fm.size());
        dalloc(n->bmap, bm.size());
        dalloc(n->biasmap, biasm.size());
        HIPX(hipMemcpy(n->fmap, fm.data(), fm.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPX(hipMemcpy(n->bmap, bm.data(), bm.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPX(hipMemcpy(n->biasmap, biasm.data(), biasm.size() * sizeof(int), hipMemcpyHostToDevice));
        n->batch_cap = (cfg.batch_size + 31) / 32 * 32;
        dalloc(n->acts, act_elems(n->batch_cap / 32));
        n->partial_blocks = 2 * n->num_cu;
        dalloc(n->partial, size_t(n->partial_blocks + kReduceSlices) * kNumParams);
        std::vector<float> p = init_params(cfg.seed);
        net_set_params(*n, p.data());
    } catch (...) {
        net_destroy(n);
        throw;
    }
    return n;
}

void net_destroy(Net *n) {
    if (!n) return;
    hipSetDevice(n->device);
    void *bufs[] = {n->master, n->adam_m, n->adam_v, n->grad, n->red, n->ffrag, n->bfrag, n->fbias,
                    n->fmap, n->bmap, n->biasmap, n->acts, n->partial};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (n->stream) hipStreamDestroy(n->stream);
    delete n;
}

void net_set_params(Net &n, const float *host_params) {
    HIPX(hipSetDevice(n.device));
    HIPX(hipMemcpy(n.master, host_params, kNumParams * sizeof(float), hipMemcpyHostToDevice));
    HIPX(hipMemset(n.adam_m, 0, kNumParams * sizeof(float)));
    HIPX(hipMemset(n.adam_v, 0, kNumParams * sizeof(float)));
    n.steps = 0;
    repack(n, n.stream);
    HIPX(hipStreamSynchronize(n.stream));
}

void net_step(Net &n, const float4 *feats, const float4 *dirs, int count, bool update, hipStream_t s,
              double *loss_out) {
    if (count <= 0) {
        if (loss_out) *loss_out = 0;
        return;
    }
    if (count > n.batch_cap) fail(LJX_ERR_INVALID_ARGUMENT, "batch larger than the configured batch_size");
    const int groups = n.batch_cap / 32;
    const int ngroups = (count + 31) / 32;
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, s, n.red, 4);
    int fgrid = std::max(1, std::min({(ngroups + 7) / 8, n.num_cu, kMaxFwdGrid}));
    hipLaunchKernelGGL(k_fwdbwd, dim3(fgrid), dim3(512), 0, s, feats, dirs, count, n.ffrag, n.fbias, n.bfrag, n.acts,
                       groups, n.red);
    hipLaunchKernelGGL(k_red_sum, dim3(1), dim3(256), 0, s, n.red, fgrid * 8);
    HIPX(hipGetLastError());
    int gpb = (ngroups + n.partial_blocks - 1) / n.partial_blocks;
    int nblk = (ngroups + gpb - 1) / gpb;
    hipLaunchKernelGGL(k_wgrad<0>, dim3(nblk), dim3(256), 0, s, feats, count, n.acts, groups, ngroups, gpb, n.partial);
    hipLaunchKernelGGL(k_wgrad<1>, dim3(nblk), dim3(256), 0, s, feats, count, n.acts, groups, ngroups, gpb, n.partial);
    hipLaunchKernelGGL(k_wgrad<2>, dim3(nblk), dim3(256), 0, s, feats, count, n.acts, groups, ngroups, gpb, n.partial);
    hipLaunchKernelGGL(k_wgrad<3>, dim3(nblk), dim3(256), 0, s, feats, count, n.acts, groups, ngroups, gpb, n.partial);
    hipLaunchKernelGGL(k_wgrad<4>, dim3(nblk), dim3(256), 0, s, feats, count, n.acts, groups, ngroups, gpb, n.partial);
    HIPX(hipGetLastError());
    float bc1 = 1.f, bc2 = 1.f;
    if (update) {
        n.steps++;
        bc1 = float(1.0 - std::pow(double(n.cfg.beta1), n.steps));
        bc2 = float(1.0 - std::pow(double(n.cfg.beta2), n.steps));
    }
    float *sliced = n.partial + size_t(n.partial_blocks) * kNumParams;
    hipLaunchKernelGGL(k_reduce_partials, dim3((kNumParams + 255) / 256, kReduceSlices), dim3(256), 0, s, n.partial,
                       nblk, sliced);
    hipLaunchKernelGGL(k_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, s, sliced, n.red, n.master, n.adam_m,
                       n.adam_v, n.grad, n.cfg.learning_rate, n.cfg.beta1, n.cfg.beta2, n.cfg.epsilon, bc1, bc2,
                       update ? 1 : 0);
    HIPX(hipGetLastError());
    if (update) repack(n, s);
    if (loss_out) {
        float r[2];
        HIPX(hipMemcpyAsync(r, n.red, sizeof(r), hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        *loss_out = r[1] > 0 ? double(r[0]) / double(r[1]) : 0.0;
    }
}

}  // namespace guide
}  // namespace ljx

using namespace ljx;

namespace {

guide::Net *as_net(void *h) {
    if (!h) fail(LJX_ERR_INVALID_ARGUMENT, "guide handle is NULL");
    return static_cast<guide::Net *>(h);
}

ljx_guide_config with_defaults(const ljx_guide_config *c) {
    ljx_guide_config d{};
    if (c) d = *c;
    if (!(d.learning_rate > 0)) d.learning_rate = 1e-3f;
    if (!(d.beta1 > 0 && d.beta1 < 1)) d.beta1 = 0.9f;
    if (!(d.beta2 > 0 && d.beta2 < 1)) d.beta2 = 0.999f;
    if (!(d.epsilon > 0)) d.epsilon = 1e-8f;
    if (d.batch_size <= 0) d.batch_size = 1 << 20;
    if (d.batch_size > (1 << 24)) fail(LJX_ERR_INVALID_ARGUMENT, "batch_size above 2^24");
    if (!c) d.seed = 1;
    return d;
}

}  // namespace

extern "C" {

int ljx_guide_create(int device, const ljx_guide_config *cfg, void **guide_out) {
    return guarded([&]() {
        if (!guide_out) fail(LJX_ERR_INVALID_ARGUMENT, "guide_out is NULL");
        *guide_out = nullptr;
        *guide_out = guide::net_create(device, with_defaults(cfg));
    });
}

int ljx_guide_destroy(void *g) {
    return guarded([&]() { guide::net_destroy(static_cast<guide::Net *>(g)); });
}

int ljx_guide_get_params(void *g, float *params, size_t n) {
    return guarded([&]() {
        guide::Net *net = as_net(g);
        if (!params || n != size_t(guide::kNumParams)) fail(LJX_ERR_INVALID_ARGUMENT, "params must hold LJX_GUIDE_NUM_PARAMS floats");
        HIPX(hipSetDevice(net->device));
        HIPX(hipStreamSynchronize(net->stream));
        HIPX(hipMemcpy(params, net->master, n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int ljx_guide_set_params(void *g, const float *params, size_t n) {
    return guarded([&]() {
        guide::Net *net = as_net(g);
        if (!params || n != size_t(guide::kNumParams)) fail(LJX_ERR_INVALID_ARGUMENT, "params must hold LJX_GUIDE_NUM_PARAMS floats");
        for (size_t i = 0; i < n; i++)
            if (!std::isfinite(params[i])) fail(LJX_ERR_INVALID_ARGUMENT, "non-finite parameter");
        guide::net_set_params(*net, params);
    });
}

int ljx_guide_forward_device(void *g, const float *features, size_t n, float *raw_out) {
    return guarded([&]() {
        guide::Net *net = as_net(g);
        if (n == 0) return;
        if (!features || !raw_out) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        if (n > size_t(1) << 30) fail(LJX_ERR_INVALID_ARGUMENT, "too many rows");
        HIPX(hipSetDevice(net->device));
        int grid = int(std::min<size_t>((n + 255) / 256, size_t(net->num_cu) * 2));
        hipLaunchKernelGGL(guide::k_forward_raw, dim3(grid), dim3(256), 0, net->stream,
                           reinterpret_cast<const float4 *>(features), int(n), net->ffrag, net->fbias, raw_out);
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(net->stream));
    });
}

int ljx_guide_gradient_device(void *g, const float *features, const float *dirs, size_t n, float *grad_out,
                              double *loss_out) {
    return guarded([&]() {
        guide::Net *net = as_net(g);
        if (!features || !dirs || !grad_out) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        if (n == 0 || n > size_t(net->batch_cap)) fail(LJX_ERR_INVALID_ARGUMENT, "n must be in [1, batch_size]");
        HIPX(hipSetDevice(net->device));
        double loss = 0;
        guide::net_step(*net, reinterpret_cast<const float4 *>(features), reinterpret_cast<const float4 *>(dirs),
                        int(n), false, net->stream, &loss);
        HIPX(hipMemcpyAsync(grad_out, net->grad, guide::kNumParams * sizeof(float), hipMemcpyDeviceToDevice,
                            net->stream));
        HIPX(hipStreamSynchronize(net->stream));
        if (loss_out) *loss_out = loss;
    });
}

int ljx_guide_train_step_device(void *g, const float *features, const float *dirs, size_t n, double *loss_out) {
    return guarded([&]() {
        guide::Net *net = as_net(g);
        if (!features || !dirs) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        if (n == 0 || n > size_t(net->batch_cap)) fail(LJX_ERR_INVALID_ARGUMENT, "n must be in [1, batch_size]");
        HIPX(hipSetDevice(net->device));
        double loss = 0;
        guide::net_step(*net, reinterpret_cast<const float4 *>(features), reinterpret_cast<const float4 *>(dirs),
                        int(n), true, net->stream, &loss);
        if (loss_out) *loss_out = loss;
    });
}

}  // extern "C"

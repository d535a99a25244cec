// Guided sampling inside the wavefront loop.
//
//   k_guide_infer    after k_extend: for each path that hit a guided material
//                    (the guided types' material buckets),
//                    network features of the vertex, the MLP forward pass on
//                    the matrix cores (64 paths per wave), the vertex's lobe
//                    set for the shading kernel, and the strategy choice: with
//                    probability guide_mix a direction drawn from the lobes
//   k_guide_collect  after a recording pass: walks each path's vertices from
//                    the last to the first, turning the radiance that arrived
//                    after a guided vertex into its training weight
//                    (luminance of f cos L_i / pdf), and appends the
//                    (features, direction, weight) samples to the batch
#pragma once

#include "ljx_kernels.h"

namespace ljx {

// Per-vertex guide stream: independent of the path's own PCG stream, so the
// BSDF draws of a guided path are the ones the unguided tracer makes.
__device__ __forceinline__ void guide_draws(const PathBufs &P, uint32_t pid, int b, float u[4]) {
    uint64_t gs = slot_info(P, int(pid)).stream() ^ 0x2545f4914f6cdd1dULL;
    uint64_t inc = pcg_inc(gs);
    uint64_t st = pcg_init_state(gs, P.seed + 0x9e3779b97f4a7c15ULL * uint64_t(b + 1));
#pragma unroll
    for (int k = 0; k < 4; k++) u[k] = float(pcg_double(st, inc));
}

__device__ __forceinline__ void dir_angles(float x, float y, float z, float &th, float &ph) {
    th = acosf(fminf(fmaxf(z, -1.f), 1.f)) * (1.f / 3.14159265358979f);
    float a = atan2f(y, x);
    if (a < 0.f) a += guide::kTwoPi;
    ph = a * (1.f / guide::kTwoPi);
}

// Orders a wave's LDS writes before its lanes read each other's (and its
// reads before the next round's writes).  The staging arrays of
// k_guide_infer are per wave, so its waves never wait for each other.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// WAVES waves per block share one copy of the network in LDS; each wave
// runs its 64 queries through the network NS sample tiles (16 queries each)
// at a time.
template <bool SPH, int WAVES, int NS>
__device__ __forceinline__ void guide_infer_body(DeviceScene S, PathBufs P, const guide::bf16x8 *ffrag,
                                                 const float *fbias) {
    using namespace guide;
    __shared__ bf16x8 sfrag[kFwdFrags * 64];
    __shared__ float sbias[kBiasFwd];
    __shared__ float sfeat[WAVES][64 * kFeat];
    __shared__ float su[WAVES][64][4];
    __shared__ int spid[WAVES][64];
    // The paths to query: the buckets of the guided material types (every
    // path there is guided) and the runtime-dispatch bucket (checked per
    // path), one index space over their entries.  Paths in other buckets
    // never read their guide direction, so the network runs on guided
    // vertices only.
    __shared__ int bstart[kNumBuckets + 1];
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int B = 0; B < kNumBuckets; B++) {
            bstart[B] = acc;
            const bool sel = B == kBucketAny || (B < kNumMatTypes && guidable(B) && ((P.guide_mask >> B) & 1u));
            if (sel) acc += P.ctr[C_BUCKET + B];
        }
        bstart[kNumBuckets] = acc;
        if (blockIdx.x == 0) atomicAdd(&P.stat[S_GUIDE_Q], (unsigned long long)acc);
    }
    // guidable types the mask leaves out: unguided
    for (int B = 0; B < kNumMatTypes; B++) {
        if (!guidable(B) || ((P.guide_mask >> B) & 1u)) continue;
        const int nb = P.ctr[C_BUCKET + B];
        for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += gridDim.x * blockDim.x)
            P.gdir[P.bucket[size_t(B) * P.capacity + k]] = make_float4(0.f, 0.f, 0.f, -1.f);
    }
    load_lds(sfrag, ffrag, kFwdFrags * 64, sbias, fbias, kBiasFwd);  // ends with a barrier
    const int n = bstart[kNumBuckets];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    for (int base = blockIdx.x * 64 * WAVES; base < n; base += gridDim.x * 64 * WAVES) {
        int i = base + w * 64 + lane;
        int pidg = -1;
        float f[kFeat] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (i < n) {
            int B = 0;
            while (i >= bstart[B + 1]) B++;
            uint32_t pid = P.bucket[size_t(B) * P.capacity + (i - bstart[B])];
            float4 h = P.hit[pid];
            int slot = __float_as_int(h.w);
            if (slot >= 0) {
                float4 p0 = S.prims[3 * slot];
                int sid = S.prim_shape[__float_as_int(p0.w)];
                const ShapeRec &sh = S.shapes[sid];
                int mid = sh.material_id;
                if (mid >= 0 && ((P.guide_mask >> S.materials[mid].type) & 1u)) {
                    float4 ro = P.ray_o[pid], rd = P.ray_d[pid];
                    float px = ro.x + h.x * rd.x, py = ro.y + h.x * rd.y, pz = ro.z + h.x * rd.z;
                    float nx, ny, nz;
                    if (SPH && sh.type == 0) {
                        nx = px - float(sh.center[0]);
                        ny = py - float(sh.center[1]);
                        nz = pz - float(sh.center[2]);
                    } else {
                        float4 p1 = S.prims[3 * slot + 1], p2 = S.prims[3 * slot + 2];
                        float ax = p1.x - p0.x, ay = p1.y - p0.y, az = p1.z - p0.z;
                        float bx = p2.x - p0.x, by = p2.y - p0.y, bz = p2.z - p0.z;
                        nx = ay * bz - az * by;
                        ny = az * bx - ax * bz;
                        nz = ax * by - ay * bx;
                    }
                    float il = rsqrtf(fmaxf(nx * nx + ny * ny + nz * nz, 1e-30f));
                    nx *= il, ny *= il, nz *= il;
                    // the normal on the side of the outgoing direction wo = -d
                    if (nx * rd.x + ny * rd.y + nz * rd.z > 0.f) nx = -nx, ny = -ny, nz = -nz;
                    f[0] = fminf(fmaxf((px - P.box_lo[0]) * P.box_scale[0], 0.f), 1.f);
                    f[1] = fminf(fmaxf((py - P.box_lo[1]) * P.box_scale[1], 0.f), 1.f);
                    f[2] = fminf(fmaxf((pz - P.box_lo[2]) * P.box_scale[2], 0.f), 1.f);
                    dir_angles(nx, ny, nz, f[3], f[4]);
                    dir_angles(-rd.x, -rd.y, -rd.z, f[5], f[6]);
                    int b = P.depth[pid];
                    guide_draws(P, pid, b, su[w][lane]);
                    if (b < P.rec_depth) {
                        size_t r = size_t(b) * P.rec_stride + pid;
                        P.rec_feat[2 * r] = make_float4(f[0], f[1], f[2], f[3]);
                        P.rec_feat[2 * r + 1] = make_float4(f[4], f[5], f[6], 0.f);
                    }
                    pidg = int(pid);
                }
            }
            if (pidg < 0) P.gdir[pid] = make_float4(0.f, 0.f, 0.f, -1.f);
        }
#pragma unroll
        for (int k = 0; k < kFeat; k++) sfeat[w][lane * kFeat + k] = f[k];
        spid[w][lane] = pidg;
        wave_lds_sync();
#pragma unroll 1
        for (int half = 0; half < 4 / NS; half++) {
        // the weight fragments are read again for each half, not kept in
        // registers across them
        if (NS < 4) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        f32x4 o[NS][3];
        mlp_forward<NS>(sfeat[w] + 16 * NS * half * kFeat, sfrag, sbias, o, nullptr);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int sl = 16 * NS * half + 16 * s + j;
            const int pid = spid[w][sl];
            Lobe L0, L1;
            lane_lobes(o[s], L0, L1);
            float lm = quad_max(fmaxf(L0.logit, L1.logit));
            float zs = quad_sum(__expf(L0.logit - lm) + __expf(L1.logit - lm));
            float log_z = lm + __logf(zs);
            LobeEval E0 = eval_lobe(L0, log_z), E1 = eval_lobe(L1, log_z);
            if (pid >= 0) {
                float4 *dst = P.glob + size_t(pid) * kLobes + 2 * g;
                dst[0] = make_float4(E0.kmx, E0.kmy, E0.kmz, E0.c);
                dst[1] = make_float4(E1.kmx, E1.kmy, E1.kmz, E1.c);
            }
            // lobe choice: the four lanes of the sample make the same draw
            float w0 = __expf(E0.logw), pair = w0 + __expf(E1.logw);
            float c0 = __shfl(pair, j), c1 = __shfl(pair, j + 16), c2 = __shfl(pair, j + 32), c3 = __shfl(pair, j + 48);
            float x = su[w][sl][1] * (c0 + c1 + c2 + c3);
            int gs = 3;
            if (x < c0) gs = 0;
            else if (x < c0 + c1) gs = 1;
            else if (x < c0 + c1 + c2) gs = 2;
            float below = gs == 0 ? 0.f : (gs == 1 ? c0 : (gs == 2 ? c0 + c1 : c0 + c1 + c2));
            float w0s = __shfl(w0, j + 16 * gs);
            bool second = (x - below) >= w0s;
            const LobeEval &E = second ? E1 : E0;  // meaningful in the owning lane
            float mx = __shfl(E.mx, j + 16 * gs), my = __shfl(E.my, j + 16 * gs), mz = __shfl(E.mz, j + 16 * gs);
            float kap = __shfl(E.kappa, j + 16 * gs);
            if (pid >= 0 && g == 0) {
                if (su[w][sl][0] < P.guide_mix) {
                    float dx, dy, dz;
                    sample_vmf(mx, my, mz, kap, su[w][sl][2], su[w][sl][3], dx, dy, dz);
                    P.gdir[pid] = make_float4(dx, dy, dz, 1.f);
                } else {
                    P.gdir[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
                }
            }
        }
        }
        wave_lds_sync();
    }
}

// Eight waves per block share one copy of the network (2 blocks and 16 waves
// per CU), each wave running its 64 queries as two halves of 32 so that it
// fits 128 registers: 11% faster than four waves of 64-query tiles at two
// waves per SIMD.
template <bool SPH>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_guide_infer(
    DeviceScene S, PathBufs P, int cur, const guide::bf16x8 *ffrag, const float *fbias) {
    guide_infer_body<SPH, 8, 2>(S, P, ffrag, fbias);
}

// Clears the record arrays of a recording pass (a fill kernel rather than a
// memset, so the pass can be captured).
__global__ void k_guide_clear(float4 *a, float4 *b, float4 *c, size_t n) {
    size_t stride = size_t(gridDim.x) * blockDim.x;
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        a[i] = z;
        b[i] = z;
        c[i] = z;
    }
}

// Walks path i's recorded vertices from its last back to its first and
// calls take(r, dir, weight) for each guided vertex with a positive finite
// weight w = lum(radiance that arrived after the vertex / its throughput).
template <class F>
__device__ __forceinline__ void guide_path_samples(const PathBufs &P, int i, F &&take) {
    const int last = P.depth[i];
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int b = P.rec_depth - 1; b >= 0; b--) {
        if (b > last) continue;  // vertices past the one the path ended at hold nothing of this pass
        const size_t r = size_t(b) * P.rec_stride + i;
        const float4 d = P.rec_dir[r];
        if (d.w > 0.5f) {
            const float4 be = P.rec_beta[r];
            const float wr = be.x > 0.f ? sx / be.x : 0.f;
            const float wg = be.y > 0.f ? sy / be.y : 0.f;
            const float wb = be.z > 0.f ? sz / be.z : 0.f;
            const float wgt = 0.212671f * wr + 0.715160f * wg + 0.072169f * wb;
            if (wgt > 0.f && isfinite(wgt)) take(r, d, wgt);
        }
        const float4 c = P.rec_c[r];
        sx += c.x, sy += c.y, sz += c.z;
    }
}

// Training samples of a recording pass, in a fixed order (paths in slot
// order, each path's vertices from last to first): the samples of each
// block of 256 paths are counted, the counts scanned, then each block writes
// its samples at its offset, so training sees the same batches on every run.
__global__ __launch_bounds__(256) void k_guide_count(PathBufs P, int *blk_count) {
    const int n = P.pass_ns * P.pass_nq;
    const int i = blockIdx.x * 256 + threadIdx.x;
    int c = 0;
    if (i < n) guide_path_samples(P, i, [&](size_t, float4, float) { c++; });
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    __shared__ int part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) blk_count[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of the block counts in place; count[0] gets the total.
__global__ __launch_bounds__(1024) void k_guide_scan(int *blk_count, int nblk, int *count) {
    __shared__ int part[1024];
    int carry = 0;
    for (int base = 0; base < nblk; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nblk ? blk_count[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int a = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nblk) blk_count[i] = carry + part[threadIdx.x] - v;
        const int tot = part[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) count[0] = carry;
}

__global__ __launch_bounds__(256) void k_guide_collect(PathBufs P, const int *blk_off, float4 *out_feat,
                                                       float4 *out_dir, int cap) {
    const int n = P.pass_ns * P.pass_nq;
    const int i = blockIdx.x * 256 + threadIdx.x;
    int c = 0;
    if (i < n) guide_path_samples(P, i, [&](size_t, float4, float) { c++; });
    // exclusive scan of the per-thread counts over the block, in thread order
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = c;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    __shared__ int wsum[4];
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int k = blk_off[blockIdx.x] + x - c;
    for (int q = 0; q < w; q++) k += wsum[q];
    if (i < n)
        guide_path_samples(P, i, [&](size_t r, float4 d, float wgt) {
            if (k < cap) {
                out_feat[2 * size_t(k)] = P.rec_feat[2 * r];
                out_feat[2 * size_t(k) + 1] = P.rec_feat[2 * r + 1];
                out_dir[k] = make_float4(d.x, d.y, d.z, wgt);
            }
            k++;
        });
}

}  // namespace ljx

// Wavefront path-tracing kernels (surface transport).
//
// One bounce of lajolla's path_tracing (src/path_tracing.h:7-325) is split
// into stages that each run over a compacted SoA queue of path ids:
//
//   k_raygen      sample_primary + the two screen-jitter draws
//   k_extend      closest-hit traversal, 64 rays per wave fetch; appends each
//                 path to the bucket of the hit material's type (material
//                 sort) or to MISS
//   k_shade<MT>   one kernel per material type: emission + MIS of the hit
//                 reached by the previous BSDF sample, Russian roulette,
//                 throughput update, then NEE (light pick, point, shadow ray
//                 with its pending contribution) and BSDF sampling
//   k_miss        environment term of escaped paths
//   k_shadow      any-hit traversal; an occlusion byte per shadow ray
//   k_shadow_resolve  adds the pending NEE contributions of unoccluded rays
//   k_accumulate  per-pixel sum of the pass's samples in sample order
//
// Each path draws from its own PCG stream (stream = sample * W * H + pixel)
// in the reference's per-path order, so results do not depend on launch
// shape or queue order.
#pragma once

#include "ljx_guide.h"
#include "ljx_shade.h"

namespace ljx {

constexpr int kNumMatTypes = 9;
constexpr int kBucketMiss = 9;
constexpr int kBucketNoMat = 10;
constexpr int kBucketAny = 11;  // unsorted hits (below the sort threshold), shaded by runtime dispatch
constexpr int kNumBuckets = 12;
// counter slots
constexpr int C_EXT = 0;      // C_EXT + (0|1): extension queues
constexpr int C_SH = 2;       // shadow queue
constexpr int C_BUCKET = 4;   // C_BUCKET + bucket
constexpr int C_XFETCH = 16;  // C_XFETCH + (0|1): rays taken from extension queue 0|1 by the traversal
constexpr int C_FLQ = 18;     // volumetric free-flight queue: its length, then (C_FLQ + 1) paths taken from it
constexpr int C_TOTAL = 16 + 4;

// One job of a batched render: its camera, image and place in the batch's
// global work index space (sample-major over its 8x8-tiled pixel grid).
struct BatchJob {
    CameraRec camera;
    long long work_base;
    int W, H, ntx8, npix8;
    int spp, pad;
    uint64_t seed;
    double *accum;    // W * H * 3
    float *out;       // W * H * 3
};
// 64-bit statistics slots
constexpr int S_EXT_RAYS = 0, S_SH_RAYS = 1, S_NODES = 2, S_PRIMS = 3, S_SH_NODES = 4, S_SH_PRIMS = 5,
              S_GUIDE_Q = 6, S_NUM = 8;

struct PathBufs {
    uint64_t *rng;
    float4 *tp;       // throughput rgb, eta_scale
    float4 *rad;      // radiance rgb
    float4 *prev;     // f (with cosine) of the last BSDF sample, its solid-angle pdf
    int *depth;       // index b of the vertex a path's current ray will hit (0: primary)
    float4 *ray_o;    // origin, tnear
    float4 *ray_d;    // direction, tfar
    float4 *hit;      // t, u, v, slot (int bits; -1: miss)
    uint32_t *extq0, *extq1;
    uint32_t *bucket; // kNumBuckets * capacity
    uint32_t *shq;
    float4 *sh_o, *sh_d, *sh_c;
    uint8_t *sh_occ;  // per shadow ray: 1 when occluded
    uint8_t *alive;   // per path slot: 1 when shading gave it another ray (cleared as the queue is built)
    int *alive_blk;   // per kScanBlock slots: live paths, then their exclusive prefix
    // volumetric path state
    int *medium;      // current medium id
    float4 *nee_c;    // nee_p_cache xyz, dir_pdf
    float4 *mtp;      // multi_trans_pdf rgb, never_scatter (1/0)
    float4 *sh_x;     // shadow walk: pdf_nee, pdf_dir, NEE stream state (lo, hi bits)
    double4 *vf_t;    // free flight: transmittance rgb, distance travelled
    double4 *vf_p;    // free flight: transmittance pdf rgb, 1 if it ended in a real collision
    uint32_t *flq;    // free flight: the paths of the volumetric queue that are in a medium
    int *ctr;
    unsigned long long *stat;
    int *spill;
    double *accum;    // W * H * 3
    int capacity;
    // image / pass
    int W, H, ntx8;
    int pass_s0, pass_ns, pass_q0, pass_nq;
    int spp, max_depth, rr_depth;
    int volpath;
    int vol_flight;  // free flight runs in k_vol_flight (scenes with grid media), else inside k_vol_shade
    int max_null;
    uint64_t seed;
    // learned guiding (path integrator)
    float4 *glob;      // capacity x 8 lobes: kappa * mean xyz, log constant (pdf = exp(dot + c))
    float4 *gdir;      // guide-sampled direction xyz; w: 1 from the guide, 0 from the BSDF, -1 not guided
    float4 *rec_feat;  // rec_depth x capacity x 2: network features of each vertex
    float4 *rec_dir;   // rec_depth x capacity: sampled direction xyz; w: 1 if a training sample
    float4 *rec_beta;  // rec_depth x capacity: throughput at the vertex
    float4 *rec_c;     // rec_depth x capacity: radiance the vertex adds (emission, NEE, environment)
    int rec_depth;     // vertices recorded per path (0: not recording)
    int rec_stride;    // record arrays are [vertex][slot]: the pass's slot count
    uint32_t guide_mask;  // material types sampled from the mixture
    float guide_mix;   // probability of the guide in the mixture
    float box_lo[3], box_scale[3];  // position features: (p - lo) * scale
    // batched rendering: slot i of a pass holds global work item work0 + i
    const BatchJob *jobs;  // null for a single render
    int njobs;
    int sort_threshold;    // bounces with fewer extension rays skip the material sort
    long long work0;
};

// Shadow-queue entries carry the path slot in the low kSlotBits bits and the
// vertex index in the rest (clamped there: only the guiding record, a few
// vertices deep, reads it).
constexpr int kSlotBits = 26;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1;
// Material types never guided: their lobes are near-specular transmission.
__host__ __device__ constexpr bool guidable(int mt) { return mt != 2 && mt != 5; }

// The camera path held by slot i: its job (0 for a single render), sample
// index, pixel, and its PCG stream, path_stream(s * W * H + y * W + x).
struct SlotInfo {
    int job, s, x, y, W, H;
    bool valid;
    __device__ uint64_t stream() const {
        return path_stream(uint64_t(s) * uint64_t(W) * uint64_t(H) + uint64_t(y) * W + x);
    }
};
__device__ __forceinline__ SlotInfo slot_info(const PathBufs &P, int i) {
    SlotInfo r;
    int q, ntx8;
    if (P.jobs) {
        long long w = P.work0 + i;
        int lo = 0, hi = P.njobs - 1;  // last job with work_base <= w
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (P.jobs[mid].work_base <= w) lo = mid;
            else hi = mid - 1;
        }
        const BatchJob &J = P.jobs[lo];
        long long local = w - J.work_base;
        r.job = lo;
        r.s = int(local / J.npix8);
        q = int(local % J.npix8);
        ntx8 = J.ntx8;
        r.W = J.W;
        r.H = J.H;
    } else {
        r.job = 0;
        r.s = P.pass_s0 + i / P.pass_nq;
        q = P.pass_q0 + i % P.pass_nq;
        ntx8 = P.ntx8;
        r.W = P.W;
        r.H = P.H;
    }
    int tile = q >> 6, w = q & 63;
    r.x = (tile % ntx8) * 8 + (w & 7);
    r.y = (tile / ntx8) * 8 + (w >> 3);
    r.valid = r.x < r.W && r.y < r.H;
    return r;
}

// Counter resets folded into kernels that do not read those counters.
__global__ void k_reset_counters(int *ctr, int first, int count) {
    int i = threadIdx.x;
    if (i < count) ctr[first + i] = 0;
}

// The next bounce's extension queue: the slots shading marked alive, in slot
// order (a stable compaction), so that neighbouring lanes of the next
// bounce's kernels read neighbouring path state and trace rays from
// neighbouring pixels.  Three launches over the pass's n slots: counts per
// kScanBlock slots, their exclusive scan (which also sets the queue's
// length), and the scatter, which clears the flags it read.
constexpr int kScanBlock = 4096;  // 256 threads x 16 slots
__device__ __forceinline__ int alive_in(uint4 f) {
    // flags are 0 or 1 bytes
    return __popc(f.x) + __popc(f.y) + __popc(f.z) + __popc(f.w);
}
__device__ __forceinline__ int block_excl_scan(int v, int &total) {
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int k = 0; k < 4; k++) {
        if (k < w) before += wsum[k];
        total += wsum[k];
    }
    __syncthreads();
    return before + x - v;
}
// Clears the flags of a pass's n slots (rounded up to 16); the scatter keeps
// them clear from then on.
__global__ __launch_bounds__(256) void k_alive_clear(PathBufs P, int n) {
    const int n16 = (n + 15) / 16;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += gridDim.x * blockDim.x)
        reinterpret_cast<uint4 *>(P.alive)[i] = make_uint4(0, 0, 0, 0);
}
__global__ __launch_bounds__(256) void k_alive_count(PathBufs P, int n) {
    const int s0 = blockIdx.x * kScanBlock + threadIdx.x * 16;
    const int c = s0 < n ? alive_in(*reinterpret_cast<const uint4 *>(P.alive + s0)) : 0;
    int total;
    block_excl_scan(c, total);
    if (threadIdx.x == 0) P.alive_blk[blockIdx.x] = total;
}
__global__ __launch_bounds__(1024) void k_alive_scan(PathBufs P, int nblk, int cur) {
    // 1024 threads x 16 blocks each: nblk <= 16384 (n <= 2^26)
    __shared__ int part[1024];
    const int b0 = threadIdx.x * 16;
    int v[16], sum = 0;
    for (int k = 0; k < 16; k++) {
        v[k] = b0 + k < nblk ? P.alive_blk[b0 + k] : 0;
        sum += v[k];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int y = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += y;
        __syncthreads();
    }
    int run = part[threadIdx.x] - sum;
    for (int k = 0; k < 16; k++) {
        if (b0 + k < nblk) P.alive_blk[b0 + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 1023) P.ctr[C_EXT + (cur ^ 1)] = part[1023];
}
__global__ __launch_bounds__(256) void k_alive_scatter(PathBufs P, int n, int cur) {
    uint32_t *next = cur ? P.extq0 : P.extq1;
    const int s0 = blockIdx.x * kScanBlock + threadIdx.x * 16;
    uint4 f = make_uint4(0, 0, 0, 0);
    if (s0 < n) f = *reinterpret_cast<const uint4 *>(P.alive + s0);
    int total;
    int k = P.alive_blk[blockIdx.x] + block_excl_scan(alive_in(f), total);
    if (s0 < n && (f.x | f.y | f.z | f.w)) {
        const uint32_t w[4] = {f.x, f.y, f.z, f.w};
        for (int j = 0; j < 16; j++)
            if ((w[j >> 2] >> (8 * (j & 3))) & 0xffu) next[k++] = uint32_t(s0 + j);
        *reinterpret_cast<uint4 *>(P.alive + s0) = make_uint4(0, 0, 0, 0);
    }
}

__global__ void k_fill_f64(double *p, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// compact = 0: every slot holds a valid pixel (W and H multiples of 8), so the
// extension queue is the identity and needs no atomics.
__global__ void k_raygen(DeviceScene S, PathBufs P, int compact) {
    int n = P.pass_ns * P.pass_nq;
    int stride = gridDim.x * blockDim.x;
    if (!compact && blockIdx.x == 0 && threadIdx.x == 0) P.ctr[C_EXT + 0] = n;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        int i = base + threadIdx.x;
        SlotInfo si{};
        bool valid = false;
        if (i < n) {
            si = slot_info(P, i);
            valid = si.valid;
        }
        if (valid) {
            uint64_t stream = si.stream();
            uint64_t inc = pcg_inc(stream);
            uint64_t st = pcg_init_state(stream, P.jobs ? P.jobs[si.job].seed : P.seed);
            double r0 = pcg_double(st, inc), r1 = pcg_double(st, inc);
            D3 o, d;
            const CameraRec &cam = P.jobs ? P.jobs[si.job].camera : S.camera;
            sample_primary(cam, (si.x + r0) / si.W, (si.y + r1) / si.H, o, d);
            P.rng[i] = st;
            P.ray_o[i] = make_float4(float(o.x), float(o.y), float(o.z), 0.0f);
            P.ray_d[i] = make_float4(float(d.x), float(d.y), float(d.z), __builtin_huge_valf());
            P.tp[i] = make_float4(1.f, 1.f, 1.f, 1.f);
            P.prev[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            P.depth[i] = 0;
        }
        if (i < n) P.rad[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (compact) {
            int k = wave_append(valid, &P.ctr[C_EXT + 0]);
            if (k >= 0) P.extq0[k] = uint32_t(i);
        } else if (i < n) {
            P.extq0[i] = uint32_t(i);
        }
    }
}

// Closest-hit traversal of the extension queue.  A wave takes 64 rays at a
// time with one atomic and walks them until the last one ends, so waves that
// drew short rays take more and the launch's tail stays short; each hit's
// distance is refined while its triangle is in cache, and the wave appends
// its paths to the buckets of the materials they hit (the material sort)
// with one atomic per bucket.  (Refilling lanes as their rays end, as
// k_shadow does, measured slower here: the refill points run the finishing
// work divergently and scatter the bucket order that shading reads.)
// The fetch counter of a launch starts at zero: k_reset_counters clears
// every counter at a pass's start and each launch clears the one the next
// launch uses.
template <bool STATS, bool SPH>
__device__ __forceinline__ int extend_finish(const DeviceScene &S, const PathBufs &P, uint32_t pid,
                                             const Tracer<false, STATS, SPH> &tr, bool sorted) {
    int bucket;
    if (tr.found) {
        TraceHit h = tr.hit;
        // the direction again for the FP64 distance (kept out of the
        // traversal's registers when no spheres need it there)
        const float4 d2 = SPH ? make_float4(tr.dx, tr.dy, tr.dz, 0.f) : P.ray_d[pid];
        h.t = refine_t<SPH>(S.prims, h.slot, h.t, tr.r.ox, tr.r.oy, tr.r.oz, d2.x, d2.y, d2.z);
        P.hit[pid] = make_float4(h.t, h.u, h.v, __int_as_float(h.slot));
        const int sb = S.slot_bucket[h.slot];
        // textured materials are shaded by the runtime-dispatch kernel
        bucket = (sorted || sb >= kNumMatTypes) ? sb : kBucketAny;
    } else {
        P.hit[pid] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
        bucket = kBucketMiss;
    }
    // volumetric paths decide between medium and surface events after
    // free-flight sampling: one queue
    return P.volpath ? 0 : bucket;
}

// Triangle-only scenes: capped at 64 registers, eight waves per SIMD (the
// LDS stacks allow eight), at the cost of a 16-byte spill; scenes with
// spheres need the FP64 root solve's registers and keep their own count.
template <bool STATS, bool SPH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SPH ? 1 : 8))) void k_extend(DeviceScene S, PathBufs P, int cur) {
    __shared__ int stack[kStackLds * 256];
    const uint32_t *q = cur ? P.extq1 : P.extq0;
    const int n = P.ctr[C_EXT + cur];
    const bool sorted = n >= P.sort_threshold;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.ctr[C_SH] = 0;
        P.ctr[C_XFETCH + (cur ^ 1)] = 0;
        atomicAdd(&P.stat[S_EXT_RAYS], (unsigned long long)n);
    }
    int *fetch = &P.ctr[C_XFETCH + cur];
    const TraceScene ts{S.qnodes, S.prims, S.spheres};
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int sstride = gridDim.x * blockDim.x;
    int *spill = P.spill + gtid;
    const int lane = __lane_id();
    uint32_t nn = 0, np = 0;
    while (true) {
        int base = 0;
        if (lane == 0) base = atomicAdd(fetch, 64);
        base = __shfl(base, 0);
        if (base >= n) break;
        const int i = base + lane;
        int bucket = -1;
        uint32_t pid = 0;
        if (i < n) {
            pid = q[i];
            const float4 o = P.ray_o[pid], d = P.ray_d[pid];
            Tracer<false, STATS, SPH> tr;
            tr.start(o.x, o.y, o.z, d.x, d.y, d.z, o.w, d.w);
            if (S.num_nodes > 0)
                while (!tr.step(ts, stack + threadIdx.x, blockDim.x, spill, sstride, nn, np)) {
                }
            bucket = extend_finish<STATS, SPH>(S, P, pid, tr, sorted);
        }
        const int k = wave_append_keyed(bucket, &P.ctr[C_BUCKET]);
        if (bucket >= 0) P.bucket[size_t(bucket) * P.capacity + k] = pid;
    }
    if (STATS) {
        atomicAdd(&P.stat[S_NODES], (unsigned long long)nn);
        atomicAdd(&P.stat[S_PRIMS], (unsigned long long)np);
    }
}

// Pending NEE contribution and the next ray of one path vertex.
template <int MT, bool SPH, bool GUIDE>
__device__ __forceinline__ void shade_surface(const DeviceScene &S, const PathBufs &P, uint32_t pid, bool &push_ext,
                                              bool &push_sh, float4 &sh_o, float4 &sh_d, float4 &sh_c) {
    float4 h = P.hit[pid];
    float4 ro = P.ray_o[pid], rd = P.ray_d[pid];
    D3 org = f3d(ro), dir = f3d(rd);
    int b = P.depth[pid];
    uint64_t st = P.rng[pid];
    const uint64_t inc = pcg_inc(slot_info(P, int(pid)).stream());
    float4 tp4 = P.tp[pid];
    D3 tp{tp4.x, tp4.y, tp4.z};
    double eta_scale = tp4.w;
    float4 rad4 = P.rad[pid];
    D3 rad{rad4.x, rad4.y, rad4.z};
    Vertex v = make_vertex<SPH>(S, __float_as_int(h.w), h.x, h.y, h.z, org, dir);
    // constant materials are read in place; the runtime-dispatch kernel, which
    // also takes textured ones, works on a copy with the textures evaluated
    // (the footprint is nonzero on camera rays only, path_tracing.h:14-16)
    MatRec mat_tex;
    if constexpr (MT < 0) {
        mat_tex = S.materials[v.material_id];
        if (mat_tex.textured) {
            double fp = 0;
            if (b == 0) {
                const SlotInfo si = slot_info(P, int(pid));
                fp = (0.25 / double(max(si.W, si.H))) * sqrt(dist2(org, v.pos)) / v.inv_uv;
            }
            resolve_textures(S, mat_tex, v.tu, v.tv, fp);
        }
    }
    const MatRec &mat = MT < 0 ? mat_tex : S.materials[v.material_id];
    bool alive = true;
    push_ext = false;
    push_sh = false;
    // guided vertex: directions come from the mixture (1 - a) BSDF + a guide
    bool guided = false;
    float4 gd = make_float4(0.f, 0.f, 0.f, -1.f);
    const float4 *lobes = nullptr;
    const double alpha = GUIDE ? double(P.guide_mix) : 0.0;
    if (GUIDE && guidable(MT >= 0 ? MT : int(mat.type))) {
        gd = P.gdir[pid];
        guided = gd.w >= 0.f;
        lobes = P.glob + size_t(pid) * guide::kLobes;
    }
    const bool rec = GUIDE && b < P.rec_depth;
    const size_t rix = size_t(b) * P.rec_stride + pid;
    D3 rad_in = rad;
    if (b == 0) {
        // path_tracing.h:58-61
        if (v.area_light_id >= 0)
            rad = rad + tp * light_emission(S, S.lights[v.area_light_id], -dir, LightPoint{v.pos, v.g.ng});
    } else {
        // path_tracing.h:243-322 for the vertex reached by the previous sample
        float4 pv = P.prev[pid];
        D3 f{pv.x, pv.y, pv.z};
        double pdf = pv.w;
        double G = fabs(dot(dir, v.g.ng)) / dist2(v.pos, org);
        double p2 = pdf * G;
        if (v.area_light_id >= 0) {
            const LightRec &l = S.lights[v.area_light_id];
            LightPoint lp{v.pos, v.g.ng};
            D3 L = light_emission(S, l, -dir, lp);
            D3 C2 = G * f * L;
            double p1 = S.light_pmf[v.area_light_id] * pdf_light_point(S, l, lp, org);
            double w2 = (p2 * p2) / (p1 * p1 + p2 * p2);
            C2 = C2 / p2;
            rad = rad + tp * C2 * w2;
        }
        double rr = 1;
        if (b + 1 >= P.rr_depth) {
            rr = fmin(maxc((1 / eta_scale) * tp), 0.95);
            if (pcg_double(st, inc) > rr) alive = false;
        }
        if (alive) tp = tp * (G * f) / (p2 * rr);
    }
    if (rec) {
        D3 add = rad - rad_in;
        P.rec_c[rix] = make_float4(float(add.x), float(add.y), float(add.z), 0.f);
    }
    if (alive && !(P.max_depth == -1 || b + 3 <= P.max_depth + 1)) alive = false;
    if (alive && S.num_lights == 0) {
        // the reference asserts on an empty light table (src/table_dist.cpp:29); end the path
        alive = false;
    }
    if (alive) {
        // next-event estimation (path_tracing.h:96-207)
        double lu0 = pcg_double(st, inc), lu1 = pcg_double(st, inc);
        double lw = pcg_double(st, inc), sw = pcg_double(st, inc);
        int lid = sample_cdf_guided(S.light_cdf, S.num_lights, lw, S.light_guide, S.light_guide_m);
        const LightRec &l = S.lights[lid];
        D3 dir_light, L;
        double G = 0, tfar, lpdf;
        if (l.type == 0) {
            LightPoint lp = sample_shape(S, S.shapes[l.shape_id], v.pos, lu0, lu1, sw);
            dir_light = normalize(lp.pos - v.pos);
            double dl = sqrt(dist2(lp.pos, v.pos));
            tfar = (1 - S.eps) * dl;
            G = fmax(-dot(dir_light, lp.n), 0.0) / dist2(lp.pos, v.pos);
            lpdf = pdf_shape(S.shapes[l.shape_id], lp, v.pos);
            L = dot(lp.n, -dir_light) > 0 ? l.intensity : D3{0, 0, 0};
        } else {
            env_nee(l, lu0, lu1, dir_light, lpdf, L);
            tfar = __builtin_huge_val();
            G = 1;
        }
        double p1 = S.light_pmf[lid] * lpdf;
        D3 dir_view = -dir;
        if (G > 0 && p1 > 0) {
            D3 f = bsdf_eval<MT>(mat, dir_view, dir_light, v.g);
            D3 C1 = G * f * L;
            double pb = bsdf_pdf<MT>(mat, dir_view, dir_light, v.g);
            if (guided)
                pb = (1 - alpha) * pb +
                     alpha * double(guide::mixture_pdf(lobes, float(dir_light.x), float(dir_light.y), float(dir_light.z)));
            double p2 = pb * G;
            double w1 = (p1 * p1) / (p1 * p1 + p2 * p2);
            C1 = C1 / p1;
            D3 c = tp * C1 * w1;
            if (c.x != 0 || c.y != 0 || c.z != 0) {
                push_sh = true;
                sh_o = make_float4(float(v.pos.x), float(v.pos.y), float(v.pos.z), float(S.eps));
                sh_d = make_float4(float(dir_light.x), float(dir_light.y), float(dir_light.z), float(tfar));
                sh_c = make_float4(float(c.x), float(c.y), float(c.z),
                                   __int_as_float(int(pid | (uint32_t(min(b, (1 << (32 - kSlotBits)) - 1)) << kSlotBits))));
            }
        }
        // BSDF sampling (path_tracing.h:209-259)
        double bu0 = pcg_double(st, inc), bu1 = pcg_double(st, inc), bw = pcg_double(st, inc);
        BsdfSample bs;
        if (guided && gd.w > 0.5f) {
            // drawn from the guide by k_guide_infer; the BSDF draws above keep
            // the path's stream aligned with the unguided tracer
            bs.valid = true;
            bs.dir = normalize(D3{double(gd.x), double(gd.y), double(gd.z)});
            bs.eta = 0;
        } else {
            bs = bsdf_sample<MT>(mat, dir_view, v.g, bu0, bu1, bw);
        }
        if (!bs.valid) {
            alive = false;
        } else {
            if (bs.eta != 0) eta_scale /= (bs.eta * bs.eta);
            D3 f = bsdf_eval<MT>(mat, dir_view, bs.dir, v.g);
            double pdf = bsdf_pdf<MT>(mat, dir_view, bs.dir, v.g);
            if (guided) {
                // one-sample MIS over the two strategies (balance heuristic):
                // the path weight is f / mixture pdf
                pdf = (1 - alpha) * pdf +
                      alpha * double(guide::mixture_pdf(lobes, float(bs.dir.x), float(bs.dir.y), float(bs.dir.z)));
                if (f.x == 0 && f.y == 0 && f.z == 0) pdf = 0;  // nothing more can be added
            }
            if (pdf <= 0) {
                alive = false;
            } else {
                P.prev[pid] = make_float4(float(f.x), float(f.y), float(f.z), float(pdf));
                P.ray_o[pid] = make_float4(float(v.pos.x), float(v.pos.y), float(v.pos.z), float(S.eps));
                P.ray_d[pid] = make_float4(float(bs.dir.x), float(bs.dir.y), float(bs.dir.z), __builtin_huge_valf());
                P.depth[pid] = b + 1;
                push_ext = true;
                if (rec && guided) {
                    P.rec_dir[rix] = make_float4(float(bs.dir.x), float(bs.dir.y), float(bs.dir.z), 1.f);
                    P.rec_beta[rix] = make_float4(float(tp.x), float(tp.y), float(tp.z), 0.f);
                }
            }
        }
    }
    P.rng[pid] = st;
    P.tp[pid] = make_float4(float(tp.x), float(tp.y), float(tp.z), float(eta_scale));
    P.rad[pid] = make_float4(float(rad.x), float(rad.y), float(rad.z), 0.f);
}

// MT = -1: the unsorted bucket, with the material type dispatched per path.
template <int MT, bool SPH, bool GUIDE>
__device__ __forceinline__ void shade_queue(const DeviceScene &S, const PathBufs &P, int cur) {
    constexpr int B = MT >= 0 ? MT : kBucketAny;
    int n = P.ctr[C_BUCKET + B];
    const uint32_t *q = P.bucket + size_t(B) * P.capacity;
    int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        int i = base + threadIdx.x;
        bool pe = false, ps = false;
        float4 so, sd, sc;
        uint32_t pid = 0;
        if (i < n) {
            pid = q[i];
            shade_surface<MT, SPH, GUIDE>(S, P, pid, pe, ps, so, sd, sc);
        }
        if (pe) P.alive[pid] = 1;
        int j = wave_append(ps, &P.ctr[C_SH]);
        if (j >= 0) {
            P.shq[j] = pid;
            P.sh_o[j] = so;
            P.sh_d[j] = sd;
            P.sh_c[j] = sc;
        }
    }
}

// The shading kernel of one material type (or, MT = -1, of the runtime-dispatch
// bucket).  Occupancy of these FP64-heavy kernels is set by their register
// count: the 2-wave variant caps it at 256 registers for the material types
// whose code fits there with little or no spilling (Lambertian, RoughPlastic,
// DisneyDiffuse, DisneyMetal, DisneySheen) and for DisneyBSDF, which spills
// there but, running beside the other types' kernels, is faster at two waves
// per SIMD than at one (shading of the 1080p large-mesh render 42.1 -> 40.6
// ms); the others run one wave per SIMD.
template <int MT, bool SPH, bool GUIDE>
__global__ __launch_bounds__(256) void k_shade(DeviceScene S, PathBufs P, int cur) {
    shade_queue<MT, SPH, GUIDE>(S, P, cur);
}
template <int MT, bool SPH, bool GUIDE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_shade_w2(DeviceScene S, PathBufs P,
                                                                                         int cur) {
    shade_queue<MT, SPH, GUIDE>(S, P, cur);
}

// Escaped paths: environment emission with MIS (path_tracing.h:17-28, 284-302).
__global__ __launch_bounds__(256) void k_miss(DeviceScene S, PathBufs P) {
    int n = P.ctr[C_BUCKET + kBucketMiss];
    const uint32_t *q = P.bucket + size_t(kBucketMiss) * P.capacity;
    if (S.envmap_light_id < 0) return;
    const LightRec &env = S.lights[S.envmap_light_id];
    int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t pid = q[i];
        float4 rd = P.ray_d[pid];
        D3 dir = f3d(rd);
        float4 tp4 = P.tp[pid];
        D3 tp{tp4.x, tp4.y, tp4.z};
        float4 r4 = P.rad[pid];
        D3 rad{r4.x, r4.y, r4.z};
        // the direction's lat-long coordinates serve the radiance and the pdf
        // (src/lights/envmap.inl:22-72)
        const D3 ld = xform3(env.to_local, dir);
        double eu, ev;
        env_uv(ld, eu, ev);
        D3 L = env_texel_bilinear(env, eu, ev) * env.scale;
        D3 rad_in = rad;
        int b = P.depth[pid];
        if (b == 0) {
            rad = rad + tp * L;
        } else {
            float4 pv = P.prev[pid];
            D3 f{pv.x, pv.y, pv.z};
            double p2 = pv.w;  // G = 1
            D3 C2 = f * L;
            const double se = sqrt(clampd(1 - ld.y * ld.y, 0, 1));
            const double epdf = se > 0 ? pdf_env_uv(env, eu, ev) / (2 * kPi * kPi * se) : 0.0;
            double p1 = S.light_pmf[S.envmap_light_id] * epdf;
            double w2 = (p2 * p2) / (p1 * p1 + p2 * p2);
            C2 = C2 / p2;
            rad = rad + tp * C2 * w2;
        }
        P.rad[pid] = make_float4(float(rad.x), float(rad.y), float(rad.z), 0.f);
        if (b < P.rec_depth) {
            D3 add = rad - rad_in;
            P.rec_c[size_t(b) * P.rec_stride + pid] = make_float4(float(add.x), float(add.y), float(add.z), 0.f);
        }
    }
}

// Any-hit traversal of the shadow queue, rays assigned to lanes by a grid
// stride; writes whether each is occluded (k_shadow_resolve adds the
// contributions).  A lane whose ray ended takes its next one as soon as
// kShadowRefill lanes of its wave are idle, instead of waiting for the
// wave's longest ray (3.6% faster on the 1080p large-mesh render than
// starting the lanes' rays together).
constexpr int kShadowRefill = 16;
template <bool STATS, bool SPH>
__global__ __launch_bounds__(256) void k_shadow(DeviceScene S, PathBufs P) {
    __shared__ int stack[kStackLds * 256];
    const int n = P.ctr[C_SH];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&P.stat[S_SH_RAYS], (unsigned long long)n);
        for (int b = 0; b < kNumBuckets; b++) P.ctr[C_BUCKET + b] = 0;
    }
    const TraceScene ts{S.qnodes, S.prims, S.spheres};
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    int *spill = P.spill + gtid;
    uint32_t nn = 0, np = 0;
    Tracer<true, STATS, SPH> tr;
    bool busy = false, fin = false;
    // each wave walks its own contiguous share of the queue (neighbouring
    // rays come from neighbouring paths), its idle lanes taking the next
    // rays of it in order
    const int nwaves = stride >> 6;
    const int per = (n + nwaves - 1) / nwaves;
    int cursor = (gtid >> 6) * per;
    const int end = min(n, cursor + per);
    int i = 0;
    while (true) {
        const uint64_t idle = __ballot(!busy);
        const int nidle = __popcll(idle);
        if (nidle >= kShadowRefill) {
            if (fin) P.sh_occ[i] = tr.found ? 1 : 0;
            fin = false;
            const int k = cursor + int(__builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0)));
            if (!busy && k < end) {
                i = k;
                const float4 o = P.sh_o[i], d = P.sh_d[i];
                tr.start(o.x, o.y, o.z, d.x, d.y, d.z, o.w, d.w);
                busy = true;
            }
            cursor = min(end, cursor + nidle);
            if (__ballot(busy) == 0) break;
        }
        if (busy) {
            if (S.num_nodes == 0 || tr.step(ts, stack + threadIdx.x, blockDim.x, spill, stride, nn, np)) {
                busy = false;
                fin = true;
            }
        }
    }
    if (STATS) {
        atomicAdd(&P.stat[S_SH_NODES], (unsigned long long)nn);
        atomicAdd(&P.stat[S_SH_PRIMS], (unsigned long long)np);
    }
}

// After k_shadow: an unoccluded shadow ray adds its pending NEE contribution
// to its path (and to the recorded radiance of its vertex when guiding
// records the pass).
__global__ __launch_bounds__(256) void k_shadow_resolve(PathBufs P) {
    const int n = P.ctr[C_SH];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (P.sh_occ[i]) continue;
        const float4 c = P.sh_c[i];
        const uint32_t bits = uint32_t(__float_as_int(c.w));
        const uint32_t pid = bits & kSlotMask;
        const float4 r = P.rad[pid];
        P.rad[pid] = make_float4(r.x + c.x, r.y + c.y, r.z + c.z, 0.f);
        const int b = int(bits >> kSlotBits);
        if (b < P.rec_depth) {
            float4 *rc = &P.rec_c[size_t(b) * P.rec_stride + pid];
            const float4 o = *rc;
            *rc = make_float4(o.x + c.x, o.y + c.y, o.z + c.z, 0.f);
        }
    }
}

// render.cpp:89-94 (path) and :136-145 (volpath drops non-finite samples).
__global__ void k_accumulate(PathBufs P) {
    int nq = P.pass_nq;
    int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nq; j += stride) {
        SlotInfo si = slot_info(P, j);
        if (!si.valid) continue;
        const int x = si.x, y = si.y;
        double r = 0, g = 0, b = 0;
        for (int k = 0; k < P.pass_ns; k++) {
            float4 v = P.rad[k * nq + j];
            if (P.volpath && !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z))) continue;
            r += v.x;
            g += v.y;
            b += v.z;
        }
        size_t o = (size_t(y) * P.W + x) * 3;
        P.accum[o] += r;
        P.accum[o + 1] += g;
        P.accum[o + 2] += b;
    }
}

// Batched renders: per-slot atomic sums into each job's image.
__global__ void k_accumulate_batch(PathBufs P, int n) {
    int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        SlotInfo si = slot_info(P, i);
        if (!si.valid) continue;
        float4 v = P.rad[i];
        if (P.volpath && !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z))) continue;
        double *a = P.jobs[si.job].accum + (size_t(si.y) * si.W + si.x) * 3;
        atomicAdd(a, double(v.x));
        atomicAdd(a + 1, double(v.y));
        atomicAdd(a + 2, double(v.z));
    }
}

__global__ void k_job_clear(const BatchJob *jobs, int j) {
    const BatchJob &J = jobs[j];
    size_t n = size_t(J.W) * J.H * 3;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) J.accum[i] = 0.0;
}

__global__ void k_job_resolve(const BatchJob *jobs, int j) {
    const BatchJob &J = jobs[j];
    size_t n = size_t(J.W) * J.H * 3;
    double inv = 1.0 / double(J.spp);
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) J.out[i] = float(J.accum[i] * inv);
}

// Auxiliary integrators (src/render.cpp:12-69): one ray through each pixel
// centre; depth, shading normal, mean curvature, the camera ray
// differential (radius, spread) or the mipmap level (none with constant
// textures: black).  Misses are black.
template <bool SPH>
__global__ __launch_bounds__(256) void k_aux(DeviceScene S, PathBufs P, int mode, float *out) {
    __shared__ int stack[kStackLds * 256];
    const int n = P.W * P.H;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    int *spill = P.spill + gtid;
    uint32_t nn = 0, np = 0;
    for (int i = gtid; i < n; i += stride) {
        int x = i % P.W, y = i / P.W;
        D3 o, d;
        sample_primary(S.camera, (x + 0.5) / P.W, (y + 0.5) / P.H, o, d);
        TraceHit h;
        bool hit = S.num_nodes > 0 &&
                   trace_ray<false, false, SPH>(TraceScene{S.qnodes, S.prims, S.spheres}, float(o.x), float(o.y),
                                                float(o.z), float(d.x), float(d.y), float(d.z), 0.f,
                                                __builtin_huge_valf(), stack + threadIdx.x, blockDim.x, spill, stride,
                                                h, nn, np);
        D3 c{0, 0, 0};
        if (hit) {
            Vertex v = make_vertex<SPH>(S, h.slot, h.t, h.u, h.v, o, d);
            if (mode == 0) {
                double dist = sqrt(dist2(v.pos, o));
                c = D3{dist, dist, dist};
            } else if (mode == 1) {
                c = v.g.sf.n;
            } else if (mode == 2) {
                double k = vertex_mean_curvature(S, h.slot, double(h.u), double(h.v), v);
                c = D3{k, k, k};
            } else if (mode == 3) {
                c = D3{0, 0.25 / double(max(P.W, P.H)), 0};  // init_ray_differential (src/ray.h:35-37)
            } else if (mode == 4 && v.material_id >= 0) {
                // the level the camera ray's footprint selects in the material's first colour
                // texture when it is an image (render.cpp:44-60; get_texture, material.cpp:121-123)
                const MatRec &m = S.materials[v.material_id];
                if (m.type != 6 && m.tex[0] && S.textures[m.tex[0] - 1].type == 1) {
                    const TexRec &t = S.textures[m.tex[0] - 1];
                    double fp = (0.25 / double(max(P.W, P.H))) * sqrt(dist2(v.pos, o)) / v.inv_uv;
                    double sf = max(t.w[0], t.h[0]) * fmax(t.uscale, t.vscale) * fp;
                    double level = log2(sf < double(1e-8f) ? double(1e-8f) : sf);
                    c = D3{level, level, level};
                }
            }
        }
        out[3 * size_t(i)] = float(c.x);
        out[3 * size_t(i) + 1] = float(c.y);
        out[3 * size_t(i) + 2] = float(c.z);
    }
}

__global__ void k_resolve(const double *accum, float *out, size_t n, double inv_spp) {
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = float(accum[i] * inv_spp);
}

}  // namespace ljx

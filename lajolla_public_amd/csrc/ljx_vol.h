// Volumetric path tracing on the wavefront loop (homogeneous and
// heterogeneous media).
//
// lajolla's vol_path_tracing is a homework stub returning zero
// (src/vol_path_tracing.h:59-64); this follows the handout's final
// algorithm (handouts/homework2.tex): homogenised free flight with one-sample
// spectral MIS over RGB (:713-758), next-event estimation with ratio tracking
// through index-matched surfaces (:770-812, :460-507), update_medium
// (:402-409), the MIS caches dir_pdf / nee_p_cache / multi_trans_pdf /
// never_scatter (:505-557), Russian roulette from rr_depth (:331-400), and the
// max_null_collisions cap (:760-764).
//
//   k_vol_flight  free flight (delta tracking) in scenes with grid media
//   k_vol_shade   free flight in homogeneous media, emission with MIS,
//                 index-matched pass-through, phase or BSDF sampling, NEE
//                 light-point sampling, RR
//   k_vol_shadow  the NEE shadow walk: closest hits through index-matched
//                 surfaces with ratio tracking in every medium on the way
//
// All NEE draws come from a sub-stream seeded by one draw of the path stream
// (stream id path_stream ^ 0x5bd1e9955bd1e995), so the shadow kernel can
// consume them after the shade kernel has moved on.
#pragma once

#include "ljx_kernels.h"

#include <vector>

namespace ljx {

constexpr bool kVolpathEnabled = true;
constexpr uint64_t kNeeStreamXor = 0x5bd1e9955bd1e995ULL;
constexpr int kMaxShadowSegments = 256;

__global__ void k_reset_stats(unsigned long long *stat) {
    if (threadIdx.x < S_NUM) stat[threadIdx.x] = 0;
}

// sigma_t = sigma_s + sigma_a at p
__device__ __forceinline__ D3 sigma_t_at(const MediumRec &m, D3 p) {
    D3 s, a;
    medium_sigmas(m, p, s, a);
    return s + a;
}
__device__ __forceinline__ D3 sigma_s_at(const MediumRec &m, D3 p) {
    D3 s, a;
    medium_sigmas(m, p, s, a);
    return s;
}

// exp(-maj t) per channel.  A grey majorant (every channel equal: grid
// media, whose sigma_t is the density alone) gives three identical
// exponentials, computed once.
__device__ __forceinline__ D3 maj_exp(D3 maj, bool grey, double t) {
    if (grey) {
        const double e = exp(-1.0 * maj.x * t);
        return D3{e, e, e};
    }
    return vexp(-1.0 * maj * t);
}

// update_medium (homework2.tex:402-409)
__device__ __forceinline__ int update_medium(D3 dir, const Vertex &v, int medium) {
    if (v.interior_medium_id != v.exterior_medium_id)
        return dot(dir, v.g.ng) > 0 ? v.exterior_medium_id : v.interior_medium_id;
    return medium;
}

// Free-flight sampling of one path in a medium (homework2.tex:745-769, delta
// tracking; its first draw picks the channel).  It is the first use of the
// path's stream in a bounce, so in scenes with grid media (long loops of
// null collisions) k_vol_flight runs it for every path of the queue before
// k_vol_shade, which takes over the stream state; otherwise k_vol_shade runs
// it in place.
struct Flight {
    D3 T, trans_dir_pdf;
    double accum_t;
    bool scatter;
};
// The loop of free_flight, resumable one collision at a time.
struct FlightRun {
    const MediumRec *m;
    D3 org, dir, maj;
    double t_hit, max_maj;
    int ch, it;
    bool grey;
    Flight fl;
};
__device__ __forceinline__ void flight_begin(FlightRun &r, const DeviceScene &S, int medium, D3 org, D3 dir,
                                             double t_hit, uint64_t &st, uint64_t inc) {
    r.fl = Flight{{1, 1, 1}, {1, 1, 1}, 0.0, false};
    r.m = &S.media[medium];
    r.org = org;
    r.dir = dir;
    r.t_hit = t_hit;
    r.maj = medium_majorant(*r.m, org, dir, __builtin_huge_val());
    r.max_maj = maxc(r.maj);
    r.grey = r.maj.x == r.maj.y && r.maj.y == r.maj.z;
    const double u = pcg_double(st, inc);
    r.ch = min(max(int(u * 3), 0), 2);
    r.it = 0;
}
// One pass of the loop; true when the flight has ended.
__device__ __forceinline__ bool flight_step(FlightRun &r, const PathBufs &P, uint64_t &st, uint64_t inc) {
    Flight &fl = r.fl;
    const D3 maj = r.maj;
    const double max_maj = r.max_maj;
    const int ch = r.ch;
    const double mc = comp(maj, ch);
    if (mc <= 0) return true;
    if (r.it >= P.max_null) return true;
    const double t = -log(1 - pcg_double(st, inc)) / mc;
    const double dt = r.t_hit - fl.accum_t;
    fl.accum_t = fmin(fl.accum_t + t, r.t_hit);
    if (t < dt) {
        D3 sig_t = sigma_t_at(*r.m, r.org + fl.accum_t * r.dir);
        D3 real_prob = sig_t / maj;
        D3 e = maj_exp(maj, r.grey, t);
        if (pcg_double(st, inc) < comp(real_prob, ch)) {
            fl.scatter = true;
            fl.T = fl.T * e / max_maj;
            fl.trans_dir_pdf = fl.trans_dir_pdf * e * maj * real_prob / max_maj;
            return true;
        }
        D3 sig_n = maj - sig_t;
        fl.T = fl.T * e * sig_n / max_maj;
        fl.trans_dir_pdf = fl.trans_dir_pdf * e * maj * (1.0 - real_prob) / max_maj;
        r.it++;
        return false;
    }
    D3 e = maj_exp(maj, r.grey, dt);
    fl.T = fl.T * e;
    fl.trans_dir_pdf = fl.trans_dir_pdf * e;
    return true;
}
__device__ __forceinline__ Flight free_flight(const DeviceScene &S, const PathBufs &P, int medium, D3 org, D3 dir,
                                              double t_hit, uint64_t &st, uint64_t inc) {
    FlightRun r;
    flight_begin(r, S, medium, org, dir, t_hit, st, inc);
    while (!flight_step(r, P, st, inc)) {
    }
    return r.fl;
}

// Distance to the surface a ray hit (infinite on a miss), from the vertex
// position the shading kernel computes.
template <bool SPH>
__device__ __forceinline__ double hit_distance(const DeviceScene &S, float4 h, D3 org, D3 dir) {
    const int slot = __float_as_int(h.w);
    if (slot < 0) return __builtin_huge_val();
    Vertex v = make_vertex<SPH>(S, slot, h.x, h.y, h.z, org, dir);
    return sqrt(dist2(org, v.pos));
}

// The paths of the volumetric queue that are in a medium, for k_vol_flight
// (the order does not matter: each path's flight uses its own stream).  Each
// block of 1,024 threads appends with one atomic per round (a wave per
// atomic contended on the one counter).
__global__ __launch_bounds__(1024) void k_vol_flight_queue(PathBufs P) {
    __shared__ int wcount[16];
    __shared__ int bbase;
    const int n = P.ctr[C_BUCKET + 0];
    const int stride = gridDim.x * blockDim.x;
    const int w = threadIdx.x >> 6;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const int i = base + threadIdx.x;
        uint32_t pid = 0;
        bool in_medium = false;
        if (i < n) {
            pid = P.bucket[i];
            in_medium = P.medium[pid] >= 0;
        }
        const uint64_t mask = __ballot(in_medium);
        if ((threadIdx.x & 63) == 0) wcount[w] = __popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            int total = 0;
            for (int k = 0; k < 16; k++) total += wcount[k];
            bbase = total > 0 ? atomicAdd(&P.ctr[C_FLQ], total) : 0;
        }
        __syncthreads();
        if (in_medium) {
            int k = bbase + int(__builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0)));
            for (int q = 0; q < w; q++) k += wcount[q];
            P.flq[k] = pid;
        }
        __syncthreads();
    }
}

// Free flight for every path of the flight queue, in a kernel small enough
// for three waves per SIMD.  The loops' lengths vary widely between paths, so
// a lane whose flight ended takes its next path as soon as kFlightRefill
// lanes of its wave are idle (one atomic per wave for all of them) rather
// than waiting for the wave's longest flight.
constexpr int kFlightRefill = 16;
template <bool SPH>
__global__ __launch_bounds__(256) void k_vol_flight(DeviceScene S, PathBufs P) {
    const int n = P.ctr[C_FLQ];
    int *fetch = &P.ctr[C_FLQ + 1];
    const int lane = __lane_id();
    FlightRun r;
    uint64_t st = 0, inc = 0;
    uint32_t pid = 0;
    bool busy = false, fin = false, more = true;  // more: wave-uniform
    while (true) {
        const uint64_t idle = __ballot(!busy);
        const int nidle = __popcll(idle);
        if ((more && nidle >= kFlightRefill) || nidle == 64) {
            if (fin) {
                P.rng[pid] = st;
                P.vf_t[pid] = make_double4(r.fl.T.x, r.fl.T.y, r.fl.T.z, r.fl.accum_t);
                P.vf_p[pid] = make_double4(r.fl.trans_dir_pdf.x, r.fl.trans_dir_pdf.y, r.fl.trans_dir_pdf.z,
                                           r.fl.scatter ? 1.0 : 0.0);
                fin = false;
            }
            if (!more) break;  // every lane idle and the queue drained
            int base = 0;
            if (lane == 0) base = atomicAdd(fetch, nidle);
            base = __shfl(base, 0);
            if (base + nidle >= n) more = false;
            const int i = base + int(__builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32),
                                                               __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0)));
            if (!busy && i < n) {
                pid = P.flq[i];
                const float4 ro = P.ray_o[pid], rd = P.ray_d[pid];
                const D3 org = f3d(ro), dir = f3d(rd);
                st = P.rng[pid];
                inc = pcg_inc(slot_info(P, int(pid)).stream());
                flight_begin(r, S, P.medium[pid], org, dir, hit_distance<SPH>(S, P.hit[pid], org, dir), st, inc);
                busy = true;
            }
        }
        if (busy && flight_step(r, P, st, inc)) {
            busy = false;
            fin = true;
        }
    }
}

template <bool SPH, int DMT>
__device__ __forceinline__ void vol_shade(const DeviceScene &S, const PathBufs &P, uint32_t pid, bool &push_ext,
                                          bool &push_sh, float4 &sh_o, float4 &sh_d, float4 &sh_c, float4 &sh_x) {
    float4 h = P.hit[pid];
    float4 ro = P.ray_o[pid], rd = P.ray_d[pid];
    D3 org = f3d(ro), dir = f3d(rd);
    int bounces = P.depth[pid];
    uint64_t st = P.rng[pid];
    const uint64_t stream = slot_info(P, int(pid)).stream();
    uint64_t inc = pcg_inc(stream);
    float4 tp4 = P.tp[pid];
    D3 tp{tp4.x, tp4.y, tp4.z};
    double eta_scale = tp4.w;
    float4 r4 = P.rad[pid];
    D3 rad{r4.x, r4.y, r4.z};
    int medium = P.medium[pid];
    float4 nc = P.nee_c[pid];
    D3 nee_p_cache{nc.x, nc.y, nc.z};
    double dir_pdf = nc.w;
    float4 mt = P.mtp[pid];
    D3 multi_trans_pdf{mt.x, mt.y, mt.z};
    bool never_scatter = mt.w != 0.f;
    push_ext = false;
    push_sh = false;

    int slot = __float_as_int(h.w);
    bool hit = slot >= 0;
    Vertex v;
    if (hit) v = make_vertex<SPH>(S, slot, h.x, h.y, h.z, org, dir);
    bool scatter = false;
    bool alive = true;
    D3 p_scatter = org;
    if (medium >= 0) {
        D3 T, trans_dir_pdf;
        double accum_t;
        if (P.vol_flight) {
            // the free flight k_vol_flight sampled (it advanced the stream state)
            const double4 ft = P.vf_t[pid], fp = P.vf_p[pid];
            T = D3{ft.x, ft.y, ft.z};
            trans_dir_pdf = D3{fp.x, fp.y, fp.z};
            accum_t = ft.w;
            scatter = fp.w != 0.0;
        } else {
            const Flight fl = free_flight(S, P, medium, org, dir, hit ? sqrt(dist2(org, v.pos)) : __builtin_huge_val(),
                                          st, inc);
            T = fl.T;
            trans_dir_pdf = fl.trans_dir_pdf;
            accum_t = fl.accum_t;
            scatter = fl.scatter;
        }
        if (scatter) never_scatter = false;
        if (isfinite(accum_t)) p_scatter = org + accum_t * dir;
        multi_trans_pdf = multi_trans_pdf * trans_dir_pdf;
        tp = tp * T / avgc(trans_dir_pdf);
    }
    if (!scatter && !hit) alive = false;
    if (alive && !scatter && v.area_light_id >= 0) {
        const LightRec &l = S.lights[v.area_light_id];
        LightPoint lp{v.pos, v.g.ng};
        D3 Le = light_emission(S, l, -dir, lp);
        if (never_scatter) {
            rad = rad + tp * Le;
        } else {
            double pdf_nee = S.light_pmf[v.area_light_id] * pdf_light_point(S, l, lp, nee_p_cache);
            double G = fabs(dot(dir, v.g.ng)) / dist2(nee_p_cache, v.pos);
            double dp = dir_pdf * avgc(multi_trans_pdf) * G;
            double w = (dp * dp) / (dp * dp + pdf_nee * pdf_nee);
            rad = rad + tp * Le * w;
        }
    }
    if (alive && P.max_depth != -1 && bounces == P.max_depth - 1) alive = false;
    bool passed = false;
    if (alive && !scatter && v.material_id == -1) {
        // index-matched interface: continue straight through
        medium = update_medium(dir, v, medium);
        P.ray_o[pid] = make_float4(float(v.pos.x), float(v.pos.y), float(v.pos.z), float(S.eps));
        P.ray_d[pid] = make_float4(float(dir.x), float(dir.y), float(dir.z), __builtin_huge_valf());
        bounces++;
        push_ext = true;
        passed = true;
    }
    if (alive && !passed) {
        D3 dir_view = -dir;
        // NEE sub-stream
        uint32_t k = pcg_next(st, inc);
        uint64_t nstream = stream ^ kNeeStreamXor;
        uint64_t ninc = pcg_inc(nstream);
        uint64_t nst = pcg_init_state(nstream, uint64_t(k));
        D3 pos = scatter ? p_scatter : v.pos;
        double lu0 = pcg_double(nst, ninc), lu1 = pcg_double(nst, ninc);
        double lw = pcg_double(nst, ninc), sw = pcg_double(nst, ninc);
        if (S.num_lights > 0) {
            int lid = sample_cdf_guided(S.light_cdf, S.num_lights, lw, S.light_guide, S.light_guide_m);
            const LightRec &l = S.lights[lid];
            LightPoint lp = sample_light_point(S, l, pos, lu0, lu1, sw);
            D3 dir_light = normalize(lp.pos - pos);
            double G = fmax(-dot(dir_light, lp.n), 0.0) / dist2(pos, lp.pos);
            double pdf_nee = S.light_pmf[lid] * pdf_light_point(S, l, lp, pos);
            if (G > 0 && pdf_nee > 0) {
                D3 f;
                double pdf_f;
                D3 pre = tp;
                if (scatter) {
                    const MediumRec &m = S.media[medium];
                    double pv = phase_value(m, dir_view, dir_light);
                    f = D3{pv, pv, pv};
                    pdf_f = pv;
                    pre = pre * sigma_s_at(m, p_scatter);
                } else {
                    MatRec mat = S.materials[v.material_id];
                    if (mat.textured) resolve_textures(S, mat, v.tu, v.tv, 0.0);
                    f = bsdf_eval<DMT>(mat, dir_view, dir_light, v.g);
                    pdf_f = bsdf_pdf<DMT>(mat, dir_view, dir_light, v.g);
                }
                D3 L = light_emission(S, l, -dir_light, lp);
                D3 K = pre * G * f * L;
                if (K.x != 0 || K.y != 0 || K.z != 0) {
                    push_sh = true;
                    sh_o = make_float4(float(pos.x), float(pos.y), float(pos.z), __int_as_float(medium));
                    sh_d = make_float4(float(lp.pos.x), float(lp.pos.y), float(lp.pos.z), __int_as_float(bounces));
                    sh_c = make_float4(float(K.x), float(K.y), float(K.z), __int_as_float(int(pid)));
                    sh_x = make_float4(float(pdf_nee), float(pdf_f * G), __uint_as_float(uint32_t(nst)),
                                       __uint_as_float(uint32_t(nst >> 32)));
                }
            }
        }
        if (scatter) {
            const MediumRec &m = S.media[medium];
            double u0 = pcg_double(st, inc), u1 = pcg_double(st, inc);
            D3 nd = phase_sample(m, dir_view, u0, u1);
            double pv = phase_value(m, dir_view, nd);
            if (pv <= 0) {
                alive = false;
            } else {
                tp = tp * (D3{pv, pv, pv} / pv) * sigma_s_at(m, p_scatter);
                dir_pdf = pv;
                nee_p_cache = p_scatter;
                multi_trans_pdf = D3{1, 1, 1};
                P.ray_o[pid] = make_float4(float(p_scatter.x), float(p_scatter.y), float(p_scatter.z), 0.f);
                P.ray_d[pid] = make_float4(float(nd.x), float(nd.y), float(nd.z), __builtin_huge_valf());
            }
        } else {
            MatRec mat = S.materials[v.material_id];
            if (mat.textured) resolve_textures(S, mat, v.tu, v.tv, 0.0);
            double u0 = pcg_double(st, inc), u1 = pcg_double(st, inc), w = pcg_double(st, inc);
            BsdfSample bs = bsdf_sample<DMT>(mat, dir_view, v.g, u0, u1, w);
            if (!bs.valid) {
                alive = false;
            } else {
                if (bs.eta != 0) eta_scale /= (bs.eta * bs.eta);
                D3 f = bsdf_eval<DMT>(mat, dir_view, bs.dir, v.g);
                double pdf = bsdf_pdf<DMT>(mat, dir_view, bs.dir, v.g);
                if (pdf <= 0) {
                    alive = false;
                } else {
                    tp = tp * f / pdf;
                    dir_pdf = pdf;
                    nee_p_cache = v.pos;
                    multi_trans_pdf = D3{1, 1, 1};
                    never_scatter = false;
                    medium = update_medium(bs.dir, v, medium);
                    P.ray_o[pid] = make_float4(float(v.pos.x), float(v.pos.y), float(v.pos.z), float(S.eps));
                    P.ray_d[pid] = make_float4(float(bs.dir.x), float(bs.dir.y), float(bs.dir.z), __builtin_huge_valf());
                }
            }
        }
        if (alive && bounces >= P.rr_depth) {
            double rr = fmin(maxc((1 / eta_scale) * tp), 0.95);
            if (pcg_double(st, inc) > rr) alive = false;
            else tp = tp / rr;
        }
        if (alive) {
            bounces++;
            push_ext = true;
        }
    }
    P.depth[pid] = bounces;
    P.rng[pid] = st;
    P.medium[pid] = medium;
    P.tp[pid] = make_float4(float(tp.x), float(tp.y), float(tp.z), float(eta_scale));
    P.rad[pid] = make_float4(float(rad.x), float(rad.y), float(rad.z), 0.f);
    P.nee_c[pid] = make_float4(float(nee_p_cache.x), float(nee_p_cache.y), float(nee_p_cache.z), float(dir_pdf));
    P.mtp[pid] = make_float4(float(multi_trans_pdf.x), float(multi_trans_pdf.y), float(multi_trans_pdf.z),
                             never_scatter ? 1.f : 0.f);
}

template <bool SPH, int DMT>
__global__ __launch_bounds__(256) void k_vol_shade(DeviceScene S, PathBufs P, int cur) {
    int n = P.ctr[C_BUCKET + 0];
    const uint32_t *q = P.bucket;
    int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        int i = base + threadIdx.x;
        bool pe = false, ps = false;
        float4 so, sd, sc, sx;
        uint32_t pid = 0;
        if (i < n) {
            pid = q[i];
            vol_shade<SPH, DMT>(S, P, pid, pe, ps, so, sd, sc, sx);
        }
        if (pe) P.alive[pid] = 1;
        int j = wave_append(ps, &P.ctr[C_SH]);
        if (j >= 0) {
            P.shq[j] = pid;
            P.sh_o[j] = so;
            P.sh_d[j] = sd;
            P.sh_c[j] = sc;
            P.sh_x[j] = sx;
        }
    }
}

// NEE segments, first pass: one any-hit query per segment that passes
// through index-matched boundaries.  A surface with a material anywhere on
// the way to the light blocks the sample whatever the media, so this
// settles most segments before the FP64 walk: sh_occ = 1 when such a surface
// blocks the segment, 2 when it crosses index-matched boundaries only (the
// walk takes closest hits through them), 0 when nothing lies on it (the
// walk's first closest-hit query, over the same interval, would find
// nothing).  Traversal only, so it runs at the any-hit kernel's occupancy
// and with its lane refill; the walk's FP64 state would cut that to two
// waves per SIMD.  Triangle-only scenes cap it at eight waves per SIMD (64
// registers, a 16-byte spill).
template <bool STATS, bool SPH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SPH ? 1 : 8))) void k_vol_occlude(DeviceScene S, PathBufs P) {
    __shared__ int stack[kStackLds * 256];
    const int n = P.ctr[C_SH];
    const TraceScene ts{S.qnodes, S.prims, S.spheres, S.slot_passthru};
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    int *spill = P.spill + gtid;
    uint32_t nn = 0, np = 0;
    unsigned long long segs = 0;
    Tracer<true, STATS, SPH, true> tr;
    bool busy = false, fin = false;
    // each wave walks its own contiguous share of the queue (k_shadow)
    const int nwaves = stride >> 6;
    const int per = (n + nwaves - 1) / nwaves;
    int cursor = (gtid >> 6) * per;
    const int end = min(n, cursor + per);
    int i = 0;
    while (true) {
        const uint64_t idle = __ballot(!busy);
        const int nidle = __popcll(idle);
        if (nidle >= kShadowRefill) {
            if (fin) P.sh_occ[i] = tr.found ? 1 : (tr.passed ? 2 : 0);
            fin = false;
            const int k = cursor + int(__builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0)));
            cursor = min(end, cursor + nidle);
            if (!busy && k < end) {
                i = k;
                const float4 so = P.sh_o[i], sd = P.sh_d[i];
                const D3 p{so.x, so.y, so.z}, pp{sd.x, sd.y, sd.z};
                const D3 dir_light = normalize(pp - p);
                const double dl = sqrt(dist2(p, pp));
                tr.start(float(p.x), float(p.y), float(p.z), float(dir_light.x), float(dir_light.y),
                         float(dir_light.z), float(S.eps), float((1 - S.eps) * dl));
                busy = true;
                if (S.num_nodes > 0) segs++;
            }
            if (__ballot(busy) == 0) break;
        }
        if (busy) {
            if (S.num_nodes == 0 || tr.step(ts, stack + threadIdx.x, blockDim.x, spill, stride, nn, np)) {
                busy = false;
                fin = true;
            }
        }
    }
    atomicAdd(&P.stat[S_SH_RAYS], segs);
    if (STATS) {
        atomicAdd(&P.stat[S_SH_NODES], (unsigned long long)nn);
        atomicAdd(&P.stat[S_SH_PRIMS], (unsigned long long)np);
    }
}

// The NEE shadow walk with ratio tracking (homework2.tex:770-812), after
// k_vol_occlude: blocked segments are skipped, and only segments that cross
// index-matched boundaries take closest hits.
template <bool STATS, bool SPH>
__global__ __launch_bounds__(256) void k_vol_shadow(DeviceScene S, PathBufs P) {
    __shared__ int stack[kStackLds * 256];
    int n = P.ctr[C_SH];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int b = 0; b < kNumBuckets; b++) P.ctr[C_BUCKET + b] = 0;
    }
    int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    int stride = gridDim.x * blockDim.x;
    int *spill = P.spill + gtid;
    uint32_t nn = 0, np = 0;
    unsigned long long segs = 0;
    for (int i = gtid; i < n; i += stride) {
        float4 so = P.sh_o[i], sd = P.sh_d[i], sc = P.sh_c[i], sx = P.sh_x[i];
        int pid = __float_as_int(sc.w);
        int medium = __float_as_int(so.w);
        int bounces = __float_as_int(sd.w);
        D3 p{so.x, so.y, so.z}, pp{sd.x, sd.y, sd.z};
        uint64_t nst = uint64_t(__float_as_uint(sx.z)) | (uint64_t(__float_as_uint(sx.w)) << 32);
        const uint64_t nstream = slot_info(P, pid).stream() ^ kNeeStreamXor;
        uint64_t ninc = pcg_inc(nstream);
        D3 dir_light = normalize(pp - p);
        D3 T{1, 1, 1}, p_nee{1, 1, 1}, p_dir{1, 1, 1};
        D3 p0 = p;
        bool blocked = false;
        int shadow_bounces = 0;
        // k_vol_occlude's verdict (the NEE stream is not written back, so
        // skipping a blocked segment's walk leaves every other draw unchanged)
        bool crossings = false;
        if (S.num_nodes > 0) {
            const uint8_t occ = P.sh_occ[i];
            if (occ == 1) continue;
            crossings = occ == 2;
        }
        for (int seg = 0; seg < kMaxShadowSegments; seg++) {
            double dl = sqrt(dist2(p0, pp));
            TraceHit h;
            if (crossings) segs++;
            bool hit = crossings &&
                       trace_ray<false, STATS, SPH>(TraceScene{S.qnodes, S.prims, S.spheres}, float(p0.x), float(p0.y),
                                                    float(p0.z), float(dir_light.x), float(dir_light.y),
                                                    float(dir_light.z), float(S.eps), float((1 - S.eps) * dl),
                                                    stack + threadIdx.x, blockDim.x, spill, stride, h, nn, np);
            Vertex v;
            double next_t = dl;
            if (hit) {
                v = make_vertex<SPH>(S, h.slot, h.t, h.u, h.v, p0, dir_light);
                next_t = sqrt(dist2(p0, v.pos));
            }
            if (medium >= 0) {
                const MediumRec &m = S.media[medium];
                D3 maj = medium_majorant(m, p0, dir_light, (1 - S.eps) * dl);
                double max_maj = maxc(maj);
                const bool grey = maj.x == maj.y && maj.y == maj.z;
                double u = pcg_double(nst, ninc);
                int ch = min(max(int(u * 3), 0), 2);
                double accum_t = 0;
                for (int it = 0;; it++) {
                    double mc = comp(maj, ch);
                    if (mc <= 0) break;
                    if (it >= P.max_null) break;
                    double t = -log(1 - pcg_double(nst, ninc)) / mc;
                    double dt = next_t - accum_t;
                    accum_t = fmin(accum_t + t, next_t);
                    if (t < dt) {
                        D3 sig_t = sigma_t_at(m, p0 + accum_t * dir_light);
                        D3 sig_n = maj - sig_t;
                        D3 e = maj_exp(maj, grey, t);
                        T = T * e * sig_n / max_maj;
                        p_nee = p_nee * e * maj / max_maj;
                        D3 real_prob = sig_t / maj;
                        p_dir = p_dir * e * maj * (1.0 - real_prob) / max_maj;
                        if (maxc(T) <= 0) break;
                    } else {
                        D3 e = maj_exp(maj, grey, dt);
                        T = T * e;
                        p_nee = p_nee * e;
                        p_dir = p_dir * e;
                        break;
                    }
                }
            }
            if (!hit) break;
            if (v.material_id >= 0) {
                blocked = true;
                break;
            }
            shadow_bounces++;
            if (P.max_depth != -1 && bounces + shadow_bounces + 1 >= P.max_depth) {
                blocked = true;
                break;
            }
            medium = update_medium(dir_light, v, medium);
            p0 = p0 + next_t * dir_light;
            if (seg == kMaxShadowSegments - 1) blocked = true;
        }
        if (!blocked && maxc(T) > 0) {
            double pdf_nee = double(sx.x) * avgc(p_nee);
            if (pdf_nee > 0) {
                double pdf_dir = double(sx.y) * avgc(p_dir);
                double w = (pdf_nee * pdf_nee) / (pdf_nee * pdf_nee + pdf_dir * pdf_dir);
                D3 c = D3{sc.x, sc.y, sc.z} * T * (w / pdf_nee);
                float4 r = P.rad[pid];
                P.rad[pid] = make_float4(float(r.x + c.x), float(r.y + c.y), float(r.z + c.z), 0.f);
            }
        }
    }
    atomicAdd(&P.stat[S_SH_RAYS], segs);
    if (STATS) {
        atomicAdd(&P.stat[S_SH_NODES], (unsigned long long)nn);
        atomicAdd(&P.stat[S_SH_PRIMS], (unsigned long long)np);
    }
}

// Volpath-specific path state initialisation after raygen.
__global__ void k_vol_init(DeviceScene S, PathBufs P) {
    int n = P.pass_ns * P.pass_nq;
    int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        P.medium[i] = P.jobs ? P.jobs[slot_info(P, i).job].camera.medium_id : S.camera.medium_id;
        float4 o = P.ray_o[i];
        P.nee_c[i] = make_float4(o.x, o.y, o.z, 0.f);
        P.mtp[i] = make_float4(1.f, 1.f, 1.f, 1.f);
    }
}

template <class Alloc, class Bufs>
void vol_alloc(bool, size_t capacity, PathBufs &P, Bufs &bufs, Alloc alloc) {
    auto take = [&](size_t bytes) {
        void *p = alloc(bytes);
        bufs.push_back({p, bytes});
        return p;
    };
    P.medium = static_cast<int *>(take(capacity * sizeof(int)));
    P.nee_c = static_cast<float4 *>(take(capacity * sizeof(float4)));
    P.mtp = static_cast<float4 *>(take(capacity * sizeof(float4)));
    P.sh_x = static_cast<float4 *>(take(capacity * sizeof(float4)));
    P.vf_t = static_cast<double4 *>(take(capacity * sizeof(double4)));
    P.vf_p = static_cast<double4 *>(take(capacity * sizeof(double4)));
    P.flq = static_cast<uint32_t *>(take(capacity * sizeof(uint32_t)));
}

inline void vol_launch_check() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("volpath kernel launch: ") + hipGetErrorString(e));
}

// basic: the scene uses no material types beyond lajolla's three basic ones.
inline void vol_shade_all(const DeviceScene &S, PathBufs &P, int cur, int grid, hipStream_t s, bool sph, bool basic) {
    if (P.vol_flight) {
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, P.ctr, C_FLQ, 2);
        hipLaunchKernelGGL(k_vol_flight_queue, dim3(std::max(1, grid / 4)), dim3(1024), 0, s, P);
        if (sph) hipLaunchKernelGGL(k_vol_flight<true>, dim3(grid), dim3(256), 0, s, S, P);
        else hipLaunchKernelGGL(k_vol_flight<false>, dim3(grid), dim3(256), 0, s, S, P);
        vol_launch_check();
    }
    if (basic) {
        if (sph) hipLaunchKernelGGL((k_vol_shade<true, kMtBasic>), dim3(grid), dim3(256), 0, s, S, P, cur);
        else hipLaunchKernelGGL((k_vol_shade<false, kMtBasic>), dim3(grid), dim3(256), 0, s, S, P, cur);
    } else {
        if (sph) hipLaunchKernelGGL((k_vol_shade<true, -1>), dim3(grid), dim3(256), 0, s, S, P, cur);
        else hipLaunchKernelGGL((k_vol_shade<false, -1>), dim3(grid), dim3(256), 0, s, S, P, cur);
    }
    vol_launch_check();
}

inline void vol_shadow(const DeviceScene &S, PathBufs &P, int grid, hipStream_t s, bool stats, bool sph) {
    if (stats) {
        if (sph) hipLaunchKernelGGL((k_vol_occlude<true, true>), dim3(grid), dim3(256), 0, s, S, P);
        else hipLaunchKernelGGL((k_vol_occlude<true, false>), dim3(grid), dim3(256), 0, s, S, P);
    } else {
        if (sph) hipLaunchKernelGGL((k_vol_occlude<false, true>), dim3(grid), dim3(256), 0, s, S, P);
        else hipLaunchKernelGGL((k_vol_occlude<false, false>), dim3(grid), dim3(256), 0, s, S, P);
    }
    vol_launch_check();
    if (stats) {
        if (sph) hipLaunchKernelGGL((k_vol_shadow<true, true>), dim3(grid), dim3(256), 0, s, S, P);
        else hipLaunchKernelGGL((k_vol_shadow<true, false>), dim3(grid), dim3(256), 0, s, S, P);
    } else {
        if (sph) hipLaunchKernelGGL((k_vol_shadow<false, true>), dim3(grid), dim3(256), 0, s, S, P);
        else hipLaunchKernelGGL((k_vol_shadow<false, false>), dim3(grid), dim3(256), 0, s, S, P);
    }
    vol_launch_check();
}

}  // namespace ljx

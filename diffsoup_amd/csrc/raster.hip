// Tile-binned rasterizer with stochastic opacity masking and a packed
// depth|triangle-id test, plus the opacity auxiliary gradient that re-walks
// the same bins.
//
// Pipeline per launch (all on one stream, no host synchronisation):
//   k_clear   zero the per-tile counters and the stats block
//   k_setup   per (view, triangle): cull, conservative rectangle, record,
//             count the tiles the rectangle touches
//   k_scan    exclusive scan of the per-tile counts (one workgroup)
//   k_emit    per (view, triangle): claim a slot in each touched tile's bin
//   k_resolve per tile: stage bin records in LDS, each lane owns one pixel
//             and keeps the smallest packed key of the fragments that pass
//             the mask; one coalesced float4 write of rast per pixel.
// Because every pixel is owned by exactly one lane, the z-buffer lives in
// registers and needs no atomics; the min over keys makes the result
// independent of bin order.  When the (triangle, tile) pairs exceed the bin
// capacity the resolve kernel walks every triangle of the view instead,
// which is slower but gives the same image.
//
// Reference behaviour reproduced: src/cuda/rasterize.cu:43-137 (rects),
// :170-277 (fragments), :319-365 (depth test), :923-971 (opacity aux),
// src/cuda/multires.cu:130-156 (per-fragment alpha), and
// python/diffsoup/multires.py:52-102 (the composition).
#include "ds_kernels.h"
#include "ds_math.h"
#include "ds_scan.h"

namespace ds {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ int view_of_tile(const Views& v, int tile) {
    int b = 0;
    while (b + 1 < v.n && v.tile_base[b + 1] <= tile) ++b;
    return b;
}

__global__ void k_fill_i32(int* p, long long n, int value) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_fill_f32(float* p, long long n, float value) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_clear(int* tile_count, int n_tiles, long long* stats) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_tiles) tile_count[i] = 0;
    if (i < kStatCount) stats[i] = 0;
}

__device__ __forceinline__ void tile_span(const int4 r, int ts, int& tx0, int& tx1, int& ty0, int& ty1) {
    tx0 = r.z / ts; tx1 = (r.w - 1) / ts;
    ty0 = r.x / ts; ty1 = (r.y - 1) / ts;
}

__global__ void k_setup(Views v, int V, const float4* __restrict__ pos, int T,
                        const int* __restrict__ tri, TriRecord* __restrict__ recs,
                        int* __restrict__ tile_count, long long* __restrict__ stats) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= (long long)v.n * T) return;
    const int b = static_cast<int>(i / T);
    const int t = static_cast<int>(i - (long long)b * T);
    const int i0 = tri[3 * t + 0], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
    TriRecord r;
    r.rect = make_int4(0, 0, 0, 0);
    if (i0 < 0 || i0 >= V || i1 < 0 || i1 >= V || i2 < 0 || i2 >= V) {
        r.p0 = r.p1 = r.p2 = make_float4(0.f, 0.f, 0.f, 1.f);
        recs[i] = r;
        stats[kStatBadIndex] = 1;
        return;
    }
    const float4* P = pos + (long long)b * V;
    r.p0 = P[i0]; r.p1 = P[i1]; r.p2 = P[i2];
    const Rect q = triangle_rect(r.p0, r.p1, r.p2, v.H[b], v.W[b]);
    r.rect = make_int4(q.y0, q.y1, q.x0, q.x1);
    recs[i] = r;
    if (q.y1 <= q.y0) return;
    int tx0, tx1, ty0, ty1;
    tile_span(r.rect, v.tile, tx0, tx1, ty0, ty1);
    int* cnt = tile_count + v.tile_base[b];
    const int tw = v.tiles_x[b];
    for (int ty = ty0; ty <= ty1; ++ty)
        for (int tx = tx0; tx <= tx1; ++tx) atomicAdd(&cnt[ty * tw + tx], 1);
}

// Exclusive scan of n ints by one 1024-thread workgroup, 4 items per thread
// per round.  off[n] receives the total; totals above `capacity` set the
// overflow stat so that k_emit writes nothing and k_resolve falls back.
// Exclusive scan of the per-tile counts into bin offsets, one workgroup
// (block_scan_1024).  off[n] is the total clamped to the bin capacity; the
// total and the overflow flag go to stats.
__global__ void __launch_bounds__(1024) k_scan(const int* __restrict__ cnt, int* __restrict__ off,
                                               int n, long long capacity, long long* stats) {
    const long long total = block_scan_1024<int>(cnt, n, off, false);
    if (threadIdx.x == 0) {
        off[n] = total > capacity ? static_cast<int>(capacity) : static_cast<int>(total);
        stats[kStatPairs] = total;
        if (total > capacity) stats[kStatOverflow] = 1;
    }
}

__global__ void k_emit(Views v, int T, const TriRecord* __restrict__ recs, int* __restrict__ tile_count,
                       const int* __restrict__ tile_off, int* __restrict__ bins,
                       const long long* __restrict__ stats) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= (long long)v.n * T) return;
    if (stats[kStatOverflow]) return;
    const int4 r = recs[i].rect;
    if (r.y <= r.x) return;
    const int b = static_cast<int>(i / T);
    const int t = static_cast<int>(i - (long long)b * T);
    int tx0, tx1, ty0, ty1;
    tile_span(r, v.tile, tx0, tx1, ty0, ty1);
    const int base = v.tile_base[b];
    const int tw = v.tiles_x[b];
    for (int ty = ty0; ty <= ty1; ++ty)
        for (int tx = tx0; tx <= tx1; ++tx) {
            const int tile = base + ty * tw + tx;
            const int slot = atomicSub(&tile_count[tile], 1) - 1;
            bins[tile_off[tile] + slot] = t;
        }
}

// Opacity at one lattice level (src/cuda/multires.cu:130-156 with C = 1).
__device__ __forceinline__ float lattice_alpha(const AlphaSource& a, int t, float b0, float b1) {
    const Corners c = lattice_corners(b0, b1, static_cast<uint32_t>(a.level));
    const float* src = a.alpha + (long long)t * a.S;
    float alpha = 0.f;
    alpha += c.w[0] * src[c.idx[0]];
    alpha += c.w[1] * src[c.idx[1]];
    alpha += c.w[2] * src[c.idx[2]];
    return alpha;
}

// Shared traversal state for one tile: which triangles to visit and where
// this lane's pixel sits.
struct TileWalk {
    int tile, b, px, py, qx0, qy0;
    float ndc_x, ndc_y;
    int start, n;
    bool overflow;
};

// interleave: with views of one tile grid, block i takes tile i / n of view
// i % n, so that the views' busiest tiles (their centres, where the scene
// projects) are spread over the launch instead of arriving together at its
// end.  The resolve gains from it (116 -> 110 us on eight 800x800 views);
// the opacity-gradient walk, which gains more from neighbouring tiles
// sharing records in the cache, keeps the tile order (70 vs 74 us).
template <int TS>
__device__ __forceinline__ TileWalk tile_walk_at(const Views& v, int T, const int* tile_off,
                                                 const long long* stats, int tile) {
    TileWalk w;
    w.tile = tile;
    w.b = view_of_tile(v, tile);
    const int lt = tile - v.tile_base[w.b];
    const int tx = lt % v.tiles_x[w.b], ty = lt / v.tiles_x[w.b];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    constexpr int QPR = TS / 8;
    w.qx0 = tx * TS + (wave % QPR) * 8;
    w.qy0 = ty * TS + (wave / QPR) * 8;
    w.px = w.qx0 + lane % 8;
    w.py = w.qy0 + lane / 8;
    w.ndc_x = pixel_ndc(w.px, v.W[w.b]);
    w.ndc_y = pixel_ndc(w.py, v.H[w.b]);
    w.overflow = stats[kStatOverflow] != 0;
    if (w.overflow) {
        w.start = 0;
        w.n = T;
    } else {
        w.start = tile_off[tile];
        w.n = tile_off[tile + 1] - w.start;
    }
    return w;
}

template <int TS>
__device__ __forceinline__ TileWalk tile_walk(const Views& v, int T, const int* tile_off,
                                              const long long* stats, bool interleave) {
    const int i = static_cast<int>(blockIdx.x);
    return tile_walk_at<TS>(v, T, tile_off, stats, interleave && v.interleave ? (i % v.n) * v.tile_base[1] + i / v.n : i);
}

// One chunk of a tile's bin, staged in LDS.  Rectangles are packed as two
// 16-bit pairs, (y0 | y1 << 16, x0 | x1 << 16): views are at most 65535
// pixels a side, so every bound fits, and the 8 bytes saved per entry keep
// the 16x16 resolve and opacity kernels at eight workgroups per CU.
template <int TS>
struct TileStage {
    float4 p0[TS * TS], p1[TS * TS], p2[TS * TS];
    uint2 rect[TS * TS];
    int tri[TS * TS];
};

__device__ __forceinline__ uint2 pack_rect(int4 r) {
    return make_uint2(static_cast<uint32_t>(r.x) | static_cast<uint32_t>(r.y) << 16,
                      static_cast<uint32_t>(r.z) | static_cast<uint32_t>(r.w) << 16);
}

template <int TS>
__device__ __forceinline__ void stage_chunk(TileStage<TS>& st, const TileWalk& w, int c0, int m, int T,
                                            const int* bins, const TriRecord* recs) {
    if (static_cast<int>(threadIdx.x) < m) {
        const int t = w.overflow ? c0 + threadIdx.x : bins[w.start + c0 + threadIdx.x];
        const TriRecord r = recs[(long long)w.b * T + t];
        st.p0[threadIdx.x] = r.p0;
        st.p1[threadIdx.x] = r.p1;
        st.p2[threadIdx.x] = r.p2;
        st.rect[threadIdx.x] = pack_rect(r.rect);
        st.tri[threadIdx.x] = t;
    }
}

// Dense (triangle, pixel) candidates of one wave's 8x8 quad.  The staged
// triangles are taken 32 at a time; lane i < 32 intersects triangle g0 + i's
// rectangle with the quad, a wave scan of those areas numbers the candidate
// pixels, and each pass hands one candidate to every lane (the triangle by a
// binary search over the prefix, the pixel row-major inside the
// intersection).  Every in-rectangle pixel of every triangle is visited
// exactly once, as the per-pixel walk over the bin visits it, but the lanes
// stay busy: a triangle covers ~10 of a quad's 64 pixels.
struct CandGroup {
    int start[32];    // exclusive prefix of the candidate counts
    // (x0 | y0 << 16, width | staged index << 4 | M << 20) of each
    // intersection, M = ceil(1024 * rcp(width)) from the hardware reciprocal;
    // for a candidate number n < 64 and width w <= 8, (n * M) >> 10 == n / w
    // exactly whenever rcp(w) is within 2 ulp of 1 / w (checked over all n, w)
    uint2 box[32];
    float ndcx[8];    // pixel-centre NDC of the quad's columns / rows
    float ndcy[8];
};

__device__ __forceinline__ void cand_setup(CandGroup& c, const TileWalk& w, int lane) {
    if (lane < 8) c.ndcx[lane] = w.ndc_x;
    if (lane % 8 == 0) c.ndcy[lane / 8] = w.ndc_y;
}

struct NoPost {
    __device__ __forceinline__ void operator()() const {}
};

// f runs once per candidate (pixel, triangle) of the staged chunk, on the
// lanes that hold one; post runs after each pass of up to 64 of them, on the
// whole wave.
template <int TS, class F, class G = NoPost>
__device__ __forceinline__ void for_each_candidate(const TileStage<TS>& st, int m, const TileWalk& w, CandGroup& c,
                                                   int lane, F&& f, G&& post = NoPost{}) {
    for (int g0 = 0; g0 < m; g0 += 32) {
        int area = 0, ix0 = 0, iy0 = 0, wd = 1;
        const int j = g0 + lane;
        if (lane < 32 && j < m) {
            const uint2 r = st.rect[j];
            ix0 = max(static_cast<int>(r.y & 0xffffu), w.qx0);
            iy0 = max(static_cast<int>(r.x & 0xffffu), w.qy0);
            const int ix1 = min(static_cast<int>(r.y >> 16), w.qx0 + 8);
            const int iy1 = min(static_cast<int>(r.x >> 16), w.qy0 + 8);
            if (ix1 > ix0 && iy1 > iy0) {
                wd = ix1 - ix0;
                area = wd * (iy1 - iy0);
            }
        }
        const int inc = scan32_inclusive(area, lane);
        const int total = __builtin_amdgcn_readlane(inc, 31);
        if (lane < 32) {
            c.start[lane] = inc - area;
            c.box[lane] = make_uint2(static_cast<uint32_t>(ix0) | static_cast<uint32_t>(iy0) << 16,
                                     static_cast<uint32_t>(wd) | static_cast<uint32_t>(j) << 4 |
                                         static_cast<uint32_t>(ceilf(1024.f * __builtin_amdgcn_rcpf(static_cast<float>(wd)))) << 20);
        }
        wave_lds_sync();
        for (int k0 = 0; k0 < total; k0 += kWave) {
            const int k = k0 + lane;
            if (k < total) {
                int lo = 0, hi = 31;   // last entry whose candidates start at or before k
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (c.start[mid] <= k) lo = mid; else hi = mid - 1;
                }
                const uint2 bx = c.box[lo];
                const int bw = static_cast<int>(bx.y & 0xfu);
                const int local = k - c.start[lo];
                const int dy = (local * static_cast<int>(bx.y >> 20)) >> 10, dx = local - dy * bw;
                const int px = static_cast<int>(bx.x & 0xffffu) + dx, py = static_cast<int>(bx.x >> 16) + dy;
                f(static_cast<int>((bx.y >> 4) & 0xffffu), px, py, (py - w.qy0) * 8 + (px - w.qx0),
                  c.ndcx[px - w.qx0], c.ndcy[py - w.qy0]);
            }
            post();
        }
        wave_lds_sync();
    }
}

// Eight waves per SIMD for 8x8 and 16x16 tiles: the record emission of the
// training step would otherwise take the walk past 64 VGPRs (seven waves);
// at 64 nothing spills (32x32 tiles would spill, and keep seven).
template <int TS, bool SHADE>
__global__ void __launch_bounds__(TS * TS) __attribute__((amdgpu_waves_per_eu(TS == 32 ? 7 : 8)))
k_resolve(Views v, int T, AlphaSource a, int mask_mode, MaskSeed mseed,
          const TriRecord* __restrict__ recs, const int* __restrict__ tile_off,
          const int* __restrict__ bins, long long* __restrict__ stats, float4* __restrict__ rast,
          int* __restrict__ tile_cov, BackgroundShade bg) {
    constexpr int NT = TS * TS;
    constexpr int NW = NT / kWave;
    __shared__ TileStage<TS> st;
    __shared__ float wave_loss[NW];
    __shared__ int wave_cnt[NW];
    __shared__ int wave_cov[NW];
    __shared__ CandGroup cg[NW];
    __shared__ unsigned long long zb[NW][kWave];   // per-wave z-buffer of packed keys
    const TileWalk w = tile_walk<TS>(v, T, tile_off, stats, true);
    const int wl = threadIdx.x % kWave, wq = threadIdx.x / kWave;
    const bool inside = w.px < v.W[w.b] && w.py < v.H[w.b];
    const long long p = v.pix_base[w.b] + (long long)w.py * v.W[w.b] + w.px;
    // the background shading's target values are loaded now, so that their
    // latency overlaps the candidate walk
    float tg[3] = {0.f, 0.f, 0.f};
    if constexpr (SHADE) {
        if (inside)
            for (int c = 0; c < 3; ++c) tg[c] = bg.target[p * 3 + c];
    }
    const bool emitting = SHADE && bg.aux_rec != nullptr;
    // Background pixels are shaded here: constant colour, the MSE term's
    // gradient and the per-pixel L1 for the opacity gradient.  Returns the
    // squared error.
    auto shade_background = [&]() {
        float* col = bg.color + p * 3;
        float* gc = bg.grad_color + p * 3;
        float l = 0.f, l1 = 0.f;
        for (int c = 0; c < 3; ++c) {
            const float d = bg.background - tg[c];
            col[c] = bg.background;
            gc[c] = bg.grad_scale * d;
            l1 += fabsf(d);
            l += d * d;
        }
        bg.pix_loss[p] = l1;
        return l;
    };

    if (w.n == 0) {
        // No triangle touches the tile (block-uniform): every pixel is
        // background, no fragment is covered and no record is emitted.
        if (emitting && wl == 0) bg.aux_n[(long long)w.tile * NW + wq] = 0;
        if (inside) rast[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        if constexpr (SHADE) {
            const float l = wave_sum_f32(inside ? shade_background() : 0.f);
            if (wl == 0) wave_loss[wq] = l;
            __syncthreads();
            if (threadIdx.x == 0) {
                float sum = 0.f;
                for (int k = 0; k < NW; ++k) sum += wave_loss[k];
                bg.loss_part[w.tile] = sum;
                bg.tile_ncov[w.tile] = 0;
            }
        }
        if (threadIdx.x == 0) tile_cov[w.tile] = 0;
        return;
    }

    const unsigned long long seed = mseed.device ? *mseed.device : mseed.value;
    const uint64_t view_hash = mask_view_hash(seed, static_cast<uint32_t>(w.b));
    CandGroup& cgw = cg[wq];
    unsigned long long* zw = zb[wq];
    cand_setup(cgw, w, wl);
    zw[wl] = static_cast<unsigned long long>(kKeyEmpty);

    float bb0 = 0.f, bb1 = 0.f, bd = 0.f;
    int bt = -1;
    int covered = 0;

    // Fragment records for the opacity gradient (training step): a covered
    // fragment whose rounded depth is not beyond its pixel's current best is
    // emitted with its depth; the pixel's final surface is never beyond the
    // current best, so every fragment that is not behind the final surface is
    // among them (the gradient pass applies the exact test).  Each wave owns
    // its quad's slots and stores a pass's records after the pass, in lane
    // order, behind a wave-uniform count.
    const long long aux_base = emitting ? ((long long)w.tile * NW + wq) * bg.aux_per_quad : 0;
    int aux_n = 0;
    bool emit = false;
    int4 erec;

    for (int c0 = 0; c0 < w.n; c0 += NT) {
        const int m = min(NT, w.n - c0);
        __syncthreads();
        stage_chunk<TS>(st, w, c0, m, T, bins, recs);
        __syncthreads();
        for_each_candidate<TS>(st, m, w, cgw, wl, [&](int j, int px, int py, int pi, float ndc_x, float ndc_y) {
            Hit h;
            if (!cover(ndc_y, ndc_x, st.p0[j], st.p1[j], st.p2[j], h)) return;
            float d;
            if (!fragment_depth(h, d)) return;
            ++covered;
            const int t = st.tri[j];
            const bool testable = depth_testable(d);
            unsigned long long key = 0, cur = 0;
            if (testable) {
                key = static_cast<unsigned long long>(depth_key(d, static_cast<uint32_t>(t)));
                cur = zw[pi];
            }
            // pixels past the view's edge carry no loss and emit nothing
            if (emitting && (!testable || (key >> 32) <= (cur >> 32)) && px < v.W[w.b] && py < v.H[w.b]) {
                emit = true;
                erec = make_int4(t | pi << 24, __float_as_int(h.b0), __float_as_int(h.b1), __float_as_int(d));
            }
            // keys are positive, so unsigned order is the signed order; a key
            // not below the current minimum cannot win whatever the mask says
            if (!testable || key >= cur) return;
            const float alpha = lattice_alpha(a, t, h.b0, h.b1);
            const float u = mask_mode == kMaskHash ? mask_threshold_in_view(view_hash, t, py, px) : 0.5f;
            if (alpha < u) return;
            atomicMin(&zw[pi], key);
        }, [&]() {
            if (!emitting) return;
            const unsigned long long em = __ballot(emit);
            if (em) {
                const int pos = aux_n + __popcll(em & ((1ull << wl) - 1ull));
                if (emit && pos < bg.aux_per_quad) bg.aux_rec[aux_base + pos] = erec;
                aux_n += __popcll(em);
                emit = false;
            }
        });
    }
    wave_lds_sync();
    {   // this lane's pixel: the winning triangle's hit, recomputed exactly as in the walk
        const unsigned long long kb = zw[wl];
        if (kb != static_cast<unsigned long long>(kKeyEmpty)) {
            bt = static_cast<int>(kb & 0xffffffffull);
            const TriRecord& r = recs[(long long)w.b * T + bt];
            Hit h;
            cover(w.ndc_y, w.ndc_x, r.p0, r.p1, r.p2, h);
            fragment_depth(h, bd);
            bb0 = h.b0;
            bb1 = h.b1;
        }
    }

    if (emitting && wl == 0) {
        bg.aux_n[(long long)w.tile * NW + wq] = aux_n;
        if (aux_n > bg.aux_per_quad)
            atomicAdd(reinterpret_cast<unsigned long long*>(stats + kStatAuxWalkQuads), 1ull);
    }
    if (inside) {
        rast[p] = bt >= 0 ? make_float4(bb0, bb1, bd, static_cast<float>(bt + 1))
                          : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
    if constexpr (SHADE) {
        // Covered pixels are listed per tile in lane order, each with its rank
        // among the covered pixels of its triangle (one atomic per distinct
        // triangle per wave).
        const float l = wave_sum_f32(inside && bt < 0 ? shade_background() : 0.f);
        const bool cov = inside && bt >= 0;
        const unsigned long long m = __ballot(cov);
        const unsigned long long below = m & ((1ull << lane) - 1ull);
        // rank within the triangle: group equal triangle ids of the wave, then
        // every group's first lane claims the group's slots with one atomic
        // (all groups in the same instruction)
        int leader = -1, local = 0, group = 0;
        unsigned long long todo = m;
        while (todo) {
            const int l0 = __ffsll(static_cast<long long>(todo)) - 1;
            const int tl = __builtin_amdgcn_readlane(bt, l0);
            const unsigned long long same = __ballot(cov && bt == tl);
            if (cov && bt == tl) {
                leader = l0;
                local = __popcll(same & ((1ull << lane) - 1ull));
                group = __popcll(same);
            }
            todo &= ~same;
        }
        int base = 0;
        if (cov && lane == leader) base = atomicAdd(&bg.tri_count[bt], group);
        base = __shfl(base, leader < 0 ? lane : leader, kWave);
        const int rank = base + local;
        if (lane == 0) {
            wave_loss[wv] = l;
            wave_cnt[wv] = __popcll(m);
        }
        __syncthreads();
        int off = 0;
        for (int k = 0; k < wv; ++k) off += wave_cnt[k];
        if (cov) bg.tile_list[(long long)w.tile * NT + off + __popcll(below)] = make_int2(static_cast<int>(p), rank);
        if (threadIdx.x == 0) {
            float sum = 0.f;
            int n = 0;
            for (int k = 0; k < NW; ++k) { sum += wave_loss[k]; n += wave_cnt[k]; }
            bg.loss_part[w.tile] = sum;
            bg.tile_ncov[w.tile] = n;
        }
    }
    const int cv = wave_sum_i32(covered);
    if (lane == 0) wave_cov[wv] = cv;
    __syncthreads();
    if (threadIdx.x == 0) {
        int n = 0;
        for (int k = 0; k < NW; ++k) n += wave_cov[k];
        tile_cov[w.tile] = n;
    }
}

// Sum of the per-tile covered-fragment counts into the stats block.
// stats[covered] = sum of the per-tile covered-fragment counts: each thread
// issues all its loads (a strided run of up to 32) before adding.
__global__ void __launch_bounds__(1024) k_sum_covered(const int* __restrict__ tile_cov, int n,
                                                      long long* __restrict__ stats) {
    __shared__ long long red[16];
    long long s = 0;
    for (int base = 0; base < n; base += 32 * 1024) {
        int v[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int i = base + k * 1024 + static_cast<int>(threadIdx.x);
            v[k] = i < n ? tile_cov[i] : 0;
        }
#pragma unroll
        for (int k = 0; k < 32; ++k) s += v[k];
    }
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int k = 0; k < 16; ++k) t += red[k];
        stats[kStatCovered] = t;
    }
}

// list[tri_off[t] + rank] = pixel for every covered pixel of every tile, and
// (when px_list is given) the same pixels in tile order:
// px_list[tile_px_off[tile] + k] = pixel.
// Workgroups from n_tiles on run the shading ranges' jobs instead, one wave
// per job (grid-stride over the job list).
template <int NT>
__global__ void __launch_bounds__(NT) k_fill_lists(Views v, const int2* __restrict__ tile_list,
                                                   const int* __restrict__ tile_ncov,
                                                   const float4* __restrict__ rast,
                                                   const float* __restrict__ target,
                                                   const int* __restrict__ tri_off, PixRec* __restrict__ list,
                                                   const int* __restrict__ tile_px_off, int4* __restrict__ px_list,
                                                   int n_tiles, ShadeRanges ranges) {
    if (static_cast<int>(blockIdx.x) >= n_tiles) {
        constexpr int WPB = NT / kWave;
        const long long waves = (long long)(gridDim.x - n_tiles) * WPB;
        const long long jobs = shade_range_jobs(ranges);
        for (long long job = (long long)(blockIdx.x - n_tiles) * WPB + threadIdx.x / kWave; job < jobs; job += waves)
            shade_range_job(ranges, job, threadIdx.x % kWave);
        return;
    }
    const int k = threadIdx.x;
    if (k >= tile_ncov[blockIdx.x]) return;
    const int2 e = tile_list[(long long)blockIdx.x * NT + k];
    const float4 r = rast[e.x];
    const float* tg = target + (long long)e.x * 3;
    const int t = static_cast<int>(r.w) - 1;
    float4* dst = reinterpret_cast<float4*>(list + tri_off[t] + e.y);
    dst[0] = make_float4(r.x, r.y, __int_as_float(e.x), __int_as_float(t));
    dst[1] = make_float4(tg[0], tg[1], tg[2], __int_as_float(tri_off[t + 1]));
    if (px_list) px_list[tile_px_off[blockIdx.x] + k] = make_int4(e.x, view_of_tile(v, blockIdx.x), t, 0);
}

// The opacity gradient of one fragment that is not strictly behind its
// pixel's visible surface (src/cuda/rasterize.cu:923-971): the pixel's L1
// loss over the probability of the fragment's mask outcome (kept for the
// visible triangle, dropped for any other), scattered onto the triangle's
// opacity lattice (src/cuda/multires.cu:178-202).
// The estimator's factor of one fragment: loss / alpha for the pixel's
// visible triangle, -loss / (1 - alpha) for any other (both clamped at 1e-5).
__device__ __forceinline__ float aux_factor(const AlphaSource& a, int t, const Corners& c, int tri_pix, float loss) {
    constexpr float eps = 1e-5f;
    const float* src = a.alpha + (long long)t * a.S;
    float alpha = 0.f;
    alpha += c.w[0] * src[c.idx[0]];
    alpha += c.w[1] * src[c.idx[1]];
    alpha += c.w[2] * src[c.idx[2]];
    if (tri_pix >= 0 && t == tri_pix) return loss / (alpha > eps ? alpha : eps);
    const float om = 1.f - alpha;
    return -(loss / (om > eps ? om : eps));
}

__device__ __forceinline__ void aux_fragment(const AlphaSource& a, int t, float b0, float b1, int tri_pix, float loss,
                                             float* __restrict__ grad_alpha) {
    const Corners c = lattice_corners(b0, b1, static_cast<uint32_t>(a.level));
    const float g = aux_factor(a, t, c, tri_pix, loss);
    float* dst = grad_alpha + (long long)t * a.S;
    unsafeAtomicAdd(&dst[c.idx[0]], g * c.w[0]);
    unsafeAtomicAdd(&dst[c.idx[1]], g * c.w[1]);
    unsafeAtomicAdd(&dst[c.idx[2]], g * c.w[2]);
}

template <int TS>
struct AuxShared {
    static constexpr int NW = TS * TS / kWave;
    TileStage<TS> st;
    CandGroup cg[NW];
    float pix_depth[NW][kWave], pix_l1[NW][kWave];
    int pix_tri[NW][kWave];
};

// One tile's opacity gradient regenerated from its bins: every wave whose
// flag is set walks its quad's fragments; the others only join the block's
// staging of the candidates.
template <int TS>
__device__ void aux_walk_tile(AuxShared<TS>& sh, const Views& v, int T, const AlphaSource& a,
                              const TriRecord* __restrict__ recs, const int* __restrict__ tile_off,
                              const int* __restrict__ bins, const long long* __restrict__ stats,
                              const float4* __restrict__ rast, const float* __restrict__ pix_loss,
                              float* __restrict__ grad_alpha, int tile, bool walk, bool skip_visible) {
    constexpr int NT = TS * TS;
    const TileWalk w = tile_walk_at<TS>(v, T, tile_off, stats, tile);
    const int wl = threadIdx.x % kWave, wq = threadIdx.x / kWave;
    CandGroup& cgw = sh.cg[wq];
    cand_setup(cgw, w, wl);
    {
        const bool inside_view = w.px < v.W[w.b] && w.py < v.H[w.b];
        float depth_pix = 0.f, loss = 0.f;
        int tri_pix = -1;
        if (inside_view) {
            const long long p = v.pix_base[w.b] + (long long)w.py * v.W[w.b] + w.px;
            const float4 rp = rast[p];
            depth_pix = rp.z;
            tri_pix = static_cast<int>(rp.w) - 1;
            loss = pix_loss[p];
        }
        sh.pix_depth[wq][wl] = depth_pix;
        sh.pix_l1[wq][wl] = loss;
        sh.pix_tri[wq][wl] = tri_pix;
    }
    for (int c0 = 0; c0 < w.n; c0 += NT) {
        const int m = min(NT, w.n - c0);
        __syncthreads();
        stage_chunk<TS>(sh.st, w, c0, m, T, bins, recs);
        __syncthreads();
        if (!walk) continue;
        for_each_candidate<TS>(sh.st, m, w, cgw, wl, [&](int j, int px, int py, int pi, float ndc_x, float ndc_y) {
            Hit h;
            if (!cover(ndc_y, ndc_x, sh.st.p0[j], sh.st.p1[j], sh.st.p2[j], h)) return;
            float d;
            if (!fragment_depth(h, d)) return;
            const int tri_pix = sh.pix_tri[wq][pi];
            if (tri_pix >= 0 && d > sh.pix_depth[wq][pi]) return;
            if (skip_visible && sh.st.tri[j] == tri_pix) return;
            aux_fragment(a, sh.st.tri[j], h.b0, h.b1, tri_pix, sh.pix_l1[wq][pi], grad_alpha);
        });
    }
}

// Score-function gradient of the per-pixel L1 loss w.r.t. fragment opacity
// (src/cuda/rasterize.cu:923-971): every fragment of the fragment list that
// is not strictly behind the visible surface contributes.  Without records,
// one tile per block, its fragments regenerated from the bins.  With the
// training-step resolve's per-quad records, a block takes tiles_per_block
// tiles: the records of all their quads are processed as one flat list, and
// the tiles with a quad whose records overflowed are then walked for those
// quads only.
template <int TS>
__global__ void __launch_bounds__(TS * TS)
k_opacity_aux(Views v, int T, AlphaSource a, const TriRecord* __restrict__ recs,
              const int* __restrict__ tile_off, const int* __restrict__ bins,
              const long long* __restrict__ stats, const float4* __restrict__ rast,
              const float* __restrict__ pix_loss, float* __restrict__ grad_alpha,
              const int4* __restrict__ aux_rec, const int* __restrict__ aux_n, int per_quad,
              int tiles_per_block, int skip_visible) {
    constexpr int NT = TS * TS;
    constexpr int NW = NT / kWave;
    constexpr int QPR = TS / 8;
    __shared__ AuxShared<TS> sh;
    if (!aux_rec) {
        aux_walk_tile<TS>(sh, v, T, a, recs, tile_off, bins, stats, rast, pix_loss, grad_alpha,
                          static_cast<int>(blockIdx.x), true, false);
        return;
    }
    __shared__ int seg_off[NT];
    __shared__ int wave_tot[NW];
    __shared__ int seg_over[NT];
    const int n_tiles = v.tile_base[v.n];
    const int tile0 = static_cast<int>(blockIdx.x) * tiles_per_block;
    const int ntile = min(tiles_per_block, n_tiles - tile0);
    const int nseg = ntile * NW;   // <= NT: tiles_per_block <= NT / NW
    const long long seg0 = (long long)tile0 * NW;
    const int i = threadIdx.x, wl = i % kWave, wq = i / kWave;
    int cnt = 0, over = 0;
    if (i < nseg) {
        cnt = aux_n[seg0 + i];
        over = cnt > per_quad;
        if (over) cnt = 0;
    }
    seg_over[i] = over;
    int inc = cnt;
    for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(inc, d, kWave);
        if (wl >= d) inc += o;
    }
    if (wl == kWave - 1) wave_tot[wq] = inc;
    __syncthreads();
    int before = 0, total = 0;
    for (int k = 0; k < NW; ++k) {
        const int x = wave_tot[k];
        if (k < wq) before += x;
        total += x;
    }
    seg_off[i] = before + inc - cnt;
    __syncthreads();
    for (int r = i; r < total; r += NT) {
        int lo = 0, hi = nseg - 1;   // the last quad whose records start at or before r
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (seg_off[mid] <= r) lo = mid; else hi = mid - 1;
        }
        const int4 rec = aux_rec[(seg0 + lo) * per_quad + (r - seg_off[lo])];
        const int tile = tile0 + lo / NW, q = lo % NW;
        const int b = view_of_tile(v, tile);
        const int lt = tile - v.tile_base[b];
        const int pi = rec.x >> 24;
        const int px = (lt % v.tiles_x[b]) * TS + (q % QPR) * 8 + (pi & 7);
        const int py = (lt / v.tiles_x[b]) * TS + (q / QPR) * 8 + (pi >> 3);
        const long long p = v.pix_base[b] + (long long)py * v.W[b] + px;
        const float4 rp = rast[p];
        const int tri_pix = static_cast<int>(rp.w) - 1;
        if (tri_pix >= 0 && __int_as_float(rec.w) > rp.z) continue;
        // the visible fragment's term was added by the shading kernel
        if (skip_visible && (rec.x & 0xffffff) == tri_pix) continue;
        aux_fragment(a, rec.x & 0xffffff, __int_as_float(rec.y), __int_as_float(rec.z), tri_pix, pix_loss[p],
                     grad_alpha);
    }
    for (int t0 = 0; t0 < ntile; ++t0) {
        int any = 0;
        for (int k = 0; k < NW; ++k) any |= seg_over[t0 * NW + k];
        if (any)
            aux_walk_tile<TS>(sh, v, T, a, recs, tile_off, bins, stats, rast, pix_loss, grad_alpha, tile0 + t0,
                              seg_over[t0 * NW + wq] != 0, skip_visible != 0);
    }
}

__global__ void k_pixel_l1(long long P, int C, const float* __restrict__ color,
                           const float* __restrict__ target, float* __restrict__ pix_loss) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= P) return;
    float l = 0.f;
    for (int c = 0; c < C; ++c) l += fabsf(color[p * C + c] - target[p * C + c]);
    pix_loss[p] = l;
}

inline unsigned blocks_for(long long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

}  // namespace

size_t raster_workspace_bytes(int B, int T, int n_tiles, int64_t capacity) {
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    return up(sizeof(long long) * kStatCount) + up(sizeof(TriRecord) * (size_t)B * T) +
           up(sizeof(int) * (size_t)n_tiles) + up(sizeof(int) * ((size_t)n_tiles + 1)) +
           up(sizeof(int) * (size_t)n_tiles) + up(sizeof(int) * (size_t)(capacity > 0 ? capacity : 1));
}

RasterWorkspace raster_workspace_carve(void* ws, int B, int T, int n_tiles, int64_t capacity) {
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    char* p = static_cast<char*>(ws);
    RasterWorkspace w;
    w.stats = reinterpret_cast<long long*>(p);           p += up(sizeof(long long) * kStatCount);
    w.recs = reinterpret_cast<TriRecord*>(p);            p += up(sizeof(TriRecord) * (size_t)B * T);
    w.tile_count = reinterpret_cast<int*>(p);            p += up(sizeof(int) * (size_t)n_tiles);
    w.tile_off = reinterpret_cast<int*>(p);              p += up(sizeof(int) * ((size_t)n_tiles + 1));
    w.tile_cov = reinterpret_cast<int*>(p);              p += up(sizeof(int) * (size_t)n_tiles);
    w.bins = reinterpret_cast<int*>(p);
    w.capacity = capacity;
    return w;
}

hipError_t launch_fill_i32(int* p, long long n, int value, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned g = static_cast<unsigned>(std::min<long long>(blocks_for(n, 256), 65536));
    k_fill_i32<<<g, 256, 0, s>>>(p, n, value);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, long long n, float value, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned g = static_cast<unsigned>(std::min<long long>(blocks_for(n, 256), 65536));
    k_fill_f32<<<g, 256, 0, s>>>(p, n, value);
    return hipGetLastError();
}

hipError_t launch_raster_bin(const Views& v, int V, const float* pos, int T, const int* tri,
                             const RasterWorkspace& w, hipStream_t s, bool clear) {
    const int n_tiles = v.tile_base[v.n];
    if (clear)
        k_clear<<<blocks_for(std::max(n_tiles, (int)kStatCount), 256), 256, 0, s>>>(w.tile_count, n_tiles,
                                                                                      w.stats);
    const long long BT = (long long)v.n * T;
    if (BT > 0) {
        k_setup<<<blocks_for(BT, 256), 256, 0, s>>>(v, V, reinterpret_cast<const float4*>(pos), T, tri,
                                                    w.recs, w.tile_count, w.stats);
    }
    k_scan<<<1, 1024, 0, s>>>(w.tile_count, w.tile_off, n_tiles, w.capacity, w.stats);
    if (BT > 0) {
        k_emit<<<blocks_for(BT, 256), 256, 0, s>>>(v, T, w.recs, w.tile_count, w.tile_off, w.bins,
                                                   w.stats);
    }
    return hipGetLastError();
}

template <bool SHADE>
static hipError_t resolve_dispatch(const Views& v, int T, const AlphaSource& a, int mask_mode,
                                   MaskSeed seed, float* rast, const RasterWorkspace& w,
                                   const BackgroundShade& bg, hipStream_t s) {
    const int n_tiles = v.tile_base[v.n];
    if (n_tiles == 0) return hipSuccess;
    float4* out = reinterpret_cast<float4*>(rast);
    switch (v.tile) {
        case 8:
            k_resolve<8, SHADE><<<n_tiles, 64, 0, s>>>(v, T, a, mask_mode, seed, w.recs, w.tile_off,
                                                       w.bins, w.stats, out, w.tile_cov, bg);
            break;
        case 16:
            k_resolve<16, SHADE><<<n_tiles, 256, 0, s>>>(v, T, a, mask_mode, seed, w.recs, w.tile_off,
                                                         w.bins, w.stats, out, w.tile_cov, bg);
            break;
        case 32:
            k_resolve<32, SHADE><<<n_tiles, 1024, 0, s>>>(v, T, a, mask_mode, seed, w.recs, w.tile_off,
                                                          w.bins, w.stats, out, w.tile_cov, bg);
            break;
        default:
            return hipErrorInvalidValue;
    }
    // the training step sums the per-tile counts in its epilogue instead
    if (!SHADE) k_sum_covered<<<1, 1024, 0, s>>>(w.tile_cov, n_tiles, w.stats);
    return hipGetLastError();
}

hipError_t launch_raster_resolve(const Views& v, int T, const AlphaSource& a, int mask_mode,
                                 MaskSeed seed, float* rast, const RasterWorkspace& w, hipStream_t s) {
    return resolve_dispatch<false>(v, T, a, mask_mode, seed, rast, w, BackgroundShade{}, s);
}

hipError_t launch_raster_resolve_shade(const Views& v, int T, const AlphaSource& a, int mask_mode,
                                       MaskSeed seed, float* rast, const RasterWorkspace& w,
                                       const BackgroundShade& bg, hipStream_t s) {
    return resolve_dispatch<true>(v, T, a, mask_mode, seed, rast, w, bg, s);
}

static hipError_t aux_launch(const Views& v, int T, const AlphaSource& a, const float* rast, const float* pix_loss,
                             float* grad_alpha, const RasterWorkspace& w, const int4* rec, const int* aux_n,
                             int per_quad, bool skip_visible, hipStream_t s) {
    const int n_tiles = v.tile_base[v.n];
    if (n_tiles == 0) return hipSuccess;
    const float4* r4 = reinterpret_cast<const float4*>(rast);
    // With records, one tile per block up to 16384 tiles, then more tiles
    // per block (at most 64: one quad count per thread).  Scenes that cover
    // few tiles densely keep their busy tiles apart; scenes spread thinly
    // over many views save the launches of empty tiles (eight 800x800 views:
    // 57 us at one tile per block vs 69 us at ten; 32 mixed views, 106K
    // tiles: 228 us at 13 vs 298 us at one).
    const int tpb = rec ? std::max(1, std::min(kWave, (n_tiles + 16383) / 16384)) : 1;
    const unsigned grid = static_cast<unsigned>((n_tiles + tpb - 1) / tpb);
    switch (v.tile) {
        case 8:
            k_opacity_aux<8><<<grid, 64, 0, s>>>(v, T, a, w.recs, w.tile_off, w.bins, w.stats, r4, pix_loss,
                                                 grad_alpha, rec, aux_n, per_quad, tpb, skip_visible ? 1 : 0);
            break;
        case 16:
            k_opacity_aux<16><<<grid, 256, 0, s>>>(v, T, a, w.recs, w.tile_off, w.bins, w.stats, r4, pix_loss,
                                                   grad_alpha, rec, aux_n, per_quad, tpb, skip_visible ? 1 : 0);
            break;
        case 32:
            k_opacity_aux<32><<<grid, 1024, 0, s>>>(v, T, a, w.recs, w.tile_off, w.bins, w.stats, r4, pix_loss,
                                                    grad_alpha, rec, aux_n, per_quad, tpb, skip_visible ? 1 : 0);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_opacity_aux(const Views& v, int T, const AlphaSource& a, const float* rast,
                              const float* pix_loss, float* grad_alpha, const RasterWorkspace& w,
                              hipStream_t s) {
    return aux_launch(v, T, a, rast, pix_loss, grad_alpha, w, nullptr, nullptr, 0, false, s);
}

hipError_t launch_opacity_aux_records(const Views& v, int T, const AlphaSource& a, const float* rast,
                                      const float* pix_loss, float* grad_alpha, const RasterWorkspace& w,
                                      const int4* rec, const int* aux_n, int per_quad, bool skip_visible,
                                      hipStream_t s) {
    return aux_launch(v, T, a, rast, pix_loss, grad_alpha, w, rec, aux_n, per_quad, skip_visible, s);
}

// Workgroups for the shading ranges' jobs: one wave per job, at most 4096 waves.
static unsigned range_blocks(const ShadeRanges& r, int waves_per_block) {
    const long long jobs = (long long)r.n_waves + 1 + (r.grad ? (r.T + 63) / 64 : 0);
    const long long waves = std::min<long long>(jobs, 4096);
    return static_cast<unsigned>((waves + waves_per_block - 1) / waves_per_block);
}

hipError_t launch_fill_triangle_lists(const Views& v, const int2* tile_list, const int* tile_ncov,
                                      const float* rast, const float* target, const int* tri_off, PixRec* list,
                                      const int* tile_px_off, int4* px_list, const ShadeRanges& ranges,
                                      hipStream_t s) {
    const int n_tiles = v.tile_base[v.n];
    const float4* r4 = reinterpret_cast<const float4*>(rast);
    switch (v.tile) {
        case 8:
            k_fill_lists<64><<<n_tiles + range_blocks(ranges, 1), 64, 0, s>>>(
                v, tile_list, tile_ncov, r4, target, tri_off, list, tile_px_off, px_list, n_tiles, ranges);
            break;
        case 16:
            k_fill_lists<256><<<n_tiles + range_blocks(ranges, 4), 256, 0, s>>>(
                v, tile_list, tile_ncov, r4, target, tri_off, list, tile_px_off, px_list, n_tiles, ranges);
            break;
        case 32:
            k_fill_lists<1024><<<n_tiles + range_blocks(ranges, 16), 1024, 0, s>>>(
                v, tile_list, tile_ncov, r4, target, tri_off, list, tile_px_off, px_list, n_tiles, ranges);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_shade_ranges(ShadeRanges ranges) {
    const long long waves = (long long)gridDim.x * 4;
    const long long jobs = shade_range_jobs(ranges);
    for (long long job = blockIdx.x * 4LL + threadIdx.x / kWave; job < jobs; job += waves)
        shade_range_job(ranges, job, threadIdx.x % kWave);
}

hipError_t launch_shade_ranges(const ShadeRanges& ranges, hipStream_t s) {
    k_shade_ranges<<<range_blocks(ranges, 4), 256, 0, s>>>(ranges);
    return hipGetLastError();
}

hipError_t launch_pixel_l1(long long P, int C, const float* color, const float* target,
                           float* pix_loss, hipStream_t s) {
    if (P <= 0) return hipSuccess;
    k_pixel_l1<<<blocks_for(P, 256), 256, 0, s>>>(P, C, color, target, pix_loss);
    return hipGetLastError();
}

}  // namespace ds

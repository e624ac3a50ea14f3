// Launch-side declarations shared by the DiffSoup HIP translation units.
// Kernels live in raster.hip, multires.hip, encoding.hip, edge_grad.hip and
// shade.hip; the extern "C" boundary in capi.hip validates arguments and
// calls these launchers on the caller's stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ds {

// Orders this wave's LDS writes before its later LDS reads (lanes of one
// wave exchanging data through LDS; no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Wave-wide reductions and scans on the VALU's data-parallel primitives
// (DPP) instead of LDS permutes.  DPP controls used: row_shr:n = 0x110 + n
// and row_ror:n = 0x120 + n, both inside each 16-lane row.  Every lane of the
// wave must be active.

// Sum over the 64 lanes, the same value in every lane: each row's total by
// four rotations, then the four row totals in a fixed order.
__device__ __forceinline__ int wave_sum_i32(int x) {
    x += __builtin_amdgcn_mov_dpp(x, 0x128, 0xf, 0xf, false);
    x += __builtin_amdgcn_mov_dpp(x, 0x124, 0xf, 0xf, false);
    x += __builtin_amdgcn_mov_dpp(x, 0x122, 0xf, 0xf, false);
    x += __builtin_amdgcn_mov_dpp(x, 0x121, 0xf, 0xf, false);
    return (__builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16)) +
           (__builtin_amdgcn_readlane(x, 32) + __builtin_amdgcn_readlane(x, 48));
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ float wave_sum_f32(float x) {
    x += dpp_f32<0x128>(x);
    x += dpp_f32<0x124>(x);
    x += dpp_f32<0x122>(x);
    x += dpp_f32<0x121>(x);
    auto at = [&](int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); };
    return (at(0) + at(16)) + (at(32) + at(48));
}

// Inclusive prefix sum over lanes 0..31 (lanes 32..63 receive unspecified
// values): a shift-and-add scan inside each row, whose shifted-in lanes read
// zero, then row 0's total added to row 1.
__device__ __forceinline__ int scan32_inclusive(int x, int lane) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    const int row0 = __builtin_amdgcn_readlane(x, 15);
    return lane >= 16 ? x + row0 : x;
}

// Views per launch: the view table below travels as a kernel argument
// (about 5.3 KB at 256 views, the largest argument block of any kernel here).
constexpr int kMaxViews = 256;

// Per-launch description of the views packed into one batch.  Passed to the
// kernels by value, so a captured graph keeps its own copy.  Views may have
// different resolutions; view b's pixels occupy [pix_base[b], pix_base[b+1])
// of every packed per-pixel buffer and its tiles [tile_base[b], tile_base[b+1]).
struct Views {
    int n;
    int tile;                      // tile edge in pixels: 8, 16 or 32
    int H[kMaxViews], W[kMaxViews];
    int tiles_x[kMaxViews];
    int tile_base[kMaxViews + 1];
    int pix_base[kMaxViews + 1];
    int interleave;                // 1 when every view has the same tile grid
};

// Per-(view, triangle) setup record: clip-space corners and the conservative
// pixel rectangle (half-open; empty when culled).
struct __align__(16) TriRecord {
    float4 p0, p1, p2;
    int4 rect;                     // y0, y1, x0, x1
};

// Device-side counters written by the raster launches (int64 each).
enum Stat : int {
    kStatCovered = 0,   // fragments inside and with depth in [-1, 1], before the mask
    kStatPairs = 1,     // (triangle, tile) pairs produced by binning
    kStatOverflow = 2,  // 1 when pairs exceeded the bin capacity (fallback path ran)
    kStatBadIndex = 3,  // 1 when a triangle referenced a vertex outside [0, V)
    kStatAuxWalkQuads = 4,  // 8x8 quads whose opacity-gradient records overflowed (their bins were walked)
    kStatCount = 8,
};

// One covered pixel of the triangle-grouped list that the shading kernel
// walks: its barycentrics, the pixel, the triangle, the pixel's target
// colour and the end of the triangle's run in the list (carried along so
// that shading walks from one triangle's run to the next, and from tile to
// tile, without dependent loads of offsets or targets).
struct __align__(16) PixRec {
    float b0, b1;
    int pixel;
    int tri;
    float tg[3];
    int seg_end;
};

// Workspace carve-up for one raster launch.
struct RasterWorkspace {
    TriRecord* recs;     // [B * T]
    int* tile_count;     // [n_tiles]
    int* tile_off;       // [n_tiles + 1]
    int* tile_cov;       // [n_tiles] covered fragments per tile (summed into stats)
    int* bins;           // [capacity]
    long long* stats;    // [kStatCount]
    int64_t capacity;
};

size_t raster_workspace_bytes(int B, int T, int n_tiles, int64_t capacity);
RasterWorkspace raster_workspace_carve(void* ws, int B, int T, int n_tiles, int64_t capacity);

// Mask modes for the depth test.
enum MaskMode : int { kMaskDeterministic = 0, kMaskHash = 1 };

struct AlphaSource {
    const float* alpha;   // [T, S] post-sigmoid lattice opacities at one level
    int level;
    int S;
};

// Seed of the stochastic mask: a value, or (when device is set) the uint64
// in device memory that the kernel reads when it runs, so that a captured
// graph picks up a new seed on every replay.
struct MaskSeed {
    unsigned long long value;
    const unsigned long long* device;
};

// clear = false when the caller has already zeroed w.tile_count and w.stats.
hipError_t launch_raster_bin(const Views& v, int V, const float* pos, int T, const int* tri,
                             const RasterWorkspace& w, hipStream_t s, bool clear = true);
hipError_t launch_raster_resolve(const Views& v, int T, const AlphaSource& a, int mask_mode,
                                 MaskSeed seed, float* rast, const RasterWorkspace& w, hipStream_t s);

// Training-step variant of the resolve: background pixels get their colour,
// MSE gradient and L1 written here, covered pixels are appended to a list.
// It leaves the covered-fragment count per tile in w.tile_cov and does not
// sum it into the stats block (the caller does).
struct BackgroundShade {
    const float* target;   // [P, 3]
    float* color;          // [P, 3]
    float* grad_color;     // [P, 3]
    float* pix_loss;       // [P]
    float* loss_part;      // [n_tiles] per-tile sum of squared background error
    int2* tile_list;       // [n_tiles * tile^2] (pixel, rank within its triangle)
    int* tile_ncov;        // [n_tiles] covered pixels per tile
    int* tri_count;        // [T] covered pixels per triangle over all views (zeroed before)
    float grad_scale;      // dL/dcolor = grad_scale * (color - target)
    float background;      // constant background colour
    // Fragment records for the opacity gradient (null: none): every covered
    // fragment not known to lie behind its pixel's visible surface when the
    // walk reached it, as (triangle | pixel-in-quad << 24, b0, b1, depth bits),
    // aux_per_quad slots per 8x8 quad (tile-major, quads in wave order);
    // aux_n[quad] = the quad's count (above aux_per_quad: overflowed, only
    // the first aux_per_quad were stored).
    int4* aux_rec;
    int* aux_n;
    int aux_per_quad;
};
hipError_t launch_raster_resolve_shade(const Views& v, int T, const AlphaSource& a, int mask_mode,
                                       MaskSeed seed, float* rast, const RasterWorkspace& w,
                                       const BackgroundShade& bg, hipStream_t s);
hipError_t launch_opacity_aux(const Views& v, int T, const AlphaSource& a, const float* rast,
                              const float* pix_loss, float* grad_alpha, const RasterWorkspace& w,
                              hipStream_t s);
// The opacity gradient from the records a training-step resolve emitted
// (BackgroundShade::aux_rec); the quads of a tile whose records overflowed
// walk its bins instead (the same result).  skip_visible: the fragments of
// each pixel's visible triangle are left out (the shading kernel added them).
hipError_t launch_opacity_aux_records(const Views& v, int T, const AlphaSource& a, const float* rast,
                                      const float* pix_loss, float* grad_alpha, const RasterWorkspace& w,
                                      const int4* rec, const int* aux_n, int per_quad, bool skip_visible,
                                      hipStream_t s);
hipError_t launch_pixel_l1(long long P, int C, const float* color, const float* target,
                           float* pix_loss, hipStream_t s);

// Multires lattice gathers and scatters.
hipError_t launch_multires_color_fwd(long long P, const float* rast, int lo, int hi, int T, int C,
                                     const float* feat, int concat, float* out, hipStream_t s);
hipError_t launch_multires_color_bwd(long long P, const float* rast, int lo, int hi, int T, int C,
                                     float* grad_feat, int concat, const float* grad_out,
                                     hipStream_t s);
hipError_t launch_accumulate_fwd(int T, int lo, int hi, int target, int C, const float* feat,
                                 float* out, hipStream_t s);
hipError_t launch_accumulate_bwd(int T, int lo, int hi, int target, int C, float* grad_feat,
                                 const float* grad_out, hipStream_t s);

hipError_t launch_sh2(const Views& v, const float* rast, const float* inv_mvp, float* enc,
                      hipStream_t s);

hipError_t launch_edge_grad(const Views& v, int C, const float* color, const float* grad_color,
                            const float* rast, int V, const float* pos, float* grad_pos,
                            int T, const int* tri, hipStream_t s);

// The shading waves' ranges of the triangle-grouped pixel list, and the
// gradient rows the shading does not store whole.  The ranges split an
// estimated cost evenly, not the pixel count: a pixel costs 1 and each
// triangle's segment seg_cost more (its row staging, coarse-level reduction
// and row store, and where tiles stop at a triangle's end, on average half a
// tile of idle lanes), with the triangle index standing in for the number of
// non-empty segments before it.  Cost(t) = tri_off[t] + seg_cost t is monotone, so a 64-way search over
// triangles (one vector load per level) finds the segment holding each
// boundary.  bounds[b], b in [0, n_waves], is where shading wave b's range
// starts (bounds[0] = 0, bounds[n_waves] = the list length).  A triangle
// whose pixels all fall in one wave's range gets its gradient row stored
// whole by that wave; a triangle split between ranges is added into with
// atomics, and a triangle with no covered pixel is not visited: with grad
// non-null those rows are cleared here (rows of row floats).
struct ShadeRanges {
    const int* tri_off;   // [T + 1]
    int T;
    int n_waves;
    int* bounds;          // out [n_waves + 1]
    float* grad;          // [T, row] rows to clear, or null
    int row;
    long long seg_cost;   // the shading configuration's (shade_seg_cost)
};

// Wave-level jobs, wave-uniform: jobs [0, n_waves] are the boundaries (a
// boundary inside a triangle's segment also clears that row), then one job
// per 64 triangles clears the rows of those with no covered pixel.
__device__ __forceinline__ long long shade_range_jobs(const ShadeRanges& r) {
    return (long long)r.n_waves + 1 + (r.grad ? (r.T + 63) / 64 : 0);
}

__device__ __forceinline__ void clear_row(float* dst, int row, int lane) {
    if ((row & 3) == 0 && (reinterpret_cast<uintptr_t>(dst) & 15u) == 0) {
        float4* d4 = reinterpret_cast<float4*>(dst);
        for (int i = lane; i < row / 4; i += 64) d4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        for (int i = lane; i < row; i += 64) dst[i] = 0.f;
    }
}

__device__ __forceinline__ void shade_range_job(const ShadeRanges& r, long long job, int lane) {
    const int T = r.T;
    if (job > r.n_waves) {
        const int t0 = static_cast<int>(job - r.n_waves - 1) * 64;
        const int t = t0 + lane;
        unsigned long long empty = __ballot(t < T && r.tri_off[t] == r.tri_off[t + 1]);
        while (empty) {
            const int k = __builtin_ctzll(empty);
            empty &= empty - 1;
            clear_row(r.grad + (long long)(t0 + k) * r.row, r.row, lane);
        }
        return;
    }
    const int b = static_cast<int>(job);
    const int total = r.tri_off[T];
    const long long q = (total + r.seg_cost * T) * b / r.n_waves;
    // largest t in [0, T] with Cost(t) <= q
    int lo = 0, hi = T + 1;
    while (hi - lo > 1) {
        const int st = (hi - lo + 63) / 64;
        const int pr = lo + lane * st;
        const bool ok = pr < hi && r.tri_off[pr] + r.seg_cost * pr <= q;
        // lane 0 probes lo, which always qualifies
        lo += (63 - __builtin_clzll(__ballot(ok))) * st;
        hi = min(hi, lo + st);
    }
    int at;
    bool split = false;
    if (b == 0) {
        at = 0;
    } else if (b == r.n_waves || lo >= T) {
        at = total;
    } else {
        // the boundary's pixel inside that triangle's segment
        const int tb = r.tri_off[lo], tn = r.tri_off[lo + 1];
        const long long over = q - (tb + r.seg_cost * lo) - r.seg_cost;
        at = over <= 0 ? tb : static_cast<int>(min<long long>(tn, tb + over));
        split = at > tb && at < tn;
    }
    if (lane == 0) r.bounds[b] = at;
    if (split && r.grad) clear_row(r.grad + (long long)lo * r.row, r.row, lane);
}

// Covered pixels grouped by triangle: list[tri_off[t] + rank] = the pixel's
// PixRec (target colours from target [P,3]); with px_list, also the covered
// pixels in tile order at px_list[tile_px_off[tile] + k] = (pixel, view,
// visible triangle, 0).  Further workgroups of the same launch run the
// shading ranges' jobs (ShadeRanges).
hipError_t launch_fill_triangle_lists(const Views& v, const int2* tile_list, const int* tile_ncov,
                                      const float* rast, const float* target, const int* tri_off, PixRec* list,
                                      const int* tile_px_off, int4* px_list, const ShadeRanges& ranges,
                                      hipStream_t s);
// The shading ranges' jobs on their own (a backward without its forward's lists).
hipError_t launch_shade_ranges(const ShadeRanges& ranges, hipStream_t s);

// Over a covered-pixel list of (pixel, view, visible triangle, 0) entries,
// *n_list of them (read on the device, at most max_n); C must be 3.
hipError_t launch_edge_grad_sorted(const Views& v, int C, const float* color, const float* grad_color,
                                   const float* rast, int V, const float* pos, float* grad_pos, int T,
                                   const int* tri, const TriRecord* recs, const int4* px_list, const int* n_list,
                                   long long max_n, hipStream_t s);

hipError_t launch_fill_i32(int* p, long long n, int value, hipStream_t s);
hipError_t launch_fill_f32(float* p, long long n, float value, hipStream_t s);

// The training step's L1 + SSIM loss (ssim.hip): scratch floats needed, and
// the launch that writes dLoss/dcolour for every pixel and the loss.
size_t photo_loss_floats(const Views& views, int C);
hipError_t launch_photo_loss(const Views& views, int C, const float* target, const float* color, float* scratch,
                             float w_l1, float w_ssim, float* grad, float* loss, hipStream_t s);

}  // namespace ds

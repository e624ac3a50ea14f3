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
// walk its bins instead (the same result).
hipError_t launch_opacity_aux_records(const Views& v, int T, const AlphaSource& a, const float* rast,
                                      const float* pix_loss, float* grad_alpha, const RasterWorkspace& w,
                                      const int4* rec, const int* aux_n, int per_quad, hipStream_t s);
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

// Covered pixels grouped by triangle: list[tri_off[t] + rank] = the pixel's
// PixRec (target colours from target [P,3]); with px_list, also the covered
// pixels in tile order at px_list[tile_px_off[tile] + k] = (pixel, view,
// visible triangle, 0).
hipError_t launch_fill_triangle_lists(const Views& v, const int2* tile_list, const int* tile_ncov,
                                      const float* rast, const float* target, const int* tri_off, PixRec* list,
                                      const int* tile_px_off, int4* px_list, hipStream_t s);

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

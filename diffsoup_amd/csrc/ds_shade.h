// Launch-side interface of the fused shading kernels (shade.hip) and the
// training-step helpers (step.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ds_kernels.h"

namespace ds {

// Waves per shading workgroup (each takes one range of the pixel list).
constexpr int kShadeWavesPerBlock = 4;

// The shading kernel's per-lane weight operands ("blobs"), built once per
// step from the fp32 master weights W1 [DH, DIN], W2 [DH, DH], W3 [3, DH]:
// tiles of the six layer products L1, L2, L3 and the transposed L3T, L2T,
// L1T, (out tiles x in tiles) each, NI = ceil(DIN / 16) input and NH = DH / 16
// hidden tiles.  Entry (tile (layer, m, t), lane, e) is the A operand for
// output row 16 m + lane % 16 and input row 16 t + 4 (lane / 16) + e; a tile
// holds 4 x 64 floats (f32: step e, lane), or 64 x 4 bf16 (lane, e).
__host__ __device__ constexpr int blob_tiles(int ni, int nh) { return nh * ni * 2 + nh * nh * 2 + nh * 2; }
__host__ __device__ constexpr int blob_tile_floats(bool bf16) { return (bf16 ? 2 : 4) * 64; }

struct BlobJob {
    const float* w1;
    const float* w2;
    const float* w3;
    float* blobs;
    int ni, nh, din, dh;
    int bf16;
    __host__ __device__ long long n() const { return (long long)blob_tiles(ni, nh) * 64; }
};

// One (tile, lane) entry of the blobs.
__device__ __forceinline__ void build_blob(const BlobJob& b, int id) {
    const int NI = b.ni, NH = b.nh, DIN = b.din, DH = b.dh;
    const int L2 = NH * NI, L3 = L2 + NH * NH, L3T = L3 + NH, L2T = L3T + NH, L1T = L2T + NH * NH;
    const int tile = id / 64, lane = id % 64;
    int base, nin, layer;
    if (tile < L2) { layer = 0; base = 0; nin = NI; }
    else if (tile < L3) { layer = 1; base = L2; nin = NH; }
    else if (tile < L3T) { layer = 2; base = L3; nin = NH; }
    else if (tile < L2T) { layer = 3; base = L3T; nin = 1; }
    else if (tile < L1T) { layer = 4; base = L2T; nin = NH; }
    else { layer = 5; base = L1T; nin = NH; }
    const int m = (tile - base) / nin, t = (tile - base) % nin;
    const int o = 16 * m + lane % 16;
    float v[4];
    for (int e = 0; e < 4; ++e) {
        const int i = 16 * t + 4 * (lane / 16) + e;
        float w = 0.f;
        switch (layer) {
            case 0: w = (i < DIN) ? b.w1[o * DIN + i] : 0.f; break;
            case 1: w = b.w2[o * DH + i]; break;
            case 2: w = (o < 3) ? b.w3[o * DH + i] : 0.f; break;
            case 3: w = (i < 3) ? b.w3[i * DH + o] : 0.f; break;
            case 4: w = b.w2[i * DH + o]; break;
            default: w = (o < DIN) ? b.w1[i * DIN + o] : 0.f; break;
        }
        v[e] = w;
    }
    float* dst = b.blobs + (long long)tile * blob_tile_floats(b.bf16 != 0);
    if (b.bf16) {
        unsigned short* d16 = reinterpret_cast<unsigned short*>(dst);
        for (int e = 0; e < 4; ++e) {   // round to nearest even
            const uint32_t u = __float_as_uint(v[e]);
            d16[lane * 4 + e] = static_cast<unsigned short>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
        }
    } else {
        for (int e = 0; e < 4; ++e) dst[e * 64 + lane] = v[e];
    }
}

struct ShadeArgs {
    const float4* rast;     // [P] (b0, b1, depth, id + 1)
    const int* tri_off;     // [T + 1] start of each triangle's covered pixels in list
    const PixRec* list;     // covered pixels grouped by triangle
    int T;
    const float* feat;      // [T, S_tot, C]
    const unsigned short* feat_bf16;   // bf16 copy read instead of feat when non-null
    float* grad_feat;       // [T, S_tot, C] written (every row)
    int S_tot;
    int lo;                 // first feature level
    const float* blobs;     // per-lane weight operands (k_blobs)
    const float* b1;
    const float* b2;
    const float* b3;
    float* color;           // [P, 3]
    float* grad_color;      // [P, 3]
    float* pix_loss;        // [P]
    float grad_scale;       // 2 / (3 P)
    float* partials;        // [nblocks, nparam]
    // Two-pass use for a loss that couples pixels (L1 + SSIM): with fwd_only
    // the kernel writes colour, the MSE gradient and the per-pixel L1 and
    // skips the adjoint; with ext_grad [P, 3] it reads dLoss/dcolour from
    // there (and writes no colour, gradient, L1 or loss) and runs the adjoint.
    const float* ext_grad;
    int fwd_only;
    // SH2 inputs (configurations compiled with them): the views, to find a
    // covered pixel's view, row and column, and their inverse MVPs [B,4,4].
    Views views;
    const float* inv_mvp;
    // where each wave's range of list starts: [nblocks * kShadeWavesPerBlock
    // + 1], the ShadeRanges bounds for the same nblocks
    const int* bounds;
    // The visible fragments' opacity gradient (alpha null: not taken here):
    // the post-sigmoid opacity lattice [T, Sa] at alpha_level and its
    // gradient table (zeroed before).  For every covered pixel the fused
    // single-pass kernel adds the estimator's term of the pixel's own
    // triangle, L1 / max(alpha, 1e-5) on the three lattice corners of its
    // barycentrics, into an LDS row per triangle that is stored once; the
    // opacity-gradient pass then takes only the fragments of other triangles.
    const float* alpha;
    float* grad_alpha;
    int alpha_level;
    int Sa;
};

bool shade_supported(int nl, int c, int dh, bool sh);
size_t shade_blob_floats(int nl, int c, int dh, bool sh);
int shade_nparam(int nl, int c, int dh, bool sh);
// Resident shading workgroups per compute unit of a compiled configuration
// (the kernel's register budget: two up to hidden width 32, else one).
int shade_blocks_per_cu(int nl, int c, int dh, bool sh);
// Cost of a triangle's segment in pixels, for the shading waves' ranges.
long long shade_seg_cost(int nl, int c, int dh, bool sh);

// visible_alpha (optional) receives whether this launch took the visible
// fragments' opacity gradient (a.alpha set, the fused single-pass form, and
// the per-wave LDS rows fitting the plan).
hipError_t launch_shade(int nl, int c, int dh, bool sh, bool bf16, const ShadeArgs& a, int nblocks,
                        hipStream_t s, bool* visible_alpha = nullptr);
// The waves' ranges of the triangle-grouped list (before any shading launch).

// Whether a configuration gathers and scatters its rows in global memory
// (rows too large for the LDS plan): its gradient table must then be zeroed.
bool shade_direct_rows(int nl, int c, int dh, bool sh, bool bf16, bool feat_bf16, int S_tot);

}  // namespace ds

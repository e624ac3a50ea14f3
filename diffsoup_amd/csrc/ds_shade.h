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
};

bool shade_supported(int nl, int c, int dh, bool sh);
size_t shade_blob_floats(int nl, int c, int dh, bool sh);
int shade_nparam(int nl, int c, int dh, bool sh);
// Resident shading workgroups per compute unit of a compiled configuration
// (the kernel's register budget: two up to hidden width 32, else one).
int shade_blocks_per_cu(int nl, int c, int dh, bool sh);
hipError_t launch_shade_blobs(int nl, int c, int dh, bool sh, bool bf16, const float* w1, const float* w2,
                              const float* w3, float* blobs, hipStream_t s);
hipError_t launch_shade(int nl, int c, int dh, bool sh, bool bf16, const ShadeArgs& a, int nblocks,
                        hipStream_t s);
// The waves' ranges of the triangle-grouped list (before any shading launch).

// Whether a configuration gathers and scatters its rows in global memory
// (rows too large for the LDS plan): its gradient table must then be zeroed.
bool shade_direct_rows(int nl, int c, int dh, bool sh, bool bf16, bool feat_bf16, int S_tot);

}  // namespace ds

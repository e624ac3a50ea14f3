/*
 * diffsoup_hip.h - C-ABI of libdiffsoup_hip.so, the MI355X implementation of
 * DiffSoup's differentiable triangle-soup rasterizer.
 *
 * Conventions shared by every entry point:
 *  - pointers are device pointers unless the parameter is documented as a
 *    host array (the per-view H / W resolution arrays, the SSIM window and
 *    scales, and everything the host-side soup splitters take);
 *  - `stream` is a hipStream_t (pass torch.cuda.current_stream().cuda_stream);
 *    all work is enqueued on it, nothing synchronises the host;
 *  - nothing allocates device memory; scratch space is a caller-owned,
 *    256-byte aligned workspace sized by ds_raster_workspace_size (the SSIM
 *    maps and partial sums are caller-owned outputs too);
 *  - the return value is DS_OK (0) or a negative DS_ERR_* code, with a
 *    message in ds_last_error() (thread-local); nothing calls exit();
 *  - views of one batch may differ in resolution: view b's pixels occupy
 *    [sum_{b'<b} H[b']W[b'], ...) of every packed per-pixel buffer;
 *  - rast holds per pixel (b0, b1, depth = z/w, triangle id + 1), background
 *    all zeros, exactly as the reference's rasterization output.
 *
 * Each declaration names the reference interface it replaces
 * (kenji-tojo/diffsoup, path:line).
 */
#ifndef DIFFSOUP_HIP_H
#define DIFFSOUP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DS_OK 0
#define DS_ERR_INVALID -1   /* bad shape, size, pointer or option            */
#define DS_ERR_HIP -2       /* a HIP launch error                            */
#define DS_ERR_WORKSPACE -4 /* workspace smaller than ds_*_workspace_size    */
#define DS_ERR_UNSUPPORTED -5 /* a configuration without a compiled kernel   */

/* Device-side counters at the start of the raster workspace (int64 each). */
#define DS_STAT_COVERED 0  /* fragments inside with depth in [-1,1], pre-mask */
#define DS_STAT_PAIRS 1    /* (triangle, tile) pairs produced by binning       */
#define DS_STAT_OVERFLOW 2 /* 1 if pairs exceeded the bin capacity             */
#define DS_STAT_BAD_INDEX 3 /* 1 if a triangle indexed a vertex outside [0,V)  */
#define DS_STAT_AUX_WALK_QUADS 4 /* training step: 8x8 quads whose opacity-
                                    gradient records overflowed (ds_step_desc.
                                    aux_records) and were walked again   */

/* Library version string.  Replaces _core.__version__ (src/main.cpp:26). */
const char* ds_version(void);

/* Message of the last failing call on this thread.  Replaces the
 * CUDA_CHECK print-and-exit (src/cuda/cuda_common.cuh:19-27) and the
 * std::runtime_error shape checks (src/main.cpp:183-185). */
const char* ds_last_error(void);

/* Lattice samples per triangle at one level: 3 at level 0, else
 * (2^(L-1)+1)(2^L+1).  Same as feats_at_level (src/cuda/multires.cuh:8-10,
 * python/diffsoup/multires.py:16-23). */
uint32_t ds_feats_at_level(int level);

/* Number of screen tiles of edge `tile` (8, 16 or 32) over B views with host
 * resolution arrays H[B], W[B]; -1 on invalid input. */
int64_t ds_raster_tile_count(int B, const int* H, const int* W, int tile);

/* Bytes of workspace for ds_rasterize_fwd / ds_opacity_aux_bwd with room for
 * `capacity` (triangle, tile) pairs.  When a batch produces more pairs the
 * kernels still return the exact image through a slower path and set
 * DS_STAT_OVERFLOW; grow the capacity for the next call. */
size_t ds_raster_workspace_size(int B, const int* H, const int* W, int T, int tile,
                                int64_t capacity);

/* Byte offset of the int64 stats block (DS_STAT_*) inside the workspace. */
int ds_raster_stats_offset(void);

/* Rasterize B views of one triangle soup with per-fragment multires opacity,
 * stochastic masking and a depth test.  pos [B,V,4] clip space (16-byte
 * aligned), tri [T,3] int32, alpha_src [T, feats_at_level(level)] post-sigmoid
 * opacities, rast [P,4] output.  stochastic != 0 keeps a fragment when
 * alpha >= u with u a hash of (seed, view, triangle, row, column);
 * stochastic == 0 uses u = 0.5.
 * Replaces _core.compute_triangle_rects, compute_fragments,
 * filter_valid_fragments, multires_triangle_alpha and depth_test
 * (src/main.cpp:30-108,174-191) as composed by
 * rasterize_multires_triangle_alpha (python/diffsoup/multires.py:52-102). */
int ds_rasterize_fwd(int B, const int* H, const int* W, int V, const float* pos, int T,
                     const int* tri, int level, const float* alpha_src, int stochastic,
                     uint64_t seed, int tile, float* rast, void* workspace,
                     size_t workspace_bytes, int64_t capacity, void* stream);

/* Bytes of the fixed device scratch behind ds_rasterize_fwd_noalloc. */
#define DS_NOALLOC_SCRATCH_BYTES (64u << 20)

/* ds_rasterize_fwd without a workspace argument, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The workspace
 * is a fixed device array of DS_NOALLOC_SCRATCH_BYTES (64 MiB) per GPU,
 * held until the process ends; the call allocates nothing and takes the
 * (triangle, tile) bin capacity from what the records and tile tables leave
 * of it (bins that overflow it still give the exact image through the
 * slower path).  Batches whose records and tile tables alone exceed it
 * (views x triangles x 64 B + 12 B per tile) are rejected with
 * DS_ERR_INVALID.  The scratch is that of the device current at the call;
 * every call on one GPU shares it, so calls that overlap on one GPU (two
 * streams, two threads) race.  The stats block it leaves is not readable
 * through this entry point. */
int ds_rasterize_fwd_noalloc(int B, const int* H, const int* W, int V, const float* pos, int T,
                             const int* tri, int level, const float* alpha_src, int stochastic,
                             uint64_t seed, int tile, float* rast, void* stream);

/* Opacity auxiliary gradient: regenerates every fragment, and for those not
 * strictly behind the visible surface adds the score-function gradient
 * +l/max(a,1e-5) (visible triangle) or -l/max(1-a,1e-5) (others) onto
 * grad_alpha [T, feats_at_level(level)] (accumulated, not scaled), with l the
 * per-pixel loss pix_loss [P].  reuse_bins != 0 reuses the bins that the
 * previous ds_rasterize_fwd left in the same workspace for the same pos/tri.
 * Replaces _core.backward_opacity_aux_loss + backward_multires_triangle_alpha
 * (src/main.cpp:134-154,193-210) as composed by _OpacityAuxLossFn.forward
 * (python/diffsoup/rasterize.py:148-190). */
int ds_opacity_aux_bwd(int B, const int* H, const int* W, int V, const float* pos, int T,
                       const int* tri, int level, const float* alpha_src, int tile,
                       const float* rast, const float* pix_loss, float* grad_alpha,
                       int reuse_bins, void* workspace, size_t workspace_bytes,
                       int64_t capacity, void* stream);

/* pix_loss[p] = sum_c |color[p,c] - target[p,c]| in channel order: the
 * per-pixel L1 inside backward_opacity_aux_loss_kernel
 * (src/cuda/rasterize.cu:955-960). */
int ds_pixel_l1(int64_t P, int C, const float* color, const float* target, float* pix_loss,
                void* stream);

/* Per-pixel multires feature gather over levels [lo, hi] of feat
 * [T, S_tot, C].  concat == 0 sums the levels into out [P, C] (the
 * reference); concat != 0 writes out [P, (hi-lo+1)*C], one slice per level.
 * Background pixels are written as zeros.  Replaces
 * _core.multires_triangle_color (src/main.cpp:212-232). */
int ds_multires_color_fwd(int64_t P, const float* rast, int lo, int hi, int T, int C,
                          const float* feat, int concat, float* out, void* stream);

/* Adjoint of ds_multires_color_fwd, accumulated into grad_feat with atomics.
 * Replaces _core.backward_multires_triangle_color (src/main.cpp:234-254). */
int ds_multires_color_bwd(int64_t P, const float* rast, int lo, int hi, int T, int C,
                          float* grad_feat, int concat, const float* grad_out, void* stream);

/* Evaluate the level-summed multires function of feat [T, S_tot(lo..hi), C]
 * at the vertices of the target-level lattice: out [T, feats_at_level(target), C].
 * Replaces _core.accumulate_to_level_forward (src/main.cpp:258-280). */
int ds_accumulate_fwd(int T, int lo, int hi, int target, int C, const float* feat, float* out,
                      void* stream);

/* Adjoint of ds_accumulate_fwd accumulated into grad_feat.  Replaces
 * _core.accumulate_to_level_backward (src/main.cpp:282-304). */
int ds_accumulate_bwd(int T, int lo, int hi, int target, int C, float* grad_feat,
                      const float* grad_out, void* stream);

/* Degree-2 real SH basis of the per-pixel view direction, enc [P, 9], zero on
 * background; inv_mvp [B,4,4] row-major.  Replaces _core.encode_view_dir_sh2
 * (src/main.cpp:158-170). */
int ds_sh2(int B, const int* H, const int* W, const float* rast, const float* inv_mvp, float* enc,
           void* stream);

/* Edge gradients w.r.t. clip-space positions, accumulated into grad_pos
 * [B,V,4]; color and grad_color are [P, C].  Replaces
 * _core.backward_edge_grad (src/main.cpp:112-130). */
int ds_edge_grad_bwd(int B, const int* H, const int* W, int C, const float* color,
                     const float* grad_color, const float* rast, int V, const float* pos,
                     float* grad_pos, int T, const int* tri, void* stream);

/* ------------------------------------------------------------------------
 * Fused training step.  One call enqueues a full forward + backward of the
 * reference's training iteration (examples/02_synthetic.py:354-388) for B
 * views of possibly different resolution: projection, rasterization with
 * stochastic opacity, per-level concatenated multires features, the pixel
 * MLP (Linear-ReLU-Linear-ReLU-Linear-Sigmoid, hidden width `hidden`, input
 * (feat_hi-feat_lo+1)*channels, 3 outputs) with the residual blend of
 * ColorMLP (python/diffsoup/multires.py:185-207), a constant background,
 * the MSE loss against `target`, the opacity auxiliary gradient, edge
 * gradients, and all parameter gradients.  bf16 != 0 runs the MLP on bf16
 * MFMA with fp32 accumulation; everything else is fp32.
 * Replaces the Python composition of the reference's ops; it has no single
 * reference counterpart.  All pointers are device pointers except H and W.
 * ------------------------------------------------------------------------ */
typedef struct ds_step_desc {
    int B;                    /* views                                       */
    const int* H;             /* host [B]                                    */
    const int* W;             /* host [B]                                    */
    int tile;                 /* screen tile edge: 8, 16 or 32               */
    int V;                    /* vertices                                    */
    const float* verts;       /* [V,3] world positions                       */
    const float* mvp;         /* [B,4,4] row-major                           */
    int T;                    /* triangles                                   */
    const int* tri;           /* [T,3]                                       */
    int alpha_level;          /* opacity lattice level                       */
    const float* alpha_logits;/* [T, feats_at_level(alpha_level)]            */
    int stochastic;           /* hash mask (1) or threshold 0.5 (0)          */
    uint64_t seed;
    int feat_lo, feat_hi;     /* feature levels, concatenated in the MLP     */
    int channels;             /* channels per level                          */
    const float* feat;        /* [T, S_tot, channels]                        */
    int hidden;               /* MLP hidden width                            */
    int bf16;                 /* MLP precision                               */
    const float* w1; const float* b1;  /* [hidden, in], [hidden]             */
    const float* w2; const float* b2;  /* [hidden, hidden], [hidden]         */
    const float* w3; const float* b3;  /* [3, hidden], [3]                   */
    const float* target;      /* [P,3] packed over views                     */
    float background;         /* background colour                           */
    float* pos;               /* out [B,V,4] clip positions                  */
    float* alpha;             /* out [T,S_a] sigmoid(alpha_logits)           */
    float* rast;              /* out [P,4]                                   */
    float* color;             /* out [P,3]                                   */
    float* grad_color;        /* out [P,3] dLoss/dcolor                      */
    float* pix_loss;          /* out [P] per-pixel L1                        */
    float* grad_verts;        /* out [V,3]                                   */
    float* grad_alpha_logits; /* out [T,S_a]                                 */
    float* grad_feat;         /* out [T,S_tot,channels]                      */
    float* grad_mlp;          /* out [ds_step_mlp_params]: dW1,dW2,dW3,db1,db2,db3,loss_sum */
    float* loss;              /* out [1] the loss (MSE unless loss_kind)     */
    void* workspace;          /* 256-byte aligned                            */
    size_t workspace_bytes;
    int64_t capacity;         /* (triangle, tile) bin capacity               */
    int timed_stage;          /* DS_STAGE_* bracketed by the events, or 0    */
    void* ev_begin;           /* hipEvent_t recorded before that stage       */
    void* ev_end;             /* hipEvent_t recorded after it                */
    const uint64_t* seed_device; /* if non-null, the mask seed is read from this
                                    device uint64 when the step runs (seed is
                                    then ignored): a captured graph replays with
                                    whatever seed the caller last stored there */
    const uint16_t* feat_bf16;   /* if non-null, the features are read from this
                                    bf16 copy [T, S_tot, channels] (round to
                                    nearest even of the fp32 master) instead of
                                    feat; grad_feat stays fp32.  With the bf16
                                    MLP, channels 0-3 of the first level (the
                                    residual-blend inputs, which reach the
                                    colour directly) are still read from feat,
                                    which must then be given too */
    int loss_kind;               /* DS_LOSS_MSE (0): mean squared error; or
                                    DS_LOSS_L1_SSIM: the reference's
                                    0.8 L1 + 0.2 (1 - SSIM) / 2 (examples/
                                    02_synthetic.py:376-380), SSIM the mean over
                                    views of pytorch_msssim.ssim per view; loss
                                    then holds that value and grad_color its
                                    gradient (pix_loss stays the per-pixel L1) */
    int advance_seed;            /* with seed_device: the step adds 1 to
                                    *seed_device after its last read, so that
                                    replays of a captured step draw a new mask
                                    each without a launch of their own */
    int accumulate;              /* 0: the features of levels feat_lo..feat_hi
                                    are concatenated into the MLP input and the
                                    opacities are sigmoid(alpha_logits) at
                                    alpha_level.  1: the reference's own
                                    composition (examples/02_synthetic.py:
                                    354-364): feat [T, S_tot(feat_lo..feat_hi),
                                    channels] and alpha_logits [T, S_tot(
                                    alpha_lo..alpha_level)] are the multires
                                    sources; the step evaluates
                                    sigmoid(accumulate_to_level(...)) of both at
                                    feat_hi / alpha_level, gathers the features
                                    at level feat_hi (channels inputs), and
                                    grad_feat / grad_alpha_logits are the
                                    gradients of the sources; alpha then holds
                                    the accumulated opacities [T, S_a] */
    int alpha_lo;                /* with accumulate: first opacity level      */
    int sh2;                     /* append the 9 SH2 values of the pixel's view
                                    direction (encode_view_dir_sh2,
                                    src/cuda/encoding.cu:9-130) to the MLP
                                    input after the features                  */
    const float* inv_mvp;        /* with sh2: [B,4,4] row-major inverse MVPs  */
    int64_t aux_records;         /* fragment records per 8x8 pixel quad the
                                    workspace holds for the opacity gradient
                                    (0: 128, two per pixel): the resolve lists
                                    every fragment that may lie in front of
                                    its pixel's surface, and a quad whose
                                    fragments overflow its slots walks its
                                    tile's bins again instead (the same
                                    gradient, more time)                     */
    int64_t* advance_counter;    /* if non-null, a device int64 the step adds 1
                                    to in its last kernel: the step counter of
                                    the optimiser update that follows it on the
                                    same stream, which then reads it as
                                    already advanced (ds_adam_step_advanced),
                                    so a captured iteration needs no launch of
                                    its own for the counter                  */
} ds_step_desc;

#define DS_LOSS_MSE 0
#define DS_LOSS_L1_SSIM 1

/* Stages of ds_train_step that timed_stage can select. */
#define DS_STAGE_PREP 1     /* sigmoid, zero accumulators                    */
#define DS_STAGE_BLOBS 2    /* MLP weight operands                           */
#define DS_STAGE_PROJECT 3  /* clip-space projection                         */
#define DS_STAGE_BIN 4      /* triangle setup, tile counts, scan, bin emit   */
#define DS_STAGE_RESOLVE 5  /* depth test + background shading               */
#define DS_STAGE_SHADE 6    /* gather + MLP + loss + adjoint + scatter       */
#define DS_STAGE_EDGE 7     /* edge gradients                                */
#define DS_STAGE_AUX 8      /* opacity auxiliary gradient                    */
#define DS_STAGE_FINAL 9    /* reductions, loss, vertex / logit gradients    */

/* Length of the grad_mlp vector for a configuration (levels concatenated,
 * no SH2 inputs; with sh2 != 0, 9 SH2 inputs after the features). */
int64_t ds_step_mlp_params(int levels, int channels, int hidden);
int64_t ds_step_mlp_params_sh2(int levels, int channels, int hidden, int sh2);

/* Workspace bytes for ds_train_step; 0 if the configuration is invalid or
 * has no compiled shading kernel. */
size_t ds_step_workspace_size(const ds_step_desc* d);

/* Enqueue one fused forward + backward step (see above). */
int ds_train_step(const ds_step_desc* d, void* stream);

/* The step split at the colour, for a caller whose loss is its own (the
 * operator behind a torch.autograd.Function): ds_step_forward enqueues the
 * forward up to the colour (color [P,3], pix_loss, rast, alpha, pos; the
 * descriptor's loss_kind is ignored); ds_step_backward, on the same
 * descriptor and workspace after it, takes dLoss/dcolor [P,3] (device) and
 * enqueues the MLP adjoint, the feature scatter, the edge and opacity
 * gradients (the opacity term of the reference's opacity_aux_loss, weight 1,
 * python/diffsoup/rasterize.py:139-252, included) and every parameter
 * gradient; it does not write loss.  Replaces the autograd backward of the
 * reference's composed ops (examples/02_synthetic.py:354-385). */
int ds_step_forward(const ds_step_desc* d, void* stream);
int ds_step_backward(const ds_step_desc* d, const float* grad_color, void* stream);

/* Tile autotuner: for each candidate tile edge (8, 16 or 32; at most 8
 * candidates) whose workspace fits d->workspace_bytes, runs one warm-up step
 * and `reps` timed steps on `stream` (events; this call synchronises the
 * host), writes the mean milliseconds per step to ms[i] (-1 for a skipped
 * candidate) and leaves d->tile set to the fastest.  The steps write the
 * descriptor's outputs like any ds_train_step.  Size the workspace with the
 * largest ds_step_workspace_size over the candidates.  No counterpart in the
 * reference (its rasterizer has no tiles). */
int ds_autotune_step(ds_step_desc* d, const int* tiles, int n_tiles, int reps, float* ms, void* stream);

/* ------------------------------------------------------------------------
 * Plain path: the reference's fragment-list rasterizer, one function per
 * compiled function its Python layer binds, for callers that drive the
 * reference's composition (python/diffsoup/rasterize.py:14-126) and as the
 * comparator of the tile-binned path.  Fragment buffers are caller-owned;
 * frag_pix [N,3] int32 (view, y, x; view -1 = invalid), frag_attrs [N,4]
 * (b0, b1, z/w, triangle id + 1).  Counts come back through a blocking copy,
 * as in the reference.  scan_tmp holds ds_plain_scan_tmp_bytes(n) bytes.
 * ------------------------------------------------------------------------ */

/* Bytes of scan scratch for n items. */
size_t ds_plain_scan_tmp_bytes(int64_t n);

/* rects [B*T,4] (h0, h_len, w0, w_len), counts [B*T], inclusive int64 prefix
 * [B*T]; returns the number of candidate fragments or a negative DS_ERR_*.
 * Replaces _core.compute_triangle_rects (src/main.cpp:30-48). */
int64_t ds_plain_triangle_rects(int H, int W, int B, int V, const float* pos, int T, const int* tri,
                                int* rects, int64_t* prefix, int* counts, void* scan_tmp, void* stream);

/* Candidate fragments.  Replaces _core.compute_fragments (src/main.cpp:50-76). */
int ds_plain_fragments(int H, int W, int B, int V, const float* pos, int T, const int* tri,
                       const int64_t* prefix, const int* rects, int64_t num_frags, int* frag_pix,
                       float* frag_attrs, void* stream);

/* Order-preserving compaction of valid fragments; flags [num_frags] int32
 * scratch; returns the valid count.  Replaces _core.filter_valid_fragments
 * (src/main.cpp:96-108) with a deterministic output order. */
int64_t ds_plain_compact(int64_t num_frags, const int* frag_pix, const float* frag_attrs, int* frag_pix_out,
                         float* frag_attrs_out, int* flags, void* scan_tmp, void* stream);

/* Per-fragment opacity, summed over levels min_level..max_level of
 * alpha_src [T, S] (S = sum of ds_feats_at_level over the range); fragments
 * with id 0 get 0.  Replaces _core.multires_triangle_alpha
 * (src/main.cpp:174-191). */
int ds_plain_alpha(int64_t n, const float* frag_attrs, int min_level, int max_level, const float* alpha_src,
                   float* frag_alpha, void* stream);

/* Adjoint of ds_plain_alpha (atomic accumulation).  Replaces
 * _core.backward_multires_triangle_alpha (src/main.cpp:193-210). */
int ds_plain_alpha_bwd(int64_t n, const float* frag_attrs, int min_level, int max_level, float* grad_alpha_src,
                       const float* grad_frag_alpha, void* stream);

/* Per-fragment mask thresholds: the hash of (seed, view_base + view,
 * triangle, row, column), or 0.5 when stochastic == 0.  Replaces the
 * torch.rand / torch.full thresholds of python/diffsoup/multires.py:97-100. */
int ds_plain_thresholds(int64_t n, const int* frag_pix, const float* frag_attrs, int stochastic, uint64_t seed,
                        int view_base, float* thresh, void* stream);

/* Depth test with a global 64-bit atomicMin into zbuf [B*H*W] int64 and a
 * gather into rast [B,H,W,4].  Replaces _core.depth_test (src/main.cpp:78-94). */
int ds_plain_depth_test(int B, int H, int W, int64_t n, const int* frag_pix, const float* frag_attrs,
                        const float* frag_alpha, const float* thresh, int64_t* zbuf, float* rast, void* stream);

/* Per-fragment opacity-aux gradient accumulated into grad_frag_alpha.
 * Replaces _core.backward_opacity_aux_loss (src/main.cpp:134-154). */
int ds_plain_opacity_aux(int B, int H, int W, int C, const float* color, const float* target, const float* rast,
                         int64_t n, const int* frag_pix, const float* frag_attrs, const float* frag_alpha,
                         float* grad_frag_alpha, void* stream);

/* ------------------------------------------------------------------------
 * Fused optimiser step.  One launch updates every parameter group: groups
 * with vector_dim == 0 get torch.optim.Adam's update (element-wise second
 * moment; the optimiser of the features, opacities and MLP in
 * examples/02_synthetic.py:310-315), groups with vector_dim == k > 0 get
 * VectorAdam's (python/diffsoup/optimize.py:33-68: second moment pooled over
 * the last dimension of width k; eps is the reference's 1e-8).  *step is a
 * device int64 holding the number of steps taken so far; the call uses
 * *step + 1 for the bias corrections and then increments it, so a captured
 * graph advances it on every replay.  exp_avg has the parameter's shape,
 * exp_avg_sq has n / k entries for vector groups.  Replaces
 * torch.optim.Adam.step and VectorAdam.step.
 * ------------------------------------------------------------------------ */
typedef struct ds_adam_group {
    float* param;
    const float* grad;
    float* exp_avg;
    float* exp_avg_sq;
    int64_t n;        /* floats in param                                    */
    int vector_dim;   /* 0: Adam; k > 0: VectorAdam over vectors of k floats */
    double lr, beta1, beta2, eps;
    float* param_out; /* NULL: update param in place; else write the updated
                         values here and leave param unchanged            */
    uint16_t* param_bf16_out; /* if non-null, also write the updated values
                         rounded to bf16 (nearest even): the copy a step
                         with feat_bf16 reads                             */
    const float* lr_mult; /* if non-null, a device float read when the update
                         runs; the learning rate is multiplied by it (a
                         captured graph picks up whatever the caller last
                         stored there)                                      */
    int64_t decay_steps; /* if > 0, the learning rate is also multiplied by
                         decay_final ^ (i / decay_steps), i = *step + 1
                         clamped to [1, decay_steps]: the reference's
                         exp_decay_mult(i, decay_steps, decay_final)
                         (examples/utils.py:435-442) at iteration i, formed
                         on the device from the step counter, so replays
                         follow the schedule with no host involvement      */
    double decay_final;
} ds_adam_group;

/* At most 16 groups. */
int ds_adam_step(const ds_adam_group* groups, int n_groups, int64_t* step, void* stream);

/* ds_adam_step for a counter that something earlier on the stream has
 * already advanced for this update (ds_step_desc.advance_counter): the
 * update is iteration *step (not *step + 1) and the counter is left as it
 * is.  No counterpart in the reference: for callers that capture the step and
 * the update in one graph and would otherwise spend a launch on the counter. */
int ds_adam_step_advanced(const ds_adam_group* groups, int n_groups, int64_t* step, void* stream);

/* Per-triangle pixel counts of a raster rast [P,4] (channel 3 = triangle id
 * + 1, 0 = background): counts [T] int64 receives the bincount of the ids
 * (added to the existing values when accumulate != 0, else overwritten).
 * Replaces count_triangle_ids (python/diffsoup/rasterize.py:348-366), the
 * kernel of count_visible_triangles (examples/utils.py:445-482). */
int ds_count_triangle_ids(int64_t P, const float* rast, int T, int64_t* counts, int accumulate, void* stream);

/* Fill n floats with value (a kernel, so it is safe inside graph capture). */
int ds_fill_f32(float* p, int64_t n, float value, void* stream);

/* ------------------------------------------------------------------------
 * Structural similarity of pytorch_msssim.ssim (a symmetric window of
 * win_size = 11 taps given in the host array win, as pytorch_msssim's
 * default Gaussian,
 * "valid" filtering, a side shorter than the window left unfiltered,
 * c1 = (K1 R)^2, c2 = (K2 R)^2), the term of the
 * reference's loss 0.8 L1 + 0.2 (1 - SSIM) / 2 (examples/02_synthetic.py:
 * 376-380).  X (reference image) and Y are [P, C] packed over B views like
 * rast, 1 <= C <= 4.
 *
 * ds_ssim_tiles: the number of window tiles (the length of the partial-sum
 * array, times C); tile_base (host, B + 1 entries, may be NULL) receives
 * where each view's tiles start.
 * ds_ssim_fwd: per-tile, per-channel sums of the SSIM map into partial
 * [tiles, C], and the three derivative maps into maps [3, C, P]
 * (channel-planar; entries outside the windows are not written).
 * ds_ssim_bwd: grad [P, C] = scale[b, c] * dscale[b, c] * dSSIM_sum/dY +
 * l1_scale * dl1 * sign(Y - X), stored or added (accumulate).  scale is a
 * host array [B, C]; dscale (device [B, C]) and dl1 (device scalar) may be
 * NULL (= 1).  For the mean SSIM of one view, scale = 1 / (C H' W').
 * ------------------------------------------------------------------------ */
int64_t ds_ssim_tiles(int B, const int* H, const int* W, int win_size, int64_t* tile_base);
int ds_ssim_fwd(int B, const int* H, const int* W, int C, const float* X, const float* Y, int win_size,
                const float* win, float c1, float c2, float* maps, float* partial, void* stream);
int ds_ssim_bwd(int B, const int* H, const int* W, int C, const float* X, const float* Y, int win_size,
                const float* win, const float* maps, const float* scale, const float* dscale, float l1_scale,
                const float* dl1, float* grad, int accumulate, void* stream);

/* ------------------------------------------------------------------------
 * Soup splitting (host).  Longest-edge bisection between training phases;
 * the outputs are bit-identical to the reference's splitters, in the same
 * order.  Unlike every other entry point these run on the host, take host
 * pointers and allocate: the result lives in an opaque ds_soup_split until
 * ds_split_soup_free.  Sizes are read with ds_split_soup_counts and the
 * arrays copied out with ds_split_soup_export (any output may be NULL):
 * verts [nv,3] f32, faces [nt,3] i32, face_map [nt] (source face of each
 * face), same_as_original [nt] (1 = the source face unchanged).
 *
 * ds_split_soup: world-space edge lengths.  Replaces split_triangle_soup
 * (src/main.cpp:308-345, src/remesh.cpp:118-156): at most num_splits
 * bisections (num_splits < 0: no cap, then tau must be > 0), stopping once
 * the longest edge is shorter than tau (tau <= 0: no threshold).
 *
 * ds_split_soup_clip: lengths in NDC through one row-major mvp [4,4], x
 * scaled by aspect_wh = W/H, threshold tau_ratio in image heights; edges
 * with an end outside the NDC cube are never split; valid_faces [nt] (NULL:
 * all) selects the faces whose edges are considered.  Replaces
 * split_triangle_soup_clip (src/main.cpp:347-385, src/remesh_clip.cpp:200-244).
 * ------------------------------------------------------------------------ */
typedef struct ds_soup_split ds_soup_split;

int ds_split_soup(const float* verts, int nv, const int* faces, int nt, int num_splits, float tau,
                  ds_soup_split** out);
int ds_split_soup_clip(const float* mvp, const float* verts, int nv, const int* faces, int nt,
                       const int* valid_faces, int num_splits, float tau_ratio, float aspect_wh,
                       ds_soup_split** out);
int ds_split_soup_counts(const ds_soup_split* s, int* nv, int* nt);
int ds_split_soup_export(const ds_soup_split* s, float* verts, int* faces, int* face_map, int* same_as_original);
void ds_split_soup_free(ds_soup_split* s);

#ifdef __cplusplus
}
#endif

#endif /* DIFFSOUP_HIP_H */

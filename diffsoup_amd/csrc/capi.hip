// extern "C" boundary of libdiffsoup_hip.so.
//
// Every entry point takes raw device pointers, host-side sizes and a
// hipStream_t, validates what can be validated on the host, enqueues its
// kernels on that stream and returns a status code.  Nothing here allocates
// device memory, synchronises the stream or exits the process: scratch space
// comes from a caller-owned workspace sized by the *_workspace_size queries,
// and counters that the reference returned through blocking copies are left
// in device memory.  The declarations and the reference interfaces they
// replace are listed in include/diffsoup_hip.h.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "diffsoup_hip.h"
#include "ds_kernels.h"
#include "ds_math.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_status(hipError_t e, const char* where) {
    if (e == hipSuccess) return DS_OK;
    return fail(DS_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

constexpr int kMaxLevel = 10;

// Build the per-launch view table; returns DS_OK or an error code.
int make_views(int B, const int* H, const int* W, int tile, ds::Views& v) {
    if (B < 1 || B > ds::kMaxViews)
        return fail(DS_ERR_INVALID, "view count %d outside [1, %d]", B, ds::kMaxViews);
    if (!H || !W) return fail(DS_ERR_INVALID, "null resolution array");
    if (tile != 8 && tile != 16 && tile != 32)
        return fail(DS_ERR_INVALID, "tile size %d not in {8, 16, 32}", tile);
    memset(&v, 0, sizeof(v));
    v.n = B;
    v.tile = tile;
    long long tiles = 0, pix = 0;
    for (int b = 0; b < B; ++b) {
        if (H[b] < 1 || W[b] < 1 || H[b] > 65535 || W[b] > 65535)
            return fail(DS_ERR_INVALID, "view %d resolution %dx%d outside [1, 65535]", b, H[b], W[b]);
        v.H[b] = H[b];
        v.W[b] = W[b];
        v.tiles_x[b] = (W[b] + tile - 1) / tile;
        v.tile_base[b] = static_cast<int>(tiles);
        v.pix_base[b] = static_cast<int>(pix);
        tiles += (long long)v.tiles_x[b] * ((H[b] + tile - 1) / tile);
        pix += (long long)H[b] * W[b];
        if (pix > INT32_MAX || tiles > INT32_MAX)
            return fail(DS_ERR_INVALID, "batch exceeds 2^31 pixels");
    }
    v.tile_base[B] = static_cast<int>(tiles);
    v.pix_base[B] = static_cast<int>(pix);
    v.interleave = 1;
    for (int b = 1; b < B; ++b)
        if (v.tile_base[b + 1] - v.tile_base[b] != v.tile_base[1] || v.tiles_x[b] != v.tiles_x[0]) v.interleave = 0;
    return DS_OK;
}

int check_levels(int lo, int hi) {
    if (lo < 0 || hi < lo || hi > kMaxLevel)
        return fail(DS_ERR_INVALID, "level range [%d, %d] invalid (0 <= lo <= hi <= %d)", lo, hi,
                    kMaxLevel);
    return DS_OK;
}

}  // namespace

namespace ds {
int step_make_views(const ds_step_desc* d, Views& v) { return make_views(d->B, d->H, d->W, d->tile, v); }
int step_fail(int code, const char* msg) { return fail(code, "%s", msg); }
int plain_fail(int code, const char* msg) { return fail(code, "%s", msg); }
int optim_fail(int code, const char* msg) { return fail(code, "%s", msg); }
int visibility_fail(int code, const char* msg) { return fail(code, "%s", msg); }
int remesh_fail(int code, const char* msg) { return fail(code, "%s", msg); }
int ssim_fail(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace ds

extern "C" {

const char* ds_version(void) { return "diffsoup-mi355x 0.1.0"; }

const char* ds_last_error(void) { return g_last_error.c_str(); }

uint32_t ds_feats_at_level(int level) {
    return level < 0 ? 0u : ds::feats_at_level(static_cast<uint32_t>(level));
}

int64_t ds_raster_tile_count(int B, const int* H, const int* W, int tile) {
    ds::Views v;
    if (make_views(B, H, W, tile, v) != DS_OK) return -1;
    return v.tile_base[B];
}

size_t ds_raster_workspace_size(int B, const int* H, const int* W, int T, int tile, int64_t capacity) {
    ds::Views v;
    if (make_views(B, H, W, tile, v) != DS_OK || T < 0 || capacity < 0) return 0;
    return ds::raster_workspace_bytes(B, T, v.tile_base[B], capacity);
}

int ds_raster_stats_offset(void) { return 0; }

static int raster_common(int B, const int* H, const int* W, int V, const float* pos, int T,
                         const int* tri, int level, const float* alpha_src, int tile,
                         void* workspace, size_t workspace_bytes, int64_t capacity, ds::Views& v,
                         ds::RasterWorkspace& w) {
    int rc = make_views(B, H, W, tile, v);
    if (rc != DS_OK) return rc;
    if (V < 0 || T < 0) return fail(DS_ERR_INVALID, "negative V (%d) or T (%d)", V, T);
    if (T > (1 << 24) - 1)
        return fail(DS_ERR_INVALID, "T = %d exceeds 2^24 - 1 (ids are stored as float)", T);
    if (capacity < 0 || capacity > INT32_MAX)
        return fail(DS_ERR_INVALID, "bin capacity %lld outside [0, 2^31)", (long long)capacity);
    if (level < 0 || level > kMaxLevel) return fail(DS_ERR_INVALID, "level %d invalid", level);
    if ((V > 0 && !pos) || (T > 0 && (!tri || !alpha_src)))
        return fail(DS_ERR_INVALID, "null input pointer");
    if (pos && !aligned16(pos)) return fail(DS_ERR_INVALID, "pos must be 16-byte aligned");
    if (!workspace) return fail(DS_ERR_INVALID, "null workspace");
    const size_t need = ds::raster_workspace_bytes(B, T, v.tile_base[B], capacity);
    if (workspace_bytes < need)
        return fail(DS_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    if (reinterpret_cast<uintptr_t>(workspace) & 255u)
        return fail(DS_ERR_INVALID, "workspace must be 256-byte aligned");
    w = ds::raster_workspace_carve(workspace, B, T, v.tile_base[B], capacity);
    return DS_OK;
}

int ds_rasterize_fwd(int B, const int* H, const int* W, int V, const float* pos, int T,
                     const int* tri, int level, const float* alpha_src, int stochastic,
                     uint64_t seed, int tile, float* rast, void* workspace, size_t workspace_bytes,
                     int64_t capacity, void* stream) {
    ds::Views v;
    ds::RasterWorkspace w;
    int rc = raster_common(B, H, W, V, pos, T, tri, level, alpha_src, tile, workspace,
                           workspace_bytes, capacity, v, w);
    if (rc != DS_OK) return rc;
    if (!rast || !aligned16(rast)) return fail(DS_ERR_INVALID, "rast must be non-null, 16-byte aligned");
    hipStream_t s = static_cast<hipStream_t>(stream);
    rc = hip_status(ds::launch_raster_bin(v, V, pos, T, tri, w, s), "raster bin");
    if (rc != DS_OK) return rc;
    ds::AlphaSource a{alpha_src, level, static_cast<int>(ds::feats_at_level(level))};
    return hip_status(ds::launch_raster_resolve(v, T, a, stochastic ? ds::kMaskHash : ds::kMaskDeterministic,
                                                ds::MaskSeed{seed, nullptr}, rast, w, s),
                      "raster resolve");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free rasterization: the workspace of ds_rasterize_fwd lives in a
// file-scope device array of DS_NOALLOC_SCRATCH_BYTES, one copy per GPU held
// until the process ends.  Host code never uses the array's name as an
// address: each device's address comes from hipGetSymbolAddress, taken once
// per device (with that device current) on the first call there.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[DS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_status(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(DS_ERR_INVALID, "device %d beyond %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return hip_status(e == hipSuccess ? hipErrorInvalidSymbol : e,
                                                      "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return DS_OK;
}

}  // namespace

extern "C" {

int ds_rasterize_fwd_noalloc(int B, const int* H, const int* W, int V, const float* pos, int T, const int* tri,
                             int level, const float* alpha_src, int stochastic, uint64_t seed, int tile, float* rast,
                             void* stream) {
    ds::Views v;
    int rc = make_views(B, H, W, tile, v);
    if (rc != DS_OK) return rc;
    if (T < 0) return fail(DS_ERR_INVALID, "negative T (%d)", T);
    // everything but the bins, then as many bin entries as the rest holds
    const size_t fixed = ds::raster_workspace_bytes(B, T, v.tile_base[B], 1) - 256;
    if (fixed + 256 > (size_t)DS_NOALLOC_SCRATCH_BYTES)
        return fail(DS_ERR_INVALID,
                    "ds_rasterize_fwd_noalloc: %d views x %d triangles need %zu bytes of scratch before any bin "
                    "entry; the fixed scratch holds %zu (use ds_rasterize_fwd with a workspace)",
                    B, T, fixed, (size_t)DS_NOALLOC_SCRATCH_BYTES);
    const int64_t capacity =
        std::min<int64_t>((int64_t)((DS_NOALLOC_SCRATCH_BYTES - fixed) / sizeof(int)), INT32_MAX);
    void* ptr = nullptr;
    rc = scratch_on_current_device(&ptr);
    if (rc != DS_OK) return rc;
    return ds_rasterize_fwd(B, H, W, V, pos, T, tri, level, alpha_src, stochastic, seed, tile, rast, ptr,
                            DS_NOALLOC_SCRATCH_BYTES, capacity, stream);
}

int ds_opacity_aux_bwd(int B, const int* H, const int* W, int V, const float* pos, int T,
                       const int* tri, int level, const float* alpha_src, int tile,
                       const float* rast, const float* pix_loss, float* grad_alpha, int reuse_bins,
                       void* workspace, size_t workspace_bytes, int64_t capacity, void* stream) {
    ds::Views v;
    ds::RasterWorkspace w;
    int rc = raster_common(B, H, W, V, pos, T, tri, level, alpha_src, tile, workspace,
                           workspace_bytes, capacity, v, w);
    if (rc != DS_OK) return rc;
    if (!rast || !aligned16(rast) || !pix_loss || (T > 0 && !grad_alpha))
        return fail(DS_ERR_INVALID, "null or misaligned rast / pix_loss / grad_alpha");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!reuse_bins) {
        rc = hip_status(ds::launch_raster_bin(v, V, pos, T, tri, w, s), "aux bin");
        if (rc != DS_OK) return rc;
    }
    ds::AlphaSource a{alpha_src, level, static_cast<int>(ds::feats_at_level(level))};
    return hip_status(ds::launch_opacity_aux(v, T, a, rast, pix_loss, grad_alpha, w, s), "opacity aux");
}

int ds_pixel_l1(int64_t P, int C, const float* color, const float* target, float* pix_loss,
                void* stream) {
    if (P < 0 || C < 1) return fail(DS_ERR_INVALID, "bad P (%lld) or C (%d)", (long long)P, C);
    if (P > 0 && (!color || !target || !pix_loss)) return fail(DS_ERR_INVALID, "null pointer");
    return hip_status(ds::launch_pixel_l1(P, C, color, target, pix_loss, static_cast<hipStream_t>(stream)),
                      "pixel l1");
}

int ds_multires_color_fwd(int64_t P, const float* rast, int lo, int hi, int T, int C, const float* feat,
                          int concat, float* out, void* stream) {
    int rc = check_levels(lo, hi);
    if (rc != DS_OK) return rc;
    if (P < 0 || C < 1 || T < 0) return fail(DS_ERR_INVALID, "bad P (%lld), T (%d) or C (%d)", (long long)P, T, C);
    if (P > 0 && (!rast || !aligned16(rast) || (T > 0 && !feat) || !out))
        return fail(DS_ERR_INVALID, "null or misaligned pointer");
    return hip_status(ds::launch_multires_color_fwd(P, rast, lo, hi, T, C, feat, concat, out,
                                                    static_cast<hipStream_t>(stream)),
                      "multires color fwd");
}

int ds_multires_color_bwd(int64_t P, const float* rast, int lo, int hi, int T, int C, float* grad_feat,
                          int concat, const float* grad_out, void* stream) {
    int rc = check_levels(lo, hi);
    if (rc != DS_OK) return rc;
    if (P < 0 || C < 1 || T < 0) return fail(DS_ERR_INVALID, "bad P (%lld), T (%d) or C (%d)", (long long)P, T, C);
    if (P > 0 && (!rast || !aligned16(rast) || (T > 0 && !grad_feat) || !grad_out))
        return fail(DS_ERR_INVALID, "null or misaligned pointer");
    return hip_status(ds::launch_multires_color_bwd(P, rast, lo, hi, T, C, grad_feat, concat, grad_out,
                                                    static_cast<hipStream_t>(stream)),
                      "multires color bwd");
}

int ds_accumulate_fwd(int T, int lo, int hi, int target, int C, const float* feat, float* out,
                      void* stream) {
    int rc = check_levels(lo, hi);
    if (rc != DS_OK) return rc;
    if (target < 0 || target > kMaxLevel) return fail(DS_ERR_INVALID, "target level %d invalid", target);
    if (T < 0 || C < 1) return fail(DS_ERR_INVALID, "bad T (%d) or C (%d)", T, C);
    if (T > 0 && (!feat || !out)) return fail(DS_ERR_INVALID, "null pointer");
    return hip_status(ds::launch_accumulate_fwd(T, lo, hi, target, C, feat, out,
                                                static_cast<hipStream_t>(stream)),
                      "accumulate fwd");
}

int ds_accumulate_bwd(int T, int lo, int hi, int target, int C, float* grad_feat,
                      const float* grad_out, void* stream) {
    int rc = check_levels(lo, hi);
    if (rc != DS_OK) return rc;
    if (target < 0 || target > kMaxLevel) return fail(DS_ERR_INVALID, "target level %d invalid", target);
    if (T < 0 || C < 1) return fail(DS_ERR_INVALID, "bad T (%d) or C (%d)", T, C);
    if (T > 0 && (!grad_feat || !grad_out)) return fail(DS_ERR_INVALID, "null pointer");
    return hip_status(ds::launch_accumulate_bwd(T, lo, hi, target, C, grad_feat, grad_out,
                                                static_cast<hipStream_t>(stream)),
                      "accumulate bwd");
}

int ds_sh2(int B, const int* H, const int* W, const float* rast, const float* inv_mvp, float* enc,
           void* stream) {
    ds::Views v;
    int rc = make_views(B, H, W, 16, v);
    if (rc != DS_OK) return rc;
    if (!rast || !aligned16(rast) || !inv_mvp || !enc) return fail(DS_ERR_INVALID, "null or misaligned pointer");
    return hip_status(ds::launch_sh2(v, rast, inv_mvp, enc, static_cast<hipStream_t>(stream)), "sh2");
}

int ds_edge_grad_bwd(int B, const int* H, const int* W, int C, const float* color,
                     const float* grad_color, const float* rast, int V, const float* pos,
                     float* grad_pos, int T, const int* tri, void* stream) {
    ds::Views v;
    int rc = make_views(B, H, W, 16, v);
    if (rc != DS_OK) return rc;
    if (C < 1 || V < 0 || T < 0) return fail(DS_ERR_INVALID, "bad C (%d), V (%d) or T (%d)", C, V, T);
    if (!color || !grad_color || !rast || !aligned16(rast) || !pos || !aligned16(pos) || !grad_pos || !tri)
        return fail(DS_ERR_INVALID, "null or misaligned pointer");
    return hip_status(ds::launch_edge_grad(v, C, color, grad_color, rast, V, pos, grad_pos, T, tri,
                                           static_cast<hipStream_t>(stream)),
                      "edge grad");
}

int ds_fill_f32(float* p, int64_t n, float value, void* stream) {
    if (n < 0 || (n > 0 && !p)) return fail(DS_ERR_INVALID, "bad fill arguments");
    return hip_status(ds::launch_fill_f32(p, n, value, static_cast<hipStream_t>(stream)), "fill");
}

}  // extern "C"

// Structural similarity and the reference's photometric loss on the GPU.
//
// The reference trains with 0.8 L1 + 0.2 (1 - SSIM) / 2 (examples/
// 02_synthetic.py:376-380), SSIM from pytorch_msssim.ssim (win_size 11,
// win_sigma 1.5, K = (0.01, 0.03), data_range 1, "valid" filtering: a
// separable Gaussian over every full window; a side shorter than the window
// is not filtered).  Per window q of channel c:
//
//   mx, my, exx, eyy, exy  = Gaussian means of x, y, x^2, y^2, xy
//   sxx = exx - mx^2, syy = eyy - my^2, sxy = exy - mx my
//   S = (2 mx my + C1) / (mx^2 + my^2 + C1) * (2 sxy + C2) / (sxx + syy + C2)
//
// and SSIM is the mean of S over windows, channels and images.  Its
// derivative with respect to the second image is a transposed filter of three
// per-window maps: dSSIM/dy(p) = sum_q G(q, p) [a(q) + 2 y(p) b(q) + x(p) c(q)]
// with a = dS/dmy, b = dS/deyy, c = dS/dexy (each holding the others fixed).
//
//   k_ssim_maps  one workgroup per 64x16 tile of windows: x and y with a
//                10-pixel halo staged in LDS one channel at a time, the
//                horizontal pass of the five sums into LDS, the vertical pass
//                in registers, then S (a per-tile partial sum) and the maps
//                a, b, c written at the window's top-left pixel.
//   k_ssim_grad  one workgroup per 64x16 tile of pixels: the three maps with
//                their halo in LDS (zero outside the windows), the same two
//                passes (the window is symmetric, so the transposed filter is
//                the filter), then scale (a + 2 y b + x c) plus the L1 term
//                l1 sign(y - x), stored or added to the gradient.
//
// Images are packed like every per-pixel buffer of the library: [P, C] with
// view b's pixels at [pix_base[b], pix_base[b+1]), views of any resolution.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "diffsoup_hip.h"
#include "ds_kernels.h"

namespace ds {

int ssim_fail(int code, const char* msg);

namespace {

constexpr int kTW = 64, kTH = 16, kHalo = 10, kThreads = 256;
constexpr int kLW = kTW + kHalo, kLH = kTH + kHalo;
constexpr int kMaxTaps = 11, kMaxC = 4;

struct SsimTiles {
    int n, C;
    int H[kMaxViews], W[kMaxViews];
    int kh[kMaxViews], kw[kMaxViews];   // taps along y / x: the window, or 1 on a short side
    int tiles_x[kMaxViews];
    int tile_base[kMaxViews + 1];
    long long pix_base[kMaxViews + 1];
    float win[kMaxTaps];
    float scale[kMaxViews * kMaxC];     // per (view, channel) gradient scale
};

__device__ __forceinline__ int view_of_tile(const SsimTiles& v, int tile) {
    int lo = 0, hi = v.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (v.tile_base[mid] <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ float wave_sum(float x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__global__ __launch_bounds__(kThreads) void k_ssim_maps(SsimTiles v, const float* __restrict__ X,
                                                        const float* __restrict__ Y, float c1, float c2,
                                                        float* __restrict__ ma, float* __restrict__ mb,
                                                        float* __restrict__ mc, float* __restrict__ partial) {
    __shared__ float sx[kLH][kLW], sy[kLH][kLW];
    __shared__ float hs[5][kLH][kTW];
    __shared__ float red[kThreads / 64];
    const int tile = blockIdx.x;
    const int b = view_of_tile(v, tile);
    const int local = tile - v.tile_base[b];
    const int ty = local / v.tiles_x[b], tx = local - ty * v.tiles_x[b];
    const int H = v.H[b], W = v.W[b], kh = v.kh[b], kw = v.kw[b], C = v.C;
    const int Hv = H - kh + 1, Wv = W - kw + 1;
    const int y0 = ty * kTH, x0 = tx * kTW;
    const long long base = v.pix_base[b];
    const int tid = threadIdx.x;

    for (int c = 0; c < C; ++c) {
        for (int i = tid; i < kLH * kLW; i += kThreads) {
            const int r = i / kLW, q = i - r * kLW;
            const int gy = y0 + r, gx = x0 + q;
            float xv = 0.f, yv = 0.f;
            if (gy < H && gx < W) {
                const long long o = (base + (long long)gy * W + gx) * C + c;
                xv = X[o];
                yv = Y[o];
            }
            sx[r][q] = xv;
            sy[r][q] = yv;
        }
        __syncthreads();
        for (int i = tid; i < (kTH + kh - 1) * kTW; i += kThreads) {
            const int r = i / kTW, j = i - r * kTW;
            float ax = 0.f, ay = 0.f, axx = 0.f, ayy = 0.f, axy = 0.f;
            for (int k = 0; k < kw; ++k) {
                const float wk = kw == 1 ? 1.f : v.win[k];
                const float xv = sx[r][j + k], yv = sy[r][j + k];
                ax += wk * xv;
                ay += wk * yv;
                axx += wk * (xv * xv);
                ayy += wk * (yv * yv);
                axy += wk * (xv * yv);
            }
            hs[0][r][j] = ax;
            hs[1][r][j] = ay;
            hs[2][r][j] = axx;
            hs[3][r][j] = ayy;
            hs[4][r][j] = axy;
        }
        __syncthreads();
        float acc = 0.f;
        for (int i = tid; i < kTH * kTW; i += kThreads) {
            const int r = i / kTW, j = i - r * kTW;
            const int oy = y0 + r, ox = x0 + j;
            if (oy >= Hv || ox >= Wv) continue;
            float mx = 0.f, my = 0.f, exx = 0.f, eyy = 0.f, exy = 0.f;
            for (int k = 0; k < kh; ++k) {
                const float wk = kh == 1 ? 1.f : v.win[k];
                mx += wk * hs[0][r + k][j];
                my += wk * hs[1][r + k][j];
                exx += wk * hs[2][r + k][j];
                eyy += wk * hs[3][r + k][j];
                exy += wk * hs[4][r + k][j];
            }
            const float sxx = exx - mx * mx, syy = eyy - my * my, sxy = exy - mx * my;
            const float a1 = 2.f * mx * my + c1, b1 = mx * mx + my * my + c1;
            const float a2 = 2.f * sxy + c2, b2 = sxx + syy + c2;
            const float lum = a1 / b1, cs = a2 / b2;
            acc += lum * cs;
            const long long o = (base + (long long)oy * W + ox) * C + c;
            // dS/dmy = dL/dmy cs + L dcs/dmy; dS/deyy = -L cs / b2; dS/dexy = 2 L / b2
            ma[o] = 2.f * (mx - lum * my) / b1 * cs + lum * 2.f * (my * cs - mx) / b2;
            mb[o] = -lum * cs / b2;
            mc[o] = 2.f * lum / b2;
        }
        acc = wave_sum(acc);
        if ((tid & 63) == 0) red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            float s = 0.f;
            for (int w = 0; w < kThreads / 64; ++w) s += red[w];
            partial[(long long)tile * C + c] = s;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_ssim_grad(SsimTiles v, const float* __restrict__ X,
                                                        const float* __restrict__ Y, const float* __restrict__ ma,
                                                        const float* __restrict__ mb, const float* __restrict__ mc,
                                                        const float* __restrict__ dscale, float l1,
                                                        const float* __restrict__ dl1, float* __restrict__ grad,
                                                        int accumulate, float* __restrict__ l1_partial) {
    __shared__ float sm[3][kLH][kLW];
    __shared__ float hs[3][kLH][kTW];
    __shared__ float red[kThreads / 64];
    float l1_acc = 0.f;   // sum of |y - x| over this tile's pixels and channels
    const int tile = blockIdx.x;
    const int b = view_of_tile(v, tile);
    const int local = tile - v.tile_base[b];
    const int ty = local / v.tiles_x[b], tx = local - ty * v.tiles_x[b];
    const int H = v.H[b], W = v.W[b], kh = v.kh[b], kw = v.kw[b], C = v.C;
    const int Hv = H - kh + 1, Wv = W - kw + 1;
    const int hy = kh - 1, hx = kw - 1;
    const int y0 = ty * kTH, x0 = tx * kTW;
    const long long base = v.pix_base[b];
    const int tid = threadIdx.x;
    const float l1s = dl1 ? l1 * dl1[0] : l1;

    for (int c = 0; c < C; ++c) {
        // windows q with q <= p < q + taps: rows y0 - hy .. y0 + kTH - 1
        for (int i = tid; i < kLH * kLW; i += kThreads) {
            const int r = i / kLW, q = i - r * kLW;
            const int wy = y0 - hy + r, wx = x0 - hx + q;
            float a = 0.f, bb = 0.f, cc = 0.f;
            if (r < kTH + hy && q < kTW + hx && wy >= 0 && wy < Hv && wx >= 0 && wx < Wv) {
                const long long o = (base + (long long)wy * W + wx) * C + c;
                a = ma[o];
                bb = mb[o];
                cc = mc[o];
            }
            sm[0][r][q] = a;
            sm[1][r][q] = bb;
            sm[2][r][q] = cc;
        }
        __syncthreads();
        for (int i = tid; i < (kTH + hy) * kTW; i += kThreads) {
            const int r = i / kTW, j = i - r * kTW;
            float a = 0.f, bb = 0.f, cc = 0.f;
            for (int k = 0; k < kw; ++k) {
                const float wk = kw == 1 ? 1.f : v.win[k];
                a += wk * sm[0][r][j + k];
                bb += wk * sm[1][r][j + k];
                cc += wk * sm[2][r][j + k];
            }
            hs[0][r][j] = a;
            hs[1][r][j] = bb;
            hs[2][r][j] = cc;
        }
        __syncthreads();
        const float sc = dscale ? v.scale[b * C + c] * dscale[b * C + c] : v.scale[b * C + c];
        for (int i = tid; i < kTH * kTW; i += kThreads) {
            const int r = i / kTW, j = i - r * kTW;
            const int py = y0 + r, px = x0 + j;
            if (py >= H || px >= W) continue;
            float a = 0.f, bb = 0.f, cc = 0.f;
            for (int k = 0; k < kh; ++k) {
                const float wk = kh == 1 ? 1.f : v.win[k];
                a += wk * hs[0][r + k][j];
                bb += wk * hs[1][r + k][j];
                cc += wk * hs[2][r + k][j];
            }
            const long long o = (base + (long long)py * W + px) * C + c;
            const float xv = X[o], yv = Y[o];
            const float d = yv - xv;
            const float sg = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
            const float gval = sc * (a + 2.f * yv * bb + xv * cc) + l1s * sg;
            grad[o] = accumulate ? grad[o] + gval : gval;
            l1_acc += fabsf(d);
        }
        __syncthreads();
    }
    if (l1_partial) {
        l1_acc = wave_sum(l1_acc);
        if ((tid & 63) == 0) red[tid >> 6] = l1_acc;
        __syncthreads();
        if (tid == 0) {
            float t = 0.f;
            for (int w = 0; w < kThreads / 64; ++w) t += red[w];
            l1_partial[tile] = t;
        }
    }
}

// The training step's photometric loss from the partial sums, one workgroup:
// w_l1 * sum|y - x| / (C P) + w_ssim * (1 - SSIM) / 2, SSIM being the mean
// over views of each view's mean over channels and windows.
__global__ __launch_bounds__(1024) void k_photo_loss(SsimTiles v, const float* __restrict__ ssim_partial,
                                                     const float* __restrict__ l1_partial, int n_pix_tiles,
                                                     float w_l1, float w_ssim, float* __restrict__ loss) {
    __shared__ double rs[1024 / 64], rl[1024 / 64];
    const int C = v.C;
    double s = 0.0, l = 0.0;
    for (int i = threadIdx.x; i < v.tile_base[v.n] * C; i += blockDim.x) {
        const int b = view_of_tile(v, i / C);
        const int kh = v.kh[b], kw = v.kw[b];
        const double windows = (double)(v.H[b] - kh + 1) * (v.W[b] - kw + 1);
        s += ssim_partial[i] / windows;
    }
    for (int i = threadIdx.x; i < n_pix_tiles; i += blockDim.x) l += l1_partial[i];
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        l += __shfl_xor(l, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        rs[threadIdx.x >> 6] = s;
        rl[threadIdx.x >> 6] = l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double st = 0.0, lt = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            st += rs[w];
            lt += rl[w];
        }
        const double ssim = st / ((double)v.n * C);
        loss[0] = static_cast<float>(w_l1 * lt / ((double)C * v.pix_base[v.n]) + w_ssim * 0.5 * (1.0 - ssim));
    }
}

// Fill the tile table for windows (valid = true) or pixels; returns tiles or -1.
long long make_tiles(int B, const int* H, const int* W, int C, int taps, const float* win, bool valid,
                     SsimTiles& t) {
    if (B < 1 || B > kMaxViews || !H || !W || C < 1 || C > kMaxC || taps < 1 || taps > kMaxTaps || !(taps & 1))
        return -1;
    std::fill(reinterpret_cast<char*>(&t), reinterpret_cast<char*>(&t) + sizeof(t), 0);
    t.n = B;
    t.C = C;
    for (int k = 0; k < taps; ++k) t.win[k] = win ? win[k] : 0.f;
    long long tiles = 0, pix = 0;
    for (int b = 0; b < B; ++b) {
        if (H[b] < 1 || W[b] < 1 || H[b] > 65535 || W[b] > 65535) return -1;
        t.H[b] = H[b];
        t.W[b] = W[b];
        t.kh[b] = H[b] >= taps ? taps : 1;
        t.kw[b] = W[b] >= taps ? taps : 1;
        const int rows = valid ? H[b] - t.kh[b] + 1 : H[b];
        const int cols = valid ? W[b] - t.kw[b] + 1 : W[b];
        t.tiles_x[b] = (cols + kTW - 1) / kTW;
        t.tile_base[b] = static_cast<int>(tiles);
        t.pix_base[b] = pix;
        tiles += (long long)t.tiles_x[b] * ((rows + kTH - 1) / kTH);
        pix += (long long)H[b] * W[b];
        if (tiles > (1LL << 30)) return -1;
    }
    t.tile_base[B] = static_cast<int>(tiles);
    t.pix_base[B] = pix;
    return tiles;
}

}  // namespace

hipError_t launch_ssim_maps(int B, const int* H, const int* W, int C, const float* X, const float* Y, int taps,
                            const float* win, float c1, float c2, float* maps, float* partial, hipStream_t s) {
    SsimTiles t;
    const long long tiles = make_tiles(B, H, W, C, taps, win, true, t);
    if (tiles < 0) return hipErrorInvalidValue;
    if (tiles == 0) return hipSuccess;
    const long long n = t.pix_base[B] * C;
    k_ssim_maps<<<static_cast<unsigned>(tiles), kThreads, 0, s>>>(t, X, Y, c1, c2, maps, maps + n, maps + 2 * n,
                                                                    partial);
    return hipGetLastError();
}

hipError_t launch_ssim_grad(int B, const int* H, const int* W, int C, const float* X, const float* Y, int taps,
                            const float* win, const float* maps, const float* scale, const float* dscale, float l1,
                            const float* dl1, float* grad, int accumulate, hipStream_t s) {
    SsimTiles t;
    const long long tiles = make_tiles(B, H, W, C, taps, win, false, t);
    if (tiles < 0) return hipErrorInvalidValue;
    for (int i = 0; i < B * C; ++i) t.scale[i] = scale[i];
    if (tiles == 0) return hipSuccess;
    const long long n = t.pix_base[B] * C;
    k_ssim_grad<<<static_cast<unsigned>(tiles), kThreads, 0, s>>>(t, X, Y, maps, maps + n, maps + 2 * n, dscale, l1,
                                                                    dl1, grad, accumulate, nullptr);
    return hipGetLastError();
}

// ---- the training step's L1 + SSIM loss (views of any resolution)

size_t photo_loss_floats(const Views& views, int C) {
    SsimTiles w, p;
    const float ones[kMaxTaps] = {};
    const long long tw = make_tiles(views.n, views.H, views.W, C, kMaxTaps, ones, true, w);
    const long long tp = make_tiles(views.n, views.H, views.W, C, kMaxTaps, ones, false, p);
    if (tw < 0 || tp < 0) return 0;
    // maps [3, P, C], window partials [tw, C], pixel-tile L1 partials [tp]
    return 3 * (size_t)w.pix_base[views.n] * C + (size_t)tw * C + (size_t)tp;
}

hipError_t launch_photo_loss(const Views& views, int C, const float* target, const float* color, float* scratch,
                             float w_l1, float w_ssim, float* grad, float* loss, hipStream_t s) {
    // the 11-tap Gaussian of sigma 1.5 in float32, normalised as pytorch_msssim does
    float win[kMaxTaps];
    float sum = 0.f;
    for (int k = 0; k < kMaxTaps; ++k) {
        const float c = static_cast<float>(k - kMaxTaps / 2);
        win[k] = expf(-(c * c) / (2.f * 1.5f * 1.5f));
        sum += win[k];
    }
    for (int k = 0; k < kMaxTaps; ++k) win[k] /= sum;
    SsimTiles w, p;
    const long long tw = make_tiles(views.n, views.H, views.W, C, kMaxTaps, win, true, w);
    const long long tp = make_tiles(views.n, views.H, views.W, C, kMaxTaps, win, false, p);
    if (tw < 1 || tp < 1) return hipErrorInvalidValue;
    const long long n = w.pix_base[views.n] * C;
    float* maps = scratch;
    float* spart = maps + 3 * n;
    float* lpart = spart + tw * C;
    const float c1 = 0.01f * 0.01f, c2 = 0.03f * 0.03f;
    k_ssim_maps<<<static_cast<unsigned>(tw), kThreads, 0, s>>>(w, target, color, c1, c2, maps, maps + n,
                                                                 maps + 2 * n, spart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int b = 0; b < views.n; ++b) {
        const double windows = (double)(views.H[b] - p.kh[b] + 1) * (views.W[b] - p.kw[b] + 1);
        for (int c = 0; c < C; ++c)
            p.scale[b * C + c] = static_cast<float>(-0.5 * w_ssim / ((double)views.n * C * windows));
    }
    const float l1s = static_cast<float>(w_l1 / ((double)C * w.pix_base[views.n]));
    k_ssim_grad<<<static_cast<unsigned>(tp), kThreads, 0, s>>>(p, target, color, maps, maps + n, maps + 2 * n,
                                                                 nullptr, l1s, nullptr, grad, 0, lpart);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_photo_loss<<<1, 1024, 0, s>>>(w, spart, lpart, static_cast<int>(tp), w_l1, w_ssim, loss);
    return hipGetLastError();
}

long long ssim_window_tiles(int B, const int* H, const int* W, int taps, long long* tile_base) {
    SsimTiles t;
    const long long tiles = make_tiles(B, H, W, 1, taps, nullptr, true, t);
    if (tiles >= 0 && tile_base)
        for (int b = 0; b <= B; ++b) tile_base[b] = t.tile_base[b];
    return tiles;
}

}  // namespace ds

extern "C" {

int64_t ds_ssim_tiles(int B, const int* H, const int* W, int win_size, int64_t* tile_base) {
    const long long n = ds::ssim_window_tiles(B, H, W, win_size, reinterpret_cast<long long*>(tile_base));
    if (n < 0) ds::ssim_fail(DS_ERR_INVALID, "ssim_tiles: bad views or window size");
    return n;
}

int ds_ssim_fwd(int B, const int* H, const int* W, int C, const float* X, const float* Y, int win_size,
                const float* win, float c1, float c2, float* maps, float* partial, void* stream) {
    if (!X || !Y || !maps || !partial || !win)
        return ds::ssim_fail(DS_ERR_INVALID, "ssim_fwd: null pointer");
    const hipError_t e = ds::launch_ssim_maps(B, H, W, C, X, Y, win_size, win, c1, c2, maps, partial,
                                              static_cast<hipStream_t>(stream));
    if (e == hipErrorInvalidValue)
        return ds::ssim_fail(DS_ERR_INVALID, "ssim_fwd: bad views, channel count (1-4) or window size (odd, <= 11)");
    if (e != hipSuccess) return ds::ssim_fail(DS_ERR_HIP, hipGetErrorString(e));
    return DS_OK;
}

int ds_ssim_bwd(int B, const int* H, const int* W, int C, const float* X, const float* Y, int win_size,
                const float* win, const float* maps, const float* scale, const float* dscale, float l1_scale,
                const float* dl1, float* grad, int accumulate, void* stream) {
    if (!X || !Y || !maps || !scale || !grad || !win)
        return ds::ssim_fail(DS_ERR_INVALID, "ssim_bwd: null pointer");
    const hipError_t e = ds::launch_ssim_grad(B, H, W, C, X, Y, win_size, win, maps, scale, dscale, l1_scale, dl1,
                                              grad, accumulate, static_cast<hipStream_t>(stream));
    if (e == hipErrorInvalidValue)
        return ds::ssim_fail(DS_ERR_INVALID, "ssim_bwd: bad views, channel count (1-4) or window size (odd, <= 11)");
    if (e != hipSuccess) return ds::ssim_fail(DS_ERR_HIP, hipGetErrorString(e));
    return DS_OK;
}

}  // extern "C"

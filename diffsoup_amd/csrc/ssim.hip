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
//   k_ssim_maps  one workgroup per 60x16 tile of windows: x and y of all
//                channels with a 10-pixel halo staged in LDS in one pass over
//                contiguous row segments, then per channel the vertical pass
//                of the five sums over the 16 window rows and every staged
//                column into LDS, the horizontal pass in registers, S (a
//                per-tile partial sum) and the maps a, b, c written
//                channel-planar at the window's top-left pixel (vertical
//                first: the first pass covers 16 rows of 70 columns instead
//                of 26 rows of 60).
//   k_ssim_grad  one workgroup per 60x16 tile of pixels: per channel the
//                three maps with their halo in LDS (zero outside the windows)
//                and the same two passes, vertical first (the window is
//                symmetric, so the transposed filter is the filter); scale
//                (a + 2 y b + x c) plus the L1 term l1 sign(y - x) collects
//                in LDS for all channels and is stored (or added) in one
//                coalesced pass.
//
// The images interleave their channels, so every global access walks whole
// row segments of all channels at once: a per-channel stride would fetch
// and write back each cache line once per channel.  In LDS the filter passes
// are register-blocked: a horizontal pass forms four adjacent outputs from
// fourteen values read as four float4s, a vertical pass two rows from
// twelve, so each staged value is read about a third as often as with one
// output per thread.  The 60-pixel tile width keeps two 8-wave workgroups per
// CU within the LDS once the staged rows are padded to a 16-byte stride.
//
// Images are packed like every per-pixel buffer of the library: [P, C] with
// view b's pixels at [pix_base[b], pix_base[b+1]), views of any resolution.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "diffsoup_hip.h"
#include "ds_kernels.h"

// The library builds with -ffp-contract=off so that the raster arithmetic
// rounds like the CPU restatement.  Nothing here decides a branch or has to
// match bit for bit, so the filter taps and the SSIM terms contract into FMAs.
#pragma clang fp contract(fast)

namespace ds {

int ssim_fail(int code, const char* msg);

namespace {

constexpr int kTW = 60, kTH = 16, kHalo = 10, kThreads = 512;
constexpr int kLW = kTW + kHalo, kLH = kTH + kHalo;
constexpr int kLWp = (kLW + 3) & ~3;   // staged row stride: rows start 16-byte aligned
constexpr int kQ = 4;                  // horizontal outputs per thread (one float4 of results)
constexpr int kR = 2;                  // vertical outputs per thread
static_assert(kTW % kQ == 0 && kTH % kR == 0, "tile edges must be multiples of the blocking");
constexpr int kLanesPerRow = 16;       // lanes given to one staged row in the horizontal passes
static_assert(kTW / kQ <= kLanesPerRow, "a row's float4 groups must fit its lanes");
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

// LDS floats of the two kernels for C channels: k_ssim_maps stages x and y
// (all channels, with the halo) and the five horizontal sums of one channel;
// k_ssim_grad stages the three maps of one channel with their halo, their
// horizontal sums, the tile's x and y (all channels) and its gradient.
template <int C>
constexpr int maps_lds_floats() { return 2 * C * kLH * kLWp + 5 * kTH * kLWp; }
template <int C>
constexpr int grad_lds_floats() { return 3 * kLH * kLWp + 3 * kTH * kLWp + 3 * C * kTH * kTW; }

// The separable passes, one per axis and tap count (the window's 11 taps,
// or 1 on a side shorter than the window), with every tap loop unrolled so
// that a thread's LDS reads issue back to back.

// Vertical sums of x, y, x^2, y^2, xy for the tile's kTH window rows over
// every staged column (the horizontal pass then needs no halo rows); a
// thread forms kR rows of one column from kR + KH - 1 staged values, the
// products once per staged value.
template <int KH>
__device__ __forceinline__ void moments_v(const float* sx, const float* sy, float* vs, const float (&g)[kMaxTaps],
                                          int tid) {
    constexpr int NV = kR + KH - 1;
    for (int i = tid; i < (kTH / kR) * kLW; i += kThreads) {
        const int rg = i / kLW, j = i - rg * kLW;
        const int r0 = rg * kR;
        float xv[NV], yv[NV], xx[NV], yy[NV], xy[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            xv[k] = sx[(r0 + k) * kLWp + j];
            yv[k] = sy[(r0 + k) * kLWp + j];
            xx[k] = xv[k] * xv[k];
            yy[k] = yv[k] * yv[k];
            xy[k] = xv[k] * yv[k];
        }
#pragma unroll
        for (int rr = 0; rr < kR; ++rr) {
            float m[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < KH; ++k) {
                const float wk = KH == 1 ? 1.f : g[k];
                m[0] += wk * xv[rr + k];
                m[1] += wk * yv[rr + k];
                m[2] += wk * xx[rr + k];
                m[3] += wk * yy[rr + k];
                m[4] += wk * xy[rr + k];
            }
#pragma unroll
            for (int q = 0; q < 5; ++q) vs[(q * kTH + r0 + rr) * kLWp + j] = m[q];
        }
    }
}

// Horizontal sums of the five vertical sums, S and the three derivative maps
// of channel c for the tile's windows, kQ adjacent windows per thread from
// kQ + KW - 1 values of each sum read as float4s; returns this thread's sum
// of S.
template <int C, int KW>
__device__ __forceinline__ float ssim_h(const float* vs, const float (&g)[kMaxTaps], int c, float c1, float c2,
                                        int Hv, int Wv, int y0, int x0, int W, long long base,
                                        float* __restrict__ maps, long long P, int tid) {
    constexpr int NV = kQ + KW - 1, N4 = (NV + 3) / 4;
    float acc = 0.f;
    for (int i = tid; i < kTH * kLanesPerRow; i += kThreads) {   // one row per 16 lanes, as above
        const int r = i / kLanesPerRow, quad = i - r * kLanesPerRow;
        if (quad >= kTW / kQ) continue;
        const int j0 = quad * kQ, oy = y0 + r;
        if (oy >= Hv || x0 + j0 >= Wv) continue;
        float v[5][4 * N4];
#pragma unroll
        for (int q = 0; q < 5; ++q)
#pragma unroll
            for (int n = 0; n < N4; ++n) {
                const float4 a = *reinterpret_cast<const float4*>(vs + (q * kTH + r) * kLWp + j0 + 4 * n);
                v[q][4 * n] = a.x; v[q][4 * n + 1] = a.y; v[q][4 * n + 2] = a.z; v[q][4 * n + 3] = a.w;
            }
#pragma unroll
        for (int o = 0; o < kQ; ++o) {
            const int ox = x0 + j0 + o;
            if (ox >= Wv) break;
            float m[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                m[q] = 0.f;
#pragma unroll
                for (int k = 0; k < KW; ++k) m[q] += (KW == 1 ? 1.f : g[k]) * v[q][o + k];
            }
            const float mx = m[0], my = m[1], exx = m[2], eyy = m[3], exy = m[4];
            const float sxx = exx - mx * mx, syy = eyy - my * my, sxy = exy - mx * my;
            const float a1 = 2.f * mx * my + c1, b1 = mx * mx + my * my + c1;
            const float a2 = 2.f * sxy + c2, b2 = sxx + syy + c2;
            // both denominators are >= C1 or C2 > 0: one hardware reciprocal each
            // (1 ulp) serves every quotient below
            const float rb1 = __builtin_amdgcn_rcpf(b1), rb2 = __builtin_amdgcn_rcpf(b2);
            const float lum = a1 * rb1, cs = a2 * rb2;
            acc += lum * cs;
            const long long o_ = base + (long long)oy * W + ox;
            // dS/dmy = dL/dmy cs + L dcs/dmy; dS/deyy = -L cs / b2; dS/dexy = 2 L / b2
            maps[(0 * C + c) * P + o_] = 2.f * ((mx - lum * my) * rb1 * cs + lum * (my * cs - mx) * rb2);
            maps[(1 * C + c) * P + o_] = -lum * cs * rb2;
            maps[(2 * C + c) * P + o_] = 2.f * lum * rb2;
        }
    }
    return acc;
}

// Vertical sums of the three maps of one channel for the tile's kTH pixel
// rows over every staged column, kR rows of one column per thread.
template <int KH>
__device__ __forceinline__ void maps_v(const float* sm, float* vs, const float (&g)[kMaxTaps], int tid) {
    constexpr int NV = kR + KH - 1;
    for (int i = tid; i < (kTH / kR) * kLW; i += kThreads) {
        const int rg = i / kLW, j = i - rg * kLW;
        const int r0 = rg * kR;
#pragma unroll
        for (int mm = 0; mm < 3; ++mm) {
            float h[NV];
#pragma unroll
            for (int k = 0; k < NV; ++k) h[k] = sm[(mm * kLH + r0 + k) * kLWp + j];
#pragma unroll
            for (int rr = 0; rr < kR; ++rr) {
                float m = 0.f;
#pragma unroll
                for (int k = 0; k < KH; ++k) m += (KH == 1 ? 1.f : g[k]) * h[rr + k];
                vs[(mm * kTH + r0 + rr) * kLWp + j] = m;
            }
        }
    }
}

// Horizontal sums of the three vertical sums and the gradient of channel c
// for the tile's pixels, kQ adjacent pixels per thread from kQ + KW - 1
// values of each sum read as float4s, into tg.
template <int C, int KW>
__device__ __forceinline__ void grad_h(const float* vs, const float (&g)[kMaxTaps], int c, float sc, float l1s,
                                       const float* tx_, const float* ty_, float* tg, int tid) {
    constexpr int rowf = kTW * C, NV = kQ + KW - 1, N4 = (NV + 3) / 4;
    for (int i = tid; i < kTH * kLanesPerRow; i += kThreads) {   // one row per 16 lanes, as above
        const int r = i / kLanesPerRow, quad = i - r * kLanesPerRow;
        if (quad >= kTW / kQ) continue;
        const int j0 = quad * kQ;
        float v[3][4 * N4];
#pragma unroll
        for (int mm = 0; mm < 3; ++mm)
#pragma unroll
            for (int n = 0; n < N4; ++n) {
                const float4 a = *reinterpret_cast<const float4*>(vs + (mm * kTH + r) * kLWp + j0 + 4 * n);
                v[mm][4 * n] = a.x; v[mm][4 * n + 1] = a.y; v[mm][4 * n + 2] = a.z; v[mm][4 * n + 3] = a.w;
            }
#pragma unroll
        for (int o = 0; o < kQ; ++o) {
            float m[3];
#pragma unroll
            for (int mm = 0; mm < 3; ++mm) {
                m[mm] = 0.f;
#pragma unroll
                for (int k = 0; k < KW; ++k) m[mm] += (KW == 1 ? 1.f : g[k]) * v[mm][o + k];
            }
            const int e = r * rowf + (j0 + o) * C + c;
            const float xv = tx_[e], yv = ty_[e];
            const float d = yv - xv;
            const float sg = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
            tg[e] = sc * (m[0] + 2.f * yv * m[1] + xv * m[2]) + l1s * sg;
        }
    }
}

// Maps are channel-planar: map m of channel c at pixel p is maps[(m C + c) P + p].
template <int C>
__global__ __launch_bounds__(kThreads) void k_ssim_maps(SsimTiles v, const float* __restrict__ X,
                                                        const float* __restrict__ Y, float c1, float c2,
                                                        float* __restrict__ maps, long long P,
                                                        float* __restrict__ partial) {
    extern __shared__ float lds[];
    float* xs = lds;                       // [C][kLH][kLWp]
    float* ys = xs + C * kLH * kLWp;       // [C][kLH][kLWp]
    float* vsum = ys + C * kLH * kLWp;     // [5][kTH][kLWp]  vertical sums
    __shared__ float red[kThreads / 64];
    const int tile = blockIdx.x;
    const int b = view_of_tile(v, tile);
    const int local = tile - v.tile_base[b];
    const int ty = local / v.tiles_x[b], tx = local - ty * v.tiles_x[b];
    const int H = v.H[b], W = v.W[b], kh = v.kh[b], kw = v.kw[b];
    const int Hv = H - kh + 1, Wv = W - kw + 1;
    const int y0 = ty * kTH, x0 = tx * kTW;
    const long long base = v.pix_base[b];
    const int tid = threadIdx.x;

    // one pass over the tile's rows, all channels: each row segment of
    // kLW pixels is kLW * C consecutive floats.  The loop is unrolled so a
    // thread's loads are all in flight before the first LDS store waits.
    constexpr int rowf = kLW * C;
    constexpr int n_stage = kLH * rowf, it_stage = (n_stage + kThreads - 1) / kThreads;
    float xv[it_stage], yv[it_stage];
#pragma unroll
    for (int it = 0; it < it_stage; ++it) {
        const int i = tid + it * kThreads;
        const int r = i / rowf, e = i - r * rowf;
        const int gy = y0 + r, gx = x0 + e / C;
        xv[it] = yv[it] = 0.f;
        if (i < n_stage && gy < H && gx < W) {
            const long long o = (base + (long long)gy * W + x0) * C + e;
            xv[it] = X[o];
            yv[it] = Y[o];
        }
    }
#pragma unroll
    for (int it = 0; it < it_stage; ++it) {
        const int i = tid + it * kThreads;
        if (i >= n_stage) break;
        const int r = i / rowf, e = i - r * rowf;
        const int q = e / C, c = e - q * C;
        xs[(c * kLH + r) * kLWp + q] = xv[it];
        ys[(c * kLH + r) * kLWp + q] = yv[it];
    }
    __syncthreads();
    float g[kMaxTaps];   // the window in registers (every tap loop below is unrolled)
    for (int k = 0; k < kMaxTaps; ++k) g[k] = v.win[k];
    for (int c = 0; c < C; ++c) {
        const float* sx = xs + c * kLH * kLWp;
        const float* sy = ys + c * kLH * kLWp;
        if (kh == 1) moments_v<1>(sx, sy, vsum, g, tid);
        else moments_v<kMaxTaps>(sx, sy, vsum, g, tid);
        __syncthreads();
        float acc = kw == 1 ? ssim_h<C, 1>(vsum, g, c, c1, c2, Hv, Wv, y0, x0, W, base, maps, P, tid)
                            : ssim_h<C, kMaxTaps>(vsum, g, c, c1, c2, Hv, Wv, y0, x0, W, base, maps, P, tid);
        acc = wave_sum(acc);
        if ((tid & 63) == 0) red[tid >> 6] = acc;
        __syncthreads();   // also orders this channel's reads of vsum before the next one's writes
        if (tid == 0) {
            float t = 0.f;
            for (int w = 0; w < kThreads / 64; ++w) t += red[w];
            partial[(long long)tile * C + c] = t;
        }
    }
}

template <int C>
__global__ __launch_bounds__(kThreads) void k_ssim_grad(SsimTiles v, const float* __restrict__ X,
                                                        const float* __restrict__ Y,
                                                        const float* __restrict__ maps, long long P,
                                                        const float* __restrict__ dscale, float l1,
                                                        const float* __restrict__ dl1, float* __restrict__ grad,
                                                        int accumulate, float* __restrict__ l1_partial) {
    extern __shared__ float lds[];
    float* sm = lds;                       // [3][kLH][kLWp]  one channel's maps with the halo
    float* vsum = sm + 3 * kLH * kLWp;     // [3][kTH][kLWp]  their vertical sums
    float* tx_ = vsum + 3 * kTH * kLWp;    // [kTH][kTW * C]  the tile's x, pixel-major
    float* ty_ = tx_ + C * kTH * kTW;      // [kTH][kTW * C]  its y
    float* tg = ty_ + C * kTH * kTW;       // [kTH][kTW * C]  its gradient
    __shared__ float red[kThreads / 64];
    const int tile = blockIdx.x;
    const int b = view_of_tile(v, tile);
    const int local = tile - v.tile_base[b];
    const int tyi = local / v.tiles_x[b], txi = local - tyi * v.tiles_x[b];
    const int H = v.H[b], W = v.W[b], kh = v.kh[b], kw = v.kw[b];
    const int Hv = H - kh + 1, Wv = W - kw + 1;
    const int hy = kh - 1, hx = kw - 1;
    const int y0 = tyi * kTH, x0 = txi * kTW;
    const long long base = v.pix_base[b];
    const int tid = threadIdx.x;
    const float l1s = dl1 ? l1 * dl1[0] : l1;
    constexpr int rowf = kTW * C;
    float g[kMaxTaps];
    for (int k = 0; k < kMaxTaps; ++k) g[k] = v.win[k];

    // the tile's x and y, all channels, row segments of kTW * C floats (the
    // staging loops are unrolled so a thread's loads are in flight together)
    constexpr int n_tile = kTH * rowf, it_tile = (n_tile + kThreads - 1) / kThreads;
    {
        float xv[it_tile], yv[it_tile];
#pragma unroll
        for (int it = 0; it < it_tile; ++it) {
            const int i = tid + it * kThreads;
            const int r = i / rowf, e = i - r * rowf;
            const int gy = y0 + r, gx = x0 + e / C;
            xv[it] = yv[it] = 0.f;
            if (i < n_tile && gy < H && gx < W) {
                const long long o = (base + (long long)gy * W + x0) * C + e;
                xv[it] = X[o];
                yv[it] = Y[o];
            }
        }
#pragma unroll
        for (int it = 0; it < it_tile; ++it) {
            const int i = tid + it * kThreads;
            if (i < n_tile) {
                tx_[i] = xv[it];
                ty_[i] = yv[it];
            }
        }
    }
    constexpr int n_halo = kLH * kLW, it_halo = (n_halo + kThreads - 1) / kThreads;
    // One channel's maps with their halo: windows q with q <= p < q + taps,
    // rows y0 - hy .. y0 + kTH - 1.  The next channel's loads are issued
    // before this channel is filtered, so their latency hides behind it.
    float a[it_halo], bb[it_halo], cc[it_halo];
    auto load_maps = [&](int c) {
#pragma unroll
        for (int it = 0; it < it_halo; ++it) {
            const int i = tid + it * kThreads;
            const int r = i / kLW, q = i - r * kLW;
            const int wy = y0 - hy + r, wx = x0 - hx + q;
            a[it] = bb[it] = cc[it] = 0.f;
            if (i < n_halo && r < kTH + hy && q < kTW + hx && wy >= 0 && wy < Hv && wx >= 0 && wx < Wv) {
                const long long o = base + (long long)wy * W + wx;
                a[it] = maps[(0 * C + c) * P + o];
                bb[it] = maps[(1 * C + c) * P + o];
                cc[it] = maps[(2 * C + c) * P + o];
            }
        }
    };
    load_maps(0);
    for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int it = 0; it < it_halo; ++it) {
            const int i = tid + it * kThreads;
            if (i < n_halo) {
                const int r = i / kLW, q = i - r * kLW;
                sm[(0 * kLH + r) * kLWp + q] = a[it];
                sm[(1 * kLH + r) * kLWp + q] = bb[it];
                sm[(2 * kLH + r) * kLWp + q] = cc[it];
            }
        }
        __syncthreads();
        if (c + 1 < C) load_maps(c + 1);
        if (kh == 1) maps_v<1>(sm, vsum, g, tid);
        else maps_v<kMaxTaps>(sm, vsum, g, tid);
        __syncthreads();
        const float sc = dscale ? v.scale[b * C + c] * dscale[b * C + c] : v.scale[b * C + c];
        if (kw == 1) grad_h<C, 1>(vsum, g, c, sc, l1s, tx_, ty_, tg, tid);
        else grad_h<C, kMaxTaps>(vsum, g, c, sc, l1s, tx_, ty_, tg, tid);
        __syncthreads();   // the next channel overwrites sm and vsum
    }
    // the gradient of all channels, stored in row segments of kTW * C floats
    float l1_acc = 0.f;   // sum of |y - x| over the tile's pixels and channels
    for (int i = tid; i < kTH * rowf; i += kThreads) {
        const int r = i / rowf, e = i - r * rowf;
        const int gy = y0 + r, gx = x0 + e / C;
        if (gy >= H || gx >= W) continue;
        const long long o = (base + (long long)gy * W + x0) * C + e;
        grad[o] = accumulate ? grad[o] + tg[i] : tg[i];
        l1_acc += fabsf(ty_[i] - tx_[i]);
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

// Launch either kernel for the runtime channel count (1 to 4).
template <int C>
hipError_t run_maps(const SsimTiles& t, long long tiles, const float* X, const float* Y, float c1, float c2,
                    float* maps, float* partial, hipStream_t s) {
    const size_t bytes = sizeof(float) * maps_lds_floats<C>();
    static bool opted_in = false;   // more than 64 KiB of LDS needs the kernel's opt-in, once
    if (!opted_in && bytes > 65536) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_ssim_maps<C>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
        if (e != hipSuccess) return e;
        opted_in = true;
    }
    k_ssim_maps<C><<<static_cast<unsigned>(tiles), kThreads, bytes, s>>>(t, X, Y, c1, c2, maps, t.pix_base[t.n],
                                                                          partial);
    return hipGetLastError();
}

template <int C>
hipError_t run_grad(const SsimTiles& t, long long tiles, const float* X, const float* Y, const float* maps,
                    const float* dscale, float l1, const float* dl1, float* grad, int accumulate, float* l1_partial,
                    hipStream_t s) {
    const size_t bytes = sizeof(float) * grad_lds_floats<C>();
    static bool opted_in = false;
    if (!opted_in && bytes > 65536) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_ssim_grad<C>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
        if (e != hipSuccess) return e;
        opted_in = true;
    }
    k_ssim_grad<C><<<static_cast<unsigned>(tiles), kThreads, bytes, s>>>(t, X, Y, maps, t.pix_base[t.n], dscale,
                                                                          l1, dl1, grad, accumulate, l1_partial);
    return hipGetLastError();
}

hipError_t dispatch_maps(const SsimTiles& t, long long tiles, const float* X, const float* Y, float c1, float c2,
                         float* maps, float* partial, hipStream_t s) {
    switch (t.C) {
        case 1: return run_maps<1>(t, tiles, X, Y, c1, c2, maps, partial, s);
        case 2: return run_maps<2>(t, tiles, X, Y, c1, c2, maps, partial, s);
        case 3: return run_maps<3>(t, tiles, X, Y, c1, c2, maps, partial, s);
        case 4: return run_maps<4>(t, tiles, X, Y, c1, c2, maps, partial, s);
    }
    return hipErrorInvalidValue;
}

hipError_t dispatch_grad(const SsimTiles& t, long long tiles, const float* X, const float* Y, const float* maps,
                         const float* dscale, float l1, const float* dl1, float* grad, int accumulate,
                         float* l1_partial, hipStream_t s) {
    switch (t.C) {
        case 1: return run_grad<1>(t, tiles, X, Y, maps, dscale, l1, dl1, grad, accumulate, l1_partial, s);
        case 2: return run_grad<2>(t, tiles, X, Y, maps, dscale, l1, dl1, grad, accumulate, l1_partial, s);
        case 3: return run_grad<3>(t, tiles, X, Y, maps, dscale, l1, dl1, grad, accumulate, l1_partial, s);
        case 4: return run_grad<4>(t, tiles, X, Y, maps, dscale, l1, dl1, grad, accumulate, l1_partial, s);
    }
    return hipErrorInvalidValue;
}

// The training step's photometric loss from the partial sums, one workgroup:
// w_l1 * sum|y - x| / (C P) + w_ssim * (1 - SSIM) / 2, SSIM being the mean
// over views of each view's mean over channels and windows.  Each thread
// strides over all views' partials with four loads in flight, tracking the
// view of its (increasing) index, so the loads of different views overlap.
__global__ __launch_bounds__(1024) void k_photo_loss(SsimTiles v, const float* __restrict__ ssim_partial,
                                                     const float* __restrict__ l1_partial, int n_pix_tiles,
                                                     float w_l1, float w_ssim, float* __restrict__ loss) {
    constexpr int U = 4;
    __shared__ double rs[1024 / 64], rl[1024 / 64];
    __shared__ double inv[kMaxViews];   // 1 / windows of each view
    __shared__ int end[kMaxViews];      // end of each view's partials
    const int C = v.C;
    if ((int)threadIdx.x < v.n) {
        const int b = threadIdx.x;
        inv[b] = 1.0 / ((double)(v.H[b] - v.kh[b] + 1) * (v.W[b] - v.kw[b] + 1));
        end[b] = v.tile_base[b + 1] * C;
    }
    __syncthreads();
    const int total = v.tile_base[v.n] * C;
    double s = 0.0, l = 0.0;
    int b = 0;
    for (int i0 = threadIdx.x; i0 < total; i0 += U * blockDim.x) {
        float x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * blockDim.x;
            x[u] = i < total ? ssim_partial[i] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < total) {
                while (i >= end[b]) ++b;
                s += x[u] * inv[b];
            }
        }
    }
    for (int i0 = threadIdx.x; i0 < n_pix_tiles; i0 += U * blockDim.x) {
        float x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * blockDim.x;
            x[u] = i < n_pix_tiles ? l1_partial[i] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) l += x[u];
    }
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
    // the kernels unroll the window's 11 taps (pytorch_msssim's default size)
    if (B < 1 || B > kMaxViews || !H || !W || C < 1 || C > kMaxC || taps != kMaxTaps) return -1;
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
    return dispatch_maps(t, tiles, X, Y, c1, c2, maps, partial, s);
}

hipError_t launch_ssim_grad(int B, const int* H, const int* W, int C, const float* X, const float* Y, int taps,
                            const float* win, const float* maps, const float* scale, const float* dscale, float l1,
                            const float* dl1, float* grad, int accumulate, hipStream_t s) {
    SsimTiles t;
    const long long tiles = make_tiles(B, H, W, C, taps, win, false, t);
    if (tiles < 0) return hipErrorInvalidValue;
    for (int i = 0; i < B * C; ++i) t.scale[i] = scale[i];
    if (tiles == 0) return hipSuccess;
    return dispatch_grad(t, tiles, X, Y, maps, dscale, l1, dl1, grad, accumulate, nullptr, s);
}

// ---- the training step's L1 + SSIM loss (views of any resolution)

size_t photo_loss_floats(const Views& views, int C) {
    SsimTiles w, p;
    const float ones[kMaxTaps] = {};
    const long long tw = make_tiles(views.n, views.H, views.W, C, kMaxTaps, ones, true, w);
    const long long tp = make_tiles(views.n, views.H, views.W, C, kMaxTaps, ones, false, p);
    if (tw < 0 || tp < 0) return 0;
    // maps [3, C, P], window partials [tw, C], pixel-tile L1 partials [tp]
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
    hipError_t e = dispatch_maps(w, tw, target, color, c1, c2, maps, spart, s);
    if (e != hipSuccess) return e;
    for (int b = 0; b < views.n; ++b) {
        const double windows = (double)(views.H[b] - p.kh[b] + 1) * (views.W[b] - p.kw[b] + 1);
        for (int c = 0; c < C; ++c)
            p.scale[b * C + c] = static_cast<float>(-0.5 * w_ssim / ((double)views.n * C * windows));
    }
    const float l1s = static_cast<float>(w_l1 / ((double)C * w.pix_base[views.n]));
    if ((e = dispatch_grad(p, tp, target, color, maps, nullptr, l1s, nullptr, grad, 0, lpart, s)) != hipSuccess)
        return e;
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
        return ds::ssim_fail(DS_ERR_INVALID, "ssim_fwd: bad views, channel count (1-4) or window size (11)");
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
        return ds::ssim_fail(DS_ERR_INVALID, "ssim_bwd: bad views, channel count (1-4) or window size (11)");
    if (e != hipSuccess) return ds::ssim_fail(DS_ERR_HIP, hipGetErrorString(e));
    return DS_OK;
}

}  // extern "C"

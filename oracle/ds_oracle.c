/*
 * ds_oracle.c - CPU restatement of DiffSoup's rasterizer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the parity tests, the
 * smoke check and the benchmark's CPU-baseline leg compare the HIP library
 * against.  No product code path may load or call it.
 *
 * It follows the reference's own structure (kenji-tojo/diffsoup): per-
 * (view, triangle) rectangles, an explicit candidate-fragment list, an
 * order-preserving compaction, per-fragment opacity, a threshold per
 * fragment, a z-buffer of packed keys whose low word is the fragment's
 * position in the compacted list, and a gather of the winner's attributes.
 * It shares no code with the HIP kernels, which bin triangles into screen
 * tiles and keep the depth key per pixel in registers; agreement between the
 * two is what the parity tests check.
 *
 * Built with -ffp-contract=off and without fast-math, so every float
 * expression rounds once per operation in the order written, which is the
 * reference's order.  Every function cites the reference code it restates.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define DSO_MIN(a, b) ((a) < (b) ? (a) : (b))
#define DSO_MAX(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------------------ */
/* Threads.  One by default: every loop then runs in the order written and  */
/* every result is deterministic (the parity tests use this).  With more,   */
/* the per-pixel and per-fragment loops are spread over OpenMP threads for  */
/* the timed CPU baseline: the fragment list keeps its canonical order, the */
/* z-buffer takes an atomic minimum (order-independent), and scatters into  */
/* shared gradients become atomic adds, whose float sums then depend on the */
/* interleaving.                                                            */
/* ------------------------------------------------------------------------ */
static int dso_nthreads = 1;

void dso_set_threads(int n) {
    dso_nthreads = n < 1 ? 1 : n;
#ifdef _OPENMP
    omp_set_num_threads(dso_nthreads);
#endif
}

int dso_get_threads(void) { return dso_nthreads; }

static inline void dso_addf(float* p, float v) {
    if (dso_nthreads > 1) {
#pragma omp atomic
        *p += v;
    } else {
        *p += v;
    }
}

/* ------------------------------------------------------------------------ */
/* Stochastic-mask hash: SplitMix64 output function (Steele, Lea & Flood,   */
/* OOPSLA 2014) applied twice to (seed, view) and (triangle, row, column).  */
/* ------------------------------------------------------------------------ */
static uint64_t dso_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* One step of the SplitMix64 generator, exposed to check the mixer against
 * the generator's published output sequence. */
uint64_t dso_splitmix64_next(uint64_t* state) {
    *state += 0x9e3779b97f4a7c15ull;
    return dso_mix(*state);
}

float dso_mask_threshold(uint64_t seed, uint32_t view, uint32_t tri, uint32_t y, uint32_t x) {
    uint64_t h = dso_mix(seed + 0x9e3779b97f4a7c15ull * ((uint64_t)view + 1ull));
    h = dso_mix(h ^ (((uint64_t)tri << 32) | ((uint64_t)y << 16) | (uint64_t)x));
    return (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
}

/* ------------------------------------------------------------------------ */
/* Triangle rectangles: src/cuda/rasterize.cu:10-137.                       */
/* rects[i] = (h0, h_len, w0, w_len), counts[i] = h_len * w_len, i = b*T+t. */
/* Returns the total number of candidate fragments.                         */
/* ------------------------------------------------------------------------ */
static int dso_outcode(float x, float y, float z, float w) {
    int c = 0;
    if (x < -w) c |= 1;
    if (x > w) c |= 2;
    if (y < -w) c |= 4;
    if (y > w) c |= 8;
    if (z < -w) c |= 16;
    if (z > w) c |= 32;
    return c;
}

static void dso_ndc_extent(const float* p, float* lo, float* hi) {
    const float inv = 1.f / p[3];
    for (int k = 0; k < 3; ++k) {
        float n = p[k] * inv;
        if (!isfinite(n)) n = n > 0.f ? 1.f : -1.f;
        n = fminf(1.f, fmaxf(-1.f, n));
        lo[k] = fminf(lo[k], n);
        hi[k] = fmaxf(hi[k], n);
    }
}

int64_t dso_triangle_rects(int H, int W, int B, int V, const float* pos, int T, const int32_t* tri,
                           int32_t* rects, int64_t* counts) {
    int64_t total = 0;
    for (int b = 0; b < B; ++b) {
        for (int t = 0; t < T; ++t) {
            const int64_t i = (int64_t)b * T + t;
            int32_t* r = rects + 4 * i;
            r[0] = r[1] = r[2] = r[3] = 0;
            counts[i] = 0;
            const float* p[3];
            int ok = 1;
            for (int k = 0; k < 3; ++k) {
                const int vi = tri[3 * t + k];
                if (vi < 0 || vi >= V) ok = 0;
                p[k] = pos + ((int64_t)b * V + (vi < 0 || vi >= V ? 0 : vi)) * 4;
            }
            if (!ok) continue;
            if (!(p[0][3] > 0.f && p[1][3] > 0.f && p[2][3] > 0.f)) continue;
            if ((dso_outcode(p[0][0], p[0][1], p[0][2], p[0][3]) &
                 dso_outcode(p[1][0], p[1][1], p[1][2], p[1][3]) &
                 dso_outcode(p[2][0], p[2][1], p[2][2], p[2][3])) != 0)
                continue;
            float lo[3] = {1.f, 1.f, 1.f}, hi[3] = {-1.f, -1.f, -1.f};
            for (int k = 0; k < 3; ++k) dso_ndc_extent(p[k], lo, hi);
            if (hi[2] < -1.f || lo[2] > 1.f) continue;
            float x0f = 0.5f * (lo[0] + 1.f) * (float)W;
            float x1f = 0.5f * (hi[0] + 1.f) * (float)W;
            float y0f = 0.5f * (lo[1] + 1.f) * (float)H;
            float y1f = 0.5f * (hi[1] + 1.f) * (float)H;
            if (x0f > x1f) { float s = x0f; x0f = x1f; x1f = s; }
            if (y0f > y1f) { float s = y0f; y0f = y1f; y1f = s; }
            const int w0 = DSO_MAX((int)floorf(x0f), 0);
            const int w1 = DSO_MIN((int)floorf(x1f), W - 1) + 1;
            const int h0 = DSO_MAX((int)floorf(y0f), 0);
            const int h1 = DSO_MIN((int)floorf(y1f), H - 1) + 1;
            const int wl = DSO_MAX(w1 - w0, 0), hl = DSO_MAX(h1 - h0, 0);
            if (wl == 0 || hl == 0) continue;
            r[0] = h0; r[1] = hl; r[2] = w0; r[3] = wl;
            counts[i] = (int64_t)hl * wl;
            total += counts[i];
        }
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* Perspective-correct point-in-triangle: src/cuda/rasterize.cu:170-206.    */
/* ------------------------------------------------------------------------ */
int dso_intersect(float ndc_y, float ndc_x, const float* p0, const float* p1, const float* p2,
                  float* b0, float* b1, float* z, float* w) {
    const float q0x = p0[0] - ndc_x * p0[3], q0y = p0[1] - ndc_y * p0[3];
    const float q1x = p1[0] - ndc_x * p1[3], q1y = p1[1] - ndc_y * p1[3];
    const float q2x = p2[0] - ndc_x * p2[3], q2y = p2[1] - ndc_y * p2[3];
    const float a0 = q1x * q2y - q1y * q2x;
    const float a1 = q2x * q0y - q2y * q0x;
    const float a2 = q0x * q1y - q0y * q1x;
    const float a = a0 + a1 + a2;
    if (a == 0.f) return 0;
    const float ia = 1.f / a;
    *b0 = a0 * ia;
    *b1 = a1 * ia;
    const float b2 = 1.0f - *b0 - *b1;
    *z = *b0 * p0[2] + *b1 * p1[2] + b2 * p2[2];
    *w = *b0 * p0[3] + *b1 * p1[3] + b2 * p2[3];
    return *b0 >= 0.f && *b1 >= 0.f && *b0 + *b1 <= 1.f;
}

static float dso_pixel_ndc(int i, int n) { return -1.f + 2.f * ((float)i + 0.5f) / (float)n; }

/* ------------------------------------------------------------------------ */
/* Candidate fragments + order-preserving compaction:                       */
/* src/cuda/rasterize.cu:208-277 and :402-456 (with the block order fixed). */
/* frag_pix [N,3] = (view, y, x); frag_attrs [N,4] = (b0, b1, z/w, t+1).   */
/* Output order: view, triangle, row-major position in the rectangle.       */
/* ------------------------------------------------------------------------ */
/* One candidate: 1 and its fragment when the pixel centre is inside the
 * triangle with depth in [-1, 1] (src/cuda/rasterize.cu:245-277). */
static int dso_candidate(int H, int W, int b, const float* p0, const float* p1, const float* p2, int t, int y,
                         int x, int32_t* fp, float* fa) {
    if (x < 0 || x >= W || y < 0 || y >= H) return 0;
    const float ny = dso_pixel_ndc(y, H), nx = dso_pixel_ndc(x, W);
    float b0, b1, cz, cw;
    if (!dso_intersect(ny, nx, p0, p1, p2, &b0, &b1, &cz, &cw)) return 0;
    const float depth = cz / cw;
    if (!(depth >= -1.f && depth <= 1.f)) return 0;
    fp[0] = b;
    fp[1] = y;
    fp[2] = x;
    fa[0] = b0;
    fa[1] = b1;
    fa[2] = depth;
    fa[3] = (float)(t + 1);
    return 1;
}

/* Threaded form: every (view, triangle) fills its own slice of a candidate
 * buffer (src/cuda/rasterize.cu:208-277, one slot per rectangle pixel), then
 * one ordered pass compacts the valid ones (:402-456 with a fixed order). */
static int64_t dso_fragments_threaded(int H, int W, int B, int V, const float* pos, int T, const int32_t* tri,
                                      const int32_t* rects, int32_t* frag_pix, float* frag_attrs) {
    const int64_t BT = (int64_t)B * T;
    int64_t* start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(BT + 1));
    start[0] = 0;
    for (int64_t i = 0; i < BT; ++i) {
        const int32_t* r = rects + 4 * i;
        start[i + 1] = start[i] + ((r[1] > 0 && r[3] > 0) ? (int64_t)r[1] * r[3] : 0);
    }
    const int64_t total = start[BT];
    int32_t* cfp = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(total > 0 ? total : 1));
    float* cfa = (float*)malloc(sizeof(float) * 4 * (size_t)(total > 0 ? total : 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < BT; ++i) {
        const int b = (int)(i / T), t = (int)(i % T);
        const int32_t* r = rects + 4 * i;
        const float* bp = pos + (int64_t)b * V * 4;
        const float* p0 = bp + 4 * (int64_t)tri[3 * t + 0];
        const float* p1 = bp + 4 * (int64_t)tri[3 * t + 1];
        const float* p2 = bp + 4 * (int64_t)tri[3 * t + 2];
        for (int64_t k = 0; k < start[i + 1] - start[i]; ++k) {
            const int64_t c = start[i] + k;
            cfp[3 * c] = -1;
            dso_candidate(H, W, b, p0, p1, p2, t, r[0] + (int)(k / r[3]), r[2] + (int)(k % r[3]), cfp + 3 * c,
                          cfa + 4 * c);
        }
    }
    int64_t n = 0;
    for (int64_t c = 0; c < total; ++c) {
        if (cfp[3 * c] < 0) continue;
        memcpy(frag_pix + 3 * n, cfp + 3 * c, 3 * sizeof(int32_t));
        memcpy(frag_attrs + 4 * n, cfa + 4 * c, 4 * sizeof(float));
        ++n;
    }
    free(cfa);
    free(cfp);
    free(start);
    return n;
}

int64_t dso_fragments(int H, int W, int B, int V, const float* pos, int T, const int32_t* tri,
                      const int32_t* rects, int32_t* frag_pix, float* frag_attrs) {
    if (dso_nthreads > 1) return dso_fragments_threaded(H, W, B, V, pos, T, tri, rects, frag_pix, frag_attrs);
    int64_t n = 0;
    for (int b = 0; b < B; ++b) {
        const float* bp = pos + (int64_t)b * V * 4;
        for (int t = 0; t < T; ++t) {
            const int32_t* r = rects + 4 * ((int64_t)b * T + t);
            if (r[1] <= 0 || r[3] <= 0) continue;
            const float* p0 = bp + 4 * (int64_t)tri[3 * t + 0];
            const float* p1 = bp + 4 * (int64_t)tri[3 * t + 1];
            const float* p2 = bp + 4 * (int64_t)tri[3 * t + 2];
            for (int k = 0; k < r[1] * r[3]; ++k) {
                const int y = r[0] + k / r[3];
                const int x = r[2] + k % r[3];
                if (x < 0 || x >= W || y < 0 || y >= H) continue;
                const float ny = dso_pixel_ndc(y, H), nx = dso_pixel_ndc(x, W);
                float b0, b1, cz, cw;
                if (!dso_intersect(ny, nx, p0, p1, p2, &b0, &b1, &cz, &cw)) continue;
                const float depth = cz / cw;
                if (!(depth >= -1.f && depth <= 1.f)) continue;
                frag_pix[3 * n + 0] = b;
                frag_pix[3 * n + 1] = y;
                frag_pix[3 * n + 2] = x;
                frag_attrs[4 * n + 0] = b0;
                frag_attrs[4 * n + 1] = b1;
                frag_attrs[4 * n + 2] = depth;
                frag_attrs[4 * n + 3] = (float)(t + 1);
                ++n;
            }
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Barycentric lattice: src/cuda/multires.cu:10-128, multires.cuh:8-20.    */
/* ------------------------------------------------------------------------ */
uint32_t dso_feats_at_level(uint32_t level) {
    return level == 0u ? 3u : (((1u << (level - 1u)) + 1u) * ((1u << level) + 1u));
}

static void dso_cell(float b0, float b1, uint32_t level, uint32_t idx[3], float wt[3]) {
    const uint32_t res = 1u << level;
    const float rf = (float)res;
    float u = b0 * rf, v = b1 * rf;
    const uint32_t x = DSO_MIN((uint32_t)floorf(u), res - 1u);
    const uint32_t y = DSO_MIN((uint32_t)floorf(v), res - 1u - x);
    u = u - (float)x;
    v = v - (float)y;
    const int flip = u + v > 1.f;
    const uint32_t fu = (uint32_t)flip;
    const float ff = (float)flip;
    const uint32_t xs[3] = {x + 1u, x, x + fu};
    uint32_t ys[3] = {y, y + 1u, 0u};
    ys[2] = DSO_MIN(y + fu, res - xs[2]);
    for (int i = 0; i < 3; ++i) idx[i] = (xs[i] + ys[i]) * (xs[i] + ys[i] + 1u) / 2u + ys[i];
    wt[0] = (1.f - ff) * u + ff * (1.f - v);
    wt[1] = (1.f - ff) * v + ff * (1.f - u);
    wt[2] = 1.f - wt[0] - wt[1];
}

/* out[j] += sum over levels lo..hi and corners of w * feat (out pre-zeroed).
 * Exposed for the closed-form lattice tests. */
void dso_interp(float b0, float b1, int lo, int hi, int C, const float* feat, float* out) {
    uint32_t off = 0;
    for (int l = lo; l <= hi; ++l) {
        uint32_t idx[3];
        float wt[3];
        dso_cell(b0, b1, (uint32_t)l, idx, wt);
        const float* f = feat + (int64_t)off * C;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < C; ++j) out[j] += wt[i] * f[(int64_t)idx[i] * C + j];
        off += dso_feats_at_level((uint32_t)l);
    }
}

static void dso_interp_bwd(float b0, float b1, int lo, int hi, int C, float* grad, const float* g) {
    uint32_t off = 0;
    for (int l = lo; l <= hi; ++l) {
        uint32_t idx[3];
        float wt[3];
        dso_cell(b0, b1, (uint32_t)l, idx, wt);
        float* d = grad + (int64_t)off * C;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < C; ++j) dso_addf(&d[(int64_t)idx[i] * C + j], g[j] * wt[i]);
        off += dso_feats_at_level((uint32_t)l);
    }
}

/* Per-fragment opacity: src/cuda/multires.cu:130-156. */
void dso_multires_alpha(int64_t N, const float* frag_attrs, int level, int S, const float* alpha_src,
                        float* frag_alpha) {
#pragma omp parallel for schedule(static) if (dso_nthreads > 1)
    for (int64_t i = 0; i < N; ++i) {
        const int t = (int)frag_attrs[4 * i + 3] - 1;
        frag_alpha[i] = 0.f;
        if (t < 0) continue;
        dso_interp(frag_attrs[4 * i], frag_attrs[4 * i + 1], level, level, 1,
                   alpha_src + (int64_t)t * S, &frag_alpha[i]);
    }
}

/* Per-fragment threshold: hash of the fragment, or 0.5 when deterministic
 * (python/diffsoup/multires.py:97-100 with the hash in place of torch.rand). */
void dso_mask_thresholds(int64_t N, const int32_t* frag_pix, const float* frag_attrs, int stochastic,
                         uint64_t seed, float* thresh) {
#pragma omp parallel for schedule(static) if (dso_nthreads > 1)
    for (int64_t i = 0; i < N; ++i) {
        thresh[i] = stochastic ? dso_mask_threshold(seed, (uint32_t)frag_pix[3 * i],
                                                    (uint32_t)((int)frag_attrs[4 * i + 3] - 1),
                                                    (uint32_t)frag_pix[3 * i + 1],
                                                    (uint32_t)frag_pix[3 * i + 2])
                               : 0.5f;
    }
}

/* ------------------------------------------------------------------------ */
/* Depth test: src/cuda/rasterize.cu:307-400.  rast [B,H,W,4] pre-zeroed.   */
/* ------------------------------------------------------------------------ */
void dso_depth_test(int B, int H, int W, int64_t N, const int32_t* frag_pix, const float* frag_attrs,
                    const float* frag_alpha, const float* thresh, float* rast) {
    const int64_t P = (int64_t)B * H * W;
    int64_t* zbuf = (int64_t*)malloc(sizeof(int64_t) * (size_t)(P > 0 ? P : 1));
    for (int64_t p = 0; p < P; ++p) zbuf[p] = INT64_MAX;
#pragma omp parallel for schedule(static) if (dso_nthreads > 1)
    for (int64_t i = 0; i < N; ++i) {
        if (frag_pix[3 * i] < 0) continue;
        const int64_t pix = ((int64_t)frag_pix[3 * i] * H + frag_pix[3 * i + 1]) * W + frag_pix[3 * i + 2];
        if (frag_alpha[i] < thresh[i]) continue;
        const float zw = frag_attrs[4 * i + 2];
        if (zw <= -1.f || !isfinite(zw)) continue;
        const float z2 = zw + 2.f;
        uint32_t bits;
        memcpy(&bits, &z2, 4);
        const int64_t key = ((int64_t)(int32_t)bits << 32) | (int64_t)(uint32_t)i;
        if (dso_nthreads > 1) {
            int64_t cur = __atomic_load_n(&zbuf[pix], __ATOMIC_RELAXED);
            while (key < cur &&
                   !__atomic_compare_exchange_n(&zbuf[pix], &cur, key, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
            }
        } else if (key < zbuf[pix]) {
            zbuf[pix] = key;
        }
    }
    for (int64_t p = 0; p < P; ++p) {
        if (zbuf[p] == INT64_MAX) continue;
        const int64_t i = (int64_t)(uint32_t)(zbuf[p] & 0xffffffffll);
        for (int k = 0; k < 4; ++k) rast[4 * p + k] = frag_attrs[4 * i + k];
    }
    free(zbuf);
}

/* ------------------------------------------------------------------------ */
/* Per-pixel features: src/cuda/multires.cu:224-312 (sum of levels), plus  */
/* the per-level concatenation (each slice is the single-level gather).    */
/* ------------------------------------------------------------------------ */
void dso_multires_color(int64_t P, const float* rast, int lo, int hi, int C, const float* feat,
                        int concat, float* out) {
    const int nl = hi - lo + 1;
    const int width = concat ? nl * C : C;
    uint32_t S = 0;
    for (int l = lo; l <= hi; ++l) S += dso_feats_at_level((uint32_t)l);
#pragma omp parallel for schedule(static) if (dso_nthreads > 1)
    for (int64_t p = 0; p < P; ++p) {
        float* o = out + p * width;
        for (int j = 0; j < width; ++j) o[j] = 0.f;
        const int t = (int)rast[4 * p + 3] - 1;
        if (t < 0) continue;
        const float* f = feat + (int64_t)t * S * C;
        if (!concat) {
            dso_interp(rast[4 * p], rast[4 * p + 1], lo, hi, C, f, o);
        } else {
            uint32_t off = 0;
            for (int l = lo; l <= hi; ++l) {
                dso_interp(rast[4 * p], rast[4 * p + 1], l, l, C, f + (int64_t)off * C, o + (l - lo) * C);
                off += dso_feats_at_level((uint32_t)l);
            }
        }
    }
}

void dso_multires_color_bwd(int64_t P, const float* rast, int lo, int hi, int C, float* grad_feat,
                            int concat, const float* grad_out) {
    const int nl = hi - lo + 1;
    const int width = concat ? nl * C : C;
    uint32_t S = 0;
    for (int l = lo; l <= hi; ++l) S += dso_feats_at_level((uint32_t)l);
#pragma omp parallel for schedule(static) if (dso_nthreads > 1)
    for (int64_t p = 0; p < P; ++p) {
        const int t = (int)rast[4 * p + 3] - 1;
        if (t < 0) continue;
        float* gf = grad_feat + (int64_t)t * S * C;
        const float* g = grad_out + p * width;
        if (!concat) {
            dso_interp_bwd(rast[4 * p], rast[4 * p + 1], lo, hi, C, gf, g);
        } else {
            uint32_t off = 0;
            for (int l = lo; l <= hi; ++l) {
                dso_interp_bwd(rast[4 * p], rast[4 * p + 1], l, l, C, gf + (int64_t)off * C, g + (l - lo) * C);
                off += dso_feats_at_level((uint32_t)l);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Cross-level accumulation: src/cuda/multires.cu:313-450.                  */
/* ------------------------------------------------------------------------ */
static void dso_lattice_xy(uint32_t k, uint32_t* x, uint32_t* y) {
    const float fi = (float)k;
    const uint32_t n = (uint32_t)floorf((sqrtf(8.f * fi + 1.f) - 1.f) * 0.5f);
    const uint32_t tn = (n * (n + 1u)) >> 1;
    *y = k - tn;
    *x = n - *y;
}

void dso_accumulate_fwd(int T, int lo, int hi, int target, int C, const float* feat, float* out) {
    uint32_t S = 0;
    for (int l = lo; l <= hi; ++l) S += dso_feats_at_level((uint32_t)l);
    const uint32_t ST = dso_feats_at_level((uint32_t)target);
    const float rf = (float)(1u << target);
    for (int t = 0; t < T; ++t)
        for (uint32_t k = 0; k < ST; ++k) {
            uint32_t x, y;
            dso_lattice_xy(k, &x, &y);
            float* o = out + ((int64_t)t * ST + k) * C;
            for (int j = 0; j < C; ++j) o[j] = 0.f;
            dso_interp((float)x / rf, (float)y / rf, lo, hi, C, feat + (int64_t)t * S * C, o);
        }
}

void dso_accumulate_bwd(int T, int lo, int hi, int target, int C, float* grad_feat, const float* grad_out) {
    uint32_t S = 0;
    for (int l = lo; l <= hi; ++l) S += dso_feats_at_level((uint32_t)l);
    const uint32_t ST = dso_feats_at_level((uint32_t)target);
    const float rf = (float)(1u << target);
    for (int t = 0; t < T; ++t)
        for (uint32_t k = 0; k < ST; ++k) {
            uint32_t x, y;
            dso_lattice_xy(k, &x, &y);
            dso_interp_bwd((float)x / rf, (float)y / rf, lo, hi, C, grad_feat + (int64_t)t * S * C,
                           grad_out + ((int64_t)t * ST + k) * C);
        }
}

/* ------------------------------------------------------------------------ */
/* View-direction SH2: src/cuda/encoding.cu:9-118 (1/sqrtf for rsqrtf).     */
/* ------------------------------------------------------------------------ */
static void dso_unproject(const float* m, float x, float y, float z, float* o) {
    const float c0 = m[0] * x + m[1] * y + m[2] * z + m[3] * 1.0f;
    const float c1 = m[4] * x + m[5] * y + m[6] * z + m[7] * 1.0f;
    const float c2 = m[8] * x + m[9] * y + m[10] * z + m[11] * 1.0f;
    const float c3 = m[12] * x + m[13] * y + m[14] * z + m[15] * 1.0f;
    const float iw = 1.0f / c3;
    o[0] = c0 * iw; o[1] = c1 * iw; o[2] = c2 * iw;
}

void dso_sh2(int B, int H, int W, const float* rast, const float* inv_mvp, float* enc) {
    const float c0 = 0.28209479177387814f, c1 = 0.4886025119029199f;
    const float c2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                         -1.0925484305920792f, 0.5462742152960396f};
    for (int b = 0; b < B; ++b)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int64_t p = ((int64_t)b * H + y) * W + x;
                float* e = enc + 9 * p;
                for (int k = 0; k < 9; ++k) e[k] = 0.f;
                if ((int)rast[4 * p + 3] - 1 < 0) continue;
                const float nx = dso_pixel_ndc(x, W), ny = dso_pixel_ndc(y, H);
                float wn[3], wf[3];
                dso_unproject(inv_mvp + 16 * b, nx, ny, -1.0f, wn);
                dso_unproject(inv_mvp + 16 * b, nx, ny, 1.0f, wf);
                const float dx = wn[0] - wf[0], dy = wn[1] - wf[1], dz = wn[2] - wf[2];
                const float il = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
                const float X = dx * il, Y = dy * il, Z = dz * il;
                const float xx = X * X, yy = Y * Y, zz = Z * Z;
                e[0] = c0;
                e[1] = -c1 * Y;
                e[2] = c1 * Z;
                e[3] = -c1 * X;
                e[4] = c2[0] * (X * Y);
                e[5] = c2[1] * (Y * Z);
                e[6] = c2[2] * (2.0f * zz - xx - yy);
                e[7] = c2[3] * (X * Z);
                e[8] = c2[4] * (xx - yy);
            }
}

/* ------------------------------------------------------------------------ */
/* Edge gradients: src/cuda/rasterize.cu:489-903 (1/sqrtf for rsqrtf).      */
/* ------------------------------------------------------------------------ */
static int dso_seg(float x0, float y0, float x1, float y1, float x2, float y2, float x3, float y3,
                   float* ix, float* iy) {
    const float eps = 1e-7f;
    const float rx = x1 - x0, ry = y1 - y0, sx = x3 - x2, sy = y3 - y2;
    const float den = rx * sy - ry * sx;
    const float qx = x2 - x0, qy = y2 - y0;
    if (fabsf(den) <= eps) {
        if (fabsf(qx * ry - qy * rx) > eps) return 0;
        const float rr = rx * rx + ry * ry;
        if (rr < eps) return 0;
        const float t0 = ((x2 - x0) * rx + (y2 - y0) * ry) / rr;
        const float t1 = ((x3 - x0) * rx + (y3 - y0) * ry) / rr;
        const float tmin = fminf(t0, t1), tmax = fmaxf(t0, t1);
        if (tmax < -eps || tmin > 1.f + eps) return 0;
        const float t = fmaxf(0.f, fminf(1.f, 0.5f * (fmaxf(0.f, tmin) + fminf(1.f, tmax))));
        *ix = x0 + t * rx;
        *iy = y0 + t * ry;
        return 1;
    }
    const float t = (qx * sy - qy * sx) / den;
    const float u = (qx * ry - qy * rx) / den;
    if (t < -eps || t > 1.f + eps || u < -eps || u > 1.f + eps) return 0;
    *ix = x0 + t * rx;
    *iy = y0 + t * ry;
    return 1;
}

static int dso_edge_hit(const float* P[3], float x0, float y0, float x1, float y1, int* i0, int* i1,
                        float* s0, float* s1) {
    float rx[3], ry[3], ix = 0.f, iy = 0.f;
    for (int k = 0; k < 3; ++k) { rx[k] = P[k][0] / P[k][3]; ry[k] = P[k][1] / P[k][3]; }
    int a, c;
    if (dso_seg(x0, y0, x1, y1, rx[0], ry[0], rx[1], ry[1], &ix, &iy)) { a = 0; c = 1; }
    else if (dso_seg(x0, y0, x1, y1, rx[1], ry[1], rx[2], ry[2], &ix, &iy)) { a = 1; c = 2; }
    else if (dso_seg(x0, y0, x1, y1, rx[2], ry[2], rx[0], ry[0], &ix, &iy)) { a = 2; c = 0; }
    else return 0;
    const float* A = P[a];
    const float* Bv = P[c];
    const float dx = Bv[0] - A[0], dy = Bv[1] - A[1], dw = Bv[3] - A[3];
    const float ax = A[0] - ix * A[3], ay = A[1] - iy * A[3];
    const float bx = dx - ix * dw, by = dy - iy * dw;
    const float den = bx * bx + by * by;
    const float t = (den != 0.0f) ? -(ax * bx + ay * by) / den : 0.5f;
    *i0 = a; *i1 = c; *s0 = 1.0f - t; *s1 = t;
    return 1;
}

static int dso_normal(const float* P[3], float n[3]) {
    if (fabsf(P[0][3]) < 1e-30f || fabsf(P[1][3]) < 1e-30f || fabsf(P[2][3]) < 1e-30f) return 0;
    float X[3], Y[3], Z[3];
    for (int k = 0; k < 3; ++k) { X[k] = P[k][0] / P[k][3]; Y[k] = P[k][1] / P[k][3]; Z[k] = P[k][2] / P[k][3]; }
    const float e1x = X[1] - X[0], e1y = Y[1] - Y[0], e1z = Z[1] - Z[0];
    const float e2x = X[2] - X[0], e2y = Y[2] - Y[0], e2z = Z[2] - Z[0];
    const float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const float l2 = nx * nx + ny * ny + nz * nz;
    if (!(l2 > 0.f)) return 0;
    const float il = 1.0f / sqrtf(l2);
    n[0] = nx * il; n[1] = ny * il; n[2] = nz * il;
    return 1;
}

static void dso_db_dp(float nvx, float nvy, float nfx, float nfy, float* dbx, float* dby) {
    float il = 1.0f / sqrtf(fmaxf(1e-30f, nvx * nvx + nvy * nvy));
    nvx *= il; nvy *= il;
    il = 1.0f / sqrtf(fmaxf(1e-30f, nfx * nfx + nfy * nfy));
    nfx *= il; nfy *= il;
    const float bx = -nfy, by = nfx;
    const float den = bx * nvx + by * nvy;
    const float s = (fabsf(den) > 1e-12f) ? (bx / den) : 0.f;
    *dbx = s * nvx;
    *dby = s * nvy;
}

/* Interior / opacity-edge / intersection contribution at the mid point. */
static void dso_mid(float g[12], const float* P[3], float mx, float my, float gx, float gy, float gz,
                    int with_z) {
    float b0, b1, cz, cw;
    if (!dso_intersect(my, mx, P[0], P[1], P[2], &b0, &b1, &cz, &cw)) return;
    const float b2 = 1.f - b0 - b1;
    const float cx = mx * cw, cy = my * cw;
    const float iw = 1.f / cw, iw2 = iw * iw;
    const float gs = with_z ? gx * cx + gy * cy + gz * cz : gx * cx + gy * cy;
    const float bb[3] = {b0, b1, b2};
    for (int k = 0; k < 3; ++k) {
        g[4 * k + 0] += bb[k] * gx * iw;
        g[4 * k + 1] += bb[k] * gy * iw;
        if (with_z) g[4 * k + 2] += bb[k] * gz * iw;
        g[4 * k + 3] += -bb[k] * gs * iw2;
    }
}

void dso_edge_grad(int B, int H, int W, int C, const float* color, const float* grad_color,
                   const float* rast, int V, const float* pos, float* grad_pos, int T, const int32_t* tri) {
    static const int DX[4] = {1, -1, 0, 0}, DY[4] = {0, 0, -1, 1};
#pragma omp parallel for collapse(2) schedule(dynamic, 4) if (dso_nthreads > 1)
    for (int b = 0; b < B; ++b) {
        for (int y0 = 0; y0 < H; ++y0) {
            const float* bp = pos + (int64_t)b * V * 4;
            for (int x0 = 0; x0 < W; ++x0) {
                const int64_t p0 = ((int64_t)b * H + y0) * W + x0;
                const int it0 = (int)rast[4 * p0 + 3] - 1;
                if (it0 < 0 || it0 >= T) continue;
                const float nx0 = dso_pixel_ndc(x0, W), ny0 = dso_pixel_ndc(y0, H);
                const float depth0 = rast[4 * p0 + 2];
                const int32_t* v0 = tri + 3 * it0;
                const float* P0[3] = {bp + 4 * v0[0], bp + 4 * v0[1], bp + 4 * v0[2]};
                float g[12] = {0};
                for (int e = 0; e < 4; ++e) {
                    const int x1 = x0 + DX[e], y1 = y0 + DY[e];
                    if (y1 < 0 || y1 >= H || x1 < 0 || x1 >= W) continue;
                    const int64_t p1 = ((int64_t)b * H + y1) * W + x1;
                    const float depth1 = rast[4 * p1 + 2];
                    int it1 = (int)rast[4 * p1 + 3] - 1;
                    if (it1 >= T) it1 = -1;
                    float adj = 0.f;
                    for (int i = 0; i < C; ++i)
                        adj += 0.5f * (grad_color[p0 * C + i] + grad_color[p1 * C + i]) *
                               (color[p0 * C + i] - color[p1 * C + i]);
                    if (adj == 0.f) continue;
                    const float gx = adj * (float)DX[e] * (float)W / 2.f;
                    const float gy = adj * (float)DY[e] * (float)H / 2.f;
                    const float nx1 = dso_pixel_ndc(x1, W), ny1 = dso_pixel_ndc(y1, H);
                    const float mx = 0.5f * (nx0 + nx1), my = 0.5f * (ny0 + ny1);
                    if (it1 == it0) {
                        dso_mid(g, P0, mx, my, gx, gy, 0.f, 0);
                        continue;
                    }
                    float c1b0, c1b1, cz1, cw1;
                    if (dso_intersect(ny1, nx1, P0[0], P0[1], P0[2], &c1b0, &c1b1, &cz1, &cw1)) {
                        const float d1own = cz1 / cw1;
                        if (it1 >= 0 && depth1 < d1own) {
                            const int32_t* v1 = tri + 3 * it1;
                            const float* P1[3] = {bp + 4 * v1[0], bp + 4 * v1[1], bp + 4 * v1[2]};
                            float o0, o1, oz, ow;
                            if (dso_intersect(ny0, nx0, P1[0], P1[1], P1[2], &o0, &o1, &oz, &ow)) {
                                const float d0other = oz / ow;
                                if (depth0 < d0other) {
                                    float nv[3], nf[3];
                                    if (dso_normal(P0, nv) && dso_normal(P1, nf)) {
                                        float mgx = 0.f, mgy = 0.f, mgz = 0.f;
                                        if (DX[e] != 0) {
                                            float dbx, dbz;
                                            dso_db_dp(nv[0], nv[2], nf[0], nf[2], &dbx, &dbz);
                                            mgx = gx * dbx;
                                            mgz = gx * dbz;
                                        } else {
                                            float dby, dbz;
                                            dso_db_dp(nv[1], nv[2], nf[1], nf[2], &dby, &dbz);
                                            mgy = gy * dby;
                                            mgz = gy * dbz;
                                        }
                                        dso_mid(g, P0, mx, my, mgx, mgy, mgz, 1);
                                    }
                                }
                            }
                        } else if (it1 < 0 || depth0 < depth1) {
                            dso_mid(g, P0, mx, my, gx, gy, 0.f, 0);
                        }
                    } else if (it1 < 0 || depth0 < depth1) {
                        int i0, i1;
                        float s0, s1;
                        if (dso_edge_hit(P0, nx0, ny0, nx1, ny1, &i0, &i1, &s0, &s1)) {
                            const float cx = s0 * P0[i0][0] + s1 * P0[i1][0];
                            const float cy = s0 * P0[i0][1] + s1 * P0[i1][1];
                            const float cw = s0 * P0[i0][3] + s1 * P0[i1][3];
                            const float iw = 1.f / cw, iw2 = iw * iw;
                            const float gs = gx * cx + gy * cy;
                            g[i0 * 4 + 0] += s0 * gx * iw;
                            g[i1 * 4 + 0] += s1 * gx * iw;
                            g[i0 * 4 + 1] += s0 * gy * iw;
                            g[i1 * 4 + 1] += s1 * gy * iw;
                            g[i0 * 4 + 3] += -s0 * gs * iw2;
                            g[i1 * 4 + 3] += -s1 * gs * iw2;
                        }
                    }
                }
                float* gp = grad_pos + (int64_t)b * V * 4;
                for (int k = 0; k < 3; ++k)
                    for (int j = 0; j < 4; ++j) dso_addf(&gp[4 * (int64_t)v0[k] + j], g[4 * k + j]);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Opacity auxiliary gradient: src/cuda/rasterize.cu:923-971 and            */
/* src/cuda/multires.cu:178-202.  grad_alpha [T,S] accumulates, unscaled.   */
/* ------------------------------------------------------------------------ */
void dso_opacity_aux(int B, int H, int W, int C, const float* color, const float* target,
                     const float* rast, int64_t N, const int32_t* frag_pix, const float* frag_attrs,
                     const float* frag_alpha, int level, int S, float* grad_alpha) {
    (void)B;
    const float eps = 1e-5f;
#pragma omp parallel for schedule(static) if (dso_nthreads > 1)
    for (int64_t i = 0; i < N; ++i) {
        if (frag_pix[3 * i] < 0) continue;
        const int64_t p = ((int64_t)frag_pix[3 * i] * H + frag_pix[3 * i + 1]) * W + frag_pix[3 * i + 2];
        const float df = frag_attrs[4 * i + 2];
        const int tf = (int)frag_attrs[4 * i + 3] - 1;
        const float dp = rast[4 * p + 2];
        const int tp = (int)rast[4 * p + 3] - 1;
        if (tp >= 0 && df > dp) continue;
        float loss = 0.f;
        for (int c = 0; c < C; ++c) loss += fabsf(color[p * C + c] - target[p * C + c]);
        const float a = frag_alpha[i];
        float g = 0.f;
        if (tp >= 0 && tf == tp) g += loss / DSO_MAX(a, eps);
        else g -= loss / DSO_MAX(1.f - a, eps);
        uint32_t idx[3];
        float wt[3];
        dso_cell(frag_attrs[4 * i], frag_attrs[4 * i + 1], (uint32_t)level, idx, wt);
        float* d = grad_alpha + (int64_t)tf * S;
        for (int k = 0; k < 3; ++k) dso_addf(&d[idx[k]], g * wt[k]);
    }
}

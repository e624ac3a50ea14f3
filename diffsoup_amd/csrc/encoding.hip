// Per-pixel view direction -> real spherical-harmonic basis up to degree 2.
// Reference: src/cuda/encoding.cu:9-130.  The direction points from the far
// plane point to the near plane point of the pixel's ray (toward the camera),
// unprojected with a row-major inverse MVP.  Normalisation uses a correctly
// rounded 1/sqrt instead of an approximate reciprocal square root.
#include "ds_kernels.h"
#include "ds_math.h"

namespace ds {
namespace {

// Real SH normalisation constants: 1/(2 sqrt(pi)), sqrt(3)/(2 sqrt(pi)),
// sqrt(15)/(2 sqrt(pi)), sqrt(5)/(4 sqrt(pi)), sqrt(15)/(4 sqrt(pi)).
constexpr float kY00 = 0.28209479177387814f;
constexpr float kY1 = 0.4886025119029199f;
constexpr float kY2a = 1.0925484305920792f;
constexpr float kY2b = 0.31539156525252005f;
constexpr float kY2c = 0.5462742152960396f;

__device__ __forceinline__ void unproject(const float* m, float x, float y, float z, float out[3]) {
    const float cx = m[0] * x + m[1] * y + m[2] * z + m[3];
    const float cy = m[4] * x + m[5] * y + m[6] * z + m[7];
    const float cz = m[8] * x + m[9] * y + m[10] * z + m[11];
    const float cw = m[12] * x + m[13] * y + m[14] * z + m[15];
    const float iw = 1.0f / cw;
    out[0] = cx * iw; out[1] = cy * iw; out[2] = cz * iw;
}

__global__ void k_sh2(Views v, const float4* __restrict__ rast, const float* __restrict__ inv_mvp,
                      float* __restrict__ enc) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= v.pix_base[v.n]) return;
    int b = 0;
    while (b + 1 < v.n && v.pix_base[b + 1] <= p) ++b;
    const int local = static_cast<int>(p - v.pix_base[b]);
    const int W = v.W[b], H = v.H[b];
    const int y = local / W, x = local - y * W;
    float* e = enc + p * 9;
    if (static_cast<int>(rast[p].w) - 1 < 0) {
        for (int i = 0; i < 9; ++i) e[i] = 0.f;
        return;
    }
    const float nx = pixel_ndc(x, W), ny = pixel_ndc(y, H);
    const float* m = inv_mvp + 16 * b;
    float wn[3], wf[3];
    unproject(m, nx, ny, -1.0f, wn);
    unproject(m, nx, ny, 1.0f, wf);
    const float dx = wn[0] - wf[0], dy = wn[1] - wf[1], dz = wn[2] - wf[2];
    const float il = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    const float X = dx * il, Y = dy * il, Z = dz * il;
    const float xx = X * X, yy = Y * Y, zz = Z * Z;
    e[0] = kY00;
    e[1] = -kY1 * Y;
    e[2] = kY1 * Z;
    e[3] = -kY1 * X;
    e[4] = kY2a * (X * Y);
    e[5] = -kY2a * (Y * Z);
    e[6] = kY2b * (2.0f * zz - xx - yy);
    e[7] = -kY2a * (X * Z);
    e[8] = kY2c * (xx - yy);
}

}  // namespace

hipError_t launch_sh2(const Views& v, const float* rast, const float* inv_mvp, float* enc,
                      hipStream_t s) {
    const long long P = v.pix_base[v.n];
    if (P <= 0) return hipSuccess;
    k_sh2<<<static_cast<unsigned>((P + 255) / 256), 256, 0, s>>>(
        v, reinterpret_cast<const float4*>(rast), inv_mvp, enc);
    return hipGetLastError();
}

}  // namespace ds

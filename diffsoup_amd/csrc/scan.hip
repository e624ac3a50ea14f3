// Prefix sums of int32 counts.
//
// Up to 65536 items: one workgroup (block_scan_1024 in ds_scan.h).  Larger inputs: three launches, no host
// synchronisation - per-block sums of 4096 items, one workgroup scans the
// block sums (int64), then every block rescans its items with its carry.
// Used for the per-triangle pixel offsets of the training step and for the
// fragment offsets / compaction of the plain path.
#include "ds_kernels.h"
#include "ds_scan.h"

namespace ds {
namespace {

constexpr int kItems = 4 * 1024;
constexpr long long kOneBlock = 64 * 1024;   // single-workgroup scan up to this many items

__global__ void __launch_bounds__(1024) k_block_sums(const int* __restrict__ in, long long n,
                                                     long long* __restrict__ sums) {
    __shared__ long long red[16];
    const long long base = blockIdx.x * (long long)kItems;
    long long s = 0;
    for (int k = 0; k < 4; ++k) {
        const long long i = base + k * 1024 + threadIdx.x;
        if (i < n) s += in[i];
    }
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int k = 0; k < 16; ++k) t += red[k];
        sums[blockIdx.x] = t;
    }
}

// In-place exclusive scan of the nb block sums by one workgroup; sums[nb] = total.
__global__ void __launch_bounds__(1024) k_scan_sums(long long* __restrict__ sums, int nb) {
    __shared__ long long ws[16];
    __shared__ long long carry;
    const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + tid;
        const long long v = i < nb ? sums[i] : 0;
        long long inc = v;
        for (int d = 1; d < 64; d <<= 1) {
            const long long o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) ws[wv] = inc;
        __syncthreads();
        long long pre = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            if (k < wv) pre += ws[k];
            tot += ws[k];
        }
        const long long c = carry;
        if (i < nb) sums[i] = c + pre + inc - v;
        __syncthreads();
        if (tid == 0) carry = c + tot;
        __syncthreads();
    }
    if (tid == 0) sums[nb] = carry;
}

template <typename OUT>
__global__ void __launch_bounds__(1024) k_block_scan(const int* __restrict__ in, long long n,
                                                     const long long* __restrict__ sums, OUT* __restrict__ out,
                                                     int inclusive, int write_total) {
    __shared__ long long ws[16];
    __shared__ long long carry;
    const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    if (tid == 0) carry = sums[blockIdx.x];
    __syncthreads();
    const long long base = blockIdx.x * (long long)kItems;
    for (int k = 0; k < 4; ++k) {
        const long long i = base + k * 1024 + tid;
        const long long v = i < n ? in[i] : 0;
        long long inc = v;
        for (int d = 1; d < 64; d <<= 1) {
            const long long o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) ws[wv] = inc;
        __syncthreads();
        long long pre = 0, tot = 0;
        for (int q = 0; q < 16; ++q) {
            if (q < wv) pre += ws[q];
            tot += ws[q];
        }
        const long long c = carry;
        if (i < n) out[i] = static_cast<OUT>(c + pre + (inclusive ? inc : inc - v));
        __syncthreads();
        if (tid == 0) carry = c + tot;
        __syncthreads();
    }
    if (write_total && blockIdx.x == gridDim.x - 1 && tid == 0) out[n] = static_cast<OUT>(carry);
}

// Inputs up to kOneBlock items: one workgroup (block_scan_1024).
template <typename OUT>
__global__ void __launch_bounds__(1024) k_scan_one(const int* in, int n, OUT* out, long long* total_slot,
                                                   int inclusive, int write_total) {
    const long long total = block_scan_1024<OUT>(in, n, out, inclusive != 0);
    if (threadIdx.x == 0) {
        *total_slot = total;
        if (write_total) out[n] = static_cast<OUT>(total);
    }
}

// Two independent exclusive scans (each at most kOneBlock items), one per
// workgroup of a single launch.
__global__ void __launch_bounds__(1024) k_scan_two(const int* in0, int n0, int* out0, long long* total0,
                                                   const int* in1, int n1, int* out1, long long* total1) {
    const int* in = blockIdx.x == 0 ? in0 : in1;
    const int n = blockIdx.x == 0 ? n0 : n1;
    int* out = blockIdx.x == 0 ? out0 : out1;
    const long long total = block_scan_1024<int>(in, n, out, false);
    if (threadIdx.x == 0) {
        *(blockIdx.x == 0 ? total0 : total1) = total;
        out[n] = static_cast<int>(total);
    }
}

__global__ void k_zero_total(long long* tmp, int* out) {
    tmp[0] = 0;
    if (out) out[0] = 0;
}

}  // namespace

size_t scan_tmp_bytes(long long n) {
    return sizeof(long long) * (size_t)((n + kItems - 1) / kItems + 2);
}

long long* scan_total_slot(void* tmp, long long n) {
    return static_cast<long long*>(tmp) + (n + kItems - 1) / kItems;
}

template <typename OUT>
static hipError_t scan_impl(const int* in, long long n, OUT* out, void* tmp, int inclusive, int write_total,
                            hipStream_t s) {
    long long* sums = static_cast<long long*>(tmp);
    const int nb = static_cast<int>((n + kItems - 1) / kItems);
    if (n > 0 && n <= kOneBlock) {
        k_scan_one<OUT><<<1, 1024, 0, s>>>(in, static_cast<int>(n), out, sums + nb, inclusive, write_total);
        return hipGetLastError();
    }
    if (nb == 0) {
        k_zero_total<<<1, 1, 0, s>>>(sums, write_total ? reinterpret_cast<int*>(out) : nullptr);
        return hipGetLastError();
    }
    k_block_sums<<<nb, 1024, 0, s>>>(in, n, sums);
    k_scan_sums<<<1, 1024, 0, s>>>(sums, nb);
    k_block_scan<OUT><<<nb, 1024, 0, s>>>(in, n, sums, out, inclusive, write_total);
    return hipGetLastError();
}

hipError_t launch_exclusive_scan(const int* in, long long n, int* out, void* tmp, hipStream_t s) {
    return scan_impl<int>(in, n, out, tmp, 0, 1, s);
}

hipError_t launch_exclusive_scan_pair(const int* in0, long long n0, int* out0, void* tmp0, const int* in1,
                                      long long n1, int* out1, void* tmp1, hipStream_t s) {
    if (n0 > 0 && n0 <= kOneBlock && n1 > 0 && n1 <= kOneBlock) {
        k_scan_two<<<2, 1024, 0, s>>>(in0, static_cast<int>(n0), out0, scan_total_slot(tmp0, n0), in1,
                                      static_cast<int>(n1), out1, scan_total_slot(tmp1, n1));
        return hipGetLastError();
    }
    const hipError_t e = scan_impl<int>(in0, n0, out0, tmp0, 0, 1, s);
    if (e != hipSuccess) return e;
    return scan_impl<int>(in1, n1, out1, tmp1, 0, 1, s);
}

hipError_t launch_exclusive_scan_inplace(int* data, long long n, void* tmp, hipStream_t s) {
    return scan_impl<int>(data, n, data, tmp, 0, 0, s);
}

hipError_t launch_inclusive_scan_i64(const int* in, long long n, long long* out, void* tmp, hipStream_t s) {
    return scan_impl<long long>(in, n, out, tmp, 1, 0, s);
}

}  // namespace ds

// Prefix sums of int32 counts (scan.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace ds {

// One-workgroup (1024 threads) exclusive or inclusive scan of n int32 counts
// into out, in passes of 8192 (each pass's loads issued during the previous
// pass): a coalesced load into LDS, a run of 8
// consecutive items per thread (LDS index padded by one word per 32 so the
// runs are read without bank conflicts), one wave scan of the run totals,
// one barrier for the wave totals, and - for int output - the results staged
// back through LDS for a coalesced store.  Returns the 64-bit total (the
// same value in every thread).  Callable from a kernel launched with 1024
// threads.  in == out is safe.
template <typename OUT>
__device__ inline long long block_scan_1024(const int* in, long long n, OUT* out, bool inclusive) {
    constexpr int kPass = 8192;
    __shared__ int buf[kPass + kPass / 32];
    __shared__ long long wsum[16];
    const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    auto pad = [](int i) { return i + (i >> 5); };
    long long carry = 0;
    // the next pass's loads are issued before this pass is scanned, so their
    // latency overlaps the LDS work instead of following it
    int nxt[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int j = k * 1024 + tid;
        nxt[k] = j < n ? in[j] : 0;
    }
    for (long long base = 0; base < n; base += kPass) {
#pragma unroll
        for (int k = 0; k < 8; ++k) buf[pad(k * 1024 + tid)] = nxt[k];
        __syncthreads();
        if (base + kPass < n) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const long long j = base + kPass + k * 1024 + tid;
                nxt[k] = j < n ? in[j] : 0;
            }
        }
        int v[8];
        long long s = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = buf[pad(8 * tid + k)];
            s += v[k];
        }
        long long inc = s;
        for (int d = 1; d < 64; d <<= 1) {
            const long long o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) wsum[wv] = inc;
        __syncthreads();
        long long pre = carry, tot = 0;
        for (int q = 0; q < 16; ++q) {
            const long long x = wsum[q];
            if (q < wv) pre += x;
            tot += x;
        }
        long long run = pre + inc - s;
        if constexpr (sizeof(OUT) == sizeof(int)) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                buf[pad(8 * tid + k)] = static_cast<int>(inclusive ? run + v[k] : run);
                run += v[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int j = k * 1024 + tid;
                if (base + j < n) out[base + j] = static_cast<OUT>(buf[pad(j)]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const long long i = base + 8 * tid + k;
                if (i < n) out[i] = static_cast<OUT>(inclusive ? run + v[k] : run);
                run += v[k];
            }
        }
        carry += tot;
        __syncthreads();   // buf and wsum are reused by the next pass
    }
    return carry;
}

size_t scan_tmp_bytes(long long n);
// Device slot of tmp that holds the total after any scan of n items.
long long* scan_total_slot(void* tmp, long long n);
// out[0..n]: exclusive prefix, out[n] = total.
hipError_t launch_exclusive_scan(const int* in, long long n, int* out, void* tmp, hipStream_t s);
// Two independent scans of the kind above in one launch when both fit one
// workgroup (else two scans); tmp0 / tmp1 are separate scan_tmp_bytes areas.
hipError_t launch_exclusive_scan_pair(const int* in0, long long n0, int* out0, void* tmp0, const int* in1,
                                      long long n1, int* out1, void* tmp1, hipStream_t s);
// In-place exclusive prefix of data[0..n).
hipError_t launch_exclusive_scan_inplace(int* data, long long n, void* tmp, hipStream_t s);
// out[i] = sum of in[0..i] as int64.
hipError_t launch_inclusive_scan_i64(const int* in, long long n, long long* out, void* tmp, hipStream_t s);

}  // namespace ds

// libdqcgen.so: device-side generation of synthetic benchmark columns
// (C ABI in include/dogqc_amd_datagen.h).  Grid-stride kernels, one value
// per thread and trip, stores coalesced except for the text rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "dogqc_amd.h"
#include "dogqc_amd_datagen.h"

namespace {

constexpr int BS = 256;

__device__ __forceinline__ uint64_t rnd(uint64_t seed, uint32_t column, uint64_t row) {
    uint64_t z = (seed ^ ((uint64_t)column * 0x9E3779B97F4A7C15ull)) + (row + 1u) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// lo + floor(u * range), u from the top 32 bits; range <= 2^32
__device__ __forceinline__ int64_t uniformOf(uint64_t r, int64_t lo, uint64_t range) {
    return lo + (int64_t)(((r >> 32) * range) >> 32);
}

template <typename T>
__global__ void __launch_bounds__(BS) k_uniform(T* out, int64_t n, uint64_t seed, uint32_t column, int64_t lo,
                                                uint64_t range) {
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
        out[i] = (T)uniformOf(rnd(seed, column, (uint64_t)i), lo, range);
}

__global__ void __launch_bounds__(BS) k_decimal(double* out, int64_t n, uint64_t seed, uint32_t column, int64_t lo,
                                                uint64_t range, double divisor) {
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
        out[i] = (double)uniformOf(rnd(seed, column, (uint64_t)i), lo, range) / divisor;
}

__global__ void __launch_bounds__(BS) k_pick_u32(uint32_t* out, int64_t n, uint64_t seed, uint32_t column,
                                                 const uint32_t* table, uint64_t ntable) {
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
        out[i] = table[uniformOf(rnd(seed, column, (uint64_t)i), 0, ntable)];
}

__global__ void __launch_bounds__(BS) k_affine_perm(int32_t* out, int64_t n, uint64_t a, uint64_t b) {
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
        out[i] = (int32_t)((a * (uint64_t)i + b) % (uint64_t)n);
}

// row r of a table of nkeys rows (keys from the affine permutation a, b),
// r uniform over the first ceil(nkeys * (i + 1) / n) rows
__global__ void __launch_bounds__(BS) k_prefix_key(int32_t* out, int64_t n, uint64_t seed, uint32_t column,
                                                   int64_t nkeys, uint64_t a, uint64_t b) {
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        const unsigned __int128 num = (unsigned __int128)(uint64_t)nkeys * (uint64_t)(i + 1);
        uint64_t rows = (uint64_t)((num + (uint64_t)n - 1u) / (uint64_t)n);
        rows = rows < 1u ? 1u : (rows > (uint64_t)nkeys ? (uint64_t)nkeys : rows);
        const uint64_t r = (uint64_t)uniformOf(rnd(seed, column, (uint64_t)i), 0, rows);
        out[i] = (int32_t)((a * r + b) % (uint64_t)nkeys);
    }
}

__global__ void __launch_bounds__(BS) k_inverse_cdf(int32_t* out, int64_t n, uint64_t seed, uint32_t column,
                                                    const double* cdf, const int32_t* values, int32_t nvalues) {
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        const double u = (double)(rnd(seed, column, (uint64_t)i) >> 11) * 0x1.0p-53;
        int32_t lo = 0, hi = nvalues;   // first index with cdf > u
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (cdf[mid] > u)
                hi = mid;
            else
                lo = mid + 1;
        }
        out[i] = values[lo < nvalues ? lo : nvalues - 1];
    }
}

__global__ void __launch_bounds__(BS) k_text(uint64_t* offs, char* chars, int64_t n, int32_t width, uint64_t seed,
                                             uint32_t column, const char* vchars, const int32_t* voffs,
                                             int32_t nvocab, int32_t words) {
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i <= n; i += (int64_t)gridDim.x * BS) {
        offs[i] = (uint64_t)i * (uint64_t)width;
        if (i == n) break;
        char* row = chars + i * width;
        int32_t pos = 0;
        for (int32_t j = 0; j < words; ++j) {
            const int64_t w = uniformOf(rnd(seed, column + (uint32_t)j, (uint64_t)i), 0, (uint64_t)nvocab);
            if (j > 0 && pos < width) row[pos++] = ' ';
            for (int32_t c = voffs[w]; c < voffs[w + 1] && pos < width; ++c) row[pos++] = vchars[c];
        }
        for (; pos < width; ++pos) row[pos] = ' ';
    }
}

unsigned gridFor(int64_t n) {
    const int64_t g = (n + BS - 1) / BS;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

int check() { return hipGetLastError() == hipSuccess ? DQC_OK : DQC_EHIP; }

}  // namespace

extern "C" {

int dqc_gen_uniform(void* out, int32_t dtype, int64_t n, uint64_t seed, uint32_t column, int64_t lo, int64_t hi,
                    void* stream) {
    if (n < 0 || hi < lo || (uint64_t)(hi - lo) >= (1ull << 32) || (n > 0 && !out)) return DQC_EINVAL;
    if (n == 0) return DQC_OK;
    const uint64_t range = (uint64_t)(hi - lo) + 1u;
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case DQC_GEN_I8:
            hipLaunchKernelGGL(k_uniform<int8_t>, dim3(gridFor(n)), dim3(BS), 0, s, (int8_t*)out, n, seed, column, lo,
                               range);
            break;
        case DQC_GEN_I32:
            hipLaunchKernelGGL(k_uniform<int32_t>, dim3(gridFor(n)), dim3(BS), 0, s, (int32_t*)out, n, seed, column,
                               lo, range);
            break;
        case DQC_GEN_U32:
            hipLaunchKernelGGL(k_uniform<uint32_t>, dim3(gridFor(n)), dim3(BS), 0, s, (uint32_t*)out, n, seed,
                               column, lo, range);
            break;
        case DQC_GEN_I64:
            hipLaunchKernelGGL(k_uniform<int64_t>, dim3(gridFor(n)), dim3(BS), 0, s, (int64_t*)out, n, seed, column,
                               lo, range);
            break;
        default:
            return DQC_EINVAL;
    }
    return check();
}

int dqc_gen_decimal(double* out, int64_t n, uint64_t seed, uint32_t column, int64_t lo, int64_t hi,
                    double divisor, void* stream) {
    if (n < 0 || hi < lo || (uint64_t)(hi - lo) >= (1ull << 32) || (n > 0 && !out)) return DQC_EINVAL;
    if (n == 0) return DQC_OK;
    hipLaunchKernelGGL(k_decimal, dim3(gridFor(n)), dim3(BS), 0, (hipStream_t)stream, out, n, seed, column, lo,
                       (uint64_t)(hi - lo) + 1u, divisor);
    return check();
}

int dqc_gen_pick_u32(uint32_t* out, int64_t n, uint64_t seed, uint32_t column, const uint32_t* table,
                     int64_t ntable, void* stream) {
    if (n < 0 || ntable <= 0 || ntable > (1ll << 32) || (n > 0 && (!out || !table))) return DQC_EINVAL;
    if (n == 0) return DQC_OK;
    hipLaunchKernelGGL(k_pick_u32, dim3(gridFor(n)), dim3(BS), 0, (hipStream_t)stream, out, n, seed, column, table,
                       (uint64_t)ntable);
    return check();
}

int dqc_gen_affine_perm(int32_t* out, int64_t n, uint64_t a, uint64_t b, void* stream) {
    if (n < 0 || n >= (1ll << 31) || (n > 0 && (!out || a >= (uint64_t)n || b >= (uint64_t)n))) return DQC_EINVAL;
    if (n == 0) return DQC_OK;
    hipLaunchKernelGGL(k_affine_perm, dim3(gridFor(n)), dim3(BS), 0, (hipStream_t)stream, out, n, a, b);
    return check();
}

int dqc_gen_prefix_key(int32_t* out, int64_t n, uint64_t seed, uint32_t column, int64_t nkeys, uint64_t a,
                       uint64_t b, void* stream) {
    if (n < 0 || nkeys <= 0 || nkeys >= (1ll << 31) || a >= (uint64_t)nkeys || b >= (uint64_t)nkeys ||
        (n > 0 && !out))
        return DQC_EINVAL;
    if (n == 0) return DQC_OK;
    hipLaunchKernelGGL(k_prefix_key, dim3(gridFor(n)), dim3(BS), 0, (hipStream_t)stream, out, n, seed, column, nkeys,
                       a, b);
    return check();
}

int dqc_gen_inverse_cdf(int32_t* out, int64_t n, uint64_t seed, uint32_t column, const double* cdf,
                        const int32_t* values, int32_t nvalues, void* stream) {
    if (n < 0 || nvalues <= 0 || (n > 0 && (!out || !cdf || !values))) return DQC_EINVAL;
    if (n == 0) return DQC_OK;
    hipLaunchKernelGGL(k_inverse_cdf, dim3(gridFor(n)), dim3(BS), 0, (hipStream_t)stream, out, n, seed, column, cdf,
                       values, nvalues);
    return check();
}

int dqc_gen_text(uint64_t* offs, char* chars, int64_t n, int32_t width, uint64_t seed, uint32_t column,
                 const char* vchars, const int32_t* voffs, int32_t nvocab, int32_t words, void* stream) {
    if (n < 0 || width <= 0 || nvocab <= 0 || words < 0 || !offs || (n > 0 && (!chars || !vchars || !voffs)))
        return DQC_EINVAL;
    hipLaunchKernelGGL(k_text, dim3(gridFor(n + 1)), dim3(BS), 0, (hipStream_t)stream, offs, chars, n, width, seed,
                       column, vchars, voffs, nvocab, words);
    return check();
}

}  // extern "C"

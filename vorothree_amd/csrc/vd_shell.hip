// Device construction of a bin shell order prefix (Algorithm3DGrid::new,
// src/algorithm/algo_3d_grid.rs:47-95).
//
// The reference enumerates every offset (dx, dy, dz) in [-nx, nx] x [-ny, ny] x [-nz, nz] (z outer,
// x inner), keys it by the squared minimum bin distance (algo_3d_grid.rs:61-66) and sorts by key.
// Here only the offsets inside the box [-K, K]^3 whose key is <= T are kept (a prefix of the full
// order), ties keep enumeration order (a stable LSD radix sort over the key bits, values = the
// enumeration index), so the prefix is the same sequence a stable sort of the full table starts
// with. Keys are non-negative doubles: their bit patterns order as unsigned integers. The table
// holds the offsets only (packed, 4 bytes each); the walk recomputes each key with shell_key.
//
// HBM-bound integer work: one thread per enumerated offset; the sort is rocPRIM's onesweep radix
// sort over 63 key bits.
#include <rocprim/device/device_radix_sort.hpp>

#include "vd_internal.h"

namespace vd {
namespace {

constexpr int kBlock = 256;

struct BoxGeom {
    ShellPack pack;
    uint32_t wx, wy;      // 2K + 1 along x and y
    double cs[3];
};

BoxGeom box_geom(const ShellPack &pack, const double cs[3]) {
    BoxGeom b;
    b.pack = pack;
    for (int a = 0; a < 3; ++a) b.cs[a] = cs[a];
    b.wx = 2u * (uint32_t)pack.k[0] + 1u;
    b.wy = 2u * (uint32_t)pack.k[1] + 1u;
    return b;
}

// enumeration index (z outer, x inner) -> box coordinates x + K0, y + K1, z + K2
__device__ inline void decode(const BoxGeom &b, uint32_t i, uint32_t &x, uint32_t &y, uint32_t &z) {
    x = i % b.wx;
    const uint32_t r = i / b.wx;
    y = r % b.wy;
    z = r / b.wy;
}

__global__ void k_zero(uint32_t *p) {
    if (threadIdx.x == 0) p[0] = 0u;
}

// Keys and enumeration indices of every offset of the box; count[0] += offsets with key <= T.
__global__ void k_shell_keys(BoxGeom b, double T, uint32_t w, uint64_t *__restrict__ keys,
                             uint32_t *__restrict__ vals, uint32_t *__restrict__ count) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    bool in = false;
    if (i < w) {
        uint32_t x, y, z;
        decode(b, i, x, y, z);
        const double d2 = shell_key((int)x - b.pack.k[0], (int)y - b.pack.k[1], (int)z - b.pack.k[2], b.cs);
        keys[i] = (uint64_t)__double_as_longlong(d2);
        vals[i] = i;
        in = d2 <= T;
    }
    const uint64_t m = __ballot(in);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (uint32_t)__popcll(m));
}

// Sorted prefix -> packed offsets; next[0] = key of the first offset after the prefix (+inf if none).
__global__ void k_shell_write(BoxGeom b, uint32_t w, uint32_t count, const uint64_t *__restrict__ keys,
                              const uint32_t *__restrict__ vals, ShellCode *__restrict__ out,
                              uint64_t *__restrict__ next) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) {
        uint32_t x, y, z;
        decode(b, vals[i], x, y, z);
        out[i] = x | (y << b.pack.sy) | (z << b.pack.sz);
    } else if (i == count) {
        next[0] = count < w ? keys[count] : 0x7ff0000000000000ull;
    }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// Scratch layout: keys[w], keys_out[w], vals[w], vals_out[w], counter, radix-sort temporary.
size_t shell_scratch_bytes(uint32_t w) {
    size_t sort_bytes = 0;
    uint64_t *k = nullptr;
    uint32_t *v = nullptr;
    (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, k, k, v, v, w, 0, 63);
    return 2 * align256((size_t)w * 8) + 2 * align256((size_t)w * 4) + 256 + align256(sort_bytes);
}

hipError_t shell_count(const ShellPack &pack, const double cs[3], double T, uint32_t w, void *scratch,
                       hipStream_t st) {
    const BoxGeom b = box_geom(pack, cs);
    char *p = static_cast<char *>(scratch);
    uint64_t *keys = reinterpret_cast<uint64_t *>(p);
    uint32_t *vals = reinterpret_cast<uint32_t *>(p + 2 * align256((size_t)w * 8));
    uint32_t *count = reinterpret_cast<uint32_t *>(p + 2 * align256((size_t)w * 8) + 2 * align256((size_t)w * 4));
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, count);
    hipLaunchKernelGGL(k_shell_keys, dim3((w + kBlock - 1) / kBlock), dim3(kBlock), 0, st, b, T, w, keys, vals,
                       count);
    return hipGetLastError();
}

const uint32_t *shell_counter(const void *scratch, uint32_t w) {
    const char *p = static_cast<const char *>(scratch);
    return reinterpret_cast<const uint32_t *>(p + 2 * align256((size_t)w * 8) + 2 * align256((size_t)w * 4));
}

hipError_t shell_finish(const ShellPack &pack, const double cs[3], uint32_t w, uint32_t count, void *scratch,
                        size_t scratch_bytes, ShellCode *out, uint64_t *next, hipStream_t st) {
    const BoxGeom b = box_geom(pack, cs);
    char *p = static_cast<char *>(scratch);
    const size_t kb = align256((size_t)w * 8), vb = align256((size_t)w * 4);
    uint64_t *keys = reinterpret_cast<uint64_t *>(p);
    uint64_t *keys_out = reinterpret_cast<uint64_t *>(p + kb);
    uint32_t *vals = reinterpret_cast<uint32_t *>(p + 2 * kb);
    uint32_t *vals_out = reinterpret_cast<uint32_t *>(p + 2 * kb + vb);
    void *tmp = p + 2 * kb + 2 * vb + 256;
    size_t tmp_bytes = scratch_bytes - (2 * kb + 2 * vb + 256);
    hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys_out, vals, vals_out, w, 0, 63, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_shell_write, dim3(count / kBlock + 1), dim3(kBlock), 0, st, b, w, count, keys_out,
                       vals_out, out, next);
    return hipGetLastError();
}

}  // namespace vd

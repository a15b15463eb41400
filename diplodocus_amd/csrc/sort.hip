// Onesweep LSD radix sort of (key, value) pairs for gfx950 (after Adinets & Merrill 2022,
// "Onesweep: A Faster Least Significant Digit Radix Sort for GPUs").
//
//   1. one pass over the keys builds the digit histograms of every pass in LDS;
//   2. one workgroup per pass turns them into exclusive digit bases;
//   3. per pass, one kernel: workgroup b sorts tile b; it ranks its keys stably with
//      64-lane ballot matching (lanes holding the same digit form a peer mask; rank = lower
//      peers + earlier items of the wave + earlier waves), learns the tile's global digit
//      offsets by a decoupled look-back on per-(tile, digit) status words, and scatters
//      keys and values.
// The look-back waits on earlier tiles only for a bounded time: a predecessor that has not
// published its count within `spin_ticks` of the 100 MHz real-time clock (for instance,
// one not yet resident) is counted from its keys by the waiting lane. Every wait therefore
// ends, whatever order the hardware starts workgroups in, and no tile counter (a
// contended atomic per workgroup) is needed.
// Stable: equal keys keep their input order, so the result equals std::stable_sort by key.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "dpl_internal.h"

namespace dpl {

namespace {

// 512-thread tiles (8 waves per tile for its load, ranking and scatter latencies): the eight
// passes over the 2M 64-bit keys of the 2M-triangle build take 191 us against 212 us with
// 256 threads x 12 keys; 256 x 6 / 8 / 10 keys, 1024 x 3 and 512 x 6 / 10 / 12 keys measured
// slower. 32-bit keys: 12 per thread (the 2M-triangle Morton30 build 2.29 -> 2.26 ms).
constexpr int kSortBlock = 512;
constexpr int kSortWaves = kSortBlock / kWave;
constexpr uint32_t kStAgg = 1u << 30;
constexpr uint32_t kStInc = 2u << 30;
constexpr uint32_t kStMask = (1u << 30) - 1u;

template <typename K>
struct SortTraits {
    static constexpr int kIpt = sizeof(K) == 4 ? 12 : 8;
    static constexpr int kTile = kSortBlock * kIpt;
};

struct SortTemp {
    uint32_t* hist;    // P x RADIX
    uint32_t* base;    // P x RADIX
    uint32_t* status;  // P x T x RADIX
};

// Look-back wait before a lane counts an unpublished predecessor itself: 100 us.
constexpr int64_t kSpinTicks = 10000;

int Passes(int key_bits, int digit_bits) { return (key_bits + digit_bits - 1) / digit_bits; }

SortTemp CarveTemp(void* temp, int passes, int radix, int64_t tiles) {
    SortTemp t;
    uint32_t* p = static_cast<uint32_t*>(temp);
    t.hist = p;
    p += size_t(passes) * radix;
    t.base = p;
    p += size_t(passes) * radix;
    t.status = p;
    return t;
}

}  // namespace

__global__ void KSortClear(uint32_t* p, int64_t words) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < words; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0u;
}

template <typename K, int DB>
__global__ __launch_bounds__(kSortBlock) void KSortHistogram(const K* __restrict__ keys, int64_t n, int passes,
                                                              uint32_t* hist) {
    constexpr int RADIX = 1 << DB;
    constexpr int MAXP = (8 * int(sizeof(K)) + DB - 1) / DB;
    __shared__ uint32_t s_h[MAXP * RADIX];
    for (int i = threadIdx.x; i < passes * RADIX; i += kSortBlock) s_h[i] = 0u;
    __syncthreads();
    for (int64_t i = blockIdx.x * int64_t(kSortBlock) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kSortBlock) {
        const K k = keys[i];
        for (int p = 0; p < passes; p++) atomicAdd(&s_h[p * RADIX + int((k >> (p * DB)) & K(RADIX - 1))], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * RADIX; i += kSortBlock) {
        const uint32_t v = s_h[i];
        if (v) atomicAdd(&hist[i], v);
    }
}

template <int DB>
__global__ __launch_bounds__(kSortBlock) void KSortBases(const uint32_t* hist, uint32_t* base) {
    constexpr int RADIX = 1 << DB;
    __shared__ uint32_t s_scan[kSortBlock / kWave + 1];
    const int p = blockIdx.x;
    const uint32_t v = threadIdx.x < RADIX ? hist[p * RADIX + threadIdx.x] : 0u;
    uint32_t tot;
    const uint32_t ex = BlockExclusiveScan<kSortBlock>(v, s_scan, tot);
    if (threadIdx.x < RADIX) base[p * RADIX + threadIdx.x] = ex;
}

// Occurrences of digit r in full tile p (the look-back's fallback for a predecessor that has
// not published).
template <typename K, int DB>
__device__ uint32_t CountDigit(const K* __restrict__ kin, int64_t p, int shift, uint32_t r) {
    constexpr int TILE = SortTraits<K>::kTile;
    const K* src = kin + p * TILE;
    uint32_t c = 0;
#pragma unroll 8
    for (int i = 0; i < TILE; i++) c += uint32_t((src[i] >> shift) & K((1 << DB) - 1)) == r ? 1u : 0u;
    return c;
}

template <typename K, int DB>
__global__ __launch_bounds__(kSortBlock) void KOnesweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                        K* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n,
                                                        int pass, int shift, SortTemp tmp, int64_t tiles,
                                                        int64_t spin_ticks) {
    constexpr int RADIX = 1 << DB;
    constexpr int IPT = SortTraits<K>::kIpt;
    constexpr int TILE = SortTraits<K>::kTile;
    __shared__ uint32_t s_wh[kSortWaves][RADIX];
    __shared__ uint32_t s_start[RADIX];  // the tile's digit runs: first local position
    __shared__ uint32_t s_off[RADIX];    // the digit's global destination of that position
    __shared__ uint32_t s_scan[kSortBlock / kWave + 1];
    __shared__ K s_key[TILE];
    __shared__ uint32_t s_val[TILE];

    for (int i = threadIdx.x; i < kSortWaves * RADIX; i += kSortBlock) (&s_wh[0][0])[i] = 0u;
    __syncthreads();
    const int64_t tile = blockIdx.x;

    const int lane = LaneId(), w = threadIdx.x / kWave;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int64_t wbase = tile * TILE + int64_t(w) * (kWave * IPT);
    K key[IPT];
    uint32_t val[IPT];
    uint32_t rank[IPT];
    uint32_t dig[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const int64_t i = wbase + int64_t(j) * kWave + lane;
        const bool valid = i < n;
        key[j] = valid ? kin[i] : K(0);
        val[j] = valid ? vin[i] : 0u;
    }
    volatile uint32_t* wh = s_wh[w];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const int64_t i = wbase + int64_t(j) * kWave + lane;
        const bool valid = i < n;
        const uint32_t d = uint32_t((key[j] >> shift) & K(RADIX - 1));
        dig[j] = valid ? d : uint32_t(RADIX);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < DB; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t r = 0;
        if (valid) {
            const uint32_t before = wh[d];
            r = before + uint32_t(__popcll(peers & lt_mask));
            __builtin_amdgcn_wave_barrier();
            const int leader = __ffsll(static_cast<unsigned long long>(peers)) - 1;
            if (lane == leader) wh[d] = before + uint32_t(__popcll(peers));
        } else {
            __builtin_amdgcn_wave_barrier();
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        rank[j] = r;
    }
    __syncthreads();

    // per digit: wave prefixes and the tile's count; local digit runs by a block scan; the
    // global offsets by a decoupled look-back across tiles
    static_assert(RADIX <= kSortBlock, "one digit per thread");
    uint32_t acc = 0;
    const int r = threadIdx.x;
    if (r < RADIX) {
#pragma unroll
        for (int ww = 0; ww < kSortWaves; ww++) {
            const uint32_t t = s_wh[ww][r];
            s_wh[ww][r] = acc;
            acc += t;
        }
    }
    uint32_t tile_total;
    const uint32_t start = BlockExclusiveScan<kSortBlock>(acc, s_scan, tile_total);
    uint32_t* st = tmp.status + size_t(pass) * size_t(tiles) * RADIX;
    if (r < RADIX) {
        uint32_t excl = 0;
        if (tile == 0) {
            __hip_atomic_store(&st[r], kStInc | acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&st[size_t(tile) * RADIX + r], kStAgg | acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // eight predecessors per step (their loads in flight together), consumed in order
            // up to the first inclusive prefix or the first not yet published; a predecessor
            // still unpublished after spin_ticks is counted from its keys (predecessors are
            // full tiles: tile < tiles)
            constexpr int kLb = 8;
            int64_t p = tile - 1;
            bool waiting = false;
            uint64_t since = 0;
            for (;;) {
                uint32_t v[kLb];
#pragma unroll
                for (int q = 0; q < kLb; q++)
                    v[q] = p - q >= 0 ? __hip_atomic_load(&st[size_t(p - q) * RADIX + r], __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : kStInc;
                bool done = false, wait = false;
                int used = 0;
#pragma unroll
                for (int q = 0; q < kLb; q++) {
                    if (done || wait) continue;
                    const uint32_t flag = v[q] & ~kStMask;
                    if (flag == 0u) {
                        wait = true;
                        continue;
                    }
                    excl += v[q] & kStMask;
                    used = q + 1;
                    done = flag == kStInc;
                }
                if (done) break;
                p -= used;
                if (!wait) continue;
                const uint64_t now = __builtin_amdgcn_s_memrealtime();
                if (used > 0 || !waiting) {
                    waiting = true;
                    since = now;
                } else if (int64_t(now - since) >= spin_ticks) {
                    excl += CountDigit<K, DB>(kin, p, shift, uint32_t(r));
                    p -= 1;
                    waiting = false;
                    continue;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __hip_atomic_store(&st[size_t(tile) * RADIX + r], kStInc | (excl + acc), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_start[r] = start;
        s_off[r] = tmp.base[pass * RADIX + r] + excl;
    }
    __syncthreads();

    // stage the tile in LDS in (digit, input) order, then write each digit run contiguously
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const uint32_t d = dig[j];
        if (d >= uint32_t(RADIX)) continue;
        const uint32_t pos = s_start[d] + s_wh[w][d] + rank[j];
        s_key[pos] = key[j];
        s_val[pos] = val[j];
    }
    __syncthreads();
    const int count = int(tile_total);
    for (int k = threadIdx.x; k < count; k += kSortBlock) {
        const K kk = s_key[k];
        const uint32_t d = uint32_t((kk >> shift) & K(RADIX - 1));
        const uint32_t dst = s_off[d] + (uint32_t(k) - s_start[d]);
        if (dst >= n) continue;  // unreachable: the offsets are exact
        kout[dst] = kk;
        vout[dst] = s_val[k];
    }
}

size_t RadixSortTempBytes(int64_t n, int key_bits, int digit_bits, bool key64) {
    const int passes = Passes(key_bits, digit_bits);
    const int radix = 1 << digit_bits;
    const int64_t tile = key64 ? SortTraits<uint64_t>::kTile : SortTraits<uint32_t>::kTile;
    const int64_t tiles = (n + tile - 1) / tile;
    return (size_t(passes) * radix * 2 + size_t(passes) * size_t(tiles) * radix) * sizeof(uint32_t);
}

template <typename K, int DB>
static hipError_t SortImpl(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n, int key_bits,
                           void* temp, size_t temp_bytes, bool* result_in_alt, hipStream_t s) {
    constexpr int RADIX = 1 << DB;
    const int passes = Passes(key_bits, DB);
    const int64_t tile = SortTraits<K>::kTile;
    const int64_t tiles = (n + tile - 1) / tile;
    const size_t words = temp_bytes / sizeof(uint32_t);
    if (RadixSortTempBytes(n, key_bits, DB, sizeof(K) == 8) > temp_bytes) return hipErrorInvalidValue;
    SortTemp t = CarveTemp(temp, passes, RADIX, tiles);
    {
        const int64_t clear = int64_t(words);
        int grid = int(std::min<int64_t>(2048, (clear + 255) / 256));
        hipLaunchKernelGGL(KSortClear, dim3(grid), dim3(256), 0, s, static_cast<uint32_t*>(temp), clear);
    }
    {
        int grid = int(std::min<int64_t>(1024, (n + kSortBlock * 16 - 1) / (kSortBlock * 16)));
        if (grid < 1) grid = 1;
        hipLaunchKernelGGL((KSortHistogram<K, DB>), dim3(grid), dim3(kSortBlock), 0, s, keys, n, passes, t.hist);
    }
    hipLaunchKernelGGL(KSortBases<DB>, dim3(passes), dim3(kSortBlock), 0, s, t.hist, t.base);
    K* ki = keys;
    uint32_t* vi = vals;
    K* ko = keys_alt;
    uint32_t* vo = vals_alt;
    // DPL_SORT_LOOKBACK_SPIN_TICKS overrides the look-back wait (tests set 0 to force the
    // counting fallback on every wait)
    int64_t spin = kSpinTicks;
    if (const char* e = std::getenv("DPL_SORT_LOOKBACK_SPIN_TICKS")) spin = std::max<int64_t>(0, std::atoll(e));
    for (int p = 0; p < passes; p++) {
        hipLaunchKernelGGL((KOnesweep<K, DB>), dim3(unsigned(tiles)), dim3(kSortBlock), 0, s, ki, vi, ko, vo, n, p,
                           p * DB, t, tiles, spin);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    *result_in_alt = (passes % 2) == 1;
    return hipGetLastError();
}

hipError_t LaunchRadixSort(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, int64_t n,
                           int key_bits, int digit_bits, bool key64, void* temp, size_t temp_bytes, bool* result_in_alt,
                           hipStream_t s) {
    *result_in_alt = false;
    if (n <= 0) return hipSuccess;
    if (key64) {
        if (digit_bits == 4)
            return SortImpl<uint64_t, 4>(keys, vals, keys_alt, vals_alt, n, key_bits, temp, temp_bytes, result_in_alt, s);
        return SortImpl<uint64_t, 8>(keys, vals, keys_alt, vals_alt, n, key_bits, temp, temp_bytes, result_in_alt, s);
    }
    uint32_t* k32 = reinterpret_cast<uint32_t*>(keys);
    uint32_t* a32 = reinterpret_cast<uint32_t*>(keys_alt);
    if (digit_bits == 4)
        return SortImpl<uint32_t, 4>(k32, vals, a32, vals_alt, n, key_bits, temp, temp_bytes, result_in_alt, s);
    return SortImpl<uint32_t, 8>(k32, vals, a32, vals_alt, n, key_bits, temp, temp_bytes, result_in_alt, s);
}

}  // namespace dpl

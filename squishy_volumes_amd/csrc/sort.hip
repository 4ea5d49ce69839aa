// sort.hip - Morton block keys, stable LSD radix sort, the counting re-sort of
// keys that a transfer wrote, device-wide scans and the occupied-block list.
//
// The radix sort is the reduce-then-scan form: per pass, a histogram kernel
// counts the digits of each 4096-key tile, a device scan turns the digit-major
// tile counts into output offsets, and a scatter kernel ranks every key inside
// its tile stably.  Ranking within a wave64 uses ballots over the digit bits
// (a match of equal digits), so equal keys keep their input order and the
// result equals a stable CPU sort by key (the reference's sort_on_cpu,
// rust/crates/gpu/src/util.rs:65-69, and its radix sort,
// rust/crates/gpu/src/radix_sort/mod.rs:148-238).
#include <algorithm>

#include "device_math.hpp"
#include "internal.hpp"

namespace svamd {

namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;  // 4096 keys
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int WAVES = SORT_THREADS / 64;

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
constexpr int TOP_THREADS = 1024;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Lanes of the calling wave (among the active ones) whose digit equals d.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, int bits) {
    uint64_t peers = __ballot(1);
    // Nearly sorted keys: usually every active lane holds the same digit.
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    if (__ballot(d == d0) == peers) return peers;
    for (int b = 0; b < bits; ++b) {
        uint32_t bit = (d >> b) & 1u;
        uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread across the work-group; returns the
// exclusive prefix and writes the group total to *total (all threads).
template <int THREADS>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t *wave_sums, uint32_t &total) {
    constexpr int NW = THREADS / 64;
    uint32_t lane = lane_id();
    uint32_t wave = threadIdx.x / 64;
    uint32_t incl = wave_inclusive_scan(x);
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t w = threadIdx.x < (uint32_t)NW ? wave_sums[threadIdx.x] : 0u;
        uint32_t wi = wave_inclusive_scan(w);
        if (threadIdx.x < (uint32_t)NW) wave_sums[threadIdx.x] = wi - w;
        if (threadIdx.x == (uint32_t)NW - 1) wave_sums[NW] = wi;
    }
    __syncthreads();
    uint32_t r = wave_sums[wave] + incl - x;
    total = wave_sums[NW];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------ scans
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t *__restrict__ in, uint64_t m,
                                                               uint32_t *__restrict__ sums) {
    __shared__ uint32_t ws[SCAN_THREADS / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        uint64_t i = base + (uint64_t)k * SCAN_THREADS + threadIdx.x;
        if (i < m) s += in[i];
    }
    uint32_t total;
    block_exclusive_scan<SCAN_THREADS>(s, ws, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of sums[0, nt) in place by one work-group; *total = sum.
__global__ __launch_bounds__(TOP_THREADS) void k_scan_top(uint32_t *__restrict__ sums, uint32_t nt,
                                                           uint32_t *__restrict__ total) {
    __shared__ uint32_t ws[TOP_THREADS / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nt; base += TOP_THREADS) {
        uint32_t i = base + threadIdx.x;
        uint32_t x = i < nt ? sums[i] : 0u;
        uint32_t t;
        uint32_t ex = block_exclusive_scan<TOP_THREADS>(x, ws, t);
        if (i < nt) sums[i] = carry + ex;
        carry += t;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const uint32_t *__restrict__ in,
                                                             uint32_t *__restrict__ out, uint64_t m,
                                                             const uint32_t *__restrict__ sums) {
    __shared__ uint32_t ws[SCAN_THREADS / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        v[k] = i < m ? in[i] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<SCAN_THREADS>(s, ws, total) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        if (i < m) out[i] = run;
        run += v[k];
    }
}

// ------------------------------------------------------------- radix sort
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint32_t *__restrict__ keys, uint64_t n, int shift,
                                                              int bits, uint32_t *__restrict__ spine, uint32_t T) {
    __shared__ uint32_t h[RADIX];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
    const uint32_t mask = (1u << bits) - 1u;
    // All of the tile's loads are issued before the first is used.
    uint32_t kreg[SORT_ITEMS];
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        const uint64_t i = base + (uint64_t)k * SORT_THREADS + threadIdx.x;
        kreg[k] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        uint64_t i = base + (uint64_t)k * SORT_THREADS + threadIdx.x;
        if (i < n) {
            uint32_t d = (kreg[k] >> shift) & mask;
            uint64_t peers = match_digit(d, bits);
            uint32_t leader = (uint32_t)__ffsll((unsigned long long)peers) - 1u;
            if (lane_id() == leader) atomicAdd(&h[d], (uint32_t)__popcll(peers));
        }
    }
    __syncthreads();
    spine[(uint64_t)threadIdx.x * T + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(const uint32_t *__restrict__ keys_in,
                                                                 const uint32_t *__restrict__ vals_in,
                                                                 uint32_t *__restrict__ keys_out,
                                                                 uint32_t *__restrict__ vals_out, uint64_t n,
                                                                 int shift, int bits,
                                                                 const uint32_t *__restrict__ offsets, uint32_t T) {
    __shared__ uint32_t digit_base[RADIX];
    __shared__ uint32_t wave_cnt[WAVES][RADIX];
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t lane = lane_id();
    const uint32_t mask = (1u << bits) - 1u;
    digit_base[threadIdx.x] = offsets[(uint64_t)threadIdx.x * T + blockIdx.x];
    uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
    // All of the tile's loads are issued before the first round.
    uint32_t kreg[SORT_ITEMS], vreg[SORT_ITEMS];
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        const uint64_t i = base + (uint64_t)k * SORT_THREADS + threadIdx.x;
        kreg[k] = i < n ? keys_in[i] : 0u;
        vreg[k] = i < n ? (vals_in ? vals_in[i] : (uint32_t)i) : 0u;
    }
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) wave_cnt[w][threadIdx.x] = 0;
        __syncthreads();
        uint64_t i = base + (uint64_t)k * SORT_THREADS + threadIdx.x;
        bool valid = i < n;
        uint32_t key = 0, val = 0, d = 0, rank = 0;
        if (valid) {
            key = kreg[k];
            val = vreg[k];
            d = (key >> shift) & mask;
            uint64_t peers = match_digit(d, bits);
            rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
            uint32_t leader = (uint32_t)__ffsll((unsigned long long)peers) - 1u;
            if (lane == leader) wave_cnt[wave][d] = (uint32_t)__popcll(peers);
        }
        __syncthreads();
        {
            // Per digit: the waves' counts become their output offsets.
            uint32_t run = digit_base[threadIdx.x];
#pragma unroll
            for (int w = 0; w < WAVES; ++w) {
                uint32_t c = wave_cnt[w][threadIdx.x];
                wave_cnt[w][threadIdx.x] = run;
                run += c;
            }
            digit_base[threadIdx.x] = run;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = wave_cnt[wave][d] + rank;
            if (pos < n) {  // always true for a consistent spine
                keys_out[pos] = key;
                vals_out[pos] = val;
            }
        }
        __syncthreads();
    }
}

__global__ void k_identity_copy(const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
                                uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        keys_out[i] = keys_in[i];
        vals_out[i] = vals_in ? vals_in[i] : (uint32_t)i;
    }
}

// ------------------------------------------------------------------ keys
// Block key of a particle: Morton code of the 4^3 block holding its base cell
// floor(x/h - 1/2) (position_to_shift_quadratic, rust/crates/cpu/src/kernels.rs:46-49).
__global__ __launch_bounds__(256) void k_keys(ParticleArrays P, GridGeom G, uint64_t n, uint32_t *__restrict__ keys,
                                              DeviceScalars *ds) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t flags = P.flags(i);
    if (flags & F_TOMB) {
        keys[i] = G.tomb_key;
        return;
    }
    const float x[3] = {P.x(0, i), P.x(1, i), P.x(2, i)};
    keys[i] = particle_block_key(x, G, ds);
}

// ------------------------------------------------------------ block list
// Tiles of 4096 sorted keys read in 16 coalesced rounds of 256.  A head is
// the first key of a run of equal valid keys; the heads, in key order, are
// the occupied blocks.
__device__ __forceinline__ bool is_head(const uint32_t *__restrict__ keys, uint64_t i, uint64_t n, uint32_t tomb,
                                        uint32_t &key) {
    if (i >= n) return false;
    key = keys[i];
    return key != tomb && (i == 0 || keys[i - 1] != key);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_heads_count(const uint32_t *__restrict__ keys, uint64_t n,
                                                               uint32_t tomb, uint32_t *__restrict__ tile_counts,
                                                               DeviceScalars *ds) {
    __shared__ uint32_t ws[SCAN_THREADS / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint64_t i = base + (uint64_t)k * SCAN_THREADS;
        uint32_t key = tomb;
        cnt += is_head(keys, i, n, tomb, key) ? 1u : 0u;
        if (i < n && key != tomb && (i + 1 == n || keys[i + 1] == tomb)) ds->live_count = (uint32_t)(i + 1);
    }
    uint32_t total;
    block_exclusive_scan<SCAN_THREADS>(cnt, ws, total);
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_heads_write(const uint32_t *__restrict__ keys, uint64_t n,
                                                               uint32_t tomb, const uint32_t *__restrict__ tile_offsets,
                                                               uint32_t *__restrict__ blk_start,
                                                               uint32_t *__restrict__ blk_key, uint32_t capacity,
                                                               DeviceScalars *ds) {
    __shared__ uint32_t ws[SCAN_THREADS / 64];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
    uint32_t running = tile_offsets[blockIdx.x];
    // The head tests (and their loads) of all rounds come first.
    uint32_t kreg[SCAN_ITEMS];
    bool hreg[SCAN_ITEMS];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        kreg[k] = tomb;
        hreg[k] = is_head(keys, base + (uint64_t)k * SCAN_THREADS, n, tomb, kreg[k]);
    }
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint64_t i = base + (uint64_t)k * SCAN_THREADS;
        const uint32_t key = kreg[k];
        const bool head = hreg[k];
        const uint64_t m = __ballot(head);
        if (lane == 0) ws[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, round = 0;
#pragma unroll
        for (int w = 0; w < SCAN_THREADS / 64; ++w) {
            before += (uint32_t)w < wave ? ws[w] : 0u;
            round += ws[w];
        }
        if (head) {
            const uint32_t slot = running + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (slot < capacity) {
                blk_start[slot] = (uint32_t)i;
                blk_key[slot] = key;
            } else {
                atomicOr(&ds->status, MPM_STATUS_INDIRECT_LIMIT_EXCEEDED | (KID_BLOCKS << 16));
            }
        }
        running += round;
        __syncthreads();
    }
}

// ------------------------------------------------------------ bucket sort
// A counting sort of block keys over all 2^key_bits keys, for the re-sort of
// keys that a transfer wrote.  The input slots are grouped by block in key
// order, so the particles that stayed in their block (no BUCKET_MOVER bit)
// come in key order: each takes the slot base of its key plus its index among
// all stayers, which keeps their input order.  The movers take the slots
// after their key's stayers, a run of equal keys in a wave with one atomic.
// Five launches:
//   k_bucket_count   counts[key] and stay[key] += run lengths; the stayers of
//                    each span of BUCKET_SPAN input slots;
//   k_bucket_reduce  per tile of BUCKET_TILE keys: its particles, occupied
//                    blocks (the tomb key excluded) and stayers; per tile of
//                    BUCKET_TILE spans: their stayers;
//   k_bucket_top     the tiles' exclusive offsets (four sums), and the number
//                    of occupied blocks;
//   k_bucket_down    per key: the stayers' slot base, the movers' first slot,
//                    the block list (start, key) in key order and the live
//                    count (the tomb key is the largest: culled particles go
//                    last), the counts cleared for the next sort; per span:
//                    the stayers before it;
//   k_bucket_place   order[slot] = input slot.
// A span's slot base + j BUCKET_GROUP + thread is its item j (j < BUCKET_ITEMS).
constexpr int BUCKET_GROUP = 256;
constexpr int BUCKET_ITEMS = 8;
constexpr int BUCKET_SPAN = BUCKET_GROUP * BUCKET_ITEMS;
constexpr int BUCKET_TILE_ITEMS = 4;  // consecutive keys per thread in the key scans
constexpr int BUCKET_TILE = SCAN_THREADS * BUCKET_TILE_ITEMS;

// The runs of equal keys among the lanes of the wave that take part: f(leader
// lane, key, run mask) once per run.
template <class F>
__device__ __forceinline__ void for_each_run(bool part, uint32_t key, F f) {
    uint64_t todo = __ballot(part);
    while (todo) {
        const int l = __ffsll((unsigned long long)todo) - 1;
        const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)key, l);
        const uint64_t m = __ballot(part && key == k);
        f((uint32_t)l, k, m);
        todo &= ~m;
    }
}

__device__ __forceinline__ uint64_t lanes_below(uint32_t lane) { return (1ull << lane) - 1ull; }

__global__ __launch_bounds__(BUCKET_GROUP) void k_bucket_count(const uint32_t *__restrict__ keys, uint64_t n,
                                                               uint32_t mask, uint32_t *__restrict__ counts,
                                                               uint32_t *__restrict__ stay,
                                                               uint32_t *__restrict__ span_stay) {
    __shared__ uint32_t ws[BUCKET_GROUP / 64];
    const uint64_t base = (uint64_t)blockIdx.x * BUCKET_SPAN + threadIdx.x;
    const uint32_t lane = lane_id();
    uint32_t raw[BUCKET_ITEMS];
#pragma unroll
    for (int j = 0; j < BUCKET_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * BUCKET_GROUP;
        raw[j] = i < n ? keys[i] : 0u;
    }
    uint32_t stayers = 0;
#pragma unroll
    for (int j = 0; j < BUCKET_ITEMS; ++j) {
        const bool valid = base + (uint64_t)j * BUCKET_GROUP < n;
        const uint32_t key = raw[j] & mask;
        const bool stayer = valid && !(raw[j] & BUCKET_MOVER);
        for_each_run(valid, key, [&](uint32_t l, uint32_t k, uint64_t m) {
            if (lane == l) atomicAdd(&counts[k], (uint32_t)__popcll(m));
        });
        for_each_run(stayer, key, [&](uint32_t l, uint32_t k, uint64_t m) {
            if (lane == l) atomicAdd(&stay[k], (uint32_t)__popcll(m));
        });
        stayers += (uint32_t)__popcll(__ballot(stayer));
    }
    if (lane == 0) ws[threadIdx.x / 64] = stayers;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < BUCKET_GROUP / 64; ++w) t += ws[w];
        span_stay[blockIdx.x] = t;
    }
}

// Tiles [0, nt) over the keys, tiles [nt, nt + nst) over the spans.
__global__ __launch_bounds__(SCAN_THREADS) void k_bucket_reduce(const uint32_t *__restrict__ counts,
                                                                 const uint32_t *__restrict__ stay, uint32_t nkeys,
                                                                 uint32_t tomb, uint32_t nt,
                                                                 const uint32_t *__restrict__ span_stay,
                                                                 uint32_t nspans, uint32_t *__restrict__ tiles) {
    __shared__ uint32_t ws[SCAN_THREADS / 64 + 1];
    uint32_t total;
    if (blockIdx.x < nt) {
        const uint32_t base = blockIdx.x * (uint32_t)BUCKET_TILE + threadIdx.x;
        uint32_t parts = 0, blocks = 0, stayers = 0;
#pragma unroll
        for (int k = 0; k < BUCKET_TILE_ITEMS; ++k) {
            const uint32_t b = base + (uint32_t)(k * SCAN_THREADS);
            const uint32_t c = b < nkeys ? counts[b] : 0u;
            parts += c;
            blocks += (c != 0u && b != tomb) ? 1u : 0u;
            stayers += b < nkeys ? stay[b] : 0u;
        }
        block_exclusive_scan<SCAN_THREADS>(parts, ws, total);
        if (threadIdx.x == 0) tiles[blockIdx.x] = total;
        block_exclusive_scan<SCAN_THREADS>(blocks, ws, total);
        if (threadIdx.x == 0) tiles[nt + blockIdx.x] = total;
        block_exclusive_scan<SCAN_THREADS>(stayers, ws, total);
        if (threadIdx.x == 0) tiles[2 * nt + blockIdx.x] = total;
    } else {
        const uint32_t t = blockIdx.x - nt;
        const uint32_t base = t * (uint32_t)BUCKET_TILE + threadIdx.x;
        uint32_t stayers = 0;
#pragma unroll
        for (int k = 0; k < BUCKET_TILE_ITEMS; ++k) {
            const uint32_t g = base + (uint32_t)(k * SCAN_THREADS);
            stayers += g < nspans ? span_stay[g] : 0u;
        }
        block_exclusive_scan<SCAN_THREADS>(stayers, ws, total);
        if (threadIdx.x == 0) tiles[3 * nt + t] = total;
    }
}

// Exclusive scans, in place, of the four tile sums: particles, blocks and
// stayers over nt key tiles, stayers over nst span tiles.
__global__ __launch_bounds__(TOP_THREADS) void k_bucket_top(uint32_t *__restrict__ tiles, uint32_t nt, uint32_t nst,
                                                            DeviceScalars *ds) {
    __shared__ uint32_t ws[TOP_THREADS / 64 + 1];
    for (int a = 0; a < 4; ++a) {
        uint32_t *sums = tiles + (uint32_t)a * nt;
        const uint32_t len = a < 3 ? nt : nst;
        uint32_t carry = 0;
        for (uint32_t base = 0; base < len; base += TOP_THREADS) {
            const uint32_t i = base + threadIdx.x;
            const uint32_t x = i < len ? sums[i] : 0u;
            uint32_t t;
            const uint32_t ex = block_exclusive_scan<TOP_THREADS>(x, ws, t);
            if (i < len) sums[i] = carry + ex;
            carry += t;
        }
        if (a == 1 && threadIdx.x == 0) ds->num_blocks = carry;
    }
}

// Key tiles: thread t takes keys 4t .. 4t + 3 of its tile (nkeys is a power
// of two of at least 2, and a tile's keys either all exist or start past the
// end).
__global__ __launch_bounds__(SCAN_THREADS) void k_bucket_down(
    uint32_t *__restrict__ counts, uint32_t *__restrict__ stay, uint32_t nkeys, uint32_t tomb, uint32_t nt,
    uint32_t *__restrict__ span_stay, uint32_t nspans, const uint32_t *__restrict__ tiles,
    uint32_t *__restrict__ cursor, uint32_t *__restrict__ stay_base, uint32_t *__restrict__ blk_start,
    uint32_t *__restrict__ blk_key, uint32_t capacity, DeviceScalars *ds) {
    __shared__ uint32_t ws[SCAN_THREADS / 64 + 1];
    uint32_t total;
    if (blockIdx.x >= nt) {
        // span tile: the stayers before each span
        const uint32_t t = blockIdx.x - nt;
        const uint32_t base = t * (uint32_t)BUCKET_TILE + threadIdx.x * (uint32_t)BUCKET_TILE_ITEMS;
        uint32_t v[BUCKET_TILE_ITEMS], sum = 0;
#pragma unroll
        for (int k = 0; k < BUCKET_TILE_ITEMS; ++k) {
            const uint32_t g = base + (uint32_t)k;
            v[k] = g < nspans ? span_stay[g] : 0u;
            sum += v[k];
        }
        uint32_t run = block_exclusive_scan<SCAN_THREADS>(sum, ws, total) + tiles[3 * nt + t];
#pragma unroll
        for (int k = 0; k < BUCKET_TILE_ITEMS; ++k) {
            const uint32_t g = base + (uint32_t)k;
            if (g < nspans) span_stay[g] = run;
            run += v[k];
        }
        return;
    }
    static_assert(BUCKET_TILE_ITEMS == 4, "one uint4 per thread");
    const uint32_t base = blockIdx.x * (uint32_t)BUCKET_TILE + threadIdx.x * (uint32_t)BUCKET_TILE_ITEMS;
    const bool here = base < nkeys;
    uint4 c4 = make_uint4(0u, 0u, 0u, 0u), s4 = make_uint4(0u, 0u, 0u, 0u);
    if (here) {
        c4 = *reinterpret_cast<const uint4 *>(counts + base);
        s4 = *reinterpret_cast<const uint4 *>(stay + base);
    }
    const uint32_t c[4] = {c4.x, c4.y, c4.z, c4.w}, st[4] = {s4.x, s4.y, s4.z, s4.w};
    uint32_t parts = 0, blocks = 0, stayers = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        parts += c[k];
        blocks += (c[k] != 0u && base + (uint32_t)k != tomb) ? 1u : 0u;
        stayers += st[k];
    }
    uint32_t rp = block_exclusive_scan<SCAN_THREADS>(parts, ws, total) + tiles[blockIdx.x];
    uint32_t rb = block_exclusive_scan<SCAN_THREADS>(blocks, ws, total) + tiles[nt + blockIdx.x];
    uint32_t rs = block_exclusive_scan<SCAN_THREADS>(stayers, ws, total) + tiles[2 * nt + blockIdx.x];
    if (!here) return;
    uint32_t sb[4], cu[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t b = base + (uint32_t)k;
        // slots of key b: [rp, rp + st) its stayers, then its movers
        sb[k] = rp - rs;
        cu[k] = rp + st[k];
        if (b == tomb) {
            ds->live_count = rp;
        } else if (c[k] != 0u) {
            if (rb < capacity) {
                blk_start[rb] = rp;
                blk_key[rb] = b;
            } else {
                atomicOr(&ds->status, MPM_STATUS_INDIRECT_LIMIT_EXCEEDED | (KID_BLOCKS << 16));
            }
            ++rb;
        }
        rp += c[k];
        rs += st[k];
    }
    *reinterpret_cast<uint4 *>(stay_base + base) = make_uint4(sb[0], sb[1], sb[2], sb[3]);
    *reinterpret_cast<uint4 *>(cursor + base) = make_uint4(cu[0], cu[1], cu[2], cu[3]);
    if (c4.x | c4.y | c4.z | c4.w) *reinterpret_cast<uint4 *>(counts + base) = make_uint4(0u, 0u, 0u, 0u);
    if (s4.x | s4.y | s4.z | s4.w) *reinterpret_cast<uint4 *>(stay + base) = make_uint4(0u, 0u, 0u, 0u);
}

__global__ __launch_bounds__(BUCKET_GROUP) void k_bucket_place(const uint32_t *__restrict__ keys, uint64_t n,
                                                               uint32_t mask, const uint32_t *__restrict__ stay_base,
                                                               const uint32_t *__restrict__ span_stay_before,
                                                               uint32_t *__restrict__ cursor,
                                                               uint32_t *__restrict__ order) {
    constexpr int NW = BUCKET_GROUP / 64;
    __shared__ uint32_t ws[BUCKET_ITEMS * NW];
    const uint64_t base = (uint64_t)blockIdx.x * BUCKET_SPAN + threadIdx.x;
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
    uint32_t raw[BUCKET_ITEMS];
    uint64_t sm[BUCKET_ITEMS];
#pragma unroll
    for (int j = 0; j < BUCKET_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * BUCKET_GROUP;
        raw[j] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (int j = 0; j < BUCKET_ITEMS; ++j) {
        const bool stayer = base + (uint64_t)j * BUCKET_GROUP < n && !(raw[j] & BUCKET_MOVER);
        sm[j] = __ballot(stayer);
        if (lane == 0) ws[j * NW + (int)wave] = (uint32_t)__popcll(sm[j]);
    }
    __syncthreads();
    // stayers of the span before (item j, this wave), in slot order
    uint32_t before = span_stay_before[blockIdx.x];
    for (uint32_t q = 0; q < wave; ++q) before += ws[q];
#pragma unroll
    for (int j = 0; j < BUCKET_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * BUCKET_GROUP;
        const bool valid = i < n;
        const uint32_t key = raw[j] & mask;
        const bool stayer = (sm[j] >> lane) & 1ull;
        uint64_t slot = n;
        if (stayer) slot = (uint64_t)stay_base[key] + before + (uint32_t)__popcll(sm[j] & lanes_below(lane));
        for_each_run(valid && !stayer, key, [&](uint32_t l, uint32_t k, uint64_t m) {
            uint32_t first = 0;
            if (lane == l) first = atomicAdd(&cursor[k], (uint32_t)__popcll(m));
            first = (uint32_t)__builtin_amdgcn_readlane((int)first, (int)l);
            if ((m >> lane) & 1ull) slot = (uint64_t)first + (uint32_t)__popcll(m & lanes_below(lane));
        });
        if (slot < n) order[slot] = (uint32_t)i;
        // the stayers of item j after this wave, and of item j + 1 before it
        if (j + 1 < BUCKET_ITEMS) {
            for (uint32_t q = wave; q < (uint32_t)NW; ++q) before += ws[j * NW + (int)q];
            for (uint32_t q = 0; q < wave; ++q) before += ws[(j + 1) * NW + (int)q];
        }
    }
}

__global__ void k_copy_bytes(const float4 *__restrict__ src, float4 *__restrict__ dst, uint64_t n4) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_reset_scalars(DeviceScalars *ds) {
    if (threadIdx.x == 0) {
        ds->num_blocks = 0;
        ds->live_count = 0;
        ds->active_nodes = 0;
    }
}

inline uint32_t tiles_of(uint64_t n, int tile) { return (uint32_t)((n + tile - 1) / tile); }

}  // namespace

size_t scan_temp_bytes(uint64_t n) {
    size_t nt = tiles_of(n, SCAN_TILE);
    return std::max<size_t>(nt, 1) * sizeof(uint32_t);
}

hipError_t exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t *total, uint64_t n, void *temp,
                          hipStream_t stream) {
    uint32_t nt = tiles_of(n, SCAN_TILE);
    uint32_t *sums = static_cast<uint32_t *>(temp);
    if (n == 0) {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(TOP_THREADS), 0, stream, sums, 0u, total);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(SCAN_THREADS), 0, stream, in, n, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(TOP_THREADS), 0, stream, sums, nt, total);
    hipLaunchKernelGGL(k_scan_down, dim3(nt), dim3(SCAN_THREADS), 0, stream, in, out, n, sums);
    return hipGetLastError();
}

// Layout of the radix-sort temp buffer: keys ping-pong, values ping-pong,
// the digit-major spine and the spine scan's tile sums.
size_t radix_sort_temp_bytes(uint64_t n) {
    size_t T = std::max<uint32_t>(tiles_of(n, SORT_TILE), 1u);
    size_t spine = (size_t)RADIX * T;
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return align(n * 4) * 2 + align(spine * 4) + align(scan_temp_bytes(spine));
}

hipError_t radix_sort(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out, uint32_t *vals_out,
                      uint32_t *keys_tmp, uint32_t *vals_tmp, uint64_t n, int key_bits, void *temp,
                      hipStream_t stream) {
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    uint32_t T = std::max<uint32_t>(tiles_of(n, SORT_TILE), 1u);
    size_t spine_len = (size_t)RADIX * T;
    char *t = static_cast<char *>(temp);
    if (!keys_tmp) keys_tmp = reinterpret_cast<uint32_t *>(t);
    t += align(n * 4);
    if (!vals_tmp) vals_tmp = reinterpret_cast<uint32_t *>(t);
    t += align(n * 4);
    uint32_t *spine = reinterpret_cast<uint32_t *>(t);
    t += align(spine_len * 4);
    void *scan_tmp = t;

    int passes = (key_bits + RADIX_BITS - 1) / RADIX_BITS;
    if (n == 0) return hipSuccess;
    if (passes <= 0) {
        hipLaunchKernelGGL(k_identity_copy, dim3(std::min<uint64_t>((n + 255) / 256, 4096)), dim3(256), 0, stream,
                           keys_in, vals_in, keys_out, vals_out, n);
        return hipGetLastError();
    }
    const uint32_t *ksrc = keys_in;
    const uint32_t *vsrc = vals_in;
    for (int p = 0; p < passes; ++p) {
        bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t *kdst = to_out ? keys_out : keys_tmp;
        uint32_t *vdst = to_out ? vals_out : vals_tmp;
        int shift = p * RADIX_BITS;
        int bits = std::min(RADIX_BITS, key_bits - shift);
        hipLaunchKernelGGL(k_radix_hist, dim3(T), dim3(SORT_THREADS), 0, stream, ksrc, n, shift, bits, spine, T);
        hipError_t e = exclusive_scan(spine, spine, nullptr, spine_len, scan_tmp, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(T), dim3(SORT_THREADS), 0, stream, ksrc, vsrc, kdst, vdst, n, shift,
                           bits, spine, T);
        ksrc = kdst;
        vsrc = vdst;
    }
    return hipGetLastError();
}

hipError_t launch_keys(const ParticleArrays &P, const GridGeom &G, uint64_t n, uint32_t *keys, DeviceScalars *ds,
                       hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, P, G, n, keys, ds);
    return hipGetLastError();
}

size_t block_list_temp_bytes(uint64_t n) { return (size_t)std::max<uint32_t>(tiles_of(n, SCAN_TILE), 1u) * 4; }

hipError_t launch_block_list(const uint32_t *keys, uint64_t n, uint32_t tomb_key, uint32_t *blk_start,
                             uint32_t *blk_key, uint32_t capacity, DeviceScalars *ds, void *temp, hipStream_t stream) {
    uint32_t nt = tiles_of(n, SCAN_TILE);
    uint32_t *tile = static_cast<uint32_t *>(temp);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_heads_count, dim3(nt), dim3(SCAN_THREADS), 0, stream, keys, n, tomb_key, tile, ds);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(TOP_THREADS), 0, stream, tile, nt, &ds->num_blocks);
    hipLaunchKernelGGL(k_heads_write, dim3(nt), dim3(SCAN_THREADS), 0, stream, keys, n, tomb_key, tile, blk_start,
                       blk_key, capacity, ds);
    return hipGetLastError();
}

uint64_t bucket_span_count(uint64_t n) { return std::max<uint64_t>((n + BUCKET_SPAN - 1) / BUCKET_SPAN, 1); }
size_t bucket_tile_sums(int key_bits, uint64_t n) {
    return 3 * (size_t)tiles_of(1ull << key_bits, BUCKET_TILE) + tiles_of(bucket_span_count(n), BUCKET_TILE);
}

hipError_t bucket_sort(const uint32_t *keys, uint64_t n, int key_bits, const BucketBuffers &B, uint32_t *order,
                       uint32_t *blk_start, uint32_t *blk_key, uint32_t capacity, DeviceScalars *ds,
                       hipStream_t stream) {
    if (key_bits < 2 || key_bits > BUCKET_SORT_MAX_BITS) return hipErrorInvalidValue;
    const uint32_t nkeys = 1u << key_bits, tomb = nkeys - 1u;
    const uint32_t nt = tiles_of(nkeys, BUCKET_TILE);
    const uint32_t spans = (uint32_t)((n + BUCKET_SPAN - 1) / BUCKET_SPAN);
    const uint32_t nst = tiles_of(spans, BUCKET_TILE);
    if (n > 0)
        hipLaunchKernelGGL(k_bucket_count, dim3(spans), dim3(BUCKET_GROUP), 0, stream, keys, n, tomb, B.counts,
                           B.stay, B.span_stay);
    hipLaunchKernelGGL(k_bucket_reduce, dim3(nt + nst), dim3(SCAN_THREADS), 0, stream, B.counts, B.stay, nkeys, tomb,
                       nt, B.span_stay, spans, B.tiles);
    hipLaunchKernelGGL(k_bucket_top, dim3(1), dim3(TOP_THREADS), 0, stream, B.tiles, nt, nst, ds);
    hipLaunchKernelGGL(k_bucket_down, dim3(nt + nst), dim3(SCAN_THREADS), 0, stream, B.counts, B.stay, nkeys, tomb,
                       nt, B.span_stay, spans, B.tiles, B.cursor, B.stay_base, blk_start, blk_key, capacity, ds);
    if (n > 0)
        hipLaunchKernelGGL(k_bucket_place, dim3(spans), dim3(BUCKET_GROUP), 0, stream, keys, n, tomb, B.stay_base,
                           B.span_stay, B.cursor, order);
    return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t stream) {
    uint64_t n4 = bytes / 16;
    if (n4 == 0) return hipSuccess;
    uint32_t grid = (uint32_t)std::min<uint64_t>((n4 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy_bytes, dim3(grid), dim3(256), 0, stream, static_cast<const float4 *>(src),
                       static_cast<float4 *>(dst), n4);
    return hipGetLastError();
}

hipError_t launch_reset_scalars(DeviceScalars *ds, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(64), 0, stream, ds);
    return hipGetLastError();
}

}  // namespace svamd

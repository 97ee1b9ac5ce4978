// Device runtime for generated DogQC pipelines on MI355X (gfx950, CDNA4).
//
// Replaces the reference's CUDA runtime headers dogqc/include/hashing.cuh and
// dogqc/include/util.cuh with primitives designed for 64-lane wavefronts:
//   * 64-bit ballots and mbcnt lane prefixes instead of 32-lane masks
//     (util.cuh:3, codegen.py:82-114);
//   * keyed hash tables with lock-free CAS on a 64-bit tag word instead of a
//     CAS on a hash followed by an unsynchronised payload store
//     (hashing.cuh:81-119) or a OnceLock spin (hashing.cuh:206-272);
//   * a multimap (count -> exclusive scan -> scatter) with a deterministic
//     device-wide scan instead of per-warp atomic range allocation
//     (hashing.cuh:139-199);
//   * per-workgroup LDS aggregation tables (per-lane private accumulators or
//     LDS atomics) flushed once per workgroup into the global table, instead
//     of one global atomic per row and aggregate (cudaTranslator.py:466-493);
//   * SQL string helpers (util.cuh:61-199).
//
// Tag words: 0 = empty slot, 1 = slot being claimed (hashed mode only),
// otherwise bit 63 is set.  In "exact" mode the remaining 63 bits hold the
// packed key itself, so a tag match is a key match and no payload compare or
// lock is needed.  In "hashed" mode they hold a 63-bit hash and the keys live
// in payload columns.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef DQC_DEVICE_INLINE
#define DQC_DEVICE_INLINE __device__ __forceinline__
#endif

namespace dqc {

constexpr int WAVE = 64;
constexpr uint64_t TAG_EMPTY = 0ull;
constexpr uint64_t TAG_LOCKED = 1ull;
constexpr uint64_t TAG_TOP = 1ull << 63;

// status word bits (first word of every plan workspace)
constexpr uint32_t ST_HT_OVERFLOW = 1u << 0;    // some keyed table ran full
constexpr uint32_t ST_MM_OVERFLOW = 1u << 1;    // multimap entry array too small
constexpr uint32_t ST_RESULT_OVERFLOW = 1u << 2;
constexpr uint32_t ST_STRING_LITERAL = 1u << 3;

// ---------------------------------------------------------------- wave64

DQC_DEVICE_INLINE uint32_t lane_id() { return __lane_id(); }

DQC_DEVICE_INLINE uint64_t ballot(bool p) { return (uint64_t)__ballot((int)p); }

// number of set bits of m in lanes strictly below this lane
DQC_DEVICE_INLINE uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

DQC_DEVICE_INLINE uint32_t popc64(uint64_t m) { return (uint32_t)__popcll(m); }

DQC_DEVICE_INLINE int first_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }

// read `v` from lane `src` (any trivially copyable type, moved as 32-bit words)
template <typename T>
DQC_DEVICE_INLINE T shfl(T v, int src) {
    constexpr int N = (int)((sizeof(T) + 3) / 4);
    uint32_t w[N];
    for (int i = 0; i < N; ++i) w[i] = 0u;
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = (uint32_t)__shfl((int)w[i], src, WAVE);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

// Forward permute: this lane's value goes to lane `dst`, and the value that
// some lane sent to this one is returned (ds_permute_b32 per 32-bit word, no
// LDS memory).  Every lane of the wavefront takes part, with distinct `dst`.
template <typename T>
DQC_DEVICE_INLINE T permute_to(T v, uint32_t dst) {
    constexpr int N = (int)((sizeof(T) + 3) / 4);
    uint32_t w[N];
    for (int i = 0; i < N; ++i) w[i] = 0u;
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = (uint32_t)__builtin_amdgcn_ds_permute((int)(dst * 4u), (int)w[i]);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

// a value known to be equal in all lanes, moved to scalar registers
DQC_DEVICE_INLINE uint64_t uniform64(uint64_t v) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// inclusive prefix sum over the 64 lanes (all lanes must participate)
DQC_DEVICE_INLINE uint64_t wave_incl_scan(uint64_t v) {
    const int l = (int)lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint64_t o = __shfl_up(v, (unsigned)d, WAVE);
        if (l >= d) v += o;
    }
    return v;
}

template <typename T>
DQC_DEVICE_INLINE T wave_sum(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

template <typename T>
DQC_DEVICE_INLINE T wave_min(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) { T o = __shfl_xor(v, d, WAVE); v = o < v ? o : v; }
    return v;
}

template <typename T>
DQC_DEVICE_INLINE T wave_max(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) { T o = __shfl_xor(v, d, WAVE); v = o > v ? o : v; }
    return v;
}

// ---------------------------------------------------------------- hashing
// 64-bit finaliser of MurmurHash3 (Austin Appleby, public domain).
DQC_DEVICE_INLINE uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// chain one more key column into a running hash
DQC_DEVICE_INLINE uint64_t hash_combine(uint64_t h, uint64_t k) {
    return mix64(h ^ (k + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2)));
}

// map a hash onto [0, cap) with a multiply-shift on its upper half
// (D. Lemire, "A fast alternative to the modulo reduction", 2016)
DQC_DEVICE_INLINE uint32_t slot_of(uint64_t h, uint32_t cap) {
    return (uint32_t)(((h >> 32) * (uint64_t)cap) >> 32);
}

DQC_DEVICE_INLINE uint32_t next_slot(uint32_t s, uint32_t cap) { return (s + 1u == cap) ? 0u : s + 1u; }

// ------------------------------------------------------------ status word
DQC_DEVICE_INLINE void raise_status(uint32_t* status, uint32_t bit) { atomicOr(status, bit); }

// -------------------------------------------------------- keyed hash tables
// Global tables: tags[cap] (uint64) + structure-of-arrays payload columns.

// Probe sequences are bounded: a key is always within MAX_PROBE slots of its
// home slot, so a build or group claim that walks farther reports the table
// as too full (the run is repeated with a larger table), lookups stop there,
// and an overflowing run never degenerates into a scan of the whole table
// per row.  Claims also give up early once some row has flagged an
// overflow in the status word (the run's result is discarded anyway).
constexpr uint32_t MAX_PROBE = 2048;

DQC_DEVICE_INLINE bool overflow_raised(const uint32_t* status) {
    return (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ST_HT_OVERFLOW) != 0u;
}

// Unique join tables come in two tag widths.  64-bit tags: bit 63 set plus
// the packed key (exact) or a hash.  32-bit tags (keys packing into 32 bits):
// the packed key + 1, so a table is half the size and more of it stays in
// L2; the one key whose tag would be 0 (all 32 bits set) lives in an extra
// slot at index cap.  In both, 0 is the empty slot.

// Insert a tag; duplicates of an existing tag take the next free slot (the
// unique-join build of the reference also stores duplicate keys,
// hashing.cuh:81-100).  Starts at slot s with n slots already tried.
// Returns the slot, or -1 when the table is too full.
// Compare-and-swap on a tag word of either width; returns the old value.
template <typename Tag>
DQC_DEVICE_INLINE Tag cas_tag(Tag* p, Tag expected, Tag desired) {
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected;
}

template <typename Tag>
DQC_DEVICE_INLINE int64_t ht_insert_from(Tag* tags, uint32_t cap, Tag tag, uint32_t s, uint32_t n,
                                         const uint32_t* status) {
    for (; n < cap && n < MAX_PROBE; ++n) {
        if ((n & 31u) == 31u && overflow_raised(status)) return -1;
        if (cas_tag(&tags[s], (Tag)TAG_EMPTY, tag) == (Tag)TAG_EMPTY) return (int64_t)s;
        s = next_slot(s, cap);
    }
    return -1;
}

DQC_DEVICE_INLINE int64_t ht_insert(uint64_t* tags, uint32_t cap, uint64_t tag, uint64_t h,
                                    const uint32_t* status) {
    return ht_insert_from<uint64_t>(tags, cap, tag, slot_of(h, cap), 0u, status);
}

// Find the next slot holding `tag`, starting at *cursor (initialise the
// cursor with probe_start).  Returns -1 at the first empty slot or after a
// full cycle.  Used by join probes, which run after the build kernel ended.
DQC_DEVICE_INLINE uint32_t probe_start(uint64_t h, uint32_t cap) { return slot_of(h, cap); }

template <typename Tag>
DQC_DEVICE_INLINE int64_t ht_probe(const Tag* tags, uint32_t cap, Tag tag, uint32_t& cursor, uint32_t& visited) {
    while (visited < cap && visited < MAX_PROBE) {
        uint32_t s = cursor;
        const Tag t = tags[s];
        cursor = next_slot(s, cap);
        ++visited;
        if (t == tag) return (int64_t)s;
        if (t == (Tag)TAG_EMPTY) { visited = cap; return -1; }
    }
    return -1;
}

// A group's accumulator combined over its R copies (copy r at a[i + r*stride];
// global group-by tables keep one copy per XCD-sized group of workgroups).
template <typename T>
DQC_DEVICE_INLINE T rep_sum(const T* a, int64_t i, uint32_t R, uint64_t stride) {
    T s = a[i];
    for (uint32_t r = 1; r < R; ++r) s += a[i + (int64_t)(r * stride)];
    return s;
}

template <typename T>
DQC_DEVICE_INLINE T rep_min(const T* a, int64_t i, uint32_t R, uint64_t stride) {
    T s = a[i];
    for (uint32_t r = 1; r < R; ++r) {
        const T v = a[i + (int64_t)(r * stride)];
        s = v < s ? v : s;
    }
    return s;
}

template <typename T>
DQC_DEVICE_INLINE T rep_max(const T* a, int64_t i, uint32_t R, uint64_t stride) {
    T s = a[i];
    for (uint32_t r = 1; r < R; ++r) {
        const T v = a[i + (int64_t)(r * stride)];
        s = v > s ? v : s;
    }
    return s;
}

// The next PROBE_WINDOW slots s, s+1, ... of a unique join table, read with
// 16-byte loads (two for 64-bit tags, one for 32-bit tags) so one memory
// round trip resolves most probes, hits and misses alike.  Join tag buffers
// carry PROBE_WINDOW padding slots (the 32-bit extra slot included) so the
// window never leaves the allocation; slots at or past cap are ignored.
constexpr uint32_t PROBE_WINDOW = 4;

template <typename Tag>
struct probe_win {
    Tag t[PROBE_WINDOW];
};

DQC_DEVICE_INLINE probe_win<uint64_t> probe_window(const uint64_t* tags, uint32_t s) {
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2), aligned(8)));
    const u64x2* q = (const u64x2*)(tags + s);
    const u64x2 a = q[0], b = q[1];
    probe_win<uint64_t> w;
    w.t[0] = a.x;
    w.t[1] = a.y;
    w.t[2] = b.x;
    w.t[3] = b.y;
    return w;
}

DQC_DEVICE_INLINE probe_win<uint32_t> probe_window(const uint32_t* tags, uint32_t s) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4), aligned(4)));
    const u32x4 a = *(const u32x4*)(tags + s);
    probe_win<uint32_t> w;
    w.t[0] = a.x;
    w.t[1] = a.y;
    w.t[2] = a.z;
    w.t[3] = a.w;
    return w;
}

// Resolve a probe from its window: the slot holding `tag` (>= 0), -1 when an
// empty slot comes first, -2 when every window slot holds another tag; then
// the probe continues with ht_probe at `cursor` with `visited` slots tried.
template <typename Tag>
DQC_DEVICE_INLINE int64_t probe_in_window(const probe_win<Tag>& w, uint32_t s, uint32_t cap, Tag tag,
                                          uint32_t& cursor, uint32_t& visited) {
    int64_t r = -2;
    uint32_t n = 0;
#pragma unroll
    for (uint32_t i = 0; i < PROBE_WINDOW; ++i) {
        if (r == -2 && s + i < cap) {
            ++n;
            if (w.t[i] == tag)
                r = (int64_t)(s + i);
            else if (w.t[i] == (Tag)TAG_EMPTY)
                r = -1;
        }
    }
    visited = n;
    cursor = s + n >= cap ? 0u : s + n;
    return r;
}

// Packed unique tables: a key packing into 32 bits with one stored column of
// at most 4 bytes keeps both in one 64-bit slot, the tag (packed key + 1) in
// the low half and the column's bits in the high half, so a probe that finds
// its key has the column in the line its window load brought (one cache miss
// per probe instead of one for the tag and one for the column).  0 is the
// empty slot; the tag-0 key lives in the extra slot at index cap with tag 1.
DQC_DEVICE_INLINE uint64_t packed_slot(uint32_t tag, uint64_t word) {
    return (word << 32) | (uint64_t)(tag != 0u ? tag : 1u);
}

template <typename T>
DQC_DEVICE_INLINE T packed_payload(uint64_t slot) {
    T v;
    const uint32_t hi = (uint32_t)(slot >> 32);
    __builtin_memcpy(&v, &hi, sizeof(T));
    return v;
}

// probe_in_window over packed slots: the low halves hold the tags
DQC_DEVICE_INLINE int64_t probe_in_window_packed(const probe_win<uint64_t>& w, uint32_t s, uint32_t cap,
                                                 uint32_t tag, uint32_t& cursor, uint32_t& visited) {
    int64_t r = -2;
    uint32_t n = 0;
#pragma unroll
    for (uint32_t i = 0; i < PROBE_WINDOW; ++i) {
        if (r == -2 && s + i < cap) {
            ++n;
            if ((uint32_t)w.t[i] == tag)
                r = (int64_t)(s + i);
            else if (w.t[i] == 0ull)
                r = -1;
        }
    }
    visited = n;
    cursor = s + n >= cap ? 0u : s + n;
    return r;
}

// ht_probe over packed slots
DQC_DEVICE_INLINE int64_t ht_probe_packed(const uint64_t* slots, uint32_t cap, uint32_t tag, uint32_t& cursor,
                                          uint32_t& visited) {
    while (visited < cap && visited < MAX_PROBE) {
        const uint32_t s = cursor;
        const uint64_t t = slots[s];
        cursor = next_slot(s, cap);
        ++visited;
        if ((uint32_t)t == tag) return (int64_t)s;
        if (t == 0ull) { visited = cap; return -1; }
    }
    return -1;
}

// Group slot in exact mode: the tag is the key, so find-or-claim is one CAS
// per probed slot with no lock.  Starts at slot s with n slots tried.
// Returns -1 when the table is full.
DQC_DEVICE_INLINE int64_t agg_slot_exact_from(uint64_t* tags, uint32_t cap, uint64_t tag, uint32_t s, uint32_t n,
                                              const uint32_t* status) {
    for (; n < cap && n < MAX_PROBE; ++n) {
        if ((n & 31u) == 31u && overflow_raised(status)) return -1;
        uint64_t t = __hip_atomic_load(&tags[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t == tag) return (int64_t)s;
        if (t == TAG_EMPTY) {
            uint64_t prev = atomicCAS((unsigned long long*)&tags[s], (unsigned long long)TAG_EMPTY,
                                      (unsigned long long)tag);
            if (prev == TAG_EMPTY || prev == tag) return (int64_t)s;
        }
        s = next_slot(s, cap);
    }
    return -1;
}

DQC_DEVICE_INLINE int64_t agg_slot_exact(uint64_t* tags, uint32_t cap, uint64_t tag, uint64_t h,
                                         const uint32_t* status) {
    return agg_slot_exact_from(tags, cap, tag, slot_of(h, cap), 0u, status);
}

// Key columns of a hashed group table are written by the lane that claims
// a slot and compared by lanes on every XCD while the same kernel runs.
// Plain stores would stay in the writer's L2, which the other XCDs do not
// see, and an agent-scope release / acquire pair writes back or invalidates
// the whole L2 (buffer_wbl2 / buffer_inv) at every row.  So the keys are
// stored and loaded as agent-scope atomics instead, which go through to the
// coherence point (sc1), and the claiming lane waits for its key stores to
// complete (vmcnt 0) before it publishes the tag.
template <typename T>
DQC_DEVICE_INLINE void coherent_store(T* p, T v) {
    if constexpr (sizeof(T) % 8 == 0) {
        uint64_t w[sizeof(T) / 8];
        __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 8); ++i)
            __hip_atomic_store((uint64_t*)p + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (sizeof(T) == 4) {
        uint32_t w;
        __builtin_memcpy(&w, &v, 4);
        __hip_atomic_store((uint32_t*)p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        static_assert(sizeof(T) == 1, "key columns are 1, 4 or 8n bytes");
        uint8_t w;
        __builtin_memcpy(&w, &v, 1);
        __hip_atomic_store((uint8_t*)p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename T>
DQC_DEVICE_INLINE T coherent_load(const T* p) {
    T v;
    if constexpr (sizeof(T) % 8 == 0) {
        uint64_t w[sizeof(T) / 8];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 8); ++i)
            w[i] = __hip_atomic_load((const uint64_t*)p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_memcpy(&v, w, sizeof(T));
    } else if constexpr (sizeof(T) == 4) {
        const uint32_t w = __hip_atomic_load((const uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_memcpy(&v, &w, 4);
    } else {
        static_assert(sizeof(T) == 1, "key columns are 1, 4 or 8n bytes");
        const uint8_t w = __hip_atomic_load((const uint8_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_memcpy(&v, &w, 1);
    }
    return v;
}

// Wait until this lane's vector memory operations have completed
// (s_waitcnt vmcnt(0), gfx9 encoding: expcnt and lgkmcnt left unwaited).
DQC_DEVICE_INLINE void wait_vector_memory() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Group slot in hashed mode.  The claiming lane CASes EMPTY->LOCKED, writes
// the key columns through `write_keys(slot)` (coherent_store), waits for
// them, then publishes the tag; other lanes read the tag and compare keys
// through `keys_equal(slot)` (coherent_load, issued only after the tag they
// depend on has been read).  All are agent-scope atomics without fences.
// The claim branch precedes the re-check in program order, so lanes of one
// wavefront waiting on a slot claimed by a sibling lane always observe the
// publication on their next loop trip.
template <typename WriteKeys, typename KeysEqual>
DQC_DEVICE_INLINE int64_t agg_slot_hashed(uint64_t* tags, uint32_t cap, uint64_t tag, uint64_t h,
                                          const uint32_t* status, WriteKeys write_keys, KeysEqual keys_equal) {
    uint32_t s = slot_of(h, cap);
    uint32_t n = 0;
    while (n < cap && n < MAX_PROBE) {
        if ((n & 31u) == 31u && overflow_raised(status)) return -1;
        uint64_t t = __hip_atomic_load(&tags[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t == TAG_EMPTY) {
            uint64_t expected = TAG_EMPTY;
            if (__hip_atomic_compare_exchange_strong(&tags[s], &expected, TAG_LOCKED, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                write_keys(s);
                wait_vector_memory();
                __hip_atomic_store(&tags[s], tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return (int64_t)s;
            }
            continue;  // lost the race: re-read this slot
        }
        if (t == TAG_LOCKED) continue;  // being published by another lane
        if (t == tag && keys_equal(s)) return (int64_t)s;
        s = next_slot(s, cap);
        ++n;
    }
    return -1;
}

// LDS variant of agg_slot_exact (workgroup-private table of `cap` slots).
// Probes at most LDS_MAX_PROBE slots: a key either has a slot within that
// window of its start or never gets one (slots are never freed), so all its
// rows agree; a full table sends new keys to the global table after a short
// probe instead of a scan of every slot.
constexpr uint32_t LDS_MAX_PROBE = 64;

DQC_DEVICE_INLINE int lds_slot_exact(uint64_t* ltags, uint32_t cap, uint64_t tag, uint64_t h) {
    uint32_t s = slot_of(h, cap);
    for (uint32_t n = 0; n < cap && n < LDS_MAX_PROBE; ++n) {
        uint64_t t = __hip_atomic_load(&ltags[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (t == tag) return (int)s;
        if (t == TAG_EMPTY) {
            uint64_t expected = TAG_EMPTY;
            if (__hip_atomic_compare_exchange_strong(&ltags[s], &expected, tag, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
                return (int)s;
            if (expected == tag) return (int)s;
        }
        s = next_slot(s, cap);
    }
    return -1;
}

// ------------------------------------------------------ join filter
// Blocked Bloom filter of a unique join's build keys: `words` 32-bit words
// (a power of two), two bits of one word per key, all from the key's hash.
// The build sets them; a probe pipeline copies the filter into LDS and skips
// the hash-table probe of rows whose bits are not both set (no match is
// possible), so only matches and false positives reach the table.
DQC_DEVICE_INLINE uint32_t bloom_word(uint64_t h, uint32_t words) { return (uint32_t)h & (words - 1u); }
DQC_DEVICE_INLINE uint32_t bloom_mask(uint64_t h) {
    return (1u << ((uint32_t)(h >> 32) & 31u)) | (1u << ((uint32_t)(h >> 40) & 31u));
}
DQC_DEVICE_INLINE void bloom_set(uint32_t* f, uint32_t words, uint64_t h) {
    atomicOr(&f[bloom_word(h, words)], bloom_mask(h));
}
DQC_DEVICE_INLINE bool bloom_test(const uint32_t* f, uint32_t words, uint64_t h) {
    const uint32_t m = bloom_mask(h);
    return (f[bloom_word(h, words)] & m) == m;
}

// ------------------------------------------------- partitioned group-by
// The first pass of a partitioned group-by gives workgroup w of `nseg` its
// own segment of every partition b: `cap` records at (b * nseg + w) * cap,
// filled through an LDS cursor, and stores the segment's record count at
// counts[b * nseg + w] at its end.  The second pass gives workgroup b
// partition b: part_segments puts the segments' exclusive prefix sums in
// pref[0..nseg] (LDS) and returns the partition's records, which the rows
// 0..total-1 of the second pass then address through part_record.

// Exclusive scan over the workgroup (blockDim.x a multiple of 64, at most
// 1024); sh holds blockDim.x / 64 + 1 words.  Every thread must call it.
DQC_DEVICE_INLINE uint32_t wg_excl_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const uint32_t w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    const uint32_t incl = (uint32_t)wave_incl_scan((uint64_t)v);
    if (lane_id() == WAVE - 1) sh[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t i = 0; i < nw; ++i) {
            const uint32_t x = sh[i];
            sh[i] = run;
            run += x;
        }
        sh[nw] = run;
    }
    __syncthreads();
    const uint32_t r = sh[w] + incl - v;
    total = sh[nw];
    __syncthreads();
    return r;
}

DQC_DEVICE_INLINE uint32_t part_segments(const uint32_t* counts, uint32_t part, uint32_t nseg, uint32_t cap,
                                         uint32_t* pref, uint32_t* sh) {
    const uint32_t k = (nseg + blockDim.x - 1) / blockDim.x;
    const uint32_t first = threadIdx.x * k;
    const uint32_t* c = counts + (uint64_t)part * nseg;
    uint32_t s = 0;
    for (uint32_t j = 0; j < k; ++j)
        if (first + j < nseg) s += std::min<uint32_t>(c[first + j], cap);
    uint32_t total = 0;
    uint32_t ex = wg_excl_scan(s, sh, total);
    for (uint32_t j = 0; j < k; ++j) {
        if (first + j < nseg) {
            pref[first + j] = ex;
            ex += std::min<uint32_t>(c[first + j], cap);
        }
    }
    if (threadIdx.x == 0) pref[nseg] = total;
    __syncthreads();
    return total;
}

// Records of the thread's U rows row[u] < pref[nseg] of partition `part`:
// the segment lo with pref[lo] <= row < pref[lo + 1], at offset
// row - pref[lo].  The U binary searches advance in lockstep (the same
// number of halvings for every row), so their LDS reads overlap.
template <int U>
DQC_DEVICE_INLINE void part_records(const uint32_t* pref, uint32_t nseg, uint32_t part, uint32_t cap,
                                    const int64_t (&row)[U], const int (&act)[U], uint64_t (&rec)[U]) {
    uint32_t lo[U], hi[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        lo[u] = 0u;
        hi[u] = act[u] ? nseg : 1u;
    }
    for (uint32_t n = nseg; n > 1u; n = (n + 1u) >> 1) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (hi[u] - lo[u] > 1u) {
                const uint32_t mid = (lo[u] + hi[u]) >> 1;
                if (pref[mid] <= (uint32_t)row[u]) lo[u] = mid;
                else hi[u] = mid;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
        rec[u] = act[u] ? ((uint64_t)part * nseg + lo[u]) * cap + ((uint32_t)row[u] - pref[lo[u]]) : 0ull;
}

// A value in one 8-byte word of a record, and back.
template <typename T>
DQC_DEVICE_INLINE uint64_t to_word(T v) {
    uint64_t w = 0;
    __builtin_memcpy(&w, &v, sizeof(T));
    return w;
}
template <typename T>
DQC_DEVICE_INLINE T from_word(uint64_t w) {
    T v;
    __builtin_memcpy(&v, &w, sizeof(T));
    return v;
}

// Compare-and-swap on an LDS word; returns the old value.
template <typename Tag>
DQC_DEVICE_INLINE Tag lds_cas(Tag* p, Tag expected, Tag desired) {
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return expected;
}

// Orders this wavefront's earlier LDS writes before its later LDS accesses
// by other lanes of the same wavefront (a wavefront's LDS operations complete
// in issue order; the fences keep the compiler from reordering across).
DQC_DEVICE_INLINE void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Claim or find a workgroup LDS slot for a 32-bit key (0xffffffff = empty),
// probing at most `tries` slots of a power-of-two table; -1 if none is free.
constexpr uint32_t LKEY_EMPTY = 0xffffffffu;

DQC_DEVICE_INLINE int lds_claim32(uint32_t* ltab, uint32_t mask, uint32_t key, int tries) {
    uint32_t s = (key * 0x9e3779b1u) & mask;
    for (int n = 0; n < tries; ++n) {
        uint32_t t = __hip_atomic_load(&ltab[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (t == key) return (int)s;
        if (t == LKEY_EMPTY) {
            uint32_t expected = LKEY_EMPTY;
            if (__hip_atomic_compare_exchange_strong(&ltab[s], &expected, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) ||
                expected == key)
                return (int)s;
        }
        s = (s + 1u) & mask;
    }
    return -1;
}

// ------------------------------------------------------------ atomics
DQC_DEVICE_INLINE void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
DQC_DEVICE_INLINE void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
DQC_DEVICE_INLINE void atomic_add(int32_t* p, int32_t v) { atomicAdd(p, v); }
DQC_DEVICE_INLINE void atomic_add(uint32_t* p, uint32_t v) { atomicAdd(p, v); }
DQC_DEVICE_INLINE void atomic_add(int64_t* p, int64_t v) {
    atomicAdd((unsigned long long*)p, (unsigned long long)v);
}

template <typename T>
DQC_DEVICE_INLINE void atomic_min_fp(T* p, T v) {
    T cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (v < cur) {
        if (__hip_atomic_compare_exchange_weak(p, &cur, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT))
            break;
    }
}

template <typename T>
DQC_DEVICE_INLINE void atomic_max_fp(T* p, T v) {
    T cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (v > cur) {
        if (__hip_atomic_compare_exchange_weak(p, &cur, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT))
            break;
    }
}

DQC_DEVICE_INLINE void atomic_min(double* p, double v) { atomic_min_fp(p, v); }
DQC_DEVICE_INLINE void atomic_max(double* p, double v) { atomic_max_fp(p, v); }
DQC_DEVICE_INLINE void atomic_min(float* p, float v) { atomic_min_fp(p, v); }
DQC_DEVICE_INLINE void atomic_max(float* p, float v) { atomic_max_fp(p, v); }
DQC_DEVICE_INLINE void atomic_min(int32_t* p, int32_t v) { atomicMin(p, v); }
DQC_DEVICE_INLINE void atomic_max(int32_t* p, int32_t v) { atomicMax(p, v); }
DQC_DEVICE_INLINE void atomic_min(uint32_t* p, uint32_t v) { atomicMin(p, v); }
DQC_DEVICE_INLINE void atomic_max(uint32_t* p, uint32_t v) { atomicMax(p, v); }
DQC_DEVICE_INLINE void atomic_min(int64_t* p, int64_t v) { atomicMin((long long*)p, (long long)v); }
DQC_DEVICE_INLINE void atomic_max(int64_t* p, int64_t v) { atomicMax((long long*)p, (long long)v); }

// LDS atomics (workgroup scope); ds_add_f64 is native on gfx950
DQC_DEVICE_INLINE void lds_add(double* p, double v) { atomicAdd(p, v); }
DQC_DEVICE_INLINE void lds_add(float* p, float v) { atomicAdd(p, v); }
DQC_DEVICE_INLINE void lds_add(int32_t* p, int32_t v) { atomicAdd(p, v); }
DQC_DEVICE_INLINE void lds_add(int64_t* p, int64_t v) { atomicAdd((unsigned long long*)p, (unsigned long long)v); }
template <typename T>
DQC_DEVICE_INLINE void lds_min(T* p, T v) {
    T cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (v < cur) {
        if (__hip_atomic_compare_exchange_weak(p, &cur, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP))
            break;
    }
}
template <typename T>
DQC_DEVICE_INLINE void lds_max(T* p, T v) {
    T cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (v > cur) {
        if (__hip_atomic_compare_exchange_weak(p, &cur, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP))
            break;
    }
}

// ------------------------------------------------------------- strings
// String values in registers are [start, end) byte ranges (util.h:11-14).
struct str_t {
    const char* start;
    const char* end;
};

// result string columns store byte offsets into the source char heap
struct str_offs {
    uint64_t start;
    uint64_t end;
};

DQC_DEVICE_INLINE str_t string_scan(const uint64_t* offs, const char* chars, int64_t ix) {
    str_t s;
    s.start = chars + offs[ix];
    s.end = chars + offs[ix + 1];
    return s;
}

DQC_DEVICE_INLINE str_offs to_string_offset(const char* chars, str_t s) {
    str_offs o;
    o.start = (uint64_t)(s.start - chars);
    o.end = (uint64_t)(s.end - chars);
    return o;
}

// Bytes [p, p + n), n <= 8, as a little-endian word with zeros above byte n.
// Reads the one or two aligned 8-byte words holding those bytes: an aligned
// word that holds a byte of a buffer lies inside that buffer's allocation
// (device allocations and code-object sections are 8-byte aligned or more).
DQC_DEVICE_INLINE uint64_t load_bytes(const char* p, uint32_t n) {
    const uintptr_t a = (uintptr_t)p & ~(uintptr_t)7;
    const uint32_t off = (uint32_t)((uintptr_t)p & 7u);
    uint64_t w = *(const uint64_t*)a >> (8u * off);
    if (off + n > 8u) w |= *(const uint64_t*)(a + 8u) << (64u - 8u * off);
    return n >= 8u ? w : (w & ((1ull << (8u * n)) - 1ull));
}

// Byte equality, 8 bytes per step.
DQC_DEVICE_INLINE bool string_equals(str_t a, str_t b) {
    const int64_t n = a.end - a.start;
    if (n != b.end - b.start) return false;
    for (int64_t o = 0; o < n; o += 8) {
        const uint32_t m = n - o < 8 ? (uint32_t)(n - o) : 8u;
        if (load_bytes(a.start + o, m) != load_bytes(b.start + o, m)) return false;
    }
    return true;
}

// String hash over 8-byte chunks: the chunks' mixes are summed (chunk k
// mixed with its index), so any split of the chunks over lanes followed by
// a sum gives the same value (string_hash_wave).
DQC_DEVICE_INLINE uint64_t string_chunk_mix(uint64_t w, uint64_t k) {
    return mix64(w ^ ((k + 1u) * 0x9E3779B97F4A7C15ull));
}
DQC_DEVICE_INLINE uint64_t string_hash_final(uint64_t sum, uint64_t n) {
    return mix64(sum ^ mix64(n + 0x84222325cbf29ce4ull));
}
DQC_DEVICE_INLINE uint64_t string_hash(str_t s) {
    const uint64_t n = (uint64_t)(s.end - s.start);
    uint64_t sum = 0;
#pragma unroll 4
    for (uint64_t o = 0; o < n; o += 8) {
        const uint32_t m = n - o < 8u ? (uint32_t)(n - o) : 8u;
        sum += string_chunk_mix(load_bytes(s.start + o, m), o >> 3);
    }
    return string_hash_final(sum, n);
}

// Wavefront-cooperative string equality (the push-down form of the
// reference's stringEqualsPushdown, util.cuh:104-130): for every active lane
// whose lengths agree, all 64 lanes compare 8-byte chunks of its two
// strings (lane l the chunks l, l + 64, ...) and a ballot combines them.
// Called by every lane of the wavefront (uniform control flow); inactive
// lanes get false.
DQC_DEVICE_INLINE bool string_equals_wave(bool active, str_t a, str_t b) {
    const uint32_t lane = lane_id();
    bool result = false;
    uint64_t todo = ballot(active && (a.end - a.start) == (b.end - b.start));
    while (todo) {
        const int t = first_lane(todo);
        todo &= todo - 1u;
        const str_t x = shfl(a, t);
        const str_t y = shfl(b, t);
        const uint64_t n = (uint64_t)(x.end - x.start);
        bool same = true;
        for (uint64_t o = 8u * lane; o < n; o += 8u * WAVE) {
            const uint32_t m = n - o < 8u ? (uint32_t)(n - o) : 8u;
            same = same && load_bytes(x.start + o, m) == load_bytes(y.start + o, m);
        }
        const bool all = ballot(!same) == 0ull;
        if ((int)lane == t) result = all;
    }
    return result;
}

// Wavefront-cooperative string hash (the push-down form of the reference's
// stringHashPushDown, hashing.cuh:39-62): the strings of the active lanes
// are hashed one after another by all 64 lanes, lane l taking the chunks
// l, l + 64, ...; a wavefront sum combines them.  The value equals
// string_hash's.  Called by every lane of the wavefront (uniform control
// flow); inactive lanes get 0.
DQC_DEVICE_INLINE uint64_t string_hash_wave(bool active, str_t s) {
    const uint32_t lane = lane_id();
    uint64_t result = 0;
    uint64_t todo = ballot(active);
    while (todo) {
        const int t = first_lane(todo);
        todo &= todo - 1u;
        const str_t h = shfl(s, t);
        const uint64_t n = (uint64_t)(h.end - h.start);
        uint64_t part = 0;
        for (uint64_t o = 8u * lane; o < n; o += 8u * WAVE) {
            const uint32_t m = n - o < 8u ? (uint32_t)(n - o) : 8u;
            part += string_chunk_mix(load_bytes(h.start + o, m), o >> 3);
        }
        const uint64_t sum = wave_sum(part);
        if ((int)lane == t) result = string_hash_final(sum, n);
    }
    return result;
}

// SQL SUBSTRING(s FROM from FOR len), 1-based, clipped to the string
// (semantics of util.cuh:132-142)
DQC_DEVICE_INLINE str_t string_substring(str_t s, int from, int len) {
    str_t r;
    r.start = s.start + (from - 1);
    r.end = r.start + len;
    if (r.start < s.start) r.start = s.start;
    if (r.end > s.end) r.end = s.end;
    if (r.start > r.end) r.start = r.end;
    return r;
}

// LIKE with the semantics of stringLikeCheck (util.cuh:148-199), which is
// not textbook LIKE: the literal prefix (before the first '%') and suffix
// (after the last '%') are compared independently, each only as far as the
// string reaches (so they may overlap: 'ab' LIKE 'ab%b' and 'a' LIKE 'a_'
// hold, and a pattern without '%' holds for every string that starts and
// ends with it); the segments between the first and last '%' are then found
// left to right between prefix and suffix, an empty segment ('%%') matching
// only a literal '%' byte.  '_' matches any byte.  Every read is bounds
// checked: the string byte the reference reads at the end of the searched
// region cannot change the result, and pattern bytes past its end read as 0
// (the terminator of the pattern literal there).
DQC_DEVICE_INLINE bool string_like(str_t s, str_t p) {
    const int64_t n = s.end - s.start;
    const int64_t m = p.end - p.start;
    auto pc = [&](int64_t i) -> char { return (i >= 0 && i < m) ? p.start[i] : '\0'; };
    auto cmp = [](char c, char l) { return c == l || l == '_'; };
    int64_t lPos = 0, sPos = 0;
    int64_t lInStart = 0, lInEnd = m, sInStart = 0, sInEnd = n;
    if (pc(0) != '%') {  // prefix
        for (; lPos < m && sPos < n && pc(lPos) != '%'; ++lPos, ++sPos)
            if (!cmp(s.start[sPos], pc(lPos))) return false;
        lInStart = lPos;
        sInStart = sPos;
    }
    if (m == 0) return true;  // the empty pattern matches every string
    if (pc(m - 1) != '%') {  // suffix
        sPos = n - 1;
        lPos = m - 1;
        for (; lPos >= 0 && sPos >= 0 && pc(lPos) != '%'; --lPos, --sPos)
            if (!cmp(s.start[sPos], pc(lPos))) return false;
        lInEnd = lPos;
        sInEnd = sPos + 1;
    }
    if (lInStart < lInEnd) {  // infix segments, left to right
        lPos = lInStart + 1;
        sPos = sInStart;
        while (sPos < sInEnd && lPos < lInEnd) {
            int64_t lTrace = lPos, sTrace = sPos;
            while (sTrace < sInEnd && cmp(s.start[sTrace], pc(lTrace))) {
                ++lTrace;
                if (pc(lTrace) == '%') {
                    lPos = ++lTrace;
                    sPos = sTrace;
                    break;
                }
                ++sTrace;
            }
            ++sPos;
        }
    }
    return lPos >= lInEnd;
}

// Batch segment of row `row`: the first member k with row < prefix[k] (the
// members' row counts of the table, nondecreasing); members beyond the row's
// prefix never occur because the batched run scans the largest prefix.
DQC_DEVICE_INLINE int32_t segment_of(const int64_t* prefix, int64_t members, int64_t row) {
    int64_t lo = 0, hi = members > 0 ? members - 1 : 0;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (row < prefix[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    return (int32_t)lo;
}

// The same for a row of the workgroup tile [base, base + tile): when the
// tile's first and last rows share a segment (almost every tile) the
// workgroup-uniform answer avoids a search per lane.
DQC_DEVICE_INLINE int32_t segment_of(const int64_t* prefix, int64_t members, int64_t row, int64_t base,
                                     int64_t tile) {
    const int32_t s0 = segment_of(prefix, members, base);
    const int32_t s1 = segment_of(prefix, members, base + tile - 1);
    return s0 == s1 ? s0 : segment_of(prefix, members, row);
}

// LIKE patterns made of literals and '%' only (no '_'), holding a '%' and no
// "%%", have generated matchers.  For a string at least as long as the
// prefix and suffix together, string_like's answer is the textbook one: the
// prefix and suffix compare bytes in place, and each inner literal is found
// left to right by string_find, which reads the string with aligned 8-byte
// loads and slides a register window over its bytes (an aligned word holding
// a byte of the string lies in the same page as that byte, so the loads stay
// in mapped memory).  Shorter strings take string_like.

DQC_DEVICE_INLINE bool bytes_equal(const char* a, const char* b, uint32_t n) {
    for (uint32_t k = 0; k < n; ++k)
        if (a[k] != b[k]) return false;
    return true;
}

// One past the first occurrence of lit[0..L) in [from, end), or nullptr.
// head holds lit's first H = min(L, 8) bytes, lit[H-1] in the low byte;
// hmask covers H bytes.
DQC_DEVICE_INLINE const char* string_find(const char* from, const char* end, const char* lit, uint32_t L,
                                          uint64_t head, uint64_t hmask, uint32_t H) {
    if (end - from < (int64_t)L) return nullptr;
    const char* last = end - L;
    uint64_t win = 0;
    uint32_t have = 0;
    // the string's aligned words are loaded 8 at a time (independent loads,
    // one round trip for up to 64 bytes), then scanned byte by byte
    constexpr int BATCH = 8;
    for (const char* p0 = (const char*)((uintptr_t)from & ~(uintptr_t)7); p0 < end; p0 += 8 * BATCH) {
        uint64_t w[BATCH];
#pragma unroll
        for (int i = 0; i < BATCH; ++i) w[i] = p0 + 8 * i < end ? *(const uint64_t*)(p0 + 8 * i) : 0ull;
#pragma unroll
        for (int i = 0; i < BATCH; ++i) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const char* q = p0 + 8 * i + j;
                if (q >= from && q < end) {
                    win = (win << 8) | ((w[i] >> (8 * j)) & 0xffu);
                    ++have;
                    if (have >= H && (win & hmask) == head) {
                        const char* st = q - (H - 1);
                        if (st > last) return nullptr;
                        if (L <= H || bytes_equal(st + H, lit + H, L - H)) return st + L;
                    }
                }
            }
        }
    }
    return nullptr;
}

// Occupied slots of a hash table's tag array (nonzero tags), added to *out:
// the number of keys the table holds, measured for cardinality feedback.
template <typename Tag>
__global__ void __launch_bounds__(256) dqc_census(const Tag* tags, uint64_t n, unsigned long long* out) {
    uint64_t c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
        c += tags[i] != (Tag)0 ? 1u : 0u;
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(out, (unsigned long long)c);
}

// --------------------------------------------- partitioned unique build
// Keys whose probe run in the second pass of a partitioned build crossed
// the end of their partition's slot range wait in a spill list of records
// [tag, home slot, payload words...]; after that pass this kernel inserts
// them with the ordinary CAS probe (every slot from the home slot to the
// range end is taken, so the key lands where a sequential build would put
// it) and copies each payload word's low `size[c]` bytes to column c.
struct spill_cols {
    void* ptr[8];
    uint32_t size[8];
    uint32_t n;
};

template <typename Tag>
__global__ void __launch_bounds__(256) dqc_spill_insert(Tag* tags, uint32_t cap, const uint64_t* recs, uint32_t words,
                                                          const uint32_t* count, uint32_t rec_cap, spill_cols cols,
                                                          uint32_t* flags, uint32_t* overflow_mask,
                                                          uint32_t overflow_bit) {
    const uint32_t n = std::min<uint32_t>(*count, rec_cap);
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint64_t* r = recs + (uint64_t)i * words;
        const Tag tag = (Tag)r[0];
        const int64_t s = ht_insert_from<Tag>(tags, cap, tag, (uint32_t)r[1], 0u, flags);
        if (s < 0) {
            raise_status(flags, ST_HT_OVERFLOW);
            atomicOr(overflow_mask, overflow_bit);
            continue;
        }
        for (uint32_t c = 0; c < cols.n && c < 8u; ++c)
            __builtin_memcpy((char*)cols.ptr[c] + (uint64_t)s * cols.size[c], &r[2 + c], cols.size[c]);
    }
}

// ---------------------------------------------------- multimap scan kernels
// Bucket directory of a multimap: dir[2b] = offset of bucket b's entries,
// dir[2b+1] = its count.  The count pass increments dir[2b+1]; these three
// kernels turn counts into offsets by an exclusive scan (per-block totals, a
// scan of the block totals, and a per-block scan that adds the block's
// base) and clear the counts, which the insert pass reuses as cursors (a
// count ends equal to its bucket's size again once every entry is in).
constexpr int SCAN_BS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_BS * SCAN_ITEMS;

__device__ inline uint64_t block_excl_scan(uint64_t v, uint64_t* sh, uint64_t& total) {
    const int t = threadIdx.x;
    const int w = t / WAVE;
    uint64_t incl = wave_incl_scan(v);
    if ((int)lane_id() == WAVE - 1) sh[w] = incl;
    __syncthreads();
    if (t == 0) {
        uint64_t run = 0;
        for (int i = 0; i < SCAN_BS / WAVE; ++i) {
            uint64_t x = sh[i];
            sh[i] = run;
            run += x;
        }
        sh[SCAN_BS / WAVE] = run;
    }
    __syncthreads();
    uint64_t r = sh[w] + incl - v;
    total = sh[SCAN_BS / WAVE];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(SCAN_BS) dqc_mm_scan_reduce(const uint32_t* dir, uint64_t n,
                                                               uint64_t* block_sums) {
    __shared__ uint64_t sh[SCAN_BS / WAVE + 1];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t ix = base + (uint64_t)i * SCAN_BS + threadIdx.x;
        if (ix < n) acc += dir[2 * ix + 1];
    }
    uint64_t total;
    block_excl_scan(acc, sh, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// single workgroup: exclusive scan of the block totals, writes the grand total
__global__ void __launch_bounds__(SCAN_BS) dqc_mm_scan_blocks(uint64_t* block_sums, uint64_t nblocks,
                                                               uint64_t* total_out) {
    __shared__ uint64_t sh[SCAN_BS / WAVE + 1];
    uint64_t carry = 0;
    for (uint64_t base = 0; base < nblocks; base += SCAN_BS) {
        uint64_t ix = base + threadIdx.x;
        uint64_t v = ix < nblocks ? block_sums[ix] : 0;
        uint64_t total;
        uint64_t ex = block_excl_scan(v, sh, total);
        if (ix < nblocks) block_sums[ix] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

// With keep_counts the counts stay (an insert pass that places entries by
// precomputed ranks needs no cursors); otherwise they are cleared.
__global__ void __launch_bounds__(SCAN_BS) dqc_mm_scan_apply(uint32_t* dir, uint64_t n,
                                                              const uint64_t* block_sums, uint32_t keep_counts) {
    __shared__ uint64_t sh[SCAN_BS / WAVE + 1];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t ix = base + i;
        v[i] = ix < n ? dir[2 * ix + 1] : 0u;
        acc += v[i];
    }
    uint64_t total;
    uint64_t run = block_excl_scan(acc, sh, total) + block_sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t ix = base + i;
        if (ix < n) {
            dir[2 * ix] = (uint32_t)run;
            if (!keep_counts) dir[2 * ix + 1] = 0u;
        }
        run += v[i];
    }
}

// Chunk directory of a push-down join's long-list queue (single
// workgroup): item i, holding len[i] matches, owns chunks
// [cstart[i], cstart[i+1]) of `chunk` matches each; cstart[n] and *total are
// the number of chunks.  n = min(*count, cap) items were queued.
__global__ void __launch_bounds__(SCAN_BS) dqc_split_scan(const uint32_t* len, const uint64_t* count, uint64_t cap,
                                                          uint32_t chunk, uint64_t* cstart, uint64_t* total) {
    __shared__ uint64_t sh[SCAN_BS / WAVE + 1];
    const uint64_t n = *count < cap ? *count : cap;
    uint64_t carry = 0;
    // SCAN_ITEMS consecutive items per thread and trip: one block scan per
    // SCAN_TILE items
    for (uint64_t base = 0; base < n; base += SCAN_TILE) {
        const uint64_t b0 = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
        uint64_t v[SCAN_ITEMS];
        uint64_t sum = 0;
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const uint64_t ix = b0 + (uint64_t)i;
            v[i] = ix < n ? ((uint64_t)len[ix] + chunk - 1u) / chunk : 0u;
            sum += v[i];
        }
        uint64_t t;
        uint64_t run = carry + block_excl_scan(sum, sh, t);
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const uint64_t ix = b0 + (uint64_t)i;
            if (ix < n) cstart[ix] = run;
            run += v[i];
        }
        carry += t;
    }
    if (threadIdx.x == 0) {
        cstart[n] = carry;
        *total = carry;
    }
}

// Item owning chunk c: the last i < n with cstart[i] <= c.
DQC_DEVICE_INLINE uint32_t split_item(const uint64_t* cstart, uint64_t n, uint64_t c) {
    uint64_t lo = 0, hi = n;   // cstart[lo] <= c < cstart[hi]
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (cstart[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    return (uint32_t)lo;
}

}  // namespace dqc

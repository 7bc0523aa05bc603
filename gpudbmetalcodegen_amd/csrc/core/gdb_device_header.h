#pragma once
// Device runtime library prepended to every generated HIP source.
//
// Counterpart of codegen/core/metal_common_header.h:39-204, written for
// gfx950 (CDNA4): 64-lane waves, DPP row operations and readlane instead of
// 32-wide simd_shuffle trees, native 64-bit global/LDS atomics instead of
// emulated lo/hi 32-bit pairs, 16-byte nontemporal column loads, and the hash
// table, bitmap, LIKE and top-k helpers the operators emit calls to.

namespace gdb {
namespace codegen {

inline const char* kGdbDeviceHeader = R"GDBHIP(
// ===== gdb device runtime (gfx950, wave64) =====
typedef signed char i8;
typedef unsigned char u8;
typedef short i16;
typedef unsigned short u16;
typedef int i32;
typedef unsigned int u32;
typedef long long i64;
typedef unsigned long long u64;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

#define GDB_DEV __device__ __forceinline__
#define GDB_I64_MAX 0x7FFFFFFFFFFFFFFFll
#define GDB_I64_MIN (-0x7FFFFFFFFFFFFFFFll - 1ll)
#define GDB_EMPTY_KEY 0xFFFFFFFFu
#define GDB_EMPTY_WORD 0x00000000FFFFFFFFull
#define GDB_KNUTH 2654435769u

// ---------------------------------------------------------------------------
// Column loads. A thread owns ITEMS consecutive rows of a tile, so a column of
// T is read as sizeof(T)*ITEMS contiguous bytes with the widest aligned vector
// load; consecutive lanes read consecutive chunks (coalesced). Loads are
// nontemporal: a scan touches each column byte once, and keeping it out of the
// caches leaves L2/MALL to hash tables and bitmaps.
// ---------------------------------------------------------------------------
template <typename V>
GDB_DEV V gdb_ld_stream(const V* p) {
    return __builtin_nontemporal_load(p);
}

template <typename T, int N>
GDB_DEV void gdb_load_items(const T* __restrict__ p, T (&r)[N]) {
    constexpr int B = (int)sizeof(T) * N;
    if constexpr (B % 16 == 0) {
#pragma unroll
        for (int k = 0; k < B / 16; ++k) {
            u32x4 v = gdb_ld_stream(reinterpret_cast<const u32x4*>(p) + k);
            __builtin_memcpy(reinterpret_cast<char*>(&r[0]) + 16 * k, &v, 16);
        }
    } else if constexpr (B % 8 == 0) {
#pragma unroll
        for (int k = 0; k < B / 8; ++k) {
            u32x2 v = gdb_ld_stream(reinterpret_cast<const u32x2*>(p) + k);
            __builtin_memcpy(reinterpret_cast<char*>(&r[0]) + 8 * k, &v, 8);
        }
    } else if constexpr (B % 4 == 0) {
#pragma unroll
        for (int k = 0; k < B / 4; ++k) {
            u32 v = gdb_ld_stream(reinterpret_cast<const u32*>(p) + k);
            __builtin_memcpy(reinterpret_cast<char*>(&r[0]) + 4 * k, &v, 4);
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) r[k] = p[k];
    }
}

template <typename T, int N>
GDB_DEV void gdb_load_tail(const T* __restrict__ base, u64 r0, u64 n, T (&r)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) r[k] = (r0 + (u64)k < n) ? base[r0 + k] : (T)0;
}

// ---------------------------------------------------------------------------
// Wave64 reductions: DPP quad_perm [1,0,3,2], [2,3,0,1], row_ror:4, row_ror:8
// leave every lane with its 16-lane row total; four readlanes combine the
// rows. The wave must be fully active (block sizes are multiples of 64).
// ---------------------------------------------------------------------------
template <int CTRL>
GDB_DEV u32 gdb_dpp32(u32 v) {
    return (u32)__builtin_amdgcn_update_dpp((int)0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
GDB_DEV u64 gdb_dpp64(u64 v) {
    const u32 lo = gdb_dpp32<CTRL>((u32)v);
    const u32 hi = gdb_dpp32<CTRL>((u32)(v >> 32));
    return ((u64)hi << 32) | (u64)lo;
}
GDB_DEV u64 gdb_readlane64(u64 v, int lane) {
    const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)v, lane);
    const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(v >> 32), lane);
    return ((u64)hi << 32) | (u64)lo;
}
GDB_DEV u64 gdb_wave_sum_u64(u64 v) {
    v += gdb_dpp64<0xB1>(v);
    v += gdb_dpp64<0x4E>(v);
    v += gdb_dpp64<0x124>(v);
    v += gdb_dpp64<0x128>(v);
    return gdb_readlane64(v, 0) + gdb_readlane64(v, 16) + gdb_readlane64(v, 32) + gdb_readlane64(v, 48);
}
GDB_DEV i64 gdb_min64(i64 a, i64 b) { return a < b ? a : b; }
GDB_DEV i64 gdb_max64(i64 a, i64 b) { return a > b ? a : b; }
GDB_DEV i64 gdb_wave_min_i64(i64 v) {
    v = gdb_min64(v, (i64)gdb_dpp64<0xB1>((u64)v));
    v = gdb_min64(v, (i64)gdb_dpp64<0x4E>((u64)v));
    v = gdb_min64(v, (i64)gdb_dpp64<0x124>((u64)v));
    v = gdb_min64(v, (i64)gdb_dpp64<0x128>((u64)v));
    return gdb_min64(gdb_min64((i64)gdb_readlane64((u64)v, 0), (i64)gdb_readlane64((u64)v, 16)),
                     gdb_min64((i64)gdb_readlane64((u64)v, 32), (i64)gdb_readlane64((u64)v, 48)));
}
GDB_DEV i64 gdb_wave_max_i64(i64 v) {
    v = gdb_max64(v, (i64)gdb_dpp64<0xB1>((u64)v));
    v = gdb_max64(v, (i64)gdb_dpp64<0x4E>((u64)v));
    v = gdb_max64(v, (i64)gdb_dpp64<0x124>((u64)v));
    v = gdb_max64(v, (i64)gdb_dpp64<0x128>((u64)v));
    return gdb_max64(gdb_max64((i64)gdb_readlane64((u64)v, 0), (i64)gdb_readlane64((u64)v, 16)),
                     gdb_max64((i64)gdb_readlane64((u64)v, 32), (i64)gdb_readlane64((u64)v, 48)));
}
__host__ __device__ constexpr u32 gdb_lds_align16(u32 x) { return (x + 15u) & ~15u; }
__host__ __device__ constexpr u32 gdb_cmax(u32 a, u32 b) { return a > b ? a : b; }
GDB_DEV int gdb_lane() { return (int)(threadIdx.x & 63u); }
GDB_DEV int gdb_wave() { return (int)(threadIdx.x >> 6); }
GDB_DEV u64 gdb_lanemask_lt() { return (1ull << gdb_lane()) - 1ull; }
// Orders one wave's LDS stores before its lanes read each other's entries.
GDB_DEV void gdb_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Atomics (relaxed). Global ones are device (agent) scope, LDS ones are
// workgroup scope; results unused compile to no-return atomics.
// ---------------------------------------------------------------------------
GDB_DEV void gdb_atomic_add_u64(u64* p, u64 v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV u64 gdb_atomic_add_u64_ret(u64* p, u64 v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV void gdb_atomic_add_u32(u32* p, u32 v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV u32 gdb_atomic_inc_u32(u32* p) {
    return __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV void gdb_atomic_min_i64(i64* p, i64 v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV void gdb_atomic_max_i64(i64* p, i64 v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV void gdb_atomic_max_u64(u64* p, u64 v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV void gdb_atomic_or_u32(u32* p, u32 v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV bool gdb_cas_u64(u64* p, u64& expected, u64 desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV u64 gdb_load_relaxed_u64(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GDB_DEV void gdb_lds_add_u64(u64* p, u64 v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
GDB_DEV void gdb_lds_add_u32(u32* p, u32 v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
GDB_DEV void gdb_lds_min_i64(i64* p, i64 v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
GDB_DEV void gdb_lds_max_i64(i64* p, i64 v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
GDB_DEV bool gdb_lds_cas_u64(u64* p, u64& expected, u64 desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
}
GDB_DEV u64 gdb_lds_load_u64(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------------------------
// Bitmaps: bit k of word k>>5 (codegen/core/metal_common_header.h:103-111).
// Keys outside [0, nbits) are absent and are never set.
// ---------------------------------------------------------------------------
GDB_DEV bool gdb_bitmap_test(const u32* __restrict__ bm, u64 nbits, i64 key) {
    return (u64)key < nbits && ((bm[(u64)key >> 5] >> ((u32)key & 31u)) & 1u);
}
// Bits only go from 0 to 1 during a build, so a plain read that already sees
// the bit skips the atomic (hot keys of a skewed column would otherwise
// serialise on one word); a stale read only costs a redundant atomic.
GDB_DEV void gdb_bitmap_set(u32* __restrict__ bm, u64 nbits, i64 key) {
    if ((u64)key >= nbits) return;
    u32* w = bm + ((u64)key >> 5);
    const u32 m = 1u << ((u32)key & 31u);
    if ((__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & m) == 0u) gdb_atomic_or_u32(w, m);
}

// 16 byte flags (nonzero = set) -> 16 bits, flag k at bit k. Per 32-bit word:
// each byte's bits are ORed into its bit 0 (shifts stay inside the byte for
// bit 0), then ((v & 0x01010101) * 0x01020408) >> 24 gathers the four bit 0s
// at bits 24..27 (the ten partial products land on distinct bits: no carries).
GDB_DEV u32 gdb_pack_flags4(u32 v) {
    v |= v >> 4;
    v |= v >> 2;
    v |= v >> 1;
    return (((v & 0x01010101u) * 0x01020408u) >> 24) & 0xFu;
}
GDB_DEV u32 gdb_pack_flags16(u32x4 x) {
    return gdb_pack_flags4(x.x) | (gdb_pack_flags4(x.y) << 4) | (gdb_pack_flags4(x.z) << 8) | (gdb_pack_flags4(x.w) << 12);
}

// Records in a block-local bitmap cache entry `e` (read as `c`) that bit m of
// word w has had its global atomic issued. An entry may claim only bits whose
// global atomic was issued. A different word replaces the entry. The same
// word gains the bit by a compare-exchange against the entry as read, never
// by an OR into whatever the entry holds by then: between the read and an OR
// another lane may have replaced it with a different word, which would then
// claim a bit nobody set there. A failed exchange only loses the note (a
// later redundant atomic).
GDB_DEV void gdb_bitmap_cache_note(u64* e, u64 c, bool same, u32 w, u32 m) {
    if (same) {
        u64 expect = c;
        (void)__hip_atomic_compare_exchange_strong(e, &expect, c | (u64)m, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        __hip_atomic_store(e, ((u64)w << 32) | (u64)m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// A bitmap set behind a block-local cache (1024 entries of LDS u64: word
// index << 32 | the bits this block has set in it). A hit skips the global
// atomic; gdb_bitmap_cache_note keeps every claimed bit backed by an issued
// global atomic. tested: also skip the atomic when a plain read already sees
// the bit — for skewed keys, where hot words re-missed after an eviction
// would otherwise be hit by every block (atomics on one address serialise);
// for near-unique keys the extra read only adds latency.
template <bool TESTED>
GDB_DEV void gdb_bitmap_set_cached(u32* __restrict__ bm, u64 nbits, i64 key, u64* cache) {
    if ((u64)key >= nbits) return;
    const u32 w = (u32)((u64)key >> 5);
    const u32 m = 1u << ((u32)key & 31u);
    u64* e = cache + (w & 1023u);
    const u64 c = __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const bool same = (u32)(c >> 32) == w;
    if (same && ((u32)c & m)) return;
    if (!TESTED || (__hip_atomic_load(bm + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & m) == 0u)
        gdb_atomic_or_u32(bm + w, m);  // no return value: nothing waits on it
    gdb_bitmap_cache_note(e, c, same, w, m);
}

// ---------------------------------------------------------------------------
// Open-addressing hash tables over 64-bit words. Word 0 of a slot packs the
// 32-bit key (low half, empty = 0xFFFFFFFF) with a 32-bit companion (high
// half: a build payload, or a group's row count). Linear probing from the high
// bits of key * 2654435769 (Fibonacci hashing); probes are bounded so a full
// table reports overflow instead of spinning.
// ---------------------------------------------------------------------------
GDB_DEV u64 gdb_hash_slot(u32 key, u32 bits) {
    return (u64)((key * GDB_KNUTH) >> (32u - bits));
}

// Insert a unique key with a 32-bit payload. Returns the slot, -1 when the
// probe bound is reached (overflow), -2 when the key is already present.
GDB_DEV i64 gdb_ht_insert_unique(u64* __restrict__ tab, u32 slot_words, u32 bits, u32 key, u32 payload,
                                 u32 max_probe) {
    const u64 mask = (1ull << bits) - 1ull;
    u64 s = gdb_hash_slot(key, bits);
    const u64 mine = (u64)key | ((u64)payload << 32);
    // CAS first: at the load factors used (<= 1/2) the home slot is usually
    // empty, and a failed CAS returns the occupant, so no separate load.
    for (u32 p = 0; p < max_probe; ++p, s = (s + 1) & mask) {
        u64* w = tab + s * slot_words;
        u64 cur = GDB_EMPTY_WORD;
        if (gdb_cas_u64(w, cur, mine)) return (i64)s;
        if ((u32)cur == GDB_EMPTY_KEY) {  // empty key, other companion bits: claim it as it is
            if (gdb_cas_u64(w, cur, mine)) return (i64)s;
        }
        if ((u32)cur == key) return -2;
    }
    return -1;
}

// Find a key; returns the slot or -1. The table must be complete (built by an
// earlier launch), so plain loads are coherent.
GDB_DEV i64 gdb_ht_find(const u64* __restrict__ tab, u32 slot_words, u32 bits, u32 key, u32 max_probe) {
    const u64 mask = (1ull << bits) - 1ull;
    u64 s = gdb_hash_slot(key, bits);
    for (u32 p = 0; p < max_probe; ++p, s = (s + 1) & mask) {
        const u32 k = (u32)tab[s * slot_words];
        if (k == key) return (i64)s;
        if (k == GDB_EMPTY_KEY) return -1;
    }
    return -1;
}

// 64-bit-key table: slots of [key, payload] u64 words, empty key all ones
// (keys must not be ~0). Home slot: the high bits of key * 2^64/phi (64-bit
// Fibonacci hashing). The reference folds 64-bit keys to 32 bits with
// k ^ (k >> 32) (metal_param_binding.h:20-26); for a composite key
// (partkey << 32 | suppkey) that fold is partkey ^ suppkey, which clusters
// the pairs of structured keys into long probe chains, so it is not used.
// Insert claims the key word with a CAS and then writes the payload (a later
// phase reads it); a key already present keeps its first payload. Returns 1
// inserted, 0 present, -1 when max_probe slots are full.
GDB_DEV u64 gdb_hash_slot64(u64 key, u32 bits) { return (key * 0x9E3779B97F4A7C15ull) >> (64u - bits); }
// Insert key (payload stored by the inserting row only); *slot receives the
// key's slot, new or existing (-1 on overflow). Returns 1 when this call
// inserted the key, 0 when it was present, -1 on overflow. A plain read
// comes first: a key, once written, never changes, so a slot already
// holding it (a hot group key) is recognised without an atomic; the
// compare-and-swap decides only for slots read as empty (a stale empty read
// just falls through to it).
GDB_DEV int gdb_ht64_insert_at(u64* __restrict__ tab, u32 bits, u64 key, u64 payload, u32 max_probe, i64* slot) {
    const u64 mask = (1ull << bits) - 1ull;
    u64 s = gdb_hash_slot64(key, bits);
    for (u32 p = 0; p < max_probe; ++p, s = (s + 1) & mask) {
        const u64 seen = tab[2 * s];
        if (seen == key) {
            *slot = (i64)s;
            return 0;
        }
        if (seen != ~0ull) continue;
        u64 expected = ~0ull;
        if (gdb_cas_u64(tab + 2 * s, expected, key)) {
            tab[2 * s + 1] = payload;
            *slot = (i64)s;
            return 1;
        }
        if (expected == key) {
            *slot = (i64)s;
            return 0;
        }
    }
    *slot = -1;
    return -1;
}
// Block-level first-arrival filter for a distinct insert: an LDS set (512
// slots, ~0 = empty) of the keys this block has already sent to the global
// table. True when this lane is the block's first with the key, or when 2
// probes find no room (many distinct keys: the filter then costs two LDS
// reads a row): only those lanes touch the global table, so a few hot keys
// cost a few global operations per block instead of one per row.
GDB_DEV bool gdb_block_first(u64* seen, u64 key) {
    u32 h = (u32)((key * 0x9E3779B97F4A7C15ull) >> 55);
    for (int p = 0; p < 2; ++p, h = (h + 1) & 511u) {
        const u64 cur = __hip_atomic_load(seen + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == key) return false;
        if (cur == ~0ull) {
            u64 exp = ~0ull;
            if (__hip_atomic_compare_exchange_strong(seen + h, &exp, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                return true;
            if (exp == key) return false;
        }
    }
    return true;
}
GDB_DEV int gdb_ht64_insert(u64* __restrict__ tab, u32 bits, u64 key, u64 payload, u32 max_probe) {
    i64 slot;
    return gdb_ht64_insert_at(tab, bits, key, payload, max_probe, &slot);
}
GDB_DEV i64 gdb_ht64_find(const u64* __restrict__ tab, u32 bits, u64 key, u32 max_probe) {
    const u64 mask = (1ull << bits) - 1ull;
    u64 s = gdb_hash_slot64(key, bits);
    for (u32 p = 0; p < max_probe; ++p, s = (s + 1) & mask) {
        const u64 k = tab[2 * s];
        if (k == key) return (i64)s;
        if (k == ~0ull) return -1;
    }
    return -1;
}

// Group-by slot: [key | count<<32, sum, min, max]. Merge one partial group
// (cnt rows with the given sum/min/max) into a global table. Returns 1 when a
// new key was inserted, 0 when merged, -1 on overflow.
GDB_DEV int gdb_gb_merge(u64* __restrict__ tab, u32 bits, u32 key, u32 cnt, i64 sum, i64 mn, i64 mx,
                         u32 max_probe) {
    const u64 mask = (1ull << bits) - 1ull;
    u64 s = gdb_hash_slot(key, bits);
    for (u32 p = 0; p < max_probe; ++p, s = (s + 1) & mask) {
        u64* w = tab + s * 4;
        u64 cur = gdb_load_relaxed_u64(w);
        int inserted = 0;
        if ((u32)cur == GDB_EMPTY_KEY) {
            u64 exp = cur;
            if (gdb_cas_u64(w, exp, (u64)key | ((u64)cnt << 32))) {
                inserted = 1;
            } else {
                cur = exp;
            }
        }
        if (inserted || (u32)cur == key) {
            if (!inserted) gdb_atomic_add_u64(w, (u64)cnt << 32);
            gdb_atomic_add_u64(w + 1, (u64)sum);
            gdb_atomic_min_i64((i64*)(w + 2), mn);
            gdb_atomic_max_i64((i64*)(w + 3), mx);
            return inserted;
        }
    }
    return -1;
}

// Block-local (LDS) group-by table: the slot words of the global table
// (key | count << 32, sum, min, max) stored word-major — word j of slot s at
// tab[j * slots + s] — so a wave's 64 random slots spread over 32 bank pairs
// instead of the 8 a 32-byte slot-major layout reaches (measured ~8 conflict
// cycles per LDS instruction with it). Returns the slot index, or -1 when the
// probe window is full (caller spills to global).
GDB_DEV int gdb_lds_gb_update(u64* tab, u32 bits, u32 key, i64 v, u32 max_probe) {
    const u32 n = 1u << bits, mask = n - 1u;
    u32 s = (u32)gdb_hash_slot(key, bits);
    for (u32 p = 0; p < max_probe; ++p, s = (s + 1) & mask) {
        u64* w = tab + s;
        u64 cur = gdb_lds_load_u64(w);
        bool inserted = false;
        if ((u32)cur == GDB_EMPTY_KEY) {
            u64 exp = cur;
            if (gdb_lds_cas_u64(w, exp, (u64)key | (1ull << 32))) {
                inserted = true;
            } else {
                cur = exp;
            }
        }
        if (inserted || (u32)cur == key) {
            if (!inserted) gdb_lds_add_u64(w, 1ull << 32);
            gdb_lds_add_u64(w + n, (u64)v);
            // min only falls and max only rises: a value no better than the
            // word as read (possibly stale, never better than current) cannot
            // change it, so the atomic is skipped (hot keys of a skewed column
            // otherwise serialise four same-address LDS atomics per row)
            if (v < (i64)gdb_lds_load_u64(w + 2 * n)) gdb_lds_min_i64((i64*)(w + 2 * n), v);
            if (v > (i64)gdb_lds_load_u64(w + 3 * n)) gdb_lds_max_i64((i64*)(w + 3 * n), v);
            return (int)s;
        }
    }
    return -1;
}

// Merge a partial group (cnt rows, sum/min/max) into a word-major LDS table
// of (mask + 1) slots starting at slot s. Returns the slot or -1 after
// max_probe slots.
GDB_DEV int gdb_lds_gb_merge_at(u64* tab, u32 mask, u32 s, u32 key, u32 cnt, i64 sum, i64 mn, i64 mx, u32 max_probe) {
    const u32 n = mask + 1u;
    for (u32 p = 0; p < max_probe; ++p, s = (s + 1) & mask) {
        u64* w = tab + s;
        u64 cur = gdb_lds_load_u64(w);
        bool inserted = false;
        if ((u32)cur == GDB_EMPTY_KEY) {
            u64 exp = cur;
            if (gdb_lds_cas_u64(w, exp, (u64)key | ((u64)cnt << 32))) {
                inserted = true;
            } else {
                cur = exp;
            }
        }
        if (inserted || (u32)cur == key) {
            if (!inserted) gdb_lds_add_u64(w, (u64)cnt << 32);
            gdb_lds_add_u64(w + n, (u64)sum);
            if (mn < (i64)gdb_lds_load_u64(w + 2 * n)) gdb_lds_min_i64((i64*)(w + 2 * n), mn);
            if (mx > (i64)gdb_lds_load_u64(w + 3 * n)) gdb_lds_max_i64((i64*)(w + 3 * n), mx);
            return (int)s;
        }
    }
    return -1;
}
// Membership of key in the key array of an LDS group table built by
// gdb_lds_gb_update (same home slot and probe window): linear probing from
// the home slot until the key, an empty slot or max_probe slots.
GDB_DEV bool gdb_lds_key_member(const u32* keys, u32 bits, u32 key, u32 max_probe) {
    const u32 mask = (1u << bits) - 1u;
    u32 s = (u32)gdb_hash_slot(key, bits);
    for (u32 p = 0; p < max_probe; ++p, s = (s + 1) & mask) {
        const u32 k = keys[s];
        if (k == key) return true;
        if (k == GDB_EMPTY_KEY) return false;
    }
    return false;
}
// Hash partition of a key: the top pbits of key * 2654435769 (0 when pbits = 0).
GDB_DEV u32 gdb_partition(u32 key, u32 pbits) { return pbits ? (u32)((key * GDB_KNUTH) >> (32u - pbits)) : 0u; }

// ---------------------------------------------------------------------------
// SQL LIKE '%w1%w2%' on a fixed-width NUL-padded field: the value ends at the
// first NUL (or at width); it matches when w1 occurs and w2 occurs at or after
// the end of the leftmost w1 (the semantics of q13_comment_match,
// codegen/planning/metal_plan_builder.cpp:1249-1266, for any two words).
// ---------------------------------------------------------------------------
GDB_DEV bool gdb_match_at(const char* s, int at, const char* w, int lw) {
    for (int j = 0; j < lw; ++j)
        if (s[at + j] != w[j]) return false;
    return true;
}
GDB_DEV bool gdb_like2(const char* s, int width, const char* w1, int l1, const char* w2, int l2) {
    int len = width;
    for (int i = 0; i < width; ++i) {
        if (s[i] == 0) {
            len = i;
            break;
        }
    }
    int p = -1;
    for (int i = 0; i + l1 <= len; ++i) {
        if (gdb_match_at(s, i, w1, l1)) {
            p = i;
            break;
        }
    }
    if (p < 0) return false;
    for (int i = p + l1; i + l2 <= len; ++i)
        if (gdb_match_at(s, i, w2, l2)) return true;
    return false;
}

// Bit-parallel LIKE '%w1%w2%' (Shift-And automata, one per word; w2's runs
// only from the byte after w1's first occurrence ends, which is the earliest
// and so the only start that matters). t1/t2 are 256-entry tables in LDS: bit
// j of t[c] is set when byte j of the word is c (gdb_like2_tables builds
// them). Words of 1..31 bytes (0 = empty word, always found). The row is read
// from LDS as aligned 32-bit words realigned with v_alignbyte, so every byte
// index below is a compile-time constant; a NUL byte ends the string as in
// gdb_like2. The row's LDS buffer must have 4 readable bytes past its end.
GDB_DEV void gdb_like2_tables(u32* t1, u32* t2, const char* __restrict__ w1, int l1, const char* __restrict__ w2, int l2) {
    for (int c = (int)threadIdx.x; c < 256; c += (int)blockDim.x) {
        u32 m1 = 0, m2 = 0;
        for (int j = 0; j < l1; ++j) m1 |= ((u32)(u8)w1[j] == (u32)c ? 1u : 0u) << j;
        for (int j = 0; j < l2; ++j) m2 |= ((u32)(u8)w2[j] == (u32)c ? 1u : 0u) << j;
        t1[c] = c ? m1 : 0u;
        t2[c] = c ? m2 : 0u;
    }
}
template <int W>
GDB_DEV bool gdb_like2_bp(const char* s, const u32* t1, const u32* t2, int l1, int l2) {
    constexpr int NW = (W + 3) / 4;
    const u32 off = (u32)(reinterpret_cast<unsigned long long>(s) & 3ull);
    const u32* wp = reinterpret_cast<const u32*>(s - off);
    u32 raw[NW + 1];
#pragma unroll
    for (int k = 0; k <= NW; ++k) raw[k] = wp[k];
    u32 a[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) a[k] = __builtin_amdgcn_alignbyte(raw[k + 1], raw[k], off);
    const u32 top1 = l1 > 0 ? 1u << (l1 - 1) : 0u, top2 = l2 > 0 ? 1u << (l2 - 1) : 0u;
    u32 d1 = 0, d2 = 0, alive = 1u;
    u32 found = l1 == 0 ? 1u : 0u;
    u32 hit = (l1 == 0 && l2 == 0) ? 1u : 0u;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const u32 c = (a[j >> 2] >> (8 * (j & 3))) & 0xFFu;
        alive &= c != 0u ? 1u : 0u;
        const u32 before = found;
        d1 = ((d1 << 1) | 1u) & t1[c];
        found |= (d1 & top1) ? alive : 0u;
        d2 = ((d2 << 1) | before) & t2[c];
        hit |= (d2 & top2) ? alive : 0u;
        if (l2 == 0) hit |= found;
    }
    return hit != 0u;
}

// LIKE 'p%' on a fixed-width NUL-padded field: the field starts with the l
// bytes of p (a field shorter than l, i.e. NUL inside the prefix, fails).
// The string compares here accumulate differences instead of returning at
// the first one, so a field read in place from HBM has all its byte loads in
// flight at once rather than one load latency per byte.
GDB_DEV bool gdb_like_prefix(const char* s, int width, const char* p, int l) {
    if (l > width) return false;
    u32 d = 0;
#pragma unroll 8
    for (int i = 0; i < l; ++i) d |= (u32)(u8)(s[i] ^ p[i]);
    return d == 0;
}

// The first two bytes of a char field as a little-endian code (byte 0 low),
// one 2-byte load (fields of even width are 2-byte aligned).
GDB_DEV i64 gdb_code2(const char* p) { return (i64)*reinterpret_cast<const unsigned short*>(p); }

// String equality of a fixed-width NUL-padded field with a constant held in
// a 32-byte slot: slot[0] is the length (0..31), the bytes follow.
GDB_DEV bool gdb_str_eq(const char* s, int width, const unsigned char* __restrict__ slot) {
    const int l = (int)slot[0];
    if (l > width) return false;
    u32 d = l == width ? 0u : (u32)(u8)s[l];
#pragma unroll 8
    for (int i = 0; i < l; ++i) d |= (u32)((unsigned char)s[i] ^ slot[1 + i]);
    return d == 0;
}

// 64-bit FNV-1a over n bytes, continuing from h (the offset basis
// 0xcbf29ce484222325 starts a fresh hash); gdb_fnv64_u32 folds the four
// little-endian bytes of v. A fingerprint of a multi-column group key.
GDB_DEV u64 gdb_fnv64(u64 h, const char* p, int n) {
    for (int i = 0; i < n; ++i) h = (h ^ (u64)(unsigned char)p[i]) * 0x100000001b3ull;
    return h;
}
GDB_DEV u64 gdb_fnv64_u32(u64 h, u32 v) {
    for (int i = 0; i < 4; ++i) h = (h ^ (u64)((v >> (8 * i)) & 255u)) * 0x100000001b3ull;
    return h;
}

GDB_DEV u64 gdb_fnv64_u64(u64 h, u64 v) {
    for (int i = 0; i < 8; ++i) h = (h ^ ((v >> (8 * i)) & 255ull)) * 0x100000001b3ull;
    return h;
}

// A NUL-padded field against a string literal of l bytes: equality, and
// the SQL ordering (bytes compared up to the shorter length, then length).
GDB_DEV bool gdb_str_eq_lit(const char* s, int width, const char* p, int l) {
    if (l > width) return false;
    u32 d = l == width ? 0u : (u32)(u8)s[l];
    for (int i = 0; i < l; ++i) d |= (u32)(u8)(s[i] ^ p[i]);
    return d == 0;
}
GDB_DEV int gdb_str_len(const char* s, int width) {
    int n = width;
    for (int i = width - 1; i >= 0; --i) n = s[i] == 0 ? i : n;
    return n;
}
GDB_DEV int gdb_str_cmp_lit(const char* s, int width, const char* p, int l) {
    const int n = gdb_str_len(s, width);
    const int m = n < l ? n : l;
    for (int i = 0; i < m; ++i) {
        const int d = (int)(u8)s[i] - (int)(u8)p[i];
        if (d != 0) return d;
    }
    return n - l;
}

// SQL LIKE of a NUL-padded field against a pattern of pl bytes with '%'
// (any run) and '_' (any one byte): greedy matching that backtracks to the
// last '%' (O(len * pl) at worst).
GDB_DEV bool gdb_like_match(const char* s, int width, const char* p, int pl) {
    const int n = gdb_str_len(s, width);
    int i = 0, j = 0, star = -1, mark = 0;
    while (i < n) {
        if (j < pl && p[j] != '%' && (p[j] == '_' || p[j] == s[i])) {
            ++i;
            ++j;
        } else if (j < pl && p[j] == '%') {
            star = j++;
            mark = i;
        } else if (star >= 0) {
            j = star + 1;
            i = ++mark;
        } else {
            return false;
        }
    }
    while (j < pl && p[j] == '%') ++j;
    return j == pl;
}

// The n (<= 8) bytes of a short fixed-width field as one integer (byte 0
// lowest): an exact group key.
GDB_DEV u64 gdb_pack_bytes(const char* p, int n) {
    u64 v = 0;
    for (int i = 0; i < n; ++i) v |= (u64)(u8)p[i] << (8 * i);
    return v;
}

// Byte equality of two n-byte fields.
GDB_DEV bool gdb_bytes_eq(const char* a, const char* b, int n) {
    u32 d = 0;
    for (int i = 0; i < n; ++i) d |= (u32)(u8)(a[i] ^ b[i]);
    return d == 0;
}

// LIKE '%p' on a fixed-width NUL-padded field: the string (up to its first
// NUL) ends with the l bytes of p.
GDB_DEV bool gdb_like_suffix(const char* s, int width, const char* p, int l) {
    int len = width;
    for (int i = width - 1; i >= 0; --i) len = s[i] == 0 ? i : len;
    if (l > len) return false;
    u32 d = 0;
#pragma unroll 8
    for (int i = 0; i < l; ++i) d |= (u32)(u8)(s[len - l + i] ^ p[i]);
    return d == 0;
}

// Wave-aggregated append: every active lane gets a distinct index from the
// counter; one atomic per wave (the first active lane's), the base is
// broadcast and each lane adds the number of active lanes below it.
GDB_DEV u64 gdb_wave_append(u64* counter) {
    const u64 mask = __ballot(1);
    const int leader = __builtin_ctzll(mask);
    const u32 lane = gdb_lane();
    u64 base = 0;
    if (lane == (u32)leader) base = gdb_atomic_add_u64_ret(counter, (u64)__builtin_popcountll(mask));
    base = __shfl(base, leader);
    return base + (u64)__builtin_popcountll(mask & ((1ull << lane) - 1ull));
}

// ---------------------------------------------------------------------------
// Top-k over 128-bit ranks (hi, lo), larger is better; rank (0,0) is reserved
// for "empty". Registers hold a sorted list; insertion bubbles one entry.
// ---------------------------------------------------------------------------
GDB_DEV bool gdb_rank_gt(u64 ah, u64 al, u64 bh, u64 bl) { return ah > bh || (ah == bh && al > bl); }

template <int K>
struct GdbTopK {
    u64 h[K];
    u64 l[K];
};
template <int K>
GDB_DEV void gdb_topk_init(GdbTopK<K>& t) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        t.h[j] = 0;
        t.l[j] = 0;
    }
}
template <int K>
GDB_DEV void gdb_topk_push(GdbTopK<K>& t, u64 h, u64 l) {
    if (gdb_rank_gt(h, l, t.h[K - 1], t.l[K - 1])) {
        t.h[K - 1] = h;
        t.l[K - 1] = l;
#pragma unroll
        for (int j = K - 1; j > 0; --j) {
            if (gdb_rank_gt(t.h[j], t.l[j], t.h[j - 1], t.l[j - 1])) {
                const u64 th = t.h[j], tl = t.l[j];
                t.h[j] = t.h[j - 1];
                t.l[j] = t.l[j - 1];
                t.h[j - 1] = th;
                t.l[j - 1] = tl;
            }
        }
    }
}
template <int K>
GDB_DEV void gdb_topk_pop(GdbTopK<K>& t) {
#pragma unroll
    for (int j = 0; j < K - 1; ++j) {
        t.h[j] = t.h[j + 1];
        t.l[j] = t.l[j + 1];
    }
    t.h[K - 1] = 0;
    t.l[K - 1] = 0;
}
// Wave-wide maximum of a u64, in every lane (the row steps of the sums
// above, then the four row maxima by readlane). All 64 lanes must be active.
GDB_DEV u64 gdb_wave_max_u64(u64 v) {
#define GDB_MAXU64_STEP(C)                       \
    {                                            \
        const u64 o = gdb_dpp64<C>(v);           \
        v = o > v ? o : v;                       \
    }
    GDB_MAXU64_STEP(0xB1)
    GDB_MAXU64_STEP(0x4E)
    GDB_MAXU64_STEP(0x124)
    GDB_MAXU64_STEP(0x128)
#undef GDB_MAXU64_STEP
    const u64 a = gdb_readlane64(v, 0), b = gdb_readlane64(v, 16);
    const u64 c = gdb_readlane64(v, 32), d = gdb_readlane64(v, 48);
    const u64 ab = a > b ? a : b, cd = c > d ? c : d;
    return ab > cd ? ab : cd;
}
// The wave's top K over every lane's sorted list: K rounds of a wave
// argmax over the list heads; the lane owning a round's maximum pops it
// (entries are unique: the low word carries the group key). A round
// compares high words only; the low words decide only between lanes whose
// high words tie with the maximum (read from the one such lane, or a second
// wave maximum over the tied lanes). Lane 0 writes round r's entry to
// out[2r], out[2r + 1].
template <int K>
GDB_DEV void gdb_topk_wave_select(GdbTopK<K>& t, u64* out) {
    for (int r = 0; r < K; ++r) {
        const u64 h = t.h[0], l = t.l[0];
        const u64 mh = gdb_wave_max_u64(h);
        const u64 tied = __ballot(h == mh);
        u64 ml;
        if (__builtin_popcountll(tied) == 1)
            ml = gdb_readlane64(l, __builtin_ctzll(tied));
        else
            ml = gdb_wave_max_u64(h == mh ? l : 0ull);
        if ((mh | ml) != 0 && h == mh && l == ml) gdb_topk_pop<K>(t);
        if (gdb_lane() == 0) {
            out[2 * r] = mh;
            out[2 * r + 1] = ml;
        }
    }
}
// Words of LDS gdb_topk_block_merge needs (blocks of up to 1024 threads).
#define GDB_TOPK_LDS_WORDS(K) (34 * (K))
// Block-wide top K: every wave selects its top K (no barrier), wave 0
// selects the block's top K from the waves' lists. On return every
// thread's list holds the block's top K, and res = lds + 32K holds it as
// (h, l) pairs. All threads of the block must call it.
template <int K>
GDB_DEV void gdb_topk_block_merge(GdbTopK<K>& t, u64* lds) {
    const int nw = (int)(blockDim.x >> 6);
    u64* res = lds + 32 * K;
    gdb_topk_wave_select<K>(t, lds + 2 * K * gdb_wave());
    __syncthreads();
    if (gdb_wave() == 0) {
        GdbTopK<K> m;
        gdb_topk_init<K>(m);
        for (int e = gdb_lane(); e < nw * K; e += 64) {
            const u64 h = lds[2 * e], l = lds[2 * e + 1];
            if ((h | l) != 0) gdb_topk_push<K>(m, h, l);
        }
        gdb_topk_wave_select<K>(m, res);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
        t.h[j] = res[2 * j];
        t.l[j] = res[2 * j + 1];
    }
    __syncthreads();
}
// ===== end gdb device runtime =====
)GDBHIP";

}  // namespace codegen
}  // namespace gdb

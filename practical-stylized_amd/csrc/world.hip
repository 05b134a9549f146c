// World mode: the reference's NL-PF (world-space hash-grid neighbour reuse with MIS/SMIS and
// stylization) for Lambertian path vertices, made deterministic.
//
// Per depth d = D..0 (PathFiltering::render, lm_path_filtering.rs:1673-1774):
//   gather   paths with a depth-d vertex, compacted in path order      (kernel_gather_points :507-531)
//   grid     Teschner-hashed cell buckets, built by a stable radix sort of (bucket, path) pairs
//            instead of the CAS push-front lists (HashGrid::build_kernel :230-255), with the
//            occupied buckets' entry ranges in a compact open-addressing table; a bucket's
//            entries are visited in ascending path order, every bucket of the 27-cell
//            neighbourhood is walked (duplicates from colliding cells included, :60-123)
//   cluster  (optional) deterministic seeding / joining                 (:532-623, 1675-1710)
//   filter   evaluate_filtered_radiance + stylization                   (:757-972, 999-1060)
//   propagate the stylized radiance to the parent's indirect record     (kernel_propagate :973-998)
// then kernel_splat + kernel_update_film (:428-506) as one per-pixel deterministic kernel.
#include <hipcub/hipcub.hpp>
#include <cmath>
#include "common.hpp"
#include "scene.hpp"

namespace nlpf {

#define WP NLPF_WORLD_PLANES
constexpr float kInvPi = 0.318309886183790671538f;
}  // namespace nlpf
#include "world_bsdf.hpp"
namespace nlpf {

struct Grid {
    float pmin[3], pmax[3];
    uint32_t res[3];
    float radius;
    uint64_t table;      // bucket space of the cell hash
    uint4* ctab;         // compact table of the occupied buckets, cmask + 1 slots
    uint32_t cmask, cshift;
};

// hash (lm_path_filtering.rs:20-23), u64 arithmetic
__device__ __forceinline__ uint32_t cell_bucket(uint32_t x, uint32_t y, uint32_t z, uint64_t table) {
    const uint64_t hsh = ((uint64_t)x * 73856093ull) ^ ((uint64_t)y * 19349663ull) ^ ((uint64_t)z * 83492791ull);
    return (uint32_t)(hsh % table);
}
// Compact bucket table. The cell hash keeps the reference's bucket space (hash_table_size
// buckets, so colliding cells still share a bucket and its entries), but only the depth's
// occupied buckets are stored: one 16-byte slot {bucket, start, end, seed} each, open addressing
// with linear probing from a Fibonacci hash of the bucket id, at most a quarter full. Lookups
// then touch a table of a few MB instead of 8 bytes x 64 Mi buckets.
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;  // bucket ids are below 2^30
__device__ __forceinline__ uint32_t ctab_home(const Grid& g, uint32_t b) { return (b * 2654435769u) >> g.cshift; }
// [start, end) of bucket b's entries, probing from slot s; (0, 0) when b holds none
__device__ __forceinline__ uint2 bucket_range_from(const Grid& g, uint32_t b, uint32_t s) {
    for (;;) {  // ends: the table always has empty slots
        const uint4 v = g.ctab[s];
        if (v.x == b) return make_uint2(v.y, v.z);
        if (v.x == kEmptyKey) return make_uint2(0u, 0u);
        s = (s + 1) & g.cmask;
    }
}
__device__ __forceinline__ uint2 bucket_range(const Grid& g, uint32_t b) { return bucket_range_from(g, b, ctab_home(g, b)); }

// HashGridVar::get_cell (:46-53)
__device__ __forceinline__ void get_cell(const Grid& g, float px, float py, float pz, uint32_t c[3]) {
    const float p[3] = {px, py, pz};
    for (int i = 0; i < 3; ++i) {
        float t = (p[i] - g.pmin[i]) / (g.pmax[i] - g.pmin[i]);
        t = fminf(fmaxf(t, 0.0f), 1.0f) * (float)g.res[i];
        uint32_t u = (uint32_t)t;
        c[i] = u < g.res[i] - 1 ? u : g.res[i] - 1;
    }
}

// Lookups of the 27-cell walk around `cell` (cell c = 9 (dz+1) + 3 (dy+1) + (dx+1), dx inner),
// split so that a thread can issue the first probes of several cells as one round of loads:
// first_probe loads the home slot of cell c's bucket (bk = its bucket, kEmptyKey outside the
// grid, which then matches the empty slot it returns); probe_range finishes the lookup.
__device__ __forceinline__ uint4 first_probe(const Grid& g, const uint32_t cell[3], int c, uint32_t& bk) {
    const int cx = (int)cell[0] + c % 3 - 1, cy = (int)cell[1] + (c / 3) % 3 - 1, cz = (int)cell[2] + c / 9 - 1;
    bk = kEmptyKey;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)g.res[0] || cy >= (int)g.res[1] || cz >= (int)g.res[2])
        return make_uint4(kEmptyKey, 0u, 0u, 0u);
    bk = cell_bucket((uint32_t)cx, (uint32_t)cy, (uint32_t)cz, g.table);
    return g.ctab[ctab_home(g, bk)];
}
__device__ __forceinline__ uint2 probe_range(const Grid& g, const uint4 v, uint32_t bk) {
    if (v.x == bk) return make_uint2(v.y, v.z);
    if (v.x == kEmptyKey) return make_uint2(0u, 0u);
    return bucket_range_from(g, bk, (ctab_home(g, bk) + 1) & g.cmask);
}

// |d| without fused multiply-adds, so the radius test rounds exactly like plain IEEE code
__device__ __forceinline__ float dist3(float x, float y, float z) {
#pragma clang fp contract(off)
    return sqrtf((x * x + y * y) + z * z);
}

template <typename T>
__device__ __forceinline__ void kahan_add(T& sum, T& comp, T v) {
    T y = v - comp;
    T t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

__device__ __forceinline__ float pl(const float* rec, uint64_t N, int c, uint64_t i) { return rec[(uint64_t)c * N + i]; }

// Wavefront-wide max / sum (every lane of the wavefront must reach the call): a kernel then
// issues one atomic per wavefront on a shared counter instead of one per lane, which the
// memory system would serialize.
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
    return v;
}

// Counters that every wavefront of a launch updates (neighbour counts, occupied buckets, bucket
// and candidate maxima) are spread over NBR_SLOTS addresses, slot blockIdx % NBR_SLOTS: atomics
// on one address serialize (about 13 ns each, tens of microseconds per launch); the host sums
// or maxes the slots in the readback it makes anyway.
constexpr uint32_t NBR_SLOTS = 256;

// ---- the points of every depth, selected once per pass ----
// Depth d's points are the paths whose last depth reaches d (a last depth beyond the pass's
// levels counts as the deepest level), in path order. Paths are taken in blocks of kSelBlock:
// w_sel_count_kernel counts each block's points per depth, w_sel_offsets_kernel scans the counts
// of each depth over the blocks (one workgroup) and totals them, and w_sel_scatter_kernel writes
// each path into the list of every depth it reaches, at its block's offset plus its rank in the
// block. Depth d's list starts at base[d] = the points of the depths below d.
constexpr uint32_t kSelBlock = 1024;  // paths per block: 256 threads x 4, path b * 1024 + 256 j + thread
struct SelBases {
    uint32_t base[32];
};
__device__ __forceinline__ int32_t sel_depth(const int32_t* path_depth, uint64_t N, uint64_t t, int32_t last) {
    if (t >= N) return -1;
    const int32_t k = path_depth[t];
    return k < last ? k : last;  // k < 0: a path without vertices
}
// cnt[d * nblk + b] = points of block b at depth d
__global__ void __launch_bounds__(256) w_sel_count_kernel(const int32_t* path_depth, uint64_t N, uint32_t L,
                                                          uint32_t* cnt, uint32_t nblk) {
    __shared__ uint32_t wc[32][4];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int32_t k[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        k[j] = sel_depth(path_depth, N, (uint64_t)blockIdx.x * kSelBlock + 256u * j + threadIdx.x, (int32_t)L - 1);
    for (uint32_t d = 0; d < L; ++d) {
        uint32_t c = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) c += (uint32_t)__popcll(__ballot(k[j] >= (int32_t)d));
        if (lane == 0) wc[d][wv] = c;
    }
    __syncthreads();
    if (threadIdx.x < L) cnt[(uint64_t)threadIdx.x * nblk + blockIdx.x] = wc[threadIdx.x][0] + wc[threadIdx.x][1] +
                                                                          wc[threadIdx.x][2] + wc[threadIdx.x][3];
}
// cnt[d * nblk + b] -> exclusive prefix over the blocks; total[d] = points at depth d
__global__ void __launch_bounds__(1024) w_sel_offsets_kernel(uint32_t* cnt, uint32_t nblk, uint32_t L,
                                                             uint32_t* total) {
    __shared__ uint32_t wave_total[16];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t d = 0; d < L; ++d) {
        uint32_t running = 0;  // uniform
        for (uint32_t b0 = 0; b0 < nblk; b0 += 1024) {
            const uint32_t b = b0 + threadIdx.x;
            const uint32_t v = b < nblk ? cnt[(uint64_t)d * nblk + b] : 0u;
            uint32_t incl = v;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t u = (uint32_t)__shfl_up((int)incl, off);
                if (lane >= (uint32_t)off) incl += u;
            }
            if (lane == 63) wave_total[wv] = incl;
            __syncthreads();
            uint32_t before = running, chunk = 0;
            for (uint32_t k = 0; k < 16; ++k) {
                if (k < wv) before += wave_total[k];
                chunk += wave_total[k];
            }
            if (b < nblk) cnt[(uint64_t)d * nblk + b] = before + incl - v;
            running += chunk;
            __syncthreads();
        }
        if (threadIdx.x == 0) total[d] = running;
    }
}
__global__ void __launch_bounds__(256) w_sel_scatter_kernel(const int32_t* path_depth, uint64_t N, uint32_t L,
                                                            const uint32_t* off, uint32_t nblk, SelBases sb,
                                                            uint32_t* pts) {
    __shared__ uint32_t wc[4][4];  // [item j][wave]: this depth's points
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t t0 = (uint64_t)blockIdx.x * kSelBlock + threadIdx.x;
    int32_t k[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) k[j] = sel_depth(path_depth, N, t0 + 256u * j, (int32_t)L - 1);
    for (uint32_t d = 0; d < L; ++d) {
        uint64_t m[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            m[j] = __ballot(k[j] >= (int32_t)d);
            if (lane == 0) wc[j][wv] = (uint32_t)__popcll(m[j]);
        }
        __syncthreads();
        // rank of item (j, wave, lane) = the block's points before it in path order
        uint32_t r = sb.base[d] + off[(uint64_t)d * nblk + blockIdx.x];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t at = r;
            for (uint32_t w = 0; w < wv; ++w) at += wc[j][w];
            if (k[j] >= (int32_t)d) pts[at + (uint32_t)__popcll(m[j] & below)] = (uint32_t)(t0 + 256u * j);
            r += wc[j][0] + wc[j][1] + wc[j][2] + wc[j][3];
        }
        __syncthreads();
    }
}

// ---- grid build: a counting sort on the compact table's slots ----
// Every point finds or claims its bucket's slot and counts itself there, keeping the count it
// saw (its arrival rank); w_slot_ranges_kernel turns the counts into [start, end) ranges;
// w_scatter_kernel puts each point at start + arrival rank; w_emit_kernel then moves each entry to
// start + (entries of its bucket with a smaller point index), the order of a stable sort by
// bucket (points are in path-id order). Buckets take their ranges in whichever order the slot
// pass claims them: only the order inside a bucket reaches a result, since every walk goes
// through a bucket's range. The emit pass costs each entry a pass over its bucket, so buckets
// above kEmitMaxBucket entries take the radix-sort path instead (w_sorted_emit_kernel).
constexpr uint32_t kEmitMaxBucket = 1024;

// empty compact table (no bucket, no entries, no seed), zeroed grid counters and range cursor
__global__ void w_grid_clear_kernel(uint4* ctab, uint32_t slots, uint32_t* counters, uint32_t ncounters,
                                    uint32_t* cursor) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < slots) ctab[s] = make_uint4(kEmptyKey, 0u, 0u, kEmptyKey);
    if (s < ncounters) counters[s] = 0u;
    if (s == 0) *cursor = 0u;
}

// slot_of[t] = slot of point t's bucket, arrival[t] = points of that bucket counted before it.
// While the table is built a slot's first two words are one 64-bit word {bucket, count}: a
// point claims an empty slot with count 1 in one compare-and-swap, and a point whose swap finds
// its bucket there adds 1 to the count (two atomics; no plain probe loads). Claimed slots count
// the occupied buckets and the largest arrival rank bounds the largest bucket (both over
// NBR_SLOTS slots). With clustering, the points also start unclustered (no cluster, no joiners).
__global__ void __launch_bounds__(256) w_insert_kernel(const float* rec, uint64_t N, const uint32_t* pts, uint32_t n,
                                                       Grid g, uint32_t* slot_of, uint32_t* arrival,
                                                       uint32_t* occupied, uint32_t* max_bucket, uint32_t* cluster_id,
                                                       uint32_t* cluster_size) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t claimed = 0, size = 0;
    if (t < n) {
        const uint32_t i = pts[t];
        if (cluster_id) {
            cluster_id[i] = 0xFFFFFFFFu;
            cluster_size[i] = 0u;
        }
        uint32_t c[3];
        get_cell(g, pl(rec, N, 0, i), pl(rec, N, 1, i), pl(rec, N, 2, i), c);
        const uint32_t k = cell_bucket(c[0], c[1], c[2], g.table);
        const unsigned long long empty = (unsigned long long)kEmptyKey, mine = (1ull << 32) | k;
        uint32_t s = ctab_home(g, k), r = 0;
        for (;;) {  // ends: the table is at most a quarter full
            unsigned long long* w = reinterpret_cast<unsigned long long*>(&g.ctab[s]);
            const unsigned long long old = atomicCAS(w, empty, mine);
            if ((uint32_t)old == kEmptyKey) {
                claimed = 1;
                break;
            }
            if ((uint32_t)old == k) {
                r = (uint32_t)(atomicAdd(w, 1ull << 32) >> 32);
                break;
            }
            s = (s + 1) & g.cmask;
        }
        slot_of[t] = s;
        arrival[t] = r;
        size = r + 1;
    }
    claimed = wave_sum(claimed);
    size = wave_max(size);
    if ((threadIdx.x & 63) == 0) {
        const uint32_t slot = blockIdx.x & (NBR_SLOTS - 1);
        if (claimed) atomicAdd(occupied + slot, claimed);
        if (size) atomicMax(max_bucket + slot, size);
    }
}

// bucket ranges from the slot counts: a workgroup takes 4096 slots (lane-interleaved, 16 per
// thread), scans their counts, claims room for its total from *cursor (one atomic per
// workgroup) and writes each occupied slot's [start, end) over {count, 0}
constexpr uint32_t kRangeSlotsPerThread = 16;
__global__ void __launch_bounds__(256) w_slot_ranges_kernel(uint4* ctab, uint64_t slots, uint32_t* cursor) {
    __shared__ uint32_t wave_total[4], block_base;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t s0 = (uint64_t)blockIdx.x * 256 * kRangeSlotsPerThread + threadIdx.x;
    uint32_t cnt[kRangeSlotsPerThread];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRangeSlotsPerThread; ++k) {
        const uint64_t s = s0 + (uint64_t)k * 256;
        cnt[k] = s < slots ? ctab[s].y : 0u;
        sum += cnt[k];
    }
    uint32_t incl = sum;  // inclusive scan over the wavefront
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= (uint32_t)off) incl += v;
    }
    if (lane == 63) wave_total[wv] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = wave_total[0] + wave_total[1] + wave_total[2] + wave_total[3];
        block_base = total ? atomicAdd(cursor, total) : 0u;
    }
    __syncthreads();
    uint32_t start = block_base + incl - sum;
    for (uint32_t k = 0; k < wv; ++k) start += wave_total[k];
#pragma unroll
    for (uint32_t k = 0; k < kRangeSlotsPerThread; ++k) {
        if (cnt[k]) {
            const uint64_t s = s0 + (uint64_t)k * 256;
            ctab[s].y = start;
            ctab[s].z = start + cnt[k];
            start += cnt[k];
        }
    }
}

// order[start of t's bucket + arrival rank of t] = t
__global__ void w_scatter_kernel(const uint32_t* slot_of, const uint32_t* arrival, uint32_t n, const uint4* ctab,
                                 uint32_t* order) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) order[ctab[slot_of[t]].y + arrival[t]] = t;
}

// per-entry arrays at the entry's place in bucket order: its bucket's table slot, path id,
// position with the path id, stylization flags
__device__ __forceinline__ void emit_entry(uint32_t f, uint32_t slot, uint32_t i, const float* rec, uint64_t N,
                                           const uint8_t* new_flags, uint32_t* keys_sorted, uint32_t* ids_sorted,
                                           float4* pos_sorted, uint8_t* flags_sorted) {
    keys_sorted[f] = slot;
    ids_sorted[f] = i;
    pos_sorted[f] = make_float4(pl(rec, N, 0, i), pl(rec, N, 1, i), pl(rec, N, 2, i), __uint_as_float(i));
    flags_sorted[f] = new_flags[i];
}

// entry e of the scatter holds point t = order[e]; its place is its bucket's start plus the
// bucket's entries with a smaller point index (buckets of at most kEmitMaxBucket entries)
__global__ void w_emit_kernel(const uint32_t* order, const uint32_t* slot_of, const uint32_t* pts, uint32_t n,
                              const uint4* ctab, const float* rec, uint64_t N, const uint8_t* new_flags,
                              uint32_t* keys_sorted, uint32_t* ids_sorted, float4* pos_sorted, uint8_t* flags_sorted) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t t = order[e];
    const uint32_t s = slot_of[t];
    const uint4 v = ctab[s];
    uint32_t f = v.y;
    for (uint32_t q = v.y; q < v.z; ++q) f += order[q] < t ? 1u : 0u;
    emit_entry(f, s, pts[t], rec, N, new_flags, keys_sorted, ids_sorted, pos_sorted, flags_sorted);
}

// the radix-sort path: entries already sorted by slot (stable, so in path-id order inside a
// bucket); the last entry of each bucket writes its [start, end) over the bucket's count
__global__ void w_sorted_emit_kernel(const uint32_t* slot_sorted, const uint32_t* ids_sorted_in, uint32_t n,
                                     uint4* ctab, const float* rec, uint64_t N, const uint8_t* new_flags,
                                     uint32_t* keys_sorted, float4* pos_sorted, uint8_t* flags_sorted) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t s = slot_sorted[e];
    const uint32_t i = ids_sorted_in[e];
    keys_sorted[e] = s;
    pos_sorted[e] = make_float4(pl(rec, N, 0, i), pl(rec, N, 1, i), pl(rec, N, 2, i), __uint_as_float(i));
    flags_sorted[e] = new_flags[i];
    if (e + 1 == n || slot_sorted[e + 1] != s) {
        const uint32_t count = ctab[s].y;
        ctab[s].y = e + 1 - count;
        ctab[s].z = e + 1;
    }
}

// largest cluster (seed + joiners) of the depth, max over out_max[NBR_SLOTS]. With clustering a
// vertex's candidates are members of its own cluster, but a bucket is walked once per visited
// cell that hashes to it, so its members can be counted up to 27 times: 27 x this value bounds
// every candidate count
__global__ void w_cluster_max_kernel(const uint32_t* pts, uint32_t n, const uint32_t* cluster_id,
                                     const uint32_t* cluster_size, uint32_t* out_max) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t members = 0;
    if (t < n) {
        const uint32_t i = pts[t];
        if (cluster_id[i] == i) members = cluster_size[i] + 1u;
    }
    members = wave_max(members);
    if ((threadIdx.x & 63) == 0 && members) atomicMax(out_max + (blockIdx.x & (NBR_SLOTS - 1)), members);
}

// per-path payload of level d (indirect record from the previous propagation): paths whose last
// depth is above ind_above took one at the previous level, the others have none (zeros; the
// indirect planes hold nothing of theirs from this pass)
__global__ void w_pack_kernel(const float* rec, uint64_t N, const uint32_t* pts, uint32_t n, const float* ind,
                              const int32_t* path_depth, int32_t ind_above, const uint8_t* new_flags,
                              const uint16_t* material, NbrRec* pk) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = pts[t];
    NbrRec r;
    r.dwi_pdf = make_float4(pl(rec, N, 19, i), pl(rec, N, 20, i), pl(rec, N, 21, i), pl(rec, N, 25, i));
    r.dirr_mis = make_float4(pl(rec, N, 22, i), pl(rec, N, 23, i), pl(rec, N, 24, i), pl(rec, N, 26, i));
    r.iwi_mis = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r.iirr = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float((uint32_t)new_flags[i]));
    if (path_depth[i] > ind_above) {
        r.iwi_mis = make_float4(ind[i], ind[N + i], ind[2 * N + i], ind[6 * N + i]);
        r.iirr = make_float4(ind[3 * N + i], ind[4 * N + i], ind[5 * N + i], r.iirr.w);
    }
    r.ns_wox = make_float4(pl(rec, N, 3, i), pl(rec, N, 4, i), pl(rec, N, 5, i), pl(rec, N, 9, i));
    r.ng_woy = make_float4(pl(rec, N, 6, i), pl(rec, N, 7, i), pl(rec, N, 8, i), pl(rec, N, 10, i));
    r.woz = make_float4(pl(rec, N, 11, i), __uint_as_float((uint32_t)material[i]), 0.0f, 0.0f);
    pk[i] = r;
}


struct FilterArgs {
    const float* rec;
    uint64_t N;
    const uint32_t* pts;
    uint32_t off, count;
    const uint8_t* new_flags;
    const uint16_t* material;
    const uint32_t* cluster_id;  // nullptr when clustering is off
    const NbrRec* pk;
    const float4* pos_sorted;
    const uint8_t* flags_sorted;  // new_flags of each entry's path, in entry order
    Grid g;
    uint32_t K;
    uint32_t* nbuf;
    uint32_t stride;  // list slots per vertex in nbuf
    Pcg32* rng;
    int use_rng;      // 0 when no vertex of the launch can exceed the reuse cap: no draws
    const nlpf_style_node* nodes;
    const int32_t* roots;
    uint32_t n_materials;
    const WBsdf* bsdf;  // BSDF by material (n_bsdf entries; diffuse beyond); read by the GLOSSY kernels
    uint32_t n_bsdf;
    int depth;
    int smis, flag_check, stylize;
    float* rad_out;     // [N][3] level radiance (v.radiance)
    unsigned long long* nbr_count;   // NBR_SLOTS counters of this depth (spread to avoid contention)
    unsigned long long* pair_count;  // NBR_SLOTS counters of SMIS pdf evaluations
    // SMIS denominators shared within a cluster (wave kernel, clustering on): 0 = each vertex
    // sums its own; 1 = build: the first entry of each uncapped list sums them for the whole
    // list into smis_sum (no radiance); 2 = use smis_sum for uncapped lists
    int smis_phase;
    double* smis_sum;  // [N] sum over the list of cos_m(wi_j), by reused vertex j
};

// The centre vertex's BSDF for a neighbour's direction: the pdf, and f |cos_i| per channel with
// the base colour R included. Without GLOSSY every material is diffuse.
template <bool GLOSSY>
struct CentreBsdf {
    float3 R, ns, ng, wo;
    WBsdf b;
    GlossCentre gc;
    __device__ __forceinline__ CentreBsdf(const FilterArgs& a, uint32_t c, float3 R_, float3 ns_, float3 ng_,
                                          float3 wo_)
        : R(R_), ns(ns_), ng(ng_), wo(wo_) {
        b.kind = NLPF_BSDF_DIFFUSE;
        if (GLOSSY) {
            const uint32_t m = a.material[c];
            if (m < a.n_bsdf) b = a.bsdf[m];
            if (b.kind == NLPF_BSDF_PRINCIPLED) gc = gloss_centre(b, R);
        }
    }
    __device__ __forceinline__ float eval(float3 wi, float3& fc) const {
        if (!GLOSSY || b.kind == NLPF_BSDF_DIFFUSE) {
            float fcos;
            const float pdf = diffuse_eval(ns, ng, wo, wi, fcos);
            fc = make_float3(R.x * fcos, R.y * fcos, R.z * fcos);
            return pdf;
        }
        if (b.kind == NLPF_BSDF_PRINCIPLED) return principled_eval(b, gc, R, ns, ng, wo, wi, fc);
        fc = make_float3(0.0f, 0.0f, 0.0f);  // emission only
        return 0.0f;
    }
};

// A reused neighbour's SMIS record (and, for a principled one, its SmisGloss part): an
// emission-only neighbour has pdf_m = 0 for every direction (its wo bit cleared)
template <bool GLOSSY>
__device__ __forceinline__ SmisRec smis_rec_of(const FilterArgs& a, const NbrRec& o, SmisGloss& g) {
    SmisRec r = smis_rec(o);
    if (GLOSSY) {
        const uint32_t m = __float_as_uint(o.woz.y);
        if (m < a.n_bsdf) {
            const WBsdf b = a.bsdf[m];
            if (b.kind == NLPF_BSDF_PRINCIPLED) {
                g = smis_gloss(o, b, r.co);
                r.bits |= SMIS_GLOSSY;
            } else if (b.kind == NLPF_BSDF_EMISSION) {
                r.bits &= ~SMIS_WO_OK;
            }
        }
    }
    return r;
}
template <bool GLOSSY>
__device__ __forceinline__ double smis_term(double sum, const SmisRec& r, const SmisGloss& g, uint32_t bits, float3 wi,
                                            double wx, double wy, double wz) {
    if (GLOSSY && (bits & SMIS_GLOSSY)) return smis_add_gloss(sum, r, g, bits, wi, wx, wy, wz);
    if (bits & SMIS_SAME) return smis_add_same(sum, r, bits, wi, wx, wy, wz);
    return smis_add(sum, r, bits, wi, wx, wy, wz);
}

// Filter vertex tl of the launch on one thread; returns its candidate count and list length.
template <bool GLOSSY>
__device__ __forceinline__ void filter_vertex_thread(const FilterArgs& a, uint32_t tl, uint32_t& cnt_out,
                                                     uint32_t& used_out) {
    const uint64_t N = a.N;
    const uint32_t c = a.pts[a.off + tl];
    const float3 p = make_float3(pl(a.rec, N, 0, c), pl(a.rec, N, 1, c), pl(a.rec, N, 2, c));
    const float3 ns = make_float3(pl(a.rec, N, 3, c), pl(a.rec, N, 4, c), pl(a.rec, N, 5, c));
    const float3 ng = make_float3(pl(a.rec, N, 6, c), pl(a.rec, N, 7, c), pl(a.rec, N, 8, c));
    const float3 wo = make_float3(pl(a.rec, N, 9, c), pl(a.rec, N, 10, c), pl(a.rec, N, 11, c));
    const float3 R = make_float3(pl(a.rec, N, 12, c), pl(a.rec, N, 13, c), pl(a.rec, N, 14, c));
    const float3 le = make_float3(pl(a.rec, N, 15, c), pl(a.rec, N, 16, c), pl(a.rec, N, 17, c));
    const uint32_t myflags = a.new_flags[c];
    const uint32_t mycluster = a.cluster_id ? a.cluster_id[c] : 0u;
    Pcg32 rng = a.use_rng ? a.rng[tl] : Pcg32{0ull, 1ull};
    uint32_t* buf = a.nbuf + (uint64_t)tl * a.stride;
    const CentreBsdf<GLOSSY> cb(a, c, R, ns, ng, wo);

    float3 q = p;
    if (a.cluster_id) q = make_float3(pl(a.rec, N, 0, mycluster), pl(a.rec, N, 1, mycluster), pl(a.rec, N, 2, mycluster));
    uint32_t cell[3];
    get_cell(a.g, q.x, q.y, q.z, cell);
    uint32_t cnt = 0;
    // the first two list entries stay in registers (at small radii most lists hold one or two)
    uint32_t b0 = 0, b1 = 0;
    auto put = [&](uint32_t i, uint32_t j) {
        if (i == 0) b0 = j;
        else if (i == 1) b1 = j;
        else buf[i] = j;
    };
    auto visit = [&](const float4 qe, uint32_t fl) {
        if (!(dist3(qe.x - q.x, qe.y - q.y, qe.z - q.z) <= a.g.radius)) return;
        const uint32_t j = __float_as_uint(qe.w);
        if (a.cluster_id && a.cluster_id[j] != mycluster) return;
        if (a.flag_check && fl != myflags) return;
        // reservoir over the reuse cap (:813-821)
        if (cnt < a.K) {
            put(cnt, j);
        } else {
            const uint32_t r = pcg_next(rng) % (cnt + 1);
            if (r < a.K) put(r, j);
        }
        ++cnt;
    };
    // nine cells at a time: their bucket ranges in one round of loads, then their first entries
    // (most buckets hold one or two at small radii) in one more, then the walk in cell order
#pragma unroll
    for (int g0 = 0; g0 < 27; g0 += 9) {
        uint32_t r0[9], r1[9], bk[9];
        uint4 v[9];
#pragma unroll
        for (int u = 0; u < 9; ++u) v[u] = first_probe(a.g, cell, g0 + u, bk[u]);
#pragma unroll
        for (int u = 0; u < 9; ++u) {
            const uint2 br = probe_range(a.g, v[u], bk[u]);
            r0[u] = br.x;
            r1[u] = br.y;
        }
        float4 qf[9];
        uint32_t ff[9];
#pragma unroll
        for (int u = 0; u < 9; ++u) {
            qf[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            ff[u] = 0;
            if (r0[u] < r1[u]) {
                qf[u] = a.pos_sorted[r0[u]];
                ff[u] = a.flags_sorted[r0[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < 9; ++u) {
            if (r0[u] < r1[u]) {
                visit(qf[u], ff[u]);
                for (uint32_t e = r0[u] + 1; e < r1[u]; ++e) visit(a.pos_sorted[e], a.flags_sorted[e]);
            }
        }
    }
    const uint32_t used = cnt < a.K ? cnt : a.K;
    cnt_out = cnt;
    used_out = used;
    const float inv_used = 1.0f / (float)used;
    // hundreds of reused terms per vertex: Kahan-compensated sums keep FP32 within ~1 ulp of
    // the exact sum (a plain running sum drifts by ~n ulp)
    float dsum[3] = {0, 0, 0}, dc[3] = {0, 0, 0}, isum[3] = {0, 0, 0}, ic[3] = {0, 0, 0};
    auto get = [&](uint32_t i) { return i == 0 ? b0 : i == 1 ? b1 : buf[i]; };
    for (uint32_t k = 0; k < used; ++k) {
        const uint32_t j = get(k);
        const NbrRec& r = a.pk[j];
        const float4 dw = r.dwi_pdf;
        if (dw.w > 0.0f) {
            float3 fc;
            const float bpdf = cb.eval(make_float3(dw.x, dw.y, dw.z), fc);
            const float4 di = r.dirr_mis;
            const float w = dw.w / (dw.w + bpdf);  // mis_weight, balance heuristic (:305-316)
            const float sc = di.w * w / dw.w * inv_used;
            kahan_add(dsum[0], dc[0], fc.x * sc * di.x);
            kahan_add(dsum[1], dc[1], fc.y * sc * di.y);
            kahan_add(dsum[2], dc[2], fc.z * sc * di.z);
        }
        const float4 iw = r.iwi_mis;
        const float3 wi = make_float3(iw.x, iw.y, iw.z);
        if (sqrtf(wi.x * wi.x + wi.y * wi.y + wi.z * wi.z) == 0.0f) continue;
        float3 fc;
        const float bpdf = cb.eval(wi, fc);
        const float4 ir = r.iirr;
        const float3 reused = make_float3(fc.x * ir.x * iw.w, fc.y * ir.y * iw.w, fc.z * ir.z * iw.w);
        if (a.smis) {
            // sum_m pdf_m(wi) over every reused neighbour m, in order, in FP64 (in units of
            // pi * pdf: a diffuse term is |cos_i|)
            const double wx = wi.x, wy = wi.y, wz = wi.z;
            double sum_cos = 0.0;
            for (uint32_t m = 0; m < used; ++m) {
                SmisGloss g;
                const SmisRec r = smis_rec_of<GLOSSY>(a, a.pk[get(m)], g);
                sum_cos = smis_term<GLOSSY>(sum_cos, r, g, r.bits, wi, wx, wy, wz);
            }
            const float sum_pdf = (float)(sum_cos * 0.318309886183790671538);
            kahan_add(isum[0], ic[0], reused.x / sum_pdf);
            kahan_add(isum[1], ic[1], reused.y / sum_pdf);
            kahan_add(isum[2], ic[2], reused.z / sum_pdf);
        } else if (bpdf > 0.0f) {
            kahan_add(isum[0], ic[0], reused.x / bpdf * inv_used);
            kahan_add(isum[1], ic[1], reused.y / bpdf * inv_used);
            kahan_add(isum[2], ic[2], reused.z / bpdf * inv_used);
        }
    }
    const float3 direct = make_float3(dsum[0], dsum[1], dsum[2]);
    const float3 indirect = make_float3(isum[0], isum[1], isum[2]);
    v3<float> L{le.x + direct.x + indirect.x, le.y + direct.y + indirect.y, le.z + direct.z + indirect.z};
    if (a.stylize) {
        const uint32_t m = a.material[c];
        const int root = m < a.n_materials ? a.roots[m] : -1;
        if (root >= 0) L = apply_style<float>(a.nodes, root, (uint32_t)a.depth, myflags, L, v3<float>{p.x, p.y, p.z});
    }
    a.rad_out[3 * (uint64_t)c] = L.x;
    a.rad_out[3 * (uint64_t)c + 1] = L.y;
    a.rad_out[3 * (uint64_t)c + 2] = L.z;
    if (a.use_rng) a.rng[tl] = rng;
}

// One thread per vertex: for small neighbour sets (a few per vertex), where a wavefront per
// vertex would leave most lanes idle. The neighbour and SMIS-pair counters are summed over the
// wavefront first: one atomic per wavefront and counter.
template <bool GLOSSY>
__global__ void __launch_bounds__(256) w_filter_thread_kernel(FilterArgs a) {
    const uint32_t tl = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cnt = 0, used = 0;
    if (tl < a.count) filter_vertex_thread<GLOSSY>(a, tl, cnt, used);
    unsigned long long nc = cnt, np = a.smis ? (unsigned long long)used * used : 0ull;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        nc += __shfl_xor(nc, off);
        np += __shfl_xor(np, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (nc) atomicAdd(a.nbr_count + (blockIdx.x & (NBR_SLOTS - 1)), nc);
        if (np) atomicAdd(a.pair_count + (blockIdx.x & (NBR_SLOTS - 1)), np);
    }
}

// One wavefront per vertex (4 per workgroup).
//   gather   the 27-cell walk is uniform across the wave; the lanes test 64 bucket entries at a
//            time, a ballot and a prefix count append the accepted ones in bucket order, and
//            lane 0 replays the reservoir draws (:813-821) for the entries beyond the cap, in
//            order, so the RNG stream is the sequential one; the first WF_CAP list entries
//            live in the wave's LDS slice, the rest in the global list buffer;
//   reuse    lanes take neighbours k = lane, lane + 64, ...; for the SMIS denominators the
//            reused neighbours' pdf inputs are staged in LDS 64 at a time, and every lane sums
//            pdf_m(wi_k) over them in order, in FP64, with a wave-uniform index m;
//   reduce   per-lane Kahan sums, then a fixed shuffle tree (deterministic).
constexpr uint32_t WF_CAP = 1024;
template <bool GLOSSY>
__global__ void __launch_bounds__(256) w_filter_wave_kernel(FilterArgs a) {
    __shared__ uint32_t lbuf[4][WF_CAP];
    __shared__ SmisRec sbuf[4][64];
    __shared__ SmisGloss gsbuf[GLOSSY ? 4 : 1][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t tl = blockIdx.x * 4 + wv;
    if (tl >= a.count) return;  // uniform per wave; no workgroup barrier below
    const uint64_t N = a.N;
    const uint32_t c = a.pts[a.off + tl];
    const float3 p = make_float3(pl(a.rec, N, 0, c), pl(a.rec, N, 1, c), pl(a.rec, N, 2, c));
    const float3 ns = make_float3(pl(a.rec, N, 3, c), pl(a.rec, N, 4, c), pl(a.rec, N, 5, c));
    const float3 ng = make_float3(pl(a.rec, N, 6, c), pl(a.rec, N, 7, c), pl(a.rec, N, 8, c));
    const float3 wo = make_float3(pl(a.rec, N, 9, c), pl(a.rec, N, 10, c), pl(a.rec, N, 11, c));
    const float3 R = make_float3(pl(a.rec, N, 12, c), pl(a.rec, N, 13, c), pl(a.rec, N, 14, c));
    const float3 le = make_float3(pl(a.rec, N, 15, c), pl(a.rec, N, 16, c), pl(a.rec, N, 17, c));
    const uint32_t myflags = a.new_flags[c];
    const uint32_t mycluster = a.cluster_id ? a.cluster_id[c] : 0u;
    Pcg32 rng = a.use_rng ? a.rng[tl] : Pcg32{0ull, 1ull};  // advanced by lane 0 only
    uint32_t* gbuf = a.nbuf + (uint64_t)tl * a.stride;
    uint32_t* lb = lbuf[wv];
    SmisRec* sb = sbuf[wv];
    SmisGloss* sg = gsbuf[GLOSSY ? wv : 0];
    const CentreBsdf<GLOSSY> cb(a, c, R, ns, ng, wo);
    auto put = [&](uint32_t i, uint32_t j) {
        if (i < WF_CAP)
            lb[i] = j;
        else
            gbuf[i] = j;
    };

    float3 q = p;
    if (a.cluster_id) q = make_float3(pl(a.rec, N, 0, mycluster), pl(a.rec, N, 1, mycluster), pl(a.rec, N, 2, mycluster));
    uint32_t cell[3];
    get_cell(a.g, q.x, q.y, q.z, cell);
    uint32_t cnt = 0;  // wave-uniform
    const uint64_t below = (1ull << lane) - 1ull;
    // lane c < 27 owns cell c of the walk (dz outer, dx inner) and loads its bucket range
    uint32_t c_e0 = 0, c_n = 0;
    if (lane < 27) {
        const int cx = (int)cell[0] + lane % 3 - 1, cy = (int)cell[1] + (lane / 3) % 3 - 1,
                  cz = (int)cell[2] + lane / 9 - 1;
        if (cx >= 0 && cy >= 0 && cz >= 0 && cx < (int)a.g.res[0] && cy < (int)a.g.res[1] && cz < (int)a.g.res[2]) {
            const uint32_t b = cell_bucket((uint32_t)cx, (uint32_t)cy, (uint32_t)cz, a.g.table);
            const uint2 br = bucket_range(a.g, b);
            c_e0 = br.x;
            c_n = br.y - br.x;
        }
    }
    // inclusive prefix of the cell sizes over the lanes: the walk's candidates, cell by cell
    // and in bucket order within a cell, are the flat sequence 0 .. total-1
    uint32_t c_end = c_n;
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)c_end, off);
        if (lane >= off) c_end += v;
    }
    const uint32_t total_cand = (uint32_t)__shfl((int)c_end, 26);
    for (uint32_t base = 0; base < total_cand; base += 64) {
        const uint32_t f = base + (uint32_t)lane;
        bool ok = f < total_cand;
        uint32_t j = 0;
        if (ok) {
            // the cell holding candidate f: the first c whose running end exceeds f
            uint32_t e = 0, start = 0;
            bool found = false;
            for (int c = 0; c < 27; ++c) {
                const uint32_t ce = (uint32_t)__builtin_amdgcn_readlane((int)c_end, c);
                if (!found && f < ce) {
                    found = true;
                    e = (uint32_t)__builtin_amdgcn_readlane((int)c_e0, c) + (f - start);
                }
                start = ce;
            }
            const float4 qe = a.pos_sorted[e];
            const uint32_t fl = a.flags_sorted[e];
            j = __float_as_uint(qe.w);
            ok = dist3(qe.x - q.x, qe.y - q.y, qe.z - q.z) <= a.g.radius;
            if (ok && a.cluster_id && a.cluster_id[j] != mycluster) ok = false;
            if (ok && a.flag_check && fl != myflags) ok = false;
        }
        const uint64_t mask = __ballot(ok);
        const uint32_t total = (uint32_t)__popcll(mask);
        const uint32_t pos = cnt + (uint32_t)__popcll(mask & below);
        if (ok && pos < a.K) put(pos, j);  // the slots before the cap fill in order
        if (cnt + total > a.K) {
            // reservoir over the cap, entry by entry in walk order (after the in-order stores
            // above have completed)
            if (cnt + total > WF_CAP) __threadfence_block();
            uint64_t rest = mask;
            uint32_t at = cnt;
            while (rest) {
                const int src = __ffsll((long long)rest) - 1;
                rest &= rest - 1;
                const uint32_t jj = (uint32_t)__shfl((int)j, src);
                if (at >= a.K && lane == 0) {
                    const uint32_t r = pcg_next(rng) % (at + 1);
                    if (r < a.K) put(r, jj);
                }
                ++at;
            }
        }
        cnt += total;
    }
    // make the list written by all lanes visible to the whole wave
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    // wave-uniform by construction; readfirstlane lets the compiler keep the list loops' counters
    // in scalar registers
    const uint32_t used = (uint32_t)__builtin_amdgcn_readfirstlane((int)(cnt < a.K ? cnt : a.K));
    const float inv_used = used ? 1.0f / (float)used : 0.0f;
    auto get = [&](uint32_t i) -> uint32_t { return i < WF_CAP ? lb[i] : gbuf[i]; };
    // stage the SMIS records of list entries m0 .. m0+mn-1 in LDS (lane m holds entry m0 + m)
    auto stage = [&](uint32_t m0, uint32_t mn) {
        if ((uint32_t)lane < mn) {
            SmisGloss g;
            sb[lane] = smis_rec_of<GLOSSY>(a, a.pk[get(m0 + (uint32_t)lane)], g);
            if (GLOSSY) sg[lane] = g;
        }
    };
    // sum + pi * pdf_m(wi) over the staged records 0 .. mn-1, in order; records are broadcast
    // from LDS, so their flags are wave-uniform
    auto smis_chunk = [&](double sum, uint32_t mn, float3 wi, double wx, double wy, double wz) {
#pragma unroll 4
        for (uint32_t m = 0; m < mn; ++m) {
            const uint32_t bits = (uint32_t)__builtin_amdgcn_readfirstlane((int)sb[m].bits);
            sum = smis_term<GLOSSY>(sum, sb[m], sg[GLOSSY ? m : 0], bits, wi, wx, wy, wz);
        }
        return sum;
    };
    // With clustering, an uncapped list (cnt <= K) is every member of the vertex's cluster with
    // its flags, in the walk order around the cluster seed: each vertex j of the list has this
    // same list. Its first entry sums pdf_m(wi_j) over it once for every j (in the order used
    // below, so the sums are the ones each member would form), and the members look them up.
    const bool shared = a.smis_phase != 0 && cnt <= a.K;
    if (a.smis_phase == 1) {
        if (!shared || used == 0 || get(0) != c) return;  // uniform per wave
        for (uint32_t k0 = 0; k0 < used; k0 += 64) {
            const uint32_t k = k0 + (uint32_t)lane;
            uint32_t j = 0;
            float3 wi = make_float3(0.0f, 0.0f, 0.0f);
            if (k < used) {
                j = get(k);
                const float4 iw = a.pk[j].iwi_mis;
                wi = make_float3(iw.x, iw.y, iw.z);
            }
            const double wx = wi.x, wy = wi.y, wz = wi.z;
            double sum_cos = 0.0;
            for (uint32_t m0 = 0; m0 < used; m0 += 64) {
                const uint32_t mn =
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)(used - m0 < 64 ? used - m0 : 64));
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
                stage(m0, mn);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
                sum_cos = smis_chunk(sum_cos, mn, wi, wx, wy, wz);
            }
            if (k < used) a.smis_sum[j] = sum_cos;
        }
        if (lane == 0) atomicAdd(a.pair_count + (blockIdx.x & (NBR_SLOTS - 1)), (unsigned long long)used * used);
        return;
    }
    // hundreds of reused terms per vertex: Kahan-compensated sums keep FP32 within ~1 ulp of
    // the exact sum (a plain running sum drifts by ~n ulp)
    float dsum[3] = {0, 0, 0}, dc[3] = {0, 0, 0}, isum[3] = {0, 0, 0}, ic[3] = {0, 0, 0};
    for (uint32_t k0 = 0; k0 < used; k0 += 64) {
        const uint32_t k = k0 + (uint32_t)lane;
        const bool active = k < used;
        float3 wi = make_float3(0.0f, 0.0f, 0.0f);
        float3 reused = make_float3(0.0f, 0.0f, 0.0f);
        bool ind_ok = false;
        float bpdf = 0.0f;
        if (active) {
            const uint32_t j = get(k);
            const NbrRec& r = a.pk[j];
            const float4 dw = r.dwi_pdf;
            if (dw.w > 0.0f) {
                float3 fc;
                const float dpdf = cb.eval(make_float3(dw.x, dw.y, dw.z), fc);
                const float4 di = r.dirr_mis;
                const float w = dw.w / (dw.w + dpdf);  // mis_weight, balance heuristic (:305-316)
                const float sc = di.w * w / dw.w * inv_used;
                kahan_add(dsum[0], dc[0], fc.x * sc * di.x);
                kahan_add(dsum[1], dc[1], fc.y * sc * di.y);
                kahan_add(dsum[2], dc[2], fc.z * sc * di.z);
            }
            const float4 iw = r.iwi_mis;
            wi = make_float3(iw.x, iw.y, iw.z);
            if (sqrtf(wi.x * wi.x + wi.y * wi.y + wi.z * wi.z) != 0.0f) {
                ind_ok = true;
                float3 fc;
                bpdf = cb.eval(wi, fc);
                const float4 ir = r.iirr;
                reused = make_float3(fc.x * ir.x * iw.w, fc.y * ir.y * iw.w, fc.z * ir.z * iw.w);
            }
        }
        if (a.smis) {
            // sum_m pdf_m(wi_k) over every reused neighbour m, in order, in FP64
            const double wx = wi.x, wy = wi.y, wz = wi.z;
            double sum_cos = 0.0;
            if (shared) {
                if (active && ind_ok) sum_cos = a.smis_sum[get(k)];
            } else {
                for (uint32_t m0 = 0; m0 < used; m0 += 64) {
                    const uint32_t mn =
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(used - m0 < 64 ? used - m0 : 64));
                    // every lane has read the previous chunk before it is overwritten
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    stage(m0, mn);
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    sum_cos = smis_chunk(sum_cos, mn, wi, wx, wy, wz);
                }
            }
            const float sum_pdf = (float)(sum_cos * 0.318309886183790671538);
            if (active && ind_ok) {
                kahan_add(isum[0], ic[0], reused.x / sum_pdf);
                kahan_add(isum[1], ic[1], reused.y / sum_pdf);
                kahan_add(isum[2], ic[2], reused.z / sum_pdf);
            }
        } else if (active && ind_ok && bpdf > 0.0f) {
            kahan_add(isum[0], ic[0], reused.x / bpdf * inv_used);
            kahan_add(isum[1], ic[1], reused.y / bpdf * inv_used);
            kahan_add(isum[2], ic[2], reused.z / bpdf * inv_used);
        }
    }
    // fixed-order reduction over the lanes
    float v[6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        v[i] = dsum[i] - dc[i];
        v[3 + i] = isum[i] - ic[i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < 6; ++i) v[i] += __shfl_xor(v[i], off);
    if (lane == 0) {
        atomicAdd(a.nbr_count + (blockIdx.x & (NBR_SLOTS - 1)), (unsigned long long)cnt);
        if (a.smis && !shared)
            atomicAdd(a.pair_count + (blockIdx.x & (NBR_SLOTS - 1)), (unsigned long long)used * used);
        v3<float> L{le.x + v[0] + v[3], le.y + v[1] + v[4], le.z + v[2] + v[5]};
        if (a.stylize) {
            const uint32_t m = a.material[c];
            const int root = m < a.n_materials ? a.roots[m] : -1;
            if (root >= 0) L = apply_style<float>(a.nodes, root, (uint32_t)a.depth, myflags, L, v3<float>{p.x, p.y, p.z});
        }
        a.rad_out[3 * (uint64_t)c] = L.x;
        a.rad_out[3 * (uint64_t)c + 1] = L.y;
        a.rad_out[3 * (uint64_t)c + 2] = L.z;
        if (a.use_rng) a.rng[tl] = rng;
    }
}

// parent (depth d-1) indirect record <- child's filtered radiance (kernel_propagate :973-998,
// record fields from kernel_filter :1043-1058)
__global__ void w_propagate_kernel(const float* rec_child, const float* rec_parent, uint64_t N, const uint32_t* pts,
                                   uint32_t n, const float* rad, float* ind) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = pts[t];
    ind[i] = -pl(rec_child, N, 9, i);
    ind[N + i] = -pl(rec_child, N, 10, i);
    ind[2 * N + i] = -pl(rec_child, N, 11, i);
    ind[3 * N + i] = rad[3 * (uint64_t)i];
    ind[4 * N + i] = rad[3 * (uint64_t)i + 1];
    ind[5 * N + i] = rad[3 * (uint64_t)i + 2];
    ind[6 * N + i] = pl(rec_child, N, 18, i);
    ind[7 * N + i] = pl(rec_parent, N, 27, i);
}

// upper bound of a point's neighbour count, max over out_max[NBR_SLOTS]: the entries of the
// buckets of its 27-cell walk (around its cluster seed when clustering), i.e. the candidates the
// filter tests
__global__ void w_cand_max_kernel(const float* rec, uint64_t N, const uint32_t* pts, uint32_t n, Grid g,
                                  const uint32_t* cluster_id,
                                  uint32_t* out_max) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t total = 0;
    if (t < n) {
        const uint32_t c = pts[t];
        const uint32_t qc = cluster_id ? cluster_id[c] : c;
        uint32_t cell[3];
        get_cell(g, pl(rec, N, 0, qc), pl(rec, N, 1, qc), pl(rec, N, 2, qc), cell);
#pragma unroll
        for (int g0 = 0; g0 < 27; g0 += 9) {  // nine lookups per round of loads
            uint32_t bk[9];
            uint4 v[9];
#pragma unroll
            for (int u = 0; u < 9; ++u) v[u] = first_probe(g, cell, g0 + u, bk[u]);
#pragma unroll
            for (int u = 0; u < 9; ++u) {
                const uint2 br = probe_range(g, v[u], bk[u]);
                total += br.y - br.x;
            }
        }
    }
    total = wave_max(total);
    if ((threadIdx.x & 63) == 0 && total) atomicMax(out_max + (blockIdx.x & (NBR_SLOTS - 1)), total);
}

__global__ void w_fill_kernel(float* p, uint64_t n, float v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// kernel_splat + kernel_update_film + Film::add_sample (:428-506, film.rs:187-222) per pixel,
// branches summed in order. accumulate == 0: film[3 npix] gets the mean of the pass's samples;
// otherwise film is a 7-plane film buffer and the samples (weight 1 each) are added onto its
// pixel and weight planes.
__global__ void w_film_kernel(const float* rad0, const float* rec0, const int32_t* path_depth, const uint8_t* flags0,
                              const uint16_t* material0, uint64_t npix, uint32_t branch, uint64_t N,
                              const nlpf_style_node* nodes, const int32_t* roots, uint32_t n_materials,
                              int stylize, int accumulate, float* film) {
#pragma clang fp contract(off)
    const uint64_t px = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= npix) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (uint32_t b = 0; b < branch; ++b) {
        const uint64_t t = (uint64_t)b * npix + px;
        if (path_depth[t] < 0) continue;
        sx += rad0[3 * t];
        sy += rad0[3 * t + 1];
        sz += rad0[3 * t + 2];
    }
    const float scale = 1.0f / (float)branch;
    float ax = 0.0f, ay = 0.0f, az = 0.0f, wsum = 0.0f;
    if (accumulate) {
        ax = film[3 * px];
        ay = film[3 * px + 1];
        az = film[3 * px + 2];
        wsum = film[6 * npix + px];
    }
    for (uint32_t b = 0; b < branch; ++b) {
        const uint64_t t = (uint64_t)b * npix + px;
        if (path_depth[t] < 0) continue;
        v3<float> L{sx * scale, sy * scale, sz * scale};
        if (stylize) {
            const uint32_t m = material0[t];
            const int root = m < n_materials ? roots[m] : -1;
            if (root >= 0)
                L = apply_style<float>(nodes, root, 0u, flags0[t], L,
                                       v3<float>{pl(rec0, N, 0, t), pl(rec0, N, 1, t), pl(rec0, N, 2, t)});
        }
        if (isnan(L.x) || isnan(L.y) || isnan(L.z)) L = v3<float>{0.0f, 0.0f, 0.0f};
        ax += L.x;
        ay += L.y;
        az += L.z;
        wsum += 1.0f;
    }
    if (accumulate) {
        film[3 * px] = ax;
        film[3 * px + 1] = ay;
        film[3 * px + 2] = az;
        film[6 * npix + px] = wsum;
        return;
    }
    film[3 * px] = wsum > 0.0f ? ax / wsum : 0.0f;
    film[3 * px + 1] = wsum > 0.0f ? ay / wsum : 0.0f;
    film[3 * px + 2] = wsum > 0.0f ? az / wsum : 0.0f;
}

// reservoir streams of the batch slots, seeded as the reference's
// init_pcg32_buffer_with_seed (sampler/mod.rs:157-174)
__global__ void w_rng_init_kernel(Pcg32* rng, uint32_t n, StdKey key) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) rng[t] = reservoir_stream(key, t);
}

// ---- deterministic clustering (replaces the CAS seeding / joining of :532-623) ----
// seed per bucket = smallest unclustered path id in it (slot_of[t]: the table slot of point t's
// bucket, from the grid build)
__global__ void w_cluster_seed_kernel(const uint32_t* pts, const uint32_t* slot_of, uint32_t n, Grid g,
                                      const uint32_t* cluster_id) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = pts[t];
    if (cluster_id[i] != 0xFFFFFFFFu) return;
    atomicMin(&g.ctab[slot_of[t]].w, i);
}
__global__ void w_cluster_mark_seed_kernel(const uint32_t* pts, const uint32_t* slot_of, uint32_t n, Grid g,
                                           uint32_t* cluster_id) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = pts[t];
    if (cluster_id[i] != 0xFFFFFFFFu) return;
    if (g.ctab[slot_of[t]].w == i) cluster_id[i] = i;
}
// seed index of one seeding round over the bucket-sorted entries: seed_rank[e] = number of
// seed entries before e (exclusive scan of these flags), seed_entry[seed_rank[e]] = e
// The same pass clears the round's per-bucket seed words (the mark kernel has read them): the
// first entry of each bucket resets its slot, so every occupied bucket starts the next round
// at kEmptyKey.
__global__ void w_seed_flags_kernel(const uint32_t* keys_sorted, const uint32_t* ids_sorted, uint32_t n, Grid g,
                                    const uint32_t* cluster_id, uint32_t* flags) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > n) return;
    flags[e] = e < n && cluster_id[ids_sorted[e]] == ids_sorted[e] ? 1u : 0u;
    if (e < n && (e == 0 || keys_sorted[e - 1] != keys_sorted[e])) g.ctab[keys_sorted[e]].w = kEmptyKey;
}
// (also zeroes the counters of the round's first propose pass: ncounters words)
__global__ void w_seed_scatter_kernel(const uint32_t* flags, const uint32_t* seed_rank, uint32_t n,
                                      uint32_t* seed_entry, uint32_t* counters, uint32_t ncounters) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < ncounters; k += blockDim.x) counters[k] = 0u;
    if (e < n && flags[e]) seed_entry[seed_rank[e]] = e;
}
// every unclustered vertex proposes to the first seed (visit order of :60-123) within the
// radius that still has room; key = seed << 32 | vertex so a sort groups proposals by seed
// with the smallest vertex ids first. Only seeds can take a proposal, so the walk visits a
// bucket's seed entries (seed_entry[seed_rank[start] .. seed_rank[end]), in bucket order).
__global__ void __launch_bounds__(256) w_cluster_propose_kernel(const float* rec, uint64_t N, const uint32_t* pts, uint32_t n, Grid g,
                                         const float4* pos_sorted,
                                         const uint32_t* seed_rank, const uint32_t* seed_entry,
                                         const uint32_t* cluster_size, uint32_t cap, const uint32_t* cluster_id,
                                         unsigned long long* prop, uint32_t* n_prop, uint32_t* n_unclustered) {
    // launched with 256-thread workgroups (4 wavefronts); n_unclustered has NBR_SLOTS slots
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t < n ? pts[t] : 0u;
    const bool open = t < n && cluster_id[i] == 0xFFFFFFFFu;
    uint32_t seed = 0xFFFFFFFFu;
    if (open) {
        const float px = pl(rec, N, 0, i), py = pl(rec, N, 1, i), pz = pl(rec, N, 2, i);
        uint32_t cell[3];
        get_cell(g, px, py, pz, cell);
        // the 27-cell walk (dz outer, dx inner) nine cells at a time: their bucket lookups in one
        // round of loads, their seed-rank bounds in a second, then the seeds in walk order
        for (int g0 = 0; g0 < 27 && seed == 0xFFFFFFFFu; g0 += 9) {
            uint32_t bk[9], s0[9], s1[9];
            uint4 v[9];
#pragma unroll
            for (int u = 0; u < 9; ++u) v[u] = first_probe(g, cell, g0 + u, bk[u]);
#pragma unroll
            for (int u = 0; u < 9; ++u) {
                const uint2 br = probe_range(g, v[u], bk[u]);
                s0[u] = s1[u] = 0;
                if (br.x < br.y) {
                    s0[u] = seed_rank[br.x];
                    s1[u] = seed_rank[br.y];
                }
            }
#pragma unroll
            for (int u = 0; u < 9; ++u) {
                if (seed != 0xFFFFFFFFu) break;
                for (uint32_t si = s0[u]; si < s1[u]; ++si) {
                    const float4 qe = pos_sorted[seed_entry[si]];
                    if (!(dist3(qe.x - px, qe.y - py, qe.z - pz) <= g.radius)) continue;
                    const uint32_t j = __float_as_uint(qe.w);
                    if (cluster_size[j] < cap) {
                        seed = j;
                        break;
                    }
                }
            }
        }
    }
    // the open count goes to slot blockIdx % NBR_SLOTS, one update per wavefront; proposal slots
    // come from one atomic per workgroup on the shared allocator, split by ballot prefix counts
    // (the proposals are sorted afterwards, so their order here does not matter)
    __shared__ uint32_t wave_count[4], block_base;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t opened = wave_sum(open ? 1u : 0u);
    if (lane == 0 && opened) atomicAdd(n_unclustered + (blockIdx.x & (NBR_SLOTS - 1)), opened);
    const bool has = seed != 0xFFFFFFFFu;
    const uint64_t mask = __ballot(has);
    if (lane == 0) wave_count[wv] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = wave_count[0] + wave_count[1] + wave_count[2] + wave_count[3];
        block_base = total ? atomicAdd(n_prop, total) : 0u;
    }
    __syncthreads();
    uint32_t base = block_base;
    for (uint32_t k = 0; k < wv; ++k) base += wave_count[k];
    if (has) prop[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = ((unsigned long long)seed << 32) | i;
}
// first index of the run of proposals to `seed` that holds index e (binary search: the keys
// are sorted, seed in the high word)
__device__ __forceinline__ uint32_t run_start(const unsigned long long* prop_sorted, uint32_t e, uint32_t seed) {
    uint32_t lo = 0, hi = e;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint32_t)(prop_sorted[mid] >> 32) < seed)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
// accept the first (cap - size) proposals of each seed, in ascending vertex order (also zeroes
// the counters of the next propose pass, which the host has read: ncounters words)
__global__ void w_cluster_accept_kernel(const unsigned long long* prop_sorted, uint32_t n, uint32_t* cluster_id,
                                        uint32_t* cluster_size, uint32_t cap, uint32_t* counters,
                                        uint32_t ncounters) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < ncounters; k += blockDim.x) counters[k] = 0u;
    if (e >= n) return;
    const uint32_t seed = (uint32_t)(prop_sorted[e] >> 32);
    const uint32_t rank = e - run_start(prop_sorted, e, seed);
    const uint32_t size0 = cluster_size[seed];
    if (size0 + rank < cap) {
        cluster_id[(uint32_t)prop_sorted[e]] = seed;
    }
}
__global__ void w_cluster_size_kernel(const unsigned long long* prop_sorted, uint32_t n, uint32_t* cluster_size,
                                      uint32_t cap) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t seed = (uint32_t)(prop_sorted[e] >> 32);
    if (e + 1 < n && (uint32_t)(prop_sorted[e + 1] >> 32) == seed) return;  // last of its run
    const uint32_t proposals = e - run_start(prop_sorted, e, seed) + 1;
    const uint32_t size0 = cluster_size[seed];
    const uint32_t room = cap > size0 ? cap - size0 : 0u;
    cluster_size[seed] = size0 + (proposals < room ? proposals : room);
}

}  // namespace nlpf

// ---------------------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------------------
namespace nlpf {

// Scalar counters: 0 selected points, 1 cluster proposals, SLOT_CURSOR the grid build's range
// cursor. Slot arrays after the 8 scalars, NBR_SLOTS each: open (unclustered) vertices and
// occupied buckets (summed), candidate bound and largest bucket or cluster (maxed).
enum GridSlots : uint32_t {
    SLOT_CURSOR = 2,
    SLOT_OPEN = 8,
    SLOT_OCCUPIED = 8 + NBR_SLOTS,
    SLOT_CAND = 8 + 2 * NBR_SLOTS,
    SLOT_LARGEST = 8 + 3 * NBR_SLOTS,
    SLOT_DEPTHS = 8 + 4 * NBR_SLOTS,  // points per depth 0..31
    SLOT_END = 8 + 4 * NBR_SLOTS + 32
};

struct WorldWork {
    uint64_t N = 0, nbuf_elems = 0, rng_n = 0, cub_bytes = 0;
    // point lists of every depth (pts_cap entries) and the selection's per-block counts;
    // grid build: pts (a depth's paths), keys (slot of each point), arrival (its arrival rank),
    // order (scattered points, or slots sorted on the radix-sort path), then per entry in bucket
    // order: keys_sorted (the table slot of the entry's bucket), ids_sorted (path id), pos_sorted,
    // flags_sorted
    uint32_t *pts = nullptr, *keys = nullptr, *arrival = nullptr, *order = nullptr, *keys_sorted = nullptr,
             *ids_sorted = nullptr, *sel_cnt = nullptr;
    uint64_t pts_cap = 0;
    uint4* ctab = nullptr;  // compact bucket table, ctab_slots slots
    uint64_t ctab_slots = 0;
    uint32_t *cluster_id = nullptr, *cluster_size = nullptr, *counters = nullptr, *nbuf = nullptr;
    unsigned long long *prop = nullptr, *prop_sorted = nullptr, *nbr = nullptr;
    float4* pos_sorted = nullptr;
    uint8_t* flags_sorted = nullptr;
    NbrRec* pk = nullptr;
    float *ind = nullptr, *rad = nullptr;
    Pcg32* rng = nullptr;
    uint32_t *seed_flags = nullptr, *seed_rank = nullptr, *seed_entry = nullptr;
    double* smis_sum = nullptr;
    void* cub = nullptr;
    // pinned host memory for the readbacks (slot counters, depth counts, per-depth statistics):
    // copies into it need no staging blit
    uint32_t* host = nullptr;
    unsigned long long* host_nbr = nullptr;
    hipEvent_t readback = nullptr;  // end of the grid counters' copy to host
};

static void world_free(WorldWork& w) {
    void* ps[] = {w.pts, w.sel_cnt, w.keys, w.arrival, w.order, w.keys_sorted, w.ids_sorted, w.ctab,
                  w.cluster_id, w.cluster_size, w.counters, w.nbuf, w.prop, w.prop_sorted, w.nbr,
                  w.pos_sorted, w.flags_sorted, w.pk, w.ind, w.rad, w.rng, w.seed_flags, w.seed_rank, w.seed_entry,
                  w.smis_sum, w.cub};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (w.host) (void)hipHostFree(w.host);
    if (w.host_nbr) (void)hipHostFree(w.host_nbr);
    if (w.readback) (void)hipEventDestroy(w.readback);
    w = WorldWork{};
}

static inline dim3 grid1(uint64_t n, uint32_t bs = 256) { return dim3((uint32_t)((n + bs - 1) / bs)); }

}  // namespace nlpf

struct nlpf_world_state {
    nlpf::WorldWork w;
};

namespace nlpf {
// workspace owned by the handle (allocated on first use, grown as needed)
static nlpf_world_state* world_state(nlpf_handle* h) {
    if (!h->world) h->world = new nlpf_world_state();
    return h->world;
}
void world_release(nlpf_handle* h) {
    if (!h->world) return;
    world_free(h->world->w);
    delete h->world;
    h->world = nullptr;
}

template <typename T>
static int dalloc(nlpf_handle* h, T** p, uint64_t n) {
    NLPF_HIP(h, hipMalloc((void**)p, sizeof(T) * (n ? n : 1)));
    return NLPF_OK;
}

static int world_reserve(nlpf_handle* h, WorldWork& w, uint64_t N, uint64_t nbuf_elems, uint64_t rng_n) {
    if (w.N >= N && w.nbuf_elems >= nbuf_elems && w.rng_n >= rng_n) return NLPF_OK;
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    world_free(w);
    int rc = 0;
    rc |= dalloc(h, &w.pts, N);  // grown by nlpf_world_run when the depths hold more points
    w.pts_cap = N;
    rc |= dalloc(h, &w.sel_cnt, 32 * ((N + kSelBlock - 1) / kSelBlock));
    rc |= dalloc(h, &w.keys, N);
    rc |= dalloc(h, &w.arrival, N);
    rc |= dalloc(h, &w.order, N);
    rc |= dalloc(h, &w.keys_sorted, N);
    rc |= dalloc(h, &w.ids_sorted, N);
    // compact bucket table: at most N occupied buckets in 4 N or more slots (a power of two)
    w.ctab_slots = 64;
    while (w.ctab_slots < 4 * N) w.ctab_slots <<= 1;
    rc |= dalloc(h, &w.ctab, w.ctab_slots);
    rc |= dalloc(h, &w.cluster_id, N);
    rc |= dalloc(h, &w.cluster_size, N);
    rc |= dalloc(h, &w.counters, SLOT_END);  // 8 scalars, then the GridSlots arrays
    rc |= dalloc(h, &w.nbuf, nbuf_elems);
    rc |= dalloc(h, &w.prop, N);
    rc |= dalloc(h, &w.prop_sorted, N);
    rc |= dalloc(h, &w.nbr, 64 * NBR_SLOTS);  // neighbours, then SMIS pairs, per depth
    rc |= dalloc(h, &w.pos_sorted, N);
    rc |= dalloc(h, &w.flags_sorted, N);
    rc |= dalloc(h, &w.pk, N);
    rc |= dalloc(h, &w.ind, 8 * N);
    rc |= dalloc(h, &w.rad, 3 * N);
    rc |= dalloc(h, &w.rng, rng_n);
    rc |= dalloc(h, &w.seed_flags, N + 1);
    rc |= dalloc(h, &w.seed_rank, N + 1);
    rc |= dalloc(h, &w.seed_entry, N);
    rc |= dalloc(h, &w.smis_sum, N);
    if (rc) return rc;
    NLPF_HIP(h, hipHostMalloc((void**)&w.host, sizeof(uint32_t) * SLOT_END, hipHostMallocDefault));
    NLPF_HIP(h, hipHostMalloc((void**)&w.host_nbr, sizeof(unsigned long long) * 64 * NBR_SLOTS,
                              hipHostMallocDefault));
    NLPF_HIP(h, hipEventCreateWithFlags(&w.readback, hipEventDisableTiming));
    // temp storage for the hipcub primitives at the largest size
    size_t b2 = 0, b3 = 0, b4 = 0;
    NLPF_HIP(h, hipcub::DeviceRadixSort::SortPairs(nullptr, b2, w.keys, w.keys_sorted, w.pts, w.ids_sorted, (int)N, 0,
                                                   32, h->stream));
    NLPF_HIP(h, hipcub::DeviceRadixSort::SortKeys(nullptr, b3, w.prop, w.prop_sorted, (int)N, 0, 64, h->stream));
    NLPF_HIP(h, hipcub::DeviceScan::ExclusiveSum(nullptr, b4, w.seed_flags, w.seed_rank, (int)(N + 1), h->stream));
    w.cub_bytes = std::max(b2, std::max(b3, b4));
    NLPF_HIP(h, hipMalloc(&w.cub, w.cub_bytes));
    w.N = N;
    w.nbuf_elems = nbuf_elems;
    w.rng_n = rng_n;
    return NLPF_OK;
}


static uint32_t slot_sum(const uint32_t* v) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < NBR_SLOTS; ++i) s += v[i];
    return s;
}
static uint32_t slot_max(const uint32_t* v) {
    uint32_t m = 0;
    for (uint32_t i = 0; i < NBR_SLOTS; ++i) m = std::max(m, v[i]);
    return m;
}


}  // namespace nlpf

using namespace nlpf;

// Validate a material BSDF table (the principled subset that needs no albedo table) and keep
// its device form: alpha = max(roughness^2, 1e-4) in FP32 as from_roughness forms it
extern "C" int nlpf_set_bsdfs(nlpf_handle* h, const nlpf_bsdf* table, uint32_t n) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (n > NLPF_STY_MAX_MATERIALS) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "too many materials (max 65536)");
    if (n && !table) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "NULL BSDF table");
    std::vector<nlpf_bsdf> hv(table, table + n);
    std::vector<WBsdf> dv(n);
    int nondiffuse = 0;
    for (uint32_t m = 0; m < n; ++m) {
        const nlpf_bsdf& b = hv[m];
        const std::string at = " (material " + std::to_string(m) + ")";
        if (b.reserved != 0) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_bsdf.reserved must be 0" + at);
        if (b.kind > NLPF_BSDF_EMISSION) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "unknown BSDF kind" + at);
        WBsdf w{};
        w.kind = b.kind;
        if (b.kind == NLPF_BSDF_PRINCIPLED) {
            const float v[] = {b.metallic, b.roughness, b.specular_tint[0], b.specular_tint[1], b.specular_tint[2],
                               b.specular_ior_level, b.transmission, b.coat_weight};
            for (float x : v)
                if (!std::isfinite(x)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "non-finite principled parameter" + at);
            if (!(b.metallic >= 0.0f && b.metallic <= 1.0f))
                NLPF_FAIL(h, NLPF_E_INVALID_ARG, "metallic must lie in [0, 1]" + at);
            if (!(b.roughness >= 0.0f)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "roughness must be >= 0" + at);
            if (b.coat_weight != 0.0f)
                NLPF_FAIL(h, NLPF_E_UNSUPPORTED,
                          "principled coat (coat_weight != 0) needs the reference's precomputed albedo table" + at);
            if (b.metallic < 1.0f - MIX_EPS && (b.specular_ior_level != 0.0f || b.transmission != 0.0f))
                NLPF_FAIL(h, NLPF_E_UNSUPPORTED,
                          "principled with metallic < 1 needs specular_ior_level 0 and transmission 0 (its "
                          "specular layer reads the reference's precomputed albedo table)" + at);
            w.metallic = b.metallic;
            const float r2 = b.roughness * b.roughness;
            w.alpha = r2 > 1e-4f ? r2 : 1e-4f;
            for (int c = 0; c < 3; ++c) w.tint[c] = b.specular_tint[c];
        }
        if (b.kind != NLPF_BSDF_DIFFUSE) nondiffuse = 1;
        dv[m] = w;
    }
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    if (h->d_bsdf) NLPF_HIP(h, hipFree(h->d_bsdf));
    h->d_bsdf = nullptr;
    h->h_bsdf.clear();
    h->bsdf_nondiffuse = 0;
    if (n) {
        NLPF_HIP(h, hipMalloc(&h->d_bsdf, sizeof(WBsdf) * n));
        NLPF_HIP(h, hipMemcpy(h->d_bsdf, dv.data(), sizeof(WBsdf) * n, hipMemcpyHostToDevice));
    }
    h->h_bsdf = hv;
    h->bsdf_nondiffuse = nondiffuse;
    return NLPF_OK;
}

extern "C" int nlpf_world_run(nlpf_handle* h, const nlpf_world_desc* desc, const float* rec, const uint8_t* flags,
                              const uint8_t* new_flags, const uint16_t* material, const int32_t* path_depth,
                              float* film, float* radiance_out) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (!desc || !rec || !flags || !new_flags || !material || !path_depth || !film)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_world_run: NULL argument");
    const Config& cfg = h->cfg;
    if (desc->width == 0 || desc->height == 0 || desc->branch == 0 || desc->levels == 0)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "world desc: width, height, branch and levels must be >= 1");
    if (desc->levels != cfg.max_depth + 1)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "world desc: levels must equal config max_depth + 1");
    if (desc->branch != cfg.branch) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "world desc: branch must equal config branch");
    const uint64_t npix = (uint64_t)desc->width * desc->height;
    const uint64_t N = npix * desc->branch;
    // the compact bucket table holds 4 slots per path or more, a power of two counted in 32-bit
    // slot indices: at most 2^29 paths keep it within 2^31 slots
    if (N > (1ull << 29)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "world desc: too many paths (at most 2^29)");
    const uint32_t L = desc->levels;
    float bmin[3], bmax[3];
    for (int i = 0; i < 3; ++i) {
        bmin[i] = cfg.has_scene_bounds ? cfg.scene_bounds[0][i] : desc->aabb_min[i];
        bmax[i] = cfg.has_scene_bounds ? cfg.scene_bounds[1][i] : desc->aabb_max[i];
        if (!(bmax[i] > bmin[i])) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "world desc: empty scene AABB");
    }
    uint32_t kmax = 0;
    for (uint32_t d = 0; d < L; ++d) kmax = std::max(kmax, cfg.reuse_at(d));
    // batch = vram * 2^20 / (4 K) vertices (lm_path_filtering.rs:1714-1721)
    uint64_t batch_max = 0;
    for (uint32_t d = 0; d < L; ++d) {
        const uint64_t K = cfg.reuse_at(d);
        batch_max = std::max(batch_max, std::max<uint64_t>(1, cfg.vram * 1024 * 1024 / (K * 4)));
    }
    batch_max = std::min<uint64_t>(batch_max, N);
    const uint64_t nbuf_elems = std::max<uint64_t>(1, (uint64_t)cfg.vram * 1024 * 1024 / 4 + 1);
    nlpf_world_state* ws = world_state(h);
    WorldWork& w = ws->w;
    int rc = world_reserve(h, w, N, std::max<uint64_t>(nbuf_elems, kmax), batch_max);
    if (rc) return rc;
    hipStream_t st = h->stream;

    NLPF_HIP(h, hipMemsetAsync(w.nbr, 0, sizeof(unsigned long long) * 64 * NBR_SLOTS, st));
    hipLaunchKernelGGL(w_rng_init_kernel, grid1(batch_max), dim3(256), 0, st, w.rng, (uint32_t)batch_max,
                       std_seed_from_u64(cfg.seed));
    if (radiance_out)
        hipLaunchKernelGGL(w_fill_kernel, grid1(3 * N * L), dim3(256), 0, st, radiance_out, 3 * N * L, 0.0f);
    if ((rc = check_launch(h, "world setup"))) return rc;
    rc = record_event(h, 0);
    if (rc) return rc;
    // the point lists of every depth: counts per block, their offsets and the totals (one
    // readback), then the lists
    const uint32_t nblk = (uint32_t)((N + kSelBlock - 1) / kSelBlock);
    hipLaunchKernelGGL(w_sel_count_kernel, dim3(nblk), dim3(256), 0, st, path_depth, N, L, w.sel_cnt, nblk);
    hipLaunchKernelGGL(w_sel_offsets_kernel, dim3(1), dim3(1024), 0, st, w.sel_cnt, nblk, L,
                       w.counters + SLOT_DEPTHS);
    if ((rc = check_launch(h, "world depth counts"))) return rc;
    NLPF_HIP(h, hipMemcpyAsync(w.host + SLOT_DEPTHS, w.counters + SLOT_DEPTHS, L * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, st));
    NLPF_HIP(h, hipStreamSynchronize(st));
    uint32_t points_at[32] = {};
    SelBases sb{};
    uint64_t listed = 0;
    for (uint32_t d = 0; d < L; ++d) {
        points_at[d] = w.host[SLOT_DEPTHS + d];
        sb.base[d] = (uint32_t)listed;
        listed += points_at[d];
    }
    if (listed > 0xFFFFFFFFull) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "world desc: more than 2^32 vertices over the levels");
    if (listed > w.pts_cap) {
        (void)hipFree(w.pts);
        w.pts = nullptr;
        w.pts_cap = 0;
        if ((rc = dalloc(h, &w.pts, listed))) return rc;
        w.pts_cap = listed;
    }
    hipLaunchKernelGGL(w_sel_scatter_kernel, dim3(nblk), dim3(256), 0, st, path_depth, N, L, w.sel_cnt, nblk, sb,
                       w.pts);
    if ((rc = check_launch(h, "world point lists"))) return rc;

    nlpf_stats stt{};
    const float ext[3] = {bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]};
    const float avg_ext = (ext[0] + ext[1] + ext[2]) / 3.0f;
    for (int d = (int)L - 1; d >= 0; --d) {
        const float* recd = rec + (uint64_t)d * WP * N;
        float* rad = radiance_out ? radiance_out + (uint64_t)d * 3 * N : w.rad;
        const uint32_t* pts = w.pts + sb.base[d];  // depth d's points, in path order
        size_t tb = w.cub_bytes;
        const uint32_t n = points_at[d];
        stt.points[d] = n;
        if (n > 0) {
            Grid g;
            const float radius = avg_ext * 0.01f * cfg.radius_scale_at((uint32_t)d);
            for (int i = 0; i < 3; ++i) {
                g.pmin[i] = bmin[i];
                g.pmax[i] = bmax[i];
                const float r = std::floor(ext[i] / radius);
                if (!(r >= 1.0f) || r > 4.0e9f)
                    NLPF_FAIL(h, NLPF_E_INVALID_CONFIG, "hash grid resolution out of range for this radius_scale");
                g.res[i] = (uint32_t)r;
            }
            g.radius = radius;
            g.table = cfg.hash_table_size;
            // this depth's compact table: 4 n slots or more, a power of two
            uint32_t cbits = 6;
            while ((1ull << cbits) < 4ull * n) ++cbits;
            g.ctab = w.ctab;
            g.cmask = (uint32_t)((1ull << cbits) - 1);
            g.cshift = 32u - cbits;
            const uint64_t slots = 1ull << cbits;
            // slot arrays: occupied buckets, candidate bound, largest bucket (or cluster)
            hipLaunchKernelGGL(w_grid_clear_kernel, grid1(std::max<uint64_t>(slots, 3 * NBR_SLOTS)), dim3(256), 0, st,
                               w.ctab, (uint32_t)slots, w.counters + SLOT_OCCUPIED, 3 * NBR_SLOTS,
                               w.counters + SLOT_CURSOR);
            hipLaunchKernelGGL(w_insert_kernel, grid1(n), dim3(256), 0, st, recd, N, pts, n, g, w.keys, w.arrival,
                               w.counters + SLOT_OCCUPIED, w.counters + SLOT_LARGEST,
                               cfg.cluster ? w.cluster_id : nullptr, w.cluster_size);
            // occupied buckets and the largest bucket to the host while the payload pack runs
            uint32_t* hc = w.host + SLOT_OCCUPIED;  // occupied buckets, (candidate bound), largest bucket or cluster
            NLPF_HIP(h, hipMemcpyAsync(hc, w.counters + SLOT_OCCUPIED, sizeof(uint32_t) * 3 * NBR_SLOTS,
                                       hipMemcpyDeviceToHost, st));
            NLPF_HIP(h, hipEventRecord(w.readback, st));
            hipLaunchKernelGGL(w_pack_kernel, grid1(n), dim3(256), 0, st, recd, N, pts, n, w.ind, path_depth,
                               d + 1 < (int)L ? (int32_t)d : INT32_MAX, new_flags + (uint64_t)d * N,
                               material + (uint64_t)d * N, w.pk);
            if ((rc = check_launch(h, "world grid"))) return rc;
            NLPF_HIP(h, hipEventSynchronize(w.readback));
            const uint32_t occupied = slot_sum(hc);
            uint32_t largest = slot_max(hc + 2 * NBR_SLOTS);
            if (largest <= kEmitMaxBucket && cfg.world_grid != "sort") {
                hipLaunchKernelGGL(w_slot_ranges_kernel, grid1(slots, 256 * kRangeSlotsPerThread), dim3(256), 0, st,
                                   w.ctab, slots, w.counters + SLOT_CURSOR);
                hipLaunchKernelGGL(w_scatter_kernel, grid1(n), dim3(256), 0, st, w.keys, w.arrival, n, w.ctab,
                                   w.order);
                hipLaunchKernelGGL(w_emit_kernel, grid1(n), dim3(256), 0, st, w.order, w.keys, pts, n, w.ctab, recd,
                                   N, new_flags + (uint64_t)d * N, w.keys_sorted, w.ids_sorted, w.pos_sorted,
                                   w.flags_sorted);
            } else {
                tb = w.cub_bytes;
                NLPF_HIP(h, hipcub::DeviceRadixSort::SortPairs(w.cub, tb, w.keys, w.order, pts, w.ids_sorted,
                                                               (int)n, 0, (int)cbits, st));
                hipLaunchKernelGGL(w_sorted_emit_kernel, grid1(n), dim3(256), 0, st, w.order, w.ids_sorted, n, w.ctab,
                                   recd, N, new_flags + (uint64_t)d * N, w.keys_sorted, w.pos_sorted,
                                   w.flags_sorted);
            }
            if ((rc = check_launch(h, "world grid ranges"))) return rc;
            const uint32_t K = cfg.reuse_at((uint32_t)d);
            if (cfg.cluster) {  // the insert pass has marked every point unclustered
                for (int outer = 0;; ++outer) {
                    if (outer > 4096) NLPF_FAIL(h, NLPF_E_NUMERIC, "clustering did not converge");
                    // per-bucket seeds live in the compact table's fourth word
                    hipLaunchKernelGGL(w_cluster_seed_kernel, grid1(n), dim3(256), 0, st, pts, w.keys, n, g,
                                       w.cluster_id);
                    hipLaunchKernelGGL(w_cluster_mark_seed_kernel, grid1(n), dim3(256), 0, st, pts, w.keys, n, g,
                                       w.cluster_id);
                    // seed entries of every bucket, for the propose walks of this round
                    hipLaunchKernelGGL(w_seed_flags_kernel, grid1((uint64_t)n + 1), dim3(256), 0, st, w.keys_sorted,
                                       w.ids_sorted, n, g, w.cluster_id, w.seed_flags);
                    tb = w.cub_bytes;
                    NLPF_HIP(h, hipcub::DeviceScan::ExclusiveSum(w.cub, tb, w.seed_flags, w.seed_rank, (int)(n + 1),
                                                                 st));
                    // the propose counters (proposals at 1, open-vertex slots) are zeroed by the seed
                    // scatter for a round's first pass and by the accept pass for the next one
                    hipLaunchKernelGGL(w_seed_scatter_kernel, grid1(n), dim3(256), 0, st, w.seed_flags, w.seed_rank, n,
                                       w.seed_entry, w.counters, SLOT_OPEN + NBR_SLOTS);
                    uint32_t n_uncl = 0;
                    for (int inner = 0;; ++inner) {
                        hipLaunchKernelGGL(w_cluster_propose_kernel, grid1(n), dim3(256), 0, st, recd, N, pts, n, g,
                                           w.pos_sorted, w.seed_rank, w.seed_entry, w.cluster_size,
                                           K, w.cluster_id, w.prop, w.counters + 1, w.counters + SLOT_OPEN);
                        if ((rc = check_launch(h, "world cluster"))) return rc;
                        uint32_t* hc = w.host;  // proposal count at 1, open-vertex slots
                        NLPF_HIP(h, hipMemcpyAsync(hc, w.counters, sizeof(uint32_t) * (SLOT_OPEN + NBR_SLOTS),
                                                   hipMemcpyDeviceToHost, st));
                        NLPF_HIP(h, hipStreamSynchronize(st));
                        const uint32_t n_prop = hc[1];
                        n_uncl = slot_sum(hc + SLOT_OPEN);
                        if (n_prop == 0) break;
                        tb = w.cub_bytes;
                        NLPF_HIP(h, hipcub::DeviceRadixSort::SortKeys(w.cub, tb, w.prop, w.prop_sorted, (int)n_prop, 0,
                                                                      64, st));
                        hipLaunchKernelGGL(w_cluster_accept_kernel, grid1(n_prop), dim3(256), 0, st, w.prop_sorted,
                                           n_prop, w.cluster_id, w.cluster_size, K, w.counters,
                                           SLOT_OPEN + NBR_SLOTS);
                        hipLaunchKernelGGL(w_cluster_size_kernel, grid1(n_prop), dim3(256), 0, st, w.prop_sorted, n_prop,
                                           w.cluster_size, K);
                        if (inner > 1 << 20) NLPF_FAIL(h, NLPF_E_NUMERIC, "cluster joining did not converge");
                    }
                    if (n_uncl == 0) break;
                }
            }
            // The reference filters in batches of vram / (4 K) points (:1714-1744); batching is
            // observable only through the reservoir streams, which are per batch slot. When no
            // point's candidate count can exceed K there are no draws, and the whole depth runs
            // as one launch with list slots sized by that count: first bounded by 27 x the
            // largest bucket (27 x the largest cluster when clustering: colliding cells walk one
            // bucket, and so its members, more than once), else counted exactly when the depth
            // spans more than one batch.
            if (cfg.cluster) {
                NLPF_HIP(h, hipMemsetAsync(w.counters + SLOT_LARGEST, 0, NBR_SLOTS * sizeof(uint32_t), st));
                hipLaunchKernelGGL(w_cluster_max_kernel, grid1(n), dim3(256), 0, st, pts, n, w.cluster_id,
                                   w.cluster_size, w.counters + SLOT_LARGEST);
                if ((rc = check_launch(h, "world cluster size"))) return rc;
                NLPF_HIP(h, hipMemcpyAsync(hc + 2 * NBR_SLOTS, w.counters + SLOT_LARGEST, sizeof(uint32_t) * NBR_SLOTS,
                                           hipMemcpyDeviceToHost, st));
                NLPF_HIP(h, hipStreamSynchronize(st));
                largest = slot_max(hc + 2 * NBR_SLOTS);
            }
            const uint64_t bound = 27ull * largest;
            // kernel choice: records lie on surfaces, so about 9 of the 27 visited cells are
            // occupied and a vertex meets ~9 n / (occupied buckets) candidates; a wavefront per
            // vertex pays off from ~16 of them (with clusters, the cluster cap bounds the set)
            const double cand = 9.0 * (double)n / (double)std::max<uint32_t>(occupied, 1u);
            const bool wave_kernel = cfg.world_kernel == "wave" ||
                                     (cfg.world_kernel == "auto" &&
                                      std::min<double>(cand, cfg.cluster ? (double)K + 1.0 : cand) >= 16.0);
            // the reference's batch (vram * 2^20 / (4 K) points); the exact candidate count can only
            // merge several of them into one launch: with one batch the batched path is a single
            // launch already, and its reservoir draws happen exactly when a count exceeds K
            const uint64_t ref_batch = std::max<uint64_t>(1, cfg.vram * 1024 * 1024 / ((uint64_t)K * 4));
            uint64_t max_cand = bound;
            if (!(bound <= K && (uint64_t)n * std::max<uint64_t>(bound, 1) <= w.nbuf_elems) && n > ref_batch) {
                hipLaunchKernelGGL(w_cand_max_kernel, grid1(n), dim3(256), 0, st, recd, N, pts, n, g,
                                   cfg.cluster ? w.cluster_id : nullptr, w.counters + SLOT_CAND);
                if ((rc = check_launch(h, "world candidates"))) return rc;
                NLPF_HIP(h, hipMemcpyAsync(hc + NBR_SLOTS, w.counters + SLOT_CAND, NBR_SLOTS * sizeof(uint32_t),
                                           hipMemcpyDeviceToHost, st));
                NLPF_HIP(h, hipStreamSynchronize(st));
                max_cand = slot_max(hc + NBR_SLOTS);
            }
            const bool single = max_cand <= K && (uint64_t)n * std::max<uint64_t>(max_cand, 1) <= w.nbuf_elems;
            const uint64_t batch = single ? n : ref_batch;
            // clustering + SMIS on the wave kernel: a build pass over every batch fills the
            // shared SMIS denominators before the filter pass reads them
            const bool shared_smis = cfg.cluster && cfg.smis && wave_kernel;
            for (int phase = shared_smis ? 1 : 0; phase <= (shared_smis ? 2 : 0); ++phase) {
                for (uint64_t off = 0; off < n; off += batch) {
                    const uint32_t cnt = (uint32_t)std::min<uint64_t>(batch, n - off);
                    FilterArgs fa;
                    fa.rec = recd;
                    fa.N = N;
                    fa.pts = pts;
                    fa.off = (uint32_t)off;
                    fa.count = cnt;
                    fa.new_flags = new_flags + (uint64_t)d * N;
                    fa.material = material + (uint64_t)d * N;
                    fa.cluster_id = cfg.cluster ? w.cluster_id : nullptr;
                    fa.pk = w.pk;
                    fa.pos_sorted = w.pos_sorted;
                    fa.flags_sorted = w.flags_sorted;
                    fa.g = g;
                    fa.K = K;
                    fa.nbuf = w.nbuf;
                    fa.stride = single ? (uint32_t)std::max<uint64_t>(max_cand, 1) : K;
                    fa.rng = w.rng;
                    fa.use_rng = single ? 0 : 1;
                    fa.nodes = h->d_nodes;
                    fa.roots = h->d_roots;
                    fa.n_materials = (uint32_t)h->h_roots.size();
                    fa.bsdf = (const WBsdf*)h->d_bsdf;
                    fa.n_bsdf = (uint32_t)h->h_bsdf.size();
                    fa.depth = d;
                    fa.smis = cfg.smis;
                    fa.flag_check = !cfg.disable_flag_check;
                    fa.stylize = (d > 0 || !cfg.no_first_hit_stylization) && !h->h_roots.empty();
                    fa.rad_out = rad;
                    fa.nbr_count = w.nbr + (uint64_t)d * NBR_SLOTS;
                    fa.pair_count = w.nbr + (uint64_t)(32 + d) * NBR_SLOTS;
                    fa.smis_phase = phase;
                    fa.smis_sum = w.smis_sum;
                    // materials that are not all diffuse take the kernels with the principled and
                    // emission-only BSDFs (more LDS and registers)
                    const bool glossy = h->bsdf_nondiffuse != 0;
                    if (wave_kernel && glossy)
                        hipLaunchKernelGGL(w_filter_wave_kernel<true>, dim3((cnt + 3) / 4), dim3(256), 0, st, fa);
                    else if (wave_kernel)
                        hipLaunchKernelGGL(w_filter_wave_kernel<false>, dim3((cnt + 3) / 4), dim3(256), 0, st, fa);
                    else if (glossy)
                        hipLaunchKernelGGL(w_filter_thread_kernel<true>, grid1(cnt), dim3(256), 0, st, fa);
                    else
                        hipLaunchKernelGGL(w_filter_thread_kernel<false>, grid1(cnt), dim3(256), 0, st, fa);
                }
            }
            if ((rc = check_launch(h, "world filter"))) return rc;
        }
        if (d > 0 && n > 0)  // indirect records of depth d - 1 (only the pack of that depth reads them)
            hipLaunchKernelGGL(w_propagate_kernel, grid1(n), dim3(256), 0, st, recd, rec + (uint64_t)(d - 1) * WP * N,
                               N, pts, n, rad, w.ind);
        if ((rc = check_launch(h, "world propagate"))) return rc;
        rc = record_event(h, 1 + (L - 1 - d));
        if (rc) return rc;
    }
    const float* rad0 = radiance_out ? radiance_out : w.rad;
    // level 0 radiance: when radiance_out is NULL, w.rad still holds depth 0 (last level filtered)
    hipLaunchKernelGGL(w_film_kernel, grid1(npix), dim3(256), 0, st, rad0, rec, path_depth, flags, material, npix,
                       desc->branch, N, h->d_nodes, h->d_roots, (uint32_t)h->h_roots.size(),
                       (int)(!cfg.no_first_hit_stylization && !h->h_roots.empty()), (int)(desc->film_accumulate != 0),
                       film);
    if ((rc = check_launch(h, "world film"))) return rc;
    rc = record_event(h, 1 + L);
    if (rc) return rc;
    const unsigned long long* slots = w.host_nbr;
    NLPF_HIP(h, hipMemcpyAsync(w.host_nbr, w.nbr, sizeof(unsigned long long) * 64 * NBR_SLOTS, hipMemcpyDeviceToHost,
                               st));
    NLPF_HIP(h, hipStreamSynchronize(st));
    unsigned long long nb[32] = {}, np[32] = {};
    for (uint32_t d = 0; d < 32; ++d)
        for (uint32_t k = 0; k < NBR_SLOTS; ++k) {
            nb[d] += slots[(uint64_t)d * NBR_SLOTS + k];
            np[d] += slots[(uint64_t)(32 + d) * NBR_SLOTS + k];
        }
    stt.total_ms = event_ms(h, 0, 1 + L);
    for (uint32_t i = 0; i < L; ++i) stt.level_ms[L - 1 - i] = event_ms(h, i, i + 1);
    stt.film_ms = event_ms(h, L, L + 1);
    for (uint32_t d = 0; d < L && d < 32; ++d) {
        stt.neighbors[d] = nb[d];
        stt.smis_pairs[d] = np[d];
        stt.vertices += stt.points[d];
    }
    h->stats = stt;
    return NLPF_OK;
}

extern "C" int nlpf_pcg32_fill(uint64_t seq, uint64_t offset, uint32_t n, uint32_t* host_out) {
    if (n && !host_out) return NLPF_E_INVALID_ARG;
    Pcg32 g = pcg_seq_offset(seq, offset);
    for (uint32_t i = 0; i < n; ++i) host_out[i] = pcg_next(g);
    return NLPF_OK;
}

extern "C" int nlpf_reservoir_streams(uint64_t seed, uint64_t first, uint64_t count, uint32_t draws,
                                      uint32_t* host_out) {
    if (count && draws && !host_out) return NLPF_E_INVALID_ARG;
    const StdKey key = std_seed_from_u64(seed);
    for (uint64_t t = 0; t < count; ++t) {
        Pcg32 g = reservoir_stream(key, first + t);
        for (uint32_t k = 0; k < draws; ++k) host_out[t * draws + k] = pcg_next(g);
    }
    return NLPF_OK;
}

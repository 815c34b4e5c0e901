// boundary.hip - boundary-point extraction (src/find_boundaries.cpp).
//
// For a pixel (x >= 1; column 0 is never a local pixel in the reference's
// 8x32 work-group tiling, src/find_boundaries.cpp:50-92) with label L, the
// neighbours RIGHT (0,+1), BOTTOM_LEFT (+1,-1), BOTTOM (+1,0) and
// BOTTOM_RIGHT (+1,+1) are tested (out-of-image neighbours read label 0).
// A record is emitted when the white bits differ and both components have
// size >= min_size; BOTTOM_LEFT is skipped when the left and down neighbours
// differ in white bit and both are large (src/find_boundaries.cpp:148-162).
// Record: packed_x = x | plane<<14, packed_y = y | (neighbour white and
// local not) << 15; key = min(label values) << 32 | max(label values).
//
// Two forms: the drop-in writes the reference's 4 sparse planes in full;
// the detector's form writes, densely and in the reference's order, the
// records of the cluster pairs that pass ValidBlobFilter.
#include <algorithm>

#include "kernels.hpp"
#include "records.hpp"
#include "scan.hpp"
#include "sort.hpp"

namespace atq {

namespace {

// The records of one pixel, one slot per plane (the planes are tested in
// plane order, so slot order is the reference's record order).
struct PixelRecords {
    uint32_t mask;  // bit p: plane p holds a record
    atq_boundary_point p[4];
    uint64_t k[4];
    ATQ_DEV uint32_t n() const { return __popc(mask); }
};

// The labels that 4 pixels (x0 .. x0+3 of row y, x0 a multiple of 4) test,
// loaded together: columns x0-1 .. x0+4 of rows y and y+1 (0 outside the
// image), and for each one whether its component holds at
// least min_size pixels, all in one round of independent loads.
struct QuadLabels {
    uint32_t a[6], b[6];    // rows y and y+1
    uint32_t big_a, big_b;  // bit i: sizes[value] >= min_size
};

ATQ_DEV void load_labels(const uint32_t *__restrict__ labels, uint32_t w, uint32_t h, uint32_t x0,
                         uint32_t y, QuadLabels &q) {
    const uint32_t *r0 = labels + (size_t)y * w;
    const uint4 m0 = *reinterpret_cast<const uint4 *>(r0 + x0);
    q.a[0] = x0 > 0 ? r0[x0 - 1] : 0u;
    q.a[1] = m0.x;
    q.a[2] = m0.y;
    q.a[3] = m0.z;
    q.a[4] = m0.w;
    q.a[5] = x0 + 4 < w ? r0[x0 + 4] : 0u;
    if (y + 1 < h) {
        const uint32_t *r1 = r0 + w;
        const uint4 m1 = *reinterpret_cast<const uint4 *>(r1 + x0);
        q.b[0] = x0 > 0 ? r1[x0 - 1] : 0u;
        q.b[1] = m1.x;
        q.b[2] = m1.y;
        q.b[3] = m1.z;
        q.b[4] = m1.w;
        q.b[5] = x0 + 4 < w ? r1[x0 + 4] : 0u;
    } else {
#pragma unroll
        for (int i = 0; i < 6; i++) q.b[i] = 0u;
    }
}

// The sizes of the 12 labels (one round of loads; `size_bits` consumes them).
ATQ_DEV void load_sizes(const uint32_t *__restrict__ sizes, const QuadLabels &q, uint32_t (&sz)[12]) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        sz[i] = sizes[q.a[i] & LABEL_VALUE_MASK];
        sz[6 + i] = sizes[q.b[i] & LABEL_VALUE_MASK];
    }
}

ATQ_DEV void size_bits(QuadLabels &q, const uint32_t (&sz)[12], uint32_t min_size) {
    q.big_a = q.big_b = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        if (sz[i] >= min_size) q.big_a |= 1u << i;
        if (sz[6 + i] >= min_size) q.big_b |= 1u << i;
    }
}

ATQ_DEV void load_quad(const uint32_t *__restrict__ labels, const uint32_t *__restrict__ sizes,
                       uint32_t w, uint32_t h, uint32_t min_size, uint32_t x0, uint32_t y,
                       QuadLabels &q) {
    load_labels(labels, w, h, x0, y, q);
    uint32_t sz[12];
    load_sizes(sizes, q, sz);
    size_bits(q, sz, min_size);
}

// The records of pixel x = x0 + k (x >= 1; column 0 is never a local pixel,
// except in fixed mode): RIGHT, BOTTOM_LEFT (unless left and down differ in
// white bit and are both large), BOTTOM, BOTTOM_RIGHT, each when the white
// bits differ and both components are large (src/find_boundaries.cpp:94-170).
// In fixed mode label 0 (127 pixels, out of image) has size 0, so it never
// counts as large.
ATQ_DEV void quad_pixel_records(const QuadLabels &q, uint32_t x0, uint32_t y, int k,
                                uint32_t key_shift, PixelRecords &out, bool fx = false) {
    out.mask = 0;
    const uint32_t x = x0 + k;
    if (x == 0 && !fx) return;
    const uint32_t L = q.a[k + 1];
    const uint32_t lv = L & LABEL_VALUE_MASK, wl = L & LABEL_PIXEL_MASK;
    const bool local_big = (q.big_a >> (k + 1)) & 1u;
    auto test = [&](uint32_t T, bool t_big, int plane) {
        const uint32_t wt = T & LABEL_PIXEL_MASK;
        if (wt == wl) return;
        if (!local_big || !t_big) return;
        const uint32_t tv = T & LABEL_VALUE_MASK;
        out.mask |= 1u << plane;
        out.p[plane].packed_x = (uint16_t)(x | ((uint32_t)plane << 14));
        out.p[plane].packed_y = (uint16_t)(wt > wl ? (COLOR_DIRECTION_MASK | y) : y);
        out.k[plane] = ((uint64_t)min(lv, tv) << key_shift) | (uint64_t)max(lv, tv);
    };
    test(q.a[k + 2], (q.big_a >> (k + 2)) & 1u, 0);  // right
    {
        const uint32_t left = q.a[k], down = q.b[k + 1];
        const bool dup = (left & LABEL_PIXEL_MASK) != (down & LABEL_PIXEL_MASK) &&
                         ((q.big_a >> k) & 1u) && ((q.big_b >> (k + 1)) & 1u);
        if (!dup) test(q.b[k], (q.big_b >> k) & 1u, 1);  // bottom-left
    }
    test(q.b[k + 1], (q.big_b >> (k + 1)) & 1u, 2);  // bottom
    test(q.b[k + 2], (q.big_b >> (k + 2)) & 1u, 3);  // bottom-right
}

}  // namespace

__global__ void __launch_bounds__(256)
k_boundaries_sparse(const uint32_t *__restrict__ labels, const uint32_t *__restrict__ sizes,
                    uint32_t w, uint32_t h, atq_boundary_point *__restrict__ points,
                    uint64_t *__restrict__ keys, uint32_t min_size) {
    const size_t n = (size_t)w * h;
    const size_t p0 = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (p0 >= n) return;
    const uint32_t y = (uint32_t)(p0 / w), x0 = (uint32_t)(p0 % w);
    QuadLabels q;
    load_quad(labels, sizes, w, h, min_size, x0, y, q);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        PixelRecords r;
        quad_pixel_records(q, x0, y, k, 32, r);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const bool has = (r.mask >> p) & 1u;
            points[p * n + p0 + k] = has ? r.p[p] : atq_boundary_point{0, 0};
            keys[p * n + p0 + k] = has ? r.k[p] : ~0ull;
        }
    }
}

// Filtered compact form (the detector's). ValidBlobFilter
// (include/cluster_bounds.h:85-120) looks only at a pair's ClusterBounds
// reduction (include/cluster_bounds.h:14-83: record count, bounding box,
// gradient and moment sums), which does not depend on the records' order, so
// the pairs that fail it are found, and their records dropped, before the
// key sort:
//  1. k_pair_count: regions of 64 x 16 pixels (4 per thread) reduce their
//     records per pair in an LDS table, then add each pair's partial
//     reduction into a global open-addressing table (key min<<32|max). A
//     pair whose count is already above the upper limit 2 * (2w + 2h) gets
//     no more adds (the pairs that span the frame would otherwise queue on
//     one address); such a pair fails the filter, and every pair at or below
//     the limit is reduced exactly.
//  2. k_pair_valid applies ValidBlobFilter to every slot's reduction and
//     lists the valid pairs (each work-group its run, by one atomic add);
//     k_pair_rank gives each listed pair its rank r among them (by key).
//  3. k_bnd_filter: each tile looks its pairs up and writes the records of
//     the valid ones, keyed (r << 2 | plane), in raster order and plane
//     order within a pixel, at its offset from a look-back over the tiles
//     (so the records of the valid pairs land in the reference's order).
//     The rank keys take 1 or 2 radix passes (sort.hip, rank_sort).
// More valid pairs than the rank list holds: the records keep the packed
// pair keys ((min << lb | max) << 2 | plane) and the full-key sort runs.
// If the global table runs out of probes, counts->pair_filter_off is set and
// every record is kept, with its packed pair key.
constexpr int BND_PX_PER_THREAD = 4;
constexpr int BND_TILE = 256 * BND_PX_PER_THREAD;
constexpr uint32_t PC_RX = 64, PC_RY = 16;  // k_pair_count's regions (256 x 4 pixels)
static_assert(PC_RX * PC_RY == BND_TILE, "one pixel quad per thread");
constexpr int TILE_SLOTS = 512;   // LDS pair table of a tile
constexpr int PAIR_PROBES = 64;   // global table probe limit
constexpr uint32_t NO_SLOT = 0xFFFFFFFFu;

ATQ_DEV uint32_t pair_hash(uint64_t key) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32);
}

// Slot of `key` in the tile's LDS table (inserted when absent and `insert`),
// or NO_SLOT. Slots hold ~key, 0 = empty.
ATQ_DEV uint32_t tile_slot(uint64_t *lk, uint64_t key, bool insert) {
    const uint64_t nk = ~key;
    uint32_t h = pair_hash(key) & (TILE_SLOTS - 1);
    for (int p = 0; p < TILE_SLOTS; p++, h = (h + 1) & (TILE_SLOTS - 1)) {
        const uint64_t k = lk[h];
        if (k == nk) return h;
        if (k == 0) {
            if (!insert) return NO_SLOT;
            const uint64_t old = atomicCAS((unsigned long long *)&lk[h], 0ull,
                                           (unsigned long long)nk);
            if (old == 0 || old == nk) return h;
        }
    }
    return NO_SLOT;
}

// Slot of `key` in the global table (inserted when absent and `insert`), or
// NO_SLOT when PAIR_PROBES slots were tried. Slots hold ~key, so that a
// zero-filled table is empty (key 0 is a real pair: label 0, the unknown
// pixels, against the white component whose root is node 0).
ATQ_DEV uint32_t global_slot(const PairTable &t, uint64_t key, bool insert) {
    const uint64_t nk = ~key;
    uint32_t h = pair_hash(key) & t.mask;
    for (int p = 0; p < PAIR_PROBES; p++, h = (h + 1) & t.mask) {
        // inserting: other work-groups insert concurrently, so the load is
        // coherent; looking up (a later kernel): the table is settled, and a
        // plain load may be served by the cache
        const uint64_t k = insert ? __hip_atomic_load(&t.keys[h], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : t.keys[h];
        if (k == nk) return h;
        if (k == 0) {
            if (!insert) return NO_SLOT;
            const uint64_t old = atomicCAS((unsigned long long *)&t.keys[h], 0ull,
                                           (unsigned long long)nk);
            if (old == 0 || old == nk) return h;
        }
    }
    return NO_SLOT;
}

// A run's (or a tile's) share of a pair's ClusterBounds reduction, in the
// table's zero-based form.
struct PairPart {
    uint32_t c, nxmin, xmax, nymin, ymax;
    int32_t gx, gy;
    int64_t pg;
    ATQ_DEV void init() {
        c = nxmin = xmax = nymin = ymax = 0;
        gx = gy = 0;
        pg = 0;
    }
    ATQ_DEV void add(atq_boundary_point p, bool fx) {
        const uint32_t xx = x_with_dx(p.packed_x, fx), yy = y_with_dy(p.packed_x, p.packed_y, fx);
        const int gxx = grad_x(p.packed_x, p.packed_y, fx), gyy = grad_y(p.packed_x, p.packed_y, fx);
        c++;
        nxmin = max(nxmin, 0xFFFFu - xx);
        xmax = max(xmax, xx);
        nymin = max(nymin, 0xFFFFu - yy);
        ymax = max(ymax, yy);
        gx += gxx;
        gy += gyy;
        pg += (int64_t)xx * gxx + (int64_t)yy * gyy;
    }
};

// The gradient sums of a pair as one 64-bit word: gy * 2^32 + gx. Both sums
// stay far inside int32 (at most 2 * (2w + 2h) records of +-1 each), so the
// 64-bit sum of such words is their exact combination and splits back.
ATQ_DEV unsigned long long grad_pack(int32_t gx, int32_t gy) {
    return (unsigned long long)((int64_t)gy * 4294967296ll + (int64_t)gx);
}
ATQ_DEV void grad_unpack(unsigned long long v, int32_t *gx, int32_t *gy) {
    *gx = (int32_t)(uint32_t)v;
    *gy = (int32_t)(((int64_t)v - (int64_t)*gx) >> 32);
}

// The bounds part of an add (the pair's count already added).
ATQ_DEV void global_bounds(const PairTable &t, uint32_t h, const PairPart &q) {
    atomicMax(&t.bbox[4 * (size_t)h + 0], q.nxmin);
    atomicMax(&t.bbox[4 * (size_t)h + 1], q.xmax);
    atomicMax(&t.bbox[4 * (size_t)h + 2], q.nymin);
    atomicMax(&t.bbox[4 * (size_t)h + 3], q.ymax);
    atomicAdd(&t.grad[h], grad_pack(q.gx, q.gy));
    atomicAdd(&t.moment[h], (unsigned long long)q.pg);
}

// The pair in slot g (count c) as the reference's ClusterBounds.
ATQ_DEV atq_cluster_bounds table_bounds(const PairTable &t, uint32_t g, uint32_t c) {
    atq_cluster_bounds b;
    b.x_min = (uint16_t)(0xFFFFu - t.bbox[4 * (size_t)g + 0]);
    b.x_max = (uint16_t)t.bbox[4 * (size_t)g + 1];
    b.y_min = (uint16_t)(0xFFFFu - t.bbox[4 * (size_t)g + 2]);
    b.y_max = (uint16_t)t.bbox[4 * (size_t)g + 3];
    grad_unpack(t.grad[g], &b.gx_sum, &b.gy_sum);
    b.pxgx_pygy_sum = (int64_t)t.moment[g];
    b.start = 0;
    b.count = c;
    return b;
}

ATQ_DEV void global_add(const PairTable &t, uint64_t key, const PairPart &q, uint32_t max_count,
                        atq_counts *counts) {
    const uint32_t h = global_slot(t, key, true);
    if (h == NO_SLOT) {
        counts->pair_filter_off = 1u;
        return;
    }
    if (__hip_atomic_load(&t.cnt[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > max_count) return;
    // a pair whose count passes the limit cannot pass the filter and needs
    // no bounds; a pair that stays within it sees every add below the limit
    if (atomicAdd(&t.cnt[h], q.c) + q.c > max_count) return;
    global_bounds(t, h, q);
}

// global_add for the two pairs of the LDS table slots threadIdx.x and
// threadIdx.x + 256, their probes, counts and bounds in lockstep so that the
// two chains of dependent atomics overlap.
ATQ_DEV void global_add2(const PairTable &t, const uint64_t (&key)[2], const PairPart (&q)[2],
                         const bool (&live)[2], uint32_t max_count, atq_counts *counts,
                         uint32_t (&found)[2]) {
    uint32_t h[2], slot[2] = {NO_SLOT, NO_SLOT};
    bool done[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        h[j] = pair_hash(key[j]) & t.mask;
        done[j] = !live[j];
    }
    for (int p = 0; p < PAIR_PROBES && !(done[0] && done[1]); p++) {
        uint64_t k[2];
#pragma unroll
        for (int j = 0; j < 2; j++)
            k[j] = done[j] ? 0ull
                           : __hip_atomic_load(&t.keys[h[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            if (done[j]) continue;
            const uint64_t nk = ~key[j];
            bool found = k[j] == nk;
            if (!found && k[j] == 0) {
                const uint64_t old = atomicCAS((unsigned long long *)&t.keys[h[j]], 0ull,
                                               (unsigned long long)nk);
                found = old == 0 || old == nk;
            }
            if (found) {
                slot[j] = h[j];
                done[j] = true;
            } else {
                h[j] = (h[j] + 1) & t.mask;
            }
        }
    }
    bool add[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        found[j] = slot[j];
        if (live[j] && slot[j] == NO_SLOT) counts->pair_filter_off = 1u;
        add[j] = slot[j] != NO_SLOT &&
                 __hip_atomic_load(&t.cnt[slot[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= max_count;
    }
    uint32_t old[2] = {0, 0};
#pragma unroll
    for (int j = 0; j < 2; j++)
        if (add[j]) old[j] = atomicAdd(&t.cnt[slot[j]], q[j].c);
#pragma unroll
    for (int j = 0; j < 2; j++)
        if (add[j] && old[j] + q[j].c <= max_count) global_bounds(t, slot[j], q[j]);
}

// Walks a thread's records (4 pixels, planes in order; fed one record at
// a time) with a one-entry cache, so runs of one pair reach the LDS table
// once with their reduction: step() calls f(key, reduction) whenever a run
// ends, finish() for the last run.
// Compat records all decode as RIGHT half-pixels (dx = 1, dy = 0): the
// doubled coordinates of the thread's records are xx = 2x + 1 and yy = 2y
// with one y and x non-decreasing in walking order, so a run's x bounds are
// its first and last record's, its y bounds yy, its gradient sums (gx, 0)
// and its moment sum the signed sum of xx (|sum| <= 16 * 2^15, int32);
// only the count, the signed gradient and moment sums and the last x are
// updated per record. Fixed-mode records decode every plane's half-pixel
// and take the general reduction (PairPart::add).
struct RunWalk {
    uint64_t cur = 0;
    uint32_t c = 0, x_first = 0, x_last = 0, yy = 0;
    int32_t gs = 0, ms = 0;
    PairPart q;  // fixed mode
    // the open run's reduction (zero, the identity of every field, when
    // there is none)
    ATQ_DEV PairPart part(bool fx) const {
        PairPart r;
        r.init();
        if (!c) return r;
        if (fx) return q;
        r.c = c;
        r.nxmin = 0xFFFFu - x_first;
        r.xmax = x_last;
        r.nymin = 0xFFFFu - yy;
        r.ymax = yy;
        r.gx = gs;
        r.pg = (int64_t)ms;
        return r;
    }
    template <typename F>
    ATQ_DEV void flush(F &f, bool fx) {
        if (c) f(cur, part(fx));
    }
    template <typename F>
    ATQ_DEV void step(uint64_t key, atq_boundary_point pt, F &f, bool fx) {
        const uint32_t xx = 2u * (pt.packed_x & COORD_VALUE_MASK) + 1u;
        if (c == 0 || key != cur) {  // key 0 is a real pair
            flush(f, fx);
            cur = key;
            c = 0;
            gs = 0;
            ms = 0;
            x_first = xx;
            yy = 2u * (pt.packed_y & COORD_VALUE_MASK);
            if (fx) q.init();
        }
        if (fx) {
            q.add(pt, true);
            c++;
        } else {
            const bool b2w = black_to_white(pt.packed_y);
            c++;
            gs += b2w ? 1 : -1;
            ms += b2w ? (int32_t)xx : -(int32_t)xx;
            x_last = xx;
        }
    }
};

// The wave's sum of every lane's part (maxima for the bounds), in scalars.
// Every lane active.
ATQ_DEV PairPart wave_part_total(const PairPart &q) {
    PairPart t;
    t.c = lane_value(wave_inclusive_sum(q.c), WAVE - 1);
    t.nxmin = lane_value(wave_max_to_last(q.nxmin), WAVE - 1);
    t.xmax = lane_value(wave_max_to_last(q.xmax), WAVE - 1);
    t.nymin = lane_value(wave_max_to_last(q.nymin), WAVE - 1);
    t.ymax = lane_value(wave_max_to_last(q.ymax), WAVE - 1);
    t.gx = (int32_t)lane_value(wave_inclusive_sum((uint32_t)q.gx), WAVE - 1);
    t.gy = (int32_t)lane_value(wave_inclusive_sum((uint32_t)q.gy), WAVE - 1);
    t.pg = lane_value(wave_inclusive_sum(q.pg), WAVE - 1);
    return t;
}

// Records of the 4 pixels of this thread in tile `tile` (w is a multiple
// of 4, so the 4 pixels share a row).
ATQ_DEV void tile_records(const uint32_t *labels, const uint32_t *sizes, uint32_t w, uint32_t h,
                          uint32_t min_size, uint32_t tile, PixelRecords (&r)[BND_PX_PER_THREAD],
                          bool fx) {
    const size_t n = (size_t)w * h;
    const size_t p0 = (size_t)tile * BND_TILE + threadIdx.x * BND_PX_PER_THREAD;
#pragma unroll
    for (int k = 0; k < BND_PX_PER_THREAD; k++) r[k].mask = 0;
    if (p0 >= n) return;
    const uint32_t y = (uint32_t)(p0 / w), x0 = (uint32_t)(p0 % w);
    QuadLabels q;
    load_quad(labels, sizes, w, h, min_size, x0, y, q);
#pragma unroll
    for (int k = 0; k < BND_PX_PER_THREAD; k++) quad_pixel_records(q, x0, y, k, 32, r[k], fx);
}

// Per 4-pixel thread of k_pair_count, its records for the filter (a
// thread's records nearly always form at most one run of one pair): x =
// the record slots (bit 4k + p: pixel k, plane p) | their colour bits
// (packed_y bit 15) << 16; y = the LDS table slot of the thread's run, with
// SUM_IRREGULAR set when the thread has more than one run or its run missed
// the LDS table (the filter then derives its records from the labels).
// The work-group's LDS slot -> pair-table slot map follows at its end, for
// its occupied slots (the only ones a summary names).
constexpr uint32_t SUM_IRREGULAR = 0x80000000u;
ATQ_DEV uint32_t pc_regions_y(uint32_t h) { return (h + PC_RY - 1) / PC_RY; }

// Each work-group takes a contiguous run of regions (a vertical strip) and
// merges all their records in its LDS table before touching the global one,
// so a pair costs one global add per work-group it meets, not per region.
// FX: fixed mode (record decode and column 0).
template <bool FX>
__global__ void __launch_bounds__(256)
k_pair_count(const uint32_t *__restrict__ labels, const uint32_t *__restrict__ sizes, uint32_t w,
             uint32_t h, uint32_t min_size, PairTable table, uint32_t max_count,
             atq_counts *counts, uint2 *__restrict__ sums, uint32_t *__restrict__ wg_map) {
    __shared__ uint64_t lk[TILE_SLOTS];
    __shared__ uint32_t lc[TILE_SLOTS], lb[4][TILE_SLOTS];
    __shared__ unsigned long long lg[TILE_SLOTS], lm[TILE_SLOTS];  // packed gradient, moment
    for (int i = threadIdx.x; i < TILE_SLOTS; i += 256) {
        lk[i] = 0;
        lc[i] = lb[0][i] = lb[1][i] = lb[2][i] = lb[3][i] = 0;
        lg[i] = 0;
        lm[i] = 0;
    }
    __syncthreads();
    // regions of PC_RX x PC_RY pixels (16 threads per row, 4 pixels each),
    // numbered down the columns of regions, so that a work-group's run of
    // regions is a compact vertical strip: a pair then spans few work-groups
    // and reaches the global table from few of them
    const uint32_t ry_n = pc_regions_y(h);
    const uint32_t regions = (w + PC_RX - 1) / PC_RX * ry_n;
    const uint32_t per = (regions + gridDim.x - 1) / gridDim.x;
    const uint32_t t_end = min(regions, (blockIdx.x + 1) * per);
    auto origin = [&](uint32_t reg, uint32_t &x0, uint32_t &y) {
        x0 = reg / ry_n * PC_RX + (threadIdx.x % (PC_RX / 4)) * 4;
        y = reg % ry_n * PC_RY + threadIdx.x / (PC_RX / 4);
        return x0 < w && y < h;
    };
    // a region's records are derived one pixel at a time and walked in
    // order as pair runs (RunWalk), so that only one pixel's records are
    // live: few registers, many work-groups per CU to hide the label and
    // size loads
    for (uint32_t reg = blockIdx.x * per; reg < t_end; reg++) {
        uint32_t x0, y;
        const bool here = origin(reg, x0, y);
        QuadLabels ql;
        if (here) load_quad(labels, sizes, w, h, min_size, x0, y, ql);
        // run state: RUN_NONE, the first run's LDS slot (NO_SLOT: missed the
        // table), | SUM_IRREGULAR once a second run follows
        constexpr uint32_t RUN_NONE = 0xFFFFFFFEu;
        uint32_t run = RUN_NONE, rc = 0;
        RunWalk rw;
        auto flush = [&](uint64_t key, const PairPart &q) {
            const uint32_t sl = tile_slot(lk, key, true);
            run = run == RUN_NONE ? sl : run | SUM_IRREGULAR;
            if (sl == NO_SLOT) {
                global_add(table, key, q, max_count, counts);
                return;
            }
            atomicAdd(&lc[sl], q.c);
            atomicMax(&lb[0][sl], q.nxmin);
            atomicMax(&lb[1][sl], q.xmax);
            atomicMax(&lb[2][sl], q.nymin);
            atomicMax(&lb[3][sl], q.ymax);
            atomicAdd(&lg[sl], grad_pack(q.gx, q.gy));
            atomicAdd(&lm[sl], (unsigned long long)q.pg);
        };
        if (here) {
#pragma unroll
            for (int k = 0; k < BND_PX_PER_THREAD; k++) {
                PixelRecords pr;
                quad_pixel_records(ql, x0, y, k, 32, pr, FX);
                rc |= pr.mask << (4 * k);
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    if (!((pr.mask >> p) & 1u)) continue;
                    if (pr.p[p].packed_y & COLOR_DIRECTION_MASK) rc |= 1u << (16 + 4 * k + p);
                    rw.step(pr.k[p], pr.p[p], flush, FX);
                }
            }
        }
        // the last runs, every lane of the wave here: when they all belong
        // to one pair (a region inside one pair's boundary, the common
        // case), the wave reduces them and one lane adds the total, instead
        // of 64 lanes queueing their LDS atomics on the same words
        const bool has = rw.c != 0;
        const uint64_t hb = __builtin_amdgcn_ballot_w64(has);
        if (hb) {
            const int l0 = __builtin_ctzll(hb);
            const uint64_t k0 = lane_value(rw.cur, l0);
            if (__popcll(hb) > 1 && __builtin_amdgcn_ballot_w64(has && rw.cur != k0) == 0) {
                const PairPart tq = wave_part_total(rw.part(FX));
                uint32_t sl = 0;
                if ((int)(threadIdx.x & (WAVE - 1)) == l0) {
                    const uint32_t save = run;
                    run = RUN_NONE;
                    flush(k0, tq);
                    sl = run;
                    run = save;
                }
                sl = lane_value(sl, l0);
                if (has) run = run == RUN_NONE ? sl : run | SUM_IRREGULAR;
            } else {
                rw.flush(flush, FX);
            }
        }
        if (here)
            sums[(y * w + x0) / 4] =
                make_uint2(rc, run == RUN_NONE ? 0u : run & (SUM_IRREGULAR | (TILE_SLOTS - 1)));
    }
    __syncthreads();
    static_assert(TILE_SLOTS == 2 * 256, "two table slots per thread");
    uint64_t key[2];
    PairPart q[2];
    bool live[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int i = threadIdx.x + 256 * j;
        live[j] = lk[i] != 0;
        key[j] = ~lk[i];
        q[j].c = lc[i];
        q[j].nxmin = lb[0][i];
        q[j].xmax = lb[1][i];
        q[j].nymin = lb[2][i];
        q[j].ymax = lb[3][i];
        grad_unpack(lg[i], &q[j].gx, &q[j].gy);
        q[j].pg = (int64_t)lm[i];
    }
    uint32_t found[2];
    global_add2(table, key, q, live, max_count, counts, found);
    // only occupied slots: a summary names the slot of a run it recorded
#pragma unroll
    for (int j = 0; j < 2; j++)
        if (live[j]) wg_map[(size_t)blockIdx.x * TILE_SLOTS + threadIdx.x + 256 * j] = found[j];
}

constexpr uint32_t RANK_NONE = 0xFFFFFFFFu;     // slot_rank: the pair fails the filter
constexpr uint32_t RANK_PENDING = 0xFFFFFFFEu;  // valid, beyond the rank list

// The ranked pairs' lookup table: open addressing over vmask + 1 words, a
// word (rank << 2lb) | packed key (packed = min << lb | max), all ones when
// empty. Usable when a rank below 16384 and the empty word's all-ones rank
// field fit beside the 2lb key bits (lb <= 24, frames below 2^24 pixels);
// larger frames look their pairs up in the pair table instead.
constexpr uint64_t VTAB_EMPTY = ~0ull;
ATQ_DEV bool vtab_usable(uint32_t lb) { return 2 * lb + 15 <= 64; }
ATQ_DEV uint32_t vtab_hash(uint64_t packed, uint32_t vbits) {
    const uint32_t x = ((uint32_t)packed ^ ((uint32_t)(packed >> 32) * 0x9E3779B1u)) * 0x85EBCA77u;
    return x >> (32 - vbits);
}
ATQ_DEV uint64_t packed_pair(uint64_t key, uint32_t lb) {
    return ((key >> 32) << lb) | (key & 0xFFFFFFFFull);
}

// Verdicts of the table's slots and the list of valid pairs: slot i holds a
// valid pair when its count is in [24, max_count] and its reduction passes
// ValidBlobFilter (include/cluster_bounds.h:85-120). One work-group per
// PV_TILE slots (2 per thread, one 8-byte count load); each work-group
// takes its run of the list with one atomic add on counts->n_valid_pairs
// (zero before the frame), so the list is in no particular order: the
// ranking counts smaller keys, which does not depend on it. When more than
// `cap` pairs are valid the ranks are not used (the full-key sort runs), so
// which ones the list holds does not matter. n_valid_pairs ends as the
// number of valid pairs.
constexpr int PV_ITEMS = 2;
constexpr uint32_t PV_TILE = 256 * PV_ITEMS;
__global__ void __launch_bounds__(256)
k_pair_valid(PairTable t, uint32_t w, uint32_t h, uint32_t max_count,
             uint32_t *__restrict__ slot_rank, uint64_t *__restrict__ vkeys,
             uint32_t *__restrict__ vslots, uint32_t *__restrict__ vcnt, uint32_t cap,
             atq_counts *counts, uint64_t *next_flags, uint32_t n_flags,
             uint64_t *__restrict__ vtab, uint32_t vwords) {
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t s_prefix;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n_flags; i += gridDim.x * 256)
        next_flags[i] = 0;
    // the ranked pairs' lookup table, emptied for k_pair_rank
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; 2 * i < vwords; i += gridDim.x * 256)
        *reinterpret_cast<uint4 *>(vtab + 2 * i) = make_uint4(~0u, ~0u, ~0u, ~0u);
    const uint32_t slots = t.mask + 1;  // a power of two >= 256
    const uint32_t tile = blockIdx.x, i0 = tile * PV_TILE + threadIdx.x * PV_ITEMS;
    const uint2 c2 = i0 < slots ? *reinterpret_cast<const uint2 *>(t.cnt + i0) : make_uint2(0, 0);
    const uint32_t cs[PV_ITEMS] = {c2.x, c2.y};
    uint32_t v[PV_ITEMS], sum = 0;
#pragma unroll
    for (int k = 0; k < PV_ITEMS; k++) {
        v[k] = cs[k] >= 24 && cs[k] <= max_count &&
                       valid_blob(table_bounds(t, i0 + k, cs[k]), w, h) ? 1u : 0u;
        sum += v[k];
    }
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<uint32_t>(sum, lds4, &tot);
    if (threadIdx.x == 0) s_prefix = tot ? atomicAdd(&counts->n_valid_pairs, tot) : 0u;
    __syncthreads();
    const uint32_t prefix = s_prefix;
    if (i0 >= slots) return;
    uint32_t run = prefix + ex;
#pragma unroll
    for (int k = 0; k < PV_ITEMS; k++) {
        slot_rank[i0 + k] = v[k] ? RANK_PENDING : RANK_NONE;
        if (v[k]) {
            if (run < cap) {
                vkeys[run] = ~t.keys[i0 + k];
                vslots[run] = i0 + k;
                vcnt[run] = cs[k];
            }
            run++;
        }
    }
}

// Rank r of every listed pair among the V valid pairs (number of smaller
// keys; keys are distinct) and its start (the records of the pairs with
// smaller keys): slot_rank[its slot] = r, rank_keys[r] = its packed key
// min << lb | max, and, since in rank mode the sorted records are exactly
// the valid pairs' records grouped in key order, cluster r's bounds (the
// table's exact reduction, include/cluster_bounds.h:14-83, starting at
// start), its centre (cx, cy) (include/cluster_bounds.h:47-49) and
// extents {start, count} (src/main.cpp:700-717) and the counts of pairs,
// clusters and their points. Also decides the sort: counts->rank_shift
// (rank_shift_for) and counts->sort_passes = the radix passes the rank keys
// need, or 0 (table
// overflow, or more valid pairs than the list holds: packed keys, fallback
// sort and cluster stage). One lane per listed pair (64 per work-group);
// the 16 waves compare them against a sixteenth of each LDS tile each.
constexpr int RANK_TILE = 4096;
constexpr int RANK_WAVES = 16;
__global__ void __launch_bounds__(64 * RANK_WAVES)
k_pair_rank(PairTable table, const uint64_t *__restrict__ vkeys,
            const uint32_t *__restrict__ vslots, const uint32_t *__restrict__ vcnt,
            atq_counts *counts, uint32_t cap, uint32_t lb, uint32_t *__restrict__ slot_rank,
            uint64_t *__restrict__ rank_keys, atq_cluster_bounds *__restrict__ bounds,
            atq_cluster_extents *__restrict__ ext, float2 *__restrict__ centre,
            uint64_t *__restrict__ vtab, uint32_t vmask) {
    __shared__ uint64_t tile[RANK_TILE];
    __shared__ uint32_t tcnt[RANK_TILE];
    __shared__ uint32_t part[RANK_WAVES][64], spart[RANK_WAVES][64];
    const uint32_t V = counts->n_valid_pairs;
    const uint32_t shift = rank_shift_for(V);
    const uint32_t np = counts->pair_filter_off == 0 && V <= cap ? rank_passes(V, shift) : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counts->sort_passes = np;
        counts->rank_shift = shift;
    }
    if (np == 0 || blockIdx.x * 64 >= V) return;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 64 + lane;
    const uint64_t mine = i < V ? vkeys[i] : 0ull;
    uint32_t below = 0, start = 0;
    for (uint32_t t0 = 0; t0 < V; t0 += RANK_TILE) {
        const uint32_t nt = min((uint32_t)RANK_TILE, V - t0);
        for (uint32_t j = threadIdx.x; j < nt; j += 64 * RANK_WAVES) {
            tile[j] = vkeys[t0 + j];
            tcnt[j] = vcnt[t0 + j];
        }
        __syncthreads();
        const uint32_t per = (nt + RANK_WAVES - 1) / RANK_WAVES;
        const uint32_t q0 = min(nt, wave * per), q1 = min(nt, q0 + per);
        for (uint32_t j = q0; j < q1; j++) {
            const bool lt = tile[j] < mine;
            below += lt ? 1u : 0u;
            start += lt ? tcnt[j] : 0u;
        }
        __syncthreads();
    }
    part[wave][lane] = below;
    spart[wave][lane] = start;
    __syncthreads();
    if (wave == 0 && i < V) {
        uint32_t r = 0, st = 0;
#pragma unroll
        for (int q = 0; q < RANK_WAVES; q++) {
            r += part[q][lane];
            st += spart[q][lane];
        }
        const uint32_t slot = vslots[i], c = vcnt[i];
        slot_rank[slot] = r;
        const uint64_t packed = packed_pair(mine, lb);
        rank_keys[r] = packed;
        if (vtab_usable(lb)) {
            // keys are distinct: the first empty word on the probe path
            const uint64_t word = ((uint64_t)r << (2 * lb)) | packed;
            for (uint32_t hh = vtab_hash(packed, __popc(vmask));; hh = (hh + 1) & vmask)
                if (atomicCAS((unsigned long long *)&vtab[hh], (unsigned long long)VTAB_EMPTY,
                              (unsigned long long)word) == VTAB_EMPTY)
                    break;
        }
        atq_cluster_bounds b = table_bounds(table, slot, c);
        b.start = st;
        bounds[r] = b;
        ext[r] = atq_cluster_extents{st, c};
        centre[r] = make_float2(bounds_cx(b), bounds_cy(b));
        if (r + 1 == V) {
            counts->n_pairs = V;
            counts->n_clusters = V;
            counts->n_cluster_points = st + c;
        }
    }
}

// One work-group per 1024-pixel tile (XCD-contiguous tile order): the
// tile's records of valid pairs, keyed for the sort, densely to the tile's
// region. Raster mode (counts->rank_shift 2, and the fallback): keys
// r << 2 | plane (fallback: the packed pair key with the plane), records in
// raster order, their number to tile_counts[0][tile] (planes 1-3: 0).
// Otherwise keys r, records plane by plane, each plane in raster order, the
// counts per plane to tile_counts[plane][tile]; the gather then lays out all
// plane-0 records first, so a stable sort by r gives the reference's
// (pair, plane, y, x) order.
//
// A thread whose records form at most one run of one pair (k_pair_count's
// summary) takes its records from the summary: the run's pair-table slot
// through the counting work-group's slot map, its verdict or rank from
// slot_rank, the records rebuilt from the slot and colour masks. Other
// threads (and every thread after a pair-table overflow) derive their
// records from the labels: the first record of each run looks its pair's
// rank up in the ranked pairs' table (one 8-byte probe per run, all of the
// thread's probes in flight together; a probe that meets another pair's
// word walks on), or, with more valid pairs than ranked or frames too large
// for that table, in the pair table and its slot verdicts; after a table
// overflow every record is kept.
template <bool FX>
__global__ void __launch_bounds__(256)
k_bnd_filter(const uint32_t *__restrict__ labels, const uint32_t *__restrict__ sizes, uint32_t w,
             uint32_t h, uint32_t min_size, uint32_t lb, PairTable table,
             const uint32_t *__restrict__ slot_rank, const uint64_t *__restrict__ vtab,
             uint32_t vmask, const atq_counts *counts, uint32_t n_tiles,
             uint64_t *__restrict__ tile_keys, atq_boundary_point *__restrict__ tile_points,
             uint32_t *__restrict__ tile_counts, const uint2 *__restrict__ sums,
             const uint32_t *__restrict__ wg_map, uint32_t pc_per) {
    const uint32_t tile = xcd_tile(blockIdx.x, n_tiles);
    if (tile >= n_tiles) return;
    __shared__ uint64_t lds8[4];
    const bool keep_all = counts->pair_filter_off != 0;
    const uint32_t np = counts->sort_passes;
    const bool raster = np == 0 || counts->rank_shift != 0;  // else plane by plane
    const bool ranked = np != 0 && vtab_usable(lb);
    const size_t n = (size_t)w * h;
    const size_t p0 = (size_t)tile * BND_TILE + threadIdx.x * BND_PX_PER_THREAD;
    const bool here = p0 < n;
    const uint32_t y = here ? (uint32_t)(p0 / w) : 0u, x0 = here ? (uint32_t)(p0 % w) : 0u;
    const uint2 sm = here && !keep_all ? sums[p0 / 4] : make_uint2(0u, 0u);
    const bool irr = here && (keep_all || (sm.y & SUM_IRREGULAR) != 0);
    // a regular thread's run: its rank (or packed key when the full-key
    // sort runs)
    const uint32_t rmask = irr ? 0u : sm.x & 0xFFFFu, cmask = sm.x >> 16;
    uint32_t rank = RANK_NONE;
    uint64_t rkey = 0;
    if (rmask) {
        const uint32_t reg = x0 / PC_RX * pc_regions_y(h) + y / PC_RY;
        const uint32_t g = wg_map[(size_t)(reg / pc_per) * TILE_SLOTS + (sm.y & (TILE_SLOTS - 1))];
        if (g != NO_SLOT) {
            rank = slot_rank[g];
            if (np == 0 && rank != RANK_NONE) rkey = packed_pair(~table.keys[g], lb);
        }
    }
    const uint32_t kept = rank != RANK_NONE ? rmask : 0u;
    // keep: bit 4k+p for record (pixel k, plane p); c: the thread's kept
    // records per plane, 16 bits each
    uint32_t keep = kept;
    uint64_t c = 0;
#pragma unroll
    for (int p = 0; p < 4; p++) c += (uint64_t)__popc(kept & (0x1111u << p)) << (16 * p);
    // irregular threads: their records from the labels, one pixel's at a
    // time (counted, then after the scan derived again and written), each
    // run's verdict looked up once (a one-entry cache); few registers, so
    // that the regular threads' path keeps the occupancy
    QuadLabels ql;
    uint64_t ckey = ~0ull;
    uint32_t crank = RANK_NONE;
    auto rank_of_record = [&](uint64_t key) -> uint32_t {
        if (keep_all) return RANK_PENDING;
        const uint64_t pk = packed_pair(key, lb);
        if (pk == ckey) return crank;
        uint32_t rr = RANK_NONE;
        if (ranked) {
            const uint64_t kmask = (1ull << (2 * lb)) - 1;
            uint32_t at = vtab_hash(pk, __popc(vmask));
            uint64_t x = vtab[at];
            while (x != VTAB_EMPTY && (x & kmask) != pk) {
                at = (at + 1) & vmask;
                x = vtab[at];
            }
            rr = x == VTAB_EMPTY ? RANK_NONE : (uint32_t)(x >> (2 * lb));
        } else {
            // valid pairs beyond the rank list (or frames too large for the
            // lookup table): the pair table's slot verdicts
            const uint32_t g = global_slot(table, key, false);
            rr = g == NO_SLOT ? RANK_NONE : slot_rank[g];
        }
        ckey = pk;
        crank = rr;
        return rr;
    };
    const bool wave_irr = __ballot(irr) != 0;
    if (wave_irr && irr) {
        load_quad(labels, sizes, w, h, min_size, x0, y, ql);
#pragma unroll
        for (int k = 0; k < BND_PX_PER_THREAD; k++) {
            PixelRecords pr;
            quad_pixel_records(ql, x0, y, k, 32, pr, FX);
#pragma unroll
            for (int p = 0; p < 4; p++) {
                if (!((pr.mask >> p) & 1u) || rank_of_record(pr.k[p]) == RANK_NONE) continue;
                keep |= 1u << (4 * k + p);
                c += 1ull << (16 * p);
            }
        }
    }
    if (raster) c = (c & 0xFFFFu) + ((c >> 16) & 0xFFFFu) + ((c >> 32) & 0xFFFFu) + (c >> 48);
    // the tile's records densely to its region, plane by plane, each plane
    // in raster order
    uint64_t tot;
    const uint64_t pre = block_exclusive_scan<uint64_t>(c, lds8, &tot);
    const size_t base = (size_t)tile * (4 * BND_TILE);
    // raster mode: one count (in plane 0's field), every record in raster order
    uint32_t pos[4], plane_at = 0;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        pos[p] = plane_at + ((uint32_t)(pre >> (16 * p)) & 0xFFFFu);
        plane_at += (uint32_t)(tot >> (16 * p)) & 0xFFFFu;
    }
    if (wave_irr && irr) {
#pragma unroll
        for (int k = 0; k < BND_PX_PER_THREAD; k++) {
            PixelRecords pr;
            quad_pixel_records(ql, x0, y, k, 32, pr, FX);
#pragma unroll
            for (int p = 0; p < 4; p++) {
                if (!((keep >> (4 * k + p)) & 1u)) continue;
                const uint32_t rr = rank_of_record(pr.k[p]);
                const size_t at = base + (raster ? pos[0]++ : pos[p]++);
                tile_keys[at] = np ? (raster ? ((uint64_t)rr << 2) | (uint64_t)p : (uint64_t)rr)
                                   : (packed_pair(pr.k[p], lb) << 2) | (uint64_t)p;
                tile_points[at] = pr.p[p];
            }
        }
    }
    // a regular thread's kept records, in slot order (raster: one running
    // position; plane by plane: the plane's position plus the thread's
    // earlier records of that plane)
    uint32_t m = kept, j = 0;
    while (m) {
        const uint32_t i = (uint32_t)__builtin_ctz(m);
        m &= m - 1u;
        const uint32_t k = i >> 2, pl = i & 3u;
        uint32_t at;
        if (raster) {
            at = pos[0] + j;
        } else {
            const uint32_t pb = pl == 0 ? pos[0] : pl == 1 ? pos[1] : pl == 2 ? pos[2] : pos[3];
            at = pb + (uint32_t)__popc(kept & (0x1111u << pl) & ((1u << i) - 1u));
        }
        j++;
        tile_keys[base + at] = np ? (raster ? ((uint64_t)rank << 2) | pl : (uint64_t)rank) : (rkey << 2) | pl;
        tile_points[base + at] = atq_boundary_point{
            (uint16_t)((x0 + k) | (pl << 14)),
            (uint16_t)(((cmask >> i) & 1u) ? (COLOR_DIRECTION_MASK | y) : y)};
    }
    if (threadIdx.x < 4) tile_counts[threadIdx.x * n_tiles + tile] = (uint32_t)(tot >> (16 * threadIdx.x)) & 0xFFFFu;
}

namespace {
// Plane-major tile counts (tile_counts[plane][tile]) and their offsets.
struct TileCount {
    const uint32_t *c;
    ATQ_DEV uint32_t operator()(uint32_t i) const { return c[i]; }
};
struct TileOffset {
    uint32_t *off;
    ATQ_DEV void operator()(uint32_t i, uint32_t pre, uint32_t) const { off[i] = pre; }
};
}  // namespace

// One work-group per GATHER_TILES consecutive tiles: the sort's input is
// every tile's plane-0 records in tile order, then the plane-1 records, and
// so on (tile_off: the exclusive scan of the plane-major tile counts; null:
// each work-group sums the counts itself and work-group 0 sets n_boundary,
// cheaper than a scan launch for frames of up to GATHER_SUM_TILES tiles).
// The work-group moves its tiles' records, the four planes' runs as one
// loop (4 records' loads in flight per thread).
constexpr uint32_t GATHER_TILES = 8;
constexpr uint32_t GATHER_SUM_TILES = 2048;
ATQ_DEV uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
struct GatherLds {
    uint32_t cnt[4][GATHER_TILES], pre[4][GATHER_TILES + 1], run_start[5];
    uint32_t dst[4][GATHER_TILES];        // where a tile's plane run goes
    uint32_t plane_off[4][GATHER_TILES];  // a plane's first record in its tile's region
    uint32_t acc[8];  // without tile_off: plane totals, records before the tiles
};

// The gather of NP planes (1: raster mode, every count in plane 0's row).
template <int NP>
ATQ_DEV void gather_tiles(const uint64_t *__restrict__ tile_keys,
                          const atq_boundary_point *__restrict__ tile_points,
                          const uint32_t *__restrict__ tile_counts,
                          const uint32_t *__restrict__ tile_off, uint32_t n_tiles,
                          atq_counts *counts, uint64_t *keys, atq_boundary_point *points,
                          GatherLds &L) {
    auto &cnt = L.cnt;
    auto &pre = L.pre;
    auto &run_start = L.run_start;
    auto &dst = L.dst;
    auto &plane_off = L.plane_off;
    auto &acc = L.acc;
    const uint32_t t0 = blockIdx.x * GATHER_TILES, nt = min(n_tiles - t0, GATHER_TILES);
    if (threadIdx.x < 8) acc[threadIdx.x] = 0;
    if (threadIdx.x < 4 * GATHER_TILES) {
        const uint32_t p = threadIdx.x / GATHER_TILES, q = threadIdx.x % GATHER_TILES;
        cnt[p][q] = q < nt && p < (uint32_t)NP ? tile_counts[p * n_tiles + t0 + q] : 0u;
        dst[p][q] = q < nt && p < (uint32_t)NP && tile_off ? tile_off[p * n_tiles + t0 + q] : 0u;
    }
    __syncthreads();
    if (!tile_off) {  // every tile's counts, summed here
        // every load in flight before the first is used (n_tiles <= 256 * 8)
        constexpr int PER = GATHER_SUM_TILES / 256;
        uint32_t v[PER][NP];
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const uint32_t t = threadIdx.x + 256 * j;
#pragma unroll
            for (int p = 0; p < NP; p++) v[j][p] = t < n_tiles ? tile_counts[p * n_tiles + t] : 0u;
        }
        uint32_t tot[NP], bef[NP];
#pragma unroll
        for (int p = 0; p < NP; p++) tot[p] = bef[p] = 0;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const bool before = threadIdx.x + 256 * j < t0;
#pragma unroll
            for (int p = 0; p < NP; p++) {
                tot[p] += v[j][p];
                bef[p] += before ? v[j][p] : 0u;
            }
        }
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const uint32_t a = wave_sum(tot[p]), b = wave_sum(bef[p]);
            if ((threadIdx.x & 63) == 0) {
                atomicAdd(&acc[p], a);
                atomicAdd(&acc[4 + p], b);
            }
        }
        __syncthreads();
    }
    // per (plane p, tile q), by 32 threads: the run's start among the
    // work-group's records of plane p, the plane's first record in the
    // tile's region and, without tile_off, the run's destination
    if (threadIdx.x < 4 * GATHER_TILES) {
        const uint32_t p = threadIdx.x / GATHER_TILES, q = threadIdx.x % GATHER_TILES;
        uint32_t before_q = 0, before_p = 0;
#pragma unroll
        for (uint32_t j = 0; j < GATHER_TILES; j++) before_q += j < q ? cnt[p][j] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) before_p += j < p ? cnt[j][q] : 0u;
        pre[p][q] = before_q;
        plane_off[p][q] = before_p;
        if (q == GATHER_TILES - 1) pre[p][GATHER_TILES] = before_q + cnt[p][q];
        if (!tile_off) {
            uint32_t base = acc[4 + p];
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) base += j < p ? acc[j] : 0u;
            dst[p][q] = base + before_q;
            if (blockIdx.x == 0 && threadIdx.x == 0)
                counts->n_boundary = acc[0] + acc[1] + acc[2] + acc[3];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0;
        for (int p = 0; p < 4; p++) {
            run_start[p] = r;
            r += pre[p][GATHER_TILES];
        }
        run_start[4] = r;
    }
    __syncthreads();
    const uint32_t total = run_start[4];
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += 4 * 256) {
        uint64_t k[4];
        atq_boundary_point pt[4];
        uint32_t to[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = i0 + 256 * u;
            if (i < total) {
                uint32_t p = 0;
#pragma unroll
                for (int j = 1; j < NP; j++) p += i >= run_start[j] ? 1u : 0u;
                const uint32_t ip = i - run_start[p];
                uint32_t tq = 0;
#pragma unroll
                for (uint32_t j = 1; j < GATHER_TILES; j++) tq += ip >= pre[p][j] ? 1u : 0u;
                const uint32_t within = ip - pre[p][tq];
                const size_t src = (size_t)(t0 + tq) * (4 * BND_TILE) + plane_off[p][tq] + within;
                k[u] = tile_keys[src];
                pt[u] = tile_points[src];
                to[u] = dst[p][tq] + within;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = i0 + 256 * u;
            if (i < total) {
                keys[to[u]] = k[u];
                points[to[u]] = pt[u];
            }
        }
    }
}

__global__ void __launch_bounds__(256)
k_bnd_gather(const uint64_t *__restrict__ tile_keys, const atq_boundary_point *__restrict__ tile_points,
             const uint32_t *__restrict__ tile_counts, const uint32_t *__restrict__ tile_off,
             uint32_t n_tiles, atq_counts *counts, RsDyn out) {
    __shared__ GatherLds L;
    // the rank passes end in Y; the full-key sort reads X
    const uint32_t np = counts->sort_passes;
    const bool to_y = np != 0 && (np & 1u) == 0;
    uint64_t *keys = to_y ? out.ky : out.kx;
    atq_boundary_point *points = reinterpret_cast<atq_boundary_point *>(to_y ? out.vy : out.vx);
    if (np == 0 || counts->rank_shift != 0)
        gather_tiles<1>(tile_keys, tile_points, tile_counts, tile_off, n_tiles, counts, keys, points, L);
    else
        gather_tiles<4>(tile_keys, tile_points, tile_counts, tile_off, n_tiles, counts, keys, points, L);
}

namespace {
// k_pair_count's grid: a few regions per work-group (merged in LDS before
// the global table), at most ~7 of them at large frames, so that the grid
// is resident at once (7 work-groups of 72 VGPRs and 22 KiB of LDS share a
// CU), and at least 1024 work-groups (about 2 regions each at 1080p: fewer
// leave the CUs waiting on each region's loads, more multiply the global
// table's flushes); non-decreasing in the frame's dims.
unsigned pair_count_grid(size_t w, size_t h) {
    const unsigned regions = atq_div_up(w, PC_RX) * atq_div_up(h, PC_RY);
    return std::min(regions, std::max(1024u, regions / 7));
}
}  // namespace

size_t bnd_summary_entries(size_t max_w, size_t max_h) { return max_w * max_h / 4 + 1; }
size_t bnd_map_words(size_t max_w, size_t max_h) {
    return (size_t)TILE_SLOTS * pair_count_grid(max_w, max_h);
}

int launch_find_boundaries_sparse(hipStream_t s, const uint32_t *labels, const uint32_t *sizes,
                                  size_t w, size_t h, atq_boundary_point *points, uint64_t *keys,
                                  int min_size) {
    k_boundaries_sparse<<<atq_div_up(w * h, 1024), 256, 0, s>>>(labels, sizes, (uint32_t)w, (uint32_t)h, points, keys,
                                             (uint32_t)min_size);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_find_boundaries_filtered(hipStream_t s, const uint32_t *labels, const uint32_t *sizes,
                                    size_t w, size_t h, int min_size, uint32_t label_bits,
                                    const PairTable &table, atq_counts *counts,
                                    const PairRanks &rk, const RsDyn &out, ScanCtx &sc, bool fx) {
    const unsigned tiles = atq_div_up(w * h, BND_TILE);
    const uint32_t max_count = (uint32_t)(2 * (2 * w + 2 * h));
    const uint32_t slots = table.mask + 1;
    if (!sc.flags || atq_div_up(slots, PV_TILE) + 1 > sc.flag_count ||
        4 * tiles / (LB_TILE / 2) + 1 > sc.flag_count ||
        slots < 256 || rk.cap == 0 || rk.cap % 64 || !rk.vtab || (rk.vmask & (rk.vmask + 1)) ||
        rk.vmask + 1 < 2 * rk.cap || !rk.sums || !rk.wg_map || (w & 3))
        return ATQ_ERR_BAD_ARG;
    const unsigned regions = atq_div_up(w, PC_RX) * atq_div_up(h, PC_RY);
    const unsigned pg = pair_count_grid(w, h);
    const uint32_t pc_per = atq_div_up(regions, pg);
    if (fx)
        k_pair_count<true><<<pg, 256, 0, s>>>(labels, sizes, (uint32_t)w, (uint32_t)h,
                                              (uint32_t)min_size, table, max_count, counts,
                                              rk.sums, rk.wg_map);
    else
        k_pair_count<false><<<pg, 256, 0, s>>>(labels, sizes, (uint32_t)w, (uint32_t)h,
                                               (uint32_t)min_size, table, max_count, counts,
                                               rk.sums, rk.wg_map);
    {
        // no look-back here, but the scans' flag arrays keep alternating:
        // this kernel zeroes the array the next scan uses
        uint64_t *next = sc.flags + (size_t)((sc.calls + 1) & 1) * sc.flag_count;
        sc.calls++;
        k_pair_valid<<<atq_div_up(slots, PV_TILE), 256, 0, s>>>(
            table, (uint32_t)w, (uint32_t)h, max_count, rk.slot_rank, rk.vkeys, rk.vslots, rk.vcnt,
            rk.cap, counts, next, sc.flag_count, rk.vtab, rk.vmask + 1);
    }
    k_pair_rank<<<rk.cap / 64, 64 * RANK_WAVES, 0, s>>>(table, rk.vkeys, rk.vslots, rk.vcnt, counts,
                                                         rk.cap, label_bits, rk.slot_rank,
                                                         rk.rank_keys, out.cl.bounds, out.cl.ext,
                                                         out.cl.centre, rk.vtab, rk.vmask);
    uint32_t *tile_counts = rk.tile_counts;
    if (fx)
        k_bnd_filter<true><<<xcd_grid(tiles), 256, 0, s>>>(
            labels, sizes, (uint32_t)w, (uint32_t)h, (uint32_t)min_size, label_bits, table,
            rk.slot_rank, rk.vtab, rk.vmask, counts, tiles, rk.tile_keys, rk.tile_points, tile_counts,
            rk.sums, rk.wg_map, pc_per);
    else
        k_bnd_filter<false><<<xcd_grid(tiles), 256, 0, s>>>(
            labels, sizes, (uint32_t)w, (uint32_t)h, (uint32_t)min_size, label_bits, table,
            rk.slot_rank, rk.vtab, rk.vmask, counts, tiles, rk.tile_keys, rk.tile_points, tile_counts,
            rk.sums, rk.wg_map, pc_per);
    uint32_t *tile_off = nullptr;
    if (tiles > GATHER_SUM_TILES) {
        tile_off = tile_counts + 4 * tiles;
        int r = device_scan(s, nullptr, 4 * tiles, TileCount{tile_counts}, TileOffset{tile_off}, sc,
                            &counts->n_boundary);
        if (r) return r;
    }
    k_bnd_gather<<<atq_div_up(tiles, GATHER_TILES), 256, 0, s>>>(
        rk.tile_keys, rk.tile_points, tile_counts, tile_off, tiles, counts, out);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq

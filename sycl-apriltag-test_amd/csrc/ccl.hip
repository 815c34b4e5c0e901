// ccl.hip - block-based union-find connected-component labelling (src/segmentation.cpp).
//
// The graph is the reference's (src/segmentation.cpp:101-306): one node per
// (2x2 block, colour), colour 255 ("white", node id = the block's top-left
// pixel index) and colour 0 ("black", id + 1). For each block, the pixels
// P(-1,-1), Q(-1,0), Q(-1,1), R(-1,2), S(0,-1), S(1,-1) are tested when the
// block's neighbour mask (built from its non-127 pixels, colour-agnostic,
// 8-connected for both colours) selects them; a tested pixel of colour c
// links the block's colour-c node to the colour-c node of the block holding
// it. Components are labelled by their minimum node id (min-linking).
//
// gfx950 organisation (three kernels, the reference has five; each takes a
// table of frames and walks them with a persistent grid). The scratch array
// holds a parent word per node (w * h / 2 words; only live local roots and
// the nodes of tile-border blocks are ever written) and a state word per
// block; no per-pixel array makes a round trip through HBM but the
// thresholded input (read twice: tiles, then border probes) and the labels.
//  1. k_ccl_local: a work-group owns a TBX x TBY tile of blocks (TBY = 8 or
//     16, chosen per resolution by the autotuner), builds every block's
//     mask and probes from registers, unions the tile-internal edges in an
//     LDS union-find (lockstep walks), counts each local root's pixels and
//     publishes a state word per block, the local roots of its border
//     blocks' nodes, and every live local root (pixels or an edge leaving
//     the tile) with its count in `sizes`.
//  2. k_ccl_border: one wave per tile collects the distinct pairs of local
//     roots its border edges join and merges each once on the global parent
//     array with the reference's lock-free atomicMin union
//     (include/union_find.h:50-77), halving paths as it walks.
//  3. k_ccl_label: a work-group per tile resolves each of its local roots
//     that holds pixels once (path halving; a merged root adds its pixel
//     count into its final root's size) and writes the tile's label words
//     from the state words (white = 0x80000000|root, black = root,
//     127 -> 0).
#include <algorithm>

#include "kernels.hpp"

namespace atq {

namespace {

constexpr int TBX = 32;  // blocks per tile row (64 px); block rows per tile: TBY
constexpr uint32_t FLAG = 0x80000000u;  // "not a local root" tag on parent words

constexpr uint8_t C_WHITE = 255, C_BLACK = 0;

// Neighbour mask of a block (bits of the 4x4 grid around it, bit 5 = TL),
// src/segmentation.cpp:137-186.
ATQ_DEV uint32_t check_mask(uint8_t tl, uint8_t tr, uint8_t bl, uint32_t x, uint32_t y, uint32_t w,
                            uint32_t h) {
    uint32_t m = 0;
    if (tl == 255 || tl == 0) m |= 0x777u;
    if (tr == 255 || tr == 0) m |= 0x777u << 1;
    if (bl == 255 || bl == 0) m |= 0x777u << 4;
    if (x == 0) m &= 0xEEEEu;
    else if (x + 2 >= w) m &= 0x7777u;
    if (y == 0) m &= 0xFFF0u;
    else if (y + 2 >= h) m &= 0x0FFFu;
    return m;
}

// The six tested pixels: (dy, dx) offset from the block's TL pixel, mask
// bit, and the neighbour block holding the pixel (dby, dbx).
struct Probe { int8_t dy, dx, bit, dby, dbx; };
__constant__ Probe PROBES[6] = {
    {-1, -1, 0, -1, -1},  // P
    {-1, 0, 1, -1, 0},    // Q
    {-1, 1, 2, -1, 0},    // Q
    {-1, 2, 3, -1, 1},    // R
    {0, -1, 4, 0, -1},    // S
    {1, -1, 8, 0, -1},    // S
};

ATQ_DEV uint32_t lds_find(const uint32_t *lp, uint32_t n) {
    while (true) {
        uint32_t p = lp[n];
        if (p == n) return n;
        n = p;
    }
}

ATQ_DEV void lds_union(uint32_t *lp, uint32_t a, uint32_t b) {
    while (true) {
        a = lds_find(lp, a);
        b = lds_find(lp, b);
        if (a < b) {
            uint32_t old = atomicMin(&lp[b], a);
            if (old == b) return;
            b = old;
        } else if (b < a) {
            uint32_t old = atomicMin(&lp[a], b);
            if (old == a) return;
            a = old;
        } else {
            return;
        }
    }
}

ATQ_DEV uint32_t g_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of n by plain (cacheable) loads, once every union is done: words
// change afterwards only by path halving, which lowers a word to another
// ancestor of the same root, so a stale value still leads there (indices
// decrease along the path, so the walk ends).
ATQ_DEV uint32_t g_find_settled(const uint32_t *parent, uint32_t n) {
    while (true) {
        uint32_t p = parent[n] & ~FLAG;
        if (p == n) return n;
        n = p;
    }
}

// Root of local root n (a FLAG-free word), halving the path on the way:
// every visited word is lowered to its grandparent with atomicMin, which
// keeps it an ancestor in the same tree while other threads halve the same
// path. Called once every union is done, so the loads may be served by the
// caches (a stale word is still an ancestor of the same root).
ATQ_DEV uint32_t g_find_halving(uint32_t *parent, uint32_t n) {
    while (true) {
        const uint32_t p = parent[n];
        if (p == n) return n;
        const uint32_t gp = parent[p];
        if (gp == p) return p;
        atomicMin(&parent[n], gp);
        n = gp;
    }
}

// The roots of a and b, both paths walked (and halved) in lockstep so that
// their loads are in flight together. The loads may be served by the caches:
// a stale word is still an ancestor (words only decrease towards the root),
// so a walk ends at a node that is or was a root, and the union's atomicMin
// returns the fresh word when that node has been linked meanwhile.
ATQ_DEV void g_find_pair(uint32_t *parent, uint32_t &a, uint32_t &b) {
    bool ra = false, rb = false;  // root reached
    while (!ra || !rb) {
        const uint32_t pa = ra ? a : parent[a];
        const uint32_t pb = rb ? b : parent[b];
        ra = ra || pa == a;
        rb = rb || pb == b;
        const uint32_t ga = ra ? a : parent[pa];
        const uint32_t gb = rb ? b : parent[pb];
        if (!ra) {
            if (ga == pa) {
                a = pa;
                ra = true;
            } else {
                atomicMin(&parent[a], ga);
                a = ga;
            }
        }
        if (!rb) {
            if (gb == pb) {
                b = pb;
                rb = true;
            } else {
                atomicMin(&parent[b], gb);
                b = gb;
            }
        }
    }
}

ATQ_DEV void g_union(uint32_t *parent, uint32_t a, uint32_t b) {
    while (true) {
        g_find_pair(parent, a, b);
        if (a < b) {
            uint32_t old = atomicMin(&parent[b], a);
            if (old == b) return;
            b = old;
        } else if (b < a) {
            uint32_t old = atomicMin(&parent[a], b);
            if (old == a) return;
            a = old;
        } else {
            return;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Compat labelling. Node index of the colour-c node of the block at block
// coordinates (bx, by), i.e. pixels (2bx .. 2bx+1, 2by .. 2by+1):
//   n = by * w + 2 * bx + c   (< w * h / 2),
// the reference's node id (its top-left pixel index + c,
// src/segmentation.cpp:257-259) being gid = n + by * w. Both numberings order
// the nodes alike, so min-linking on node indices yields the reference's
// roots, and the parent array needs only w * h / 2 words.
ATQ_DEV uint32_t node_gid(uint32_t n, uint32_t w) { return n + (n / w) * w; }

constexpr uint32_t CNT_OUT = 0x80000000u;  // the local root has an edge leaving the tile
constexpr uint32_t ROOT_BITS = 11;         // local node index in a state word

// Scratch words of a frame: the parent array (w * h / 2 words, written only
// at live local roots and at the tile-border blocks' nodes), then one state
// word per block (w * h / 4 words): its two nodes' local roots and its
// pixels' colour codes.
ATQ_DEV uint32_t *ccl_state(const FrontFrame &F) { return F.parent + (size_t)F.w * F.h / 2; }

// Border entries of the tiles' edge blocks, held in the labels buffer until
// the last kernel writes the labels (the parent array keeps only live
// local roots): for every tile row, the entries of its top (side 0) and
// bottom (side 1) block rows across the whole frame; for every tile column,
// those of its left (0) and right (1) block columns down the whole frame.
// 2 * tile_rows * w + 2 * tile_cols * h words, fewer than w * h for any
// frame of at least 8 x 8 pixels. An entry is two words, the local roots of
// the block's white and black nodes (NODE_BITS bits) with the 2-bit codes
// of the block's six probes that leave the tile in their top bits (probes
// 0-2 in the first word, 3-5 in the second).
constexpr uint32_t NODE_BITS = 25;  // node indices < w * h / 2 <= 2^25
constexpr uint32_t NODE_MASK = (1u << NODE_BITS) - 1;

struct BorderMaps {
    uint32_t *base;
    uint32_t w, h, tile_rows;
    ATQ_DEV uint2 *row(uint32_t tr, uint32_t side, uint32_t bxg) const {
        return reinterpret_cast<uint2 *>(base) + (size_t)(2 * tr + side) * (w / 2) + bxg;
    }
    ATQ_DEV uint2 *col(uint32_t tc, uint32_t side, uint32_t byg) const {
        return reinterpret_cast<uint2 *>(base) + (size_t)tile_rows * w + (size_t)(2 * tc + side) * (h / 2) + byg;
    }
};

template <int TBY>
ATQ_DEV BorderMaps border_maps(const FrontFrame &F) {
    return BorderMaps{F.labels, F.w, F.h, (F.h + 2 * TBY - 1) / (2 * TBY)};
}

template <int TBY>
struct LocalLds {
    static constexpr int NODES = 2 * TBX * TBY;
    uint32_t lp[NODES];   // local parent (local node indices)
    uint32_t cnt[NODES];  // pixel count of each local root, CNT_OUT if it has an outgoing edge
    uint8_t part[NODES];  // the node is a run head that takes part in a union pair
    uint32_t flag[3];     // barrier_or
};

// Two bytes (columns x, x+1) of row y, or 127s outside the frame.
ATQ_DEV uint32_t ld_pair(const uint8_t *thr, int x, int y, uint32_t w, uint32_t h) {
    return (y >= 0 && y < (int)h && x < (int)w)
               ? (uint32_t)*reinterpret_cast<const uint16_t *>(thr + (size_t)y * w + x)
               : 0x7F7Fu;
}
ATQ_DEV uint32_t ld_byte(const uint8_t *thr, int x, int y, uint32_t w, uint32_t h) {
    return (y >= 0 && y < (int)h && x >= 0 && x < (int)w) ? (uint32_t)thr[(size_t)y * w + x] : 127u;
}
ATQ_DEV uint32_t colour_code(uint32_t v) { return v == C_WHITE ? 1u : (v == C_BLACK ? 2u : 0u); }

// Global node index of local node j of the tile at pixel (x0, y0).
ATQ_DEV uint32_t tile_node(uint32_t j, int x0, int y0, uint32_t w) {
    const uint32_t b = j >> 1;
    return (uint32_t)(y0 / 2 + (int)(b / TBX)) * w + (uint32_t)x0 + 2 * (b % TBX) + (j & 1);
}

// Work-group barrier that also returns whether any thread's pred was set.
// __syncthreads_or costs about 600 clocks here (two barriers and a
// reduction), this one a barrier and an LDS round trip: flag f[s % 3] is set
// by the threads whose pred holds, read after the barrier, and the flag of
// step s + 2 (last read before this barrier) is cleared for reuse; s is the
// work-group-uniform step counter.
ATQ_DEV bool barrier_or(uint32_t *f, uint32_t &s, bool pred) {
    if (pred) f[s % 3] = 1;
    __syncthreads();
    const bool any = f[s % 3] != 0;
    if (threadIdx.x == 0) f[(s + 2) % 3] = 0;
    s++;
    return any;
}

// One tile of frame F (32 * TBY threads; thread t owns block (t % 32, t / 32)).
//  1. Every thread loads its block's pixels and the row above them as byte
//     pairs straight into registers (the columns left and right of a pair
//     come from the neighbouring lanes, the tile's halo columns from extra
//     loads of the edge lanes), then builds the block's neighbour mask and
//     probes exactly as the reference (src/segmentation.cpp:137-255).
//  2. A block whose S probes reach its left neighbour in colour c continues
//     that neighbour's run: one ballot per block row and colour points every
//     node at the first node of its run.
//  3. The P, Q, Q, R probes join runs to the row above through an LDS
//     union-find run in rounds of hooking (atomicMin of the larger root of
//     each pair onto the smaller) and pointer jumping.
//  4. Count each local root's pixels, and mark the local roots that own an
//     edge leaving the tile.
//  5. Publish: a state word per block (both nodes' local roots and the
//     pixels' colour codes); the parent word of every node of the tile's
//     border blocks (its local root) and of every live local root (itself:
//     a root with pixels or an outgoing edge); sizes[] at live roots.
template <int TBY>
ATQ_DEV void ccl_local_tile(LocalLds<TBY> &L, const FrontFrame &F, uint32_t tile, uint32_t tiles_x) {
    constexpr int NODES = LocalLds<TBY>::NODES, NT = TBX * TBY;
    static_assert(TBX == 32 && TBY % 2 == 0 && NODES <= (1 << ROOT_BITS), "a block row is a half-wave");
    const uint8_t *__restrict__ thr = F.thr;
    const uint32_t w = F.w, h = F.h;
    uint32_t *__restrict__ parent = F.parent;
    const int tid = threadIdx.x, lane = tid & 63, bx = tid & 31, by = tid / 32;
    const int x0 = (int)(tile % tiles_x) * (2 * TBX), y0 = (int)(tile / tiles_x) * (2 * TBY);
    const int x = x0 + 2 * bx, y = y0 + 2 * by;
    const bool in_img = x < (int)w && y < (int)h;

    // 1. loads: rows y-1, y, y+1 at columns x, x+1; the tile's halo columns
    // x0-1 (edge lane 0) and x0+64 (edge lane 31)
    const uint32_t ya = ld_pair(thr, x, y - 1, w, h);
    const uint32_t yt = ld_pair(thr, x, y, w, h);
    const uint32_t yb = ld_pair(thr, x, y + 1, w, h);
    uint32_t el0 = 127u, el1 = 127u, el2 = 127u, er = 127u;
    if (bx == 0) {
        el0 = ld_byte(thr, x0 - 1, y - 1, w, h);
        el1 = ld_byte(thr, x0 - 1, y, w, h);
        el2 = ld_byte(thr, x0 - 1, y + 1, w, h);
    }
    if (bx == TBX - 1) er = ld_byte(thr, x0 + 2 * TBX, y - 1, w, h);
    L.cnt[2 * tid] = 0;
    L.cnt[2 * tid + 1] = 0;
    L.part[2 * tid] = 0;
    L.part[2 * tid + 1] = 0;
    if (tid < 3) L.flag[tid] = 0;

    // probes (2 bits each: 0 not tested or 127, 1 white, 2 black), runs;
    // the left neighbour's column x-1 (rows y-1, y, y+1) and the right
    // neighbour's column x+2 (row y-1)
    // (lane moves across the whole wave: the edge lanes bx == 0 and
    // bx == TBX - 1, whose neighbours lie in the next block row or tile, use
    // their own halo loads instead)
    const uint32_t lpk = lane_below(ya | (yt << 16));
    const uint32_t lb = lane_below(yb);
    const uint32_t rpk = lane_above(ya);
    const uint32_t pP = bx ? (lpk >> 8) & 0xFFu : el0;
    const uint32_t pS0 = bx ? lpk >> 24 : el1;
    const uint32_t pS1 = bx ? (lb >> 8) & 0xFFu : el2;
    const uint32_t pR = bx < TBX - 1 ? rpk & 0xFFu : er;
    const uint32_t probe[6] = {pP, ya & 0xFFu, ya >> 8, pR, pS0, pS1};
    uint32_t code = 0;
    if (in_img) {
        const uint32_t m = check_mask((uint8_t)(yt & 0xFFu), (uint8_t)(yt >> 8), (uint8_t)(yb & 0xFFu),
                                      (uint32_t)x, (uint32_t)y, w, h);
#pragma unroll
        for (int k = 0; k < 6; k++) {
            constexpr uint32_t BIT[6] = {0, 1, 2, 3, 4, 8};
            if (m & (1u << BIT[k])) code |= colour_code(probe[k]) << (2 * k);
        }
    }
    // edges leaving the tile: P (left or upper tiles), Q (tile above), R
    // (right or upper tiles), S (left tile)
    uint32_t out = 0, xcode = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const bool cross = k == 0   ? (by == 0 || bx == 0)
                           : k < 3  ? by == 0
                           : k == 3 ? (by == 0 || bx == TBX - 1)
                                    : bx == 0;
        const uint32_t c = (code >> (2 * k)) & 3u;
        if (cross && c) {
            out |= 1u << (c - 1);
            xcode |= c << (2 * k);
        }
    }
    // what the last steps need of the pixels, in one word: the pixels'
    // colour codes (8 bits), white and black counts (3 bits each), the
    // outgoing-edge colours (2 bits) and the codes of the probes that
    // leave the tile (12 bits)
    uint32_t info;
    {
        uint32_t codes = 0, nw = 0, nb = 0;
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const uint32_t t = (yt >> (8 * j)) & 0xFFu, u = (yb >> (8 * j)) & 0xFFu;
            codes |= colour_code(t) << (2 * j);
            codes |= colour_code(u) << (2 * j + 4);
            nw += (t == C_WHITE) + (u == C_WHITE);
            nb += (t == C_BLACK) + (u == C_BLACK);
        }
        info = codes | (nw << 8) | (nb << 11) | (out << 14) | (xcode << 16);
    }
    const uint64_t half = lane < 32 ? 0x00000000FFFFFFFFull : 0xFFFFFFFF00000000ull;
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    uint32_t head[2];
#pragma unroll
    for (uint32_t c = 0; c < 2; c++) {
        const uint32_t cc = c + 1;
        const bool left = bx > 0 && (((code >> 8) & 3u) == cc || ((code >> 10) & 3u) == cc);
        const uint64_t starts = __ballot(!left) & half & upto;
        const int run0 = (63 - __clzll((long long)starts)) & 31;
        head[c] = 2 * (by * TBX + run0) + c;
        L.lp[2 * tid + c] = head[c];
    }
    __syncthreads();

    // 3. unions with the row above (P, Q, Q, R), in rounds over the run
    // heads that take part in a pair (the local parent array restricted to
    // them is a forest of stars at the start of each round): every pair
    // hooks the larger of its two roots under the smaller with atomicMin
    // (if that root was hooked meanwhile, the pair of the smaller root and
    // the word atomicMin found joins in the next round), then pointer
    // jumping turns the forest into stars again; until no pair is left. Concurrent unions with retries would instead walk the chains
    // that simultaneous hooks of consecutive rows build, one LDS round trip
    // per step. Nodes that are not run heads keep pointing at their head.
    uint32_t A[4], B[4], vm = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int nbx = bx + (k == 0 ? -1 : (k == 3 ? 1 : 0));
        const uint32_t cd = (code >> (2 * k)) & 3u;
        uint32_t a = 0xFFFFFFFFu, b = 0xFFFFFFFFu;
        if (cd && by > 0 && nbx >= 0 && nbx < TBX) {
            a = head[cd - 1];
            b = L.lp[2 * ((by - 1) * TBX + nbx) + cd - 1];
        }
        const uint32_t pa = lane_below(a), pb = lane_below(b);
        bool dup = (lane & 31) > 0 && pa == a && pb == b;
#pragma unroll
        for (int q = 0; q < k; q++) dup = dup || (((vm >> q) & 1u) && A[q] == a && B[q] == b);
        A[k] = a;
        B[k] = b;
        if (a != 0xFFFFFFFFu && !dup) {
            vm |= 1u << k;
            L.part[a] = 1;
            L.part[b] = 1;
        }
    }
    const uint32_t n0 = 2 * tid, n1 = 2 * tid + 1;
    bool jump0 = false, jump1 = false;  // this thread's nodes are heads in a pair
    uint32_t step = 0;
    for (bool first = true;; first = false) {
        uint32_t ra[4], rb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if ((vm >> i) & 1u) {
                ra[i] = L.lp[A[i]];
                rb[i] = L.lp[B[i]];
            }
        }
        // a hook whose root was hooked meanwhile (atomicMin finds another
        // word than the root) leaves the pair (lower root, that word) to
        // join next round; every other hook is final
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (!((vm >> i) & 1u)) continue;
            if (ra[i] == rb[i]) {
                vm &= ~(1u << i);  // joined
                continue;
            }
            const uint32_t lo = min(ra[i], rb[i]), hi = max(ra[i], rb[i]);
            const uint32_t was = atomicMin(&L.lp[hi], lo);
            if (was == hi) {
                vm &= ~(1u << i);
            } else {
                A[i] = lo;
                B[i] = was;
            }
        }
        const bool again = barrier_or(L.flag, step, vm != 0);
        if (first) {
            jump0 = L.part[n0] != 0;
            jump1 = L.part[n1] != 0;
        }
        // pointer jumping, unsynchronised: this round's hooks are done, so
        // every walk ends at its root (a word another thread lowers
        // meanwhile still names an ancestor); both walks in lockstep
        uint32_t p0 = jump0 ? L.lp[n0] : n0, p1 = jump1 ? L.lp[n1] : n1;
        while (true) {
            const uint32_t g0 = jump0 ? L.lp[p0] : p0, g1 = jump1 ? L.lp[p1] : p1;
            if (g0 == p0 && g1 == p1) break;
            p0 = g0;
            p1 = g1;
        }
        if (jump0) L.lp[n0] = p0;
        if (jump1) L.lp[n1] = p1;
        __syncthreads();
        if (!again) break;
    }
    // the heads are stars: every node's local root is its head's parent
    const uint32_t r0 = L.lp[head[0]], r1 = L.lp[head[1]];

    // pixel counts and outgoing edges per local root (the roots of this
    // block's nodes are r0, r1)
    {
        const uint32_t nw = (info >> 8) & 7u, nb = (info >> 11) & 7u;
        if (nw) atomicAdd(&L.cnt[r0], nw);
        if (nb) atomicAdd(&L.cnt[r1], nb);
        if (info & (1u << 14)) atomicOr(&L.cnt[r0], CNT_OUT);
        if (info & (2u << 14)) atomicOr(&L.cnt[r1], CNT_OUT);
    }
    __syncthreads();

    // 5. publish
    if (in_img) {
        const uint32_t xcode = info >> 16;
        ccl_state(F)[(size_t)(y / 2) * (w / 2) + (uint32_t)(x / 2)] =
            r0 | (r1 << ROOT_BITS) | ((info & 0xFFu) << (2 * ROOT_BITS));
        if (by == 0 || by == TBY - 1 || bx == 0 || bx == TBX - 1) {
            // the block's border entry: both nodes' local roots, the codes
            // of its probes that leave the tile in their top bits
            const uint2 e = make_uint2(tile_node(r0, x0, y0, w) | ((xcode & 0x3Fu) << NODE_BITS),
                                       tile_node(r1, x0, y0, w) | ((xcode >> 6) << NODE_BITS));
            const BorderMaps M = border_maps<TBY>(F);
            const uint32_t tr = (uint32_t)y0 / (2 * TBY), tc = (uint32_t)x0 / (2 * TBX);
            const uint32_t bxg = (uint32_t)x / 2, byg = (uint32_t)y / 2;
            if (by == 0) *M.row(tr, 0, bxg) = e;
            if (by == TBY - 1) *M.row(tr, 1, bxg) = e;
            if (bx == 0) *M.col(tc, 0, byg) = e;
            if (bx == TBX - 1) *M.col(tc, 1, byg) = e;
        }
    }
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const int j = 2 * tid + c;
        const uint32_t cn = L.cnt[j];
        if ((c ? r1 : r0) != (uint32_t)j || cn == 0) continue;
        const uint32_t n = tile_node(j, x0, y0, w);
        parent[n] = n;
        F.sizes[node_gid(n, w)] = cn & ~CNT_OUT;
    }
    // sizes[0] stands for 127 pixels and out-of-image neighbours
    // (src/find_boundaries.cpp:49-55): node 0 is always a root
    if (tile == 0 && tid == 0) F.sizes[0] = L.cnt[0] & ~CNT_OUT;
    (void)NT;
    __syncthreads();  // the LDS is reused by the next tile
}

template <int TBY>
__global__ void __launch_bounds__(TBX * TBY) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_ccl_local(FrontTable tab) {
    __shared__ LocalLds<TBY> L;
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        const uint32_t tiles_x = (F.w + 2 * TBX - 1) / (2 * TBX);
        const uint32_t n_tiles = tiles_x * ((F.h + 2 * TBY - 1) / (2 * TBY));
        for (uint32_t it = blockIdx.x; it < xcd_grid(n_tiles); it += gridDim.x) {
            const uint32_t tile = xcd_tile(it, n_tiles);
            if (tile < n_tiles) ccl_local_tile<TBY>(L, F, tile, tiles_x);
        }
    }
}

// One wave per tile; lane k takes border block k (top row, then the left
// and right columns; and block k + 64 of tiles with more than 64 border
// blocks) and the edges that leave the tile from it, read
// from the border entries the local pass wrote: its own entry (local roots
// and the codes of its probes that leave the tile) and the entries of the
// blocks those probes reach (their local roots), one round of independent
// loads. Each edge is a pair of local roots; the wave keeps the distinct
// pairs in an LDS hash set, the lane that inserts a pair appends it to the
// wave's edge list, and then the lanes merge the listed pairs one each on
// the global parent array (lock-free atomicMin union,
// include/union_find.h:50-77), so that the unions' chains of dependent
// atomics run side by side rather than one after another in the lanes
// that found several edges. A tile border is crossed by few components, so
// a tile issues a handful of global unions.
template <int TBY>
constexpr int border_items() { return TBX + 2 * (TBY - 1); }
constexpr int BORDER_SLOTS = 512;
constexpr int BORDER_LIST = 256;  // listed edges per tile (more: merged where found)

// The border edges of one tile of frame F (one wave; lane k).
template <int TBY>
ATQ_DEV void ccl_border_tile(uint64_t *hs, uint64_t *el, uint32_t *en, const FrontFrame &F,
                             uint32_t tile, uint32_t tiles_x, uint32_t n_tiles) {
    constexpr int ITEMS = border_items<TBY>();
    static_assert(ITEMS <= 128, "at most two border blocks per lane");
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    constexpr uint64_t EMPTY = ~0ull;
    const uint32_t w = F.w, h = F.h;
    uint32_t *__restrict__ parent = F.parent;
    const BorderMaps M = border_maps<TBY>(F);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tr = tile / tiles_x, tc = tile % tiles_x;
#pragma unroll
    for (uint32_t k = lane; k < (uint32_t)((ITEMS + 63) / 64 * 64); k += 64) {
        bool live = tile < n_tiles && k < (uint32_t)ITEMS;
        int bx = 0, by = 0, x = 0, y = 0;
        if (live) {
            if (k < (uint32_t)TBX) { bx = k; by = 0; }
            else if (k < (uint32_t)(TBX + TBY - 1)) { bx = 0; by = 1 + (k - TBX); }
            else { bx = TBX - 1; by = 1 + (k - TBX - (TBY - 1)); }
            x = (int)tc * 2 * TBX + 2 * bx;
            y = (int)tr * 2 * TBY + 2 * by;
            live = x < (int)w && y < (int)h;
        }
        const uint32_t bxg = (uint32_t)x / 2, byg = (uint32_t)y / 2;
        uint2 own = make_uint2(0u, 0u), nb[6];
        bool cross[6];
        if (live) own = by == 0 ? *M.row(tr, 0, bxg) : *M.col(tc, bx == 0 ? 0 : 1, byg);
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const Probe q = PROBES[i];
            const int nbx = bx + q.dbx, nby = by + q.dby;
            const int ny = y + 2 * q.dby, nx = x + 2 * q.dbx;
            cross[i] = live && !(nbx >= 0 && nbx < TBX && nby >= 0) && ny >= 0 && nx >= 0 && nx < (int)w;
            nb[i] = make_uint2(0u, 0u);
            if (cross[i]) {
                // the row above: the bottom row of the tile row above; else the
                // facing column of the left or right tile
                if (by == 0 && q.dby < 0) nb[i] = *M.row(tr - 1, 1, (uint32_t)(nx / 2));
                else nb[i] = *M.col(q.dbx < 0 ? tc - 1 : tc + 1, q.dbx < 0 ? 1 : 0, (uint32_t)(ny / 2));
            }
        }
        const uint32_t codes = (own.x >> NODE_BITS) | ((own.y >> NODE_BITS) << 6);
        uint32_t ea[6], eb[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            ea[i] = NONE;
            eb[i] = NONE;
            const uint32_t cd = (codes >> (2 * i)) & 3u;  // 1 white, 2 black
            if (cross[i] && cd) {
                const uint32_t a = (cd == 1 ? own.x : own.y) & NODE_MASK;
                const uint32_t b = (cd == 1 ? nb[i].x : nb[i].y) & NODE_MASK;
                ea[i] = min(a, b);
                eb[i] = max(a, b);
            }
        }
#pragma unroll
        for (int i = 0; i < 6; i++) {
            if (ea[i] == NONE || ea[i] == eb[i]) continue;
            const uint64_t key = ((uint64_t)ea[i] << 32) | eb[i];
            uint32_t hsl = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & (BORDER_SLOTS - 1);
            bool mine = true;  // the set is full: merge it here
            for (int p = 0; p < BORDER_SLOTS; p++, hsl = (hsl + 1) & (BORDER_SLOTS - 1)) {
                const uint64_t old = atomicCAS((unsigned long long *)&hs[hsl], (unsigned long long)EMPTY,
                                               (unsigned long long)key);
                if (old == EMPTY) break;
                if (old == key) {
                    mine = false;
                    break;
                }
            }
            if (mine) {
                const uint32_t at = atomicAdd(en, 1u);
                if (at < (uint32_t)BORDER_LIST) el[at] = key;
                else g_union(parent, ea[i], eb[i]);
            }
        }
    }
    // the wave's LDS operations complete in order: every listed edge is in
    // place once the count is read back
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t n = min(*en, (uint32_t)BORDER_LIST);
    for (uint32_t j = lane; j < n; j += 64) {
        const uint64_t e = el[j];
        g_union(parent, (uint32_t)(e >> 32), (uint32_t)e);
    }
    (void)h;
}

// Four tiles per work-group round, one per wave.
template <int TBY>
__global__ void __launch_bounds__(256)
k_ccl_border(FrontTable tab) {
    __shared__ uint64_t set[4][BORDER_SLOTS];
    __shared__ uint64_t list[4][BORDER_LIST];
    __shared__ uint32_t list_n[4];
    const uint32_t wave = threadIdx.x >> 6, k = threadIdx.x & 63;
    uint64_t *hs = set[wave];
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        const uint32_t tiles_x = (F.w + 2 * TBX - 1) / (2 * TBX);
        const uint32_t n_tiles = tiles_x * ((F.h + 2 * TBY - 1) / (2 * TBY));
        for (uint32_t g = blockIdx.x; g < (n_tiles + 3) / 4; g += gridDim.x) {
            for (int i = k; i < BORDER_SLOTS; i += 64) hs[i] = ~0ull;
            if (k == 0) list_n[wave] = 0;
            __syncthreads();
            ccl_border_tile<TBY>(hs, list[wave], &list_n[wave], F, 4 * g + wave, tiles_x, n_tiles);
            __syncthreads();
        }
    }
}

// Labels of TPW tiles after the border merge (all 256 threads; the tiles
// are interleaved so that each thread has several independent loads in
// flight per step). Every local root that holds pixels is resolved once:
// its final root is found on the parent array (path halving, all of the
// thread's walks in lockstep), a root merged by the border pass adds its
// pixel count into the final root's size, and the final root's reference
// id goes to an LDS table; then every block's four label words are written
// from its state word (white = 0x80000000 | root, black = root, 127 -> 0),
// two blocks per thread per row, 16 bytes per store. Once a size reaches
// size_cap further adds are skipped (the detector only asks whether a size
// reaches min_component_size; the drop-in passes UINT32_MAX for exact
// sizes).
template <int TBY>
constexpr int label_tiles() { return TBY == 8 ? 2 : 1; }

template <int TBY>
ATQ_DEV void ccl_label_tiles(uint32_t *table, const FrontFrame &F, uint32_t tile0, uint32_t n_tiles,
                             uint32_t tiles_x) {
    constexpr int TPW = label_tiles<TBY>();
    constexpr int NODES = 2 * TBX * TBY, PAIRS = TBX * TBY / 2;
    constexpr int PPT = TPW * PAIRS / 256, NPT = TPW * NODES / 256;
    static_assert(TPW * PAIRS % 256 == 0, "whole pairs per thread");
    const uint32_t w = F.w, h = F.h, size_cap = F.size_cap;
    uint32_t *__restrict__ parent = F.parent;
    uint32_t *__restrict__ sizes = F.sizes;
    const uint32_t *__restrict__ state = ccl_state(F);
    const int tid = threadIdx.x;
    auto origin = [&](int sub, int &x0, int &y0) {
        const uint32_t t = tile0 + (uint32_t)sub;
        x0 = (int)(t % tiles_x) * (2 * TBX);
        y0 = (int)(t / tiles_x) * (2 * TBY);
    };
    uint2 st[PPT];
    bool in[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int i = tid + 256 * k, sub = i / PAIRS, pi = i % PAIRS;
        int x0, y0;
        origin(sub, x0, y0);
        const int by = pi / (TBX / 2), bx = 2 * (pi % (TBX / 2));
        const int x = x0 + 2 * bx, y = y0 + 2 * by;
        // w is a multiple of 4: both blocks of a pair are in the frame or neither
        in[k] = tile0 + (uint32_t)sub < n_tiles && x < (int)w && y < (int)h;
        st[k] = in[k] ? *reinterpret_cast<const uint2 *>(state + (size_t)(y / 2) * (w / 2) + (uint32_t)(x / 2))
                      : make_uint2(0u, 0u);
    }
    for (int j = tid; j < TPW * NODES; j += 256) table[j] = 0;
    __syncthreads();
    constexpr uint32_t RM = (1u << ROOT_BITS) - 1;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        if (!in[k]) continue;
        uint32_t *tb = table + ((tid + 256 * k) / PAIRS) * NODES;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t s = e ? st[k].y : st[k].x, codes = s >> (2 * ROOT_BITS);
            // codes: 2 bits per pixel, 01 white, 10 black
            if (codes & 0x55u) tb[s & RM] = 1;
            if (codes & 0xAAu) tb[(s >> ROOT_BITS) & RM] = 1;
        }
    }
    __syncthreads();
    {
        uint32_t n[NPT], cur[NPT], act = 0;
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int j = tid + 256 * k, sub = j / NODES;
            int x0, y0;
            origin(sub, x0, y0);
            n[k] = tile_node((uint32_t)(j % NODES), x0, y0, w);
            cur[k] = n[k];
            if (table[j]) act |= 1u << k;
        }
        // the pixel counts of the tile's local roots, loaded alongside the
        // walks (only a merged root's count is used, and no other
        // work-group adds into a merged root's count)
        uint32_t sz[NPT];
#pragma unroll
        for (int k = 0; k < NPT; k++) sz[k] = (act >> k) & 1u ? sizes[node_gid(n[k], w)] : 0u;
        // final roots, path halving (every word lowered stays an ancestor)
        uint32_t todo = act;
        while (todo) {
            uint32_t p[NPT], g[NPT];
#pragma unroll
            for (int k = 0; k < NPT; k++)
                if ((todo >> k) & 1u) p[k] = parent[cur[k]];
#pragma unroll
            for (int k = 0; k < NPT; k++)
                if ((todo >> k) & 1u) g[k] = p[k] == cur[k] ? p[k] : parent[p[k]];
#pragma unroll
            for (int k = 0; k < NPT; k++) {
                if (!((todo >> k) & 1u)) continue;
                if (p[k] == cur[k]) {
                    todo &= ~(1u << k);
                } else if (g[k] == p[k]) {
                    cur[k] = p[k];
                    todo &= ~(1u << k);
                } else {
                    atomicMin(&parent[cur[k]], g[k]);
                    cur[k] = g[k];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            if (!((act >> k) & 1u)) continue;
            const uint32_t rg = node_gid(cur[k], w);
            if (cur[k] != n[k] && sz[k] && g_load(&sizes[rg]) < size_cap) atomicAdd(&sizes[rg], sz[k]);
            table[tid + 256 * k] = rg;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        if (!in[k]) continue;
        const int i = tid + 256 * k, sub = i / PAIRS, pi = i % PAIRS;
        const uint32_t *tb = table + sub * NODES;
        int x0, y0;
        origin(sub, x0, y0);
        const int by = pi / (TBX / 2), bx = 2 * (pi % (TBX / 2));
        const int x = x0 + 2 * bx, y = y0 + 2 * by;
        uint32_t top[4], bot[4];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint32_t s = e ? st[k].y : st[k].x, codes = s >> (2 * ROOT_BITS);
            const uint32_t lw = LABEL_PIXEL_MASK | tb[s & RM], lb = tb[(s >> ROOT_BITS) & RM];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t c = (codes >> (2 * j)) & 3u;
                const uint32_t lab = c == 1 ? lw : (c == 2 ? lb : 0u);
                if (j < 2) top[2 * e + j] = lab;
                else bot[2 * e + j - 2] = lab;
            }
        }
        *reinterpret_cast<uint4 *>(F.labels + (size_t)y * w + x) = make_uint4(top[0], top[1], top[2], top[3]);
        *reinterpret_cast<uint4 *>(F.labels + (size_t)(y + 1) * w + x) = make_uint4(bot[0], bot[1], bot[2], bot[3]);
    }
    __syncthreads();  // the table is reused by the next tiles
}

template <int TBY>
__global__ void __launch_bounds__(256)
k_ccl_label(FrontTable tab) {
    constexpr int TPW = label_tiles<TBY>();
    __shared__ uint32_t table[TPW * 2 * TBX * TBY];
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        front_clear(F);  // the later stages' tables and counts
        const uint32_t tiles_x = (F.w + 2 * TBX - 1) / (2 * TBX);
        const uint32_t n_tiles = tiles_x * ((F.h + 2 * TBY - 1) / (2 * TBY));
        const uint32_t groups = (n_tiles + TPW - 1) / TPW;
        for (uint32_t it = blockIdx.x; it < xcd_grid(groups); it += gridDim.x) {
            const uint32_t g = xcd_tile(it, groups);
            if (g < groups) ccl_label_tiles<TBY>(table, F, g * TPW, n_tiles, tiles_x);
        }
    }
}

namespace {
template <int TBY>
void ccl_launch(hipStream_t s, const FrontTable &tab) {
    uint32_t tiles = 0;
    for (uint32_t i = 0; i < tab.n; i++) {
        const uint32_t w = tab.fr[i].w, h = tab.fr[i].h;
        tiles = std::max(tiles, (uint32_t)(atq_div_up(w, 2 * TBX) * atq_div_up(h, 2 * TBY)));
    }
    // persistent grids: the largest frame's items, capped well above what
    // the device holds at once (frames up to 8192 x 8192 loop a few times)
    const uint32_t lg = std::min<uint32_t>(xcd_grid(tiles), 8192u);
    const uint32_t bg = std::min<uint32_t>(atq_div_up(tiles, 4), 4096u);
    k_ccl_local<TBY><<<lg ? lg : 8, TBX * TBY, 0, s>>>(tab);
    k_ccl_border<TBY><<<bg ? bg : 1, 256, 0, s>>>(tab);
    const uint32_t labg = std::min<uint32_t>(xcd_grid(atq_div_up(tiles, label_tiles<TBY>())), 8192u);
    k_ccl_label<TBY><<<labg ? labg : 8, 256, 0, s>>>(tab);
}
}  // namespace

int launch_ccl_frames(hipStream_t s, const FrontFrame *frames, int n, const FrontShapes &sh) {
    if (n < 1 || n > FRONT_MAX_FRAMES || !front_shapes_supported(sh)) return ATQ_ERR_BAD_ARG;
    FrontTable tab{};
    for (int i = 0; i < n; i++) tab.fr[i] = frames[i];
    tab.n = (uint32_t)n;
    if (sh.ccl_rows == 8) ccl_launch<8>(s, tab);
    else if (sh.ccl_rows == 32) ccl_launch<32>(s, tab);
    else ccl_launch<16>(s, tab);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_ccl(hipStream_t s, const uint8_t *thr, size_t w, size_t h, uint32_t *parent,
               uint32_t *labels, uint32_t *sizes, uint32_t size_cap) {
    FrontFrame fr{};
    fr.thr = const_cast<uint8_t *>(thr);
    fr.parent = parent;
    fr.labels = labels;
    fr.sizes = sizes;
    fr.w = (uint32_t)w;
    fr.h = (uint32_t)h;
    fr.size_cap = size_cap;
    return launch_ccl_frames(s, &fr, 1, FrontShapes{});
}


// ---------------------------------------------------------------------------
// Fixed mode: pixel-level labelling, white 8-connected, black 4-connected,
// 127 pixels unlabelled (the reference's block graph instead links colour
// sets through a colour-agnostic neighbour mask, src/segmentation.cpp:139-255).
// Labels: white 0x80000000 | (root + 1), black root + 1, 127 -> 0, root the
// component's minimum pixel index; sizes[root + 1] = its pixel count
// (saturating near size_cap like the compat form), sizes[0] = 0, so label 0
// (127 pixels, out-of-image neighbours) never reads as a large component.
//
// Four steps (the compat form's first two; roots and labels apart):
//  1. k_fx_local: a work-group owns a 64 x 16 pixel tile, each wave four of
//     its rows, one pixel per lane. One ballot per row gives every pixel the
//     first pixel of its same-colour run (the horizontal links); the links to
//     the row above (up; up-left and up-right for white) go through an LDS
//     union-find, a union that repeats the previous lane's pair skipped.
//     Writes parent words (FLAG | local root for non-roots) and each local
//     root's count at sizes[root + 1].
//  2. k_fx_border: the edges that leave a tile - every 8-neighbour edge is
//     owned by its later pixel in raster order (left, up-left, up, up-right),
//     so a tile owns those of its top row (up, up-left, up-right), its left
//     column (left, up-left) and its right column (up-right) - as distinct
//     pairs of local roots, merged on the global parent array.
//  3. k_fx_flatten: each merged local root finds its final root (path
//     halving), points at it and adds its count into sizes[root + 1].
//  4. k_fx_final: the label words, one load per non-root pixel.
namespace {
constexpr int FX_TX = 64, FX_TY = 16;  // pixel tile: one row per wave lane set
constexpr int FX_PX = FX_TX * FX_TY;
constexpr int FX_BORDER_ITEMS = FX_TX + FX_TY + (FX_TY - 1);  // top, left, right (rows >= 1)

struct FxLds {
    uint32_t px[FX_TY][FX_TX / 4];  // 4 pixels per word
    uint32_t lp[FX_PX];
    uint32_t cnt[FX_PX];
};

ATQ_DEV uint8_t fx_px(const FxLds &L, int r, int c) {
    return (uint8_t)(L.px[r][c >> 2] >> (8 * (c & 3)));
}

ATQ_DEV uint32_t fx_tiles_x(uint32_t w) { return (w + FX_TX - 1) / FX_TX; }
ATQ_DEV uint32_t fx_tiles(uint32_t w, uint32_t h) { return fx_tiles_x(w) * ((h + FX_TY - 1) / FX_TY); }

ATQ_DEV void fx_local_tile(FxLds &L, const FrontFrame &F, uint32_t tile) {
    const uint32_t w = F.w, h = F.h, tx = fx_tiles_x(w);
    const int x0 = (int)(tile % tx) * FX_TX, y0 = (int)(tile / tx) * FX_TY;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        // 4 pixels per thread; w is a multiple of 4, so a word never
        // straddles the image's right edge
        const int r = tid >> 4, c4 = (tid & 15) * 4;
        const int y = y0 + r, x = x0 + c4;
        L.px[r][c4 >> 2] = (y < (int)h && x < (int)w)
                               ? *reinterpret_cast<const uint32_t *>(F.thr + (size_t)y * w + x)
                               : 0x7F7F7F7Fu;
    }
    for (int i = tid; i < FX_PX; i += 256) L.cnt[i] = 0;
    __syncthreads();
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    // runs: every pixel points at the first pixel of its same-colour run
#pragma unroll
    for (int k = 0; k < FX_TY / 4; k++) {
        const int r = 4 * wave + k;
        const uint8_t v = fx_px(L, r, lane);
        const bool cont = lane > 0 && v != 127 && fx_px(L, r, lane - 1) == v;
        const uint64_t starts = __ballot(!cont) & upto;
        L.lp[r * FX_TX + lane] = r * FX_TX + (63 - __clzll((long long)starts));
    }
    __syncthreads();
    // links to the row above, within the tile
#pragma unroll
    for (int k = 0; k < FX_TY / 4; k++) {
        const int r = 4 * wave + k;
        const uint8_t v = fx_px(L, r, lane);
        const uint32_t me = L.lp[r * FX_TX + lane];
#pragma unroll
        for (int d = -1; d <= 1; d++) {
            const int c = lane + d;
            const bool link = r > 0 && v != 127 && c >= 0 && c < FX_TX && (d == 0 || v == C_WHITE) &&
                              fx_px(L, r - 1, c) == v;
            const uint32_t ra = link ? me : 0xFFFFFFFFu;
            const uint32_t rb = link ? L.lp[(r - 1) * FX_TX + c] : 0xFFFFFFFFu;
            const uint32_t pa = lane_below(ra), pb = lane_below(rb);
            const bool dup = lane > 0 && pa == ra && pb == rb;
            if (link && !dup) lds_union(L.lp, ra, rb);
        }
    }
    __syncthreads();
    for (int i = tid; i < FX_PX; i += 256) L.lp[i] = lds_find(L.lp, i);
    __syncthreads();
    for (int i = tid; i < FX_PX; i += 256)
        if (fx_px(L, i / FX_TX, i % FX_TX) != 127) atomicAdd(&L.cnt[L.lp[i]], 1u);
    __syncthreads();
    for (int i = tid; i < FX_PX; i += 256) {
        const int y = y0 + i / FX_TX, x = x0 + i % FX_TX;
        if (y >= (int)h || x >= (int)w) continue;
        const uint32_t r = L.lp[i];
        const uint32_t gr = (uint32_t)(y0 + (int)(r / FX_TX)) * w + (uint32_t)(x0 + (int)(r % FX_TX));
        const uint32_t gi = (uint32_t)y * w + (uint32_t)x;
        F.parent[gi] = r == (uint32_t)i ? gr : (gr | FLAG);
        if (r == (uint32_t)i) F.sizes[gi + 1] = L.cnt[i];
    }
    __syncthreads();  // the LDS is reused by the next tile
}
}  // namespace

__global__ void __launch_bounds__(256)
k_fx_local(FrontTable tab) {
    __shared__ FxLds L;
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        if (blockIdx.x == 0 && threadIdx.x == 0) F.sizes[0] = 0;
        const uint32_t n_tiles = fx_tiles(F.w, F.h);
        for (uint32_t it = blockIdx.x; it < xcd_grid(n_tiles); it += gridDim.x) {
            const uint32_t tile = xcd_tile(it, n_tiles);
            if (tile < n_tiles) fx_local_tile(L, F, tile);
        }
    }
}

// The border edges of one tile (two waves, item k < FX_BORDER_ITEMS: the
// top row, the left column, then the right column below its first row).
// One round of independent loads: the item's pixel and parent word, each
// probed neighbour's pixel and parent word.
ATQ_DEV void fx_border_tile(uint64_t *hs, const FrontFrame &F, uint32_t tile, uint32_t k) {
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    constexpr uint64_t EMPTY = ~0ull;
    const uint32_t w = F.w, h = F.h, tx = fx_tiles_x(w);
    const int x0 = (int)(tile % tx) * FX_TX, y0 = (int)(tile / tx) * FX_TY;
    int x = 0, y = 0, dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0}, np = 0;
    if (k < (uint32_t)FX_TX) {  // top row: up-left, up, up-right
        x = x0 + (int)k;
        y = y0;
        np = 3;
        dx[0] = -1; dx[1] = 0; dx[2] = 1;
        dy[0] = dy[1] = dy[2] = -1;
    } else if (k < (uint32_t)(FX_TX + FX_TY)) {  // left column: left, up-left
        x = x0;
        y = y0 + (int)(k - FX_TX);
        np = y > y0 ? 2 : 1;
        dx[0] = -1; dy[0] = 0;
        dx[1] = -1; dy[1] = -1;
    } else if (k < (uint32_t)FX_BORDER_ITEMS) {  // right column below row 0: up-right
        x = x0 + FX_TX - 1;
        y = y0 + 1 + (int)(k - FX_TX - FX_TY);
        np = 1;
        dx[0] = 1; dy[0] = -1;
    }
    const bool live = np > 0 && x < (int)w && y < (int)h;
    uint8_t v = 127, pv[3] = {127, 127, 127};
    uint32_t own = 0, nb[3] = {0, 0, 0};
    if (live) {
        const uint32_t gi = (uint32_t)y * w + (uint32_t)x;
        v = F.thr[gi];
        own = g_load(&F.parent[gi]);
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int qx = x + dx[i], qy = y + dy[i];
        if (live && i < np && qx >= 0 && qx < (int)w && qy >= 0) {
            const uint32_t gq = (uint32_t)qy * w + (uint32_t)qx;
            pv[i] = F.thr[gq];
            nb[i] = g_load(&F.parent[gq]);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const bool diag = dx[i] != 0 && dy[i] != 0;
        if (v == 127 || pv[i] != v || (diag && v != C_WHITE)) continue;
        const uint32_t a = own & ~FLAG, b = nb[i] & ~FLAG;
        if (a == b) continue;
        const uint32_t ea = min(a, b), eb = max(a, b);
        const uint64_t key = ((uint64_t)ea << 32) | eb;
        uint32_t hsl = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & (BORDER_SLOTS - 1);
        bool mine = true;  // the set is full: merge it here
        for (int p = 0; p < BORDER_SLOTS; p++, hsl = (hsl + 1) & (BORDER_SLOTS - 1)) {
            const uint64_t old = atomicCAS((unsigned long long *)&hs[hsl], (unsigned long long)EMPTY,
                                           (unsigned long long)key);
            if (old == EMPTY) break;
            if (old == key) {
                mine = false;
                break;
            }
        }
        if (mine) g_union(F.parent, ea, eb);
    }
    (void)NONE;
}

// Two tiles per work-group round, 128 threads each.
__global__ void __launch_bounds__(256)
k_fx_border(FrontTable tab) {
    __shared__ uint64_t set[2][BORDER_SLOTS];
    const uint32_t half = threadIdx.x >> 7, k = threadIdx.x & 127;
    uint64_t *hs = set[half];
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        const uint32_t n_tiles = fx_tiles(F.w, F.h);
        for (uint32_t g = blockIdx.x; g < (n_tiles + 1) / 2; g += gridDim.x) {
            for (int i = k; i < BORDER_SLOTS; i += 128) hs[i] = ~0ull;
            __syncthreads();
            const uint32_t tile = 2 * g + half;
            if (tile < n_tiles) fx_border_tile(hs, F, tile, k);
            __syncthreads();
        }
    }
}

// Every local root (a parent word without FLAG that is not its own index
// is a local root merged by the border pass) gets its final root and adds
// its count into sizes[root + 1]; 4 pixels per thread.
__global__ void __launch_bounds__(256)
k_fx_flatten(FrontTable tab) {
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        uint32_t *__restrict__ parent = F.parent;
        uint32_t *__restrict__ sizes = F.sizes;
        const uint32_t quads = F.w * F.h / 4, size_cap = F.size_cap;
        for (uint32_t q = blockIdx.x * 256 + threadIdx.x; q < quads; q += gridDim.x * 256) {
            const uint4 v4 = *reinterpret_cast<const uint4 *>(parent + 4 * (size_t)q);
            const uint32_t v[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t n = 4 * q + j;
                if ((v[j] & FLAG) || v[j] == n) continue;
                const uint32_t r = g_find_halving(parent, v[j]);
                atomicMin(&parent[n], r);
                const uint32_t sz = sizes[n + 1];
                if (sz && g_load(&sizes[r + 1]) < size_cap) atomicAdd(&sizes[r + 1], sz);
            }
        }
    }
}

// Labels, 4 pixels per thread.
__global__ void __launch_bounds__(256)
k_fx_final(FrontTable tab) {
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        front_clear(F);  // the later stages' tables and counts
        const uint32_t *__restrict__ parent = F.parent;
        const uint32_t quads = F.w * F.h / 4;
        for (uint32_t q = blockIdx.x * 256 + threadIdx.x; q < quads; q += gridDim.x * 256) {
            const uint4 v4 = *reinterpret_cast<const uint4 *>(parent + 4 * (size_t)q);
            const uint32_t t4 = *reinterpret_cast<const uint32_t *>(F.thr + 4 * (size_t)q);
            const uint32_t v[4] = {v4.x, v4.y, v4.z, v4.w};
            uint32_t r[4];
#pragma unroll
            for (int j = 0; j < 4; j++) r[j] = (v[j] & FLAG) ? parent[v[j] & ~FLAG] : v[j];
            uint32_t lab[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t p = (t4 >> (8 * j)) & 0xFFu;
                lab[j] = p == C_WHITE ? (LABEL_PIXEL_MASK | (r[j] + 1))
                                      : (p == C_BLACK ? r[j] + 1 : 0u);
            }
            *reinterpret_cast<uint4 *>(F.labels + 4 * (size_t)q) = make_uint4(lab[0], lab[1], lab[2], lab[3]);
        }
    }
}

int launch_ccl_fixed_frames(hipStream_t s, const FrontFrame *frames, int n) {
    if (n < 1 || n > FRONT_MAX_FRAMES) return ATQ_ERR_BAD_ARG;
    FrontTable tab{};
    uint32_t tiles = 0, quads = 0;
    for (int i = 0; i < n; i++) {
        tab.fr[i] = frames[i];
        const uint32_t w = frames[i].w, h = frames[i].h;
        tiles = std::max(tiles, (uint32_t)(atq_div_up(w, FX_TX) * atq_div_up(h, FX_TY)));
        quads = std::max(quads, w * h / 4);
    }
    tab.n = (uint32_t)n;
    const uint32_t lg = std::min<uint32_t>(xcd_grid(tiles), 8192u);
    const uint32_t bg = std::min<uint32_t>(atq_div_up(tiles, 2), 4096u);
    const uint32_t fg = std::min<uint32_t>(atq_div_up(quads, 256), 16384u);
    k_fx_local<<<lg ? lg : 8, 256, 0, s>>>(tab);
    k_fx_border<<<bg ? bg : 1, 256, 0, s>>>(tab);
    k_fx_flatten<<<fg ? fg : 1, 256, 0, s>>>(tab);
    k_fx_final<<<fg ? fg : 1, 256, 0, s>>>(tab);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq

// Dual-cell enumeration and mesh assembly on the device (genIsoOurs state 1,
// src/surface_reconstructor.cpp:290-324; traversal include/traverse.h:12-312; VisitorExtract,
// src/visitorextract.cpp:41-257; Mesh::insert_vert / insert_tri, src/global.cpp:13-53).
//
// The recursive traversal of the reference visits every octree vertex P interior to the root that
// is a corner of some leaf, and emits the 8 leaves around P. Here every leaf emits its 8 corners as
// integer keys at the finest level; a radix sort groups equal points; the leaves of octants not
// covered by a corner entry (P on a face or edge of a larger leaf) are found by point location.
// Dual cells come out in point-key order instead of traversal order; vertices are keyed by the
// unordered pair of leaves of their edge and triangles are deduplicated by sorted vertex triple.
#include "pisurf_internal.h"

#define PISURF_DMC_QUAL __constant__ static const
#define PISURF_DMC_NAME(x) d_##x
#include "dmc_tables.h"

namespace pisurf {

__device__ __forceinline__ bool is_leaf_type(uint8_t t) { return t == 2 || t == 0; }

// point keys: Morton interleave of the 21-bit finest-level coordinates (x in bit 0), so dual cells,
// and the vertices numbered from them, come out in space-filling order
__device__ __forceinline__ uint64_t morton_spread(uint64_t v) {
    v &= 0x1fffff;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
__device__ __forceinline__ uint32_t morton_compact(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffffull;
    return (uint32_t)v;
}

// (flag[nodes] = 0 closes the array for its scan)
__global__ void k_leaf_flags(LeafLevels lv, int* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == lv.off[lv.nlev]) flag[i] = 0;
    if (i >= lv.off[lv.nlev]) return;
    int L = 0;
    while (L + 1 < lv.nlev && i >= lv.off[L + 1]) L++;
    flag[i] = is_leaf_type(lv.type[L][i - lv.off[L]]) ? 1 : 0;
}

// The distinct corner points of the leaves, in three steps. Each workgroup takes a tile of 256
// consecutive nodes of one level (a level's nodes are in Morton order, so a tile covers a compact
// piece of space and its leaves share most corners) and merges the tile's corner keys in an LDS
// hash table: one record per distinct point of the tile, holding its Morton key and, for each of
// the point's 8 octants, the tile's leaf in that octant (-1 if none). Tile b writes its records to
// its own region, slots b * kCtRecs .. b * kCtRecs + count - 1 (no shared counter: tiles of a
// frame number in the tens of thousands, and one device-wide atomic per tile serialises them),
// and its count to tcount[b]. A scan of the counts then places each tile's keys in one contiguous
// array, packed with the record's slot, for the global sort. A point shared by several tiles has a
// record in each; the sort groups them and the point pass merges them. A leaf with P as corner c
// lies in octant c ^ 7 of P, and no two leaves lie in the same octant of P, so the merge is a plain
// union, and every result taken from the records is independent of the tiles.
//
// Keys: Morton interleave of the finest-level corner coordinates (3 dp bits); corners on the root
// boundary are not points of the traversal and are skipped.
constexpr int kCtT = 256;               // nodes per tile
constexpr int kCtRecs = 8 * kCtT;       // corner keys, hence records, per tile at most
constexpr int kCtSlots = 2 * kCtRecs;   // hash slots: load factor at most 1/2
constexpr int kCtSlotBits = 12;
static_assert((1 << kCtSlotBits) == kCtSlots, "slot bits");
constexpr uint64_t kCtEmpty = ~0ull;     // above every key (3 dp <= 63 bits)

__global__ void __launch_bounds__(kCtT) k_corner_tiles(LeafLevels lv, const int* __restrict__ lidx,
                                                       double4* __restrict__ leaf_node, uint8_t* __restrict__ leaf_pos,
                                                       uint64_t* __restrict__ fkey, int* __restrict__ rleaf,
                                                       int* __restrict__ tcount) {
    __shared__ uint64_t tab[kCtSlots];
    __shared__ short uid[kCtSlots];  // record of a slot within the tile (< kCtRecs)
    __shared__ int wsum[kCtT / kWave];
    const int t = threadIdx.x;
    for (int k = t; k < kCtSlots; k += kCtT) tab[k] = kCtEmpty;
    __syncthreads();
    // this tile's level (uniform over the workgroup) and node
    const int64_t b = blockIdx.x;
    int lev = 0;
    while (lev + 1 < lv.nlev && b >= lv.tile0[lev + 1]) lev++;
    const int dp = lv.dp;
    const int64_t i = (b - lv.tile0[lev]) * kCtT + t;
    int L = -1;
    int slot[8];
#pragma unroll
    for (int c = 0; c < 8; c++) slot[c] = -1;
    if (i < lv.off[lev + 1] - lv.off[lev] && is_leaf_type(lv.type[lev][i])) {
        L = lidx[lv.off[lev] + i];
        const double4 nd = lv.node[lev][i];
        leaf_node[L] = nd;
        leaf_pos[L] = nd.w > 0.0 ? 1 : 0;  // sign of the leaf value, read by the point pass
        const int4 p = lv.ip[lev][i];
        const int sh = dp - p.w;
        const int64_t hi = ((int64_t)1 << dp) - 1;
        int64_t cx[2], cy[2], cz[2];
        uint64_t mx[2], my[2], mz[2];
#pragma unroll
        for (int o = 0; o < 2; o++) {
            cx[o] = (int64_t)(p.x + o) << sh;
            cy[o] = (int64_t)(p.y + o) << sh;
            cz[o] = (int64_t)(p.z + o) << sh;
            mx[o] = morton_spread(cx[o]);
            my[o] = morton_spread(cy[o]) << 1;
            mz[o] = morton_spread(cz[o]) << 2;
        }
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
            const int64_t X = cx[ox], Y = cy[oy], Z = cz[oz];
            if (!(X >= 1 && X <= hi && Y >= 1 && Y <= hi && Z >= 1 && Z <= hi)) continue;
            const uint64_t k = mx[ox] | my[oy] | mz[oz];
            // at most kCtRecs keys in kCtSlots slots: a free slot is always reached
            unsigned h = (unsigned)((k * 0x9E3779B97F4A7C15ull) >> (64 - kCtSlotBits));
            for (;;) {
                const uint64_t cur = tab[h];
                if (cur == k) break;
                if (cur == kCtEmpty) {
                    const uint64_t prev = atomicCAS((unsigned long long*)&tab[h], (unsigned long long)kCtEmpty,
                                                    (unsigned long long)k);
                    if (prev == kCtEmpty || prev == k) break;
                }
                h = (h + 1) & (kCtSlots - 1);
            }
            slot[c] = (int)h;
        }
    }
    __syncthreads();
    // number the occupied slots in slot order
    constexpr int per = kCtSlots / kCtT;
    uint64_t mine[per];
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < per; k++) {
        mine[k] = tab[t * per + k];
        cnt += mine[k] != kCtEmpty ? 1 : 0;
    }
    const int lane = t & (kWave - 1), w = t / kWave;
    const int incl = wave_incl_scan(cnt, lane);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int u = incl - cnt, total = 0;
#pragma unroll
    for (int k = 0; k < kCtT / kWave; k++) {
        if (k < w) u += wsum[k];
        total += wsum[k];
    }
    if (t == 0) tcount[b] = total;
    if (b == 0 && t == 0) tcount[gridDim.x] = 0;  // closes the counts for their scan
    // each record: its key, and no leaf in any octant yet
    const int4 none = make_int4(-1, -1, -1, -1);
#pragma unroll
    for (int k = 0; k < per; k++) {
        if (mine[k] == kCtEmpty) continue;
        const int64_t r = b * kCtRecs + u;
        uid[t * per + k] = (short)u;
        fkey[r] = mine[k];
        ((int4*)rleaf)[2 * r] = none;
        ((int4*)rleaf)[2 * r + 1] = none;
        u++;
    }
    // the record stores above are complete before any thread of the tile writes a leaf into them
    __threadfence_block();
    __syncthreads();
    if (L >= 0) {
#pragma unroll
        for (int c = 0; c < 8; c++)
            if (slot[c] >= 0) rleaf[(b * kCtRecs + uid[slot[c]]) * 8 + (c ^ 7)] = L;
    }
}

// the tiles' keys in one array (tile b from tbase[b] on), each packed with its record slot when
// vbits > 0 (else the slot in rval)
__global__ void __launch_bounds__(kCtT) k_records_compact(const int* __restrict__ tcount, const int* __restrict__ tbase,
                                                          const uint64_t* __restrict__ fkey, int vbits,
                                                          uint64_t* __restrict__ rkey, int* __restrict__ rval) {
    const int64_t b = blockIdx.x;
    const int n = tcount[b];
    const int64_t d0 = tbase[b];
    for (int u = threadIdx.x; u < n; u += kCtT) {
        const int64_t r = b * kCtRecs + u;
        if (vbits) {
            rkey[d0 + u] = (fkey[r] << vbits) | (uint64_t)r;
        } else {
            rkey[d0 + u] = fkey[r];
            rval[d0 + u] = (int)r;
        }
    }
}

// leaf containing the finest cell q, descending from node i of level L (the root: 0, 0)
__device__ __forceinline__ int locate_from(const TreeView& tv, const int* lidx, int L, int64_t i, int qx, int qy,
                                           int qz) {
    for (; L < tv.nlev; L++) {
        const uint8_t t = tv.type[L][i];
        if (is_leaf_type(t)) return lidx[tv.off[L] + i];
        const int sh = tv.dp - L - 1;
        const int b = ((qx >> sh) & 1) | (((qy >> sh) & 1) << 1) | (((qz >> sh) & 1) << 2);
        i = (int64_t)tv.child[L][i] + b;
    }
    return -1;
}

// the tree index: cell c of level glev (x fastest) descended from the root until level glev or a
// leaf above it
__global__ void k_tree_index(TreeView tv, int* __restrict__ gidx) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = tv.glev;
    if (c >= ((int64_t)1 << (3 * g))) return;
    const int x = (int)(c & ((1 << g) - 1)), y = (int)((c >> g) & ((1 << g) - 1)), z = (int)(c >> (2 * g));
    int L = 0;
    int64_t i = 0;
    for (; L < g; L++) {
        if (is_leaf_type(tv.type[L][i])) break;
        const int sh = g - L - 1;
        const int b = ((x >> sh) & 1) | (((y >> sh) & 1) << 1) | (((z >> sh) & 1) << 2);
        i = (int64_t)tv.child[L][i] + b;
    }
    gidx[c] = (L << 27) | (int)i;
}
// the index at level glev from the one at glev - 1 (coarse): a cell keeps its parent cell's entry
// when that is a leaf, and otherwise takes the child of the parent cell's node that holds it
__global__ void k_tree_index_refine(TreeView tv, const int* __restrict__ coarse, int* __restrict__ gidx) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = tv.glev, gc = g - 1;
    if (c >= ((int64_t)1 << (3 * g))) return;
    const int x = (int)(c & ((1 << g) - 1)), y = (int)((c >> g) & ((1 << g) - 1)), z = (int)(c >> (2 * g));
    const int e = coarse[(int64_t)(x >> 1) | ((int64_t)(y >> 1) << gc) | ((int64_t)(z >> 1) << (2 * gc))];
    const int L = e >> 27, i = e & ((1 << 27) - 1);
    if (L < gc || is_leaf_type(tv.type[gc][i])) {
        gidx[c] = e;
        return;
    }
    const int b = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
    gidx[c] = (g << 27) | (tv.child[gc][i] + b);
}
int tree_index_level(int dp) { return min(dp, 8); }
void launch_tree_index(TreeView tv, int* gidx, int* coarse, hipStream_t s) {
    const int64_t n = (int64_t)1 << (3 * tv.glev);
    if (tv.glev >= 1 && coarse) {
        TreeView tc = tv;
        tc.glev = tv.glev - 1;
        const int64_t nc = (int64_t)1 << (3 * tc.glev);
        k_tree_index<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(tc, coarse);
        k_tree_index_refine<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(tv, coarse, gidx);
        return;
    }
    k_tree_index<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(tv, gidx);
}
// where the descent to finest cell (qx, qy, qz) resumes: its index cell's node
__device__ __forceinline__ void index_start(const TreeView& tv, int qx, int qy, int qz, int& L, int64_t& i) {
    const int gs = tv.dp - tv.glev, g = tv.glev;
    const int64_t c = (int64_t)(qx >> gs) | ((int64_t)(qy >> gs) << g) | ((int64_t)(qz >> gs) << (2 * g));
    const int e = tv.gidx[c];
    L = e >> 27;
    i = e & ((1 << 27) - 1);
}

// the 8 leaves around the point whose records start at sorted position j (at most one record per
// tile, and a point lies in at most 8 tiles' leaves, so its run is at most 8 long)
__device__ __forceinline__ void gather_point(const TreeView& tv, const int* lidx, const SortedKV& kv, int64_t nent,
                                             const int* rleaf, int64_t j, int n[8]) {
    const uint64_t key = kv.key(j);
#pragma unroll
    for (int o = 0; o < 8; o++) n[o] = -1;
    for (int64_t k = j; k < nent && k < j + 8 && kv.key(k) == key; k++) {
        const int64_t r = kv.val(k);
        const int4 a = ((const int4*)rleaf)[2 * r], b = ((const int4*)rleaf)[2 * r + 1];
        n[0] = max(n[0], a.x);
        n[1] = max(n[1], a.y);
        n[2] = max(n[2], a.z);
        n[3] = max(n[3], a.w);
        n[4] = max(n[4], b.x);
        n[5] = max(n[5], b.y);
        n[6] = max(n[6], b.z);
        n[7] = max(n[7], b.w);
    }
    const int px = (int)morton_compact(key), py = (int)morton_compact(key >> 1), pz = (int)morton_compact(key >> 2);
    bool missing = false;
    for (int o = 0; o < 8; o++) missing = missing || n[o] < 0;
    if (missing) {
        // The descents start from the tree index, not the root. When a coordinate of P is a
        // multiple of the index cell size the 8 octant cells lie in different index cells, and
        // each missing octant descends from its own.
        const int gm = (1 << (tv.dp - tv.glev)) - 1;
        if ((px & gm) == 0 || (py & gm) == 0 || (pz & gm) == 0) {
            for (int o = 0; o < 8; o++)
                if (n[o] < 0) {
                    const int qx = px - 1 + (o & 1), qy = py - 1 + ((o >> 1) & 1), qz = pz - 1 + ((o >> 2) & 1);
                    int L0;
                    int64_t i0;
                    index_start(tv, qx, qy, qz, L0, i0);
                    n[o] = locate_from(tv, lidx, L0, i0, qx, qy, qz);
                }
            missing = false;
        }
    }
    if (missing) {
        // The finest cells of the 8 octants, P - 1 or P per axis, take the same child of a node
        // whose children have size 2^sh as long as no coordinate of P is a multiple of 2^sh, so
        // their descents share the path down to that node and are walked once.
        int L;
        int64_t i;
        index_start(tv, px, py, pz, L, i);
        int shared = -1;
        for (; L < tv.nlev; L++) {
            const uint8_t t = tv.type[L][i];
            if (is_leaf_type(t)) {
                shared = lidx[tv.off[L] + i];
                break;
            }
            const int sh = tv.dp - L - 1;
            if (sh <= 0) break;
            const int m = (1 << sh) - 1;
            if ((px & m) == 0 || (py & m) == 0 || (pz & m) == 0) break;
            const int b = ((px >> sh) & 1) | (((py >> sh) & 1) << 1) | (((pz >> sh) & 1) << 2);
            i = (int64_t)tv.child[L][i] + b;
        }
        for (int o = 0; o < 8; o++)
            if (n[o] < 0)
                n[o] = shared >= 0 ? shared
                                   : locate_from(tv, lidx, L, i, px - 1 + (o & 1), py - 1 + ((o >> 1) & 1),
                                                 pz - 1 + ((o >> 2) & 1));
    }
    int any = -1;
    for (int o = 0; o < 8; o++)
        if (n[o] >= 0) any = n[o];
    for (int o = 0; o < 8; o++)
        if (n[o] < 0) n[o] = any;  // unreachable for a consistent tree
}

// List of the run heads of n sorted entries (entry j with j = 0 or key(j) != key(j - 1)), in
// order: per tile of 2048 entries the heads are counted, the counts scanned, and each tile writes
// its heads at its offset; the flags are recomputed from the keys, never stored.
constexpr int kHeadT = 256, kHeadI = 8, kHeadTile = kHeadT * kHeadI;
__device__ __forceinline__ int head_flags(const SortedKV& kv, int64_t n, int64_t base, bool f[kHeadI]) {
    int c = 0;
    uint64_t prev = (base > 0 && base - 1 < n) ? kv.key(base - 1) : 0;
#pragma unroll
    for (int q = 0; q < kHeadI; q++) {
        const int64_t j = base + q;
        f[q] = false;
        if (j < n) {
            const uint64_t k = kv.key(j);
            f[q] = (j == 0 || k != prev);
            prev = k;
        }
        c += f[q];
    }
    return c;
}
__global__ void __launch_bounds__(kHeadT) k_head_count(SortedKV kv, int64_t n, int* __restrict__ tile_count) {
    __shared__ int wsum[kHeadT / kWave];
    bool f[kHeadI];
    int c = head_flags(kv, n, (int64_t)blockIdx.x * kHeadTile + (int64_t)threadIdx.x * kHeadI, f);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, kWave);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_count[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (blockIdx.x == 0 && threadIdx.x == 0) tile_count[gridDim.x] = 0;  // closes the counts for their scan
}
__global__ void __launch_bounds__(kHeadT) k_head_write(SortedKV kv, int64_t n, const int* __restrict__ tile_off,
                                                       int* __restrict__ list) {
    __shared__ int wsum[kHeadT / kWave];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t base = (int64_t)blockIdx.x * kHeadTile + (int64_t)t * kHeadI;
    bool f[kHeadI];
    const int c = head_flags(kv, n, base, f);
    const int incl = wave_incl_scan(c, lane);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int pos = tile_off[blockIdx.x] + incl - c;
    for (int k = 0; k < w; k++) pos += wsum[k];
#pragma unroll
    for (int q = 0; q < kHeadI; q++)
        if (f[q]) list[pos++] = (int)(base + q);
}


__device__ __forceinline__ int sign_index(const uint8_t* leaf_pos, const int n[8]) {
    int s = 0;
#pragma unroll
    for (int o = 0; o < 8; o++)
        if (leaf_pos[n[o]]) s |= 1 << o;
    return s;
}

// one thread per distinct point (heads[t] = its first sorted entry): flag points whose dual cell
// is non-trivial and stage the cell (8 leaves in case order, sign configuration) at slot t;
// k_points_compact then packs the staged cells
__global__ void k_points(TreeView tv, const int* __restrict__ lidx, const uint8_t* __restrict__ leaf_pos,
                         SortedKV kv, int64_t nent, const int* __restrict__ rleaf, const int* __restrict__ heads,
                         int64_t npts, int* __restrict__ dflag, int* __restrict__ stage) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == npts) dflag[t] = 0;  // closes the flags for their scan
    if (t >= npts) return;
    int n[8];
    gather_point(tv, lidx, kv, nent, rleaf, heads[t], n);
    const int s = sign_index(leaf_pos, n);
    const uint8_t code = d_kDmcCase[s];
    const bool keep = code != 0 && code != 255;
    dflag[t] = keep ? 1 : 0;
    if (!keep) return;
    int* d = stage + 9 * t;
    const uint32_t perm = d_kDmcPerm[s];
#pragma unroll
    for (int k = 0; k < 8; k++) d[k] = n[(perm >> (4 * k)) & 15];
    d[8] = s;
}

__global__ void k_points_compact(const int* __restrict__ dflag, const int* __restrict__ dpos,
                                 const int* __restrict__ stage, int64_t nent, int* __restrict__ duals) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nent || !dflag[j]) return;
    const int* src = stage + 9 * j;
    int* d = duals + 9 * (int64_t)dpos[j];
#pragma unroll
    for (int k = 0; k < 9; k++) d[k] = src[k];
}

__device__ __forceinline__ int canon_slot(uint8_t code) {
    for (int k = 0; k < PISURF_DMC_NCASES; k++)
        if (d_kDmcCanon[k] == code) return k;
    return -1;
}

__global__ void k_tri_count(const int* __restrict__ duals, int64_t nd, int* __restrict__ ntri) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d == nd) ntri[d] = 0;  // closes the counts for their scan
    if (d >= nd) return;
    const int slot = canon_slot(d_kDmcCase[duals[9 * d + 8]]);
    ntri[d] = slot < 0 ? 0 : d_kDmcNTri[slot];
}

// triangle corners in output winding (generate_mesh: mirrored cases keep the table order, the
// others are reversed); vertex key = unordered leaf pair of the corner's dual edge
__global__ void k_tri_corners(const int* __restrict__ duals, int64_t nd, const int* __restrict__ toff,
                              uint64_t* __restrict__ vkey, int* __restrict__ vval, int* __restrict__ vend) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nd) return;
    const int* n8 = duals + 9 * d;
    const int s = n8[8];
    const int slot = canon_slot(d_kDmcCase[s]);
    if (slot < 0) return;
    const bool mirror = d_kDmcMirror[s] != 0;
    for (int t = 0; t < d_kDmcNTri[slot]; t++)
        for (int kk = 0; kk < 3; kk++) {
            const int k = mirror ? kk : 2 - kk;
            const int e = d_kDmcTri[slot][t][k];
            const int la = n8[d_kDmcEdge[e][0]], lb = n8[d_kDmcEdge[e][1]];
            const int64_t idx = 3 * ((int64_t)toff[d] + t) + kk;
            const uint64_t a = (uint64_t)min(la, lb), b = (uint64_t)max(la, lb);
            vkey[idx] = (a << 32) | b;
            vval[idx] = (int)idx;
            vend[2 * idx] = la;
            vend[2 * idx + 1] = lb;
        }
}


// ---- hash-table dedupe (first occurrence wins, as Mesh::insert_vert / insert_tri do) ----
// Open addressing with linear probing over a power-of-two table of more slots than keys, so every
// probe sequence ends. A slot holds only the smallest index that inserted its key; an insert
// compares keys through key[table[h]], and takes one atomic (the claim of an empty slot, or the
// minimum on a found one) instead of a key exchange followed by a minimum.
constexpr unsigned kSlotEmpty = 0xffffffffu;
__device__ __forceinline__ uint64_t mix64(uint64_t k) {  // SplitMix64 finaliser
    k ^= k >> 30;
    k *= 0xbf58476d1ce4e5b9ull;
    k ^= k >> 27;
    k *= 0x94d049bb133111ebull;
    k ^= k >> 31;
    return k;
}
// slot[i] = table slot of key[i]; table[slot] = smallest i with that key
// An entry whose key equals that of one of the kDedupeWindow entries before it in the same wave
// takes that entry's slot without probing (nearby triangle corners repeat vertices); the table keeps
// the smallest index of each key, which such an entry never holds.
constexpr int kDedupeWindow = 31;
__global__ void k_hash_insert(const uint64_t* __restrict__ key, int64_t n, unsigned* table, uint64_t mask,
                              int* __restrict__ slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const bool act = i < n;
    const uint64_t k = act ? key[i] : 0;
    const unsigned klo = (unsigned)k, khi = (unsigned)(k >> 32);
    // root: the earliest lane among this one and the kDedupeWindow before it with the same key
    int root = lane;
#pragma unroll
    for (int o = kDedupeWindow; o >= 1; o--) {
        const unsigned lo = __shfl_up(klo, o, kWave), hi = __shfl_up(khi, o, kWave);
        if (root == lane && lane >= o && lo == klo && hi == khi) root = lane - o;
    }
    // follow roots to a lane that is its own root (pointer jumping; roots only point back)
#pragma unroll
    for (int it = 0; it < 6; it++) root = __shfl(root, root, kWave);
    int h_found = -1;
    if (act && root == lane) {
        uint64_t h = mix64(k) & mask;
        for (uint64_t probe = 0; probe <= mask; probe++) {
            // a plain read may be stale: an empty slot is claimed only through the exchange, and a
            // slot's index only decreases among indices of one key
            unsigned cur = table[h];
            if (cur == kSlotEmpty) {
                cur = atomicCAS(&table[h], kSlotEmpty, (unsigned)i);
                if (cur == kSlotEmpty) {
                    h_found = (int)h;
                    break;
                }
            }
            if (key[cur] == k) {
                if (cur > (unsigned)i) atomicMin(&table[h], (unsigned)i);
                h_found = (int)h;
                break;
            }
            h = (h + 1) & mask;
        }
    }
    const int hs = __shfl(h_found, root, kWave);
    if (act) slot[i] = hs;
}
__global__ void k_hash_first(const int* __restrict__ slot, const unsigned* __restrict__ table, int64_t n,
                             int* __restrict__ first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) first[i] = (table[slot[i]] == (unsigned)i) ? 1 : 0;
    if (i == n) first[i] = 0;  // closes the flags for their scan
}
// vertex id of every corner (rank of its key's first corner) and the bisection edge of each vertex,
// oriented as its first corner's dual cell has it
__global__ void k_vertex_ids_hashed(const int* __restrict__ slot, const unsigned* __restrict__ table,
                                    const int* __restrict__ rank, int64_t n, const int* __restrict__ vend,
                                    const double4* __restrict__ leaf_node, int* __restrict__ corner_vid,
                                    double* __restrict__ ends) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const unsigned f = table[slot[c]];
    const int vid = rank[f];
    corner_vid[c] = vid;
    if (f == (unsigned)c) {
        const double4 a = leaf_node[vend[2 * c]], b = leaf_node[vend[2 * c + 1]];
        double* e = ends + 8 * (int64_t)vid;
        e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
        e[4] = b.x; e[5] = b.y; e[6] = b.z; e[7] = b.w;
    }
}
// triangle key: sorted vertex triple, 21 bits each (vertex count below 2^21)
__global__ void k_tri_key63(const int* __restrict__ cv, int64_t nt, uint64_t* __restrict__ k) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int a = cv[3 * t], b = cv[3 * t + 1], c = cv[3 * t + 2];
    const int lo = min(a, min(b, c)), hi = max(a, max(b, c)), mid = a + b + c - lo - hi;
    k[t] = ((uint64_t)lo << 42) | ((uint64_t)mid << 21) | (uint64_t)hi;
}

// triangle dedupe keys: sorted vertex triple (first pass by the largest id, second by the pair)
__global__ void k_tri_key_hi(const int* __restrict__ cv, int64_t nt, uint32_t* __restrict__ k, int* __restrict__ v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int a = cv[3 * t], b = cv[3 * t + 1], c = cv[3 * t + 2];
    k[t] = (uint32_t)max(a, max(b, c));
    v[t] = (int)t;
}
__global__ void k_tri_key_lo(const int* __restrict__ cv, const int* __restrict__ perm, int64_t nt,
                             uint64_t* __restrict__ k, int* __restrict__ v) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    const int t = perm[j];
    const int a = cv[3 * (int64_t)t], b = cv[3 * (int64_t)t + 1], c = cv[3 * (int64_t)t + 2];
    const int lo = min(a, min(b, c)), hi = max(a, max(b, c)), mid = a + b + c - lo - hi;
    k[j] = ((uint64_t)(uint32_t)lo << 32) | (uint32_t)mid;
    v[j] = t;
}
__device__ __forceinline__ void sorted3(const int* cv, int64_t t, int& lo, int& mid, int& hi) {
    const int a = cv[3 * t], b = cv[3 * t + 1], c = cv[3 * t + 2];
    lo = min(a, min(b, c));
    hi = max(a, max(b, c));
    mid = a + b + c - lo - hi;
}
__global__ void k_tri_keep(const int* __restrict__ cv, const int* __restrict__ order, int64_t nt,
                           int* __restrict__ keep) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j == nt) keep[j] = 0;  // closes the flags for their scan
    if (j >= nt) return;
    const int t = order[j];
    int k = 1;
    if (j > 0) {
        int a0, a1, a2, b0, b1, b2;
        sorted3(cv, t, a0, a1, a2);
        sorted3(cv, order[j - 1], b0, b1, b2);
        k = (a0 == b0 && a1 == b1 && a2 == b2) ? 0 : 1;  // equal triples: the earlier triangle stays
    }
    keep[t] = k;
}
__global__ void k_tri_out(const int* __restrict__ cv, const int* __restrict__ keep, const int* __restrict__ kpos,
                          int64_t nt, int* __restrict__ tris) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt || !keep[t]) return;
    const int64_t o = kpos[t];
    tris[3 * o] = cv[3 * t] + 1;
    tris[3 * o + 1] = cv[3 * t + 1] + 1;
    tris[3 * o + 2] = cv[3 * t + 2] + 1;
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_leaf_flags(const LeafLevels& lv, int* flag, hipStream_t s) {
    k_leaf_flags<<<nblk(lv.off[lv.nlev] + 1), 256, 0, s>>>(lv, flag);
}
int64_t corner_tiles(int64_t nodes) { return (nodes + kCtT - 1) / kCtT; }
int64_t corner_tile_records() { return kCtRecs; }
void launch_corner_tiles(const LeafLevels& lv, const int* lidx, double4* leaf_node, uint8_t* leaf_pos, uint64_t* fkey,
                         int* rleaf, int* tcount, hipStream_t s) {
    if (lv.tile0[lv.nlev] > 0)
        k_corner_tiles<<<(unsigned)lv.tile0[lv.nlev], kCtT, 0, s>>>(lv, lidx, leaf_node, leaf_pos, fkey, rleaf,
                                                                    tcount);
    else
        launch_fill_i32(tcount, 1, 0, s);
}
void launch_records_compact(const int* tcount, const int* tbase, int64_t ntiles, const uint64_t* fkey, int vbits,
                            uint64_t* rkey, int* rval, hipStream_t s) {
    if (ntiles > 0) k_records_compact<<<(unsigned)ntiles, kCtT, 0, s>>>(tcount, tbase, fkey, vbits, rkey, rval);
}
void launch_hash_dedupe(const uint64_t* key, int64_t n, unsigned* table, int64_t tsize, int* slot, int* first,
                        hipStream_t s) {
    if (n <= 0) {
        launch_fill_i32(first, 1, 0, s);
        return;
    }
    (void)hipMemsetAsync(table, 0xff, tsize * sizeof(unsigned), s);  // every slot empty
    k_hash_insert<<<nblk(n), 256, 0, s>>>(key, n, table, (uint64_t)(tsize - 1), slot);
    k_hash_first<<<nblk(n + 1), 256, 0, s>>>(slot, table, n, first);
}
void launch_vertex_ids_hashed(const int* slot, const unsigned* table, const int* rank, int64_t n, const int* vend,
                              const double4* leaf_node, int* corner_vid, double* ends, hipStream_t s) {
    if (n > 0) k_vertex_ids_hashed<<<nblk(n), 256, 0, s>>>(slot, table, rank, n, vend, leaf_node, corner_vid, ends);
}
void launch_tri_key63(const int* cv, int64_t nt, uint64_t* k, hipStream_t s) {
    if (nt > 0) k_tri_key63<<<nblk(nt), 256, 0, s>>>(cv, nt, k);
}
static_assert(kHeadTile == 2048, "head_list_count assumes tiles of 2048 entries");
int64_t head_list_tmp_ints(int64_t n) { return 2 * ((n + kHeadTile - 1) / kHeadTile + 1); }
void launch_head_list(SortedKV kv, int64_t n, int* list, int* tiles, int* scan_tmp, hipStream_t s, ScanPub pub) {
    const int64_t nb = (n + kHeadTile - 1) / kHeadTile;
    int* tile_off = tiles + (nb + 1);
    if (n <= 0) {
        (void)hipMemsetAsync(tile_off, 0, sizeof(int), s);
        return;
    }
    k_head_count<<<(unsigned)nb, kHeadT, 0, s>>>(kv, n, tiles);
    exclusive_scan_i32(tiles, tile_off, nb + 1, scan_tmp, s, pub);
    k_head_write<<<(unsigned)nb, kHeadT, 0, s>>>(kv, n, tile_off, list);
}
void launch_points(const TreeView& tv, const int* lidx, const uint8_t* leaf_pos, SortedKV kv, int64_t nent,
                   const int* rleaf, const int* heads, int64_t npts, int* dflag, int* stage, hipStream_t s) {
    k_points<<<nblk(npts + 1), 256, 0, s>>>(tv, lidx, leaf_pos, kv, nent, rleaf, heads, npts, dflag, stage);
}
void launch_points_compact(const int* dflag, const int* dpos, const int* stage, int64_t nent, int* duals,
                           hipStream_t s) {
    if (nent > 0) k_points_compact<<<nblk(nent), 256, 0, s>>>(dflag, dpos, stage, nent, duals);
}
void launch_tri_count(const int* duals, int64_t nd, int* ntri, hipStream_t s) {
    k_tri_count<<<nblk(nd + 1), 256, 0, s>>>(duals, nd, ntri);
}
void launch_tri_corners(const int* duals, int64_t nd, const int* toff, uint64_t* vkey, int* vval, int* vend,
                        hipStream_t s) {
    if (nd > 0) k_tri_corners<<<nblk(nd), 256, 0, s>>>(duals, nd, toff, vkey, vval, vend);
}
void launch_tri_key_hi(const int* cv, int64_t nt, uint32_t* k, int* v, hipStream_t s) {
    if (nt > 0) k_tri_key_hi<<<nblk(nt), 256, 0, s>>>(cv, nt, k, v);
}
void launch_tri_key_lo(const int* cv, const int* perm, int64_t nt, uint64_t* k, int* v, hipStream_t s) {
    if (nt > 0) k_tri_key_lo<<<nblk(nt), 256, 0, s>>>(cv, perm, nt, k, v);
}
void launch_tri_keep(const int* cv, const int* order, int64_t nt, int* keep, hipStream_t s) {
    k_tri_keep<<<nblk(nt + 1), 256, 0, s>>>(cv, order, nt, keep);
}
void launch_tri_out(const int* cv, const int* keep, const int* kpos, int64_t nt, int* tris, hipStream_t s) {
    if (nt > 0) k_tri_out<<<nblk(nt), 256, 0, s>>>(cv, keep, kpos, nt, tris);
}

}  // namespace pisurf

// Internal declarations shared by the pisurf host code and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

namespace pisurf {

// ---- scan / sort workspace ----
// where a scan publishes its last element for the host (scan_read_begin): the slot's device
// address, the value's offset in it and the sequence number; dst NULL: nowhere
struct ScanPub {
    char* dst = nullptr;
    int off = 0;
    uint32_t seq = 0;
};
// device exclusive scan of n int32 values; tmp must hold scan_tmp_ints(n) ints
int64_t scan_tmp_ints(int64_t n);
void exclusive_scan_i32(const int* in, int* out, int64_t n, int* tmp, hipStream_t s, ScanPub pub = ScanPub());
// stable LSD radix sort of (key, value) by the low `bits` bits of key; results in k0/v0.
// k1/v1 are ping-pong buffers; tmp must hold radix_tmp_ints(n) ints.
int64_t radix_tmp_ints(int64_t n);
void radix_sort_pairs(uint32_t* k0, int* v0, uint32_t* k1, int* v1, int64_t n, int bits, int* tmp, hipStream_t s);
void radix_sort_pairs64(uint64_t* k0, int* v0, uint64_t* k1, int* v1, int64_t n, int bits, int* tmp, hipStream_t s);
// Sorted entries: packed words (vbits > 0: the key above vbits bits of value, v unused) or
// separate key / value arrays (vbits = 0). Packed, a sort pass makes one scattered store per
// entry instead of two.
struct SortedKV {
    const uint64_t* k;
    const int* v;
    int vbits;
    __device__ __forceinline__ uint64_t key(int64_t i) const { return vbits ? (k[i] >> vbits) : k[i]; }
    __device__ __forceinline__ int val(int64_t i) const {
        return vbits ? (int)(k[i] & ((1ull << vbits) - 1)) : v[i];
    }
};
// value bits to pack beside a key of `bits` bits for values 0 .. n - 1, or 0 when they do not fit
// one word
int pack_vbits(int bits, int64_t n);
// stable sort of n entries by the low `bits` bits of their key. vbits > 0: k0 holds packed words
// with keys below 2^bits, k1 is the ping-pong buffer, and the result stays where the last pass put
// it (k0 or k1); vbits = 0: key / value arrays, result in k0 / v0.
SortedKV radix_sort_entries(uint64_t* k0, int* v0, uint64_t* k1, int* v1, int64_t n, int bits, int vbits, int* tmp,
                            hipStream_t s);

// ---- read-backs ----
// wait for the stream's work so far by polling an event (no driver sleep)
hipError_t stream_wait_on(hipStream_t s, hipEvent_t ev);
// A read-back slot: kReadbackBytes of host memory mapped into the device's address space, then a
// sequence word. A one-workgroup kernel on the stream copies the bytes into the slot and then
// writes the next sequence number; the host polls that word. Slots come from a process-wide pool
// (a context holds one; they are never freed, only returned to the pool).
constexpr size_t kReadbackBytes = 16384;
struct HostRead {
    char* host;     // host address of the slot
    char* dev;      // its device address
    uint32_t next;  // last sequence number written to the slot
};
HostRead* host_read_acquire();  // nullptr when mapped memory cannot be allocated
void host_read_release(HostRead* rd);
// copy bytes (<= kReadbackBytes) of device memory to host after the stream's earlier work
hipError_t read_device(void* host, const void* dptr, size_t bytes, hipStream_t s, HostRead* rd);
// wait for the stream's earlier work
hipError_t stream_fence(hipStream_t s, HostRead* rd);
// A scan total that the scan's final kernel publishes itself, with no copy launch after it:
// scan_read_begin takes the slot's next sequence number (dst NULL when rd is NULL), the scan
// writes its last element at byte off of the slot and then that number, and scan_read_end waits
// for it (or, without a slot, reads dlast). Reads in between may use the slot (the host waits
// for a sequence number or a later one) as long as they leave the value's bytes alone: copies
// fill the slot from byte 0, so a value kept past other reads takes kScanPubFar.
constexpr int kScanPubFar = (int)kReadbackBytes - 64;
ScanPub scan_read_begin(HostRead* rd, int off = 0);
hipError_t scan_read_end(int* host, const int* dlast, ScanPub p, hipStream_t s, HostRead* rd);

// ---- grid ----
void launch_bbox(const float* xyz, int64_t n, float* partial /*>= 6*1024*/, float* out6, int* nan_count,
                 hipStream_t s);
void launch_keys(const float* xyz, int64_t n, GridDev g, uint32_t* keys, int* vals, int* n_invalid, hipStream_t s);
// off[0 .. ncell] from the sorted keys (cell + 1, 0 outside the grid)
void launch_cell_offsets(const uint32_t* skeys, int64_t n, const int* n_invalid, int64_t ncell, int* off,
                         hipStream_t s);
void launch_gather_pos(const float* xyz, const int* sidx, int64_t n, float4* spos, hipStream_t s);
void launch_neighbor_count(GridDev g, const float* q, int64_t nq, int* counts, hipStream_t s);
void launch_neighbor_write(GridDev g, const float* q, int64_t nq, const int* offsets, const int* sidx, int* ids,
                           hipStream_t s);
// *out += sum of n non-negative counts, in 64 bits
void launch_sum_counts(const int* counts, int64_t n, unsigned long long* out, hipStream_t s);
// p[0 .. n) = v (a kernel, not a memset, so the work can be captured in a graph)
void launch_fill_i32(int* p, int64_t n, int v, hipStream_t s);
// out[i] = in[i] widened to 64 bits
void launch_widen_i64(const int* in, int64_t n, int64_t* out, hipStream_t s);
void launch_box_count(GridDev g, const int* lo3, const int* n3, int64_t ncells, int* counts, hipStream_t s);
void launch_box_write(GridDev g, const int* lo3, const int* n3, int64_t ncells, const int* offsets, const int* sidx,
                      int* ids, hipStream_t s);
void launch_cell_window(GridDev g, int* start, int* end, hipStream_t s);

// ---- anisotropy ----
struct AnisoParams {
    float radius;          // r
    float nI;              // neighbour radius r * 4
    float nI2;             // nI * nI (float, as compared in the reference)
    double h2;             // Influnce2
    float xmean_delta;     // 0.9 (B-spline) / unused in poly6 mode
    int use_poly6;
    int use_ani;
    double nI3;            // pow(double(nI), 3) as the host's libm gives it (wij's denominator)
};
// one thread per sorted particle (n_invalid: particles outside the grid; these kernels handle any)
void launch_xmean(int prec, GridDev g, const float4* spos, int64_t n, int64_t n_invalid, AnisoParams ap,
                  void* xm /*R x4*/, hipStream_t s);
void launch_gmat(int prec, GridDev g, const float4* spos, int64_t n, int64_t n_invalid, AnisoParams ap, const void* xm,
                 void* rec_a, void* rec_g, void* rec_lo, unsigned long long* n_fp64_eig, hipStream_t s);

// ---- field ----
// Event counters, spread over kCtrStripes copies one cache line apart: a block adds to the copy of
// its index, so the per-wave atomics of a launch do not all serialise on one address. The host
// sums the copies.
constexpr int kCtrStripes = 64;
struct alignas(128) Counters {
    unsigned long long field_evals, grad_evals, candidates, qef_solves, sample_evals, sample_candidates,
        sample_unique, run_evals, run_candidates;
};
__device__ __forceinline__ Counters* ctr_stripe(Counters* c) { return c + (blockIdx.x & (kCtrStripes - 1)); }
// wave-reduced add: one atomic per wave (every lane of the wave must call it)
__device__ __forceinline__ void wave_count(unsigned long long* ctr, unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(ctr, v);
}
void launch_eval_points(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, const float* q,
                        int64_t n, double* f, double* grad, Counters* ctr, hipStream_t s);
// trial_ctr: counter block for the timed trial launches of shape tuning (their counts are not kept)
void launch_normals(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, const float4* spos,
                    int64_t n, int64_t n_invalid, float4* nrm, Counters* ctr, Counters* trial_ctr, hipStream_t s);
// dense sampling box: lo / hi corners, steps (intervals) and dims = steps + 1 points per axis
struct UniformBox {
    float lo[3], hi[3];
    int steps[3], dims[3];
};
void launch_uniform_grid(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, UniformBox ub,
                         int64_t total, float* out, Counters* ctr, hipStream_t s);

// ---- octree levels ----
struct NodeLevel {
    float4* c;         // centre xyz, half
    int8_t* depth;
    uint64_t* id;
    uint8_t* type;
    uint8_t* empty;
    float* curv;
    double4* node;     // vertex xyz, field value
    int* flag;         // INTERNAL -> 1 (scanned into child offsets)
    int* child;        // level-local index of the first child in the next level, -1
    int4* ip;          // integer cell coordinates at the node's depth (x, y, z, depth)
    int64_t n;         // node count, or with ndev an upper bound that sizes the arrays and grids
    const int* ndev = nullptr;  // node count in device memory (levels built without a host read)
};
// node count of a level as the kernels see it
__device__ __forceinline__ int64_t lv_count(const NodeLevel& lv) { return lv.ndev ? (int64_t)*lv.ndev : lv.n; }
struct TreeParams {
    float radius;
    float margin;      // r * smooth_factor (box margin of checkEmptyAndCalcCurv)
    float b_lo[3], b_hi[3];  // HashGrid::Bounding
    int depth_min, depth_max;
    int calc_normals;
    int curv;          // 0: emptiness only (leaf depth of a frame, where the reference discards curv)
    int qef_mode;
    float qef_svd_tol;
};
// group: lanes per node (4, 8, 16 or 64), about the number of (y, z) cell rows a node box spans.
// pass: kEmptyAndCurv both; kEmptyOnly emptiness alone (stops at the first particle found);
// kCurvOnly the curvature of the sampled nodes whose 27 samples (R = rsz bytes) change sign, the
// only nodes whose classification reads it.
constexpr int kEmptyAndCurv = 0, kEmptyOnly = 1, kCurvOnly = 2;
void launch_empty_curv(GridDev g, const float4* spos, const float4* nrm, TreeParams tp, NodeLevel lv, int group,
                       int pass, const void* samples, int rsz, hipStream_t s);
// flag[n+1], pos[n+1] (pos[n] = number of listed nodes, published to pub), list[n]: nodes to
// sample; cflag / cpos / clist: leaf-depth nodes that need only their centre value (cpos[n]
// published to cpub)
void launch_sample_list(TreeParams tp, NodeLevel lv, int* flag, int* pos, int* list, int* cflag, int* cpos,
                        int* clist, int* tmp, hipStream_t s, ScanPub pub = ScanPub(), ScanPub cpub = ScanPub());
void launch_centers(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, NodeLevel lv,
                    const int* list, const int* count, void* samples, Counters* ctr, hipStream_t s);
// the same for k listed centres (count known on the host), grouped by hash cell and evaluated in
// runs of cells; key / val / key2 / val2 hold k entries, cwork 4 (k + 2) ints. The counts of groups
// and runs are read back through rd; a failed read sets *err (left untouched otherwise).
void launch_centers_runs(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, NodeLevel lv,
                         const int* list, int64_t k, uint64_t* key, int* val, uint64_t* key2, int* val2, int* cwork,
                         int* tmp, void* samples, Counters* ctr, Counters* trial_ctr, HostRead* rd,
                         hipStream_t s, hipError_t* err);
// sample deduplication: keys for the nent = 27 * listed entries (hash cell group, then the
// lattice position modulo 2^lb per axis; the group starts at bit gshift = 3 lb), then (after the
// key sort) heads, the distinct points grouped into runs of cells, one workgroup per run between
// events e0 / e1, and the scatter to samples. head
// holds 2 (nent / 256 + 2) ints, hpos / rep / hlist nent, cwork 4 (nent + 2); kx_out is the
// number of cells of a row per run. Counts are read back through rd; a failed read sets *err.
void launch_sample_dedupe_keys(GridDev g, NodeLevel lv, const int* list, int64_t nent, const float origin[3],
                               int lb, int vbits, uint64_t* key, int* val, hipStream_t s);
// Sample records merged per tile of 256 listed nodes: the record count in *nrec (cleared here),
// every entry's record slot (ent_rec, < k * 27), per record its sort key as
// launch_sample_dedupe_keys makes it (packed with the slot when vbits > 0, else the slot in rval)
// and point (rent: node * 27 + sample; the level's node count times 27 must fit an int)
void launch_sample_tiles(GridDev g, NodeLevel lv, const int* list, int64_t k, const float origin[3], int lb,
                         int vbits, int* nrec, int* ent_rec, uint64_t* rkey, int* rval, int* rent, hipStream_t s);
// Distinct points of nrec sorted items evaluated and their values scattered to the level's samples.
// ent NULL: the items are the sample entries (nent == nrec, ent_rec and rec_head unused); else
// item value v is a record slot, ent[v] its node * 27 + sample, ent_rec the record slot of each of
// nent entries and rec_head (one int per slot) scratch.
void launch_sample_dedupe_eval(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, NodeLevel lv,
                               const int* list, SortedKV kv, const int* ent, int gshift, int64_t nrec,
                               const int* ent_rec, int64_t nent, int* rec_head, int* head, int* hpos, int* rep,
                               int* hlist, int* cwork, int64_t* nheads_out, int64_t* nruns_out, int* kx_out,
                               void* vals, void* samples, int* tmp, Counters* ctr, Counters* trial_ctr, hipEvent_t e0,
                               hipEvent_t e1, HostRead* rd, hipStream_t s, hipError_t* err);
// Phase-split deduplicated sampling (the steps of launch_sample_dedupe_eval and
// launch_centers_runs cut at each count the host reads), for sweeps over many frames:
// heads of nent sorted entries, their count at heads_count(head, nent)
int* heads_count(int* head, int64_t nent);
void launch_heads(NodeLevel lv, const int* list, SortedKV kv, int64_t nent, int* head, int* rep, int* hlist, int* hpos,
                  int* tmp, hipStream_t s);
// cell groups of nh points (hlist NULL: the sorted entries are the points), count at
// groups_count(cwork, nh)
int* groups_count(int* cwork, int nh);
void launch_cell_groups(SortedKV kv, int gshift, const int* hlist, int nh, int* cwork, int* tmp, hipStream_t s);
// runs of kx cells over ng groups: count at runs_count, run starts at runs_start
int* runs_count(int* cwork, int nh, int ng);
int* runs_start(int* cwork, int nh, int ng);
void launch_run_starts(GridDev g, SortedKV kv, int gshift, const int* hlist, int nh, int ng, int kx, int* cwork,
                       int* tmp, hipStream_t s);
// block size and cells per run for an average of p points per cell
void run_shape_for(double points_per_cell, int* kx, int* t);
void launch_sample_scatter(const int* list, SortedKV kv, const int* rep, const int* hpos, int64_t nent,
                           const void* vals, int rsz, void* samples, hipStream_t s);
// leaf-depth centres of the listed nodes keyed by cell group and sorted
SortedKV launch_center_keys_sorted(GridDev g, NodeLevel lv, const int* list, int64_t k, uint64_t* key, int* val,
                                   uint64_t* key2, int* val2, int* tmp, hipStream_t s);
// one frame's runs in the multi-frame queue: kind 0 = distinct sample points (values to out[h]),
// kind 1 = leaf-depth centres (values to slot 13 of the node's samples in out)
struct RunFrame {
    GridDev g;
    const void* ra;
    const void* rg;
    FieldConst fc;
    NodeLevel lv;
    const int* list;
    SortedKV kv;
    const int* hlist;
    const int* rstart;
    void* out;
    Counters* ctr;
    int kind;
};
// runs of every frame of a level in one launch fed by a device counter: item_off[f] = first item
// of frame f (nframes + 1 entries, device), items (nitems, device) and qhead (one int, device) are
// work space; at most max_blocks workgroups of T (128 or 256) threads
void launch_run_queue(int prec, int mode, int T, const RunFrame* frames, const int* item_off, int nframes, int nitems,
                      int2* items, int* qhead, int max_blocks, hipStream_t s);
void launch_samples(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, TreeParams tp,
                    NodeLevel lv, const int* list, const int* count, void* samples /*n*27 R*/, Counters* ctr,
                    hipStream_t s);
// node types, vertices and values, and lv.flag (1 for an internal node; lv.flag[lv.n] = 0 closes
// the flags for their scan)
void launch_classify(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, TreeParams tp,
                     NodeLevel lv, const void* samples, int32_t* signchange, Counters* ctr, hipStream_t s);
void launch_children(NodeLevel parent, const int* scanned, NodeLevel next, hipStream_t s);
// the same for a parent level with a device count (parent.ndev, parent.n its upper bound): the
// children's count 8 * scanned[parent.n] is written to *next_count
void launch_children_counted(NodeLevel parent, const int* scanned, NodeLevel next, int* next_count, hipStream_t s);
void launch_sample_grads(int prec, TreeParams tp, NodeLevel lv, const void* samples, double* out_samples,
                         double* out_grads, hipStream_t s);

// ---- device extraction ----
constexpr int kMaxLevels = 32;
struct TreeView {
    int nlev;                          // levels present
    int dp;                            // deepest level (finest integer resolution 2^dp)
    int64_t off[kMaxLevels];           // global node offset of each level
    const uint8_t* type[kMaxLevels];
    const int* child[kMaxLevels];
    // index of the tree at level glev (<= dp): for each of the (2^glev)^3 cells, the node of level
    // glev containing it, or the leaf above glev that does, as (level << 27) | level-local index
    const int* gidx = nullptr;
    int glev = 0;
};
// levels of the tree index of a tree of deepest level dp; the index (8^glev cells), built from
// the one a level above in coarse (8^(glev - 1) cells) when coarse is not NULL
int tree_index_level(int dp);
void launch_tree_index(TreeView tv, int* gidx, int* coarse, hipStream_t s);
// the levels of a built tree as the extraction's per-node passes read them
struct LeafLevels {
    int nlev;                          // levels present
    int dp;                            // deepest level (finest integer resolution 2^dp)
    int64_t off[kMaxLevels + 1];       // global index of each level's first node; off[nlev] = nodes
    int64_t tile0[kMaxLevels + 1];     // first corner tile of each level; tile0[nlev] = tiles
    const uint8_t* type[kMaxLevels];
    const double4* node[kMaxLevels];
    const int4* ip[kMaxLevels];
};
// flag[i] = 1 when global node i is a leaf, one launch for every level
void launch_leaf_flags(const LeafLevels& lv, int* flag, hipStream_t s);
// Distinct corner points per tile of corner_tile_records() / 8 nodes of a level (lv.tile0 from
// corner_tiles() of each level's node count), one launch for every level. Tile b writes its
// records to slots b * corner_tile_records() + u, u < tcount[b]: the point's Morton key (3 dp
// bits) in fkey, and in rleaf[8 slot + o] the tile's leaf in octant o of the point, or -1;
// tcount[ntiles] = 0 closes the counts for their scan. Also leaf_node[L] = vertex and value of
// leaf L, leaf_pos[L] = (value > 0).
int64_t corner_tiles(int64_t nodes);
int64_t corner_tile_records();
void launch_corner_tiles(const LeafLevels& lv, const int* lidx, double4* leaf_node, uint8_t* leaf_pos, uint64_t* fkey,
                         int* rleaf, int* tcount, hipStream_t s);
// the records' keys, tile b's from tbase[b] on (the scanned counts), packed with their slots when
// vbits > 0 (else the slots in rval)
void launch_records_compact(const int* tcount, const int* tbase, int64_t ntiles, const uint64_t* fkey, int vbits,
                            uint64_t* rkey, int* rval, hipStream_t s);
// first[i] = 1 when i is the smallest index with key[i] (open-addressing table of tsize slots, a
// power of two above the number of distinct keys, with room to spare; slot[i] = slot of key[i],
// table[slot] = smallest index of the key; n < 2^32 - 1)
void launch_hash_dedupe(const uint64_t* key, int64_t n, unsigned* table, int64_t tsize, int* slot, int* first,
                        hipStream_t s);
void launch_vertex_ids_hashed(const int* slot, const unsigned* table, const int* rank, int64_t n, const int* vend,
                              const double4* leaf_node, int* corner_vid, double* ends, hipStream_t s);
void launch_tri_key63(const int* cv, int64_t nt, uint64_t* k, hipStream_t s);
// point t (first sorted record heads[t]; the run of its key's records, rleaf as written by
// launch_corner_tiles): dflag[t] = 1 when its dual cell is non-trivial, the cell goes to
// stage[9 t ..]
void launch_points(const TreeView& tv, const int* lidx, const uint8_t* leaf_pos, SortedKV kv, int64_t nent,
                   const int* rleaf, const int* heads, int64_t npts, int* dflag, int* stage, hipStream_t s);
void launch_points_compact(const int* dflag, const int* dpos, const int* stage, int64_t nent, int* duals,
                           hipStream_t s);
void launch_tri_count(const int* duals, int64_t nd, int* ntri, hipStream_t s);
void launch_tri_corners(const int* duals, int64_t nd, const int* toff, uint64_t* vkey, int* vval, int* vend,
                        hipStream_t s);
// list[0 .. h) = the run heads of n sorted entries in order, and h in *head_list_count(tiles, n).
// tiles holds head_list_tmp_ints(n) ints; scan_tmp is exclusive_scan_i32 workspace for
// n / 2048 + 2 ints.
int64_t head_list_tmp_ints(int64_t n);
inline int* head_list_count(int* tiles, int64_t n) { return tiles + 2 * ((n + 2047) / 2048) + 1; }
// (n > 0: the head count is published to pub)
void launch_head_list(SortedKV kv, int64_t n, int* list, int* tiles, int* scan_tmp, hipStream_t s,
                      ScanPub pub = ScanPub());
void launch_tri_key_hi(const int* cv, int64_t nt, uint32_t* k, int* v, hipStream_t s);
void launch_tri_key_lo(const int* cv, const int* perm, int64_t nt, uint64_t* k, int* v, hipStream_t s);
void launch_tri_keep(const int* cv, const int* order, int64_t nt, int* keep, hipStream_t s);
void launch_tri_out(const int* cv, const int* keep, const int* kpos, int64_t nt, int* tris, hipStream_t s);

// ---- vertex placement ----
void launch_bisect(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, float half_r,
                   const double* ends, int64_t n, double* out, int* nev, Counters* ctr, hipStream_t s);
void launch_qef(int prec, GridDev g, const void* rec_a, const void* rec_g, FieldConst fc, TreeParams tp,
                const float4* cen, const double* samples, const double* grads, int64_t n, int mode, double4* out,
                Counters* ctr, hipStream_t s);

}  // namespace pisurf

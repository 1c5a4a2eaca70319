// Hash grid on the device: bounding box, cell keys, stable LSD radix sort, CSR cell offsets,
// and the neighbour / box queries of HashGrid (src/hash_grid.cpp).
#include <atomic>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "pisurf_internal.h"

namespace pisurf {

// ---------------------------------------------------------------------------------------------
// Exclusive scan (int32): tile of 2048 = 256 threads x 8 consecutive items, wave64 shuffles.
// ---------------------------------------------------------------------------------------------
constexpr int kScanT = 256, kScanI = 8, kScanTile = kScanT * kScanI;

// The thread that writes a scan's last element in its final kernel also publishes it to a
// read-back slot (pdst, non-null only for that kernel): the value at byte poff, then, once it is
// visible system-wide, the sequence number the host waits for.
__device__ __forceinline__ void scan_publish(char* pdst, int poff, uint32_t pseq, int v) {
    *(volatile int*)(pdst + poff) = v;
    __threadfence_system();
    *(volatile uint32_t*)(pdst + kReadbackBytes) = pseq;
    __threadfence_system();
}

__global__ void __launch_bounds__(kScanT) k_scan_tile(const int* __restrict__ in, int* __restrict__ out, int64_t n,
                                                      int* __restrict__ sums, char* pdst, int poff, uint32_t pseq) {
    __shared__ int wsum[kScanT / kWave];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)t * kScanI;
    int v[kScanI];
    int acc = 0;
#pragma unroll
    for (int k = 0; k < kScanI; k++) {
        const int64_t i = base + k;
        v[k] = (i < n) ? in[i] : 0;
        acc += v[k];
    }
    const int incl = wave_incl_scan(acc, lane);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int woff = 0;
    for (int k = 0; k < w; k++) woff += wsum[k];
    int run = woff + incl - acc;
#pragma unroll
    for (int k = 0; k < kScanI; k++) {
        const int64_t i = base + k;
        if (i < n) out[i] = run;
        if (pdst && i == n - 1) scan_publish(pdst, poff, pseq, run);
        run += v[k];
    }
    if (t == kScanT - 1) sums[blockIdx.x] = woff + incl;
}

__global__ void k_scan_add(int* __restrict__ out, int64_t n, const int* __restrict__ add, char* pdst, int poff,
                           uint32_t pseq) {
    const int64_t i = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
    const int a = add[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanI; k++) {
        const int64_t j = i + (int64_t)k * kScanT;
        if (j < n) {
            const int v = out[j] + a;
            out[j] = v;
            if (pdst && j == n - 1) scan_publish(pdst, poff, pseq, v);
        }
    }
}

// Second pass for up to kScanDirect tiles: each tile adds the sum of the tile totals before it,
// reduced by the block itself, so the totals need no scan of their own (one launch fewer).
constexpr int kScanDirect = 2048;
__global__ void __launch_bounds__(kScanT) k_scan_add_prefix(int* __restrict__ out, int64_t n,
                                                            const int* __restrict__ sums, char* pdst, int poff,
                                                            uint32_t pseq) {
    __shared__ int wsum[kScanT / kWave];
    const int t = threadIdx.x;
    if (blockIdx.x == 0) return;  // uniform over the block
    int a = 0;
    for (int k = t; k < (int)blockIdx.x; k += kScanT) a += sums[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, kWave);
    if ((t & 63) == 0) wsum[t >> 6] = a;
    __syncthreads();
    a = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    const int64_t i = (int64_t)blockIdx.x * kScanTile + t;
#pragma unroll
    for (int k = 0; k < kScanI; k++) {
        const int64_t j = i + (int64_t)k * kScanT;
        if (j < n) {
            const int v = out[j] + a;
            out[j] = v;
            if (pdst && j == n - 1) scan_publish(pdst, poff, pseq, v);
        }
    }
}

int64_t scan_tmp_ints(int64_t n) {
    int64_t tot = 0;
    while (n > 1) {
        n = (n + kScanTile - 1) / kScanTile;
        tot += 2 * n + 2;
    }
    return tot + 4;
}

void exclusive_scan_i32(const int* in, int* out, int64_t n, int* tmp, hipStream_t s, ScanPub pub) {
    if (n <= 0) return;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    int* sums = tmp;
    int* sums_scanned = tmp + nb + 1;
    k_scan_tile<<<dim3((unsigned)nb), dim3(kScanT), 0, s>>>(in, out, n, sums, nb == 1 ? pub.dst : nullptr, pub.off,
                                                            pub.seq);
    if (nb > 1 && nb <= kScanDirect) {
        k_scan_add_prefix<<<dim3((unsigned)nb), dim3(kScanT), 0, s>>>(out, n, sums, pub.dst, pub.off, pub.seq);
    } else if (nb > 1) {
        exclusive_scan_i32(sums, sums_scanned, nb, tmp + 2 * nb + 2, s);
        k_scan_add<<<dim3((unsigned)nb), dim3(kScanT), 0, s>>>(out, n, sums_scanned, pub.dst, pub.off, pub.seq);
    }
}

// ---------------------------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits. Per tile of 4096 keys: a digit histogram, a global
// digit-major exclusive scan of the per-tile histograms, then a stable scatter (below). (10-bit
// digits would save a pass on 21-30-bit keys, but a tile of 4096 keys then averages 4 keys per
// digit and the scatter's writes break into 32-byte pieces: a 10-bit pass of the 23 M sample
// entries took 255 us against 130 us for an 8-bit one.)
// ---------------------------------------------------------------------------------------------
constexpr int kRsT = 256, kRsRounds = 16, kRsTile = kRsT * kRsRounds, kRsDigits = 256;

template <class KT>
__global__ void __launch_bounds__(kRsT) k_rs_hist(const KT* __restrict__ keys, int64_t n, int shift,
                                                  int* __restrict__ bh, int nblocks) {
    __shared__ int h[4][kRsDigits];  // one histogram per wave: atomics of different waves do not collide
    const int t = threadIdx.x, w = t >> 6;
    h[0][t] = h[1][t] = h[2][t] = h[3][t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
#pragma unroll 4
    for (int r = 0; r < kRsRounds; r++) {
        const int64_t i = base + r * kRsT + t;
        if (i < n) atomicAdd(&h[w][(int)((keys[i] >> shift) & 255u)], 1);
    }
    __syncthreads();
    bh[(int64_t)t * nblocks + blockIdx.x] = h[0][t] + h[1][t] + h[2][t] + h[3][t];
}

// Scatter of one tile: wave w takes the tile's keys [w * 1024, (w + 1) * 1024) in 16 rounds of 64.
// Within a round, lanes of equal digit find each other with 8 ballots; a wave's running count per
// digit lives in LDS (the one leader lane per digit advances it), so a key's offset among its
// wave's keys of that digit needs no barrier. The per-wave counts then give every key its place
// in the tile sorted by digit (thread t owns consecutive digits); the tile is staged in LDS in
// that order and written out by consecutive threads, so each digit's run of keys goes to
// consecutive addresses.
template <class KT, bool VAL>
__global__ void __launch_bounds__(kRsT) k_rs_scatter(const KT* __restrict__ kin, const int* __restrict__ vin,
                                                     KT* __restrict__ kout, int* __restrict__ vout, int64_t n,
                                                     int shift, const int* __restrict__ bscan, int nblocks) {
    constexpr int DB = 8, ND = 1 << DB, DPT = ND / kRsT;
    __shared__ int base[ND];
    __shared__ int cnt[4][ND];
    __shared__ int dstart[ND];
    __shared__ int wsum[4];
    __shared__ KT sk[kRsTile];
    __shared__ int sv[VAL ? kRsTile : 1];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int d = t; d < ND; d += kRsT) {
        base[d] = bscan[(int64_t)d * nblocks + blockIdx.x];
        cnt[0][d] = cnt[1][d] = cnt[2][d] = cnt[3][d] = 0;
    }
    __syncthreads();
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int64_t tile0 = (int64_t)blockIdx.x * kRsTile;
    const int64_t chunk = tile0 + (int64_t)w * (kRsRounds * kWave);
    KT k[kRsRounds];
    int v[kRsRounds], off[kRsRounds];
#pragma unroll
    for (int r = 0; r < kRsRounds; r++) {
        const int64_t i = chunk + r * kWave + lane;
        const bool valid = i < n;
        k[r] = valid ? kin[i] : KT(0);
        v[r] = (VAL && valid) ? vin[i] : 0;
        const int d = (int)((k[r] >> shift) & (ND - 1));
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < DB; b++) {
            const unsigned long long bb = __ballot((d >> b) & 1);
            m &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int rank = __popcll(m & lt);
        const int before = cnt[w][d];
        off[r] = valid ? before + rank : -1;
        if (valid && rank == 0) cnt[w][d] = before + __popcll(m);
    }
    __syncthreads();
    // thread t owns digits [t DPT, (t + 1) DPT): per-wave offsets, then each digit's start in the
    // sorted tile
    int dtot[DPT], tsum = 0;
#pragma unroll
    for (int i = 0; i < DPT; i++) {
        const int d = t * DPT + i;
        int tot = 0;
#pragma unroll
        for (int ww = 0; ww < 4; ww++) {
            const int c = cnt[ww][d];
            cnt[ww][d] = tot;
            tot += c;
        }
        dtot[i] = tot;
        tsum += tot;
    }
    const int inc = wave_incl_scan(tsum, lane);
    if (lane == kWave - 1) wsum[w] = inc;
    __syncthreads();
    int run = inc - tsum;
    for (int ww = 0; ww < w; ww++) run += wsum[ww];
#pragma unroll
    for (int i = 0; i < DPT; i++) {
        dstart[t * DPT + i] = run;
        run += dtot[i];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsRounds; r++) {
        if (off[r] < 0) continue;
        const int d = (int)((k[r] >> shift) & (ND - 1));
        const int tp = dstart[d] + cnt[w][d] + off[r];
        sk[tp] = k[r];
        if (VAL) sv[tp] = v[r];
    }
    __syncthreads();
    const int nt = (int)min((int64_t)kRsTile, n - tile0);
    for (int e = t; e < nt; e += kRsT) {
        const KT key = sk[e];
        const int d = (int)((key >> shift) & (ND - 1));
        const int pos = base[d] + (e - dstart[d]);
        kout[pos] = key;
        if (VAL) vout[pos] = sv[e];
    }
}

int64_t radix_tmp_ints(int64_t n) {
    const int64_t nb = (n + kRsTile - 1) / kRsTile;
    return 2 * kRsDigits * nb + scan_tmp_ints(kRsDigits * nb) + 16;
}

// Sorts by bits [lo_bit, lo_bit + bits) of the keys; v0 / v1 NULL: keys only. Returns the buffer
// holding the result (k0 or k1; with values, v0 or v1 alike).
template <class KT>
static KT* radix_passes(KT* k0, int* v0, KT* k1, int* v1, int64_t n, int lo_bit, int bits, int* tmp, hipStream_t s) {
    const int nb = (int)((n + kRsTile - 1) / kRsTile);
    int* bh = tmp;
    int* bscan = tmp + kRsDigits * (int64_t)nb;
    int* stmp = bscan + kRsDigits * (int64_t)nb;
    const int passes = (bits + 7) / 8;
    KT *ks = k0, *kd = k1;
    int *vs = v0, *vd = v1;
    for (int p = 0; p < passes; p++) {
        const int shift = lo_bit + 8 * p;
        k_rs_hist<KT><<<nb, kRsT, 0, s>>>(ks, n, shift, bh, nb);
        exclusive_scan_i32(bh, bscan, kRsDigits * (int64_t)nb, stmp, s);
        if (vs) k_rs_scatter<KT, true><<<nb, kRsT, 0, s>>>(ks, vs, kd, vd, n, shift, bscan, nb);
        else k_rs_scatter<KT, false><<<nb, kRsT, 0, s>>>(ks, nullptr, kd, nullptr, n, shift, bscan, nb);
        KT* tk = ks; ks = kd; kd = tk;
        int* tv = vs; vs = vd; vd = tv;
    }
    return ks;
}

template <class KT>
static void radix_sort_impl(KT* k0, int* v0, KT* k1, int* v1, int64_t n, int bits, int* tmp, hipStream_t s) {
    if (n <= 1) return;
    if (radix_passes<KT>(k0, v0, k1, v1, n, 0, bits, tmp, s) != k0) {
        (void)hipMemcpyAsync(k0, k1, n * sizeof(KT), hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(v0, v1, n * sizeof(int), hipMemcpyDeviceToDevice, s);
    }
}


void radix_sort_pairs(uint32_t* k0, int* v0, uint32_t* k1, int* v1, int64_t n, int bits, int* tmp, hipStream_t s) {
    radix_sort_impl<uint32_t>(k0, v0, k1, v1, n, bits, tmp, s);
}

void radix_sort_pairs64(uint64_t* k0, int* v0, uint64_t* k1, int* v1, int64_t n, int bits, int* tmp, hipStream_t s) {
    radix_sort_impl<uint64_t>(k0, v0, k1, v1, n, bits, tmp, s);
}

int pack_vbits(int bits, int64_t n) {
    int vb = 1;
    while (vb < 31 && ((int64_t)1 << vb) < n) vb++;
    return (((int64_t)1 << vb) >= n && bits + vb <= 64) ? vb : 0;
}

SortedKV radix_sort_entries(uint64_t* k0, int* v0, uint64_t* k1, int* v1, int64_t n, int bits, int vbits, int* tmp,
                            hipStream_t s) {
    if (vbits <= 0) {
        radix_sort_impl<uint64_t>(k0, v0, k1, v1, n, bits, tmp, s);
        return SortedKV{k0, v0, 0};
    }
    if (n <= 1) return SortedKV{k0, nullptr, vbits};
    return SortedKV{radix_passes<uint64_t>(k0, nullptr, k1, nullptr, n, vbits, bits, tmp, s), nullptr, vbits};
}

// ---------------------------------------------------------------------------------------------
// Small read-backs (counts, bounding box) in the frame's level loop. The wait polls an event
// instead of blocking in the HIP runtime, so a count that is ready a few microseconds after the copy
// is seen then, not after an interrupt wake-up. A context's counts come back through its read-back
// slot (host-mapped memory written by a small kernel), which saves the copy and the event of an
// ordinary asynchronous transfer: about 8 us per read-back.
// ---------------------------------------------------------------------------------------------
hipError_t stream_wait_on(hipStream_t s, hipEvent_t ev) {
    hipError_t e;
    if ((e = hipEventRecord(ev, s)) != hipSuccess) return e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) __builtin_ia32_pause();
    return e;
}

namespace {
std::mutex g_slot_m;
std::vector<HostRead*> g_slot_free;
constexpr size_t kSlotBytes = kReadbackBytes + 256;  // the sequence word sits at kReadbackBytes

// bytes (a multiple of 4 when word) to the slot, then, after every thread's stores are visible
// system-wide, the sequence number
__global__ void __launch_bounds__(256) k_publish(const void* __restrict__ src, int bytes, char* dst, uint32_t seq) {
    if ((bytes & 3) == 0) {
        for (int i = threadIdx.x; i < bytes / 4; i += 256) ((uint32_t*)dst)[i] = ((const uint32_t*)src)[i];
    } else {
        for (int i = threadIdx.x; i < bytes; i += 256) dst[i] = ((const char*)src)[i];
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        *(volatile uint32_t*)(dst + kReadbackBytes) = seq;
        __threadfence_system();
    }
}

// Poll the sequence word until it reaches want (or a later number: a scan's published value can
// be read after later reads of the slot). Every few thousand polls the stream is queried, so that
// a device error ends the wait, and a stream that has finished without the word showing ends it
// too.
hipError_t slot_wait(HostRead* rd, uint32_t want, hipStream_t s) {
    volatile uint32_t* w = (volatile uint32_t*)(rd->host + kReadbackBytes);
    auto reached = [&] { return (int32_t)(*w - want) >= 0; };
    for (uint32_t i = 1;; i++) {
        if (reached()) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return hipSuccess;
        }
        __builtin_ia32_pause();
        if ((i & 4095) == 0) {
            const hipError_t e = hipStreamQuery(s);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return e;
            for (int k = 0; k < (1 << 16); k++)
                if (reached()) {
                    std::atomic_thread_fence(std::memory_order_acquire);
                    return hipSuccess;
                }
            return hipErrorUnknown;
        }
    }
}
}  // namespace

HostRead* host_read_acquire() {
    std::lock_guard<std::mutex> lk(g_slot_m);
    if (!g_slot_free.empty()) {
        HostRead* r = g_slot_free.back();
        g_slot_free.pop_back();
        return r;
    }
    HostRead* r = new (std::nothrow) HostRead();
    if (!r) return nullptr;
    void* h = nullptr;
    void* d = nullptr;
    if (hipHostMalloc(&h, kSlotBytes, hipHostMallocMapped | hipHostMallocCoherent | hipHostMallocPortable) != hipSuccess) {
        delete r;
        return nullptr;
    }
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
        // kept, like every slot, until the process ends
        delete r;
        return nullptr;
    }
    std::memset(h, 0, kSlotBytes);
    r->host = (char*)h;
    r->dev = (char*)d;
    r->next = 0;
    return r;
}

void host_read_release(HostRead* rd) {
    if (!rd) return;
    std::lock_guard<std::mutex> lk(g_slot_m);
    g_slot_free.push_back(rd);
}

hipError_t read_device(void* host, const void* dptr, size_t bytes, hipStream_t s, HostRead* rd) {
    if (bytes > kReadbackBytes || !rd) return hipErrorInvalidValue;
    const uint32_t want = ++rd->next;
    k_publish<<<1, 256, 0, s>>>(dptr, (int)bytes, rd->dev, want);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = slot_wait(rd, want, s)) != hipSuccess) return e;
    std::memcpy(host, rd->host, bytes);
    return hipSuccess;
}

ScanPub scan_read_begin(HostRead* rd, int off) {
    ScanPub p;
    if (rd && off >= 0 && off + 4 <= (int)kReadbackBytes) {
        p.dst = rd->dev;
        p.off = off;
        p.seq = ++rd->next;
    }
    return p;
}

hipError_t scan_read_end(int* host, const int* dlast, ScanPub p, hipStream_t s, HostRead* rd) {
    if (!p.dst) return read_device(host, dlast, sizeof(int), s, rd);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = slot_wait(rd, p.seq, s)) != hipSuccess) return e;
    std::memcpy(host, rd->host + p.off, sizeof(int));
    return hipSuccess;
}

hipError_t stream_fence(hipStream_t s, HostRead* rd) {
    if (!rd) return hipErrorInvalidValue;
    const uint32_t want = ++rd->next;
    k_publish<<<1, 256, 0, s>>>(nullptr, 0, rd->dev, want);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return slot_wait(rd, want, s);
}

// ---------------------------------------------------------------------------------------------
// Tight bounding box (SurfReconstructor::loadRootBox, src/surface_reconstructor.cpp:30-49):
// min/max are exact, so the reduction order does not matter. NaN coordinates are counted.
// ---------------------------------------------------------------------------------------------
constexpr int kBbT = 256, kBbBlocks = 1024;

__global__ void __launch_bounds__(kBbT) k_bbox_partial(const float* __restrict__ xyz, int64_t n,
                                                       float* __restrict__ part, int* __restrict__ nan_count) {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    int nans = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBbT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBbT) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float v = xyz[3 * i + a];
            if (v != v) { nans++; continue; }
            mn[a] = fminf(mn[a], v);
            mx[a] = fmaxf(mx[a], v);
        }
    }
    __shared__ float smn[3][kBbT], smx[3][kBbT];
    __shared__ int sn[kBbT];
    for (int a = 0; a < 3; a++) { smn[a][threadIdx.x] = mn[a]; smx[a][threadIdx.x] = mx[a]; }
    sn[threadIdx.x] = nans;
    __syncthreads();
    for (int o = kBbT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            for (int a = 0; a < 3; a++) {
                smn[a][threadIdx.x] = fminf(smn[a][threadIdx.x], smn[a][threadIdx.x + o]);
                smx[a][threadIdx.x] = fmaxf(smx[a][threadIdx.x], smx[a][threadIdx.x + o]);
            }
            sn[threadIdx.x] += sn[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int a = 0; a < 3; a++) { part[6 * blockIdx.x + 2 * a] = smn[a][0]; part[6 * blockIdx.x + 2 * a + 1] = smx[a][0]; }
        if (sn[0]) atomicAdd(nan_count, sn[0]);
    }
}

__global__ void __launch_bounds__(kBbT) k_bbox_final(const float* __restrict__ part, int nb, float* __restrict__ out) {
    __shared__ float s[6][kBbT];
    for (int a = 0; a < 3; a++) { s[2 * a][threadIdx.x] = FLT_MAX; s[2 * a + 1][threadIdx.x] = -FLT_MAX; }
    for (int b = threadIdx.x; b < nb; b += kBbT)
        for (int a = 0; a < 3; a++) {
            s[2 * a][threadIdx.x] = fminf(s[2 * a][threadIdx.x], part[6 * b + 2 * a]);
            s[2 * a + 1][threadIdx.x] = fmaxf(s[2 * a + 1][threadIdx.x], part[6 * b + 2 * a + 1]);
        }
    __syncthreads();
    for (int o = kBbT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int a = 0; a < 3; a++) {
                s[2 * a][threadIdx.x] = fminf(s[2 * a][threadIdx.x], s[2 * a][threadIdx.x + o]);
                s[2 * a + 1][threadIdx.x] = fmaxf(s[2 * a + 1][threadIdx.x], s[2 * a + 1][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

void launch_bbox(const float* xyz, int64_t n, float* partial, float* out6, int* nan_count, hipStream_t s) {
    int nb = (int)std::min<int64_t>(kBbBlocks, (n + kBbT - 1) / kBbT);
    if (nb < 1) nb = 1;
    k_bbox_partial<<<nb, kBbT, 0, s>>>(xyz, n, partial, nan_count);
    k_bbox_final<<<1, kBbT, 0, s>>>(partial, nb, out6);
}

// ---------------------------------------------------------------------------------------------
// Cell keys (CalcHashList, src/hash_grid.cpp:87-101). key+1 is sorted as an unsigned value so the
// out-of-range key -1 sorts first, as the signed keys do in the reference.
// ---------------------------------------------------------------------------------------------
__global__ void k_keys(const float* __restrict__ xyz, int64_t n, GridDev g, uint32_t* __restrict__ keys,
                       int* __restrict__ vals, int* __restrict__ n_invalid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ix, iy, iz;
    cell_of(g, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], ix, iy, iz);
    const bool valid = ix >= 0 && ix < g.n[0] && iy >= 0 && iy < g.n[1] && iz >= 0 && iz < g.n[2];
    const int key = valid ? (iz * g.n[1] + iy) * g.n[0] + ix : -1;
    keys[i] = (uint32_t)(key + 1);
    vals[i] = (int)i;
    if (!valid) atomicAdd(n_invalid, 1);
}

void launch_keys(const float* xyz, int64_t n, GridDev g, uint32_t* keys, int* vals, int* n_invalid, hipStream_t s) {
    if (n <= 0) return;
    k_keys<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(xyz, n, g, keys, vals, n_invalid);
}

// off[c] = number of valid particles in cells below c (FindStartEnd's counts, used by the
// reference as positions into the full sorted list): the first sorted position whose stored key
// (cell + 1, 0 outside the grid) exceeds c, less the particles outside. One thread per cell,
// binary search over the sorted keys.
__global__ void k_cell_offsets(const uint32_t* __restrict__ skeys, int64_t n, const int* __restrict__ n_invalid,
                               int64_t ncell, int* __restrict__ off) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    int64_t lo = 0, hi = n;  // first position with skeys[p] >= c + 1
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)skeys[mid] < c + 1) lo = mid + 1;
        else hi = mid;
    }
    off[c] = (int)(lo - *n_invalid);
}

void launch_cell_offsets(const uint32_t* skeys, int64_t n, const int* n_invalid, int64_t ncell, int* off,
                         hipStream_t s) {
    k_cell_offsets<<<(unsigned)((ncell + 1 + 255) / 256), 256, 0, s>>>(skeys, n, n_invalid, ncell, off);
}

__global__ void k_gather_pos(const float* __restrict__ xyz, const int* __restrict__ sidx, int64_t n,
                             float4* __restrict__ spos) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = sidx[p];
    spos[p] = make_float4(xyz[3 * (int64_t)i], xyz[3 * (int64_t)i + 1], xyz[3 * (int64_t)i + 2], 0.0f);
}

void launch_gather_pos(const float* xyz, const int* sidx, int64_t n, float4* spos, hipStream_t s) {
    if (n <= 0) return;
    k_gather_pos<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(xyz, sidx, n, spos);
}

// dense StartList/EndList view: start[c] = off[c], end[c] = off[c+1]
__global__ void k_cell_window(GridDev g, int* __restrict__ start, int* __restrict__ end) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.ncell) return;
    start[c] = g.off[c];
    end[c] = g.off[c + 1];
}
void launch_cell_window(GridDev g, int* start, int* end, hipStream_t s) {
    if (g.ncell <= 0) return;
    k_cell_window<<<(unsigned)((g.ncell + 255) / 256), 256, 0, s>>>(g, start, end);
}

// ---------------------------------------------------------------------------------------------
// GetPIdxList for query points: count pass, then write pass (ids are particle indices).
// ---------------------------------------------------------------------------------------------
__global__ void k_nb_count(GridDev g, const float* __restrict__ q, int64_t nq, int* __restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    int ix, iy, iz;
    cell_of(g, q[3 * i], q[3 * i + 1], q[3 * i + 2], ix, iy, iz);
    int c = 0;
    for (int r = 0; r < 9; r++) {
        int lo, hi;
        if (row_range(g, ix, iy, iz, r, lo, hi)) c += hi - lo;
    }
    counts[i] = c;
}
__global__ void k_nb_write(GridDev g, const float* __restrict__ q, int64_t nq, const int* __restrict__ offsets,
                           const int* __restrict__ sidx, int* __restrict__ ids) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    int ix, iy, iz;
    cell_of(g, q[3 * i], q[3 * i + 1], q[3 * i + 2], ix, iy, iz);
    int o = offsets[i];
    for (int r = 0; r < 9; r++) {
        int lo, hi;
        if (!row_range(g, ix, iy, iz, r, lo, hi)) continue;
        for (int j = lo; j < hi; j++) ids[o++] = sidx[j];
    }
}
void launch_neighbor_count(GridDev g, const float* q, int64_t nq, int* counts, hipStream_t s) {
    if (nq <= 0) return;
    k_nb_count<<<(unsigned)((nq + 255) / 256), 256, 0, s>>>(g, q, nq, counts);
}
// 64-bit total of n non-negative counts, added to *out (one atomic per wave)
__global__ void __launch_bounds__(256) k_sum_counts(const int* __restrict__ c, int64_t n,
                                                    unsigned long long* __restrict__ out) {
    unsigned long long v = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v += (unsigned)c[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, v);
}
__global__ void k_fill_i32(int* __restrict__ p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_i32(int* p, int64_t n, int v, hipStream_t s) {
    if (n <= 0) return;
    k_fill_i32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
}
__global__ void k_widen_i64(const int* __restrict__ in, int64_t n, int64_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}
void launch_widen_i64(const int* in, int64_t n, int64_t* out, hipStream_t s) {
    if (n <= 0) return;
    k_widen_i64<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, n, out);
}
void launch_sum_counts(const int* counts, int64_t n, unsigned long long* out, hipStream_t s) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 2048);
    k_sum_counts<<<(unsigned)blocks, 256, 0, s>>>(counts, n, out);
}
void launch_neighbor_write(GridDev g, const float* q, int64_t nq, const int* offsets, const int* sidx, int* ids,
                           hipStream_t s) {
    if (nq <= 0) return;
    k_nb_write<<<(unsigned)((nq + 255) / 256), 256, 0, s>>>(g, q, nq, offsets, sidx, ids);
}

// GetInBoxParticles (src/hash_grid.cpp:164-194): cells enumerated x-outer, y, z-inner over
// [lo, lo + n) per axis; out-of-range cells contribute nothing.
__device__ __forceinline__ int box_cell(const GridDev& g, const int* lo3, const int* n3, int64_t t) {
    const int z = lo3[2] + (int)(t % n3[2]);
    const int y = lo3[1] + (int)((t / n3[2]) % n3[1]);
    const int x = lo3[0] + (int)(t / ((int64_t)n3[2] * n3[1]));
    if (x < 0 || x >= g.n[0] || y < 0 || y >= g.n[1] || z < 0 || z >= g.n[2]) return -1;
    return (z * g.n[1] + y) * g.n[0] + x;
}
__global__ void k_box_count(GridDev g, const int* lo3, const int* n3, int64_t ncells, int* counts) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    const int c = box_cell(g, lo3, n3, t);
    counts[t] = (c < 0) ? 0 : g.off[c + 1] - g.off[c];
}
__global__ void k_box_write(GridDev g, const int* lo3, const int* n3, int64_t ncells, const int* offsets,
                            const int* sidx, int* ids) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    const int c = box_cell(g, lo3, n3, t);
    if (c < 0) return;
    int o = offsets[t];
    for (int j = g.off[c]; j < g.off[c + 1]; j++) ids[o++] = sidx[j];
}
void launch_box_count(GridDev g, const int* lo3, const int* n3, int64_t ncells, int* counts, hipStream_t s) {
    if (ncells <= 0) return;
    k_box_count<<<(unsigned)((ncells + 255) / 256), 256, 0, s>>>(g, lo3, n3, ncells, counts);
}
void launch_box_write(GridDev g, const int* lo3, const int* n3, int64_t ncells, const int* offsets, const int* sidx,
                      int* ids, hipStream_t s) {
    if (ncells <= 0) return;
    k_box_write<<<(unsigned)((ncells + 255) / 256), 256, 0, s>>>(g, lo3, n3, ncells, offsets, sidx, ids);
}

}  // namespace pisurf

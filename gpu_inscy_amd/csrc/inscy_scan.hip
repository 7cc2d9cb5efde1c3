// Exclusive prefix sums and the stable key sort of the engine (the reference
// writes its own scans too: inclusive_scan_points/nodes/any, util.cu:74-249).
//
// Scan: reduce, then scan.  Tiles of 4096 ints (256 threads x 16 consecutive
// items); pass 1 sums each tile, pass 2 scans the tile sums in one block,
// pass 3 scans each tile again from its offset.  Inputs of one tile take
// pass 3 alone.  HBM traffic: 8 bytes read and 4 written per item.
//
// Sort: least-significant-digit radix sort of (uint16 key, int value) pairs by
// 8-bit digits, stable.  Per digit: a histogram of every tile of 2048 items,
// the exclusive scan of the [digit][tile] counts (tile order within a digit,
// so equal digits keep their order across tiles), and a scatter that ranks
// each item among the equal digits before it in its tile (wave ballots over
// the 8 digit bits, then the waves of the block in order, step by step).
#include "inscy_internal.h"

namespace inscy {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

size_t align256(size_t b) { return (b + 255) & ~(size_t) 255; }

int scan_tiles(int items) { return (items + kScanTile - 1) / kScanTile; }

// exclusive prefix of v over the block's threads (in thread order); `total`
// receives the block's sum
__device__ __forceinline__ int block_exclusive(int v, int &total) {
    __shared__ int wsum[kScanThreads / 64];
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int pre = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        pre += w < wid ? wsum[w] : 0;
        total += wsum[w];
    }
    __syncthreads();  // wsum is reused by the next call
    return pre + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int *__restrict__ in, int items,
                                                              int *__restrict__ tsum) {
    const size_t at = (size_t) blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {  // striped: coalesced, order-free
        const size_t i = at + (size_t) j * kScanThreads + threadIdx.x;
        s += i < (size_t) items ? in[i] : 0;
    }
    int total;
    (void) block_exclusive(s, total);
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

// exclusive scan of the tile sums in place, one block
__global__ __launch_bounds__(1024) void k_scan_tiles(int *__restrict__ tsum, int nt) {
    __shared__ int wsum[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    for (int base = 0; base < nt; base += 1024) {
        const int i = base + (int) threadIdx.x;
        const int v = i < nt ? tsum[i] : 0;
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        int pre = carry, tot = 0;
        for (int w = 0; w < 16; ++w) {
            pre += w < wid ? wsum[w] : 0;
            tot += wsum[w];
        }
        if (i < nt) tsum[i] = pre + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
}

// each thread scans its 16 consecutive items; tile offset from tsum (or 0)
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const int *__restrict__ in, int items,
                                                            const int *__restrict__ tsum, int *__restrict__ out) {
    const size_t at = (size_t) blockIdx.x * kScanTile + (size_t) threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = at + j;
        v[j] = i < (size_t) items ? in[i] : 0;
        s += v[j];
    }
    int total;
    int run = block_exclusive(s, total) + (tsum ? tsum[blockIdx.x] : 0);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = at + j;
        if (i < (size_t) items) out[i] = run;
        run += v[j];
    }
}

constexpr int kSortThreads = 256;
constexpr int kSortSteps = 8;
constexpr int kSortTile = kSortThreads * kSortSteps;
constexpr int kDigits = 256;

int sort_tiles(int n) { return (n + kSortTile - 1) / kSortTile; }

// item s*256 + t of a tile is thread t's item of step s
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint16_t *__restrict__ keys, int n, int shift,
                                                            int ntile, int *__restrict__ hist) {
    __shared__ int h[kDigits];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t at = (size_t) blockIdx.x * kSortTile;
#pragma unroll
    for (int s = 0; s < kSortSteps; ++s) {
        const size_t i = at + (size_t) s * kSortThreads + threadIdx.x;
        if (i < (size_t) n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1);
    }
    __syncthreads();
    hist[(size_t) threadIdx.x * ntile + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint16_t *__restrict__ keys,
                                                               const int *__restrict__ vals, int n, int shift,
                                                               int ntile, const int *__restrict__ off,
                                                               uint16_t *__restrict__ keys_out,
                                                               int *__restrict__ vals_out) {
    __shared__ int run[kDigits];
    __shared__ int wcnt[kSortThreads / 64][kDigits];
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    run[threadIdx.x] = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; ++w) wcnt[w][threadIdx.x] = 0;
    const int base = off[(size_t) threadIdx.x * ntile + blockIdx.x];  // this tile's first slot of digit threadIdx.x
    __shared__ int toff[kDigits];
    toff[threadIdx.x] = base;
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    const size_t at = (size_t) blockIdx.x * kSortTile;
    for (int s = 0; s < kSortSteps; ++s) {
        const size_t i = at + (size_t) s * kSortThreads + threadIdx.x;
        const bool valid = i < (size_t) n;
        const int key = valid ? keys[i] : 0;
        const int val = valid ? vals[i] : 0;
        const int dg = (key >> shift) & 0xFF;
        // lanes of the wave holding the same digit
        unsigned long long peer = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bb = __ballot(valid && ((dg >> b) & 1));
            peer &= ((dg >> b) & 1) ? bb : ~bb;
        }
        const int rank_w = __popcll(peer & below);
        if (valid && rank_w == 0) wcnt[wid][dg] = __popcll(peer);
        __syncthreads();
        if (valid) {
            int prior = run[dg] + rank_w;
            for (int w = 0; w < wid; ++w) prior += wcnt[w][dg];
            const int dst = toff[dg] + prior;
            keys_out[dst] = (uint16_t) key;
            vals_out[dst] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int w = 0; w < kSortThreads / 64; ++w) {
            add += wcnt[w][threadIdx.x];
            wcnt[w][threadIdx.x] = 0;
        }
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

}  // namespace

size_t scan_temp_bytes(int items) { return align256((size_t) std::max(scan_tiles(std::max(items, 1)), 1) * 4); }

hipError_t launch_exclusive_sum(void *temp, size_t temp_bytes, const int *in, int *out, int items, hipStream_t st) {
    if (items <= 0) return hipSuccess;
    const int nt = scan_tiles(items);
    if (nt == 1) {
        k_scan_down<<<1, kScanThreads, 0, st>>>(in, items, nullptr, out);
        return hipGetLastError();
    }
    if (temp_bytes < scan_temp_bytes(items)) return hipErrorInvalidValue;
    int *tsum = static_cast<int *>(temp);
    k_scan_reduce<<<nt, kScanThreads, 0, st>>>(in, items, tsum);
    k_scan_tiles<<<1, 1024, 0, st>>>(tsum, nt);
    k_scan_down<<<nt, kScanThreads, 0, st>>>(in, items, tsum, out);
    return hipGetLastError();
}

size_t sort_temp_bytes(int n) {
    const size_t cells = (size_t) kDigits * sort_tiles(std::max(n, 1));
    return 2 * align256(cells * 4) + scan_temp_bytes((int) cells);
}

hipError_t launch_sort_pairs_u16(const uint16_t *keys, uint16_t *keys_tmp, uint16_t *keys_out, const int *vals,
                                 int *vals_tmp, int *vals_out, int n, int bits, void *temp, size_t temp_bytes,
                                 hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (temp_bytes < sort_temp_bytes(n)) return hipErrorInvalidValue;
    const int nt = sort_tiles(n);
    const size_t cells = (size_t) kDigits * nt;
    char *t = static_cast<char *>(temp);
    int *hist = reinterpret_cast<int *>(t);
    int *off = reinterpret_cast<int *>(t + align256(cells * 4));
    void *stemp = t + 2 * align256(cells * 4);
    const int passes = bits <= 8 ? 1 : 2;
    const uint16_t *kin = keys;
    const int *vin = vals;
    for (int p = 0; p < passes; ++p) {
        // the last pass writes the outputs; with two passes the first goes to the temporaries
        uint16_t *kout = p + 1 == passes ? keys_out : keys_tmp;
        int *vout = p + 1 == passes ? vals_out : vals_tmp;
        k_sort_hist<<<nt, kSortThreads, 0, st>>>(kin, n, 8 * p, nt, hist);
        hipError_t e = launch_exclusive_sum(stemp, scan_temp_bytes((int) cells), hist, off, (int) cells, st);
        if (e != hipSuccess) return e;
        k_sort_scatter<<<nt, kSortThreads, 0, st>>>(kin, vin, n, 8 * p, nt, off, kout, vout);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        kin = kout;
        vin = vout;
    }
    return hipSuccess;
}

}  // namespace inscy

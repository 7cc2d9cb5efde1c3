// Allocation-free neighbour counts: the grid-index pass of the depth-first
// driver (k_grid_keys, scan, k_grid_scatter, k_grid_count) over buffers carved
// from one file-scope device array, for callers that cannot allocate device
// memory (code under graph capture, serving loops).
//
// Every call on a device shares that device's array and leaves its contents
// behind; calls that overlap on one device (different streams, or threads)
// race on it.  Everything a call reads from the array is written by the same
// call first: the count array is filled with -1 and the grid histogram with 0
// by kernels (no memset nodes), before the passes that read them.
#include "inscy_internal.h"

#include <atomic>

namespace inscy {

// The array itself.  In host code `scratch` names a host-side stand-in, not
// device memory: host code uses only the address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::atomic<unsigned char *> g_scratch_addr[kMaxDevices];

__global__ void k_sub_columns(const float *__restrict__ X, int n, int d, SubspaceArg S, int ecells, float ecs,
                              float *__restrict__ Xt, uint16_t *__restrict__ cellE) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int t = 0; t < S.k; ++t) {
        // the clamp and the cell formula of k_discretise
        float x = X[(size_t) p * d + S.dims[t]];
        x = fminf(fmaxf(x, 0.0f), 1.0f);
        Xt[(size_t) t * n + p] = x;
        cellE[(size_t) t * n + p] = (uint16_t) min((int) __fdiv_rn(x, ecs), ecells - 1);
    }
}

// out[i] = cnt[queries[i]], or -1 for a query id outside [0, n)
__global__ void k_gather_checked(const int *__restrict__ cnt, int n, const int *__restrict__ queries, int nq,
                                 int *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const int q = queries[i];
    out[i] = (q >= 0 && q < n) ? cnt[q] : -1;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t) 255; }

// index dimensions of the subspace: the first j of its k positions with
// C^j <= max(C, min(8n, 2^20)) cells (every dimension is matched by the
// identity level order; which ones index changes the speed, not the counts)
GridArg noalloc_grid(int n, int k, int ecells) {
    GridArg G{};
    G.C = ecells;
    G.j = 1;
    const long long cap = std::max<long long>(ecells, std::min<long long>(8LL * n, 1LL << 20));
    long long cells = ecells;
    while (G.j < 3 && G.j < k && cells * ecells <= cap) {
        cells *= ecells;
        G.j++;
    }
    G.checkmask = k >= 64 ? ~0ull : ((1ull << k) - 1);
    for (int a = 0; a < G.j; ++a) {
        G.tI[a] = a;
        G.checkmask &= ~(1ull << a);
    }
    return G;
}

long long cells_of(const GridArg &G) {
    long long c = 1;
    for (int a = 0; a < G.j; ++a) c *= G.C;
    return c;
}

}  // namespace

size_t noalloc_scratch_need(int n, int k, int ecells) {
    const GridArg G = noalloc_grid(n, k, ecells);
    const size_t ncell = (size_t) cells_of(G);
    const size_t kn = (size_t) k * n;
    return align256(kn * 4) + align256(kn * 2)          // the subspace's columns and eps-cells
           + 4 * align256((size_t) n * 4)                // keys, rank, order, counts
           + align256(kn * 4) + align256(kn * 2)         // sorted coordinates and cells
           + 2 * align256((ncell + 1) * 4)               // grid histogram and cell starts
           + align256(scan_temp_bytes((int) ncell + 1));
}

hipError_t noalloc_scratch(unsigned char **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    unsigned char *p = g_scratch_addr[dev].load(std::memory_order_acquire);
    if (!p) {
        // once per device, with that device current: this device's copy of the array
        e = hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[dev].store(p, std::memory_order_release);
    }
    *out = p;
    return hipSuccess;
}

hipError_t launch_neighbour_counts_noalloc(unsigned char *base, const float *X, int n, int d, float eps,
                                           float s_star, const SubspaceArg &S, const int *queries, int nq, int *out,
                                           hipStream_t st) {
    const int k = S.k;
    const int ecells = (int) ceil(1.0 / (double) eps);
    const GridArg G = noalloc_grid(n, k, ecells);
    const size_t ncell = (size_t) cells_of(G);
    const size_t kn = (size_t) k * n;
    unsigned char *at = base;
    auto take = [&](size_t bytes) {
        unsigned char *p = at;
        at += align256(bytes);
        return p;
    };
    float *Xt = reinterpret_cast<float *>(take(kn * 4));
    uint16_t *cellE = reinterpret_cast<uint16_t *>(take(kn * 2));
    GridBuffers B{};
    B.keys = reinterpret_cast<int *>(take((size_t) n * 4));
    B.rank = reinterpret_cast<int *>(take((size_t) n * 4));
    B.order = reinterpret_cast<int *>(take((size_t) n * 4));
    int *cnt = reinterpret_cast<int *>(take((size_t) n * 4));
    B.Ys = reinterpret_cast<float *>(take(kn * 4));
    B.Es = reinterpret_cast<uint16_t *>(take(kn * 2));
    B.hist = reinterpret_cast<int *>(take((ncell + 1) * 4));
    B.off = reinterpret_cast<int *>(take((ncell + 1) * 4));
    B.scan_bytes = scan_temp_bytes((int) ncell + 1);
    B.scan_temp = take(B.scan_bytes);
    if ((size_t) (at - base) > kNoAllocScratchBytes) return hipErrorInvalidValue;  // checked by the caller first

    const unsigned bs = 256;
    k_sub_columns<<<(unsigned) ((n + bs - 1) / bs), bs, 0, st>>>(X, n, d, S, ecells, 1.0f / (float) ecells, Xt,
                                                                 cellE);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_fill_i32(cnt, -1, (size_t) n, st);
    if (e != hipSuccess) return e;
    // the subspace's columns renumbered 0..k-1: a data set of k dimensions
    SubspaceArg Sk{};
    Sk.k = k;
    for (int t = 0; t < k; ++t) Sk.dims[t] = t;
    const DevData D{Xt, nullptr, cellE, nullptr, nullptr, n, k, 1, ecells};
    e = launch_subspace_counts(D, Sk, G, s_star, INT_MAX, B, nullptr, 0, cnt, nullptr, KernelTimer{}, st);
    if (e != hipSuccess) return e;
    if (nq > 0) {
        k_gather_checked<<<(unsigned) ((nq + bs - 1) / bs), bs, 0, st>>>(cnt, n, queries, nq, out);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace inscy

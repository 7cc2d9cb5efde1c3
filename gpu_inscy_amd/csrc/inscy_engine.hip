// C ABI of the MI355X INSCY engine (include/inscy.h).
#include "inscy_engine.h"

#include <atomic>
#include <mutex>

namespace inscy {

static thread_local std::string g_error;
static thread_local std::vector<Engine::TraceRow> g_trace;
static bool g_trace_on = false;
static bool g_phi_all_exact = false;
static bool g_generic_walk = false;
static long long g_keep_budget = -1;
static bool g_profile_on = false;
static int g_max_batch = 0;  // inscy_set_tuning: child subspaces per batch (0: the default)
static int g_workers = 0;    // inscy_set_tuning: run_many worker streams (0: the default)

static int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

// The exception in flight as a return code and message: nothing thrown inside
// the library crosses the C ABI.
static int fail_current() {
    try {
        throw;
    } catch (const Failure &f) {
        return fail(f.code, f.what());
    } catch (const std::bad_alloc &) {
        return fail(INSCY_ERR_MEMORY, "host allocation failed");
    } catch (const std::system_error &e) {
        return fail(INSCY_ERR_DEVICE, std::string("host thread or system call failed: ") + e.what());
    } catch (const std::exception &e) {
        return fail(INSCY_ERR_DEVICE, std::string("internal error: ") + e.what());
    } catch (...) {
        return fail(INSCY_ERR_DEVICE, "unknown internal error");
    }
}

static int check_params(int64_t n, int32_t d, const inscy_params *p, int32_t mode) {
    if (!p) return fail(INSCY_ERR_ARGUMENT, "params is NULL");
    if (n < 1 || n > (int64_t) INT_MAX / 2) return fail(INSCY_ERR_ARGUMENT, "n out of range");
    if (d < 1 || d > kMaxDims) return fail(INSCY_ERR_ARGUMENT, "d must be in [1, 128]");
    if (!(p->neighborhood_size > 0.0f) || !std::isfinite(p->neighborhood_size))
        return fail(INSCY_ERR_ARGUMENT, "neighborhood_size must be > 0");
    if (p->number_of_cells < 1 || p->number_of_cells > 255)
        return fail(INSCY_ERR_ARGUMENT, "number_of_cells must be in [1, 255]");
    if (mode < INSCY_MODE_EXACT || mode > INSCY_MODE_COLLECT) return fail(INSCY_ERR_ARGUMENT, "unknown mode");
    // A permuted SCY-tree level order runs the depth-first driver over
    // s-connection chains (inscy_perm.hip), with neighbour-count density or,
    // for rectangular = false, phi summed in the neighbourhood tree's leaf order.
    if (p->entropy_order != 0) {
        if (mode == INSCY_MODE_LEARNED || mode == INSCY_MODE_COLLECT)
            return fail(INSCY_ERR_UNSUPPORTED, "entropy_order != 0 runs the exact drivers only");
        if (d > 255) return fail(INSCY_ERR_ARGUMENT, "d must be at most 255");
    }
    return INSCY_OK;
}

// The learned pruner's network: bf16 weights on the host and one copy per
// device, uploaded on first use after each inscy_set_pruner.  The copies are
// held until the process ends.
struct PrunerState {
    std::mutex mu;
    std::vector<uint16_t> w;  // empty: no network loaded
    float thr = 0.f;
    int min_depth = 1;
    int version = 0;
    std::map<int, std::pair<int, uint16_t *>> dev;  // device -> (version, parameters)
};
static PrunerState &pruner_state() {
    static PrunerState *p = new PrunerState;
    return *p;
}

// float32 -> bf16, round to nearest even (the device kernels round alike)
static uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t) ((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// the current device's copy of the network (null when none is loaded)
static const uint16_t *pruner_on_device(float &thr, int &min_depth) {
    PrunerState &P = pruner_state();
    std::lock_guard<std::mutex> g(P.mu);
    if (P.w.empty()) return nullptr;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    auto &e = P.dev[dev];
    if (!e.second) HIP_OK(hipMalloc(&e.second, P.w.size() * sizeof(uint16_t)));
    if (e.first != P.version) {
        HIP_OK(hipMemcpy(e.second, P.w.data(), P.w.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
        e.first = P.version;
    }
    thr = P.thr;
    min_depth = P.min_depth;
    return e.second;
}

static int run_impl(const float *d_X, int64_t n, int32_t d, const inscy_params *p, int32_t mode, hipStream_t user,
                    inscy_result **out, int mem_share = 1) {
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = user;
    bool own = false;
    try {
        if (!st) {
            HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            own = true;
        }
        auto res = std::make_unique<inscy_result>();
        {
            Engine eng(*p, (int) n, d, st);
            eng.trace_on = g_trace_on;
            eng.phi_all_exact = g_phi_all_exact;
            eng.generic_walk = g_generic_walk;
            eng.keep_budget_override = g_keep_budget;
            eng.max_batch_override = g_max_batch;
            eng.profile_on = g_profile_on;
            eng.mem_share = mem_share;
            // phi density (rectangular = false) is implemented in the
            // level-synchronous driver only; it serves every mode
            eng.levels_mode = (mode != INSCY_MODE_EXACT || p->rectangular == 0) && p->entropy_order == 0;
            eng.collect = mode == INSCY_MODE_COLLECT;
            if (mode == INSCY_MODE_LEARNED) {
                eng.pruner = pruner_on_device(eng.prune_thr, eng.prune_min_depth);
                if (!eng.pruner) throw Failure(INSCY_ERR_ARGUMENT, "INSCY_MODE_LEARNED needs a network (inscy_set_pruner)");
            }
            eng.setup(d_X);
            if (p->entropy_order != 0)
                eng.run_perm(*res);
            else if (eng.levels_mode)
                eng.run_levels(*res);
            else
                eng.run(*res);
            res->stats = eng.stats;
            if (g_trace_on) g_trace = eng.trace;
        }
        if (own) (void) hipStreamDestroy(st);
        res->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = res.release();
        g_error.clear();
        return INSCY_OK;
    } catch (...) {
        if (own) (void) hipStreamDestroy(st);
        return fail_current();
    }
}

// Many independent data sets: worker threads, each with its own stream and
// engine, take the sets from a queue ordered by estimated cost (n^2 d,
// largest first) so that the longest runs start first and the streams overlap
// on the device.  Each engine budgets a 1/workers share of the free device
// memory, so the engines running at once do not claim the same bytes.
static int run_many_impl(const float *const *X, const int64_t *n, const int32_t *d, const inscy_params *params,
                         int32_t count, int32_t mode, bool host_input, inscy_result **out) {
    for (int32_t i = 0; i < count; ++i) out[i] = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(INSCY_ERR_DEVICE, "no HIP device");
    std::vector<int> order(count);
    for (int32_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (double) n[a] * (double) n[a] * d[a] > (double) n[b] * (double) n[b] * d[b];
    });
    int workers = 8;
    if (const char *w = std::getenv("INSCY_WORKERS")) workers = std::max(1, std::atoi(w));
    if (g_workers > 0) workers = g_workers;
    workers = std::min<int>(workers, count);
    std::atomic<int> next{0};
    std::vector<int> rc(count, INSCY_OK);
    std::vector<std::string> msg(count);
    auto work = [&]() {
        if (hipSetDevice(dev) != hipSuccess) return;
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return;
        for (int q = next++; q < count; q = next++) {
            const int i = order[q];
            if (host_input) {
                float *dX = nullptr;
                const size_t bytes = (size_t) n[i] * d[i] * sizeof(float);
                if (dev_alloc(&dX, bytes) != hipSuccess) {
                    rc[i] = INSCY_ERR_MEMORY;
                    msg[i] = "device allocation of the input failed";
                    continue;
                }
                if (hipMemcpyAsync(dX, X[i], bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
                    rc[i] = INSCY_ERR_DEVICE;
                    msg[i] = "input copy failed";
                } else {
                    rc[i] = run_impl(dX, n[i], d[i], &params[i], mode, st, &out[i], workers);
                }
                (void) hipStreamSynchronize(st);
                dev_free(dX);
            } else {
                rc[i] = run_impl(X[i], n[i], d[i], &params[i], mode, st, &out[i], workers);
            }
            if (rc[i] != INSCY_OK) msg[i] = g_error;
        }
        (void) hipStreamDestroy(st);
    };
    std::vector<std::thread> th;
    th.reserve(workers);
    for (int w = 1; w < workers; ++w) {
        try {
            th.emplace_back(work);
        } catch (const std::system_error &) {
            break;  // fewer workers: the queue is drained by those running
        }
    }
    work();
    for (auto &t : th) t.join();
    for (int32_t i = 0; i < count; ++i) {
        if (rc[i] != INSCY_OK || !out[i]) {
            const int code = rc[i] != INSCY_OK ? rc[i] : INSCY_ERR_DEVICE;
            const std::string m = "data set " + std::to_string(i) + ": " + (msg[i].empty() ? "not run" : msg[i]);
            for (int32_t j = 0; j < count; ++j) {
                delete out[j];
                out[j] = nullptr;
            }
            return fail(code, m);
        }
    }
    g_error.clear();
    return INSCY_OK;
}

}  // namespace inscy

using namespace inscy;

extern "C" {

int inscy_run_many(const float *const *X, const int64_t *n, const int32_t *d, const inscy_params *params,
                   int32_t count, int32_t mode, inscy_result **out) {
    if (!X || !n || !d || !params || !out || count < 1) return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    for (int32_t i = 0; i < count; ++i) {
        int rc = check_params(n[i], d[i], &params[i], mode);
        if (rc) return fail(rc, "data set " + std::to_string(i) + ": " + g_error);
        if (!X[i]) return fail(INSCY_ERR_ARGUMENT, "data set " + std::to_string(i) + ": X is NULL");
    }
    return run_many_impl(X, n, d, params, count, mode, true, out);
}

int inscy_run_many_device(const float *const *d_X, const int64_t *n, const int32_t *d, const inscy_params *params,
                          int32_t count, int32_t mode, inscy_result **out) {
    if (!d_X || !n || !d || !params || !out || count < 1) return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    for (int32_t i = 0; i < count; ++i) {
        int rc = check_params(n[i], d[i], &params[i], mode);
        if (rc) return fail(rc, "data set " + std::to_string(i) + ": " + g_error);
        if (!d_X[i]) return fail(INSCY_ERR_ARGUMENT, "data set " + std::to_string(i) + ": X is NULL");
    }
    return run_many_impl(d_X, n, d, params, count, mode, false, out);
}


int inscy_run_device(const float *d_X, int64_t n, int32_t d, const inscy_params *params, int32_t mode, void *stream,
                     inscy_result **out) {
    if (!out) return fail(INSCY_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    int rc = check_params(n, d, params, mode);
    if (rc) return rc;
    if (!d_X) return fail(INSCY_ERR_ARGUMENT, "X is NULL");
    return run_impl(d_X, n, d, params, mode, (hipStream_t) stream, out);
}

int inscy_run(const float *X, int64_t n, int32_t d, const inscy_params *params, int32_t mode, inscy_result **out) {
    if (!out) return fail(INSCY_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    int rc = check_params(n, d, params, mode);
    if (rc) return rc;
    if (!X) return fail(INSCY_ERR_ARGUMENT, "X is NULL");
    float *dX = nullptr;
    size_t bytes = (size_t) n * d * sizeof(float);
    hipError_t e = dev_alloc(&dX, bytes);
    if (e != hipSuccess) return fail(INSCY_ERR_MEMORY, std::string("device allocation: ") + hipGetErrorString(e));
    e = hipMemcpy(dX, X, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dev_free(dX);
        return fail(INSCY_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    rc = run_impl(dX, n, d, params, mode, nullptr, out);
    dev_free(dX);
    return rc;
}

int32_t inscy_result_num_subspaces(const inscy_result *res) { return res ? (int32_t) res->subspaces.size() : 0; }

int32_t inscy_result_subspace(const inscy_result *res, int32_t i, int32_t *dims_out) {
    if (!res || i < 0 || i >= (int32_t) res->subspaces.size()) return -1;
    const auto &s = res->subspaces[i];
    if (dims_out) std::copy(s.begin(), s.end(), dims_out);
    return (int32_t) s.size();
}

const int32_t *inscy_result_labels(const inscy_result *res, int32_t i) {
    if (!res || i < 0 || i >= (int32_t) res->subspaces.size() || !res->labels) return nullptr;
    return res->labels.get() + (size_t) i * res->n;
}

const int32_t *inscy_result_labels_all(const inscy_result *res) {
    if (!res || res->subspaces.empty()) return nullptr;
    return res->labels.get();
}

int64_t inscy_result_num_points(const inscy_result *res) { return res ? res->n : 0; }

int inscy_result_stats(const inscy_result *res, inscy_stats *out) {
    if (!res || !out) return fail(INSCY_ERR_ARGUMENT, "NULL argument");
    *out = res->stats;
    return INSCY_OK;
}

void inscy_result_free(inscy_result *res) { delete res; }

const char *inscy_last_error(void) { return g_error.c_str(); }

int inscy_set_pruner(const float *weights, float threshold, int32_t min_depth) {
    if (weights && (!std::isfinite(threshold) || min_depth < 0))
        return fail(INSCY_ERR_ARGUMENT, "threshold must be finite and min_depth >= 0");
    PrunerState &P = pruner_state();
    std::lock_guard<std::mutex> g(P.mu);
    if (!weights) {
        P.w.clear();
        ++P.version;
        return INSCY_OK;
    }
    std::vector<uint16_t> w(kMlpParams);
    for (int i = 0; i < kMlpParams; ++i) {
        if (!std::isfinite(weights[i])) return fail(INSCY_ERR_ARGUMENT, "weights must be finite");
        w[i] = bf16_rne(weights[i]);
    }
    P.w.swap(w);
    P.thr = threshold;
    P.min_depth = min_depth;
    ++P.version;
    return INSCY_OK;
}

int64_t inscy_result_pairs(const inscy_result *res, const float **features, const uint8_t **labels,
                           const int32_t **info, const int64_t **cost) {
    if (!res) return 0;
    if (features) *features = res->pair_feat.data();
    if (labels) *labels = res->pair_label.data();
    if (info) *info = res->pair_info.data();
    if (cost) *cost = res->pair_cost.data();
    return (int64_t) res->pair_label.size();
}

int inscy_mlp_forward(const uint16_t *d_params, const float *d_features, int64_t count, float *d_prob, void *stream) {
    if (!d_params || !d_features || !d_prob || count < 0) return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    hipError_t e = launch_mlp_predict(d_features, count, d_params, d_prob, (hipStream_t) stream);
    return e == hipSuccess ? INSCY_OK : fail(INSCY_ERR_DEVICE, hipGetErrorString(e));
}

int inscy_mlp_gradient(const uint16_t *d_params, const float *d_features, const float *d_labels,
                       const float *d_weights, int64_t count, float inv_weight_sum, float *d_partials,
                       int32_t nblocks, float *d_grad, void *stream) {
    if (!d_params || !d_features || !d_labels || !d_weights || !d_partials || !d_grad || count < 1 || nblocks < 1 ||
        !std::isfinite(inv_weight_sum))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    hipError_t e = launch_mlp_grad(d_features, d_labels, d_weights, count, inv_weight_sum, d_params, d_partials,
                                   nblocks, d_grad, (hipStream_t) stream);
    return e == hipSuccess ? INSCY_OK : fail(INSCY_ERR_DEVICE, hipGetErrorString(e));
}

int inscy_mlp_adam(float *d_master, const float *d_grad, float *d_m, float *d_v, uint16_t *d_params, int32_t step,
                   float lr, void *stream) {
    if (!d_master || !d_grad || !d_m || !d_v || !d_params || step < 1 || !(lr > 0.0f))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    hipError_t e = launch_mlp_adam(d_master, d_grad, d_m, d_v, d_params, step, lr, (hipStream_t) stream);
    return e == hipSuccess ? INSCY_OK : fail(INSCY_ERR_DEVICE, hipGetErrorString(e));
}

int inscy_mlp_round(const float *d_master, uint16_t *d_params, void *stream) {
    if (!d_master || !d_params) return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    hipError_t e = launch_mlp_round(d_master, d_params, (hipStream_t) stream);
    return e == hipSuccess ? INSCY_OK : fail(INSCY_ERR_DEVICE, hipGetErrorString(e));
}

int inscy_release_memory(void) {
    device_cache().release(-1);
    host_cache().release(-1);
    return INSCY_OK;
}

int inscy_profile(int on) {
    g_profile_on = on != 0;
    return INSCY_OK;
}

int inscy_set_tuning(int32_t max_batch, int32_t workers) {
    if (max_batch < 0 || max_batch > 65536 || workers < 0 || workers > 64)
        return fail(INSCY_ERR_ARGUMENT, "max_batch must be in [0, 65536] and workers in [0, 64]");
    g_max_batch = max_batch;
    g_workers = workers;
    g_error.clear();
    return INSCY_OK;
}

// Test hooks (not in the public header): the engine's exclusive prefix sum
// and its stable (uint16 key, int value) sort on device arrays, with the
// scratch sizes they need.
int64_t inscy_debug_scan_temp(int64_t items) {
    if (items <= 0 || items >= INT_MAX) return -1;
    return (int64_t) scan_temp_bytes((int) items);
}

int inscy_debug_exclusive_sum(const int *in, int *out, int64_t items, void *temp, int64_t temp_bytes, void *stream) {
    if (!in || !out || items <= 0 || items >= INT_MAX || temp_bytes < inscy_debug_scan_temp(items) ||
        (!temp && temp_bytes > 0))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    hipError_t e = launch_exclusive_sum(temp, (size_t) temp_bytes, in, out, (int) items, (hipStream_t) stream);
    return e == hipSuccess ? INSCY_OK : fail(INSCY_ERR_DEVICE, hipGetErrorString(e));
}

int64_t inscy_debug_sort_temp(int64_t n) {
    if (n <= 0 || n >= INT_MAX) return -1;
    return (int64_t) sort_temp_bytes((int) n);
}

int inscy_debug_sort_pairs(const uint16_t *keys, uint16_t *keys_tmp, uint16_t *keys_out, const int *vals,
                           int *vals_tmp, int *vals_out, int64_t n, int32_t bits, void *temp, int64_t temp_bytes,
                           void *stream) {
    if (!keys || !keys_tmp || !keys_out || !vals || !vals_tmp || !vals_out || !temp || n <= 0 || n >= INT_MAX ||
        bits < 1 || bits > 16 || temp_bytes < inscy_debug_sort_temp(n))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    hipError_t e = launch_sort_pairs_u16(keys, keys_tmp, keys_out, vals, vals_tmp, vals_out, (int) n, bits, temp,
                                         (size_t) temp_bytes, (hipStream_t) stream);
    return e == hipSuccess ? INSCY_OK : fail(INSCY_ERR_DEVICE, hipGetErrorString(e));
}

// Test hook (not in the public header): the density thresholds the engine
// compares with, out = {c(k), alpha(k), expDen(k), omega(k)} (host arithmetic
// only, no device call).
int inscy_debug_thresholds(int32_t k, float eps, int32_t n, float *out) {
    if (!out || k < 0 || k > kMaxDims || n < 1) return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    out[0] = ball(k);
    out[1] = alpha_thr(k, eps, n);
    out[2] = expden(k, eps, n);
    out[3] = omega_thr(k);
    return INSCY_OK;
}

// Test hook (not in the public header): record one trace row per evaluated
// restricted tree — (mask, lo, hi, points, survivors, pass, redundant).
int inscy_debug_trace(int on) {
    g_trace_on = on != 0;
    return INSCY_OK;
}

// Test hook (not in the public header): decide every phi density by the
// leaf-order recomputation instead of the double sum with its rounding band.
int inscy_debug_phi_exact(int on) {
    g_phi_all_exact = on != 0;
    return INSCY_OK;
}

// Test hook (not in the public header): run every neighbour walk of the
// level-synchronous driver for k known only at run time, as for a batch that
// mixes dimensionalities (no launch compiled for its k).
int inscy_debug_generic_walk(int on) {
    g_generic_walk = on != 0;
    return INSCY_OK;
}

// Test hook (not in the public header): bytes for the grid-index sorts and
// for the sorted columns a run keeps across batches (each); < 0: the batch
// budget.
int inscy_debug_keep_budget(int64_t bytes) {
    g_keep_budget = bytes;
    return INSCY_OK;
}

// Test hook (not in the public header): the ordered compaction of the query
// flags.  out[0, *nsel) receives the positions of the nonzero bytes of
// flags[0, nb) in ascending order.  Device pointers on `stream`: flags
// readable up to nb rounded up to 16 bytes, out of nb ints, temp of
// inscy_debug_flag_compact_temp(nb) bytes.
int64_t inscy_debug_flag_compact_temp(int64_t nb) {
    if (nb <= 0 || nb >= INT_MAX) return -1;
    return (int64_t) select_temp_bytes((int) nb);
}

int inscy_debug_flag_compact(const uint8_t *flags, int64_t nb, void *temp, int64_t temp_bytes, int *out, int *nsel,
                             void *stream) {
    if (!flags || !temp || !out || !nsel || nb <= 0 || nb >= INT_MAX ||
        temp_bytes < inscy_debug_flag_compact_temp(nb))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    hipError_t e = launch_flag_compact(flags, (size_t) nb, temp, (size_t) temp_bytes, out, nsel, (hipStream_t) stream);
    return e == hipSuccess ? INSCY_OK : fail(INSCY_ERR_DEVICE, hipGetErrorString(e));
}

int64_t inscy_debug_trace_rows(int64_t *out, int64_t cap) {
    int64_t m = (int64_t) g_trace.size();
    if (out)
        for (int64_t i = 0; i < std::min(m, cap); ++i) {
            const auto &t = g_trace[i];
            int64_t *o = out + 7 * i;
            o[0] = t.mask;
            o[1] = t.lo;
            o[2] = t.hi;
            o[3] = t.pre;
            o[4] = t.post;
            o[5] = t.pass;
            o[6] = t.redundant;
        }
    return m;
}

int inscy_discretise(const float *d_X, int64_t n, int32_t d, int32_t cells, float eps, uint8_t *d_cell_c,
                     uint16_t *d_cell_e, uint64_t *d_near_mask) {
    if (!d_X || !d_cell_c || !d_cell_e || !d_near_mask || n < 1 || d < 1 || d > kMaxDims || cells < 1 ||
        cells > 255 || !(eps > 0.0f))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    try {
        int *bad = nullptr;
        HIP_OK(hipMalloc(&bad, sizeof(int)));
        HIP_OK(hipMemset(bad, 0, sizeof(int)));
        HIP_OK(launch_discretise(d_X, (int) n, d, cells, eps, nullptr, d_cell_c, d_cell_e, d_near_mask, (d + 63) / 64,
                                 bad, 0));
        int h = 0;
        HIP_OK(hipMemcpy(&h, bad, sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipFree(bad));
        if (h) return fail(INSCY_ERR_ARGUMENT, "X must be finite with every value in [0, 1]");
        return INSCY_OK;
    } catch (...) {
        return fail_current();
    }
}

// Kernel-level entry points: the production engine's subspace pass on its own.
namespace {
dmask_t mask_from(const int32_t *dims, int32_t k, int d) {
    if (k > kMaxSubDims) throw Failure(INSCY_ERR_ARGUMENT, "a subspace has at most 64 dimensions");
    dmask_t m = 0;
    for (int i = 0; i < k; ++i) {
        if (dims[i] < 0 || dims[i] >= d || (i && dims[i] <= dims[i - 1]))
            throw Failure(INSCY_ERR_ARGUMENT, "dims must be ascending ids in [0, d)");
        m |= dbit(dims[i]);
    }
    return m;
}
}  // namespace

int inscy_neighbour_counts(const float *d_X, int64_t n, int32_t d, float eps, const int32_t *dims, int32_t k,
                           const int32_t *d_queries, int64_t nq, int32_t *d_out) {
    if (!d_X || !dims || k < 1 || k > d || !d_queries || !d_out || n < 1 || d < 1 || d > kMaxDims || !(eps > 0.0f))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    try {
        inscy_params p{eps, 1.0f, 1, 1, 1.0f, 1, 1, 0};
        Engine eng(p, (int) n, d, nullptr);
        eng.setup(d_X);
        const int *cnt = eng.exact_counts(mask_from(dims, k, d));
        HIP_OK(launch_gather(cnt, d_queries, (int) nq, d_out, nullptr));
        HIP_OK(hipDeviceSynchronize());
        return INSCY_OK;
    } catch (...) {
        return fail_current();
    }
}

int inscy_neighbour_counts_noalloc(const float *d_X, int64_t n, int32_t d, float eps, const int32_t *dims, int32_t k,
                                   const int32_t *d_queries, int64_t nq, int32_t *d_out, void *stream) {
    if (!d_X || !dims || k < 1 || k > d || (nq > 0 && (!d_queries || !d_out)) || nq < 0 || nq > INT_MAX || n < 1 ||
        n > (int64_t) INT_MAX / 2 || d < 1 || d > kMaxDims || !(eps > 0.0f) || !std::isfinite(eps))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    const double ec = std::ceil(1.0 / (double) eps);
    if (ec > 65535.0) return fail(INSCY_ERR_UNSUPPORTED, "neighborhood_size below 1/65535 is not supported");
    try {
        (void) mask_from(dims, k, d);
        const size_t need = noalloc_scratch_need((int) n, k, (int) ec);
        if (need > kNoAllocScratchBytes)
            return fail(INSCY_ERR_ARGUMENT, "n and k need " + std::to_string(need) + " bytes of scratch; the allocation-free "
                        "entry point holds " + std::to_string(kNoAllocScratchBytes) + " (use inscy_neighbour_counts)");
        SubspaceArg S{};
        S.k = k;
        for (int t = 0; t < k; ++t) S.dims[t] = dims[t];
        unsigned char *base = nullptr;
        HIP_OK(noalloc_scratch(&base));
        HIP_OK(launch_neighbour_counts_noalloc(base, d_X, (int) n, d, eps, sqrt_threshold(eps), S, d_queries, (int) nq,
                                               d_out, (hipStream_t) stream));
        return INSCY_OK;
    } catch (...) {
        return fail_current();
    }
}

int inscy_dense_components(const float *d_X, int64_t n, int32_t d, const inscy_params *p, const int32_t *dims,
                           int32_t k, int32_t *d_comp) {
    if (!d_X || !p || !dims || k < 1 || k > d || !d_comp || n < 1 || d < 1 || d > kMaxDims ||
        !(p->neighborhood_size > 0.0f) || p->number_of_cells < 1 || p->number_of_cells > 255)
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    if (!p->rectangular) return fail(INSCY_ERR_UNSUPPORTED, "rectangular=False is not implemented");
    try {
        Engine eng(*p, (int) n, d, nullptr);
        eng.setup(d_X);
        const int *comp = eng.subspace(mask_from(dims, k, d)).comp;
        HIP_OK(hipMemcpyAsync(d_comp, comp, (size_t) n * sizeof(int), hipMemcpyDeviceToDevice, nullptr));
        HIP_OK(hipDeviceSynchronize());
        return INSCY_OK;
    } catch (...) {
        return fail_current();
    }
}

}  // extern "C"

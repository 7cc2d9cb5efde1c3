// Internal declarations shared by the HIP kernels and the host driver of the
// MI355X INSCY engine.  Not part of the public ABI (include/inscy.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace inscy {

constexpr int kMaxDims = 64;  // subspace bitmasks are uint64

// Subspace argument passed by value to kernels: k ascending dimension ids.
struct SubspaceArg {
    int k;
    int dims[kMaxDims];
};

// Device-resident discretised data set.  All per-dimension arrays are
// dimension-major ([d][n]) so that a wave reading one dimension of 64
// consecutive points issues one coalesced load.
struct DevData {
    const float *Xt;        // [d][n] coordinates
    const uint8_t *cellC;   // [d][n] cell in the c-grid (ScyTreeNode.cu:100-102)
    const uint16_t *cellE;  // [d][n] cell in the ceil(1/eps)-grid (inscy_map.cpp:243-244)
    const int *sid1;        // [d][n] point ids bucketed by cellE of that dimension
    const int *off1;        // [d][ecells+1] bucket starts
    int n, d, cells, ecells;
};

// Largest float s with sqrtf(s) <= eps: the distance test fl(sqrt(sum)) <= eps
// of ClusteringCpu.cu:52-81 is then sum <= s_star without a device sqrt.
float sqrt_threshold(float eps);

// Launchers (all asynchronous on `st`; return hipError_t of the launch).
hipError_t launch_discretise(const float *X, int n, int d, int cells, float eps, float *Xt, uint8_t *cellC,
                             uint16_t *cellE, uint64_t *nearm, int *bad, hipStream_t st);
hipError_t launch_build_index(const uint16_t *cellE, int n, int d, int ecells, int *hist, int *off1, int *sid1,
                              hipStream_t st);
hipError_t launch_flags(const int *spx, const uint64_t *spm, const int *nsp, int nsp_max, const uint8_t *cellC_m,
                        int m, int *flags, hipStream_t st);
hipError_t launch_filter(const int *P, const int *nP, int nP_max, const int *spx, const uint64_t *spm,
                         const int *nsp, int nsp_max, const uint8_t *cellC_m, int m, int lo, int hi, int *outP,
                         int *outx, uint64_t *outm, int *counters, hipStream_t st);
hipError_t launch_counts(const DevData &D, const SubspaceArg &S, const int *queries, const int *nq, int nq_max,
                         float s_star, int *cnt, unsigned long long *cand, hipStream_t st);
hipError_t launch_weak(const int *P, const int *nP, int nP_max, const int *cnt, int thr, int *W, int *nW,
                       int *stamp, int tid, hipStream_t st);
hipError_t launch_path_prune(const DevData &D, uint64_t smask, const int *W, const int *nW, int nW_max,
                             const int *spx, const uint64_t *spm, const int *nsp, int nsp_max, const int *stamp,
                             int tid, unsigned long long *bits, uint64_t *table, uint64_t tmask, int *outx,
                             uint64_t *outm, int *nout, hipStream_t st);
hipError_t launch_all_queries(const int *cnt, int n, int *queries, int *nq, hipStream_t st);
hipError_t launch_components(const DevData &D, const SubspaceArg &S, float s_star, const int *cnt, int thr,
                             int *parent, int *comp, int *csize, hipStream_t st);
hipError_t launch_redundancy(const int *W, const int *nW, int nW_max, int const *const *compH,
                             int const *const *csizeH, uint8_t const *const *inresH, int nH, int *mn,
                             hipStream_t st);
hipError_t launch_seed(const int *W, const int *nW, int nW_max, const int *comp, uint8_t *lab, hipStream_t st);
hipError_t launch_collect(const uint8_t *lab, int n, const int *csize, int *out, int *nout, hipStream_t st);
hipError_t launch_join_check(const int *compS, const uint8_t *keep, const int *csizeS, int n, float r,
                             int const *const *compH, int const *const *csizeH, uint8_t const *const *inresH,
                             int nH, uint8_t *drop, hipStream_t st);
hipError_t launch_labels(const int *comp, const uint8_t *inres, int n, int *out, hipStream_t st);
hipError_t launch_fill_u8(uint8_t *p, uint8_t v, size_t count, hipStream_t st);
hipError_t launch_fill_i32(int *p, int v, size_t count, hipStream_t st);
hipError_t launch_fill_u64(uint64_t *p, uint64_t v, size_t count, hipStream_t st);
hipError_t launch_set_flags(uint8_t *flags, const int *ids, int count, hipStream_t st);
hipError_t launch_iota(int *p, int n, hipStream_t st);
hipError_t launch_root_paths(const uint64_t *nearm, int n, int *outx, uint64_t *outm, int *nout, hipStream_t st);

}  // namespace inscy

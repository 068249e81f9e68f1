// Variable-size batched LU / solve.
//
// MAGMA launches every step of a variable-size batch sized by the batch
// maximum and lets finished matrices exit early (src/zgetrf_vbatched.cpp:20-124),
// with host syncs to learn the maxima (magmablas/vbatched_check.cu:55,
// src/zgetrf_vbatched.cpp:366-371). Here the sizes are grouped once into
// n-buckets:
//   * square n <= 128: the register-tile kernel instance for ceil(n/16)
//     (fused with the solve when nrhs <= 16);
//   * larger or rectangular: the blocked path with a panel width per bucket,
// each bucket launched over an index map of its members. A plan holds the
// buckets and their device maps, so executing it allocates nothing, does not
// synchronise, and can be captured once into a hipGraph and replayed.
//
// Per-bucket kernel choices come from the caller (the project's autotuner) or
// default to the built-in rule.
#include "magma_internal.h"
#include "getrf_kernels.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

namespace mgi {

constexpr int kNumBuckets = 22;

// Bucket of a square order n (1..2048); -1 outside.
int vb_bucket_of(int n)
{
    if (n < 1) return -1;
    if (n <= 128) return (n + 15) / 16 - 1;                       // 0..7
    if (n <= 512) return 8 + (n - 128 + 63) / 64 - 1;            // 8..13
    if (n <= 1024) return 14 + (n - 512 + 127) / 128 - 1;        // 14..17
    if (n <= 2048) return 18 + (n - 1024 + 255) / 256 - 1;       // 18..21
    return -1;
}

// Largest order in a bucket.
int vb_bucket_max(int b)
{
    if (b < 8) return 16 * (b + 1);
    if (b < 14) return 128 + 64 * (b - 7);
    if (b < 18) return 512 + 128 * (b - 13);
    return 1024 + 256 * (b - 17);
}

struct VBucket {
    int id;          // bucket id, or -1 for the rectangular group
    int variant;     // resolved kernel choice (kVb* below)
    int count;
    int max_m, max_n, max_nrhs;
    int offset;      // into the concatenated map
};

enum VbVariant { kVbAuto = 0, kVbTile = 1, kVbBlocked32 = 2, kVbBlocked16 = 3, kVbBlocked8 = 4, kVbBlocked2x32 = 5 };

// panel width of a blocked variant (64 = the two-level form)
static int variant_nb(int v)
{
    return v == kVbBlocked2x32 ? 64 : v == kVbBlocked32 ? 32 : v == kVbBlocked16 ? 16 : v == kVbBlocked8 ? 8 : 0;
}

// Group a batch (host copies of the sizes) into buckets. Matrices with a zero
// dimension are left out (nothing to factor). Returns false if some size is
// beyond what the kernels handle.
static bool build_buckets(const int* h_m, const int* h_n, const int* h_nrhs, int batch, const int* variants,
                          bool solve, std::vector<VBucket>& out, std::vector<int>& map)
{
    std::vector<std::vector<int>> members(kNumBuckets + 1);
    for (int i = 0; i < batch; ++i) {
        const int m = h_m ? h_m[i] : h_n[i];
        const int n = h_n[i];
        if (m <= 0 || n <= 0) continue;
        if (m == n) {
            const int b = vb_bucket_of(n);
            if (b < 0) return false;
            members[b].push_back(i);
        } else {
            if (!getrf_blocked_supported(m)) return false;
            members[kNumBuckets].push_back(i);
        }
    }
    out.clear();
    map.clear();
    for (int b = 0; b <= kNumBuckets; ++b) {
        if (members[b].empty()) continue;
        VBucket vb{};
        vb.id = b < kNumBuckets ? b : -1;
        vb.count = (int)members[b].size();
        vb.offset = (int)map.size();
        for (int i : members[b]) {
            const int m = h_m ? h_m[i] : h_n[i];
            vb.max_m = std::max(vb.max_m, m);
            vb.max_n = std::max(vb.max_n, h_n[i]);
            vb.max_nrhs = std::max(vb.max_nrhs, h_nrhs ? h_nrhs[i] : 0);
            map.push_back(i);
        }
        int v = (variants && vb.id >= 0) ? variants[vb.id] : kVbAuto;
        const bool tile_ok = vb.id >= 0 && vb.max_n <= 128 && (!solve || vb.max_nrhs <= 16);
        if (v == kVbTile && !tile_ok) v = kVbAuto;
        if (variant_nb(v) != 0 && !blocked_nb_valid(vb.max_m, variant_nb(v))) v = kVbAuto;
        if (v == kVbAuto) v = tile_ok ? kVbTile : kVbBlocked32;
        if (variant_nb(v) != 0 && !blocked_nb_valid(vb.max_m, variant_nb(v))) {
            v = blocked_nb_valid(vb.max_m, 16) ? kVbBlocked16 : kVbBlocked8;
        }
        vb.variant = v;
        out.push_back(vb);
    }
    return true;
}

// Launch every bucket (asynchronous, allocation-free). d_map holds the
// concatenated member lists; m_arr may be null for square systems. Bucket i
// goes to streams[lane[i]] when lanes are given (the buckets touch disjoint
// systems and share no scratch, so they may run concurrently), else to
// stream.
constexpr int kVbLanes = 8;          // most concurrent streams of a captured plan
constexpr int kVbLanesDefault = 4;
constexpr int kVbSplitDefault = 4;   // pieces per blocked bucket

// integer from the environment, or fallback when unset
static int env_int(const char* name, int fallback)
{
    const char* e = std::getenv(name);
    return e ? std::atoi(e) : fallback;
}

static magma_int_t exec_buckets(const std::vector<VBucket>& buckets, const int* d_map, const int* m_arr,
                                const int* n_arr, const int* nrhs_arr, double** dA_array, const int* ldda_arr,
                                int** ipiv_array, double** dB_array, const int* lddb_arr, int* info_array,
                                bool solve, hipStream_t stream, const hipStream_t* streams = nullptr,
                                const int* lane = nullptr)
{
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        const VBucket& vb = buckets[bi];
        if (streams && lane) stream = streams[lane[bi]];
        const int* map = d_map + vb.offset;
        magma_int_t rc;
        if (vb.variant == kVbTile) {
            TileArgs a{0, 0, 0, 0, map, n_arr, nrhs_arr, ldda_arr, lddb_arr};
            const int nt = (vb.max_n + 15) / 16;
            rc = solve ? gesv_tile_launch_args(nt, a, dA_array, ipiv_array, info_array, dB_array, vb.count, stream)
                       : getrf_tile_launch_args(nt, a, dA_array, ipiv_array, info_array, vb.count, stream);
        } else {
            SizeArgs sz{0, 0, 0, m_arr ? m_arr : n_arr, n_arr, ldda_arr, map};
            rc = getrf_blocked_launch_sz(sz, vb.max_m, vb.max_n, dA_array, ipiv_array, info_array, vb.count, stream,
                                         variant_nb(vb.variant));
            if (rc == MAGMA_SUCCESS && solve && vb.max_nrhs > 0) {
                RsSizes rs{0, 0, 0, 0, vb.max_n, n_arr, nrhs_arr, ldda_arr, lddb_arr, map};
                rc = getrs_generic_launch_t<double>(0, rs, vb.max_nrhs, dA_array, ipiv_array, dB_array, vb.count,
                                                    stream);
            }
        }
        if (rc != MAGMA_SUCCESS) return rc;
    }
    return MAGMA_SUCCESS;
}

__global__ void vb_fill_kernel(int* p, int v, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

static magma_int_t copy_sizes(const int* d, int batch, std::vector<int>& h, hipStream_t stream)
{
    h.resize(batch);
    MGI_HIP_CHECK(hipMemcpyAsync(h.data(), d, sizeof(int) * batch, hipMemcpyDeviceToHost, stream));
    return MAGMA_SUCCESS;
}

}  // namespace mgi

// ---------------------------------------------------------------------------
// Plan object (no MAGMA counterpart)
// ---------------------------------------------------------------------------
struct magma_vbatched_plan {
    int device;
    int batch;
    bool solve;
    std::vector<mgi::VBucket> buckets;
    int* d_map;
    hipGraph_t graph;
    hipGraphExec_t graph_exec;
};

using namespace mgi;

extern "C" {

// include/magma_zvbatched.h:57. Sizes are read back to the host to form the
// buckets (MAGMA reads its maxima back the same way); host-synchronous at
// return.
magma_int_t magma_dgetrf_vbatched(magma_int_t* m, magma_int_t* n, double** dA_array, magma_int_t* ldda,
                                  magma_int_t** ipiv_array, magma_int_t* info_array,
                                  magma_int_t batchCount, magma_queue_t queue)
{
    if (batchCount < 0) {
        magma_xerbla(__func__, 7);
        return -7;
    }
    if (batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    hipStream_t s = queue->stream;
    std::vector<int> hm, hn, hl;
    magma_int_t rc;
    if ((rc = copy_sizes(m, batchCount, hm, s)) || (rc = copy_sizes(n, batchCount, hn, s)) ||
        (rc = copy_sizes(ldda, batchCount, hl, s)))
        return rc;
    MGI_HIP_CHECK(hipStreamSynchronize(s));
    // magmablas/vbatched_check.cu:18-62 argument codes
    magma_int_t arginfo = 0;
    for (int i = 0; i < batchCount && arginfo == 0; ++i) if (hm[i] < 0) arginfo = -1;
    for (int i = 0; i < batchCount && arginfo == 0; ++i) if (hn[i] < 0) arginfo = -2;
    for (int i = 0; i < batchCount && arginfo == 0; ++i) if (hl[i] < std::max(1, hm[i])) arginfo = -4;
    if (arginfo != 0) {
        magma_xerbla(__func__, -arginfo);
        return arginfo;
    }
    std::vector<VBucket> buckets;
    std::vector<int> map;
    if (!build_buckets(hm.data(), hn.data(), nullptr, batchCount, nullptr, false, buckets, map))
        return MAGMA_ERR_NOT_SUPPORTED;
    vb_fill_kernel<<<ceildiv(batchCount, 256), 256, 0, s>>>(info_array, 0, batchCount);
    MGI_LAUNCH_CHECK();
    if (!map.empty()) {
        int* d_map = (int*)queue_scratch(queue, sizeof(int) * map.size());
        if (d_map == nullptr) return MAGMA_ERR_DEVICE_ALLOC;
        MGI_HIP_CHECK(hipMemcpyAsync(d_map, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, s));
        rc = exec_buckets(buckets, d_map, m, n, nullptr, dA_array, ldda, ipiv_array, nullptr, nullptr, info_array,
                          false, s);
        if (rc != MAGMA_SUCCESS) return rc;
    }
    MGI_HIP_CHECK(hipStreamSynchronize(s));
    return 0;
}

// include/magma_zvbatched.h:73. No argument checks and no host sync: one
// blocked launch sequence over the whole batch (sizes stay on the device),
// sized by the caller's maxima. The workspace holds nothing; a query returns
// a minimal size.
magma_int_t magma_dgetrf_vbatched_max_nocheck_work(magma_int_t* m, magma_int_t* n, magma_int_t max_m,
                                                   magma_int_t max_n, magma_int_t max_minmn, magma_int_t max_mxn,
                                                   double** dA_array, magma_int_t* ldda,
                                                   magma_int_t** dipiv_array, magma_int_t* info_array,
                                                   void* work, magma_int_t* lwork, magma_int_t batchCount,
                                                   magma_queue_t queue)
{
    (void)max_minmn;
    (void)max_mxn;
    (void)work;
    if (lwork == nullptr) return -12;
    if (*lwork < 0) {
        *lwork = 128;
        return 0;
    }
    if (batchCount <= 0 || max_m <= 0 || max_n <= 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    SizeArgs sz{0, 0, 0, m, n, ldda, nullptr};
    return getrf_blocked_launch_sz(sz, max_m, max_n, dA_array, dipiv_array, info_array, batchCount, queue->stream);
}

magma_int_t magma_dgetrs_vbatched(magma_trans_t trans, magma_int_t* n, magma_int_t* nrhs, magma_int_t max_n,
                                  magma_int_t max_nrhs, double** dA_array, magma_int_t* ldda,
                                  magma_int_t** dipiv_array, double** dB_array, magma_int_t* lddb,
                                  magma_int_t batchCount, magma_queue_t queue)
{
    if (trans != MagmaNoTrans && trans != MagmaTrans && trans != MagmaConjTrans) {
        magma_xerbla(__func__, 1);
        return -1;
    }
    if (max_n < 0) {
        magma_xerbla(__func__, 4);
        return -4;
    }
    if (max_nrhs < 0) {
        magma_xerbla(__func__, 5);
        return -5;
    }
    if (batchCount <= 0 || max_n == 0 || max_nrhs == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    RsSizes rs{0, 0, 0, 0, max_n, n, nrhs, ldda, lddb, nullptr};
    return getrs_generic_launch_t<double>(trans == MagmaNoTrans ? 0 : 1, rs, max_nrhs, dA_array, dipiv_array,
                                          dB_array, batchCount, queue->stream);
}

// Reads n and nrhs back to the host to form the buckets, then runs
// asynchronously.
magma_int_t magma_dgesv_vbatched(magma_int_t* n, magma_int_t* nrhs, magma_int_t max_n, magma_int_t max_nrhs,
                                 double** dA_array, magma_int_t* ldda, magma_int_t** dipiv_array,
                                 double** dB_array, magma_int_t* lddb, magma_int_t* dinfo_array,
                                 magma_int_t batchCount, magma_queue_t queue)
{
    (void)max_n;
    (void)max_nrhs;
    if (batchCount < 0) {
        magma_xerbla(__func__, 11);
        return -11;
    }
    if (batchCount == 0) return 0;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipSetDevice(queue->device);
    hipStream_t s = queue->stream;
    std::vector<int> hn, hr;
    magma_int_t rc;
    if ((rc = copy_sizes(n, batchCount, hn, s)) || (rc = copy_sizes(nrhs, batchCount, hr, s))) return rc;
    MGI_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < batchCount; ++i) {
        if (hn[i] < 0) { magma_xerbla(__func__, 1); return -1; }
        if (hr[i] < 0) { magma_xerbla(__func__, 2); return -2; }
    }
    std::vector<VBucket> buckets;
    std::vector<int> map;
    if (!build_buckets(nullptr, hn.data(), hr.data(), batchCount, nullptr, true, buckets, map))
        return MAGMA_ERR_NOT_SUPPORTED;
    vb_fill_kernel<<<ceildiv(batchCount, 256), 256, 0, s>>>(dinfo_array, 0, batchCount);
    MGI_LAUNCH_CHECK();
    if (map.empty()) return 0;
    int* d_map = (int*)queue_scratch(queue, sizeof(int) * map.size());
    if (d_map == nullptr) return MAGMA_ERR_DEVICE_ALLOC;
    MGI_HIP_CHECK(hipMemcpyAsync(d_map, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, s));
    return exec_buckets(buckets, d_map, nullptr, n, nrhs, dA_array, ldda, dipiv_array, dB_array, lddb, dinfo_array,
                        true, s);
}

// ---- plan API --------------------------------------------------------------
magma_int_t magma_dgesv_vbatched_plan_create(const magma_int_t* h_n, const magma_int_t* h_nrhs,
                                             magma_int_t batchCount, const magma_int_t* h_variants,
                                             magma_queue_t queue, magma_vbatched_plan_t* plan_ptr)
{
    if (plan_ptr == nullptr) return -6;
    *plan_ptr = nullptr;
    if (batchCount < 0) return -3;
    if (queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    for (int i = 0; i < batchCount; ++i)
        if (h_n[i] < 0 || (h_nrhs && h_nrhs[i] < 0)) return -1;
    auto* p = new (std::nothrow) magma_vbatched_plan();
    if (p == nullptr) return MAGMA_ERR_HOST_ALLOC;
    p->device = queue->device;
    p->batch = batchCount;
    p->solve = h_nrhs != nullptr;
    p->d_map = nullptr;
    p->graph = nullptr;
    p->graph_exec = nullptr;
    std::vector<int> map;
    if (!build_buckets(nullptr, h_n, h_nrhs, batchCount, h_variants, p->solve, p->buckets, map)) {
        delete p;
        return MAGMA_ERR_NOT_SUPPORTED;
    }
    hipSetDevice(queue->device);
    if (!map.empty()) {
        if (hipMalloc(&p->d_map, sizeof(int) * map.size()) != hipSuccess) {
            delete p;
            return MAGMA_ERR_DEVICE_ALLOC;
        }
        if (hipMemcpy(p->d_map, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice) != hipSuccess) {
            hipFree(p->d_map);
            delete p;
            return MAGMA_ERR_UNKNOWN;
        }
    }
    *plan_ptr = p;
    return 0;
}

magma_int_t magma_vbatched_plan_num_buckets(magma_vbatched_plan_t plan)
{
    return plan ? (magma_int_t)plan->buckets.size() : 0;
}

// Bucket description: id (-1 for rectangular), variant, count, max_n, max_nrhs.
magma_int_t magma_vbatched_plan_bucket(magma_vbatched_plan_t plan, magma_int_t i, magma_int_t* info5)
{
    if (plan == nullptr || i < 0 || i >= (magma_int_t)plan->buckets.size() || info5 == nullptr) return -1;
    const VBucket& b = plan->buckets[i];
    info5[0] = b.id;
    info5[1] = b.variant;
    info5[2] = b.count;
    info5[3] = b.max_n;
    info5[4] = b.max_nrhs;
    return 0;
}

magma_int_t magma_dgesv_vbatched_plan_exec(magma_vbatched_plan_t plan, magma_int_t* n, magma_int_t* nrhs,
                                           double** dA_array, magma_int_t* ldda, magma_int_t** dipiv_array,
                                           double** dB_array, magma_int_t* lddb, magma_int_t* dinfo_array,
                                           magma_queue_t queue)
{
    if (plan == nullptr || queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    hipStream_t s = queue->stream;
    vb_fill_kernel<<<ceildiv(plan->batch, 256), 256, 0, s>>>(dinfo_array, 0, plan->batch);
    MGI_LAUNCH_CHECK();
    if (plan->buckets.empty()) return 0;
    return exec_buckets(plan->buckets, plan->d_map, nullptr, n, plan->solve ? nrhs : nullptr, dA_array, ldda,
                        dipiv_array, dB_array, lddb, dinfo_array, plan->solve, s);
}

// Record plan_exec with these arguments into a hipGraph owned by the plan
// (replacing an earlier one). Replays reuse the same device pointers.
magma_int_t magma_dgesv_vbatched_plan_capture(magma_vbatched_plan_t plan, magma_int_t* n, magma_int_t* nrhs,
                                              double** dA_array, magma_int_t* ldda, magma_int_t** dipiv_array,
                                              double** dB_array, magma_int_t* lddb, magma_int_t* dinfo_array,
                                              magma_queue_t queue)
{
    if (plan == nullptr || queue == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    if (plan->graph_exec) { hipGraphExecDestroy(plan->graph_exec); plan->graph_exec = nullptr; }
    if (plan->graph) { hipGraphDestroy(plan->graph); plan->graph = nullptr; }
    // Record on a private stream: the caller's stream may be the legacy
    // default stream, which cannot be captured. Work already queued on the
    // caller's stream is finished first so the recording sees settled inputs.
    MGI_HIP_CHECK(hipStreamSynchronize(queue->stream));
    hipStream_t s = nullptr;
    MGI_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        hipStreamDestroy(s);
        hip_report(e, "hipStreamBeginCapture", __FILE__, __LINE__);
        return MAGMA_ERR_UNKNOWN;
    }
    magma_int_t rc = MAGMA_SUCCESS;
    vb_fill_kernel<<<ceildiv(plan->batch, 256), 256, 0, s>>>(dinfo_array, 0, plan->batch);
    // The buckets are independent: they are recorded on up to kVbLanes
    // forked streams (greedy balance of count * n^3), so the graph lets a
    // small bucket run beside a large one; a blocked bucket is recorded as
    // `split` pieces of consecutive members so the panel phase of one piece
    // overlaps the update of another. The lanes join back into s.
    const int nlanes = std::max(1, std::min(kVbLanes, env_int("MAGMA_VB_LANES", kVbLanesDefault)));
    const int split = std::max(1, std::min(16, env_int("MAGMA_VB_SPLIT", kVbSplitDefault)));
    std::vector<VBucket> pieces;
    for (const VBucket& vb : plan->buckets) {
        const int np = vb.variant == kVbTile ? 1 : std::min(split, vb.count);
        for (int p = 0; p < np; ++p) {
            VBucket pc = vb;
            const int lo = (int)((long long)vb.count * p / np), hi = (int)((long long)vb.count * (p + 1) / np);
            pc.offset = vb.offset + lo;
            pc.count = hi - lo;
            if (pc.count > 0) pieces.push_back(pc);
        }
    }
    const size_t npc = pieces.size();
    std::vector<int> lane(npc, 0);
    {
        std::vector<double> load(nlanes, 0.0);
        std::vector<size_t> order(npc);
        for (size_t i = 0; i < npc; ++i) order[i] = i;
        auto cost = [&](size_t i) {
            const double m = pieces[i].max_n;
            return (double)pieces[i].count * m * m * m;
        };
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost(x) > cost(y); });
        for (size_t i : order) {
            const int best = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            lane[i] = best;
            load[best] += cost(i);
        }
    }
    hipStream_t ls[kVbLanes] = {};
    hipEvent_t fork = nullptr, joins[kVbLanes] = {};
    bool lanes_ok = hipEventCreateWithFlags(&fork, hipEventDisableTiming) == hipSuccess &&
                    hipEventRecord(fork, s) == hipSuccess;
    for (int q = 0; q < nlanes && lanes_ok; ++q) {
        lanes_ok = hipStreamCreateWithFlags(&ls[q], hipStreamNonBlocking) == hipSuccess &&
                   hipEventCreateWithFlags(&joins[q], hipEventDisableTiming) == hipSuccess &&
                   hipStreamWaitEvent(ls[q], fork, 0) == hipSuccess;
    }
    if (!lanes_ok) rc = MAGMA_ERR_UNKNOWN;
    if (rc == MAGMA_SUCCESS && npc > 0)
        rc = exec_buckets(pieces, plan->d_map, nullptr, n, plan->solve ? nrhs : nullptr, dA_array, ldda,
                          dipiv_array, dB_array, lddb, dinfo_array, plan->solve, s, ls, lane.data());
    for (int q = 0; q < nlanes; ++q) {
        if (ls[q] && joins[q] && hipEventRecord(joins[q], ls[q]) == hipSuccess) (void)hipStreamWaitEvent(s, joins[q], 0);
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    for (int q = 0; q < nlanes; ++q) {
        if (joins[q]) hipEventDestroy(joins[q]);
        if (ls[q]) hipStreamDestroy(ls[q]);
    }
    if (fork) hipEventDestroy(fork);
    hipStreamDestroy(s);
    if (rc != MAGMA_SUCCESS) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) { hip_report(e, "hipStreamEndCapture", __FILE__, __LINE__); return MAGMA_ERR_UNKNOWN; }
    plan->graph = g;
    MGI_HIP_CHECK(hipGraphInstantiate(&plan->graph_exec, g, nullptr, nullptr, 0));
    return 0;
}

magma_int_t magma_vbatched_plan_replay(magma_vbatched_plan_t plan, magma_queue_t queue)
{
    if (plan == nullptr || queue == nullptr || plan->graph_exec == nullptr) return MAGMA_ERR_NOT_INITIALIZED;
    MGI_HIP_CHECK(hipGraphLaunch(plan->graph_exec, queue->stream));
    return 0;
}

void magma_vbatched_plan_destroy(magma_vbatched_plan_t plan)
{
    if (plan == nullptr) return;
    if (plan->graph_exec) hipGraphExecDestroy(plan->graph_exec);
    if (plan->graph) hipGraphDestroy(plan->graph);
    if (plan->d_map) hipFree(plan->d_map);
    delete plan;
}

}  // extern "C"

// Batched OS-SART: many independent reconstructions of different volume,
// view-count and detector sizes advance together.
//
// One subset step of every problem is five phases -- z-prefix of x, forward
// projection, SART-weighted v-prefix of the residual, y-slab backprojection,
// x-slab backprojection fused with the SART update -- and each phase is ONE
// persistent launch over all problems: a grid sized to fill the GPU once,
// whose wavefronts pull (problem, work item) pairs from an atomic queue until
// the phase drains.  Small problems fill the gaps that large ones leave, where
// per-problem launches would leave most of the 256 CUs idle at every phase
// tail.  One iteration (every subset of every problem) is captured in a HIP
// graph and replayed; the queue counters are cleared by a kernel at the head
// of the graph.
//
// The arithmetic is that of the single-problem OS-SART (src/DD3_GPU_recon.h:
// 30-65; weighting src/SARTWeighting.h:22-91): the same per-wavefront bodies
// and the same float operations in the same order, so a batched iteration
// reproduces the per-problem iterations bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/ctlab_mi355x.h"
#include "dd_device.hpp"

#include <type_traits>
#include "plan.hpp"
#include "recon.hpp"
#include "status.hpp"

namespace ctlab {

// One problem's subset-k step: everything its phases read and write.
struct BatchProb {
    DevGeom g;                              // subset plan geometry
    CTLAB_GMEM void* cz;                    // z-prefix integral image (T, nvox)
    CTLAB_GMEM void* P;                     // projection v-prefix (T, subset nprj)
    CTLAB_GMEM float* prj;                  // subset projection work buffer
    const CTLAB_GMEM float* b;              // measured data of the subset, subset order
    const CTLAB_GMEM float* rowsum;         // A_s 1
    CTLAB_GMEM float* vol;                  // backprojection work volume
    CTLAB_GMEM float* x;                    // reconstruction, updated in place
    const CTLAB_GMEM float* colsum;         // A_s^T 1
    const CTLAB_GMEM int* views_y;
    const CTLAB_GMEM int* views_x;
    int ny, nx;
    const CTLAB_GMEM ColRec* rec;
    const CTLAB_GMEM unsigned short* start;
    int SJ, SR;
};

enum Phase { PH_ZPREFIX = 0, PH_PROJ, PH_PREFIX, PH_BACK_Y, PH_BACK_X, PH_COUNT };

// columns per wavefront and group in the streaming phases (z-prefix, v-prefix)
constexpr int kChunkColumns = 16;

// The descriptors are never written while a phase runs: read them through the
// constant address space so that they land in scalar registers.
template <typename T>
__device__ __forceinline__ T load_const(const T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(4))) T*)p;
#else
    return *p;
#endif
}

__device__ __forceinline__ BatchProb load_prob(const BatchProb* p)
{
    return load_const(p);
}

// Work queue of one launch: problem k owns items [off[k], off[k+1]).
struct Queue {
    const int* off;
    int* counter;
    int nprob;
};

// Pulls groups of 4*CHUNK consecutive items per block (one atomic) until the
// queue is empty; wavefront w takes items base + w*CHUNK ... + CHUNK-1 and
// calls f(k, local item) for each.  With CHUNK = 1 the four wavefronts of a
// block work on four neighbouring items at once (neighbouring views or slabs,
// which read nearly the same data through the CU's L1), as the blocks of the
// single-problem kernels do.  Groups arrive in increasing order, so each
// wavefront's problem index only moves forward.  The whole block sees the
// same group and leaves together once the counter passes the total.
template <int CHUNK, class F>
__device__ __forceinline__ void drain(const Queue q, F&& f)
{
    __shared__ int s_base;
    const int total = load_const(q.off + q.nprob);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int k = 0;
    for (;;) {
        if (threadIdx.x == 0) s_base = atomicAdd(q.counter, 4 * CHUNK);
        __syncthreads();
        const int base = __builtin_amdgcn_readfirstlane(s_base);
        __syncthreads();
        if (base >= total) break;
#pragma unroll
        for (int c = 0; c < CHUNK; ++c) {
            const int it = base + wv * CHUNK + c;
            if (it >= total) break;
            while (it >= load_const(q.off + k + 1)) ++k;
            f(k, it - load_const(q.off + k));
        }
    }
}

__global__ void bk_clear(int* __restrict__ c, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}

// item = voxel column (y*XN + x) of problem k
template <typename T>
__global__ void __launch_bounds__(256) bk_zprefix(const BatchProb* __restrict__ pb, Queue q)
{
    const int lane = threadIdx.x & 63;
    drain<kChunkColumns>(q, [&](int k, int col) {
        const BatchProb b = load_prob(pb + k);
        zprefix_wave(b.x, (CTLAB_GMEM T*)b.cz, b.g.mask, (long)col, b.g.ZN, lane);
    });
}

// item = view + PN * tile: concurrent wavefronts take neighbouring views of one
// column tile, which read nearly the same slab rows
template <typename T, int NU>
__global__ void __launch_bounds__(256) bk_proj(const BatchProb* __restrict__ pb, Queue q)
{
    __shared__ T sG[4][NU * kGStride];
    __shared__ __align__(16) double sFT[4][2 * (NU + 1)];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    drain<1>(q, [&](int k, int it) {
        const BatchProb b = load_prob(pb + k);
        if (b.g.cone)
            proj_walk_wave<T, NU, true>(b.g, (const CTLAB_GMEM T*)b.cz, b.prj, it % b.g.PN, it / b.g.PN, lane,
                                        sG[wv], sFT[wv]);
        else
            proj_walk_wave<T, NU, false>(b.g, (const CTLAB_GMEM T*)b.cz, b.prj, it % b.g.PN, it / b.g.PN, lane,
                                         sG[wv], sFT[wv]);
    });
}

// item = view * DNU + column: SART-weighted residual and its v-prefix
template <typename T>
__global__ void __launch_bounds__(256) bk_prefix(const BatchProb* __restrict__ pb, Queue q)
{
    const int lane = threadIdx.x & 63;
    drain<kChunkColumns>(q, [&](int k, int it) {
        const BatchProb b = load_prob(pb + k);
        prjprefix_wave<T>(b.g, b.prj, (CTLAB_GMEM T*)b.P, (long)it, lane, b.b, b.rowsum);
    });
}

// item = slab + NS * row tile.  y-slab pass overwrites vol; the x-slab pass
// adds its part and applies the SART update to x:
//   x += colSum > 1e-7 ? lambda * vol / colSum : 0, then clamp at 0 if nonneg
// (src/SARTWeighting.h:56-91, in the order of the unfused update kernel).
template <typename T, int NM, int YSLAB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
bk_back(const BatchProb* __restrict__ pb, Queue q, float lambda, int nonneg)
{
    constexpr int yslab = YSLAB;
    __shared__ T sG[4][72];
    __shared__ typename BackAcc<T, NM>::type sA[4][(NM + 1) * 64];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    drain<1>(q, [&](int k, int it) {
        const BatchProb b = load_prob(pb + k);
        const DevGeom& g = b.g;
        const int NS = yslab ? g.YN : g.XN, NR = yslab ? g.XN : g.YN;
        const int j = it % NS, rt = it / NS;
        T acc[NM];
        auto accum = [&](auto zm) {
            back_walk_accum<T, NM, decltype(zm)::value>(g, (const CTLAB_GMEM T*)b.P, yslab ? b.views_y : b.views_x,
                                                        yslab ? b.ny : b.nx, NR, j, rt, b.rec, b.start, b.SJ, b.SR,
                                                        lane, sG[wv], sA[wv], acc);
        };
        switch (zmode_of(g)) {
        case kZFan: accum(std::integral_constant<int, kZFan>()); break;
        case kZCone: accum(std::integral_constant<int, kZCone>()); break;
        default: accum(std::integral_constant<int, kZCone64>()); break;
        }
        CTLAB_GMEM float* __restrict__ vol = b.vol;
        if constexpr (YSLAB) {
            back_store_tile<T, NM>(g, acc, 1, NR, j, rt, lane, [&](long idx, float v) { vol[idx] = v; });
        } else {
            CTLAB_GMEM float* __restrict__ x = b.x;
            const CTLAB_GMEM float* __restrict__ cs = b.colsum;
            back_store_tile<T, NM>(g, acc, 0, NR, j, rt, lane, [&](long idx, float v) {
                const float t = vol[idx] + v;
                const float w = cs[idx];
                float r = x[idx] + (w > 1.0e-7f ? lambda * (t / w) : 0.0f);
                if (nonneg && r < 0.0f) r = 0.0f;
                x[idx] = r;
            });
        }
    });
}

}  // namespace ctlab

using namespace ctlab;

namespace {

constexpr int kTiles[3] = {8, 12, 16};

int chk(hipError_t e, const char* what) { return hip_check(e, what); }

int tile_index(int t)
{
    for (int i = 0; i < 3; ++i)
        if (kTiles[i] == t) return i;
    return -1;
}

}  // namespace

struct ctlab_batch {
    int device = 0;
    int n = 0;
    int sat_bits = 64;
    int K = 0;                              // largest subset count
    std::vector<ctlab_recon*> R;
    std::vector<BatchProb> desc;            // [K][n]
    BatchProb* d_desc = nullptr;
    std::vector<int> off;                   // [K][PH_COUNT][n+1]
    int* d_off = nullptr;
    int* d_cnt = nullptr;                   // [K][PH_COUNT]
    int nu = 16, nm = 16;
    int ncu = 0;
    std::vector<float*> x;                  // volumes the descriptors point at
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    float g_lambda = 0.0f;
    int g_nonneg = -1;
};

namespace {

void destroy(ctlab_batch* B)
{
    if (!B) return;
    DeviceGuard dg(B->device);
    if (B->graph) hipGraphExecDestroy(B->graph);
    if (B->stream) hipStreamDestroy(B->stream);
    if (B->d_desc) hipFree(B->d_desc);
    if (B->d_off) hipFree(B->d_off);
    if (B->d_cnt) hipFree(B->d_cnt);
    delete B;
}

// per-problem item counts of every (subset, phase) for the current tiles
void build_offsets(ctlab_batch* B)
{
    const int n = B->n;
    B->off.assign((size_t)B->K * PH_COUNT * (n + 1), 0);
    for (int k = 0; k < B->K; ++k)
        for (int ph = 0; ph < PH_COUNT; ++ph) {
            int* o = &B->off[((size_t)k * PH_COUNT + ph) * (n + 1)];
            for (int i = 0; i < n; ++i) {
                const ctlab_recon* r = B->R[i];
                long cnt = 0;
                if (k < r->osNum) {
                    const GeomIn& in = r->sub[k]->in;
                    switch (ph) {
                    case PH_ZPREFIX: cnt = (long)in.XN * in.YN; break;
                    case PH_PROJ: cnt = (long)((in.DNU + B->nu - 1) / B->nu) * in.PN; break;
                    case PH_PREFIX: cnt = (long)in.PN * in.DNU; break;
                    case PH_BACK_Y: cnt = (long)in.YN * ((in.XN + B->nm - 1) / B->nm); break;
                    default: cnt = (long)in.XN * ((in.YN + B->nm - 1) / B->nm); break;
                    }
                }
                o[i + 1] = o[i] + (int)cnt;
            }
        }
}

void fill_desc(ctlab_batch* B)
{
    const int n = B->n;
    B->desc.assign((size_t)B->K * n, BatchProb{});
    for (int k = 0; k < B->K; ++k)
        for (int i = 0; i < n; ++i) {
            const ctlab_recon* r = B->R[i];
            BatchProb& d = B->desc[(size_t)k * n + i];
            d.x = B->x.empty() ? nullptr : gmem(B->x[i]);
            if (k >= r->osNum) continue;   // no items in this step
            const ctlab_plan* S = r->sub[k];
            d.g = S->dev();
            d.cz = gmem(S->d_cz);
            d.P = gmem(S->d_P);
            d.prj = gmem(r->d_prj);
            d.b = gmem(r->d_bperm + r->sub_off[k]);
            d.rowsum = gmem(r->d_rowsum + r->sub_off[k]);
            d.vol = gmem(r->d_vol);
            d.colsum = gmem(r->d_colsum + (long)k * r->nvox);
            d.ny = (int)S->views_y.size();
            d.nx = (int)S->views_x.size();
            d.views_y = gmem(S->d_views);
            d.views_x = gmem(S->d_views + d.ny);
            d.rec = gmem((const ColRec*)S->d_rec);
            d.start = gmem(S->d_start);
            d.SJ = S->SJ;
            d.SR = S->SR;
        }
}

int upload_tables(ctlab_batch* B)
{
    fill_desc(B);
    build_offsets(B);
    if (chk(hipMemcpy(B->d_desc, B->desc.data(), B->desc.size() * sizeof(BatchProb), hipMemcpyHostToDevice),
            "upload batch descriptors") ||
        chk(hipMemcpy(B->d_off, B->off.data(), B->off.size() * sizeof(int), hipMemcpyHostToDevice),
            "upload batch offsets"))
        return -1;
    return 0;
}

// persistent grid: enough blocks to fill every CU at the kernel's occupancy
template <typename K>
int grid_for(K kernel, int ncu)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return per_cu * ncu;
}

template <typename T, int NU, int NM>
int enqueue_step(ctlab_batch* B, int k, float lambda, int nonneg, hipStream_t s, hipEvent_t* ev)
{
    const int n = B->n;
    auto queue = [&](int ph) {
        Queue q;
        q.off = B->d_off + ((size_t)k * PH_COUNT + ph) * (n + 1);
        q.counter = B->d_cnt + k * PH_COUNT + ph;
        q.nprob = n;
        return q;
    };
    auto total = [&](int ph) { return B->off[((size_t)k * PH_COUNT + ph) * (n + 1) + n]; };
    const BatchProb* d = B->d_desc + (size_t)k * n;
    auto mark = [&](int i) { if (ev) hipEventRecord(ev[i], s); };
    mark(0);
    if (total(PH_ZPREFIX) > 0)
        hipLaunchKernelGGL(bk_zprefix<T>, dim3(grid_for(bk_zprefix<T>, B->ncu)), dim3(256), 0, s, d,
                           queue(PH_ZPREFIX));
    if (total(PH_PROJ) > 0)
        hipLaunchKernelGGL((bk_proj<T, NU>), dim3(grid_for(bk_proj<T, NU>, B->ncu)), dim3(256), 0, s, d,
                           queue(PH_PROJ));
    mark(1);
    if (total(PH_PREFIX) > 0)
        hipLaunchKernelGGL(bk_prefix<T>, dim3(grid_for(bk_prefix<T>, B->ncu)), dim3(256), 0, s, d,
                           queue(PH_PREFIX));
    if (total(PH_BACK_Y) > 0)
        hipLaunchKernelGGL((bk_back<T, NM, 1>), dim3(grid_for(bk_back<T, NM, 1>, B->ncu)), dim3(256), 0, s, d,
                           queue(PH_BACK_Y), lambda, nonneg);
    if (total(PH_BACK_X) > 0)
        hipLaunchKernelGGL((bk_back<T, NM, 0>), dim3(grid_for(bk_back<T, NM, 0>, B->ncu)), dim3(256), 0, s, d,
                           queue(PH_BACK_X), lambda, nonneg);
    mark(2);
    return chk(hipGetLastError(), "batch launch");
}

template <typename T>
int enqueue_step_t(ctlab_batch* B, int k, float lambda, int nonneg, hipStream_t s, hipEvent_t* ev)
{
    const int a = tile_index(B->nu), c = tile_index(B->nm);
    switch (a * 3 + c) {
    case 0: return enqueue_step<T, 8, 8>(B, k, lambda, nonneg, s, ev);
    case 1: return enqueue_step<T, 8, 12>(B, k, lambda, nonneg, s, ev);
    case 2: return enqueue_step<T, 8, 16>(B, k, lambda, nonneg, s, ev);
    case 3: return enqueue_step<T, 12, 8>(B, k, lambda, nonneg, s, ev);
    case 4: return enqueue_step<T, 12, 12>(B, k, lambda, nonneg, s, ev);
    case 5: return enqueue_step<T, 12, 16>(B, k, lambda, nonneg, s, ev);
    case 6: return enqueue_step<T, 16, 8>(B, k, lambda, nonneg, s, ev);
    case 7: return enqueue_step<T, 16, 12>(B, k, lambda, nonneg, s, ev);
    default: return enqueue_step<T, 16, 16>(B, k, lambda, nonneg, s, ev);
    }
}

// one iteration: clear the queue counters, then every subset step
int enqueue_iteration(ctlab_batch* B, float lambda, int nonneg, hipStream_t s, hipEvent_t* ev)
{
    const int nc = B->K * PH_COUNT;
    hipLaunchKernelGGL(bk_clear, dim3((nc + 255) / 256), dim3(256), 0, s, B->d_cnt, nc);
    for (int k = 0; k < B->K; ++k) {
        hipEvent_t* e = ev ? ev + 3 * k : nullptr;
        const int rc = B->sat_bits == 32 ? enqueue_step_t<float>(B, k, lambda, nonneg, s, e)
                                         : enqueue_step_t<double>(B, k, lambda, nonneg, s, e);
        if (rc) return rc;
    }
    return 0;
}

int set_volumes(ctlab_batch* B, float* const* d_x)
{
    bool same = !B->x.empty();
    for (int i = 0; i < B->n && same; ++i) same = B->x[i] == d_x[i];
    if (same) return 0;
    B->x.assign(d_x, d_x + B->n);
    fill_desc(B);
    return chk(hipMemcpy(B->d_desc, B->desc.data(), B->desc.size() * sizeof(BatchProb), hipMemcpyHostToDevice),
               "upload batch descriptors");
}

}  // namespace

extern "C" {

ctlab_batch* ctlab_batch_create(ctlab_recon* const* recons, int n)
{
    clear_error();
    if (!recons || n < 1) {
        set_error(CTLAB_ERR_ARG, "need n >= 1 reconstructions");
        return nullptr;
    }
    for (int i = 0; i < n; ++i) {
        const ctlab_recon* r = recons[i];
        if (!r || !r->have_data) {
            set_error(CTLAB_ERR_ARG, "reconstruction %d is NULL or has no data", i);
            return nullptr;
        }
        if (r->device != recons[0]->device || r->full->sat_bits != recons[0]->full->sat_bits) {
            set_error(CTLAB_ERR_ARG, "all reconstructions must share one device and one SAT precision");
            return nullptr;
        }
        if (is_blocked(r->full->in)) {
            set_error(CTLAB_ERR_GEOMETRY,
                      "reconstruction %d: the batch takes volumes of at most %d slices and detectors of at most "
                      "%d rows (ZN %d, DNV %d); run it with ctlab_os_sart",
                      i, kBlock, kBlock, r->full->in.ZN, r->full->in.DNV);
            return nullptr;
        }
        for (const ctlab_plan* S : r->sub)
            if (S->sat_bits != r->full->sat_bits) {
                set_error(CTLAB_ERR_ARG, "reconstruction %d: subset plans differ in SAT precision", i);
                return nullptr;
            }
    }
    ctlab_batch* B = new ctlab_batch();
    B->device = recons[0]->device;
    B->n = n;
    B->sat_bits = recons[0]->full->sat_bits;
    B->R.assign(recons, recons + n);
    for (const ctlab_recon* r : B->R) B->K = std::max(B->K, r->osNum);
    DeviceGuard dg(B->device);
    hipDeviceProp_t prop;
    if (chk(hipGetDeviceProperties(&prop, B->device), "device properties")) { destroy(B); return nullptr; }
    B->ncu = prop.multiProcessorCount;
    if (chk(hipMalloc(&B->d_desc, sizeof(BatchProb) * (size_t)B->K * n), "hipMalloc") ||
        chk(hipMalloc(&B->d_off, sizeof(int) * (size_t)B->K * PH_COUNT * (n + 1)), "hipMalloc") ||
        chk(hipMalloc(&B->d_cnt, sizeof(int) * (size_t)B->K * PH_COUNT), "hipMalloc") ||
        chk(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking), "stream")) {
        destroy(B);
        return nullptr;
    }
    // the largest phase (z-prefix columns or prefix rows) must fit the int queue
    long items = 0;
    for (const ctlab_recon* r : B->R) {
        const GeomIn& in = r->full->in;
        items += std::max((long)in.XN * in.YN, (long)in.PN * in.DNU);
    }
    if (items > (1L << 30)) {
        set_error(CTLAB_ERR_ARG, "batch too large for the work queue");
        destroy(B);
        return nullptr;
    }
    if (upload_tables(B)) { destroy(B); return nullptr; }
    return B;
}

void ctlab_batch_destroy(ctlab_batch* B) { destroy(B); }

int ctlab_batch_set_tiles(ctlab_batch* B, int nu, int nm)
{
    clear_error();
    if (!B || tile_index(nu) < 0 || tile_index(nm) < 0) {
        set_error(CTLAB_ERR_ARG, "tile sizes must be 8, 12 or 16");
        return CTLAB_ERR_ARG;
    }
    if (B->nu == nu && B->nm == nm) return CTLAB_OK;
    DeviceGuard dg(B->device);
    B->nu = nu;
    B->nm = nm;
    if (B->graph) { hipGraphExecDestroy(B->graph); B->graph = nullptr; }
    if (chk(hipDeviceSynchronize(), "sync")) return ctlab_last_error();
    build_offsets(B);
    if (chk(hipMemcpy(B->d_off, B->off.data(), B->off.size() * sizeof(int), hipMemcpyHostToDevice),
            "upload batch offsets"))
        return ctlab_last_error();
    return CTLAB_OK;
}

int ctlab_batch_get_tiles(const ctlab_batch* B, int* nu, int* nm)
{
    clear_error();
    if (!B || !nu || !nm) {
        set_error(CTLAB_ERR_ARG, "NULL argument");
        return CTLAB_ERR_ARG;
    }
    *nu = B->nu;
    *nm = B->nm;
    return CTLAB_OK;
}

int ctlab_batch_os_sart(ctlab_batch* B, float* const* d_x, int iters, float lambda, int nonneg, int use_graph,
                        void* stream)
{
    clear_error();
    if (!B || !d_x || iters < 0) {
        set_error(CTLAB_ERR_ARG, "need batch, volumes and iters >= 0");
        return CTLAB_ERR_ARG;
    }
    for (int i = 0; i < B->n; ++i)
        if (!d_x[i]) {
            set_error(CTLAB_ERR_ARG, "volume %d is NULL", i);
            return CTLAB_ERR_ARG;
        }
    DeviceGuard dg(B->device);
    hipStream_t s = (hipStream_t)stream;
    // the descriptors are read by kernels already queued: finish them first
    if (chk(hipStreamSynchronize(B->stream), "sync") || chk(hipStreamSynchronize(s), "sync")) return ctlab_last_error();
    if (set_volumes(B, d_x)) return ctlab_last_error();
    if (!use_graph) {
        for (int it = 0; it < iters; ++it)
            if (enqueue_iteration(B, lambda, nonneg, s, nullptr)) return ctlab_last_error();
        return CTLAB_OK;
    }
    if (!B->graph || B->g_lambda != lambda || B->g_nonneg != nonneg) {
        if (B->graph) { hipGraphExecDestroy(B->graph); B->graph = nullptr; }
        hipGraph_t gr = nullptr;
        if (chk(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal), "begin capture"))
            return ctlab_last_error();
        const int rc = enqueue_iteration(B, lambda, nonneg, B->stream, nullptr);
        hipError_t e = hipStreamEndCapture(B->stream, &gr);
        if (rc || chk(e, "end capture")) { if (gr) hipGraphDestroy(gr); return ctlab_last_error(); }
        e = hipGraphInstantiate(&B->graph, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        if (chk(e, "graph instantiate")) return ctlab_last_error();
        B->g_lambda = lambda;
        B->g_nonneg = nonneg;
    }
    hipEvent_t ev;
    if (chk(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event")) return ctlab_last_error();
    hipEventRecord(ev, s);
    hipStreamWaitEvent(B->stream, ev, 0);
    for (int it = 0; it < iters; ++it)
        if (chk(hipGraphLaunch(B->graph, B->stream), "graph launch")) { hipEventDestroy(ev); return ctlab_last_error(); }
    hipEventRecord(ev, B->stream);
    hipStreamWaitEvent(s, ev, 0);
    hipEventDestroy(ev);
    return CTLAB_OK;
}

// Times one iteration per tile size (projector phases with nu = t, backprojector
// phases with nm = t, t = 8, 12, 16) on the given volumes, which are updated,
// and keeps the fastest of each.  ms_out (6 floats or NULL): projection phases
// (z-prefix + projection) for nu = 8/12/16, then backprojection phases
// (prefix + both passes + update) for nm = 8/12/16.
int ctlab_batch_autotune(ctlab_batch* B, float* const* d_x, float lambda, float* ms_out)
{
    clear_error();
    if (!B || !d_x) {
        set_error(CTLAB_ERR_ARG, "need batch and volumes");
        return CTLAB_ERR_ARG;
    }
    for (int i = 0; i < B->n; ++i)
        if (!d_x[i]) {
            set_error(CTLAB_ERR_ARG, "volume %d is NULL", i);
            return CTLAB_ERR_ARG;
        }
    DeviceGuard dg(B->device);
    if (chk(hipDeviceSynchronize(), "sync") || set_volumes(B, d_x)) return ctlab_last_error();
    const int nev = 3 * B->K;
    std::vector<hipEvent_t> ev(nev, nullptr);
    for (auto& e : ev)
        if (chk(hipEventCreate(&e), "event")) return ctlab_last_error();
    double fwd[3] = {0, 0, 0}, bwd[3] = {0, 0, 0};
    int rc = 0;
    const int nu0 = B->nu, nm0 = B->nm;
    for (int pass = 0; pass < 4 && !rc; ++pass) {   // pass 0 warms up
        const int t = kTiles[pass == 0 ? 2 : pass - 1];
        rc = ctlab_batch_set_tiles(B, t, t);
        if (rc) break;
        rc = enqueue_iteration(B, lambda, 0, B->stream, ev.data());
        if (rc || chk(hipStreamSynchronize(B->stream), "autotune run")) { rc = -1; break; }
        if (pass == 0) continue;
        double f = 0, b = 0;
        for (int k = 0; k < B->K; ++k) {
            float a = 0, c = 0;
            hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]);
            hipEventElapsedTime(&c, ev[3 * k + 1], ev[3 * k + 2]);
            f += a;
            b += c;
        }
        fwd[pass - 1] = f;
        bwd[pass - 1] = b;
    }
    for (auto& e : ev) hipEventDestroy(e);
    if (rc) {
        ctlab_batch_set_tiles(B, nu0, nm0);
        return ctlab_last_error() ? ctlab_last_error() : CTLAB_ERR_HIP;
    }
    const int bu = (int)(std::min_element(fwd, fwd + 3) - fwd);
    const int bm = (int)(std::min_element(bwd, bwd + 3) - bwd);
    if (ms_out)
        for (int i = 0; i < 3; ++i) {
            ms_out[i] = (float)fwd[i];
            ms_out[3 + i] = (float)bwd[i];
        }
    return ctlab_batch_set_tiles(B, kTiles[bu], kTiles[bm]);
}

int ctlab_batch_size(const ctlab_batch* B) { return B ? B->n : 0; }

}  // extern "C"

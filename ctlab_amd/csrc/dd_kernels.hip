// Branchless distance-driven projector / backprojector kernels for gfx950.
//
// Projector (DD3Proj_gpu prjMode 0, src/DD3_GPU_Proj.h:41-44; fan-beam
// DD2_multiGPU method 0, src/multiSlices_ker.cuh:35-38).
//   For every view and slab the footprint integral of detector cell (u,v) is a
//   difference of a summed-area table (SAT) of the slab read at the four
//   footprint corners (the reference reads an FP64 SAT, src/utilities.cuh:
//   231-254).  The SAT is built in two separable passes:
//     z: one pass over the volume per call writes the inclusive z-prefix of
//        every voxel column (kernel zprefix);
//     row axis: inside the projector each wavefront walks the rows of the
//        slab in increasing row coordinate, keeping the running row-sum of the
//        z-prefix rows in registers (lane = z), so the SAT row at any
//        fractional row coordinate is S + frac * c.  Differencing two such
//        rows gives the z-prefix of the column footprint, which lane v then
//        reads at its two row-boundary z coordinates (cross-lane gathers).
//   The walk starts from a local origin for every (view, column tile, slab),
//   so the running sums stay small; in FP64 the cancellation of the final
//   differences is far below FP32 output rounding.
// Backprojector (DD3Back_gpu prjMode 0, src/DD3_GPU_Back.h:40-44; DD2_multiGPU
// method 2): the exact transpose.  The per-column v-prefix of y*L (kernel
// prjprefix) is read at each voxel's z-boundaries mapped onto the detector
// rows; the row-axis overlap weights come from a piecewise-linear prefix over
// detector columns evaluated at the integer row lattice points.  Each
// wavefront owns a tile of voxels of one slab and loops over all views of one
// slab orientation, so no atomics are needed.
#include "dd_device.hpp"
#include "plan.hpp"

#include <algorithm>
#include <type_traits>

namespace ctlab {

// One wavefront per voxel column (lane = z; ZN <= 64), looping over columns
// when the grid holds fewer wavefronts: coalesced read of the column,
// wavefront scan, coalesced write.
template <typename T>
__global__ void __launch_bounds__(256)
zprefix_kernel(const float* __restrict__ vol, T* __restrict__ cz,
               const unsigned char* __restrict__ mask, long ncol, int ZN)
{
    for (long col = blockIdx.x * 4L + (threadIdx.x >> 6); col < ncol; col += gridDim.x * 4L)
        zprefix_wave(gmem(vol), gmem(cz), gmem(mask), col, ZN, threadIdx.x & 63);
}

// Projector walk.  Block = 4 wavefronts = 4 consecutive views of one tile of
// NU detector columns (neighbouring views read nearly the same slab rows);
// each wavefront loops over all slabs of its view.  Blocks are dealt to the 8
// XCDs round-robin; XCD k takes the view groups [k*G8, (k+1)*G8) of every
// tile (G8 = ceil(groups / 8)), view group fastest, so that the blocks an XCD
// runs at once hold neighbouring views and share slab rows through its L2
// (dealing consecutive groups to different XCDs left each XCD's blocks 32
// views apart: 36 % L2 hits at cfg2).
template <typename T, int NU>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(8))) dd_proj_walk(DevGeom g, const T* __restrict__ cz, float* __restrict__ prj,
                                                      int ntiles, int G8)
{
    __shared__ T sG[4][NU * kGStride];
    __shared__ __align__(16) double sFT[4][2 * (NU + 1)];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned w = blockIdx.x >> 3;                       // index within the XCD
    const int tile = (int)(w / (unsigned)G8);
    const int grp = (int)(blockIdx.x & 7) * G8 + (int)(w % (unsigned)G8);
    const int p = grp * 4 + wv;
    if (tile >= ntiles || p >= g.PN) return;
    if (g.cone) proj_walk_wave<T, NU, true>(g, gmem(cz), gmem(prj), p, tile, threadIdx.x & 63, sG[wv], sFT[wv]);
    else proj_walk_wave<T, NU, false>(g, gmem(cz), gmem(prj), p, tile, threadIdx.x & 63, sG[wv], sFT[wv]);
}

// v-prefix of y*L for the backprojector (the same values as prjprefix_wave):
// one wavefront per kPrefixCols consecutive (view, column) rows, whose loads
// are all issued before the first scan, so that they are in flight together.
constexpr int kPrefixCols = 4;
template <typename T>
__global__ void __launch_bounds__(256)
prjprefix_kernel(DevGeom g, const float* __restrict__ prj, T* __restrict__ P)
{
    const long n = (long)g.PN * g.DNU;
    const int lane = threadIdx.x & 63, DNV = g.DNV;
    const int vl = min(lane, DNV - 1);
    const double ezv = g.ez[vl];
    for (long idx0 = (blockIdx.x * 4L + (threadIdx.x >> 6)) * kPrefixCols; idx0 < n;
         idx0 += gridDim.x * 4L * kPrefixCols) {
        // every load of the wave's rows first (rows past the last read the last one)
        float y[kPrefixCols];
        double e2[kPrefixCols], ta[kPrefixCols];
#pragma unroll
        for (int c = 0; c < kPrefixCols; ++c) {
            const long idx = min(idx0 + c, n - 1);
            y[c] = prj[idx * DNV + vl];
            e2[c] = g.exy2[idx];
            ta[c] = g.tau[idx];
        }
#pragma unroll
        for (int c = 0; c < kPrefixCols; ++c) {
            const long idx = idx0 + c;
            const double L = g.dx * sqrt(e2[c] + ezv * ezv) * fabs(ta[c]);
            const T r = wave_inclusive_scan(lane < DNV ? (T)((double)y[c] * L) : T(0));
            if (idx < n && lane < DNV) P[idx * DNV + lane] = r;
        }
    }
}

// First walk-order column whose upper boundary lies above row r0 = 8*r8, per
// (view, slab, r8); 0xFFFF when every column lies below r0.
__global__ void __launch_bounds__(256)
back_start_kernel(DevGeom g, const ColRec* __restrict__ rec, unsigned short* __restrict__ start, int SJ, int SR)
{
    const long n = (long)g.PN * SJ * SR;
    for (long idx = blockIdx.x * 256L + threadIdx.x; idx < n; idx += gridDim.x * 256L) {
        const int r8 = (int)(idx % SR);
        const long pj = idx / SR;
        const int j = (int)(pj % SJ);
        const int p = (int)(pj / SJ);
        const ViewParams vp = g.vp[p];
        const int DNU = g.DNU;
        const double r0 = 8.0 * r8;
        if (j >= vp.NS || r0 >= vp.NR) { start[idx] = 0xFFFF; continue; }
        const double rs = fabs((vp.s0 + j * g.dx) / g.dx);
        const ColRec* R = rec + (long)p * (DNU + 1);
        // the records' slopes carry the sign of the source-to-slab distance
        auto A = [&](int q) { return vp.rbase + rs * (q < DNU ? R[q].klo : R[DNU - 1].khi); };
        if (A(DNU) <= r0) { start[idx] = 0xFFFF; continue; }
        int lo = 1, hi = DNU;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (A(mid) > r0) hi = mid; else lo = mid + 1;
        }
        start[idx] = (unsigned short)(lo - 1);
    }
}

// Block = 4 wavefronts = 4 consecutive slabs of one tile of NM rows (their
// detector-column ranges nearly coincide); each wavefront loops over the views
// of one slab orientation; lane = z.  Per view: the start column comes from
// the table above (prefetched one view ahead), then columns are consumed in
// walk order with the next column's record and P row in flight.
//
// Both slab orientations run in ONE launch: the first nby blocks take the
// y-slab views and write vol_y, the rest take the x-slab views and write vol_x
// (a scratch volume); back_add then sums them.  One grid instead of two
// dependent launches leaves one tail of partly filled CUs instead of two.
template <typename T, int NM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
dd_back_walk(DevGeom g, const T* __restrict__ P, float* __restrict__ vol_y, float* __restrict__ vol_x,
             const int* __restrict__ views_y, int ny, const int* __restrict__ views_x, int nx, int nby,
             int ntr_y, int ntr_x, const ColRec* __restrict__ rec, const unsigned short* __restrict__ start,
             int SJ, int SR)
{
    __shared__ T sG[4][72];
    __shared__ typename BackAcc<T, NM>::type sA[4][(NM + 1) * 64];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ys = blockIdx.x < (unsigned)nby;
    const int b = ys ? blockIdx.x : blockIdx.x - nby;
    const int NS = ys ? g.YN : g.XN, NR = ys ? g.XN : g.YN, ntr = ys ? ntr_y : ntr_x;
    const int rt = b % ntr;
    const int j = (b / ntr) * 4 + wv;
    if (j >= NS) return;
    auto walk = [&](auto zm) {
        back_walk_wave<T, NM, decltype(zm)::value>(g, gmem(P), gmem(ys ? vol_y : vol_x), gmem(ys ? views_y : views_x),
                                                   ys ? ny : nx, ys, 0, NR, j, rt, gmem(rec), gmem(start), SJ, SR,
                                                   threadIdx.x & 63, sG[wv], sA[wv]);
    };
    switch (zmode_of(g)) {
    case kZFan: walk(std::integral_constant<int, kZFan>()); break;
    case kZCone: walk(std::integral_constant<int, kZCone>()); break;
    default: walk(std::integral_constant<int, kZCone64>()); break;
    }
}

// vol += add (n floats)
__global__ void __launch_bounds__(256) back_add(float* __restrict__ vol, const float* __restrict__ add, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        vol[i] += add[i];
}

// Row-block staging for volumes deeper than 64 slices and detectors with more
// than 64 rows: copy a block of nd consecutive elements out of every row of
// length ns (voxel columns, detector columns), and write one back.
__global__ void __launch_bounds__(256) rows_gather(const float* __restrict__ src, float* __restrict__ dst, long nrows,
                                                   int ns, int off, int nd)
{
    const long n = nrows * nd;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const long r = i / nd;
        dst[i] = src[r * ns + off + (i - r * nd)];
    }
}

__global__ void __launch_bounds__(256) rows_scatter(const float* __restrict__ src, float* __restrict__ dst, long nrows,
                                                    int ns, int nd, int off, int add)
{
    const long n = nrows * ns;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const long r = i / ns;
        const long o = r * nd + off + (i - r * ns);
        const float v = src ? src[i] : 0.0f;
        dst[o] = add ? dst[o] + v : v;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned stream_grid(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

// Grids of the kernels that loop over their items (one item per thread, or per
// wavefront: 4 per block): every item's block when there are few enough,
// else kMaxBlocks.  A launch holds at most 2^32 work-items, which a volume of
// 8192 x 8192 columns reaches with one wavefront per column.
constexpr long kMaxBlocks = 1L << 20;
static unsigned item_grid(long n, int per_block)
{
    return (unsigned)std::max<long>(1, std::min<long>((n + per_block - 1) / per_block, kMaxBlocks));
}

hipError_t launch_rows_gather(const float* src, float* dst, long nrows, int ns, int off, int nd, hipStream_t s)
{
    hipLaunchKernelGGL(rows_gather, dim3(stream_grid(nrows * nd)), dim3(256), 0, s, src, dst, nrows, ns, off, nd);
    return hipGetLastError();
}

hipError_t launch_rows_scatter(const float* src, float* dst, long nrows, int ns, int nd, int off, int add,
                               hipStream_t s)
{
    hipLaunchKernelGGL(rows_scatter, dim3(stream_grid(nrows * ns)), dim3(256), 0, s, src, dst, nrows, ns, nd, off,
                       add);
    return hipGetLastError();
}

template <typename T, int NU>
static void launch_proj_walk(const DevGeom& g, const T* cz, float* d_prj, hipStream_t s)
{
    const int ntiles = (g.DNU + NU - 1) / NU;
    const int G8 = ((g.PN + 3) / 4 + 7) / 8;                  // view groups per XCD
    hipLaunchKernelGGL((dd_proj_walk<T, NU>), dim3(8u * (unsigned)G8 * (unsigned)ntiles), dim3(256), 0, s, g, cz,
                       d_prj, ntiles, G8);
}

template <typename T, int NM>
static void launch_back_walk(const DevGeom& g, const T* P, float* d_vol, float* d_vol_x, const int* vy, int ny,
                             const int* vx, int nx, const ColRec* rec, const unsigned short* start, int SJ, int SR,
                             hipStream_t s)
{
    const int ntr_y = (g.XN + NM - 1) / NM, ntr_x = (g.YN + NM - 1) / NM;
    const int nby = ntr_y * ((g.YN + 3) / 4), nbx = ntr_x * ((g.XN + 3) / 4);
    hipLaunchKernelGGL((dd_back_walk<T, NM>), dim3((unsigned)(nby + nbx)), dim3(256), 0, s, g, P, d_vol, d_vol_x,
                       vy, ny, vx, nx, nby, ntr_y, ntr_x, rec, start, SJ, SR);
}

template <typename T>
hipError_t launch_proj(const DevGeom& g, const float* d_vol, void* d_cz, float* d_prj, hipStream_t s)
{
    if (is_fan2d(g) && g.ZN <= kFanProjSlices) return launch_fan2d_proj<T>(g, d_vol, d_cz, d_prj, s);
    T* cz = static_cast<T*>(d_cz);
    const long ncol = (long)g.XN * g.YN;
    hipLaunchKernelGGL(zprefix_kernel<T>, dim3(item_grid(ncol, 4)), dim3(256), 0, s,
                       d_vol, cz, (const unsigned char*)g.mask, ncol, g.ZN);
    switch (g.tile_nu) {
    case 8: launch_proj_walk<T, 8>(g, cz, d_prj, s); break;
    case 12: launch_proj_walk<T, 12>(g, cz, d_prj, s); break;
    default: launch_proj_walk<T, 16>(g, cz, d_prj, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_back_start(const DevGeom& g, const void* d_rec, unsigned short* d_start, int SJ, int SR,
                             hipStream_t s)
{
    const long n = (long)g.PN * SJ * SR;
    hipLaunchKernelGGL(back_start_kernel, dim3(item_grid(n, 256)), dim3(256), 0, s, g,
                       (const ColRec*)d_rec, d_start, SJ, SR);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_back(const DevGeom& g, const float* d_prj, void* d_P, float* d_vol, float* d_scratch,
                       const int* d_views_y, int ny, const int* d_views_x, int nx, const void* d_rec,
                       const unsigned short* d_start, int SJ, int SR, hipStream_t s)
{
    const ColRec* rec = (const ColRec*)d_rec;
    T* P = static_cast<T*>(d_P);
    const long ncu = (long)g.PN * g.DNU;
    hipLaunchKernelGGL(prjprefix_kernel<T>, dim3(item_grid(ncu, 4 * kPrefixCols)), dim3(256), 0, s,
                       g, d_prj, P);
    // y-slab views (slab = y, rows along x) into d_vol, x-slab views into d_scratch, one launch
    switch (g.tile_nm) {
    case 8: launch_back_walk<T, 8>(g, P, d_vol, d_scratch, d_views_y, ny, d_views_x, nx, rec, d_start, SJ, SR, s); break;
    case 12: launch_back_walk<T, 12>(g, P, d_vol, d_scratch, d_views_y, ny, d_views_x, nx, rec, d_start, SJ, SR, s); break;
    default: launch_back_walk<T, 16>(g, P, d_vol, d_scratch, d_views_y, ny, d_views_x, nx, rec, d_start, SJ, SR, s); break;
    }
    const long nvox = (long)g.XN * g.YN * g.ZN;
    hipLaunchKernelGGL(back_add, dim3((unsigned)std::min<long>((nvox + 255) / 256, 8192)), dim3(256), 0, s, d_vol,
                       d_scratch, nvox);
    return hipGetLastError();
}

template hipError_t launch_proj<double>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_proj<float>(const DevGeom&, const float*, void*, float*, hipStream_t);
template hipError_t launch_back<double>(const DevGeom&, const float*, void*, float*, float*, const int*, int,
                                        const int*, int, const void*, const unsigned short*, int, int, hipStream_t);
template hipError_t launch_back<float>(const DevGeom&, const float*, void*, float*, float*, const int*, int,
                                       const int*, int, const void*, const unsigned short*, int, int, hipStream_t);

}  // namespace ctlab
